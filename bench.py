"""Benchmark of the MI355X structure-factor path.

Headline workload (the precision configuration): fcc a = 4.0 A, 40x40x40
cells (256,000 atoms, 3 types) with Warren-Cowley-like short-range order
from 10^5 seeded swap-MC steps; 3D block h,k,l in [-4, 4], 161^3 points;
X-ray 4-Gaussian form factors.  One step = one full Fourier run
(dc_four_run_dev: gather + exact phase reduction, phase GEMM, FP64
reduction with form factors, Friedel mirror) with the crystal resident in
HBM and outputs left in HBM.

Metric: atom.hkl pairs/s = sum_types N_type x 161^3 / wall time of a step,
for the split-precision MFMA path that holds max|dI|/max(I) <= 1e-5 on every
input, FP16x3 (FP16 hi/lo halves on both sides, three products per pair on
v_mfma_f32_16x16x32_f16, the same rate as the BF16 instruction); BF16x3,
FP16x2 and the plain FP32 MFMA path run the same workload beside it.  The CPU leg (the only part that runs the oracle)
times DISCUS's turbo algorithm on the host cores and checks a bounded sample
of the GPU outputs against the exact FP64 sum (max|dI|/max(I)).

The same JSON line carries the other workloads (each timed with the
library's HIP events on its stream):
  lots     fcc 100^3 cells (4M atoms, 10% substitution, correlated
           displacements), 64 periodic box lots of 8^3 cells, 8 hk-layers of
           401^2, FP64 <F> and cross-lot variance
  adjoint  orthorhombic 4/5/6 A, 32,768 atoms, 4 types, 65^3 block:
           chi^2 forward and adjoint (dL/d(x,y,z,occ,B) of every atom)
  batch    96 trial structures of log-uniform size in [2k, 60k] atoms, each
           against its own 256^2 layer: one population step (persistent
           GEMM + fused reduction + adjoint) replayed as a hipGraph, the
           generation's arrays resident in HBM (dc_batch_chi2_dev)
  post     the post-processing tail of four_run on a 161^3 grid
  nufft    the NUFFT technique on the headline workload

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--prec fp16x3]
  python bench.py --calibrate-copy 1073741824   (plain copy kernel of known
           size, for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE counters)
"""
import argparse
import ctypes as C
import dataclasses
import json
import os
import re
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "atom·hkl pairs/s, split-BF16 MFMA vs plain FP32 HIP, at ≤1e-5 rel err vs FP64"
PEAK = {"bf16x3": 2500.0, "fp16x3": 2500.0, "fp16x2": 2500.0, "fp32": 157.3}   # TFLOP/s dense
KSPLIT = {"bf16x3": 3, "fp16x3": 3, "fp16x2": 2, "fp32": 1}
HBM_PEAK = 8000.0   # GB/s
SHIFT = np.array([0.37, -0.21, 0.113])   # general origin shift of the coherent-sum check
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "phase_gemm_fcc40_161cube_traffic.md")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_info():
    info = {"nproc": os.cpu_count()}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in ("Model name", "Thread(s) per core", "Core(s) per socket", "Socket(s)"):
                info[k.strip()] = v.strip()
    except Exception:
        pass
    return info


def traffic_from_profile(prec):
    """HBM bytes per launch of the phase GEMM from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (already scaled by the copy-kernel
    calibration there); None when absent."""
    if not os.path.exists(TRAFFIC_FILE):
        return None
    pat = re.compile(r"^\|\s*" + re.escape(prec) + r"\s*\|.*\|\s*([0-9.eE+]+)\s*\|\s*$")
    for line in open(TRAFFIC_FILE):
        m = pat.match(line.strip())
        if m:
            return float(m.group(1))
    return None


def calibrate_copy(nbytes):
    import torch
    x = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda:0")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    log(f"copied {nbytes} bytes x 3")


LOTS_CHECK = {}   # sampled output and inputs of the lots run, checked in the CPU leg


def bench_lots(_lib, X, FR, torch, prec):
    c = X.fcc_correlated(n=100, a=4.0, subst=0.10, sigma=0.08, seed=3)
    eck, vi, inc = FR.grid_block((-10.0, -10.0, 0.0), (10.0, 10.0, 3.5), (401, 401, 8))
    ls = (8, 8, 8)
    lbeg = X.lot_origins(c.icc, 64, seed=11, lperiod=True, ls=ls)
    ntot = int(np.prod(inc))
    dev = torch.device("cuda:0")
    bufs = {k: torch.empty(n * ntot, dtype=torch.float64, device=dev)
            for k, n in (("csf", 2), ("dsi", 1), ("fmean", 2), ("dvar", 1))}
    ctx = _lib.Context(device=0, precision=prec)
    ctx.crystal_set(c)
    desc, keep = ctx.run_desc(eck, vi, inc, c.rten(), FR.formtab(c), ilots=2, ls=ls, lbeg=lbeg, lperiod=True,
                              four_friedel=True)
    stats = []
    for it in range(4):
        ctx.four_run_dev(desc, csf=bufs["csf"], dsi=bufs["dsi"], fmean=bufs["fmean"], dvar=bufs["dvar"])
        torch.cuda.synchronize()
        if it:
            stats.append(ctx.stats())
    ctx.close()
    # bounded sample for the check in the CPU leg: layer l = 0, every 10th point (h, k step 0.5)
    dsi = bufs["dsi"].cpu().numpy().reshape(inc, order="F")
    dvar = bufs["dvar"].cpu().numpy().reshape(inc, order="F")
    LOTS_CHECK.update(sample=dsi[::10, ::10, 0].copy(), sample_dvar=dvar[::10, ::10, 0].copy(), crystal=c, lbeg=lbeg,
                      ls=ls)
    ms = float(np.median([s["total_ms"] for s in stats]))
    gemm = float(np.median([s["gemm_ms"] for s in stats]))
    red = float(np.median([s["reduce_ms"] for s in stats]))
    pairs = stats[0]["pairs"]
    # lot reduction (all lots of a batch in one pass): per pixel it reads the
    # FP32 slice planes (8 B per slice) and reads + writes the FP64 accumulators
    # dsi (8 B), Welford mean (16 B) and M2 (8 B), and writes csf (16 B): 80 B
    nsl = stats[0]["slices"]
    red_bytes = 8.0 * ntot * nsl + 80.0 * ntot
    return {"workload": "fcc 100^3 cells (4M atoms, 2 types, 10% subst., corr. displ. 0.08 A), 64 periodic box "
                        "lots of 8^3 cells, 8 hk-layers of 401^2 (h,k in [-10,10], l = 0..3.5)",
            "precision": prec, "pairs": pairs, "ms": ms, "value": pairs / (ms * 1e-3), "gemm_ms": gemm,
            "reduce_ms": red, "reduce_hbm_GBps": red_bytes / (red * 1e-3) / 1e9 if red > 0 else None,
            "reduce_hbm_frac": red_bytes / (red * 1e-3) / 1e9 / HBM_PEAK if red > 0 else None}


def bench_adjoint(_lib, X, FR):
    from diffusecode_amd.refine import scattering_table
    cr = X.orthorhombic()
    occ, bdw = X.refinement_parameters(cr)
    cf = scattering_table(cr.scat)
    eck, vi, inc = FR.grid_block((-2.0, -2.0, -2.0), (2.0, 2.0, 2.0), (65, 65, 65))
    out = {"workload": "orthorhombic 4/5/6 A, 16^3 cells x 8 sites (32,768 atoms, 4 types), 65^3 block, "
                       "chi^2 gradient wrt 3N positions, N occupancies, N B"}
    ctx = _lib.Context(device=0)
    for prec in ("fp32", "bf16x3"):
        ctx.set_precision(prec)
        d, keep = ctx.model_desc(cr.pos, cr.iscat, occ, bdw, cf, eck, vi, inc, cr.rten())
        icalc, _ = ctx.model_intensity(d)
        target = icalc.ravel(order="F") * 1.01
        rows = []
        for it in range(4):
            ctx.model_chi2(d, target)
            if it:
                rows.append(ctx.stats())
        fwd = float(np.median([s["forward_ms"] for s in rows]))
        adj = float(np.median([s["adjoint_ms"] for s in rows]))
        out[prec] = {"forward_ms": fwd, "adjoint_ms": adj, "adjoint_over_forward": adj / fwd,
                     "pairs_per_forward": rows[0]["pairs"], "adjoint_walk": int(rows[0]["adjoint_walk"])}
    ctx.close()
    return out


def bench_nufft(_lib, torch, c, steps=3):
    """The NUFFT technique (four_run_nufft, fourier_sup.f90:327-690) on the
    headline workload: fcc 40^3 cells, 161^3 block h,k,l in [-4,4] (step 0.05
    -> scales 2, positions 2 pi * 2 * pos / 40), per type a 3D type-1 NUFFT at
    tol 1e-7 as DISCUS sets it (two types per transform), form factors, sum, |F|^2 (dc_four_run_nufft_dev,
    outputs in HBM).  Accuracy of I against the exact FP64 oracle on two
    subsampled l-layers, as for the headline (checked in the CPU leg)."""
    from diffusecode_amd import fourier as FR
    eck, vi, inc = FR.grid_block((-4, -4, -4), (4, 4, 4), (161, 161, 161))
    cf = FR.formtab(c)
    n = int(np.prod(inc))
    dev = torch.device("cuda:0")
    d_dsi = torch.empty(n, dtype=torch.float64, device=dev)
    d_csf = torch.empty(2 * n, dtype=torch.float64, device=dev)
    ctx = _lib.Context(device=0)
    ctx.crystal_set(c)
    desc, keep = ctx.run_desc(eck, vi, inc, c.rten(), cf)
    L = _lib.lib()
    torch.cuda.synchronize()

    def run():
        ctx.check(L.dc_four_run_nufft_dev(ctx.h, C.byref(desc), None, d_csf.data_ptr(), d_dsi.data_ptr()))
    run()
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        run()
    ctx.synchronize()
    ms = (time.perf_counter() - t0) / steps * 1e3
    dsi = d_dsi.cpu().numpy().reshape(inc, order="F")
    ctx.close()
    natoms = int(np.count_nonzero(c.iscat > 0))
    return {"ms": round(ms, 3), "value": natoms * n / (ms * 1e-3), "unit": "atom·hkl pairs/s (full grid)",
            "tol": 1e-7, "max_rel_err_sampled": None,
            "_samples": {0.0: dsi[::4, ::4, 80].copy(), 1.0: dsi[::4, ::4, 100].copy()},
            "note": "four_run_nufft on the headline workload: types in pairs, real spread of each type "
                    "(exponential-of-semicircle, w=8) into one component of a 324^3 complex grid, pruned 3D transform "
                    "(hipFFT 1D passes, gather-transposes of the needed modes), Hermitian separation and deconvolution, "
                    "form factor; sum over types and |F|^2; FP64"}


def bench_post(_lib, FR, torch, steps=5):
    """Post-processing tail of four_run on the headline grid (161^3): four_weight
    into dsi3d, Lanczos filter of dsi, four_conv of csf/dsi/dsi3d, Lanczos
    filter of dsi3d (fourier_sup.f90:279-290), all on device buffers."""
    inc = (161, 161, 161)
    eck, vi, _ = FR.grid_block((-4, -4, -4), (4, 4, 4), inc)
    n = int(np.prod(inc))
    dev = torch.device("cuda:0")
    g = torch.Generator(device="cpu").manual_seed(3)
    base_dsi = torch.rand(n, dtype=torch.float64, generator=g).to(dev)
    base_csf = torch.randn(2 * n, dtype=torch.float64, generator=g).to(dev)
    d_dsi, d_csf = base_dsi.clone(), base_csf.clone()
    d_d3 = torch.empty(n, dtype=torch.float64, device=dev)
    rten = np.eye(3) / 16.0
    ctx = _lib.Context(device=0)
    desc = ctx.post_desc(inc, eck, vi, rten, 1.0 / 64.0, fave=0.0, nlots=1, icc=(40, 40, 40), lots_off=True,
                         res_sigma=(1.5, 1.5, 1.5), res_tr=np.eye(3) * 0.5, lanczos=True, damp=0.5, width=4)
    torch.cuda.synchronize()
    ctx.four_post_dev(desc, d_csf, d_dsi, d_d3)          # plans, buffers
    ctx.synchronize()
    t = 0.0
    for _ in range(steps):
        d_dsi.copy_(base_dsi)
        d_csf.copy_(base_csf)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ctx.four_post_dev(desc, d_csf, d_dsi, d_d3)
        ctx.synchronize()
        t += time.perf_counter() - t0
    ctx.close()
    ms = t / steps * 1e3
    return {"ms": round(ms, 3), "grid": list(inc), "pixels_per_s": n / (ms * 1e-3),
            "steps": "four_weight + Lanczos(dsi) + four_conv(csf, dsi, dsi3d) + Lanczos(dsi3d), FP64, hipFFT Z2Z"}


def bench_batch(_lib, X, FR, prec, torch):
    """Population step (chi^2 of 96 members against their own layers): the
    generation's arrays (positions, occupancies, B, targets) resident in HBM,
    steps issued back to back as a DIFFEV loop on the device would; the first
    call plans, autotunes and captures the graph."""
    from diffusecode_amd.population import Population
    crs = X.population()
    eck, vi, inc = FR.grid_block((-4.0, -4.0, 0.0), (4.0, 4.0, 0.0), (256, 256, 1))
    pop = Population.from_crystals(crs, eck, vi, inc, precision=prec)
    desc, keep = pop.desc()
    pos, occ, bdw = pop.params()
    dev = torch.device("cuda:0")
    d_pos = torch.from_numpy(np.ascontiguousarray(pos.reshape(-1, 3))).to(dev)
    d_occ = torch.from_numpy(np.ascontiguousarray(occ)).to(dev)
    d_bdw = torch.from_numpy(np.ascontiguousarray(bdw)).to(dev)
    d_t = torch.ones((len(crs), pop.npixels), dtype=torch.float64, device=dev)
    n = pop.natoms
    d_grad = torch.empty((n, 5), dtype=torch.float64, device=dev)
    out = {"workload": f"{len(crs)} members (fcc/bcc/rocksalt, log-uniform 2k-60k atoms, {n} atoms), "
                       "own 256^2 hk0 layer each, per-atom B", "precision": prec,
           "inputs": "device-resident generation arrays, steps back to back"}
    for want_grad in (False, True):
        g = d_grad if want_grad else None
        pop.ctx.batch_chi2_dev(desc, d_pos, d_occ, d_bdw, d_t, grad=g)      # plan, tune, capture
        rows = []
        for it in range(6):
            pop.ctx.batch_chi2_dev(desc, d_pos, d_occ, d_bdw, d_t, grad=g)
            rows.append(pop.ctx.stats())
        ms = float(np.median([s["forward_ms"] for s in rows]))
        tot = float(np.median([s["total_ms"] for s in rows]))
        key = "forward_adjoint" if want_grad else "forward"
        pairs = rows[0]["pairs"]
        out[key] = {"step_ms": ms, "pairs_per_s": pairs / (ms * 1e-3), "call_ms": tot,
                    "info": pop.ctx.batch_info()}
        if not want_grad:
            fl = 8.0 * KSPLIT[prec] * pairs
            out[key]["mfma_frac_of_peak"] = round(fl / (ms * 1e-3) / 1e12 / PEAK[prec], 4)
            out[key]["gemm_ms_tuned"] = rows[0]["gemm_ms"]
    pop.ctx.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--prec", default="fp16x3")
    ap.add_argument("--fp32-steps", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--others", default="fp32,fp16x2,bf16x3,fp16x3",
                    help="schemes timed beside the headline one (fp32 is always timed: it is the ratio's base)")
    ap.add_argument("--no-coherent", action="store_true", help="skip the coherent-sum check runs")
    ap.add_argument("--extras", default="lots,adjoint,batch,post,nufft,peaks",
                    help="comma list of the extra measurements to run")
    ap.add_argument("--cpu-layers", type=int, default=40)
    ap.add_argument("--calibrate-copy", type=int, default=0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); "
              "this path shards over hkl slabs, one slab per GPU, which is not measured here.", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return

    import torch
    from diffusecode_amd import _lib, crystal as X, fourier as FR

    t0 = time.time()
    c = X.fcc_ordered(n=40, a=4.0, ntypes=3, mc_steps=100_000, seed=2)
    eck, vi, inc = FR.grid_block((-4, -4, -4), (4, 4, 4), (161, 161, 161))
    cf = FR.formtab(c)
    log(f"crystal {c.natoms} atoms types {c.type_counts().tolist()} built in {time.time() - t0:.1f}s")
    ntot = int(np.prod(inc))
    natoms = int(np.count_nonzero(c.iscat > 0))
    pairs = natoms * ntot

    dev = torch.device("cuda:0")
    d_dsi = torch.empty(ntot, dtype=torch.float64, device=dev)
    d_csf = torch.empty(2 * ntot, dtype=torch.float64, device=dev)
    ctx = _lib.Context(device=0)
    ctx.crystal_set(c)
    desc, keep = ctx.run_desc(eck, vi, inc, c.rten(), cf)

    def run_steps(prec, steps, warmup):
        ctx.set_precision(prec)
        for _ in range(warmup):
            ctx.four_run_dev(desc, csf=d_csf, dsi=d_dsi)
        torch.cuda.synchronize()
        gemm_ms, launches = 0.0, 0
        t = time.perf_counter()
        for _ in range(steps):
            ctx.four_run_dev(desc, csf=d_csf, dsi=d_dsi)
            s = ctx.stats()
            gemm_ms += s["gemm_ms"]
            launches += s["gemm_launches"]
        torch.cuda.synchronize()
        dt = time.perf_counter() - t
        return dt / steps, gemm_ms / max(launches, 1), launches / steps

    log(f"timing {args.prec}: {args.steps} steps after {args.warmup} warmup")
    t_step, gemm_launch_ms, launches_per_step = run_steps(args.prec, args.steps, args.warmup)
    value = pairs / t_step
    log(f"{args.prec}: {t_step * 1e3:.2f} ms/step, {value / 1e12:.2f} Tpairs/s, gemm {gemm_launch_ms:.2f} ms/launch")
    dsi_split = d_dsi.view(inc[2], inc[1], inc[0]).cpu().numpy()

    fp32_steps = args.fp32_steps if args.fp32_steps is not None else max(2, args.steps // 3)
    t32, g32, _ = run_steps("fp32", fp32_steps, 1)
    v32 = pairs / t32
    log(f"fp32: {t32 * 1e3:.2f} ms/step, {v32 / 1e12:.2f} Tpairs/s, gemm {g32:.2f} ms/launch")
    # the other split schemes on the same workload: FP16x2 (A split, B single
    # FP16) and BF16x3 / FP16x3, whichever is not the headline
    others = {}
    for prec in ("fp16x2", "bf16x3", "fp16x3"):
        if prec == args.prec or prec not in args.others.split(","):
            continue
        tt, gg_ms, _ = run_steps(prec, fp32_steps, 1)
        others[prec] = {"t": tt, "gemm_ms": gg_ms, "dsi": d_dsi.view(inc[2], inc[1], inc[0]).cpu().numpy()}
        log(f"{prec}: {tt * 1e3:.2f} ms/step, {pairs / tt / 1e12:.2f} Tpairs/s, gemm {gg_ms:.2f} ms/launch")

    # coherent sums: the same crystal shifted by a general vector.  Its sites
    # are no longer at 0 and 1/2, so at the Bragg peaks every cell adds the
    # same inexact phase and the per-product error of a scheme adds up over
    # all atoms; one run per scheme, the same two sampled layers
    moved = dataclasses.replace(c, pos=c.pos + SHIFT)
    ctx.crystal_set(moved)
    coherent_samples = {}
    for prec in (() if args.no_coherent else ("fp16x3", "bf16x3", "fp16x2", "fp32")):
        ctx.set_precision(prec)
        ctx.four_run_dev(desc, csf=d_csf, dsi=d_dsi)
        torch.cuda.synchronize()
        dm = d_dsi.view(inc[2], inc[1], inc[0]).cpu().numpy()
        coherent_samples[prec] = {0.0: dm[80, ::4, ::4].T.copy(), 1.0: dm[100, ::4, ::4].T.copy()}

    # dominant kernel: the phase GEMM; algorithmic MFMA flops = 8 x k_split per
    # computed atom.hkl pair (the GEMM computes the Friedel-reduced region)
    fnum0 = (inc[0] + 1) // 2
    pairs_computed_per_launch = natoms * fnum0 * inc[1] * inc[2] / launches_per_step
    achieved = 8.0 * KSPLIT[args.prec] * pairs_computed_per_launch / (gemm_launch_ms * 1e-3) / 1e12
    roofline = {"bound": "mfma", "achieved": round(achieved, 2), "peak": PEAK[args.prec], "unit": "TFLOP/s",
                "frac": round(achieved / PEAK[args.prec], 4), "traffic": traffic_from_profile(args.prec),
                "kernel": f"phase_gemm_kernel<{args.prec}>", "avg_launch_ms": round(gemm_launch_ms, 3),
                "flops_per_launch": 8.0 * KSPLIT[args.prec] * pairs_computed_per_launch,
                "fp32_kernel": {"avg_launch_ms": round(g32, 3),
                                "achieved": round(8.0 * pairs_computed_per_launch / (g32 * 1e-3) / 1e12, 2),
                                "peak": PEAK["fp32"]}}
    for prec, o in others.items():
        roofline[f"{prec}_kernel"] = {"avg_launch_ms": round(o["gemm_ms"], 3),
                                      "achieved": round(8.0 * KSPLIT[prec] * pairs_computed_per_launch
                                                        / (o["gemm_ms"] * 1e-3) / 1e12, 2),
                                      "peak": PEAK[prec]}

    # bounded sample of the GPU output for the check in the CPU leg: two
    # l-layers (l = 0 and l = 1) subsampled every 4th point (h, k step 0.2)
    head_samples = {0.0: dsi_split[80, ::4, ::4].T.copy(), 1.0: dsi_split[100, ::4, ::4].T.copy()}
    other_samples = {p: {0.0: o["dsi"][80, ::4, ::4].T.copy(), 1.0: o["dsi"][100, ::4, ::4].T.copy()}
                     for p, o in others.items()}
    max_err = None
    other_err = {p: None for p in others}
    coherent = None
    ctx.close()

    extras = {}
    want_nufft = not args.no_extras and "nufft" in {e.strip() for e in args.extras.split(",")}
    nufft_samples = None
    if want_nufft:
        extras["nufft"] = bench_nufft(_lib, torch, c)
        nufft_samples = extras["nufft"].pop("_samples")
        log(f"nufft: {json.dumps(extras['nufft'])}")
    want = set() if args.no_extras else {e.strip() for e in args.extras.split(",") if e.strip()}
    if "lots" in want:
        extras["lots"] = bench_lots(_lib, X, FR, torch, args.prec)
        log(f"lots: {extras['lots']['ms']:.2f} ms, {extras['lots']['value'] / 1e12:.2f} Tpairs/s")
    if "adjoint" in want:
        extras["adjoint"] = bench_adjoint(_lib, X, FR)
        log(f"adjoint: {json.dumps(extras['adjoint'])}")
    if "batch" in want:
        extras["batch"] = bench_batch(_lib, X, FR, args.prec, torch)
        log(f"batch: {json.dumps(extras['batch'])}")
    if "post" in want:
        extras["post"] = bench_post(_lib, FR, torch)
        log(f"post: {json.dumps(extras['post'])}")
    if "peaks" in want:
        # the device's ceilings measured with plain MFMA chains and HBM streams
        # (dc_measure_peaks); the fractions above stay priced against the
        # published peaks, these say what a plain stream of the same
        # instruction reaches on this part
        pctx = _lib.Context(device=0)
        try:
            pk = pctx.measure_peaks(1 << 32)
        finally:
            pctx.close()
        mkey = {"bf16x3": "mfma_bf16_tflops", "fp16x3": "mfma_f16_tflops", "fp16x2": "mfma_f16_tflops",
                "fp32": "mfma_f32_tflops"}[args.prec]
        roofline["measured_ceilings"] = {k: round(v, 2) if isinstance(v, float) else v for k, v in pk.items()}
        roofline["frac_of_measured_ceiling"] = round(roofline["achieved"] / pk[mkey], 4)
        log(f"peaks: {json.dumps(roofline['measured_ceilings'])}")

    cpu = None
    if not args.no_cpu:
        # CPU leg: the oracle as the timed baseline and as the checker of the samples
        from oracle import oracle as O
        if not args.no_check:
            refs = {}
            for lval in (0.0, 1.0):
                e2, v2, i2 = FR.grid_block((-4, -4, lval), (4, 4, lval), (41, 41, 1))
                refs[lval] = O.four_run(c, e2, v2, i2, cf)["dsi"][:, :, 0]

            def sampled_err(samples):
                errs = [(np.abs(samples[l] - refs[l]).max(), refs[l].max()) for l in refs]
                return float(max(e for e, _ in errs) / max(m for _, m in errs))
            max_err = sampled_err(head_samples)
            log(f"{args.prec} sampled max|dI|/max(I) vs FP64 oracle: {max_err:.3e}")
            for p, smp in other_samples.items():
                other_err[p] = sampled_err(smp)
                log(f"{p} sampled max|dI|/max(I): {other_err[p]:.3e}")
            moved = dataclasses.replace(c, pos=c.pos + SHIFT)
            mrefs = {}
            for lval in (0.0, 1.0):
                e2, v2, i2 = FR.grid_block((-4, -4, lval), (4, 4, lval), (41, 41, 1))
                mrefs[lval] = O.four_run(moved, e2, v2, i2, cf)["dsi"][:, :, 0]
            coherent = {"workload": "the headline crystal shifted by t = (0.37, -0.21, 0.113) (all 256k atoms in "
                                    "phase at the Bragg peaks), layers l = 0 and 1 every 4th point vs FP64"}
            for prec, smp in coherent_samples.items():  # (empty with --no-coherent)
                errs = [(np.abs(smp[l] - mrefs[l]).max(), mrefs[l].max()) for l in mrefs]
                coherent[prec] = float(max(e for e, _ in errs) / max(m for _, m in errs))
            log(f"coherent sums, max|dI|/max(I): {json.dumps(coherent)}")
            if LOTS_CHECK and "lots" in extras:
                cl = LOTS_CHECK["crystal"]
                e3, v3, i3 = FR.grid_block((-10.0, -10.0, 0.0), (10.0, 10.0, 0.0), (41, 41, 1))
                lr = O.four_run(cl, e3, v3, i3, FR.formtab(cl), ilots=O.LOT_BOX, ls=LOTS_CHECK["ls"],
                                lbeg=LOTS_CHECK["lbeg"], lperiod=True, want_var=True)
                ref, refv = lr["dsi"][:, :, 0], lr["dvar"][:, :, 0]
                extras["lots"]["max_rel_err_sampled"] = float(np.abs(LOTS_CHECK["sample"] - ref).max() / ref.max())
                # the FP64 cross-lot variance <|F|^2> - |<F>|^2 of the same run
                extras["lots"]["dvar_max_rel_err_sampled"] = float(np.abs(LOTS_CHECK["sample_dvar"] - refv).max()
                                                                   / np.abs(refv).max())
                log(f"lots sampled max|dI|/max(I): {extras['lots']['max_rel_err_sampled']:.3e}, "
                    f"dvar {extras['lots']['dvar_max_rel_err_sampled']:.3e}")
            if nufft_samples is not None:
                extras["nufft"]["max_rel_err_sampled"] = sampled_err(nufft_samples)
                log(f"nufft sampled max|dI|/max(I): {extras['nufft']['max_rel_err_sampled']:.3e}")
        nthr = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        nl = args.cpu_layers
        st, lmn, off = O.fourier_lmn(eck, vi, inc)
        num_s = [inc[0], inc[1], nl]
        fnum_s = [(inc[0] + 1) // 2, inc[1], nl]
        # DISCUS's own four_strucf when its build is present (compiled from the
        # reference sources, oracle/build_ref.py), else the oracle's bit-exact
        # restatement of it
        ref = O.ref_available()
        t = time.perf_counter()
        for t1 in range(1, c.nscat + 1):
            sel = c.iscat == t1
            if ref:
                O.ref_strucf(c.pos[sel], vi[:, 0], vi[:, 1], vi[:, 2], off, lmn, num_s, fnum=fnum_s,
                             nthreads=nthr)
            else:
                O.strucf_turbo(c.pos[sel], vi[:, 0], vi[:, 1], vi[:, 2], off, lmn, num_s, fnum=fnum_s,
                               nthreads=nthr)
        dt = time.perf_counter() - t
        cpairs = natoms * inc[0] * inc[1] * nl
        what = ("DISCUS's four_strucf (four_strucf_OMP_mod.f90:19-188 with its OpenMP kernel, compiled from the "
                "reference sources with flang -O3 -ffast-math -fopenmp, oracle/build_ref.py)" if ref else
                "DISCUS turbo restatement (oracle/dc_oracle.c, OpenMP over abscissa rows)")
        cpu = {"value": cpairs / dt, "unit": "atom·hkl pairs/s", "cores": nthr, "kind": "reference" if ref else "port",
               "sample": f"{what}, Friedel-halved like DISCUS, on {nl} of 161 l-layers of the same workload, all "
                         f"{natoms} atoms: {cpairs:.3e} pairs in {dt:.1f}s",
               "host": cpu_info()}
        log(f"cpu baseline ({cpu['kind']}): {cpu['value'] / 1e9:.2f} Gpairs/s on {nthr} threads")

    line = {
        "metric": METRIC, "value": value, "unit": "atom·hkl pairs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": t_step * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": args.prec, "data": "synthetic",
        "scheme_note": "value is the split scheme named in dtype; fp16x3 (three FP16 MFMA products, the same "
                       "MFMA rate as split BF16) is the default because it holds 1e-5 on every input, coherent "
                       "sums and Bragg-subtracted runs included; bf16x3 holds 1e-5 on direct structure factors "
                       "and up to 4e-5 on Bragg-subtracted runs",
        "config": {"workload": "fcc a=4.0A 40^3 cells (256k atoms, 3 types, 1e5 swap-MC steps), "
                               "3D block h,k,l in [-4,4] at 161^3 points, lots off, Friedel on",
                   "atoms": int(c.natoms), "grid": inc, "pairs_per_step": pairs, "precision": args.prec},
        "fp32_value": v32, "fp32_ms_per_step": t32 * 1e3, "speedup_vs_fp32": value / v32,
        "max_rel_err_sampled": max_err, "coherent_sum_max_rel_err": coherent,
        "roofline": roofline, "cpu_baseline": cpu,
    }
    for p, o in others.items():
        line[f"{p}_value"] = pairs / o["t"]
        line[f"{p}_ms_per_step"] = o["t"] * 1e3
        line[f"{p}_speedup_vs_fp32"] = (pairs / o["t"]) / v32
        line[f"{p}_max_rel_err_sampled"] = other_err[p]
    line.update(extras)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
