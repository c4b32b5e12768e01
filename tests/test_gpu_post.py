"""GPU post-processing (dc_four_post / dc_four_fill / dc_four_accumulate_dev /
dc_qval) against the numpy restatement in oracle/post.py, through the C ABI.

Bars: FFT-based steps (four_conv, Lanczos) max|d|/max|ref| <= 1e-12 (FP64
transforms of different libraries differ in rounding only); four_weight,
the point-group average and the Bragg fill are bit-exact (same operations in
the same order); qval <= 1e-13 relative (exp/atan2/sqrt rounding)."""
import os

import numpy as np
import pytest

from diffusecode_amd import _lib
from diffusecode_amd import crystal as X
from diffusecode_amd import fourier as FR
from oracle import oracle as O
from oracle import post as P

pytestmark = pytest.mark.gpu

TR = [0.7, 0.1, 0.0, -0.05, 0.9, 0.0, 0.0, 0.02, 1.1]
SIG = [1.0, 1.3, 0.8]
RTEN = (np.eye(3) / 5.64 ** 2).ravel(order="F")
VR = 1.0 / 5.64 ** 3


@pytest.fixture(scope="module")
def ctx():
    c = _lib.Context()
    yield c
    c.close()


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def fields(num, seed):
    rng = np.random.default_rng(seed)
    csf = rng.standard_normal(num) + 1j * rng.standard_normal(num)
    dsi = rng.random(num) + 0.1
    return csf, dsi


def grid(num, step=0.1, lo=None):
    lo = [-(n // 2) * step for n in num] if lo is None else lo
    hi = [l + (n - 1) * step for l, n in zip(lo, num)]
    eck, vi, inc = FR.grid_block(lo, hi, num)
    return eck, vi


@pytest.mark.parametrize("num", [(33, 1, 1), (32, 1, 1), (17, 13, 1), (12, 19, 1), (13, 18, 1), (9, 11, 7), (10, 8, 6)])
def test_conv_matches_oracle(ctx, num):
    print(f"case four_conv num={num}", flush=True)
    csf, dsi = fields(num, sum(num))
    eck, vi = grid(num)
    d = ctx.post_desc(num, eck, vi, RTEN, VR, fave=1.0, lots_off=True, res_sigma=SIG, res_tr=TR)
    c, s, _ = ctx.four_post(d, csf, dsi, want_dsi3d=False)
    rc, rs, _ = P.four_conv(num, SIG, TR, csf, dsi, lots_off=True, fave=1.0)
    assert rel(s, rs) <= 1e-12
    assert rel(c, rc) <= 1e-12


def test_conv_lots_on_rescales_csf_only(ctx):
    num = (15, 11, 1)
    csf, dsi = fields(num, 4)
    eck, vi = grid(num)
    d = ctx.post_desc(num, eck, vi, RTEN, VR, fave=1.0, lots_off=False, res_sigma=SIG, res_tr=TR)
    c, s, _ = ctx.four_post(d, csf, dsi, want_dsi3d=False)
    rc, rs, _ = P.four_conv(num, SIG, TR, csf, dsi, lots_off=False, fave=1.0)
    assert rel(s, rs) <= 1e-12 and rel(c, rc) <= 1e-12


@pytest.mark.parametrize("num", [(41, 1, 1), (40, 1, 1), (21, 23, 1), (22, 20, 1), (13, 15, 11), (12, 14, 10),
                                 (15, 13, 5), (15, 4, 13), (3, 13, 15)])
def test_lanczos_matches_oracle(ctx, num):
    print(f"case lanczos num={num}", flush=True)
    _, dsi = fields(num, 7 + sum(num))
    eck, vi = grid(num)
    d = ctx.post_desc(num, eck, vi, RTEN, VR, fave=1.0, lanczos=True, rscale=1.0, damp=0.5, width=4)
    _, s, _ = ctx.four_post(d, None, dsi, want_dsi3d=False)
    ref = P.four_filter(dsi, num, 1.0, 0.5, 4)
    assert rel(s, ref) <= 1e-12


def test_weight_bragg(ctx):
    num = (21, 21, 9)
    eck, vi = grid(num, step=0.25, lo=[-2.5, -2.5, -1.0])
    _, dsi = fields(num, 11)
    d = ctx.post_desc(num, eck, vi, RTEN, VR, fave=0.0, nlots=1, icc=(6, 6, 6), diff_scale=1.0)
    _, s, d3 = ctx.four_post(d, None, dsi)
    ref, scalef = P.four_weight(dsi, num, eck.ravel(order="F"), vi.ravel(order="F"), RTEN, VR, 0.0, 1, (6, 6, 6), (1, 1, 1))
    assert scalef == pytest.approx(216 * 0.25 ** 3, rel=1e-12)     # ncells * |det(vi)|
    assert np.array_equal(s, dsi)
    assert rel(d3, ref) <= 1e-15
    assert not np.array_equal(d3, dsi)            # some Bragg pixels were rescaled


def test_full_tail_sequence(ctx):
    """four_weight -> Lanczos(dsi) -> four_conv(csf, dsi, dsi3d) -> Lanczos(dsi3d)."""
    num = (13, 15, 11)
    eck, vi = grid(num, step=0.25, lo=[-1.5, -1.75, -1.25])
    csf, dsi = fields(num, 12)
    d = ctx.post_desc(num, eck, vi, RTEN, VR, fave=0.0, nlots=2, ls=(3, 3, 3), lots_off=False,
                      res_sigma=SIG, res_tr=TR, lanczos=True, damp=0.5, width=4)
    c, s, d3 = ctx.four_post(d, csf, dsi)
    r3, _ = P.four_weight(dsi, num, eck.ravel(order="F"), vi.ravel(order="F"), RTEN, VR, 0.0, 2, (1, 1, 1), (3, 3, 3))
    rs = P.four_filter(dsi, num, 1.0, 0.5, 4)
    rc, rs, r3 = P.four_conv(num, SIG, TR, csf, rs, r3, lots_off=False, fave=0.0)
    r3 = P.four_filter(r3, num, 1.0, 0.5, 4)
    assert rel(s, rs) <= 1e-12 and rel(c, rc) <= 1e-12 and rel(d3, r3) <= 1e-12


def test_post_argument_errors(ctx):
    num = (1, 9, 9)
    eck, vi = grid(num)
    _, dsi = fields(num, 1)
    d = ctx.post_desc(num, eck, vi, RTEN, VR, fave=1.0, res_sigma=SIG, res_tr=TR)
    with pytest.raises(_lib.DCError) as e:
        ctx.four_post(d, None, dsi, want_dsi3d=False)
    assert e.value.ier_num == _lib.DC_ERR_ARG
    d = ctx.post_desc((9, 9, 1), *grid((9, 9, 1)), RTEN, VR, fave=1.0, lanczos=True, rscale=2.0)
    with pytest.raises(_lib.DCError):
        ctx.four_post(d, None, np.ones((9, 9, 1)), want_dsi3d=False)


def c4v_ops():
    r4 = np.array([[0, -1, 0], [1, 0, 0], [0, 0, 1]], float)
    mx = np.diag([-1.0, 1, 1])
    ops, g = [], np.eye(3)
    for _ in range(4):
        ops += [g.copy(), g @ mx]
        g = g @ r4
    return ops


@pytest.mark.parametrize("num", [(9, 9, 1), (11, 8, 5), (12, 12, 7)])
def test_symm_bit_exact(ctx, num):
    csf, dsi = fields(num, 21 + num[0])
    eck, vi = grid(num)
    ops = c4v_ops()
    d, keep = ctx.symm_desc(num, eck, vi, fave=0.0, four_symm=True, ops=ops)
    c, s = ctx.four_fill(d, csf, dsi)
    mats = np.stack(ops[1:], axis=2).ravel(order="F")
    e, v = eck.ravel(order="F"), vi.ravel(order="F")
    assert np.array_equal(c, P.four_symm(csf, num, e, v, mats, len(ops)))
    assert np.array_equal(s, P.four_symm(dsi, num, e, v, mats, len(ops), unit=1000.0))


@pytest.mark.parametrize("step,lo", [(0.25, (-2.0, -1.5, -1.0)), (1.0, (-3.0, -1.0, 0.0))])
def test_fill_bragg_bit_exact(ctx, step, lo):
    """step 0.25: isolated Bragg pixels (parallel kernel).  step 1.0: every
    pixel is a Bragg pixel, each fill reads already filled neighbours: the
    sequential kernel, in DISCUS's loop order."""
    num = (9, 7, 5)
    eck, vi = grid(num, step=step, lo=list(lo))
    csf, dsi = fields(num, 31)
    d, keep = ctx.symm_desc(num, eck, vi, fave=-1.0, four_symm=False)
    c, s = ctx.four_fill(d, csf, dsi)
    e, v = eck.ravel(order="F"), vi.ravel(order="F")
    rc = P.four_fill_bragg(csf, num, e, v)
    rs = P.four_fill_bragg(dsi, num, e, v)
    assert not np.array_equal(rs, dsi)
    assert np.array_equal(c, rc) and np.array_equal(s, rs)


class HipBuf:
    """Device memory from the HIP runtime the library uses (no torch: a second
    HIP runtime initialised after the library's may see no device)."""
    _rt = None

    def __init__(self, host):
        import ctypes as C
        if HipBuf._rt is None:
            HipBuf._rt = C.CDLL("libamdhip64.so")
        self.C = C
        self.n = host.nbytes
        self.p = C.c_void_p()
        assert HipBuf._rt.hipMalloc(C.byref(self.p), C.c_size_t(self.n)) == 0
        assert HipBuf._rt.hipMemcpy(self.p, host.ctypes.data_as(C.c_void_p), C.c_size_t(self.n), 1) == 0

    def get(self, like):
        out = np.empty_like(like)
        assert HipBuf._rt.hipMemcpy(out.ctypes.data_as(self.C.c_void_p), self.p, self.C.c_size_t(self.n), 2) == 0
        return out

    def free(self):
        HipBuf._rt.hipFree(self.p)

    @property
    def addr(self):
        return self.p.value


def test_accumulate_finished(ctx):
    num = (9, 9, 3)
    eck, vi = grid(num)
    ops = c4v_ops()
    d, keep = ctx.symm_desc(num, eck, vi, fave=0.0, four_symm=True, ops=ops)
    n = int(np.prod(num))
    acc_c, acc_d = HipBuf(np.zeros(2 * n)), HipBuf(np.zeros(n))
    ref_c = np.zeros(num, complex)
    ref_d = np.zeros(num)
    bufs = []
    for seed in (1, 2):
        csf, dsi = fields(num, seed)
        tc = HipBuf(np.ascontiguousarray(csf.ravel(order="F").view(np.float64)))
        td = HipBuf(np.ascontiguousarray(dsi.ravel(order="F")))
        bufs += [tc, td]
        ctx.four_accumulate_dev("accum", d, tc.addr, td.addr, acc_c.addr, acc_d.addr)
        ref_c = ref_c + csf
        ref_d = ref_d + dsi
    ctx.four_accumulate_dev("finished", d, tc.addr, td.addr, acc_c.addr, acc_d.addr)
    ctx.synchronize()
    mats = np.stack(ops[1:], axis=2).ravel(order="F")
    e, v = eck.ravel(order="F"), vi.ravel(order="F")
    got_c = tc.get(np.zeros(2 * n)).view(np.complex128).reshape(num, order="F")
    got_d = td.get(np.zeros(n)).reshape(num, order="F")
    assert np.array_equal(got_c, P.four_symm(ref_c, num, e, v, mats, 8))
    assert np.array_equal(got_d, P.four_symm(ref_d, num, e, v, mats, 8, unit=1000.0))
    with pytest.raises(_lib.DCError) as err:
        ctx.four_accumulate_dev("finished", d, tc.addr, td.addr, None, None)
    assert err.value.ier_num == _lib.DC_ERR_NOSUM
    for b in bufs + [acc_c, acc_d]:
        b.free()


@pytest.mark.parametrize("value", ["inten", "ampli", "phase", "real", "imag", "sq", "f2aver", "faver2", "faver",
                                   "iq", "f2averb", "faver2b", "faverb"])
@pytest.mark.parametrize("laver", [False, True])
def test_qval_matches_oracle(ctx, value, laver):
    cr = X.rocksalt(n=3, seed=2)
    num = (21, 17, 3)
    eck, vi = grid(num, step=0.2)
    csf, dsi = fields(num, 41)
    acsf, _ = fields(num, 42)
    cf = FR.formtab(cr, ldbw=True)
    cp = FR.formtab(cr, ldbw=False)
    rten = cr.rten()
    st, istl = O.stltab(eck[:, 0], vi[:, 0], vi[:, 1], vi[:, 2], num, rten.ravel(order="F"))
    amount = [float((cr.iscat == t + 1).sum()) for t in range(cr.nscat)]
    nreal = float((cr.iscat > 0).sum())
    d, keep = ctx.qval_desc(num, eck, vi, rten, value, laver=laver, cfact=cf, cfact_pure=cp, amount=amount,
                            n_real_atoms=nreal, u2aver=0.01, faver_sign=1.0)
    got = ctx.qval(d, csf, acsf, dsi)
    ref = P.qval(_lib.QVAL[value], laver, csf, acsf, dsi, istl, cf, cp, amount, nreal, 0.01, 0.001)
    assert rel(got, ref) <= 1e-13


def test_end_to_end_run_then_post(ctx):
    """dc_four_run on a rocksalt crystal, then dc_four_post, against the
    oracle run followed by the oracle tail."""
    cr = X.rocksalt(n=4, seed=5)
    num = (15, 13, 11)
    eck, vi, inc = FR.grid_block((-1.75, -1.5, -1.25), (1.75, 1.5, 1.25), num)
    cf = FR.formtab(cr)
    ctx.crystal_set(cr)
    desc, keep = ctx.run_desc(eck, vi, inc, cr.rten(), cf)
    gpu = ctx.four_run(desc)
    ref = O.four_run(cr, eck, vi, inc, cf)
    rten = cr.rten().ravel(order="F")
    vr = float(np.sqrt(np.linalg.det(cr.rten())))
    d = ctx.post_desc(num, eck, vi, rten, vr, fave=0.0, nlots=1, icc=cr.icc, lots_off=True,
                      res_sigma=SIG, res_tr=TR, lanczos=True, width=4)
    c, s, d3 = ctx.four_post(d, gpu["csf"], gpu["dsi"])
    r3, _ = P.four_weight(ref["dsi"], num, eck.ravel(order="F"), vi.ravel(order="F"), rten, vr, 0.0, 1, cr.icc, (1, 1, 1))
    rs = P.four_filter(ref["dsi"], num, 1.0, 0.5, 4)
    rc, rs, r3 = P.four_conv(num, SIG, TR, ref["csf"], rs, r3, lots_off=True, fave=0.0)
    r3 = P.four_filter(r3, num, 1.0, 0.5, 4)
    # the run carries the FP32 path's error (<= 1e-6 of max I; F to ~1e-6 of max |F|)
    assert rel(s, rs) <= 2e-6 and rel(d3, r3) <= 2e-6 and rel(c, rc) <= 1e-5


def test_fortran_host_post_and_nufft():
    """A Fortran host fills the bind(C) descriptors of diffuse_turbo_mod and
    calls dc_nufft1, dc_four_post, dc_four_fill and dc_qval (closed forms)."""
    import os
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = os.path.join(root, "diffusecode_amd", "_build", "fortran", "example_post")
    assert os.path.exists(exe), "build the Fortran examples first (__graft_entry__.build)"
    print("case fortran post/nufft", flush=True)
    p = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(p.stdout, p.stderr, flush=True)
    assert p.returncode == 0 and "PASS" in p.stdout


def test_fourier_menu_filter_resolution_and_nufft(ctx):
    """The fourier menu mirror: 'set technique,nufft', 'set filter,lanczos',
    a resolution, and 'run' against the oracle run followed by the oracle tail."""
    from diffusecode_amd.fourier import Fourier
    cr = X.rocksalt(n=4, seed=12)
    f = Fourier(cr, ctx=ctx)
    f.ll(-2.0, -2.0, 0.0)
    f.lr(2.0, -2.0, 0.0)
    f.ul(-2.0, 2.0, 0.0)
    f.nabs(25)
    f.nord(25)
    f.set_symm("ignore")
    f.set_options(filter="lanczos", damp=0.5, width=4)
    f.resolution(0.02, 0.03)
    out = f.run()
    eck, vi, inc = f.eck, f.vi, f.inc
    cf = FR.formtab(cr)
    ref = O.four_run(cr, eck, vi, inc, cf, four_friedel=False)
    rs = P.four_filter(ref["dsi"], inc, 1.0, 0.5, 4)
    rc, rs, _ = P.four_conv(inc, f.res_sigma, f.res_tr.ravel(order="F"), ref["csf"], rs, lots_off=True)
    assert rel(out["dsi"], rs) <= 2e-6
    # the NUFFT technique on a block grid of the same crystal (step 1/4 = scales 1 for 4 cells)
    g = Fourier(cr, ctx=ctx)
    g.ll(-2.0, -2.0, 0.0)
    g.lr(2.0, -2.0, 0.0)
    g.ul(-2.0, 2.0, 0.0)
    g.nabs(17)
    g.nord(17)
    g.set_options(technique="nufft")
    out = g.run()
    ref = O.four_run(cr, g.eck, g.vi, g.inc, cf, four_friedel=False)
    assert rel(out["dsi"], ref["dsi"]) <= 1e-6


def test_fourier_output_hdf5(ctx, tmp_path):
    """'output' with 'outfile type hdf5' (graphic.f90:606-636): dc_qval of the
    run's pixels on the GPU times the output scale, written as a Yell 1.0 file;
    the file's data is the FP64 oracle's intensity (x scale) and its step
    sizes follow 'abscissa k / ordinate h'."""
    from diffusecode_amd import build as B
    if B.hdf5_root() is None or not os.path.exists(B.YELL_LIB):
        pytest.skip("no HDF5 installation / libdiffuse_yell.so not built")
    from diffusecode_amd.fourier import Fourier
    from h5read import H5
    cr = X.rocksalt(n=3, seed=5)
    f = Fourier(cr, ctx=ctx)
    f.ll(-1.5, -1.5, 0.0)
    f.lr(1.5, -1.5, 0.0)
    f.ul(-1.5, 1.5, 0.0)
    f.nabs(13)
    f.nord(11)
    f.abscissa("k")
    f.ordinate("h")
    f.run()
    path = tmp_path / "run.h5"
    q = f.output_hdf5(path, value="inten", scale=2.0)
    ref = O.four_run(cr, f.eck, f.vi, f.inc, FR.formtab(cr))
    assert rel(q, 2.0 * ref["dsi"]) <= 1e-6
    h5 = H5()
    fid = h5.open(path)
    try:
        data, r = h5.doubles(fid, "data")
        steps, _ = h5.doubles(fid, "step_sizes")
        cell, _ = h5.doubles(fid, "unit_cell")
    finally:
        h5.L.H5Fclose(fid)
    assert r["dims"] == (13, 11, 1)
    np.testing.assert_array_equal(data, q)
    np.testing.assert_array_equal(steps, [f.vi[1, 0], f.vi[0, 1], f.vi[2, 2]])
    np.testing.assert_array_equal(cell, list(cr.lattice))
