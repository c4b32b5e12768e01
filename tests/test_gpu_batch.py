"""GPU parity of population scoring (dc_batch_chi2 through the C ABI): every
member's intensity, chi^2 and gradient against the FP64 oracle, hipGraph
replays against eager execution (bit for bit, with other work between the
replays), and the autotuner's choice.

Bars: max|dI|/max(I) <= 1e-6 (FP32), <= 1e-5 (BF16x3) per member; chi^2
relative 1e-4; gradients max|dg| <= 1e-4 (FP32), 1e-3 (BF16x3) of max|g|
per member and parameter kind.  The chi^2 and gradient bars are wider than
the single-model ones because the weights 1/T (T down to 1e-3 max T) turn
the forward's absolute intensity error (~1e-7 max I in FP32) into ~1e-4
relative residual errors on the weak pixels of these Bragg-dominated
members; the adjoint's own error is ~1e-6."""
import numpy as np
import pytest

from diffusecode_amd import _lib
from diffusecode_amd import crystal as X
from diffusecode_amd import fourier as FR
from diffusecode_amd.population import Population
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def make_population(nmembers=6, nmin=300, nmax=3000, n=(33, 33, 1), seed=5, precision="fp32", ctx=None):
    crs = X.population(nmembers=nmembers, nmin=nmin, nmax=nmax, seed=seed)
    eck, vi, inc = FR.grid_block((-3.0, -3.0, 0.0), (3.0, 3.0, 0.0), n)
    return Population.from_crystals(crs, eck, vi, inc, ctx=ctx, precision=precision)


def oracle_member(pop, m, pos=None, occ=None, bdw=None):
    mb = pop.members[m]
    return O.model_f(mb["pos"] if pos is None else pos, mb["types"], mb["occ"] if occ is None else occ,
                     mb["bdw"] if bdw is None else bdw, mb["cfact"], pop.eck, pop.vi, pop.inc, mb["rten"])


def targets(pop, seed=3):
    rng = np.random.default_rng(seed)
    T = np.zeros((pop.npixels, len(pop.members)), order="F")
    for m, mb in enumerate(pop.members):
        p2, o2, b2 = X.perturb_parameters(mb["pos"], mb["occ"], mb["bdw"], seed=seed + m)
        I = (np.abs(oracle_member(pop, m, p2, o2, b2)) ** 2).ravel(order="F")
        T[:, m] = I * (1.0 + 0.01 * rng.standard_normal(I.size))
    return T


def grad_rel(g, ref):
    """max over parameter kinds of max|g - ref| / max|ref|; a kind whose
    reference is identically zero (a member without B) must come out zero."""
    out = 0.0
    for k in range(5):
        scale = np.abs(ref[:, k]).max()
        d = np.abs(g[:, k] - ref[:, k]).max()
        out = max(out, d / scale if scale > 0 else (0.0 if d == 0 else np.inf))
    return out


@pytest.fixture(scope="module")
def ctx():
    c = _lib.Context()
    yield c
    c.close()


def check_population(ctx, prec, bar, build):
    """Score the population build(ctx) returns in precision prec and check
    every member's intensities (max|dI|/max(I) <= bar), chi^2 and gradients
    against the oracle."""
    ctx.set_precision(prec)
    try:
        pop = build(ctx)
        T = targets(pop)
        W = 1.0 / np.maximum(T, 1e-3 * T.max(axis=0))
        chi, I, g = pop.score(T, W, want_icalc=True, want_grad=True)
    finally:
        ctx.set_precision("fp32")
    for m, mb in enumerate(pop.members):
        F = oracle_member(pop, m)
        Iref = (np.abs(F) ** 2).ravel(order="F")
        e = np.abs(I[:, m] - Iref).max() / Iref.max()
        chi_ref = float((W[:, m] * (Iref - T[:, m]) ** 2).sum())
        gref = O.model_grad(mb["pos"], mb["types"], mb["occ"], mb["bdw"], mb["cfact"], pop.eck, pop.vi, pop.inc,
                            mb["rten"], F, 2.0 * W[:, m] * (Iref - T[:, m]))
        gm = g[pop.atom_beg[m]:pop.atom_beg[m + 1]]
        ge = grad_rel(gm, gref)
        print(f"  {prec} member {m} ({len(mb['types'])} atoms): dI {e:.2e} chi2 {abs(chi[m] - chi_ref) / chi_ref:.2e} grad {ge:.2e}")
        assert e <= bar
        assert abs(chi[m] - chi_ref) <= 1e-4 * chi_ref
        assert ge <= (1e-4 if prec == "fp32" else 1e-3)


@pytest.mark.parametrize("prec,bar", [("fp32", 1e-6), ("bf16x3", 1e-5), ("fp16x3", 1e-5)])
def test_batch_matches_oracle(ctx, prec, bar):
    check_population(ctx, prec, bar, lambda c: make_population(ctx=c))


@pytest.mark.parametrize("prec,bar", [("fp32", 1e-6), ("fp16x3", 1e-5)])
def test_batch_many_types(ctx, prec, bar):
    """Members with ten atom types: more than the population reduction holds
    per pixel in LDS (eight), so its sums and Q planes read the form factors
    from global memory."""
    def build(c):
        crs = [X.fcc_ordered(n=4, ntypes=10, mc_steps=0, seed=11 + k) for k in range(3)]
        eck, vi, inc = FR.grid_block((-3.0, -3.0, 0.0), (3.0, 3.0, 0.0), (33, 33, 1))
        pop = Population.from_crystals(crs, eck, vi, inc, ctx=c, precision=prec)
        assert all(len(np.unique(mb["types"])) == 10 for mb in pop.members)
        return pop
    check_population(ctx, prec, bar, build)


def test_graph_replays_equal_eager():
    """Forced tile (no autotune): graph replays, with other GPU work between
    them, give the same bits as eager launches for three generations."""
    a = _lib.Context()
    b = _lib.Context()
    try:
        a.set_gemm_tile(0)
        b.set_gemm_tile(0)
        b.set_graphs(False)
        pa = make_population(ctx=a)
        pb = make_population(ctx=b)
        T = targets(pa)
        pos0, occ0, bdw0 = pa.params()
        rng = np.random.default_rng(1)
        for gen in range(3):
            pos = pos0 + rng.normal(0.0, 1e-3, pos0.shape)
            occ = occ0 * rng.uniform(0.95, 1.0, occ0.shape)
            ra = pa.score(T, pos=pos, occ=occ, bdw=bdw0, want_icalc=True)
            # unrelated eager work on the same context between replays
            cr = X.rocksalt(n=2)
            a.crystal_set(cr)
            rb = pb.score(T, pos=pos, occ=occ, bdw=bdw0, want_icalc=True)
            assert a.batch_info()["graph"] and not b.batch_info()["graph"]
            np.testing.assert_array_equal(ra[0], rb[0])
            np.testing.assert_array_equal(ra[1], rb[1])
            np.testing.assert_array_equal(ra[2], rb[2])
    finally:
        a.close()
        b.close()


def test_split_column_walk_in_graph():
    """BF16x3 batch gradients on a short 3D grid (17 x 33 x 33: 5 column tiles)
    with the reverse pass walking column ranges of 3 (A resident in LDS, a
    ragged last range) inside the captured graph: the gradient matches the
    FP64 oracle, the one-tile-per-work-group pass to summation-order
    rounding, and graph replays repeat the eager bits."""
    a = _lib.Context(precision="bf16x3")
    b = _lib.Context(precision="bf16x3")
    c = _lib.Context(precision="bf16x3")
    try:
        for x in (a, b, c):
            x.set_gemm_tile(0)
        a.set_adjoint_walk(3)
        b.set_adjoint_walk(3)
        b.set_graphs(False)
        c.set_adjoint_walk(1)
        crs = X.population(nmembers=3, nmin=300, nmax=1500, seed=8)
        eck, vi, inc = FR.grid_block((-2.0, -2.0, -1.0), (2.0, 2.0, 1.0), (17, 33, 33))
        pops = [Population.from_crystals(crs, eck, vi, inc, ctx=x, precision="bf16x3") for x in (a, b, c)]
        T = targets(pops[0])
        W = 1.0 / np.maximum(T, 1e-3 * T.max(axis=0))
        ra = [pops[0].score(T, W) for _ in range(3)]
        assert a.stats()["adjoint_walk"] == 3 and a.batch_info()["graph"]
        rb = pops[1].score(T, W)
        rc = pops[2].score(T, W)
        assert c.stats()["adjoint_walk"] == 1
        for r in ra[1:]:
            np.testing.assert_array_equal(r[2], ra[0][2])
        np.testing.assert_array_equal(ra[0][2], rb[2])
        g, g1 = ra[0][2], rc[2]
        pop = pops[0]
        for m, mb in enumerate(pop.members):
            F = oracle_member(pop, m)
            Iref = (np.abs(F) ** 2).ravel(order="F")
            gref = O.model_grad(mb["pos"], mb["types"], mb["occ"], mb["bdw"], mb["cfact"], pop.eck, pop.vi, pop.inc,
                                mb["rten"], F, 2.0 * W[:, m] * (Iref - T[:, m]))
            sl = slice(pop.atom_beg[m], pop.atom_beg[m + 1])
            ge = grad_rel(g[sl], gref)
            g1e = grad_rel(g[sl], g1[sl])
            print(f"  member {m} ({len(mb['types'])} atoms): grad vs fp64 {ge:.2e}, vs walk 1 {g1e:.1e}")
            assert ge <= 1e-3
            assert g1e <= 1e-5
    finally:
        a.close()
        b.close()
        c.close()


def test_autotuner_choice(ctx):
    pop = make_population(nmembers=4, ctx=ctx)
    T = targets(pop)
    chi, _, _ = pop.score(T, want_grad=False)
    info = ctx.batch_info()
    print("  autotune:", info)
    assert 0 <= info["tile"] < len(_lib.Context.gemm_tiles())
    assert info["graph"] and info["tune_ms"] > 0.0
    chi2, _, _ = pop.score(T, want_grad=False)      # replay of the same shape
    np.testing.assert_array_equal(chi, chi2)


@pytest.mark.parametrize("prec", ["fp32", "fp16x3"])
def test_batch_device_arrays_match_host(prec):
    """dc_batch_chi2_dev (generation arrays in HBM) gives the host entry
    point's chi^2, intensities and gradients bit for bit, on two generations."""
    import torch
    dev = torch.device("cuda:0")
    hc, dc_ = _lib.Context(precision=prec), _lib.Context(precision=prec)
    try:
        for c in (hc, dc_):
            c.set_autotune(False)   # same tile in both contexts
        pop = make_population(ctx=hc)
        T = targets(pop)
        W = 1.0 / np.maximum(T, 1e-3 * T.max(axis=0))
        pos, occ, bdw = pop.params()
        desc, keep = pop.desc()
        M, P, n = len(pop.members), pop.npixels, pop.natoms
        d_t = torch.from_numpy(np.ascontiguousarray(T.T)).to(dev)
        d_w = torch.from_numpy(np.ascontiguousarray(W.T)).to(dev)
        d_i = torch.zeros((M, P), dtype=torch.float64, device=dev)
        d_g = torch.zeros((n, 5), dtype=torch.float64, device=dev)
        rng = np.random.default_rng(1)
        for gen in range(2):
            p = pos + (0.01 * gen) * rng.standard_normal(pos.shape)
            chi, I, g = pop.score(T, W, pos=p, want_icalc=True, want_grad=True)
            d_pos = torch.from_numpy(np.ascontiguousarray(p)).to(dev)
            d_occ = torch.from_numpy(np.ascontiguousarray(occ)).to(dev)
            d_bdw = torch.from_numpy(np.ascontiguousarray(bdw)).to(dev)
            chi_d = dc_.batch_chi2_dev(desc, d_pos, d_occ, d_bdw, d_t, d_w, icalc=d_i, grad=d_g)
            assert np.array_equal(chi, chi_d), (gen, chi, chi_d)
            assert np.array_equal(I, d_i.cpu().numpy().T), gen
            assert np.array_equal(g, d_g.cpu().numpy()), gen
    finally:
        hc.close()
        dc_.close()
