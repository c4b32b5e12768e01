"""GPU type-1 NUFFT (dc_nufft1, dc_four_strucf_nufft) against the exact sums
of oracle/nufft.py, through the C ABI.

FINUFFT's contract is a relative l2 error of about tol; the bars here are
relative l2 error <= 5 tol and max|error| / max|f| <= 20 tol (the measured
values are printed)."""
import numpy as np
import pytest

from diffusecode_amd import _lib
from diffusecode_amd import crystal as X
from oracle import nufft as NU
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = _lib.Context()
    yield c
    c.close()


def errs(got, ref):
    l2 = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
    mx = float(np.abs(got - ref).max() / np.abs(ref).max())
    return l2, mx


@pytest.mark.parametrize("dim,N,tol", [(1, (101,), 1e-9), (1, (64,), 1e-12), (2, (37, 30), 1e-9), (2, (20, 25), 1e-5),
                                       (3, (24, 21, 17), 1e-7), (3, (16, 16, 16), 1e-9)])
@pytest.mark.parametrize("iflag", [1, -1])
def test_nufft1_vs_direct(ctx, dim, N, tol, iflag):
    print(f"case nufft1 dim={dim} N={N} tol={tol} iflag={iflag}", flush=True)
    rng = np.random.default_rng(dim * 100 + len(N) + int(-np.log10(tol)))
    M = 3000
    pts = [rng.uniform(-3 * np.pi, 3 * np.pi, M) for _ in range(dim)]
    c = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    got = ctx.nufft1(pts, c, iflag, tol, N)
    ref = NU.nufft1_direct(pts, c, iflag, N)
    l2, mx = errs(got, ref)
    print(f"  rel l2 {l2:.2e}  rel max {mx:.2e}")
    assert l2 <= 5 * tol and mx <= 20 * tol


def test_nufft1_no_points(ctx):
    got = ctx.nufft1([np.zeros(0)], np.zeros(0, complex), 1, 1e-9, (9,))
    assert np.all(got == 0)


def test_strucf_nufft_is_structure_factor(ctx):
    """four_strucf_3d on a rocksalt crystal with DISCUS's scaling
    x = 2 pi scales pos / ncells, against the exact FP64 structure factor at
    h = k scales / ncells (tol 1e-7 as DISCUS sets it)."""
    cr = X.rocksalt(n=6, seed=3)
    ncells, scales = 6.0, 2.0
    sel = cr.iscat == 1
    pos = cr.pos[sel]
    pts = [2 * np.pi * scales * pos[:, d] / ncells for d in range(3)]
    inc, ls = (13, 11, 9), (1, 2, 1)
    got = ctx.four_strucf_nufft(pts, 1.0, inc, ls)
    ref = NU.strucf_nufft(3, pts, 1.0, inc, ls)
    l2, mx = errs(got, ref)
    print(f"  strucf 3D rel l2 {l2:.2e}  rel max {mx:.2e}")
    assert l2 <= 1e-6 and mx <= 1e-6
    # and the same grid through the exact oracle in hkl: h = (k * ls) * scales / ncells
    N = [(inc[d] - 1) * ls[d] + 1 for d in range(3)]
    step = [ls[d] * scales / ncells for d in range(3)]
    xm = np.array([-(N[d] // 2) * scales / ncells for d in range(3)])
    F = O.strucf_exact(pos, xm, [step[0], 0, 0], [0, step[1], 0], [0, 0, step[2]], inc)
    assert float(np.abs(got - F).max() / np.abs(F).max()) <= 1e-6


@pytest.mark.parametrize("dim", [1, 2])
def test_strucf_nufft_low_dim(ctx, dim):
    rng = np.random.default_rng(dim)
    M = 500
    pts = [rng.uniform(-np.pi, np.pi, M) for _ in range(dim)]
    inc = (41, 33, 1)[:dim] + (1,) * (3 - dim)
    ls = (3, 2, 1)[:dim] + (1,) * (3 - dim)
    got = ctx.four_strucf_nufft(pts, 0.7, inc, ls)
    ref = NU.strucf_nufft(dim, pts, 0.7, inc, ls)
    l2, mx = errs(got, ref)
    print(f"  strucf {dim}D rel l2 {l2:.2e}  rel max {mx:.2e}")
    assert l2 <= 1e-8 and mx <= 1e-8


# ---------------------------------------------------------------- run level
from diffusecode_amd import fourier as FR


def _run(ctx, cr, lo, hi, n, fave=0.0, type_occ=None, sel=None):
    eck, vi, inc = FR.grid_block(lo, hi, n)
    cf = FR.formtab(cr)
    ctx.crystal_set(cr)
    kw = {}
    if fave:
        kw = dict(fave=fave, sel_cell=np.ones(int(np.prod(cr.icc)), np.uint8) if sel is None else sel)
    desc, keep = ctx.run_desc(eck, vi, inc, cr.rten(), cf, **kw)
    return ctx.four_run_nufft(desc, type_occ), (eck, vi, inc, cf)


@pytest.mark.parametrize("step,half", [(1.0 / 3.0, 2.0), (1.0, 2.0)])
def test_run_nufft_matches_exact_run(ctx, step, half):
    """step 1/3: scales 2, lscales 1; step 1: scales 6 -> lscales 2, scales 3.
    dsi against the exact FP64 oracle run (lots off, Friedel off in the oracle
    sense: the full grid)."""
    print(f"case run_nufft step={step}", flush=True)
    cr = X.rocksalt(n=6, seed=7)
    n = int(round(2 * half / step)) + 1
    got, (eck, vi, inc, cf) = _run(ctx, cr, (-half,) * 3, (half,) * 3, (n, n, n))
    ref = O.four_run(cr, eck, vi, inc, cf, four_friedel=False)
    err = float(np.abs(got["dsi"] - ref["dsi"]).max() / ref["dsi"].max())
    print(f"  run_nufft max|dI|/max(I) {err:.2e}")
    assert err <= 1e-6


def test_run_nufft_type_occ_weights(ctx):
    """DISCUS weights each atom by cr_occ in the transform and again in cfact:
    type_occ scales type t's F by occ_t."""
    cr = X.rocksalt(n=4, seed=8)
    a, _ = _run(ctx, cr, (-1.0,) * 3, (1.0,) * 3, (7, 7, 7))
    b, _ = _run(ctx, cr, (-1.0,) * 3, (1.0,) * 3, (7, 7, 7), type_occ=[0.5, 0.5])
    assert np.allclose(b["csf"], 0.5 * a["csf"], rtol=1e-12, atol=1e-12 * np.abs(a["csf"]).max())


def test_run_nufft_fave_zeroes_bragg(ctx):
    cr = X.rocksalt(n=6, seed=9)
    got, (eck, vi, inc, cf) = _run(ctx, cr, (-2.0,) * 3, (2.0,) * 3, (13, 13, 13), fave=50.0)
    bragg = np.zeros(inc, bool)
    bragg[::3, ::3, ::3] = True
    assert np.all(got["csf"][bragg] == 0) and np.all(got["dsi"][bragg] == 0)
    assert np.count_nonzero(got["dsi"][~bragg]) == np.count_nonzero(~bragg)


def test_run_nufft_errors(ctx):
    cr = X.rocksalt(n=6, seed=10)
    # a step 0.3 with fave: 1/0.3 is not an integer -> -187 (scales 2 is integer)
    with pytest.raises(_lib.DCError) as e:
        _run(ctx, cr, (-1.5,) * 3, (1.5,) * 3, (11, 11, 11), fave=50.0)
    assert e.value.ier_num in (-186, -187)
    # step 0.25 (0.25 * 6 cells is not an integer, 1/0.25 is): an off-centre corner -> -188
    with pytest.raises(_lib.DCError) as e:
        _run(ctx, cr, (-1.75, -2.0, -2.0), (2.25, 2.0, 2.0), (17, 17, 17), fave=50.0)
    assert e.value.ier_num == -188
    # the centred grid passes the checks
    got, _ = _run(ctx, cr, (-2.0,) * 3, (2.0,) * 3, (17, 17, 17), fave=50.0)
    assert got["dsi"].shape == (17, 17, 17)
    # oblique increments are refused
    eck, vi, inc = FR.grid_block((-1, -1, 0), (1, 1, 0), (9, 9, 1))
    vi = vi.copy()
    vi[1, 0] = 0.05
    cf = FR.formtab(cr)
    ctx.crystal_set(cr)
    desc, keep = ctx.run_desc(eck, vi, inc, cr.rten(), cf)
    with pytest.raises(_lib.DCError) as e:
        ctx.four_run_nufft(desc)
    assert e.value.ier_num == _lib.DC_ERR_ARG


@pytest.mark.parametrize("case", ["fcc 3 types", "ortho 4 types", "ortho, type 2 empty"])
def test_run_nufft_type_pairs(ctx, case):
    """The run level transforms two types at a time (real strengths, one in
    each component of the complex grid, separated by Hermitian symmetry):
    3 types (a pair and a single), 4 types (two pairs), and 4 types whose
    type 2 has no atoms.  Distinct type_occ per type catch a swapped
    component: the reference is the exact run with each type's form-factor
    column scaled by its type_occ."""
    import dataclasses
    print(f"case run_nufft pairs {case}", flush=True)
    if case.startswith("fcc"):
        cr = X.fcc_ordered(n=6, ntypes=3, mc_steps=3000, seed=4)
    else:
        cr = X.orthorhombic(ncell=(6, 6, 6), seed=6)
        if "empty" in case:
            cr = dataclasses.replace(cr, iscat=np.where(cr.iscat == 2, 0, cr.iscat).astype(np.int32))
    occ = np.array([0.3, 0.7, 0.9, 0.55][:cr.nscat])
    got, (eck, vi, inc, cf) = _run(ctx, cr, (-2.0,) * 3, (2.0,) * 3, (13, 13, 13), type_occ=occ)
    ref = O.four_run(cr, eck, vi, inc, cf * occ[None, :], four_friedel=False)
    err = float(np.abs(got["dsi"] - ref["dsi"]).max() / ref["dsi"].max())
    print(f"  {case}: max|dI|/max(I) {err:.2e}")
    assert err <= 1e-6
