"""GPU parity of the refinement model (dc_model_* through the C ABI) against
the FP64 oracle: forward intensity with per-atom occupancy and B, the
hand-written adjoint against the oracle's direct FP64 gradient, and chi^2
gradients against FP64 central differences (h = 1e-6) of the oracle chi^2.

Bars: forward max|dI|/max(I) <= 1e-6 (FP32) and <= 1e-5 (BF16x3); gradients
max|dg| <= 2e-5 * max|g| per parameter kind (x, y, z, occ, B) with the FP32
adjoint and <= 1e-4 with the split BF16x3 adjoint (products carry ~2^-16
relative error before the FP32 accumulation over the grid)."""
import numpy as np
import pytest

from diffusecode_amd import _lib
from diffusecode_amd import crystal as X
from diffusecode_amd import fourier as FR
from diffusecode_amd.refine import RefineModel, scattering_table
from oracle import oracle as O

pytestmark = pytest.mark.gpu

GRAD_BAR = 2e-5
GRAD_BAR_PREC = {"fp32": 2e-5, "bf16x3": 1e-4}


def problem(ncell=(4, 4, 3), n=(33, 25, 17), lo=(-2.0, -1.5, -1.0), hi=(2.0, 1.5, 1.0), seed=4):
    cr = X.orthorhombic(ncell=ncell, seed=seed)
    occ, bdw = X.refinement_parameters(cr, seed=seed + 1)
    cf = scattering_table(cr.scat)
    eck, vi, inc = FR.grid_block(lo, hi, n)
    return cr, occ, bdw, cf, eck, vi, inc


@pytest.fixture(scope="module")
def ctx():
    c = _lib.Context()
    yield c
    c.close()


def rel_i(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("prec,bar", [("fp32", 1e-6), ("bf16x3", 1e-5)])
@pytest.mark.parametrize("n", [(33, 25, 17), (33, 33, 1), (17, 9, 5)])
def test_model_forward(ctx, prec, bar, n):
    cr, occ, bdw, cf, eck, vi, inc = problem(n=n)
    ctx.set_precision(prec)
    try:
        d, keep = ctx.model_desc(cr.pos, cr.iscat, occ, bdw, cf, eck, vi, inc, cr.rten())
        icalc, fcalc = ctx.model_intensity(d)
    finally:
        ctx.set_precision("fp32")
    F = O.model_f(cr.pos, cr.iscat, occ, bdw, cf, eck, vi, inc, cr.rten())
    e = rel_i(icalc, np.abs(F) ** 2)
    print(f"  {prec} {n}: max|dI|/max(I) = {e:.3e}")
    assert e <= bar


def test_model_forward_without_b(ctx):
    cr, occ, bdw, cf, eck, vi, inc = problem()
    d, keep = ctx.model_desc(cr.pos, cr.iscat, occ, None, cf, eck, vi, inc, cr.rten())
    icalc, fcalc = ctx.model_intensity(d)
    F = O.model_f(cr.pos, cr.iscat, occ, None, cf, eck, vi, inc, cr.rten())
    assert rel_i(icalc, np.abs(F) ** 2) <= 1e-6


def grad_err(g, ref):
    return [np.abs(g[:, k] - ref[:, k]).max() / np.abs(ref[:, k]).max() for k in range(5)]


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
@pytest.mark.parametrize("n", [(33, 25, 17), (33, 33, 1), (17, 9, 5), (40, 7, 3)])
def test_model_vjp_matches_fp64_gradient(ctx, n, prec):
    cr, occ, bdw, cf, eck, vi, inc = problem(n=n)
    rng = np.random.default_rng(9)
    d, keep = ctx.model_desc(cr.pos, cr.iscat, occ, bdw, cf, eck, vi, inc, cr.rten())
    ctx.set_precision(prec)
    try:
        icalc, fcalc = ctx.model_intensity(d)
        gi = rng.standard_normal(icalc.size)
        g = ctx.model_vjp(d, fcalc, gi)
    finally:
        ctx.set_precision("fp32")
    F = O.model_f(cr.pos, cr.iscat, occ, bdw, cf, eck, vi, inc, cr.rten())
    ref = O.model_grad(cr.pos, cr.iscat, occ, bdw, cf, eck, vi, inc, cr.rten(), F, gi)
    e = grad_err(g, ref)
    print("  vjp", prec, n, " ".join(f"{x:.2e}" for x in e))
    assert max(e) <= GRAD_BAR_PREC[prec]


_WALK_CASE = {}


def walk_case():
    """4,096 atoms (128 atom tiles of 32) on a 17 x 33 x 33 grid: 2 stages of
    16 abscissa rows and 1,089 columns = 5 column tiles of 256."""
    if not _WALK_CASE:
        cr, occ, bdw, cf, eck, vi, inc = problem(ncell=(8, 8, 8), n=(17, 33, 33))
        F = O.model_f(cr.pos, cr.iscat, occ, bdw, cf, eck, vi, inc, cr.rten())
        gi = np.random.default_rng(21).standard_normal(F.size)
        ref = O.model_grad(cr.pos, cr.iscat, occ, bdw, cf, eck, vi, inc, cr.rten(), F, gi)
        _WALK_CASE.update(prob=(cr, occ, bdw, cf, eck, vi, inc), gi=gi, ref=ref)
    return _WALK_CASE


@pytest.mark.parametrize("prec", ["fp32", "bf16x3", "fp16x2"])
@pytest.mark.parametrize("walk,expect", [(1, 1), (2, 2), (3, 3), (5, 5), (32, 5), (0, None)])
def test_model_vjp_column_walk(ctx, prec, walk, expect):
    """Reverse pass on a short grid (FP32 and split): one work-group per column tile
    (walk 1), column ranges with A resident in LDS (2 and 3 leave a ragged
    last range, 5 and 32 walk all five column tiles in one range), and the
    planned walk.  Each matches the FP64 gradient to the scheme's bar (FP16x2
    is the lower-precision mode whose A-only split leaves ~1e-4 on random
    structures), and the walks agree with one another to FP32 summation-order
    rounding."""
    case = walk_case()
    cr, occ, bdw, cf, eck, vi, inc = case["prob"]
    d, keep = ctx.model_desc(cr.pos, cr.iscat, occ, bdw, cf, eck, vi, inc, cr.rten())
    ctx.set_precision(prec)
    ctx.set_adjoint_walk(walk)
    try:
        icalc, fcalc = ctx.model_intensity(d)
        g = ctx.model_vjp(d, fcalc, case["gi"])
        used = ctx.stats()["adjoint_walk"]
        ctx.set_adjoint_walk(1)
        g1 = ctx.model_vjp(d, fcalc, case["gi"])
    finally:
        ctx.set_adjoint_walk(0)
        ctx.set_precision("fp32")
    e = grad_err(g, case["ref"])
    e1 = grad_err(g, g1)
    print(f"  {prec} walk {walk} (used {used}): vs fp64", " ".join(f"{x:.2e}" for x in e),
          "| vs walk 1", " ".join(f"{x:.1e}" for x in e1))
    if expect is not None:
        assert used == expect
    assert max(e) <= {"fp32": GRAD_BAR_PREC["fp32"], "bf16x3": GRAD_BAR_PREC["bf16x3"], "fp16x2": 1e-3}[prec]
    assert max(e1) <= 1e-5


def test_adjoint_walk_knob_range(ctx):
    with pytest.raises(RuntimeError):
        ctx.set_adjoint_walk(33)
    with pytest.raises(RuntimeError):
        ctx.set_adjoint_walk(-1)
    ctx.set_adjoint_walk(0)


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_model_chi2_gradient_vs_central_differences(ctx, prec):
    """200 random parameters: GPU chi^2 gradient against FP64 central differences (h = 1e-6)."""
    cr, occ, bdw, cf, eck, vi, inc = problem()
    rten = cr.rten()
    p2, o2, b2 = X.perturb_parameters(cr.pos, occ, bdw, seed=12)
    Ft = O.model_f(p2, cr.iscat, o2, b2, cf, eck, vi, inc, rten)
    rng = np.random.default_rng(13)
    It = (np.abs(Ft) ** 2).ravel(order="F")
    target = It * (1.0 + 0.01 * rng.standard_normal(It.size))
    wt = 1.0 / np.maximum(target, 1e-3 * target.max())
    d, keep = ctx.model_desc(cr.pos, cr.iscat, occ, bdw, cf, eck, vi, inc, rten)
    ctx.set_precision(prec)
    try:
        chi, icalc, g = ctx.model_chi2(d, target, wt)
    finally:
        ctx.set_precision("fp32")
    F0 = O.model_f(cr.pos, cr.iscat, occ, bdw, cf, eck, vi, inc, rten)
    I0 = (np.abs(F0) ** 2).ravel(order="F")
    chi_ref = float((wt * (I0 - target) ** 2).sum())
    assert abs(chi - chi_ref) <= (1e-5 if prec == "fp32" else 1e-4) * chi_ref
    n = len(cr.iscat)
    atoms = rng.integers(0, n, 200)
    kinds = rng.integers(0, 5, 200)
    fd = O.model_fd(cr.pos, cr.iscat, occ, bdw, cf, eck, vi, inc, rten, F0, target, wt, atoms, kinds, h=1e-6)
    ref = O.model_grad(cr.pos, cr.iscat, occ, bdw, cf, eck, vi, inc, rten, F0, 2.0 * wt * (I0 - target))
    worst = 0.0
    for k in range(5):
        sel = kinds == k
        if not sel.any():
            continue
        scale = np.abs(ref[:, k]).max()
        e = np.abs(g[atoms[sel], k] - fd[sel]).max() / scale
        worst = max(worst, e)
        print(f"  {prec} kind {k}: max|g - fd|/max|g| = {e:.2e} over {sel.sum()} parameters")
    assert worst <= GRAD_BAR_PREC[prec]


def test_model_vacancies_and_type_order(ctx):
    """Vacancies get zero gradient; caller order is kept when types interleave."""
    cr, occ, bdw, cf, eck, vi, inc = problem(ncell=(2, 2, 2), n=(17, 13, 9))
    types = cr.iscat.copy()
    types[::7] = 0
    d, keep = ctx.model_desc(cr.pos, types, occ, bdw, cf, eck, vi, inc, cr.rten())
    icalc, fcalc = ctx.model_intensity(d)
    gi = np.ones(icalc.size)
    g = ctx.model_vjp(d, fcalc, gi)
    assert np.all(g[::7] == 0.0)
    F = O.model_f(cr.pos, types, occ, bdw, cf, eck, vi, inc, cr.rten())
    ref = O.model_grad(cr.pos, types, occ, bdw, cf, eck, vi, inc, cr.rten(), F, gi)
    assert max(grad_err(g, ref)) <= GRAD_BAR


def test_model_rejects_non_orthogonal_grid_with_b(ctx):
    cr, occ, bdw, cf, eck, vi, inc = problem(ncell=(2, 2, 2))
    rten = X.reciprocal_metric(4.0, 5.0, 6.0, 90.0, 100.0, 90.0)
    d, keep = ctx.model_desc(cr.pos, cr.iscat, occ, bdw, cf, eck, vi, inc, rten)
    with pytest.raises(_lib.DCError) as e:
        ctx.model_intensity(d)
    assert e.value.ier_num == _lib.DC_ERR_ARG


def test_autograd_functions(ctx):
    torch = pytest.importorskip("torch")
    cr, occ, bdw, cf, eck, vi, inc = problem(ncell=(2, 2, 2), n=(17, 13, 9))
    m = RefineModel(cr.iscat, cf, eck, vi, inc, cr.rten(), ctx=ctx)
    pos = torch.tensor(cr.pos, dtype=torch.float64, requires_grad=True)
    o = torch.tensor(occ, dtype=torch.float64, requires_grad=True)
    b = torch.tensor(bdw, dtype=torch.float64, requires_grad=True)
    I = m.intensity(pos, o, b)
    rng = np.random.default_rng(2)
    w = torch.tensor(rng.random(tuple(inc)), dtype=torch.float64)
    (I * w).sum().backward()
    icalc, fcalc = m.intensity_np(cr.pos, occ, bdw)
    g = m.vjp_np(cr.pos, occ, bdw, fcalc, w.numpy().ravel(order="F"))
    np.testing.assert_array_equal(pos.grad.numpy(), g[:, :3])   # deterministic reductions
    np.testing.assert_array_equal(o.grad.numpy(), g[:, 3])
    np.testing.assert_array_equal(b.grad.numpy(), g[:, 4])
    # chi^2 through the one-call path equals chi^2 built from I with torch ops
    target = torch.tensor(icalc * 1.1, dtype=torch.float64)
    for t in (pos, o, b):
        t.grad = None
    chi = m.chi2(pos, o, b, target)
    chi.backward()
    g1 = [t.grad.clone() for t in (pos, o, b)]
    for t in (pos, o, b):
        t.grad = None
    chi2 = ((m.intensity(pos, o, b) - target) ** 2).sum()
    chi2.backward()
    assert abs(chi.item() - chi2.item()) <= 1e-9 * chi2.item()
    for a, t in zip(g1, (pos, o, b)):
        assert (a - t.grad).abs().max().item() <= 1e-6 * t.grad.abs().max().item()


@pytest.mark.parametrize("prec,bar", [("fp32", 1e-5), ("bf16x3", 1e-4)])
def test_model_jvp_vs_central_differences_and_vjp(ctx, prec, bar):
    """Jacobian-vector products along random directions: against FP64 central
    differences of the oracle intensity (h = 1e-6, per pixel, relative to the
    largest |dI|), and dual to the vjp: <G, J v> = <J^T G, v>."""
    cr, occ, bdw, cf, eck, vi, inc = problem(ncell=(3, 3, 2), n=(25, 17, 9))
    rten = cr.rten()
    rng = np.random.default_rng(21)
    n = len(cr.iscat)
    dirs = rng.standard_normal((3, n, 5))
    dirs[:, :, :3] *= 0.01
    d, keep = ctx.model_desc(cr.pos, cr.iscat, occ, bdw, cf, eck, vi, inc, rten)
    ctx.set_precision(prec)
    try:
        icalc, fcalc = ctx.model_intensity(d)
        di = ctx.model_jvp(d, fcalc, dirs)
        gi = rng.standard_normal(icalc.size)
        g = ctx.model_vjp(d, fcalc, gi)
    finally:
        ctx.set_precision("fp32")
    h = 1e-6
    for k in range(dirs.shape[0]):
        v = dirs[k]
        Fp = O.model_f(cr.pos + h * v[:, :3], cr.iscat, occ + h * v[:, 3], bdw + h * v[:, 4], cf, eck, vi, inc, rten)
        Fm = O.model_f(cr.pos - h * v[:, :3], cr.iscat, occ - h * v[:, 3], bdw - h * v[:, 4], cf, eck, vi, inc, rten)
        fd = ((np.abs(Fp) ** 2 - np.abs(Fm) ** 2) / (2 * h)).ravel(order="F")
        e = np.abs(di[:, k] - fd).max() / np.abs(fd).max()
        dual = abs(gi @ di[:, k] - (g * v).sum()) / (np.abs(gi) @ np.abs(di[:, k]))
        print(f"  {prec} direction {k}: max|dI - fd|/max|fd| = {e:.2e}, duality {dual:.2e}")
        assert e <= bar
        assert dual <= bar


@pytest.mark.parametrize("prec,bar", [("fp16x3", 1e-5), ("bf16x3", 1e-5)])
@pytest.mark.parametrize("scale", [1e-6, 1.0, 1e4, "mixed"])
def test_model_forward_weight_range(ctx, prec, bar, scale):
    """The FP16 schemes scale each slice's B amplitudes (the per-atom weights
    occ) by a power of two into [0.5, 1) and undo it in the epilogue: weights
    of 1e-6 or 1e4, beyond FP16's normal range, and weights spanning four
    decades inside one slice keep the split bar against the FP64 oracle."""
    cr, occ, bdw, cf, eck, vi, inc = problem(n=(17, 9, 5))
    if scale == "mixed":
        occ = occ * np.logspace(-4, 0, occ.size)[np.random.default_rng(3).permutation(occ.size)]
    else:
        occ = occ * scale
    ctx.set_precision(prec)
    try:
        d, keep = ctx.model_desc(cr.pos, cr.iscat, occ, bdw, cf, eck, vi, inc, cr.rten())
        icalc, fcalc = ctx.model_intensity(d)
    finally:
        ctx.set_precision("fp32")
    F = O.model_f(cr.pos, cr.iscat, occ, bdw, cf, eck, vi, inc, cr.rten())
    e = rel_i(icalc, np.abs(F) ** 2)
    print(f"  {prec} occ x {scale}: max|dI|/max(I) = {e:.3e}")
    assert e <= bar


def _run_fortran_example(name):
    import os
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = os.path.join(root, "diffusecode_amd", "_build", "fortran", name)
    assert os.path.exists(exe), "build the Fortran examples first (__graft_entry__.build)"
    print(f"case fortran {name}", flush=True)
    p = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(p.stdout, p.stderr, flush=True)
    assert p.returncode == 0 and "PASS" in p.stdout


def test_fortran_host_refine():
    """A Fortran host fills dc_model_desc (bind(C)) and calls
    dc_model_intensity, dc_model_jvp (REFINE's derivative columns, copied into
    a gl_data(:,:,:,k)-shaped array) and dc_model_chi2: the columns against
    central differences, the adjoint gradient against the columns."""
    _run_fortran_example("example_refine")


def test_fortran_host_diffev():
    """A Fortran host fills dc_batch_desc (bind(C)) for a generation of three
    members and scores it with dc_batch_chi2 twice (the second generation
    replays the captured graph); each member against dc_model_chi2 alone."""
    _run_fortran_example("example_diffev")
