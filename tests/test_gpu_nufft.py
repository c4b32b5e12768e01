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
