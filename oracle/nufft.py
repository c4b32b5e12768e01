"""CPU checker of the NUFFT technique: the exact type-1 sums that FINUFFT
approximates, in numpy FP64.

TEST INFRASTRUCTURE ONLY: tests/ and bench.py use this module as the checker.
The product package (diffusecode_amd) never imports it.

  nufft1_direct   fk(k) = sum_j c_j exp(iflag i k . x_j), k_d = -N_d/2 ..
                  (N_d-1)/2, k1 fastest: the definition FINUFFT's type 1
                  computes to tolerance tol (finufft{1,2,3}d1 as called in
                  discus/prog/fourier_finufft.f90:62, 124, 210; FINUFFT
                  itself is a dependency vendored under finufft/, whose
                  documented definition this restates).
                  Pinned by (tests/test_oracle_post.py): a single point gives
                  exp(iflag i k x); points on the lattice 2 pi m / L give
                  the Dirichlet kernel; iflag = -1 is the conjugate of +1
                  for real weights; a brute-force cmath loop on three points.
  strucf_nufft    four_strucf_1d/2d/3d (fourier_finufft.f90:19-245): modes
                  (inc-1)*lscales+1 per axis, every lscales-th kept, weight
                  occ.  Pinned by: with positions 2 pi s x / n (the scaling
                  of four_getatm_anis_iam, fourier_sup.f90:2144-2216) it equals
                  the exact structure factor at h = k s / n.
"""
from __future__ import annotations

import numpy as np


def modes(n):
    return np.arange(n) - n // 2


def nufft1_direct(pts, c, iflag, N, chunk=2048):
    """pts: list of dim coordinate arrays; c complex (M,); N: mode counts."""
    dim = len(pts)
    N = list(N)[:dim]
    ks = [modes(n).astype(float) for n in N]
    out = np.zeros(tuple(N), np.complex128, order="F")
    c = np.asarray(c, np.complex128)
    M = c.shape[0]
    for s in range(0, M, chunk):
        e = min(M, s + chunk)
        f = [np.exp(1j * iflag * np.outer(np.asarray(pts[d][s:e], float), ks[d])) for d in range(dim)]  # (m, N_d)
        if dim == 1:
            out += (c[s:e, None] * f[0]).sum(0)
        elif dim == 2:
            out += np.einsum("j,ja,jb->ab", c[s:e], f[0], f[1])
        else:
            out += np.einsum("j,ja,jb,jc->abc", c[s:e], f[0], f[1], f[2])
    return out


def strucf_nufft(dim, pos, occ, inc, lscales):
    """pos: list of dim scaled coordinate arrays; returns tcsf of shape inc[:dim]."""
    N = [(inc[d] - 1) * lscales[d] + 1 for d in range(dim)]
    c = np.full(len(pos[0]), complex(occ, 0.0))
    f = nufft1_direct(pos, c, 1, N)
    sl = tuple(slice(0, None, lscales[d]) for d in range(dim))
    return f[sl]
