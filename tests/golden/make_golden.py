"""Writes the known-answer fixtures of this directory.

The reference holds no golden vectors for the structure-factor path (its tests
do not cover it, and its code is Fortran), so every expected value here comes
from what the mathematics fixes, evaluated in this script with exact rational
or integer arithmetic or with a closed formula in FP64 -- never from the oracle
or from the library under test:

* lattice_dirichlet.json -- a perfect N1 x N2 x N3 lattice (one site per cell):
  F(h) = prod_a exp(i pi h_a (N_a - 1)) sin(pi N_a h_a) / sin(pi h_a), the
  lattice sum over the cells (DISCUS's F with a unit form factor), on a
  general 3D grid that does not meet a Bragg point.
* rocksalt_selection.json -- rocksalt (sites of two types on the fcc
  sublattices) at integer hkl: F = N_cells 4 (fA + fB) for h, k, l all even,
  N_cells 4 (fA - fB) all odd, 0 mixed; constant form factors fA, fB.  The
  grid is symmetric, so DISCUS's Friedel halving applies to it.
* turbo_phase_kat.json -- DISCUS's integer phase pipeline
  (four_strucf_OMP_mod.f90:135-170): per atom and vector v, the increment
  NINT(2^22 (t - INT(t))) with t = v.x, the phase word
  iarg = sum lmn(m) inc(m) + (j-1) inc_u + (i-1) inc_v + (h-1) inc_w, and the
  table index IAND(ISHFT(iarg, -6), 2^16 - 1) of cex (fourier_sup.f90:2602-2632).
  Coordinates and vectors are dyadic rationals, so DISCUS's FP64 products are
  exact and equal the rational ones here; the atoms include exact
  half-integer ties of NINT (rounded away from zero) and negative increments.

Run from the repository root:  python tests/golden/make_golden.py
"""
import json
import math
import os
from fractions import Fraction

HERE = os.path.dirname(os.path.abspath(__file__))


def dirichlet():
    n = (5, 4, 3)
    corner = (-1.03, -0.47, 0.2)
    step = (0.137, 0.101, 0.09)
    inc = (15, 11, 5)
    eck = [[corner[a] + (step[a] * (inc[a] - 1) if c == a + 1 else 0.0) for c in range(4)] for a in range(3)]
    vi = [[step[a] if c == a else 0.0 for c in range(3)] for a in range(3)]
    re, im = [], []
    # Fortran order: abscissa fastest
    for k in range(inc[2]):
        for j in range(inc[1]):
            for i in range(inc[0]):
                idx = (i, j, k)
                f = complex(1.0, 0.0)
                for a in range(3):
                    h = corner[a] + idx[a] * step[a]
                    f *= complex(math.cos(math.pi * h * (n[a] - 1)), math.sin(math.pi * h * (n[a] - 1))) * \
                        (math.sin(math.pi * n[a] * h) / math.sin(math.pi * h))
                re.append(f.real)
                im.append(f.imag)
    return {"what": "perfect lattice, one site per cell at the cell origin, unit form factor",
            "cells": list(n), "sites": [[0.0, 0.0, 0.0]], "types": [1], "lattice": [4.0, 4.0, 4.0, 90.0, 90.0, 90.0],
            "eck": eck, "vi": vi, "inc": list(inc), "form_factor": [1.0],
            "F_re": re, "F_im": im}


def rocksalt():
    n = 3
    fa, fb = 1.7, 0.6
    inc = (5, 5, 5)
    eck = [[-2.0, 2.0 if a == 0 else -2.0, 2.0 if a == 1 else -2.0, 2.0 if a == 2 else -2.0] for a in range(3)]
    vi = [[1.0 if c == a else 0.0 for c in range(3)] for a in range(3)]
    sites = [[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0], [.5, 0, 0], [0, .5, 0], [0, 0, .5], [.5, .5, .5]]
    types = [1, 1, 1, 1, 2, 2, 2, 2]
    ncell = n ** 3
    re = []
    for k in range(inc[2]):
        for j in range(inc[1]):
            for i in range(inc[0]):
                odd = sum((x - 2) % 2 for x in (i, j, k))
                re.append(4 * (fa + fb) * ncell if odd == 0 else 4 * (fa - fb) * ncell if odd == 3 else 0.0)
    return {"what": "rocksalt at integer hkl, constant form factors per type",
            "cells": [n, n, n], "sites": sites, "types": types, "lattice": [5.64, 5.64, 5.64, 90.0, 90.0, 90.0],
            "eck": eck, "vi": vi, "inc": list(inc), "form_factor": [fa, fb],
            "F_re": re, "F_im": [0.0] * len(re)}


def nint(q: Fraction) -> int:
    """Fortran NINT: nearest integer, ties away from zero."""
    a = abs(q)
    r = math.floor(a + Fraction(1, 2))
    return r if q >= 0 else -r


def turbo_kat():
    # dyadic coordinates: t = v.x is exact in FP64; the last two atoms put
    # 2^22 frac(t) exactly on a half-integer (ties of NINT)
    atoms = [(0.3125, -1.71875, 2.4375), (-0.59375, 0.84375, -3.15625),
             (2.0 ** -20, 0.0, 0.0),                     # 2^22 t_u = +0.5 -> 1
             (-3 * 2.0 ** -20, 5 * 2.0 ** -19, 0.5)]     # t_u -> -1.5 -> -2, t_v -> 65533.5 -> 65534
    uin, vin, win = (0.125, 0.0, 0.0), (0.0, -0.0625, 0.03125), (0.0, 0.0, 0.25)
    off = [(2.0 ** -10, 0.0, 0.0), (0.0, 2.0 ** -10, 0.0), (0.0, 0.0, 2.0 ** -9)]
    vecs = [uin, vin, win] + off
    lmn = (-40, 17, 3, 3, -5, 0)
    num = (7, 5, 3)
    idx_all = []
    incs = []
    for x in atoms:
        xf = [Fraction(c) for c in x]
        inc = []
        for v in vecs:
            t = sum(Fraction(a) * b for a, b in zip(v, xf))
            fr = t - math.trunc(t)
            inc.append(nint(fr * 2 ** 22))
        incs.append(inc)
        idx = []
        # tcsf(j, i, h) with j abscissa (fastest), i ordinate, h top
        for h in range(num[2]):
            for i in range(num[1]):
                for j in range(num[0]):
                    iarg = sum(m * c for m, c in zip(lmn, inc)) + j * inc[0] + i * inc[1] + h * inc[2]
                    idx.append((iarg >> 6) & 0xFFFF)
        idx_all.append(idx)
    return {"what": "DISCUS integer phase pipeline: NINT(2^22 frac(v.x)) increments and 2^16 table indices",
            "atoms": [list(a) for a in atoms], "uin": list(uin), "vin": list(vin), "win": list(win),
            "off_shift": [[off[c][r] for c in range(3)] for r in range(3)],   # off_shift(r, c) = vector c
            "lmn": list(lmn), "num": list(num), "increments": incs, "table_index": idx_all}


def main():
    for name, make in (("lattice_dirichlet.json", dirichlet), ("rocksalt_selection.json", rocksalt),
                       ("turbo_phase_kat.json", turbo_kat)):
        with open(os.path.join(HERE, name), "w") as f:
            json.dump(make(), f, separators=(",", ":"))
        print("wrote", name)


if __name__ == "__main__":
    main()
