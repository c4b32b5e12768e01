"""Writes ref_strucf_vectors.npz and ref_lmn_vectors.json.

ref_strucf_vectors.npz: input/output vectors of DISCUS's own
four_strucf (discus/prog/four_strucf_OMP_mod.f90:19-68 with its OpenMP,
serial and 1-D kernels and tcsf_form_iso, fourier_form.f90:147-177), run from
oracle/_ref/libref_strucf.so, which oracle/build_ref.py compiles from the
reference sources with the reference's own flags.

Each case holds the kernel's inputs (atoms, uin/vin/win, off_shift, lmn, num,
fnum, the form-factor column and istl bins when lform is on) and DISCUS's
tcsf.  The cases cover a 2-D layer at an lmn-encoded corner, a 3-D block, a
Friedel-reduced block (fnum < num), a 1-D line (the 1-D kernel), a form
factor through istl, atoms far from the origin and at negative coordinates
(the NINT / INT arithmetic of the increments), and two thread counts.

ref_lmn_vectors.json: DISCUS's own fourier_lmn (discus/prog/fourier_lmn_mod.f90:4-279,
same build): the corner multiples lmn, the offset vectors off and ier_num of
layers, volumes, lines, points, oblique layers and a grid with a zero
increment.

Run from the repository root after `python oracle/build_ref.py`:
    python tests/golden/make_ref_vectors.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402


def grid(lo, hi, n):
    lo, hi = np.asarray(lo, float), np.asarray(hi, float)
    n = [int(v) for v in n]
    eck = np.zeros((3, 4))
    vi = np.zeros((3, 3))
    eck[:, 0] = lo
    for a in range(3):
        if n[a] > 1:
            vi[a, a] = (hi[a] - lo[a]) / (n[a] - 1)
    eck[:, 1] = lo + vi[:, 0] * (n[0] - 1)
    eck[:, 2] = lo + vi[:, 1] * (n[1] - 1)
    eck[:, 3] = lo + vi[:, 2] * (n[2] - 1)
    return eck, vi, n


def cases():
    rng = np.random.default_rng(20261203)
    out = []
    # rocksalt-like 4x4x4 cells, one sublattice, vacancies and displacements,
    # 2-D layer with an lmn-encoded corner
    cells = np.stack(np.meshgrid(*[np.arange(4)] * 3, indexing="ij"), -1).reshape(-1, 3).astype(float)
    fcc = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    pos = (cells[:, None, :] + fcc[None]).reshape(-1, 3)
    pos = pos[rng.random(len(pos)) > 0.05]
    pos = pos + rng.normal(0.0, 0.02, size=pos.shape)
    out.append(("layer_hk03", pos, grid((-2.5, -2.5, 0.3), (2.5, 2.5, 0.3), (41, 41, 1)), None, 1))
    # 3-D block, random positions in a 6 x 5 x 4 box
    pos3 = rng.random((300, 3)) * np.array([6.0, 5.0, 4.0])
    out.append(("block", pos3, grid((-1.3, -0.9, -0.7), (1.1, 1.5, 0.9), (13, 11, 9)), None, 4))
    # Friedel-reduced block: fnum(1) < num(1)
    out.append(("block_fnum", pos3, grid((-1.0, -1.0, -1.0), (1.0, 1.0, 1.0), (9, 9, 9)), (5, 9, 9), 3))
    # 1-D line along a*: the 1-D kernel
    out.append(("line", pos3, grid((-3.0, 0.25, 0.5), (3.0, 0.25, 0.5), (301, 1, 1)), None, 2))
    # atoms far from the origin and at negative coordinates
    far = rng.random((200, 3)) * 20.0 - np.array([1000.0, -700.5, 333.25])
    out.append(("far_atoms", far, grid((-2.0, -2.0, 1.0), (2.0, 2.0, 1.0), (33, 33, 1)), None, 1))
    # form factor through istl bins
    out.append(("form_factor", pos3, grid((-1.5, -1.5, 0.0), (1.5, 1.5, 0.0), (31, 31, 1)), "ff", 2))
    return out


def lmn_grids():
    """Grids for DISCUS's fourier_lmn: layers, volumes, lines along and off the
    axes, single points, grids through hkl = 0 and away from it, oblique
    increment vectors, and a zero increment (ier_num -132)."""
    rng = np.random.default_rng(5)
    out = []
    for lo, hi, n in [((-5, -5, 0), (5, 5, 0), (101, 101, 1)), ((-4, -4, -4), (4, 4, 4), (161, 161, 161)),
                      ((-10, -10, 0.5), (10, 10, 0.5), (401, 401, 1)), ((0, 0, 1.5), (3, 2, 1.5), (31, 21, 1)),
                      ((-2, -3, 0.25), (2, 3, 3.75), (41, 61, 15)), ((-3, 0.25, 0.5), (3, 0.25, 0.5), (301, 1, 1)),
                      ((0.3, -2, 1.1), (0.3, 2, 1.1), (1, 81, 1)), ((0.5, 0.5, 0.5), (0.5, 0.5, 0.5), (1, 1, 1)),
                      ((1, 2, 3), (1, 2, 3), (1, 1, 1)), ((-1.5, -1.5, -1.5), (1.5, 1.5, 1.5), (7, 7, 7))]:
        eck, vi, inc = grid(lo, hi, n)
        out.append((eck, vi, inc))
    for _ in range(6):
        # oblique 2-D layers: eck(:,1) + i u + j v with general u, v
        u = rng.normal(0, 0.05, 3)
        v = rng.normal(0, 0.05, 3)
        c = rng.normal(0, 2.0, 3)
        n = [int(rng.integers(5, 40)), int(rng.integers(5, 40)), 1]
        eck = np.zeros((3, 4))
        eck[:, 0] = c
        eck[:, 1] = c + (n[0] - 1) * u
        eck[:, 2] = c + (n[1] - 1) * v
        eck[:, 3] = c
        vi = np.stack([u, v, np.zeros(3)], axis=1)
        out.append((eck, vi, n))
    eck, vi, inc = grid((-1, -1, 0), (1, 1, 0), (11, 11, 1))
    vi = vi.copy()
    vi[:, 1] = 0.0          # zero increment along an axis with 11 points
    out.append((eck, vi, inc))
    return out


def main():
    if not O.ref_available():
        raise SystemExit("oracle/_ref/libref_strucf.so missing: run python oracle/build_ref.py")
    rng = np.random.default_rng(7)
    data = {}
    names = []
    for name, pos, (eck, vi, num), extra, nthr in cases():
        st, lmn, off = O.fourier_lmn(eck, vi, num)
        assert st == 0, name
        fnum = np.asarray(num, np.int32)
        cf = istl = None
        if isinstance(extra, tuple):
            fnum = np.asarray(extra, np.int32)
        elif extra == "ff":
            ncf = 64
            s = np.arange(ncf) / ncf
            cf = (2.0 * np.exp(-3.0 * s * s) + 0.5) * np.exp(1j * 0.3 * s)
            istl = rng.integers(0, ncf, size=tuple(num)).astype(np.int32)
        tcsf = O.ref_strucf(pos, vi[:, 0], vi[:, 1], vi[:, 2], off, lmn, num, fnum=fnum, cfact_col=cf, istl=istl,
                            nthreads=nthr, variant=0)
        p = name + "/"
        data[p + "pos"] = np.asarray(pos, np.float64)
        data[p + "uvw"] = np.stack([vi[:, 0], vi[:, 1], vi[:, 2]])
        data[p + "off"] = np.asarray(off, np.float64)
        data[p + "lmn"] = np.asarray(lmn, np.int32)
        data[p + "num"] = np.asarray(num, np.int32)
        data[p + "fnum"] = fnum
        data[p + "eck"] = np.asarray(eck, np.float64)
        data[p + "vi"] = np.asarray(vi, np.float64)
        if cf is not None:
            data[p + "cfact"] = cf
            data[p + "istl"] = istl
        data[p + "tcsf"] = tcsf
        names.append(name)
    data["cases"] = np.array(names)
    np.savez_compressed(os.path.join(HERE, "ref_strucf_vectors.npz"), **data)
    print("wrote", len(names), "cases")
    lmn_cases = []
    for eck, vi, inc in lmn_grids():
        st, lmn, off = O.ref_fourier_lmn(eck, vi, inc)
        lmn_cases.append({"eck": np.asarray(eck).tolist(), "vi": np.asarray(vi).tolist(), "inc": [int(v) for v in inc],
                          "ier_num": st, "lmn": [int(v) for v in lmn], "off": np.asarray(off).tolist()})
    with open(os.path.join(HERE, "ref_lmn_vectors.json"), "w") as f:
        json.dump({"source": "DISCUS fourier_lmn (discus/prog/fourier_lmn_mod.f90:4-279), compiled from the "
                             "reference sources by oracle/build_ref.py", "cases": lmn_cases}, f, indent=1)
    print("wrote", len(lmn_cases), "fourier_lmn cases")


if __name__ == "__main__":
    main()
