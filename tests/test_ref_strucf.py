"""The oracle's restatement of DISCUS's turbo kernel against DISCUS's own
four_strucf, compiled from the reference sources (oracle/build_ref.py):

* against the committed vectors of tests/golden/ref_strucf_vectors.npz
  (written by tests/golden/make_ref_vectors.py from that build), bit for bit:
  a 2-D layer at an lmn-encoded corner, a 3-D block, a Friedel-reduced block,
  a 1-D line (the 1-D kernel), atoms far from the origin, a form factor
  through istl bins;
* where the build is present (this needs the reference sources), against the
  library itself: the serial, OpenMP and dispatcher entries at several thread
  counts, and the fixtures regenerated from it.

The cex table of the build is filled as four_cexpt does
(fourier_sup.f90:2602-2632, restated in oracle/ref_strucf_bind.f90, because
fourier_sup.f90 needs FFTW3 and HDF5); everything else is the reference's
compiled code."""
import os

import numpy as np
import pytest

from oracle import oracle as O

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_strucf_vectors.npz")


def _cases():
    z = np.load(FIX)
    out = []
    for name in z["cases"]:
        p = str(name) + "/"
        d = {k[len(p):]: z[k] for k in z.files if k.startswith(p)}
        out.append((str(name), d))
    return out


def _turbo(d):
    u, v, w = d["uvw"]
    return O.strucf_turbo(d["pos"], u, v, w, d["off"], d["lmn"], d["num"], fnum=d["fnum"],
                          cfact_col=d.get("cfact"), istl=d.get("istl"))


def test_fixture_cases_present():
    names = [n for n, _ in _cases()]
    assert {"layer_hk03", "block", "block_fnum", "line", "far_atoms", "form_factor"} <= set(names)


@pytest.mark.parametrize("name,d", _cases(), ids=[n for n, _ in _cases()])
def test_oracle_turbo_matches_discus_vectors(name, d):
    """Bit for bit: DISCUS's integer phase increments, its table index, its
    summation order per pixel and the form-factor multiply."""
    F = _turbo(d)
    assert F.shape == d["tcsf"].shape
    assert np.array_equal(F, d["tcsf"]), (name, np.abs(F - d["tcsf"]).max())
    # Friedel-reduced case: DISCUS zeroes tcsf beyond fnum
    f = d["fnum"]
    if tuple(f) != tuple(d["num"]):
        assert np.all(d["tcsf"][f[0]:, :, :] == 0)


needs_ref = pytest.mark.skipif(not O.ref_available(), reason="DISCUS's kernel not built (oracle/build_ref.py)")


@needs_ref
def test_reference_entries_and_threads_agree():
    """four_strucf_serial, four_strucf_omp and the dispatcher give the same
    bits at 1, 3 and 8 threads (each pixel row is summed by one thread in
    atom order)."""
    name, d = _cases()[1]
    u, v, w = d["uvw"]
    outs = [O.ref_strucf(d["pos"], u, v, w, d["off"], d["lmn"], d["num"], nthreads=t, variant=var)
            for var in (0, 1, 2) for t in (1, 3, 8)]
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])


@needs_ref
def test_fixtures_regenerate_from_reference():
    for name, d in _cases():
        u, v, w = d["uvw"]
        F = O.ref_strucf(d["pos"], u, v, w, d["off"], d["lmn"], d["num"], fnum=d["fnum"], cfact_col=d.get("cfact"),
                         istl=d.get("istl"), nthreads=2)
        assert np.array_equal(F, d["tcsf"]), name


@needs_ref
def test_oracle_turbo_matches_reference_on_random_crystals():
    """Fresh random inputs (not in the fixtures): 3-D blocks with general
    increment vectors and off-diagonal metrics of the grid."""
    rng = np.random.default_rng(11)
    for trial in range(4):
        pos = rng.random((120, 3)) * rng.uniform(2, 40, size=3) - rng.uniform(-50, 50, size=3)
        u, v, w = rng.normal(0, 0.07, size=(3, 3))
        off = rng.normal(0, 0.3, size=(3, 3))
        lmn = rng.integers(-30, 30, size=6)
        num = [int(rng.integers(2, 12)), int(rng.integers(2, 9)), int(rng.integers(1, 6))]
        a = O.strucf_turbo(pos, u, v, w, off, lmn, num)
        b = O.ref_strucf(pos, u, v, w, off, lmn, num, nthreads=4)
        assert np.array_equal(a, b), trial


LMN_FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_lmn_vectors.json")


def _lmn_cases():
    import json
    return json.load(open(LMN_FIX))["cases"]


@pytest.mark.parametrize("case", _lmn_cases(), ids=[f"lmn{i}" for i in range(len(_lmn_cases()))])
def test_oracle_fourier_lmn_matches_discus_vectors(case):
    """The oracle's fourier_lmn (in __float128) gives DISCUS's ier_num and
    integer corner multiples exactly, and its offset vectors to a few units in
    the last place (tests/golden/ref_lmn_vectors.json, from DISCUS's own
    fourier_lmn, which works in REAL(16) through matrix_mod's inverse: the two
    quad-precision evaluation orders differ in the last bits of some oblique
    offsets, up to 1e-14 of the largest offset, far below what moves an
    integer phase increment)."""
    st, lmn, off = O.fourier_lmn(np.array(case["eck"]), np.array(case["vi"]), case["inc"])
    assert st == case["ier_num"]
    if st == 0:
        assert [int(v) for v in lmn] == case["lmn"]
        ref = np.array(case["off"])
        np.testing.assert_allclose(np.asarray(off), ref, rtol=1e-14, atol=1e-14 * np.abs(ref).max())


@needs_ref
def test_oracle_fourier_lmn_matches_reference_on_random_layers():
    rng = np.random.default_rng(23)
    for trial in range(40):
        u, v = rng.normal(0, 0.08, 3), rng.normal(0, 0.08, 3)
        c = rng.normal(0, 3.0, 3)
        n = [int(rng.integers(2, 60)), int(rng.integers(1, 60)), 1]
        eck = np.zeros((3, 4))
        eck[:, 0] = c
        eck[:, 1] = c + (n[0] - 1) * u
        eck[:, 2] = c + (n[1] - 1) * v
        eck[:, 3] = c
        vi = np.stack([u, v, np.zeros(3)], axis=1)
        a = O.fourier_lmn(eck, vi, n)
        b = O.ref_fourier_lmn(eck, vi, n)
        assert a[0] == b[0], trial
        if a[0] == 0:
            assert np.array_equal(np.asarray(a[1]), b[1]), (trial, a[1], b[1])
            np.testing.assert_allclose(np.asarray(a[2]), b[2], rtol=1e-14, atol=1e-14 * np.abs(b[2]).max())
