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
