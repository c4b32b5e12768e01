"""The oracle against the committed known-answer fixtures (tests/golden):
closed-form lattice sums, rocksalt selection rules at integer hkl (through the
whole run, Friedel halving included) and DISCUS's integer phase pipeline with
exact NINT ties.  The fixtures come from exact arithmetic in
tests/golden/make_golden.py, not from the oracle."""
import math

import numpy as np
import pytest

from golden_cases import load, structure_case
from oracle import oracle as O


@pytest.mark.parametrize("name", ["lattice_dirichlet.json", "rocksalt_selection.json"])
def test_oracle_run_matches_closed_form(name):
    c, eck, vi, inc, cfact, F = structure_case(name)
    out = O.four_run(c, eck, vi, inc, cfact)
    scale = np.abs(F).max()
    assert np.abs(out["csf"] - F).max() <= 1e-11 * scale
    assert np.abs(out["dsi"] - np.abs(F) ** 2).max() <= 1e-11 * scale ** 2


def test_oracle_exact_strucf_matches_dirichlet():
    c, eck, vi, inc, cfact, F = structure_case("lattice_dirichlet.json")
    got = O.strucf_exact(c.pos, eck[:, 0], vi[:, 0], vi[:, 1], vi[:, 2], list(inc))
    assert np.abs(got - F).max() <= 1e-11 * np.abs(F).max()


def test_oracle_turbo_matches_integer_pipeline():
    d = load("turbo_phase_kat.json")
    num = d["num"]
    want = np.zeros(int(np.prod(num)), np.complex128)
    for idx in d["table_index"]:
        k = np.array(idx)
        want += np.cos(2 * math.pi * k / 65536) + 1j * np.sin(2 * math.pi * k / 65536)
    got = O.strucf_turbo(np.array(d["atoms"]), d["uin"], d["vin"], d["win"], np.array(d["off_shift"]), d["lmn"], num)
    np.testing.assert_allclose(got.ravel(order="F"), want, rtol=0, atol=1e-13)


def test_fixture_ties_are_present():
    """The KAT must exercise NINT's ties (away from zero) in both signs."""
    inc = load("turbo_phase_kat.json")["increments"]
    assert inc[2][0] == 1 and inc[3][0] == -2 and inc[3][1] == 65534
