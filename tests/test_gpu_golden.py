"""The HIP path against the committed known-answer fixtures (tests/golden):
a closed-form lattice sum on a general 3D grid and the rocksalt selection
rules at integer hkl (Friedel halving applies), through dc_four_run in every
precision.  Bars: max|dI|/max(I) and max|dF|/max|F| <= 1e-6 (FP32), 1e-5
(BF16x3, FP16x3), 3e-4 (FP16x2, whose B operand is a single FP16)."""
import numpy as np
import pytest

from diffusecode_amd import _lib
from golden_cases import structure_case

pytestmark = pytest.mark.gpu
BAR = {"fp32": 1e-6, "bf16x3": 1e-5, "fp16x3": 1e-5, "fp16x2": 3e-4}


@pytest.fixture(scope="module")
def ctx():
    c = _lib.Context()
    yield c
    c.close()


@pytest.mark.parametrize("prec", ["fp32", "bf16x3", "fp16x3", "fp16x2"])
@pytest.mark.parametrize("name", ["lattice_dirichlet.json", "rocksalt_selection.json"])
def test_run_matches_closed_form(ctx, name, prec):
    print(f"case golden {name} prec={prec}", flush=True)
    c, eck, vi, inc, cfact, F = structure_case(name)
    ctx.set_precision(prec)
    try:
        ctx.crystal_set(c)
        desc, keep = ctx.run_desc(eck, vi, inc, c.rten(), cfact)
        out = ctx.four_run(desc)
    finally:
        ctx.set_precision("fp32")
    I = np.abs(F) ** 2
    ei = np.abs(out["dsi"] - I).max() / I.max()
    ef = np.abs(out["csf"] - F).max() / np.abs(F).max()
    print(f"  max|dI|/max(I) = {ei:.2e}, max|dF|/max|F| = {ef:.2e}", flush=True)
    assert ei <= BAR[prec] and ef <= BAR[prec]
