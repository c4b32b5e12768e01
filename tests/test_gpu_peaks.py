"""dc_measure_peaks: the device's ceilings measured with plain instruction
streams (independent MFMA chains, HBM read and copy).  No counterpart in
DISCUS; the figures put the kernels' roofline fractions in context.

Bounds: a plain stream cannot beat the published peak of its instruction
(MI355X: 2.5 PFLOP/s dense BF16/FP16, 157.3 TFLOP/s FP32 MFMA, 8 TB/s HBM)
by more than clock and cache slack, and a working part reaches a good share
of it; the bars are deliberately wide (0.3x .. 1.05x for MFMA, 0.2x .. 1.25x
for the HBM streams over a 1 GiB buffer, four times the 256 MB last-level
cache)."""
import pytest

from diffusecode_amd import _lib

pytestmark = pytest.mark.gpu

PUBLISHED = {"mfma_bf16_tflops": 2500.0, "mfma_f16_tflops": 2500.0, "mfma_f32_tflops": 157.3,
             "hbm_read_gbs": 8000.0, "hbm_copy_gbs": 8000.0}


@pytest.fixture(scope="module")
def ctx():
    c = _lib.Context()
    yield c
    c.close()


def test_ceilings_are_plausible(ctx):
    p = ctx.measure_peaks(1 << 30)
    print("  " + ", ".join(f"{k} {v:.1f}" if isinstance(v, float) else f"{k} {v}" for k, v in p.items()))
    assert p["hbm_bytes"] == 1 << 30
    assert p["num_cu"] == ctx.device_info()[1]
    for k, pub in PUBLISHED.items():
        lo, hi = (0.3, 1.05) if k.startswith("mfma") else (0.2, 1.25)
        assert lo * pub <= p[k] <= hi * pub, (k, p[k])


def test_argument_checks(ctx):
    for bad in (1000, (1 << 24) - 16, (1 << 24) + 8, 1 << 37):
        with pytest.raises(_lib.DCError):
            ctx.measure_peaks(bad)
    # the context stays usable after a rejected call
    p = ctx.measure_peaks(1 << 24)
    assert p["hbm_bytes"] == 1 << 24 and p["mfma_bf16_tflops"] > 0
