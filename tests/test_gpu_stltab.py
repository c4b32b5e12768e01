"""four_stltab on the GPU (dc_four_stltab) against the oracle, integer for
integer: the sin(theta)/lambda bins every run of the library uses
(fourier_sup.f90:2637-2687, h rounded to FP32 by REAL() of default kind)."""
import numpy as np
import pytest

from diffusecode_amd import crystal as X
from diffusecode_amd import fourier as FR
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from diffusecode_amd import _lib
    c = _lib.Context()
    yield c
    c.close()


def _grids():
    cub = lambda a: X.reciprocal_metric(a, a, a)  # noqa: E731
    yield "rocksalt hk0 101^2", FR.grid_block((-5, -5, 0), (5, 5, 0), (101, 101, 1)), cub(5.64)
    yield "fcc 161^3", FR.grid_block((-4, -4, -4), (4, 4, 4), (161, 161, 161)), cub(4.0)
    yield "lots 8 x 401^2", FR.grid_block((-10, -10, 0), (10, 10, 3.5), (401, 401, 8)), cub(4.0)
    yield "triclinic layer", FR.grid_block((-3.1, -2.2, 0.7), (2.9, 3.3, 0.7), (121, 97, 1)), \
        X.reciprocal_metric(5.1, 6.3, 7.7, 84.0, 97.0, 101.0)
    xm, u = np.array([0.001, 0.0, 0.0]), np.array([0.002, 0.0, 0.0])
    eck = np.zeros((3, 4))
    eck[:, 0] = xm
    eck[:, 1] = xm + 3999 * u
    vi = np.zeros((3, 3))
    vi[:, 0] = u
    yield "near-tie line", (eck, vi, [4000, 1, 1]), np.eye(3)


@pytest.mark.parametrize("case", range(5))
def test_stltab_bins_bit_exact(ctx, case):
    name, (eck, vi, inc), rten = list(_grids())[case]
    print(f"case stltab {name}", flush=True)
    st, ref = O.stltab(eck[:, 0], vi[:, 0], vi[:, 1], vi[:, 2], inc, rten)
    assert st == 0
    got = ctx.four_stltab(eck[:, 0], vi[:, 0], vi[:, 1], vi[:, 2], inc, rten)
    nd = int(np.count_nonzero(got != ref))
    print(f"  bins {ref.size}, differing {nd}", flush=True)
    assert nd == 0


def test_stltab_overflow(ctx):
    """A bin beyond CFPKT gives four_stltab's -3 (ER_FOUR)."""
    from diffusecode_amd import _lib
    eck, vi, inc = FR.grid_block((0, 0, 0), (200, 0, 0), (3, 1, 1))
    with pytest.raises(_lib.DCError) as e:
        ctx.four_stltab(eck[:, 0], vi[:, 0], vi[:, 1], vi[:, 2], inc, X.reciprocal_metric(4.0, 4.0, 4.0))
    assert e.value.ier_num == -3 and e.value.ier_typ == 9
