"""dc_strucf_dev_noalloc (the layer call that allocates nothing) against
dc_strucf through the C ABI, on two calls in a row, the second with a
smaller input (the first, 20,000 atoms on 201 x 201 x 31, runs its FP32
slices in batches that fit the 128 MiB scratch array; dc_strucf runs them in
one), and a problem too large for the array is refused before any kernel."""
import numpy as np
import pytest

from diffusecode_amd import crystal as X
from diffusecode_amd import fourier as FR
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _case(n, shape, seed):
    rng = np.random.default_rng(seed)
    pos = rng.random((n, 3)) * np.array([6.0, 5.0, 4.0])
    lo = [-(k // 2) * 0.13 for k in shape]
    eck, vi, num = FR.grid_block(lo, [l + (k - 1) * 0.13 for l, k in zip(lo, shape)], shape)
    st, lmn, off = O.fourier_lmn(eck, vi, num)
    assert st == 0
    return pos, eck, vi, num, lmn, off


@pytest.mark.parametrize("prec", ["fp32", "fp16x3"])
def test_noalloc_matches_strucf(prec):
    import torch
    torch.cuda.init()
    from diffusecode_amd import _lib
    dev = torch.device("cuda:0")
    ctx = _lib.Context(device=0)
    try:
        ctx.set_precision(prec)
        c = X.rocksalt(n=6)
        cf = FR.formtab(c)[:, 0]
        rt = c.rten()
        for k, (n, shape) in enumerate(((20000, (201, 201, 31)), (3000, (33, 29, 1)))):
            print(f"case noalloc {prec} call {k}: {n} atoms, grid {shape}", flush=True)
            pos, eck, vi, num, lmn, off = _case(n, shape, seed=k)
            ref = ctx.strucf(pos, vi[:, 0], vi[:, 1], vi[:, 2], off, lmn, num, cfact_col=cf, xm=eck[:, 0], rten=rt)
            x = torch.tensor(np.asfortranarray(pos).ravel(order="F"), dtype=torch.float64, device=dev)
            cfd = torch.tensor(np.ascontiguousarray(cf).view(np.float64), dtype=torch.float64, device=dev)
            out = torch.full((2 * int(np.prod(num)),), np.nan, dtype=torch.float64, device=dev)
            torch.cuda.synchronize()
            ctx.strucf_dev_noalloc(x.data_ptr(), n, n, vi[:, 0], vi[:, 1], vi[:, 2], off, lmn, num,
                                   d_cfact=cfd.data_ptr(), ncfact=cf.size, xm=eck[:, 0], rten=rt,
                                   d_tcsf=out.data_ptr())
            got = out.cpu().numpy().view(np.complex128).reshape(tuple(num), order="F")
            err = float(np.abs(got - ref).max() / np.abs(ref).max())
            print(f"  max|d|/max vs dc_strucf = {err:.2e}", flush=True)
            assert np.isfinite(got).all()
            assert err <= 1e-12
    finally:
        ctx.close()


def test_noalloc_refuses_oversized():
    import torch
    torch.cuda.init()
    from diffusecode_amd import _lib
    dev = torch.device("cuda:0")
    ctx = _lib.Context(device=0)
    try:
        num = [401, 401, 161]          # one plane of 25.9M pixels (207 MB) exceeds the 128 MiB array
        eck, vi, _ = FR.grid_block((-4, -4, -4), (4, 4, 4), num)
        st, lmn, off = O.fourier_lmn(eck, vi, num)
        x = torch.zeros(30, dtype=torch.float64, device=dev)
        out = torch.zeros(2 * int(np.prod(num)), dtype=torch.float64, device=dev)
        with pytest.raises(_lib.DCError) as e:
            ctx.strucf_dev_noalloc(x.data_ptr(), 10, 10, vi[:, 0], vi[:, 1], vi[:, 2], off, lmn, num,
                                   d_tcsf=out.data_ptr())
        assert e.value.ier_num == _lib.DC_ERR_ARG
    finally:
        ctx.close()
