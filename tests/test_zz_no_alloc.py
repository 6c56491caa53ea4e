"""The allocation-free entry point rvgp_colsumsq_noalloc against
rvgp_colsumsq (same kernels and summation order: equal bit for bit), called
through the C ABI twice in a row on the library's fixed scratch array, the
second call with a smaller input, in FP64 and FP32."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _colsumsq_ws(Z):
    import torch
    from rvgp_amd import _lib
    M, k = Z.shape
    out = torch.empty(k, dtype=torch.float64, device=Z.device)
    wsb = _lib.lib().rvgp_colsumsq_ws_bytes(M, k)
    ws = torch.empty(wsb, dtype=torch.uint8, device=Z.device)
    _lib.call("rvgp_colsumsq", int(Z.dtype == torch.float64), M, k, _lib.ptr(Z), Z.stride(0),
              _lib.ptr(out), _lib.ptr(ws), wsb, _lib.stream())
    return out


def _colsumsq_noalloc(Z):
    import torch
    from rvgp_amd import _lib
    M, k = Z.shape
    out = torch.empty(k, dtype=torch.float64, device=Z.device)
    _lib.call("rvgp_colsumsq_noalloc", int(Z.dtype == torch.float64), M, k, _lib.ptr(Z),
              Z.stride(0), _lib.ptr(out), _lib.stream())
    return out


@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_noalloc_equals_workspace_entry_point(dtype):
    import torch
    dt = getattr(torch, dtype)
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    # the headline residual shape (300k x 576) first, then a smaller one: the
    # second call reuses the scratch the first one filled
    for M, k in ((300000, 576), (4099, 37)):
        Z = torch.randn((M, k), generator=g, device="cuda", dtype=torch.float64).to(dt)
        a = _colsumsq_noalloc(Z)
        b = _colsumsq_ws(Z)
        torch.cuda.synchronize()
        assert torch.equal(a, b), (M, k)
        ref = (Z.double() ** 2).sum(0)
        assert torch.allclose(a, ref, rtol=1e-12, atol=0)


def test_noalloc_rejects_sizes_beyond_its_scratch():
    import torch
    Z = torch.zeros((40000, 1100), dtype=torch.float32, device="cuda")  # 1024 x 1100 > 2^20
    with pytest.raises(ValueError, match="scratch"):
        _colsumsq_noalloc(Z)
