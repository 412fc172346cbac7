"""lmx_dscaled_residual_noalloc_dev (the entry point that allocates
nothing: its accumulators are a fixed device array of the library) against
lmx_dscaled_residual_dev, through the C ABI: the same bits on two calls in
a row, the second with a smaller input, and a clear error above its size
limit."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _call(lib, name, A, X, B, eta):
    fn = getattr(lib, name)
    fn.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64] + [ctypes.c_void_p, ctypes.c_int64] * 3 + \
        [ctypes.c_void_p]
    fn.restype = ctypes.c_int
    n, k = X.shape
    return fn(None, n, k, A.data_ptr(), A.stride(0), X.data_ptr(), X.stride(0), B.data_ptr(), B.stride(0),
              eta.data_ptr())


def test_noalloc_matches_allocating_entry(lmx_gpu):
    import torch
    lib = lmx_gpu.raw()
    g = torch.Generator(device="cuda").manual_seed(7)
    for n, k in ((3000, 16), (700, 3)):   # the second call is smaller
        A = torch.rand(n, n, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
        X = torch.rand(n, k, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
        B = torch.rand(n, k, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
        e_ref = torch.full((k,), -1.0, dtype=torch.float64, device="cuda")
        e_new = torch.full((k,), -2.0, dtype=torch.float64, device="cuda")
        assert _call(lib, "lmx_dscaled_residual_dev", A, X, B, e_ref) == 0
        assert _call(lib, "lmx_dscaled_residual_noalloc_dev", A, X, B, e_new) == 0
        torch.cuda.synchronize()
        assert torch.equal(e_ref, e_new), (e_ref, e_new)
        assert bool((e_new > 0).all())


def test_noalloc_refuses_above_its_scratch(lmx_gpu):
    import torch
    lib = lmx_gpu.raw()
    n, k = 64, 17
    A = torch.zeros(n, n, dtype=torch.float64, device="cuda")
    X = torch.zeros(n, k, dtype=torch.float64, device="cuda")
    eta = torch.zeros(k, dtype=torch.float64, device="cuda")
    assert _call(lib, "lmx_dscaled_residual_noalloc_dev", A, X, X, eta) == lmx_gpu.LMX_ERR_ARG
    assert "16" in lmx_gpu.last_error()
