"""Argument checks of the device-pointer wrappers (lmx/__init__.py): a bad
tensor must raise in Python before any pointer reaches the library.  These
run on the CPU: a CPU tensor or a non-tensor is refused without a GPU."""
import numpy as np
import pytest


def test_dev_wrappers_refuse_host_tensors():
    import torch
    import lmx
    A = torch.zeros(4, 4, dtype=torch.float64)
    ipiv = torch.zeros(4, dtype=torch.int32)
    info = torch.zeros(1, dtype=torch.int32)
    with pytest.raises(ValueError, match="CUDA"):
        lmx.getrf_dev(A, ipiv, info, stream=0)
    with pytest.raises(ValueError, match="CUDA"):
        lmx.lu_dev(A, ipiv, info, stream=0)
    with pytest.raises(ValueError, match="CUDA"):
        lmx.dgemm_dev(A, A, A, stream=0)
    with pytest.raises(TypeError):
        lmx.potrf_dev(np.zeros((4, 4)), info, stream=0)


def test_swaps_to_perm_matches_oracle(orc):
    import lmx
    rng = np.random.default_rng(0)
    n = 50
    ipiv = np.array([rng.integers(k, n) for k in range(n)])
    np.testing.assert_array_equal(lmx.swaps_to_perm(ipiv), orc.swaps_to_perm(ipiv))


def test_bf16_bits_round_to_nearest_even():
    import lmx
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -2.5, 1.0 + 2.0 ** -7], np.float32)
    y = lmx.from_bf16_bits(lmx.to_bf16_bits(x))
    np.testing.assert_array_equal(y, np.array([1.0, 1.0, 1.0 + 2.0 ** -6, -2.5, 1.0 + 2.0 ** -7], np.float32))
