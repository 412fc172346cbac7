"""BF16 matrix-core GEMM (v_mfma_f32_16x16x32_bf16, FP32 accumulation)
through the C ABI: exact products on integer data (which pins the operand
and accumulator lane layouts), all four transposes, both the vector and the
scalar load paths and both tile sizes, and the error bound of the split
(bf16x3) FP32 GEMM against an FP64 product."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [(16, 16, 32), (64, 64, 64), (37, 53, 71), (100, 96, 200), (2944, 2944, 64), (2945, 2939, 70)]


def _op(X, t):
    return X.T if t == "T" else X


@pytest.mark.parametrize("m,n,k", SHAPES)
@pytest.mark.parametrize("ta,tb", [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")])
def test_gemm_bf16_exact_integers(lmx_gpu, m, n, k, ta, tb):
    # small integers are exact in BF16 and every partial sum is exact in FP32
    rng = np.random.default_rng(m * 7 + n * 3 + k + (ta == "T") * 2 + (tb == "T"))
    A = rng.integers(-8, 9, (k, m) if ta == "T" else (m, k)).astype(np.float32)
    B = rng.integers(-8, 9, (n, k) if tb == "T" else (k, n)).astype(np.float32)
    C0 = rng.integers(-8, 9, (m, n)).astype(np.float32)
    ref = 2.0 * (_op(A, ta).astype(np.float64) @ _op(B, tb)) - 1.0 * C0
    C = lmx_gpu.gemm_bf16(lmx_gpu.to_bf16_bits(A), lmx_gpu.to_bf16_bits(B), C0, alpha=2.0, beta=-1.0,
                          transa=ta, transb=tb)
    np.testing.assert_array_equal(C, ref.astype(np.float32))


@pytest.mark.parametrize("m,n,k", [(64, 64, 256), (300, 200, 256), (2944, 2944, 256), (1030, 517, 33)])
@pytest.mark.parametrize("ta,tb", [("N", "N"), ("N", "T"), ("T", "N")])
def test_sgemm_bf16x3_error_bound(lmx_gpu, m, n, k, ta, tb):
    # hi*hi + hi*lo + lo*hi: each operand is hi + lo to within 2^-17 of itself
    # and the dropped lo*lo is below 2^-18 of the product; FP32 sums add k u32.
    import torch
    rng = np.random.default_rng(m + n + k)
    A = rng.uniform(-1, 1, (k, m) if ta == "T" else (m, k)).astype(np.float32)
    B = rng.uniform(-1, 1, (n, k) if tb == "T" else (k, n)).astype(np.float32)
    C0 = rng.uniform(-1, 1, (m, n)).astype(np.float32)
    dA, dB, dC = (torch.tensor(x, device="cuda") for x in (A, B, C0))
    lmx_gpu.sgemm_bf16x3_dev(dA, dB, dC, alpha=-1.0, beta=1.0, transa=ta, transb=tb)
    C = dC.cpu().numpy()
    Ad, Bd = _op(A, ta).astype(np.float64), _op(B, tb).astype(np.float64)
    ref = C0 - Ad @ Bd
    bound = (2.0 ** -16 + (k + 2) * 2.0 ** -24) * (np.abs(Ad) @ np.abs(Bd) + np.abs(C0))
    err = np.abs(C - ref)
    assert np.all(err <= bound), float((err / bound).max())
    # and it is clearly better than a single BF16 rounding
    assert err.max() < 2.0 ** -10 * np.abs(Ad).max() * np.abs(Bd).max() * k


def test_sgemm_bf16x3_integers_exact(lmx_gpu):
    import torch
    rng = np.random.default_rng(5)
    m, n, k = 512, 384, 256
    A = rng.integers(-100, 101, (m, k)).astype(np.float32)   # lo = 0 for |x| <= 256
    B = rng.integers(-100, 101, (k, n)).astype(np.float32)
    dA, dB = torch.tensor(A, device="cuda"), torch.tensor(B, device="cuda")
    dC = torch.zeros(m, n, dtype=torch.float32, device="cuda")
    lmx_gpu.sgemm_bf16x3_dev(dA, dB, dC)
    np.testing.assert_array_equal(dC.cpu().numpy(), (A.astype(np.float64) @ B).astype(np.float32))


def test_gemm_bf16_dev_torch(lmx_gpu):
    import torch
    m, n, k = 256, 192, 128
    g = torch.Generator(device="cuda").manual_seed(3)
    A = torch.randint(-16, 17, (m, k), device="cuda", generator=g).to(torch.bfloat16)
    B = torch.randint(-16, 17, (k, n), device="cuda", generator=g).to(torch.bfloat16)
    C = torch.zeros(m, n, dtype=torch.float32, device="cuda")
    lmx_gpu.gemm_bf16_dev(A, B, C)
    ref = A.double() @ B.double()
    assert torch.equal(C.double(), ref)
