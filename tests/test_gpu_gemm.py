"""GPU GEMM (lmx_dgemm / lmx_sgemm through the C ABI) against the oracle's
matMulP / gemm, the reference KATs and exact integer products."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
EPS = 2.220446049250313e-16


def bound(A, B, k, alpha=1.0, beta=0.0, C0=None, eps=EPS):
    g = 2 * max(k, 1) * eps
    b = g * abs(alpha) * (np.abs(A) @ np.abs(B))
    if C0 is not None:
        b = b + 2 * eps * abs(beta) * np.abs(C0)
    return b + 1e-300


def test_kat_2x2(lmx_gpu, orc):
    # test/Test/BLAS.hs:132-144
    A = np.array([[1.0, 2.0], [3.0, 4.0]])
    B = np.array([[5.0, 6.0], [7.0, 8.0]])
    np.testing.assert_array_equal(lmx_gpu.matMul(A, B), [[19, 22], [43, 50]])
    np.testing.assert_array_equal(lmx_gpu.matMul(A, B), orc.matMulP(A, B))


def test_kat_gemm_3x3(lmx_gpu, orc):
    # test/Test/BLAS.hs:156-172: gemm 2.0 A B 0.5 I
    A = np.arange(1, 10, dtype=float).reshape(3, 3)
    B = 9.0 - np.arange(9, dtype=float).reshape(3, 3)
    C = lmx_gpu.gemm(2.0, A, B, 0.5, np.eye(3))
    np.testing.assert_array_equal(C, [[60.5, 48, 36], [168, 138.5, 108], [276, 228, 180.5]])
    np.testing.assert_array_equal(C, orc.gemm(2.0, A, B, 0.5, np.eye(3)))


@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (5, 7, 9), (16, 16, 4), (33, 17, 9), (64, 64, 16),
                                   (65, 129, 31), (128, 128, 16), (200, 300, 17), (513, 257, 129),
                                   (1000, 900, 700)])
def test_exact_integer_products(lmx_gpu, m, n, k):
    rng = np.random.default_rng(m + 7 * n + 13 * k)
    A = rng.integers(-8, 9, (m, k)).astype(float)
    B = rng.integers(-8, 9, (k, n)).astype(float)
    exact = (A.astype(np.int64) @ B.astype(np.int64)).astype(float)
    np.testing.assert_array_equal(lmx_gpu.dgemm(A, B), exact)
    np.testing.assert_array_equal(lmx_gpu.sgemm(A, B).astype(float), exact)


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N"), ("N", "T"), ("T", "T")])
@pytest.mark.parametrize("m,n,k", [(7, 5, 3), (130, 70, 100), (300, 200, 257)])
def test_dgemm_vs_oracle(lmx_gpu, orc, ta, tb, m, n, k):
    rng = np.random.default_rng(m * n + k)
    A = rng.uniform(-1, 1, (k, m) if ta == "T" else (m, k))
    B = rng.uniform(-1, 1, (n, k) if tb == "T" else (k, n))
    C0 = rng.uniform(-1, 1, (m, n))
    Ao = A.T if ta == "T" else A
    Bo = B.T if tb == "T" else B
    C = lmx_gpu.dgemm(A, B, C0, alpha=0.75, beta=-1.5, transa=ta, transb=tb)
    want = orc.gemm(0.75, Ao, Bo, -1.5, C0)
    assert np.all(np.abs(C - want) <= 2 * bound(Ao, Bo, k, 0.75, -1.5, C0))
    # matMulP with the reference's exact rounding order: agreement within the bound
    P = lmx_gpu.dgemm(A, B, transa=ta, transb=tb)
    assert np.all(np.abs(P - orc.matMulP(Ao, Bo)) <= 2 * bound(Ao, Bo, k))


def test_beta_zero_does_not_read_c(lmx_gpu):
    A = np.ones((40, 30))
    B = np.ones((30, 20))
    C0 = np.full((40, 20), np.nan)
    np.testing.assert_array_equal(lmx_gpu.dgemm(A, B, C0, beta=0.0), np.full((40, 20), 30.0))


def test_sgemm_tolerance(lmx_gpu):
    rng = np.random.default_rng(1)
    A = rng.uniform(-1, 1, (300, 500)).astype(np.float32)
    B = rng.uniform(-1, 1, (500, 200)).astype(np.float32)
    C = lmx_gpu.sgemm(A, B)
    ref = A.astype(float) @ B.astype(float)
    eps32 = 2.0 ** -23
    assert np.all(np.abs(C - ref) <= 2 * 500 * eps32 * (np.abs(A) @ np.abs(B)))


def test_device_dgemm_8192_checksum(lmx_gpu):
    # size-independent property at the 8192^3 shape: (A B) 1 = A (B 1)
    import torch
    n = 8192
    g = torch.Generator(device="cuda").manual_seed(5)
    A = torch.rand(n, n, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    B = torch.rand(n, n, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    C = torch.empty_like(A)
    lmx_gpu.dgemm_dev(A, B, C)
    torch.cuda.synchronize()
    one = torch.ones(n, 1, dtype=torch.float64, device="cuda")
    lhs = C @ one
    rhs = A @ (B @ one)
    scale = (A.abs() @ (B.abs() @ one)).max().item()
    assert (lhs - rhs).abs().max().item() <= 4 * n * EPS * scale
    # spot rows exactly against a float64 reference
    ref = A[:128].cpu().numpy() @ B.cpu().numpy()
    err = np.abs(C[:128].cpu().numpy() - ref).max()
    assert err <= 2 * n * EPS * np.abs(A[:128].cpu().numpy()).sum(1).max()


@pytest.mark.parametrize("m,n,k", [(3000, 3000, 520), (3001, 2999, 517)])
def test_large_tile_paths_exact(lmx_gpu, m, n, k):
    # 128 x 128 tiles: the NN kernel when rows are 16-byte aligned (first
    # shape), the general kernel otherwise (second); small integers keep every
    # product and partial sum exact in both precisions, beta = 1 included
    rng = np.random.default_rng(m + n + k)
    A = rng.integers(-8, 9, (m, k)).astype(float)
    B = rng.integers(-8, 9, (k, n)).astype(float)
    C0 = rng.integers(-8, 9, (m, n)).astype(float)
    exact = A @ B - C0  # integer-valued, |entries| < 2^24: exact in float64 and float32
    np.testing.assert_array_equal(lmx_gpu.dgemm(A, B, C0, alpha=1.0, beta=-1.0), exact)
    np.testing.assert_array_equal(lmx_gpu.sgemm(A, B, C0, alpha=1.0, beta=-1.0).astype(float), exact)


@pytest.mark.parametrize("m,n,k", [(3000, 3000, 520), (200, 300, 64), (3000, 2999, 516)])
def test_nt_kernel_exact(lmx_gpu, m, n, k):
    # op(B) = B^T with B's rows 16-byte aligned: the NN kernel loads B's tiles
    # along k like A's (128 and 64 tiles; n need not be a multiple of 4)
    rng = np.random.default_rng(m + 2 * n + k)
    A = rng.integers(-8, 9, (m, k)).astype(float)
    Bt = rng.integers(-8, 9, (n, k)).astype(float)
    C0 = rng.integers(-8, 9, (m, n)).astype(float)
    exact = A @ Bt.T + 2 * C0
    np.testing.assert_array_equal(lmx_gpu.dgemm(A, Bt, C0, alpha=1.0, beta=2.0, transb="T"), exact)
    np.testing.assert_array_equal(lmx_gpu.sgemm(A, Bt, C0, alpha=1.0, beta=2.0, transb="T").astype(float), exact)


@pytest.mark.parametrize("n,k", [(4100, 520), (1000, 64), (333, 36)])
def test_syrk_lower_tiles_exact(lmx_gpu, n, k):
    # A A^T on lower tiles only (the Cholesky update), mirrored by lmx_dsyrk
    rng = np.random.default_rng(n + k)
    A = rng.integers(-8, 9, (n, k)).astype(float)
    C0 = rng.integers(-8, 9, (n, n)).astype(float)
    C0 = C0 + C0.T
    np.testing.assert_array_equal(lmx_gpu.syrk(A, "N", alpha=-1.0, beta=1.0, c=C0), C0 - A @ A.T)
