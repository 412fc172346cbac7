"""Mixed-precision LU / Cholesky (FP32 factor on MFMA) with FP64 iterative
refinement, through the C ABI.  Pinned by the FP64 oracle: normwise
backward error <= 10 n eps and the solution matching the FP64 solve to
within kappa-scaled eps (conditionTolerance, test/Test/Residuals.hs:218)."""
import numpy as np
import pytest
import scipy.linalg as sl

pytestmark = pytest.mark.gpu
EPS = 2.220446049250313e-16


def haar(n, rng):
    # QR of a Gaussian with the sign fix (Mezzadri 2007)
    Q, R = np.linalg.qr(rng.standard_normal((n, n)))
    return Q * np.sign(np.diag(R))


def randsvd(n, kappa, seed, spd=False):
    rng = np.random.default_rng(seed)
    U = haar(n, rng)
    V = U if spd else haar(n, rng)
    sigma = kappa ** (-np.arange(n) / (n - 1))
    return (U * sigma) @ V.T


@pytest.mark.parametrize("kappa", [1e2, 1e4, 1e6])
@pytest.mark.parametrize("n,nrhs", [(300, 1), (1000, 4)])
def test_gesv_mixed_f32(lmx_gpu, orc, kappa, n, nrhs):
    A = randsvd(n, kappa, int(kappa) + n)
    B = np.random.default_rng(n).uniform(-1, 1, (n, nrhs))
    X, ri = lmx_gpu.gesv_mixed(A, B, lmx_gpu.PREC_F32)
    assert ri.converged == 1 and ri.fell_back == 0
    assert orc.scaled_residual_linear(A, X, B) <= 10 * n * EPS
    assert ri.backward_error <= np.sqrt(n) * EPS * 2
    X64 = sl.solve(A, B)
    err = np.abs(X - X64).max() / np.abs(X64).max()
    assert err <= 10 * orc.condition_tolerance(kappa, n)


def test_gesv_mixed_f32_many_rows_per_wave(lmx_gpu, orc):
    # n = 9300: the FP32 triangular solves of the refinement run 2 rows per
    # wave in their first row update and end in a ragged 84-row block
    n = 9300
    rng = np.random.default_rng(93)
    A = rng.uniform(-1, 1, (n, n))
    B = rng.uniform(-1, 1, (n, 1))
    X, ri = lmx_gpu.gesv_mixed(A, B, lmx_gpu.PREC_F32)
    assert ri.converged == 1 and ri.fell_back == 0
    assert orc.scaled_residual_linear(A, X, B) <= 10 * n * EPS


@pytest.mark.parametrize("kappa", [1e2, 1e4, 1e6])
@pytest.mark.parametrize("n,nrhs", [(300, 1), (1000, 4)])
def test_posv_mixed_f32(lmx_gpu, orc, kappa, n, nrhs):
    A = randsvd(n, kappa, int(kappa) + n + 1, spd=True)
    A = (A + A.T) / 2
    B = np.random.default_rng(n + 1).uniform(-1, 1, (n, nrhs))
    X, ri = lmx_gpu.posv_mixed(A, B, lmx_gpu.PREC_F32)
    assert ri.converged == 1
    assert orc.scaled_residual_linear(A, X, B) <= 10 * n * EPS
    X64 = sl.solve(A, B, assume_a="pos")
    err = np.abs(X - X64).max() / np.abs(X64).max()
    assert err <= 10 * orc.condition_tolerance(kappa, n)


def test_mixed_leaves_inputs_unchanged(lmx_gpu):
    n = 200
    A = randsvd(n, 1e3, 7)
    B = np.ones((n, 2))
    A0, B0 = A.copy(), B.copy()
    lmx_gpu.gesv_mixed(A, B)
    np.testing.assert_array_equal(A, A0)
    np.testing.assert_array_equal(B, B0)


def test_mixed_falls_back_beyond_fp32_range(lmx_gpu, orc):
    # kappa ~ 1e12: the FP32 factor cannot drive refinement; the FP64 path answers
    n = 200
    A = randsvd(n, 1e12, 3)
    b = np.random.default_rng(3).uniform(-1, 1, n)
    x, ri = lmx_gpu.gesv_mixed(A, b)
    assert ri.fell_back == 1
    assert orc.scaled_residual_linear(A, x, b) <= 10 * n * EPS


def test_device_mixed_large(lmx_gpu):
    import torch
    n, k = 4096, 16
    g = torch.Generator(device="cuda").manual_seed(1)
    A = torch.rand(n, n, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    B = torch.rand(n, k, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    X = torch.zeros_like(B)
    ri = lmx_gpu.gesv_mixed_dev(A, B, X, lmx_gpu.PREC_F32)
    torch.cuda.synchronize()
    assert ri.converged == 1 and ri.fell_back == 0
    R = B - A @ X
    num = R.abs().max(dim=0).values
    den = A.abs().sum(1).max() * X.abs().max(dim=0).values + B.abs().max(dim=0).values
    assert (num / den).max().item() <= 10 * n * EPS


# ---- the BF16 factor (trailing updates on BF16 tiles, GMRES-IR) ----------

@pytest.mark.parametrize("kappa", [1e2, 1e4, 1e6])
@pytest.mark.parametrize("n,nrhs", [(300, 1), (1000, 4)])
def test_gesv_mixed_bf16(lmx_gpu, orc, kappa, n, nrhs):
    A = randsvd(n, kappa, int(kappa) + n)
    B = np.random.default_rng(n).uniform(-1, 1, (n, nrhs))
    X, ri = lmx_gpu.gesv_mixed(A, B, lmx_gpu.PREC_BF16)
    assert ri.converged == 1 and ri.fell_back == 0
    assert orc.scaled_residual_linear(A, X, B) <= 10 * n * EPS
    X64 = sl.solve(A, B)
    err = np.abs(X - X64).max() / np.abs(X64).max()
    assert err <= 10 * orc.condition_tolerance(kappa, n)


@pytest.mark.parametrize("kappa", [1e2, 1e4, 1e6])
@pytest.mark.parametrize("n,nrhs", [(300, 1), (1000, 4)])
def test_posv_mixed_bf16(lmx_gpu, orc, kappa, n, nrhs):
    A = randsvd(n, kappa, int(kappa) + n + 1, spd=True)
    A = (A + A.T) / 2
    B = np.random.default_rng(n + 1).uniform(-1, 1, (n, nrhs))
    X, ri = lmx_gpu.posv_mixed(A, B, lmx_gpu.PREC_BF16)
    assert ri.converged == 1 and ri.fell_back == 0
    assert orc.scaled_residual_linear(A, X, B) <= 10 * n * EPS
    X64 = sl.solve(A, B, assume_a="pos")
    err = np.abs(X - X64).max() / np.abs(X64).max()
    assert err <= 10 * orc.condition_tolerance(kappa, n)


def test_bf16_falls_back_beyond_range(lmx_gpu, orc):
    n = 200
    A = randsvd(n, 1e13, 4)
    b = np.random.default_rng(4).uniform(-1, 1, n)
    x, ri = lmx_gpu.gesv_mixed(A, b, lmx_gpu.PREC_BF16)
    assert ri.fell_back == 1
    assert orc.scaled_residual_linear(A, x, b) <= 10 * n * EPS


def test_bf16_factor_is_not_the_fp32_factor(lmx_gpu):
    # the BF16 trailing updates change the factor (and the refinement path):
    # the iteration counts differ from the FP32 factor's on the same matrix
    n = 1000
    A = randsvd(n, 1e4, 11)
    b = np.random.default_rng(11).uniform(-1, 1, (n, 1))
    _, r32 = lmx_gpu.gesv_mixed(A, b, lmx_gpu.PREC_F32)
    _, r16 = lmx_gpu.gesv_mixed(A, b, lmx_gpu.PREC_BF16)
    assert r32.converged == 1 and r16.converged == 1
    assert (r32.iterations, r32.backward_error) != (r16.iterations, r16.backward_error)
