"""GPU FP32 LU (lmx_sgetrf) through the C ABI: lu at e = Float
(src/.../Solve/LU.hs:121, generic over Fractional e), the factor the
mixed-precision solvers refine with.

The pivots of an FP32 factorisation of a general matrix depend on FP32
rounding, so the pivot order is pinned with matrices whose pivots are
fixed by construction: A = Q (L U) with |l_ij| <= 1/2 below the diagonal
and 1 <= |u_jj| <= 2.  At step j the candidates are l_ij u_jj, so row j of
L U wins by a factor of two over every other row in exact arithmetic.  The
multipliers and the off-diagonal part of U shrink like 1/sqrt(i), which
keeps L and U well conditioned, so FP32 rounding cannot close that margin.
LAPACK's sgetrf (scipy) recovers the same pivots on these matrices
(test_known_pivots.py), which pins the construction.  The pivot-rule cases (first of equal magnitudes,
NaN never chosen below the diagonal) follow rawPivotSearch
(Internal/Kernel.hs:1541-1556), as the FP64 tests do."""
import numpy as np
import pytest
from conftest import swaps_to_perm

pytestmark = pytest.mark.gpu
EPS32 = float(np.finfo(np.float32).eps)


@pytest.mark.parametrize("n", [1, 5, 31, 32, 33, 200, 513, 1024, 1025, 1100, 2100, 4096])
def test_sgetrf_known_pivots(lmx_gpu, known_pivots, n):
    A, L, U, sigma = known_pivots(n, n)
    LU, ipiv, info = lmx_gpu.sgetrf(A)
    assert info == 0
    assert LU.dtype == np.float32
    np.testing.assert_array_equal(swaps_to_perm(ipiv), sigma)
    Lh = np.tril(LU.astype(float), -1) + np.eye(n)
    Uh = np.triu(LU.astype(float))
    # LAPACK's sgetrf reaches 1e-6 on these matrices
    assert np.abs(Lh - L).max() <= 1e-5
    assert np.abs(Uh - U).max() <= 2e-5


@pytest.mark.parametrize("n", [300, 1500, 3000])
def test_sgetrf_backward_error_and_multipliers(lmx_gpu, n):
    rng = np.random.default_rng(n + 7)
    A = rng.uniform(-1, 1, (n, n)).astype(np.float32)
    LU, ipiv, info = lmx_gpu.sgetrf(A)
    assert info == 0
    L = np.tril(LU.astype(float), -1) + np.eye(n)
    U = np.triu(LU.astype(float))
    # partial pivoting: no multiplier exceeds 1 in magnitude
    assert np.abs(np.tril(LU, -1)).max() <= 1.0
    perm = swaps_to_perm(ipiv)
    R = A.astype(float)[perm] - L @ U
    growth = np.abs(U).max() / np.abs(A).max()
    assert np.abs(R).max() / np.abs(A).max() <= 4 * n * EPS32 * max(growth, 1.0)


def test_sgetrf_ties_across_workgroups(lmx_gpu):
    # equal magnitudes in rows 3 and 1500 (different panel workgroups): the
    # first one wins
    n = 2100
    rng = np.random.default_rng(5)
    A = rng.uniform(-0.5, 0.5, (n, n)).astype(np.float32)
    A[3, 0], A[1500, 0] = -1.0, 1.0
    _, ipiv, _ = lmx_gpu.sgetrf(A)
    assert ipiv[0] == 3
    A[3, 0], A[1500, 0] = 0.25, -1.0
    A[1900, 0] = 1.0
    _, ipiv, _ = lmx_gpu.sgetrf(A)
    assert ipiv[0] == 1500


@pytest.mark.parametrize("n", [70, 2100])
def test_sgetrf_nan_pivot_rule(lmx_gpu, n):
    rng = np.random.default_rng(8)
    A = rng.uniform(-1, 1, (n, n)).astype(np.float32)
    A[n - 5, 0] = np.nan
    A[9, 0] = 2.0
    _, ipiv, _ = lmx_gpu.sgetrf(A)
    assert ipiv[0] == 9
    B = rng.uniform(-1, 1, (n, n)).astype(np.float32)
    B[0, 0] = np.nan
    B[n - 3, 0] = 5.0
    _, ipiv, _ = lmx_gpu.sgetrf(B)
    assert ipiv[0] == 0


@pytest.mark.parametrize("n,zero_col", [(5, 2), (700, 650), (2100, 1300)])
def test_sgetrf_singular_info(lmx_gpu, n, zero_col):
    rng = np.random.default_rng(n)
    A = rng.uniform(-1, 1, (n, n)).astype(np.float32)
    A[:, zero_col] = 0.0
    _, _, info = lmx_gpu.sgetrf(A)
    assert info == zero_col + 1


def test_sgetrf_matches_dgetrf_pivots_on_exact_data(lmx_gpu):
    # integer permutation-scaled data: every operation is exact in both
    # precisions, so the FP32 and FP64 factors agree bit for bit
    n = 600
    rng = np.random.default_rng(12)
    D = np.diag(rng.choice([-4.0, -2.0, 2.0, 4.0], n))
    A = np.eye(n)[rng.permutation(n)] @ D
    LU32, ip32, i32 = lmx_gpu.sgetrf(A)
    LU64, ip64, i64 = lmx_gpu.getrf(A)
    assert i32 == i64 == 0
    np.testing.assert_array_equal(ip32, ip64)
    np.testing.assert_array_equal(LU32.astype(float), LU64)
