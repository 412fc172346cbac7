"""GPU LU factor/solve through the C ABI against LAPACK golden factors, the
oracle's luSolveP / lu, the reference KATs and the backward-error bound
eta <= 10 n eps (scaledResidualLinear, test/Test/Residuals.hs:69-78)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
EPS = 2.220446049250313e-16


def eta(orc, A, x, b):
    return orc.scaled_residual_linear(A, x, b)


@pytest.mark.parametrize("n", [1, 2, 3, 5, 10, 12, 33, 63, 64, 65, 100, 130])
def test_getrf_vs_lapack_golden(lmx_gpu, orc, golden, n):
    A = golden[f"A_{n}"]
    LU, ipiv, info = lmx_gpu.getrf(A)
    assert info == 0
    np.testing.assert_array_equal(ipiv, golden[f"lapack_ipiv_{n}"])
    np.testing.assert_allclose(LU, golden[f"lapack_lu_{n}"], rtol=0, atol=1e-12 * max(1, n))
    # same pivots as the oracle's luSolveP factorisation
    LUo, swo = orc.lu_p_factor(A)
    np.testing.assert_array_equal(ipiv, swo)


@pytest.mark.parametrize("n", [200, 511, 512, 513, 1000, 1100, 2100])
def test_getrf_vs_oracle(lmx_gpu, orc, n):
    rng = np.random.default_rng(n)
    A = rng.uniform(-1, 1, (n, n))
    LU, ipiv, info = lmx_gpu.getrf(A)
    LUo, swo = orc.lu_p_factor(A)
    assert info == 0
    np.testing.assert_array_equal(ipiv, swo)
    assert np.abs(LU - LUo).max() <= 1e-10
    perm = lmx_gpu.swaps_to_perm(ipiv)
    assert orc.scaled_residual_lu(A, LU, perm) < 10


def test_lu_kat_3x3(lmx_gpu):
    # test/Test/Solve.hs:107-119, closed form x = [1, -1, 0]
    A = np.array([[2.0, 1, 1], [4, 3, 3], [8, 7, 9]])
    x = lmx_gpu.luSolve(A, np.ones(3))
    np.testing.assert_allclose(x, [1, -1, 0], atol=4 * EPS)
    LU, perm = lmx_gpu.lu(A)
    np.testing.assert_array_equal(perm, [2, 0, 1])


@pytest.mark.parametrize("n", [1, 2, 3, 31, 32, 33, 64, 100, 127, 128, 129, 257, 500, 1000])
@pytest.mark.parametrize("nrhs", [1, 16])
def test_gesv_backward_error_and_oracle(lmx_gpu, orc, n, nrhs):
    rng = np.random.default_rng(n * 31 + nrhs)
    A = rng.uniform(-1, 1, (n, n))
    B = rng.uniform(-1, 1, (n, nrhs))
    X, LU, ipiv, info = lmx_gpu.gesv(A, B)
    assert info == 0
    assert eta(orc, A, X, B) <= 10 * n * EPS
    # parity with the reference algorithm: the oracle's luSolveP, column by column
    kappa = np.linalg.cond(A, np.inf)
    for j in range(min(nrhs, 3)):
        xo = orc.luSolveP(A, B[:, j])
        err = np.abs(X[:, j] - xo).max() / max(np.abs(xo).max(), 1e-300)
        assert err <= 10 * orc.condition_tolerance(kappa, n) + 1e-14


@pytest.mark.parametrize("n", [3, 5, 10, 12])
def test_gesv_vs_mpmath(lmx_gpu, orc, golden, n):
    A, b, xe = golden[f"A_{n}"], golden[f"B_{n}"][:, 0], golden[f"mp_x_{n}"]
    x = lmx_gpu.luSolveP(A, b)
    kappa = np.linalg.cond(A, np.inf)
    assert np.abs(x - xe).max() / np.abs(xe).max() <= 10 * orc.condition_tolerance(kappa, n)


@pytest.mark.parametrize("n", [6, 70, 600])
def test_permutation_matrix_exact(lmx_gpu, n):
    rng = np.random.default_rng(n)
    P = np.eye(n)[rng.permutation(n)]
    b = rng.uniform(-1, 1, n)
    np.testing.assert_array_equal(lmx_gpu.luSolveP(P, b), P.T @ b)
    LU, perm = lmx_gpu.lu(P)
    np.testing.assert_array_equal(LU, np.eye(n))
    np.testing.assert_array_equal(P[perm], np.eye(n))


def test_hilbert5(lmx_gpu, orc):
    # test/Test/Solve.hs:157-165
    n = 5
    H = 1.0 / (np.arange(n)[:, None] + np.arange(n)[None, :] + 1)
    b = H @ np.arange(1, n + 1, dtype=float)
    x = lmx_gpu.luSolve(H, b)
    assert eta(orc, H, x, b) < 1e-4
    assert eta(orc, H, x, b) <= 10 * n * EPS


def test_tie_breaking_first_max(lmx_gpu, orc):
    # rawPivotSearch keeps the first of equal |a_ik| (Internal/Kernel.hs:1541-1556)
    A = np.array([[1.0, 2.0], [-1.0, 3.0]])
    LU, ipiv, info = lmx_gpu.getrf(A)
    np.testing.assert_array_equal(ipiv, [0, 1])
    np.testing.assert_array_equal(LU, orc.lu_p_factor(A)[0])
    # ties across panel workgroups: equal magnitudes in rows 3 and 900
    n = 1000
    rng = np.random.default_rng(3)
    A = rng.uniform(-0.5, 0.5, (n, n))
    A[3, 0], A[900, 0] = -1.0, 1.0
    LU, ipiv, info = lmx_gpu.getrf(A)
    assert ipiv[0] == 3
    np.testing.assert_array_equal(ipiv, orc.lu_p_factor(A)[1])


def test_nan_pivot_rule(lmx_gpu, orc):
    # rawPivotSearch (Internal/Kernel.hs:1541-1556) never selects a NaN below
    # row k, and keeps row k when a_kk itself is NaN
    n = 70
    rng = np.random.default_rng(8)
    A = rng.uniform(-1, 1, (n, n))
    A[5, 0] = np.nan
    A[9, 0] = 2.0
    _, ipiv, _ = lmx_gpu.getrf(A)
    assert ipiv[0] == 9
    assert ipiv[0] == orc.lu_p_factor(A)[1][0]
    B = rng.uniform(-1, 1, (n, n))
    B[0, 0] = np.nan
    B[3, 0] = 5.0
    _, ipiv, _ = lmx_gpu.getrf(B)
    assert ipiv[0] == 0
    assert ipiv[0] == orc.lu_p_factor(B)[1][0]


@pytest.mark.parametrize("n,zero_col", [(5, 2), (100, 40), (700, 650)])
def test_singular_info(lmx_gpu, n, zero_col):
    rng = np.random.default_rng(n)
    A = rng.uniform(-1, 1, (n, n))
    A[:, zero_col] = 0.0
    LU, ipiv, info = lmx_gpu.getrf(A)
    assert info == zero_col + 1


def test_getrs_reuses_factors(lmx_gpu, orc):
    rng = np.random.default_rng(11)
    n = 300
    A = rng.uniform(-1, 1, (n, n))
    LU, ipiv, info = lmx_gpu.getrf(A)
    for nrhs in (1, 5, 64):
        B = rng.uniform(-1, 1, (n, nrhs))
        X = lmx_gpu.getrs(LU, ipiv, B)
        assert eta(orc, A, X, B) <= 10 * n * EPS


def test_device_getrs_with_swap_list(lmx_gpu, orc):
    import torch
    rng = np.random.default_rng(12)
    n, k = 700, 4
    A = rng.uniform(-1, 1, (n, n))
    B = rng.uniform(-1, 1, (n, k))
    dA = torch.tensor(A, device="cuda")
    dB = torch.tensor(B, device="cuda")
    ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    lmx_gpu.getrf_dev(dA, ipiv, info)
    lmx_gpu.getrs_dev(dA, ipiv, dB)
    torch.cuda.synchronize()
    assert info.item() == 0
    assert eta(orc, A, dB.cpu().numpy(), B) <= 10 * n * EPS


@pytest.mark.parametrize("n", [4096, 8192])
def test_device_gesv_large(lmx_gpu, orc, n):
    # the n = 8192, 16 right-hand-side workload: backward error, size-independent
    import torch
    k = 16
    g = torch.Generator(device="cuda").manual_seed(n)
    A = torch.rand(n, n, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    B = torch.rand(n, k, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    LU = A.clone()
    X = B.clone()
    ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    lmx_gpu.gesv_dev(LU, ipiv, X, info)
    torch.cuda.synchronize()
    assert info.item() == 0
    R = B - A @ X
    num = R.abs().max(dim=0).values
    den = A.abs().sum(1).max() * X.abs().max(dim=0).values + B.abs().max(dim=0).values
    assert (num / den).max().item() <= 10 * n * EPS
    # pivots are a valid swap list
    iv = ipiv.cpu().numpy()
    assert np.all(iv >= np.arange(n)) and np.all(iv < n)


# ---- the generic lu: maximumBy's tie rule (Solve/LU.hs:121-167) ----------

def test_lu_tie_kat_2x2(lmx_gpu, orc):
    # maximumBy keeps the later of equal |a_ik| (Solve/LU.hs:128-135): the
    # reference's lu gives perm [1, 0] and LU [[-1, 3], [-1, 5]] here
    A = np.array([[1.0, 2.0], [-1.0, 3.0]])
    LU, perm = lmx_gpu.lu(A)
    np.testing.assert_array_equal(perm, [1, 0])
    np.testing.assert_array_equal(LU, [[-1.0, 3.0], [-1.0, 5.0]])
    LUo, permo = orc.lu(A)
    np.testing.assert_array_equal(perm, permo)
    np.testing.assert_array_equal(LU, LUo)
    LUf, permf = lmx_gpu.lu_float(A)
    np.testing.assert_array_equal(permf, [1, 0])
    np.testing.assert_array_equal(LUf, LUo.astype(np.float32))


@pytest.mark.parametrize("n,zero_cols", [(7, 0), (33, 1), (64, 0), (300, 2), (1000, 0), (2100, 3)])
def test_lu_ties_bit_exact_vs_oracle(lmx_gpu, orc, tie_blocks, n, zero_cols):
    # every block column is a tie between two rows scattered over the panel
    # workgroups (512 rows each); zero columns are n-way ties at 0 and skip
    # their elimination.  Exact arithmetic: the factors must be the oracle's bits.
    A = tie_blocks(n, 7 * n + zero_cols, zero_cols)
    LU, perm = lmx_gpu.lu(A)
    LUo, permo = orc.lu(A)
    np.testing.assert_array_equal(perm, permo)
    np.testing.assert_array_equal(LU, LUo)
    # the same matrix under luSolveP's rule (getrf): the first row of a tie.
    # (With a zero column luSolveP's packed path divides by zero and its
    # later pivots see NaN, where getrf reports info and skips the column,
    # so that comparison is made on the nonsingular matrices only.)
    LUg, ipiv, info = lmx_gpu.getrf(A)
    if zero_cols == 0:
        LUp, swp = orc.lu_p_factor(A)
        assert info == 0
        np.testing.assert_array_equal(ipiv, swp)
        np.testing.assert_array_equal(LUg, LUp)
        if n >= 33:
            assert not np.array_equal(lmx_gpu.swaps_to_perm(ipiv), perm)
    else:
        assert info > 0


@pytest.mark.parametrize("n", [64, 1000])
def test_lu_float_ties_bit_exact(lmx_gpu, orc, tie_blocks, n):
    # lu at e = Float: the integer data stays exact in FP32
    A = tie_blocks(n, 3 * n)
    LU, perm = lmx_gpu.lu_float(A)
    LUo, permo = orc.lu(A)
    np.testing.assert_array_equal(perm, permo)
    np.testing.assert_array_equal(LU, LUo.astype(np.float32))


def test_lu_nan_rule_vs_oracle(lmx_gpu, orc, tie_blocks):
    # maximumBy never moves past a NaN candidate to a later row and never
    # leaves a NaN at row k (compare gives GT whenever NaN is involved)
    A = tie_blocks(40, 11)
    c = int(np.flatnonzero(A[:, 0])[0])  # a row of column 0's tie
    below = [i for i in range(40) if A[i, 0] == 0 and i > 0]
    B = A.copy()
    B[below[-1], 0] = np.nan            # NaN below row 0: never chosen
    C = A.copy()
    C[0, 0] = np.nan                    # NaN on the diagonal: kept
    for M in (B, C):
        LU, perm = lmx_gpu.lu(M)
        LUo, permo = orc.lu(M)
        np.testing.assert_array_equal(perm, permo)
        np.testing.assert_array_equal(LU, LUo)
    assert c >= 0


@pytest.mark.parametrize("n", [3, 10, 33, 64, 100, 130])
def test_lu_golden_vs_oracle(lmx_gpu, orc, golden, n):
    # random data has no ties: lu's pivots are LAPACK's and the oracle's
    A = golden[f"A_{n}"]
    LU, perm = lmx_gpu.lu(A)
    LUo, permo = orc.lu(A)
    np.testing.assert_array_equal(perm, permo)
    np.testing.assert_allclose(LU, LUo, rtol=0, atol=1e-12 * n)


def test_lu_dev_matches_host(lmx_gpu, orc, tie_blocks):
    import torch
    n = 700
    A = tie_blocks(n, 5, 1)
    dA = torch.tensor(A, device="cuda")
    perm = torch.zeros(n, dtype=torch.int32, device="cuda")
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    lmx_gpu.lu_dev(dA, perm, info)
    torch.cuda.synchronize()
    LUo, permo = orc.lu(A)
    np.testing.assert_array_equal(perm.cpu().numpy(), permo)
    np.testing.assert_array_equal(dA.cpu().numpy(), LUo)
    assert info.item() > 0  # the zero column is reported (1-based)


@pytest.mark.parametrize("n", [16384, 20000])
def test_getrs_dev_large_swap_list(lmx_gpu, n):
    # the swap-list conversion of getrs_dev above n = 16384
    import torch
    g = torch.Generator(device="cuda").manual_seed(n)
    A = torch.rand(n, n, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    B = torch.rand(n, 2, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    LU = A.clone()
    ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    lmx_gpu.getrf_dev(LU, ipiv, info)
    X = B.clone()
    lmx_gpu.getrs_dev(LU, ipiv, X)
    torch.cuda.synchronize()
    assert info.item() == 0
    R = B - A @ X
    num = R.abs().max(dim=0).values
    den = A.abs().sum(1).max() * X.abs().max(dim=0).values + B.abs().max(dim=0).values
    assert (num / den).max().item() <= 10 * n * EPS


def test_dev_wrappers_refuse_bad_tensors(lmx_gpu):
    # checked in Python before the pointers reach the library
    import torch
    n = 64
    A = torch.rand(n, n, dtype=torch.float64, device="cuda")
    ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    with pytest.raises(ValueError, match="dtype"):
        lmx_gpu.getrf_dev(A.float(), ipiv, info)
    with pytest.raises(ValueError, match="square"):
        lmx_gpu.getrf_dev(A[:, :32], ipiv, info)
    with pytest.raises(ValueError, match="stride"):
        lmx_gpu.getrf_dev(A.t(), ipiv, info)
    with pytest.raises(ValueError, match="elements"):
        lmx_gpu.getrf_dev(A, ipiv[:10], info)
    with pytest.raises(ValueError, match="dtype"):
        lmx_gpu.getrf_dev(A, ipiv.long(), info)
    with pytest.raises(ValueError, match="rows"):
        lmx_gpu.getrs_dev(A, ipiv, torch.zeros(n + 1, 2, dtype=torch.float64, device="cuda"))
    with pytest.raises(ValueError, match="conform"):
        lmx_gpu.dgemm_dev(A, A[:32], torch.zeros(n, n, dtype=torch.float64, device="cuda"))
    X = torch.zeros(n, 3, dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError, match="columns"):
        lmx_gpu.gesv_mixed_dev(A, torch.zeros(n, 2, dtype=torch.float64, device="cuda"), X)
