"""Pins the CPU oracle (oracle/lmx_oracle.c) before it is trusted as the
checker: reference KATs, LAPACK factors, closed forms, exact rational
emulation of the reference's rounding, and mpmath solutions.  No GPU."""
from fractions import Fraction

import numpy as np
import pytest
import scipy.linalg as sl

EPS = 2.220446049250313e-16


# ---------------------------------------------------------------- helpers

def fma_exact(a, b, c):
    """Correctly rounded fma(a, b, c): exact rational result, rounded once
    (int/int true division in CPython is correctly rounded)."""
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def rawkernel_element(a_row, b_col, tail):
    """Reference rounding of one C element of rawGemmKernel from C = 0
    (Internal/Kernel.hs:145-806): an ascending FMA chain, or -- for the last
    n%4 columns -- per-64-k-tile mul+add partial sums added to C."""
    k = len(a_row)
    if not tail:
        acc = 0.0
        for kk in range(k):
            acc = fma_exact(a_row[kk], b_col[kk], acc)
        return acc
    c = 0.0
    for bk in range(0, k, 64):
        part = 0.0
        for kk in range(bk, min(bk + 64, k)):
            part = part + a_row[kk] * b_col[kk]
        c = c + part
    return c


# ---------------------------------------------------------------- GEMM

def test_matmul_kat_2x2(orc):
    # test/Test/BLAS.hs:132-144
    A = np.array([[1.0, 2.0], [3.0, 4.0]])
    B = np.array([[5.0, 6.0], [7.0, 8.0]])
    np.testing.assert_array_equal(orc.matMulP(A, B), [[19, 22], [43, 50]])
    np.testing.assert_array_equal(orc.matMulGeneric(A, B), [[19, 22], [43, 50]])


def test_gemm_kat_3x3(orc):
    # test/Test/BLAS.hs:156-172: gemm 2.0 A B 0.5 I
    A = np.arange(1, 10, dtype=float).reshape(3, 3)
    B = 9.0 - np.arange(9, dtype=float).reshape(3, 3)
    C = orc.gemm(2.0, A, B, 0.5, np.eye(3))
    np.testing.assert_array_equal(C, [[60.5, 48, 36], [168, 138.5, 108], [276, 228, 180.5]])


@pytest.mark.parametrize("m,k,n", [(5, 7, 9), (97, 101, 99), (130, 150, 131), (1, 1, 1), (3, 0, 4)])
def test_matmulp_exact_integers(orc, m, k, n):
    # small integers: every product and partial sum is exact, so any order
    # gives the exact product -- pins indexing / transposition.
    rng = np.random.default_rng(m * 1000 + k * 10 + n)
    A = rng.integers(-8, 9, (m, k)).astype(float)
    B = rng.integers(-8, 9, (k, n)).astype(float)
    exact = (A.astype(np.int64) @ B.astype(np.int64)).astype(float)
    np.testing.assert_array_equal(orc.matMulP(A, B), exact)


@pytest.mark.parametrize("m,k,n,rows", [(6, 150, 11, [0, 3, 5]), (97, 130, 99, [0, 50, 96])])
def test_matmulp_rounding_bit_exact(orc, m, k, n, rows):
    # Exact rational emulation of the reference's per-element rounding;
    # (97,130,99) takes the packed path (min >= 96), (6,150,11) the unpacked.
    rng = np.random.default_rng(7 + n)
    A = rng.uniform(-1, 1, (m, k))
    B = rng.uniform(-1, 1, (k, n))
    C = orc.matMulP(A, B)
    n4 = n - n % 4
    sensitive = False
    for i in rows:
        for j in range(n):
            want = rawkernel_element(A[i], B[:, j], j >= n4)
            assert C[i, j] == want, (i, j)
            other = rawkernel_element(A[i], B[:, j], not (j >= n4))
            sensitive |= other != want
    # the check can tell an FMA chain from mul+add partial sums
    assert sensitive


def test_matmulp_error_bound(orc):
    rng = np.random.default_rng(3)
    A = rng.uniform(-1, 1, (200, 300))
    B = rng.uniform(-1, 1, (300, 150))
    C = orc.matMulP(A, B)
    gamma = 300 * EPS / (1 - 300 * EPS)
    bound = gamma * (np.abs(A) @ np.abs(B))
    assert np.all(np.abs(C - A @ B) <= 2 * bound)


def test_gemm_generic_rounding_bit_exact(orc):
    rng = np.random.default_rng(11)
    m, k, n = 4, 40, 5
    A, B, C0 = rng.uniform(-1, 1, (m, k)), rng.uniform(-1, 1, (k, n)), rng.uniform(-1, 1, (m, n))
    alpha, beta = 0.7, -1.3
    C = orc.gemm(alpha, A, B, beta, C0)
    for i in range(m):
        for j in range(n):
            acc = beta * C0[i, j]
            for l in range(k):
                acc = acc + (alpha * A[i, l]) * B[l, j]
            assert C[i, j] == acc


# ---------------------------------------------------------------- LU

KAT_A = np.array([[2.0, 1, 1], [4, 3, 3], [8, 7, 9]])


def test_lu_kat_3x3(orc):
    # test/Test/Solve.hs:107-119; closed form x = [1, -1, 0]
    b = np.ones(3)
    for x in (orc.luSolveP(KAT_A, b), orc.luSolve(KAT_A, b)):
        np.testing.assert_allclose(x, [1, -1, 0], atol=4 * EPS)
    LU, sw = orc.lu_p_factor(KAT_A)
    lu, piv = sl.lu_factor(KAT_A)
    np.testing.assert_array_equal(sw, [2, 2, 2])
    np.testing.assert_array_equal(sw, piv)
    np.testing.assert_allclose(LU, lu, rtol=0, atol=4 * EPS)
    np.testing.assert_allclose(LU[:, 0], [8, 0.25, 0.5])


def test_lu_tie_breaking(orc):
    # generic lu: maximumBy keeps the later of equal |a_ik| (Solve/LU.hs:135);
    # luSolveP: rawPivotSearch keeps the first (Kernel.hs:1541-1556).
    A = np.array([[1.0, 2.0], [-1.0, 3.0]])
    LU, perm = orc.lu(A)
    np.testing.assert_array_equal(perm, [1, 0])
    np.testing.assert_array_equal(LU, [[-1, 3], [-1, 5]])
    LUp, sw = orc.lu_p_factor(A)
    np.testing.assert_array_equal(sw, [0, 1])
    np.testing.assert_array_equal(LUp, [[1, 2], [-1, 5]])


def _lu_last_max_replay(A):
    """Outer-product LU with partial pivoting (GVL4 Algorithm 3.4.1) written
    with numpy row operations; the pivot is the last row of largest |a_ik|
    (argmax over the reversed column).  On tie-block matrices every value is
    an integer, which is asserted, so the replay is exact."""
    A = np.array(A, dtype=np.float64)
    n = A.shape[0]
    perm = np.arange(n)
    for k in range(n - 1):
        col = np.abs(A[k:, k])
        p = k + (len(col) - 1 - int(np.argmax(col[::-1])))
        A[[k, p]] = A[[p, k]]
        perm[[k, p]] = perm[[p, k]]
        if A[k, k] != 0:
            mult = A[k + 1:, k] / A[k, k]
            A[k + 1:, k] = mult
            A[k + 1:, k + 1:] -= np.outer(mult, A[k, k + 1:])
    return A, perm


@pytest.mark.parametrize("n,zero_cols", [(2, 0), (7, 0), (12, 1), (40, 0), (90, 3)])
def test_lu_tie_blocks_exact(orc, tie_blocks, n, zero_cols):
    # every pivot search of a block column is a tie; the generic lu takes the
    # later row (maximumBy), and the arithmetic is exact integer work
    A = tie_blocks(n, 100 + n, zero_cols)
    LU, perm = orc.lu(A)
    LUr, permr = _lu_last_max_replay(A)
    np.testing.assert_array_equal(perm, permr)
    np.testing.assert_array_equal(LU, LUr)
    L = np.tril(LU, -1) + np.eye(n)
    U = np.triu(LU)
    assert np.all(np.isin(np.tril(LU, -1), [-1.0, 0.0, 1.0]))
    np.testing.assert_array_equal(A[perm], L @ U)
    # the first-max rule of luSolveP picks other rows on the same matrix
    if n >= 7:
        _, sw = orc.lu_p_factor(A)
        assert not np.array_equal(orc.swaps_to_perm(sw), perm)


@pytest.mark.parametrize("n", [1, 2, 3, 5, 10, 12, 33, 63, 64, 65, 100, 130])
def test_lu_p_factor_vs_lapack(orc, golden, n):
    A = golden[f"A_{n}"]
    LU, sw = orc.lu_p_factor(A)
    np.testing.assert_array_equal(sw, golden[f"lapack_ipiv_{n}"])
    np.testing.assert_allclose(LU, golden[f"lapack_lu_{n}"], rtol=0, atol=1e-12 * max(1, n))
    perm = orc.swaps_to_perm(sw)
    assert orc.scaled_residual_lu(A, LU, perm) < 10


@pytest.mark.parametrize("n", [3, 5, 10, 33, 64])
def test_lu_generic_vs_lapack(orc, golden, n):
    A = golden[f"A_{n}"]
    LU, perm = orc.lu(A)
    lu, piv = sl.lu_factor(A)
    np.testing.assert_array_equal(perm, orc.swaps_to_perm(piv))
    np.testing.assert_allclose(LU, lu, rtol=0, atol=1e-12 * n)


@pytest.mark.parametrize("n", [64, 70, 97])
def test_lu_blocked_equals_unblocked(orc, n):
    rng = np.random.default_rng(n)
    A = rng.uniform(-1, 1, (n, n))
    LUb, swb = orc.lu_panel_factor(A, 32)
    LUu, swu = orc.lu_unblocked_factor(A)
    np.testing.assert_array_equal(swb[: n - 1], swu[: n - 1])
    np.testing.assert_allclose(LUb, LUu, rtol=0, atol=1e-12)


@pytest.mark.parametrize("n", [6, 70])
def test_lu_permutation_matrix(orc, n):
    # A = permutation matrix: pivots are forced and x = A^T b exactly
    rng = np.random.default_rng(n)
    P = np.eye(n)[rng.permutation(n)]
    b = rng.uniform(-1, 1, n)
    np.testing.assert_array_equal(orc.luSolveP(P, b), P.T @ b)
    np.testing.assert_array_equal(orc.luSolve(P, b), P.T @ b)
    LU, sw = orc.lu_p_factor(P)
    np.testing.assert_array_equal(LU, np.eye(n))
    perm = orc.swaps_to_perm(sw)
    np.testing.assert_array_equal(P[perm], np.eye(n))


def test_lu_hilbert5(orc):
    # test/Test/Solve.hs:157-165 (threshold 1e-4); also within 10 n eps
    n = 5
    H = 1.0 / (np.arange(n)[:, None] + np.arange(n)[None, :] + 1)
    x0 = np.arange(1, n + 1, dtype=float)
    b = H @ x0
    for x in (orc.luSolve(H, b), orc.luSolveP(H, b)):
        eta = orc.scaled_residual_linear(H, x, b)
        assert eta < 1e-4 and eta <= 10 * n * EPS


@pytest.mark.parametrize("n", [3, 5, 10, 12])
def test_lu_vs_mpmath(orc, golden, n):
    A, b, xe = golden[f"A_{n}"], golden[f"B_{n}"][:, 0], golden[f"mp_x_{n}"]
    kappa = np.linalg.cond(A, np.inf)
    for x in (orc.luSolveP(A, b), orc.luSolve(A, b)):
        err = np.abs(x - xe).max() / np.abs(xe).max()
        assert err <= 10 * orc.condition_tolerance(kappa, n)


@pytest.mark.parametrize("n", [10, 100, 300])
def test_lusolvep_backward_error(orc, n):
    rng = np.random.default_rng(n + 5)
    A = rng.uniform(-1, 1, (n, n))
    b = rng.uniform(-1, 1, n)
    assert orc.scaled_residual_linear(A, orc.luSolveP(A, b), b) <= 10 * n * EPS
    if n <= 100:
        assert orc.scaled_residual_linear(A, orc.luSolve(A, b), b) <= 10 * n * EPS


def test_lu_solve_factored_matches_lusolvep(orc):
    rng = np.random.default_rng(9)
    n = 120
    A = rng.uniform(-1, 1, (n, n))
    b = rng.uniform(-1, 1, n)
    LU, sw = orc.lu_p_factor(A)
    np.testing.assert_array_equal(orc.lu_solve_factored(LU, sw, b), orc.luSolveP(A, b))


# ---------------------------------------------------------------- triangular

def test_triangular_integer_solutions(orc):
    rng = np.random.default_rng(5)
    n = 9
    L = np.tril(rng.integers(-3, 4, (n, n)), -1).astype(float) + np.eye(n)
    U = L.T.copy()
    x = rng.integers(-5, 6, n).astype(float)
    np.testing.assert_array_equal(orc.forwardSubUnit(L, L @ x), x)
    np.testing.assert_array_equal(orc.backSubUnit(U, U @ x), x)
    D = np.diag(rng.choice([1.0, 2.0, 4.0], n))
    np.testing.assert_array_equal(orc.forwardSub(L @ D, (L @ D) @ x), x)
    np.testing.assert_array_equal(orc.backSub(D @ U, (D @ U) @ x), x)


def test_triangular_vs_scipy(orc):
    rng = np.random.default_rng(6)
    n = 40
    T = rng.uniform(-1, 1, (n, n)) + np.diag(rng.uniform(1, 10, n))
    L, U = np.tril(T), np.triu(T)
    b = rng.uniform(-1, 1, n)
    np.testing.assert_allclose(orc.forwardSub(L, b), sl.solve_triangular(L, b, lower=True), rtol=1e-12)
    np.testing.assert_allclose(orc.backSub(U, b), sl.solve_triangular(U, b, lower=False), rtol=1e-12)
    np.testing.assert_allclose(orc.forwardSubUnit(L, b),
                               sl.solve_triangular(L, b, lower=True, unit_diagonal=True), rtol=1e-12)
    np.testing.assert_allclose(orc.backSubUnit(U, b),
                               sl.solve_triangular(U, b, lower=False, unit_diagonal=True), rtol=1e-12)
    # unit variants ignore the stored diagonal
    np.testing.assert_array_equal(orc.forwardSubUnit(L, b), orc.forwardSubUnit(np.tril(L, -1), b))


# ---------------------------------------------------------------- Cholesky

@pytest.mark.parametrize("n", [1, 3, 10, 33, 63, 64, 65, 100, 130])
def test_cholesky_vs_lapack(orc, golden, n):
    S = golden[f"S_{n}"]
    L = golden[f"lapack_chol_{n}"]
    Gp = orc.cholesky_p_factor(S)
    np.testing.assert_allclose(np.tril(Gp), L, rtol=0, atol=1e-13)
    np.testing.assert_allclose(orc.cholesky(S), L, rtol=0, atol=1e-13)
    assert orc.scaled_residual_cholesky(S, np.tril(Gp)) < 10


def test_cholesky_p_upper_block_residue(orc):
    # panelCholFactor (Solve/Cholesky.hs:296-312) subtracts the whole
    # nBelow x jb product, so the strict upper triangle of every diagonal
    # block after the first holds -sum_k g_ik g_jk; the substitutions never
    # read it.  A restatement that clipped the subtraction would miss this.
    n = 100
    rng = np.random.default_rng(2)
    G = rng.standard_normal((n, n))
    S = G @ G.T / n + np.eye(n)
    Gp = orc.cholesky_p_factor(S)
    assert np.all(np.triu(Gp[:32, :32], 1) == 0)
    blk = Gp[32:64, 32:64]
    assert np.abs(np.triu(blk, 1)).max() > 0
    assert np.all(np.triu(Gp, 1)[:, 64:][:32] == 0)
    L = np.tril(Gp)
    i, j = 32, 40
    np.testing.assert_allclose(Gp[i, j], -(L[i, :32] @ L[j, :32]), rtol=1e-12)


@pytest.mark.parametrize("n", [10, 100])
def test_cholesky_tridiagonal_closed_form(orc, n):
    # tridiag(-1, 2, -1): g_ii = sqrt((i+2)/(i+1)), g_{i+1,i} = -sqrt((i+1)/(i+2))
    T = 2 * np.eye(n) - np.eye(n, k=1) - np.eye(n, k=-1)
    i = np.arange(n)
    want = np.diag(np.sqrt((i + 2) / (i + 1))) + np.diag(-np.sqrt((i[:-1] + 1) / (i[:-1] + 2)), -1)
    np.testing.assert_allclose(np.tril(orc.cholesky_p_factor(T)), want, rtol=0, atol=4 * EPS)
    np.testing.assert_allclose(orc.cholesky(T), want, rtol=0, atol=4 * EPS)


def test_cholesky_diagonal_closed_form(orc):
    d = np.array([4.0, 9.0, 16.0, 2.0, 0.25] * 14)
    n = len(d)
    np.testing.assert_array_equal(np.tril(orc.cholesky_p_factor(np.diag(d))), np.diag(np.sqrt(d)))
    b = np.arange(n, dtype=float)
    np.testing.assert_allclose(orc.choleskySolveP(np.diag(d), b), b / d, rtol=EPS)


@pytest.mark.parametrize("n", [3, 5, 10, 12])
def test_cholesky_vs_mpmath(orc, golden, n):
    S, b, xe = golden[f"S_{n}"], golden[f"B_{n}"][:, 0], golden[f"mp_xs_{n}"]
    kappa = np.linalg.cond(S, np.inf)
    for x in (orc.choleskySolveP(S, b), orc.choleskySolve(S, b)):
        err = np.abs(x - xe).max() / np.abs(xe).max()
        assert err <= 10 * orc.condition_tolerance(kappa, n)


@pytest.mark.parametrize("n", [10, 100, 200])
def test_cholsolvep_backward_error(orc, n):
    rng = np.random.default_rng(n)
    G = rng.standard_normal((n, n))
    S = G @ G.T / n + np.eye(n)
    b = rng.uniform(-1, 1, n)
    assert orc.scaled_residual_linear(S, orc.choleskySolveP(S, b), b) <= 10 * n * EPS
    assert orc.scaled_residual_linear(S, orc.choleskySolve(S, b), b) <= 10 * n * EPS


def test_cholesky_reads_lower_triangle_only(orc):
    rng = np.random.default_rng(4)
    n = 70
    G = rng.standard_normal((n, n))
    S = G @ G.T / n + np.eye(n)
    junk = S + np.triu(rng.uniform(-5, 5, (n, n)), 1)
    b = rng.uniform(-1, 1, n)
    np.testing.assert_array_equal(orc.choleskySolveP(junk, b), orc.choleskySolveP(S, b))
    np.testing.assert_array_equal(orc.cholesky(junk), orc.cholesky(S))


# ---- level 1/2, norms, SYRK, det, inverse iteration ----------------------

from fractions import Fraction  # noqa: E402


def _fma_exact(a, b, c):
    # one rounding of the exact a*b + c: the semantics of fma()
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def _raw_dot_emulated(x, y):
    # rawDot# (Internal/Kernel.hs:89-110) in exact arithmetic with explicit
    # roundings: four FMA lanes, ((a+b)+c)+d, then mul+add for the tail
    n = len(x)
    n4 = n - n % 4
    acc = [0.0] * 4
    for i in range(0, n4, 4):
        for lane in range(4):
            acc[lane] = _fma_exact(x[i + lane], y[i + lane], acc[lane])
    s = float(Fraction(float(Fraction(acc[0]) + Fraction(acc[1]))) + Fraction(acc[2]))
    s = float(Fraction(s) + Fraction(acc[3]))
    for i in range(n4, n):
        p = float(Fraction(x[i]) * Fraction(y[i]))
        s = float(Fraction(s) + Fraction(p))
    return s


@pytest.mark.parametrize("n", [1, 3, 4, 7, 8, 13, 64, 101])
def test_raw_dot_rounding_bit_exact(orc, n):
    rng = np.random.default_rng(n)
    x, y = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    assert orc.rawDot(x, y) == _raw_dot_emulated(x, y)


def test_matvecp_rows_are_raw_dots_and_exact_integers(orc):
    rng = np.random.default_rng(2)
    A, x = rng.uniform(-1, 1, (9, 37)), rng.uniform(-1, 1, 37)
    y = orc.matvecP(A, x)
    for i in range(9):
        assert y[i] == _raw_dot_emulated(A[i], x)
    Ai, xi = rng.integers(-50, 51, (20, 30)).astype(float), rng.integers(-50, 51, 30).astype(float)
    np.testing.assert_array_equal(orc.matvecP(Ai, xi), Ai @ xi)


def test_gemv_generic_mul_add_chain(orc):
    rng = np.random.default_rng(3)
    A, x, y = rng.uniform(-1, 1, (6, 11)), rng.uniform(-1, 1, 11), rng.uniform(-1, 1, 6)
    out = orc.gemv(0.5, A, x, -2.0, y)
    for i in range(6):
        acc = 0.0
        for j in range(11):
            acc = float(Fraction(acc) + Fraction(float(Fraction(A[i, j]) * Fraction(x[j]))))
        assert out[i] == float(Fraction(0.5 * acc) + Fraction(-2.0 * y[i]))


def test_norms_closed_forms_and_numpy(orc):
    A = np.array([[1.0, -2.0], [3.0, 4.0]])
    assert orc.norm1(A) == 6.0 and orc.normInf(A) == 7.0 and orc.normFrob(A) == np.sqrt(30.0)
    rng = np.random.default_rng(4)
    B = rng.uniform(-1, 1, (50, 70))
    assert abs(orc.norm1(B) - np.linalg.norm(B, 1)) <= 100 * EPS * np.linalg.norm(B, 1)
    assert abs(orc.normInf(B) - np.linalg.norm(B, np.inf)) <= 100 * EPS * np.linalg.norm(B, np.inf)
    assert abs(orc.normFrob(B) - np.linalg.norm(B)) <= 100 * EPS * np.linalg.norm(B)
    # a transposed operand would swap norm1 and normInf
    assert orc.norm1(B) != orc.normInf(B)


def test_matmulatap_exact_and_symmetric(orc):
    rng = np.random.default_rng(5)
    A = rng.integers(-9, 10, (70, 45)).astype(float)
    C = orc.matMulAtAP(A)
    np.testing.assert_array_equal(C, A.T @ A)
    R = rng.uniform(-1, 1, (100, 67))
    D = orc.matMulAtAP(R)
    np.testing.assert_array_equal(D, D.T)
    assert np.abs(D - R.T @ R).max() <= 110 * EPS * (np.abs(R).T @ np.abs(R)).max()


def test_det_closed_forms(orc):
    # the reference's lu takes the later row of a tie: perm [1, 0], U = [[-1, 3], [0, 5]]
    assert orc.det(np.array([[1.0, 2.0], [-1.0, 3.0]])) == 5.0
    rng = np.random.default_rng(6)
    for n in (1, 5, 9):
        P = np.eye(n)[rng.permutation(n)]
        assert orc.det(P) == round(np.linalg.det(P))
    U = np.triu(rng.uniform(1, 2, (8, 8)))
    prod = 1.0
    for i in range(8):
        prod = prod * U[i, i]
    assert orc.det(U) == prod
    H = 1.0 / (np.arange(4)[:, None] + np.arange(4)[None, :] + 1)
    assert abs(orc.det(H) - 1.0 / 6048000.0) <= 1e-9 / 6048000.0 * 1e3
    B = rng.uniform(-1, 1, (30, 30))
    assert abs(orc.det(B) - np.linalg.det(B)) <= 1e-11 * abs(np.linalg.det(B))
    assert orc.det(np.zeros((0, 0))) == 1.0


def test_inverse_iteration_known_spectrum(orc):
    n = 12
    D = np.diag(np.arange(1.0, n + 1))
    lam, q, it = orc.inverseIteration(D, 2.3, np.ones(n), 100, 1e-13)
    assert abs(lam - 2.0) <= 1e-12 and abs(abs(q[1]) - 1.0) <= 1e-12 and it < 100
    rng = np.random.default_rng(7)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    ev = np.linspace(-3, 5, n)
    S = (Q * ev) @ Q.T
    S = (S + S.T) / 2
    lam, q, it = orc.inverseIteration(S, 0.9, np.ones(n), 200, 1e-14)
    target = ev[np.argmin(np.abs(ev - 0.9))]
    assert abs(lam - target) <= 1e-10
    assert np.linalg.norm(S @ q - lam * q) <= 1e-6  # the Rayleigh quotient converges first
    lam0, q0, it0 = orc.inverseIteration(S, 0.9, np.ones(n), 0, 1e-14)
    assert lam0 == 0.0 and it0 == 0 and np.array_equal(q0, np.ones(n))


def test_matmulppar_partition(orc):
    # gridDims (BLAS/Level3.hs:276-294) slices of widths divisible by 4 leave
    # every element's arithmetic as matMulP's; exact integers agree anywhere
    rng = np.random.default_rng(8)
    A, B = rng.uniform(-1, 1, (256, 100)), rng.uniform(-1, 1, (100, 256))
    np.testing.assert_array_equal(orc.matMulPPar(A, B, 8), orc.matMulP(A, B))
    Ai, Bi = rng.integers(-9, 10, (301, 77)).astype(float), rng.integers(-9, 10, (77, 203)).astype(float)
    for t in (1, 3, 4, 6, 16):
        np.testing.assert_array_equal(orc.matMulPPar(Ai, Bi, t), Ai @ Bi)
