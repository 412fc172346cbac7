"""GPU Cholesky and triangular solves through the C ABI against LAPACK
golden factors, the oracle's choleskySolveP / cholesky / substitutions,
closed forms and the backward-error bound."""
import numpy as np
import pytest
import scipy.linalg as sl

pytestmark = pytest.mark.gpu
EPS = 2.220446049250313e-16


def spd(n, seed):
    rng = np.random.default_rng(seed)
    G = rng.standard_normal((n, n))
    return G @ G.T / n + np.eye(n)


@pytest.mark.parametrize("n", [1, 3, 10, 33, 63, 64, 65, 100, 130])
def test_potrf_vs_lapack_golden(lmx_gpu, orc, golden, n):
    S = golden[f"S_{n}"]
    G, info = lmx_gpu.potrf(S)
    assert info == 0
    np.testing.assert_allclose(np.tril(G), golden[f"lapack_chol_{n}"], rtol=0, atol=1e-13)
    # strict upper triangle untouched
    np.testing.assert_array_equal(np.triu(G, 1), np.triu(S, 1))
    np.testing.assert_allclose(lmx_gpu.cholesky(S), np.tril(orc.cholesky_p_factor(S)), rtol=0, atol=1e-13)


@pytest.mark.parametrize("n", [1, 2, 64, 65, 127, 128, 129, 300, 1000])
@pytest.mark.parametrize("nrhs", [1, 16])
def test_posv_backward_error_and_oracle(lmx_gpu, orc, n, nrhs):
    S = spd(n, n + nrhs)
    B = np.random.default_rng(n).uniform(-1, 1, (n, nrhs))
    X, G, info = lmx_gpu.posv(S, B)
    assert info == 0
    assert orc.scaled_residual_linear(S, X, B) <= 10 * n * EPS
    kappa = np.linalg.cond(S, np.inf)
    xo = orc.choleskySolveP(S, B[:, 0])
    assert np.abs(X[:, 0] - xo).max() / np.abs(xo).max() <= 10 * orc.condition_tolerance(kappa, n)


@pytest.mark.parametrize("n", [10, 100, 300])
def test_tridiagonal_closed_form(lmx_gpu, n):
    T = 2 * np.eye(n) - np.eye(n, k=1) - np.eye(n, k=-1)
    i = np.arange(n)
    want = np.diag(np.sqrt((i + 2) / (i + 1))) + np.diag(-np.sqrt((i[:-1] + 1) / (i[:-1] + 2)), -1)
    np.testing.assert_allclose(lmx_gpu.cholesky(T), want, rtol=0, atol=8 * EPS)


@pytest.mark.parametrize("n,bad", [(5, 2), (100, 70), (400, 130)])
def test_not_positive_definite_info(lmx_gpu, n, bad):
    S = spd(n, 1)
    S[bad, bad] = -50.0
    G, info = lmx_gpu.potrf(S)
    assert info == bad + 1


def test_reads_lower_triangle_only(lmx_gpu):
    n = 150
    S = spd(n, 4)
    junk = S + np.triu(np.random.default_rng(4).uniform(-5, 5, (n, n)), 1)
    b = np.ones(n)
    np.testing.assert_array_equal(lmx_gpu.choleskySolve(junk, b), lmx_gpu.choleskySolve(S, b))


def test_potrs_reuses_factor(lmx_gpu, orc):
    n = 257
    S = spd(n, 9)
    G, info = lmx_gpu.potrf(S)
    B = np.random.default_rng(2).uniform(-1, 1, (n, 33))
    X = lmx_gpu.potrs(G, B)
    assert orc.scaled_residual_linear(S, X, B) <= 10 * n * EPS


@pytest.mark.parametrize("n", [4096, 8192])
def test_device_posv_large(lmx_gpu, n):
    import torch
    k = 16
    g = torch.Generator(device="cuda").manual_seed(n)
    G = torch.randn(n, n, dtype=torch.float64, device="cuda", generator=g)
    S = G @ G.T / n + torch.eye(n, dtype=torch.float64, device="cuda")
    B = torch.rand(n, k, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    L = S.clone()
    X = B.clone()
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    lmx_gpu.posv_dev(L, X, info)
    torch.cuda.synchronize()
    assert info.item() == 0
    R = B - S @ X
    num = R.abs().max(dim=0).values
    den = S.abs().sum(1).max() * X.abs().max(dim=0).values + B.abs().max(dim=0).values
    assert (num / den).max().item() <= 10 * n * EPS


# ---------------------------------------------------------------- triangular

@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("diag", ["N", "U"])
@pytest.mark.parametrize("n,nrhs", [(1, 1), (7, 3), (64, 1), (65, 2), (200, 70), (300, 1), (1100, 3), (2100, 9)])
def test_trsm_vs_scipy(lmx_gpu, uplo, trans, diag, n, nrhs):
    rng = np.random.default_rng(n * 7 + nrhs)
    T = rng.uniform(-1, 1, (n, n)) / np.sqrt(n) + np.diag(rng.uniform(1, 2, n))
    B = rng.uniform(-1, 1, (n, nrhs))
    X = lmx_gpu.trsm(T, B, uplo, trans, diag)
    want = sl.solve_triangular(T, B, lower=(uplo == "L"), trans=(1 if trans == "T" else 0),
                               unit_diagonal=(diag == "U"))
    np.testing.assert_allclose(X, want, rtol=1e-11, atol=1e-11)


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("n,nrhs", [(10000, 1), (13500, 3)])
def test_trsm_many_rows_per_wave(lmx_gpu, uplo, n, nrhs):
    # the row updates after each 1024-row block run 2 (n = 10000) and 3
    # (n = 13500) rows per wave; 13500 also ends in a ragged 188-row block
    rng = np.random.default_rng(n + nrhs)
    T = rng.uniform(-1, 1, (n, n)) / n
    T[np.diag_indices(n)] = rng.uniform(1, 2, n)
    B = rng.uniform(-1, 1, (n, nrhs))
    X = lmx_gpu.trsm(T, B, uplo, "N", "N")
    want = sl.solve_triangular(T, B, lower=(uplo == "L"))
    np.testing.assert_allclose(X, want, rtol=1e-11, atol=1e-11)


def test_reference_substitutions_vs_oracle(lmx_gpu, orc):
    rng = np.random.default_rng(5)
    n = 90
    T = rng.uniform(-1, 1, (n, n)) / np.sqrt(n) + np.diag(rng.uniform(1, 10, n))
    L, U = np.tril(T), np.triu(T)
    b = rng.uniform(-1, 1, n)
    for name, M in [("forwardSub", L), ("backSub", U), ("forwardSubUnit", L), ("backSubUnit", U)]:
        x = getattr(lmx_gpu, name)(M, b)
        xo = getattr(orc, name)(M, b)
        np.testing.assert_allclose(x, xo, rtol=1e-11, atol=1e-12)


def test_triangular_integer_solutions_exact(lmx_gpu):
    rng = np.random.default_rng(6)
    n = 70
    L = np.tril(rng.integers(-2, 3, (n, n)), -1).astype(float) * 0.5 ** 4 + np.eye(n)
    x = rng.integers(-5, 6, n).astype(float)
    np.testing.assert_allclose(lmx_gpu.forwardSubUnit(L, L @ x), x, rtol=0, atol=1e-9)
