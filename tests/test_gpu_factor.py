"""Resident factors (lmx_factor_*) through the C ABI: a reused factor gives
the same bits as a fresh solve, FP64 handles match the oracle's luSolveP /
choleskySolveP within conditionTolerance, low-precision handles reach
eta <= 10 n eps, and several right-hand-side blocks reuse one factor."""
import numpy as np
import pytest
import scipy.linalg as sl

pytestmark = pytest.mark.gpu
EPS = 2.220446049250313e-16


def _spd(n, seed):
    G = np.random.default_rng(seed).standard_normal((n, n))
    return G @ G.T / n + np.eye(n)


@pytest.mark.parametrize("n", [1, 5, 64, 65, 300, 1500])
def test_lu_factor_f64_vs_oracle(lmx_gpu, orc, n):
    rng = np.random.default_rng(n)
    A = rng.uniform(-1, 1, (n, n))
    f = lmx_gpu.luFactor(A)
    assert f.info == 0
    kappa = np.linalg.cond(A, np.inf)
    for k in (1, 3, 16):
        B = rng.uniform(-1, 1, (n, k))
        X = lmx_gpu.luSolveWith(f, B)
        assert orc.scaled_residual_linear(A, X, B) <= 10 * n * EPS
        xo = orc.luSolveP(A, B[:, 0])
        assert np.abs(X[:, 0] - xo).max() / np.abs(xo).max() <= 10 * orc.condition_tolerance(kappa, n) + 1e-15
    f.close()


@pytest.mark.parametrize("n", [3, 100, 900])
def test_cholesky_factor_f64_vs_oracle(lmx_gpu, orc, n):
    A = _spd(n, n)
    f = lmx_gpu.choleskyFactor(A)
    assert f.info == 0
    b = np.random.default_rng(1).uniform(-1, 1, n)
    x = lmx_gpu.choleskySolveWith(f, b)
    assert orc.scaled_residual_linear(A, x, b) <= 10 * n * EPS
    xo = orc.choleskySolveP(A, b)
    assert np.abs(x - xo).max() / np.abs(xo).max() <= 10 * orc.condition_tolerance(np.linalg.cond(A), n)


@pytest.mark.parametrize("prec", [0, 1, 2], ids=["f64", "f32", "bf16"])
@pytest.mark.parametrize("kind", [0, 1], ids=["lu", "chol"])
def test_reused_factor_matches_fresh_solve(lmx_gpu, prec, kind):
    n = 700
    rng = np.random.default_rng(kind * 10 + prec)
    A = _spd(n, 3) if kind == 1 else rng.uniform(-1, 1, (n, n))
    B = rng.uniform(-1, 1, (n, 4))
    f = lmx_gpu.Factor(A, kind, prec)
    X1, r1 = f.solve(B)
    X2, r2 = f.solve(B)          # the same handle again
    np.testing.assert_array_equal(X1, X2)
    g = lmx_gpu.Factor(A, kind, prec)  # a fresh factor of the same matrix
    X3, _ = g.solve(B)
    np.testing.assert_array_equal(X1, X3)
    if prec != 0:
        ref = (lmx_gpu.posv_mixed if kind == 1 else lmx_gpu.gesv_mixed)(A, B, prec)[0]
        np.testing.assert_array_equal(X1, ref)  # the one-shot mixed solve, bit for bit
        assert r1.converged == 1
    eta = max(lmx_gpu.scaled_residual_linear(A, X1[:, c], B[:, c]) for c in range(4))
    assert eta <= 10 * n * EPS
    q = f.query()
    assert q["n"] == n and q["kind"] == kind and q["precision"] == prec and q["bytes"] > 0
    f.close()
    g.close()


def test_factor_dev_many_blocks(lmx_gpu):
    import torch
    n = 4096
    g = torch.Generator(device="cuda").manual_seed(2)
    A = torch.rand(n, n, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    f = lmx_gpu.Factor(A, lmx_gpu.FACTOR_LU, lmx_gpu.PREC_F32)
    for k in (1, 8, 64):
        B = torch.rand(n, k, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
        X = torch.zeros_like(B)
        ri = f.solve_dev(B, X)
        torch.cuda.synchronize()
        assert ri.converged == 1 and ri.fell_back == 0
        R = B - A @ X
        den = A.abs().sum(1).max() * X.abs().max(0).values + B.abs().max(0).values
        assert (R.abs().max(0).values / den).max().item() <= 10 * n * EPS
    f.close()


def test_singular_factor_reports_info(lmx_gpu):
    n = 50
    A = np.random.default_rng(5).uniform(-1, 1, (n, n))
    A[:, 7] = 0.0
    f = lmx_gpu.luFactor(A)
    assert f.info == 8
    f.close()
