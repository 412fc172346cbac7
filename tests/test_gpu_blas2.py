"""GEMV, norms, the residual and its normwise backward error, SYRK
(matMulAtAP), det and inverse iteration through the C ABI, against the
oracle's restatements of the reference (oracle/oracle.py)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
EPS = 2.220446049250313e-16


@pytest.mark.parametrize("m,n", [(1, 1), (7, 5), (64, 64), (300, 1000), (1000, 257), (4097, 4100)])
def test_matvec_vs_oracle(lmx_gpu, orc, m, n):
    rng = np.random.default_rng(m + n)
    A, x = rng.uniform(-1, 1, (m, n)), rng.uniform(-1, 1, n)
    y = lmx_gpu.matvec(A, x)
    yo = orc.matvecP(A, x)
    assert np.all(np.abs(y - yo) <= 2 * n * EPS * (np.abs(A) @ np.abs(x)))
    Ai, xi = rng.integers(-60, 61, (m, n)).astype(float), rng.integers(-60, 61, n).astype(float)
    np.testing.assert_array_equal(lmx_gpu.matvec(Ai, xi), orc.matvecP(Ai, xi))


def test_gemv_alpha_beta_vs_oracle(lmx_gpu, orc):
    rng = np.random.default_rng(9)
    m, n = 333, 222
    A, x, y = rng.uniform(-1, 1, (m, n)), rng.uniform(-1, 1, n), rng.uniform(-1, 1, m)
    out = lmx_gpu.gemv(0.75, A, x, -1.5, y)
    ref = orc.gemv(0.75, A, x, -1.5, y)
    assert np.all(np.abs(out - ref) <= 4 * n * EPS * (np.abs(A) @ np.abs(x) + np.abs(y)))
    # beta = 0 does not read y (BLAS); the reference would form 0 * y
    ynan = np.full(m, np.nan)
    assert np.all(np.isfinite(lmx_gpu.gemv(1.0, A, x, 0.0, ynan)))


def test_gemv_dev(lmx_gpu, orc):
    import torch
    rng = np.random.default_rng(10)
    m, n = 2000, 1500
    A, x = rng.uniform(-1, 1, (m, n)), rng.uniform(-1, 1, n)
    dA, dx = torch.tensor(A, device="cuda"), torch.tensor(x, device="cuda")
    dy = torch.zeros(m, dtype=torch.float64, device="cuda")
    lmx_gpu.gemv_dev(dA, dx, dy)
    y = dy.cpu().numpy()
    assert np.all(np.abs(y - orc.matvecP(A, x)) <= 2 * n * EPS * (np.abs(A) @ np.abs(x)))


@pytest.mark.parametrize("m,n", [(1, 1), (3, 70), (70, 3), (513, 1029), (4096, 4096)])
def test_norms_vs_oracle(lmx_gpu, orc, m, n):
    rng = np.random.default_rng(m * n)
    A = rng.uniform(-1, 1, (m, n))
    for mine, ref in ((lmx_gpu.norm1, orc.norm1), (lmx_gpu.normInf, orc.normInf), (lmx_gpu.normFrob, orc.normFrob)):
        r = ref(A)
        assert abs(mine(A) - r) <= 2 * max(m, n) * EPS * r
    assert lmx_gpu.norm(A, "M") == np.abs(A).max()
    Ai = rng.integers(-100, 101, (m, n)).astype(float)   # exact integer sums
    assert lmx_gpu.norm1(Ai) == orc.norm1(Ai) and lmx_gpu.normInf(Ai) == orc.normInf(Ai)


def test_vector_norms(lmx_gpu, orc):
    x = np.random.default_rng(1).uniform(-1, 1, 777)
    assert abs(lmx_gpu.vnorm2(x) - np.linalg.norm(x)) <= 10 * EPS * np.linalg.norm(x)
    assert abs(lmx_gpu.vnorm1(x) - np.abs(x).sum()) <= 800 * EPS * np.abs(x).sum()
    assert lmx_gpu.vnormInf(x) == np.abs(x).max()
    assert lmx_gpu.normInf(np.array([[1.0, -2.0], [3.0, 4.0]])) == 7.0
    assert lmx_gpu.norm1(np.array([[1.0, -2.0], [3.0, 4.0]])) == 6.0


@pytest.mark.parametrize("n,k", [(5, 1), (300, 1), (300, 3), (1000, 16), (1000, 17), (2049, 40)])
def test_scaled_residual_vs_oracle(lmx_gpu, orc, n, k):
    rng = np.random.default_rng(n + k)
    A, X, B = rng.uniform(-1, 1, (n, n)), rng.uniform(-1, 1, (n, k)), rng.uniform(-1, 1, (n, k))
    eta = lmx_gpu.scaled_residual_linear(A, X, B)
    for c in range(k):
        ref = orc.scaled_residual_linear(A, X[:, c], B[:, c])
        assert abs(eta[c] - ref) <= 4 * n * EPS * ref
    # a solved system: the backward error of a GPU solve is tiny
    x = lmx_gpu.luSolveP(A, B[:, 0])
    e = lmx_gpu.scaled_residual_linear(A, x, B[:, 0])
    assert e <= 10 * n * EPS and abs(e - orc.scaled_residual_linear(A, x, B[:, 0])) <= 0.5 * e + 1e-300


def test_residual_and_scaled_residual_dev(lmx_gpu):
    import torch
    n, k = 3000, 5
    g = torch.Generator(device="cuda").manual_seed(4)
    A = torch.rand(n, n, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    X = torch.rand(n, k, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    B = torch.rand(n, k, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    R = torch.empty_like(B)
    lmx_gpu.residual_dev(A, X, B, R)
    ref = B - A @ X
    bound = 2 * n * EPS * (A.abs() @ X.abs() + B.abs())
    assert bool(((R - ref).abs() <= bound).all())
    eta = torch.zeros(k, dtype=torch.float64, device="cuda")
    lmx_gpu.scaled_residual_dev(A, X, B, eta)
    den = A.abs().sum(1).max() * X.abs().max(0).values + B.abs().max(0).values + n * EPS
    want = (ref.abs().max(0).values / den)
    assert bool(((eta - want).abs() <= 1e-12 * want).all())


@pytest.mark.parametrize("m,n", [(10, 7), (200, 64), (1000, 700), (3000, 2100)])
def test_matmulatap_vs_oracle(lmx_gpu, orc, m, n):
    rng = np.random.default_rng(m + 3 * n)
    A = rng.uniform(-1, 1, (m, n))
    C = lmx_gpu.matMulAtAP(A)
    Co = orc.matMulAtAP(A)
    np.testing.assert_array_equal(C, C.T)
    assert np.all(np.abs(C - Co) <= 2 * m * EPS * (np.abs(A).T @ np.abs(A)))
    Ai = rng.integers(-20, 21, (m, n)).astype(float)
    np.testing.assert_array_equal(lmx_gpu.matMulAtAP(Ai), Ai.T @ Ai)


def test_syrk_alpha_beta_and_dev(lmx_gpu):
    import torch
    rng = np.random.default_rng(12)
    n, k = 500, 300
    A = rng.uniform(-1, 1, (n, k))
    C0 = rng.uniform(-1, 1, (n, n))
    C0 = C0 + C0.T
    C = lmx_gpu.syrk(A, "N", alpha=-0.5, beta=2.0, c=C0)
    ref = -0.5 * (A @ A.T) + 2.0 * C0
    assert np.abs(C - ref).max() <= 4 * k * EPS * (0.5 * np.abs(A) @ np.abs(A).T + 2 * np.abs(C0)).max()
    dA = torch.tensor(A, device="cuda")
    dC = torch.zeros(k, k, dtype=torch.float64, device="cuda")
    lmx_gpu.syrk_dev(dA, dC, trans="T")
    assert np.abs(dC.cpu().numpy() - A.T @ A).max() <= 4 * n * EPS * (np.abs(A).T @ np.abs(A)).max()


def test_det_closed_forms(lmx_gpu, orc, tie_blocks):
    assert lmx_gpu.det(np.array([[1.0, 2.0], [-1.0, 3.0]])) == 5.0
    rng = np.random.default_rng(13)
    for n in (1, 6, 100, 1100):
        P = np.eye(n)[rng.permutation(n)]
        assert lmx_gpu.det(P) == orc.det(P) == round(np.linalg.det(P))
    U = np.triu(rng.uniform(0.9, 1.1, (700, 700)))
    assert lmx_gpu.det(U) == orc.det(U)
    # exact integer factors: the same bits as the oracle
    for n in (40, 999):
        T = tie_blocks(n, n)
        assert lmx_gpu.det(T) == orc.det(T)


@pytest.mark.parametrize("n", [10, 200, 1500])
def test_det_vs_oracle(lmx_gpu, orc, n):
    rng = np.random.default_rng(n)
    A = rng.uniform(-1, 1, (n, n)) / np.sqrt(n) * 2.0
    d, do = lmx_gpu.det(A), orc.det(A)
    assert np.sign(d) == np.sign(do)
    assert abs(d - do) <= 1e-11 * n * abs(do)


def test_det_dev(lmx_gpu, orc):
    import torch
    rng = np.random.default_rng(21)
    A = rng.uniform(-1, 1, (300, 300)) / 8
    out = torch.zeros(1, dtype=torch.float64, device="cuda")
    dA = torch.tensor(A, device="cuda")
    lmx_gpu.det_dev(dA, out)
    assert abs(out.item() - orc.det(A)) <= 1e-10 * abs(orc.det(A))
    np.testing.assert_array_equal(dA.cpu().numpy(), A)  # A is not modified


@pytest.mark.parametrize("n,mu", [(12, 0.9), (300, 1.7), (2000, -0.35)])
def test_inverse_iteration_vs_oracle(lmx_gpu, orc, n, mu):
    rng = np.random.default_rng(n)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    ev = np.linspace(-3, 5, n)
    S = (Q * ev) @ Q.T
    S = (S + S.T) / 2
    q0 = np.ones(n)
    lam, q, it = lmx_gpu.inverseIteration(S, mu, q0, 200, 1e-13)
    lo, qo, ito = orc.inverseIteration(S, mu, q0, 200, 1e-13)
    target = ev[np.argmin(np.abs(ev - mu))]
    assert abs(lam - target) <= 1e-9 and abs(lo - target) <= 1e-9
    assert abs(it - ito) <= 1
    assert min(np.abs(q - qo).max(), np.abs(q + qo).max()) <= 1e-6
    assert np.linalg.norm(S @ q - lam * q) <= 1e-5
    # max_iter = 0 returns (0, q0)
    lam0, q0r, it0 = lmx_gpu.inverseIteration(S, mu, q0, 0, 1e-13)
    assert lam0 == 0.0 and it0 == 0 and np.array_equal(q0r, q0)
