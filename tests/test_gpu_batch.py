"""Batched solves (lmx_batch_plan_*) through the C ABI: every problem of an
uneven-size batch against the oracle's luSolveP / choleskySolveP
(src/.../Solve/LU.hs:265-304, Solve/Cholesky.hs:241-273), its pivots against
the oracle's swap list, and the backward error eta <= 10 n eps
(scaledResidualLinear, test/Test/Residuals.hs:69-78)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
EPS = 2.220446049250313e-16
LU, CH = 0, 1

SIZES = [1, 2, 3, 7, 15, 16, 17, 31, 32, 33, 47, 63, 64, 65, 100, 127, 128, 129, 200, 255, 256, 257, 300,
         383, 384, 385, 450, 511, 512]


def spd(n, rng):
    # SPD test matrices as the workload defines them: G G^T / n + I
    G = rng.standard_normal((n, n))
    return G @ G.T / n + np.eye(n)


def make_batch(seed, sizes, kinds=(LU, CH), nrhs_max=8):
    rng = np.random.default_rng(seed)
    probs = []
    for i, n in enumerate(sizes):
        kind = kinds[i % len(kinds)]
        nrhs = int(rng.integers(1, nrhs_max + 1))
        A = spd(n, rng) if kind == CH else rng.uniform(-1, 1, (n, n))
        if kind == CH:
            # the upper triangle must never be read
            A = np.tril(A) + np.triu(np.full((n, n), np.nan), 1)
        B = rng.uniform(-1, 1, (n, nrhs))
        probs.append((kind, A, B))
    return probs


def full(kind, A):
    return np.tril(A) + np.tril(A, -1).T if kind == CH else A


def check_against_oracle(orc, probs, X, factor_tol=10.0):
    for (kind, A, B), x in zip(probs, X):
        n = A.shape[0]
        Af = full(kind, A)
        assert orc.scaled_residual_linear(Af, x, B) <= 10 * n * EPS, (kind, n)
        kappa = np.linalg.cond(Af, np.inf)
        for j in range(min(B.shape[1], 2)):
            xo = orc.luSolveP(Af, B[:, j]) if kind == LU else orc.choleskySolveP(Af, B[:, j])
            err = np.abs(x[:, j] - xo).max() / max(np.abs(xo).max(), 1e-300)
            assert err <= factor_tol * orc.condition_tolerance(kappa, n) + 1e-14, (kind, n, err)


@pytest.mark.parametrize("flags", [0, 2])
def test_batch_fp64_uneven_sizes(lmx_gpu, orc, flags):
    probs = make_batch(1, SIZES)
    X, info, iters = lmx_gpu.solve_batched(probs, lmx_gpu.PREC_F64, flags)
    assert (info == 0).all()
    check_against_oracle(orc, probs, X)


def test_batch_fp64_pivots_and_factors(lmx_gpu, orc):
    rng = np.random.default_rng(5)
    sizes = [5, 33, 64, 100, 257, 512]
    probs = [(LU, rng.uniform(-1, 1, (n, n)), rng.uniform(-1, 1, (n, 2))) for n in sizes]
    X, info, _, fac = lmx_gpu.solve_batched(probs, lmx_gpu.PREC_F64, lmx_gpu.BATCH_FACTORS, return_factors=True)
    assert (info == 0).all()
    for (kind, A, B), (LUf, ipiv) in zip(probs, fac):
        LUo, swo = orc.lu_p_factor(A)
        np.testing.assert_array_equal(ipiv, swo)
        assert np.abs(LUf - LUo).max() <= 1e-10 * max(1, A.shape[0] / 64)
        perm = lmx_gpu.swaps_to_perm(ipiv)
        assert orc.scaled_residual_lu(A, LUf, perm) < 10


def test_batch_pivot_rule_ties_and_nan(lmx_gpu, orc):
    # rawPivotSearch (Internal/Kernel.hs:1541-1556): first maximum wins, a NaN
    # below the diagonal is never chosen, a NaN on the diagonal keeps its row
    rng = np.random.default_rng(9)
    A1 = rng.uniform(-0.5, 0.5, (70, 70))
    A1[3, 0], A1[60, 0] = -1.0, 1.0
    A2 = rng.uniform(-1, 1, (70, 70))
    A2[5, 0], A2[9, 0] = np.nan, 2.0
    A3 = rng.uniform(-1, 1, (40, 40))
    A3[0, 0], A3[3, 0] = np.nan, 5.0
    probs = [(LU, A, np.ones((A.shape[0], 1))) for A in (A1, A2, A3)]
    _, _, _, fac = lmx_gpu.solve_batched(probs, lmx_gpu.PREC_F64, lmx_gpu.BATCH_FACTORS, return_factors=True)
    assert fac[0][1][0] == 3
    assert fac[1][1][0] == 9
    assert fac[2][1][0] == 0
    for (_, A, _), (_, ipiv) in zip(probs[:2], fac[:2]):
        assert ipiv[0] == orc.lu_p_factor(A)[1][0]


def test_batch_singular_and_not_spd_report_info(lmx_gpu, orc):
    rng = np.random.default_rng(3)
    good = rng.uniform(-1, 1, (50, 50))
    sing = rng.uniform(-1, 1, (80, 80))
    sing[:, 37] = 0.0
    notspd = spd(40, rng)
    notspd[20, 20] = -5.0
    B = [rng.uniform(-1, 1, (A.shape[0], 1)) for A in (good, sing, notspd)]
    probs = [(LU, good, B[0]), (LU, sing, B[1]), (CH, notspd, B[2])]
    X, info, _ = lmx_gpu.solve_batched(probs, lmx_gpu.PREC_F64)
    assert info[0] == 0
    assert info[1] == 38
    assert 1 <= info[2] <= 21
    np.testing.assert_array_equal(X[1], B[1])  # untouched
    assert orc.scaled_residual_linear(good, X[0], B[0]) <= 10 * 50 * EPS


@pytest.mark.parametrize("flags", [0, 2])
def test_batch_mixed_refines_to_fp64(lmx_gpu, orc, flags):
    probs = make_batch(2, SIZES)
    X, info, iters = lmx_gpu.solve_batched(probs, lmx_gpu.PREC_F32, flags)
    assert (info == 0).all()
    assert (iters >= 0).all() and (iters <= 30).all()  # converged without falling back
    check_against_oracle(orc, probs, X)


def test_batch_mixed_falls_back_on_hard_problem(lmx_gpu, orc):
    # kappa ~ 1e9 > 1/eps_f32: FP32 refinement cannot converge, the solve
    # falls back to the FP64 factorisation
    rng = np.random.default_rng(4)
    n = 120
    Q1, _ = np.linalg.qr(rng.standard_normal((n, n)))
    Q2, _ = np.linalg.qr(rng.standard_normal((n, n)))
    A = (Q1 * np.logspace(0, -9, n)) @ Q2.T
    easy = rng.uniform(-1, 1, (60, 60))
    B = [rng.uniform(-1, 1, (n, 3)), rng.uniform(-1, 1, (60, 1))]
    X, info, iters = lmx_gpu.solve_batched([(LU, A, B[0]), (LU, easy, B[1])], lmx_gpu.PREC_F32)
    assert (info == 0).all()
    assert iters[0] == -1 and iters[1] >= 0
    assert orc.scaled_residual_linear(A, X[0], B[0]) <= 10 * n * EPS
    assert orc.scaled_residual_linear(easy, X[1], B[1]) <= 10 * 60 * EPS


@pytest.mark.parametrize("prec", [0, 1])
def test_batch_result_is_the_same_alone_and_in_the_mix(lmx_gpu, prec):
    # a problem's arithmetic does not depend on which workgroup runs it or
    # on its neighbours: bitwise the same result alone and in a batch
    probs = make_batch(7, [300, 17, 512, 64, 129, 8, 250, 33])
    Xall, info, _ = lmx_gpu.solve_batched(probs, prec)
    assert (info == 0).all()
    for i in (0, 2, 5):
        Xi, _, _ = lmx_gpu.solve_batched([probs[i]], prec)
        np.testing.assert_array_equal(Xi[0], Xall[i])


def test_batch_graph_replays_and_autotune(lmx_gpu, orc):
    import torch
    probs = make_batch(11, [64, 200, 512, 31, 400, 100] * 3)
    shapes = [(k, A.shape[0], B.shape[1]) for k, A, B in probs]
    plist, na, nb, npv = lmx_gpu.batch_layout(shapes)
    a = torch.tensor(np.concatenate([A.ravel() for _, A, _ in probs]), device="cuda")
    b = torch.tensor(np.concatenate([B.ravel() for _, _, B in probs]), device="cuda")
    plan = lmx_gpu.BatchPlan(plist, na, nb, 0, lmx_gpu.PREC_F32, lmx_gpu.BATCH_GRAPH | lmx_gpu.BATCH_AUTOTUNE)
    try:
        q = plan.query()
        assert q["tuned_ms"][q["variant"]] == min(q["tuned_ms"]) > 0
        A1, B1 = a.clone(), b.clone()
        info = torch.zeros(len(probs), dtype=torch.int32, device="cuda")
        results = []
        for rep in range(3):
            A1.copy_(a)
            B1.copy_(b)
            plan.execute(A1, B1, info)
            torch.cuda.synchronize()
            assert (info.cpu().numpy() == 0).all()
            results.append(B1.cpu().numpy().copy())
            _ = (A1.sum() + B1.sum()).item()  # eager work between replays
        for r in results[1:]:
            np.testing.assert_array_equal(r, results[0])
    finally:
        plan.close()
    bh = results[0]
    X = [bh[p.b_offset:p.b_offset + p.n * p.nrhs].reshape(p.n, p.nrhs) for p in plist]
    check_against_oracle(orc, probs, X)


def test_batch_workload_sizes_backward_error(lmx_gpu):
    # 2000 problems of the batched workload's distribution (n log-uniform in
    # [8, 512], 1-8 right-hand sides, LU on U[-1,1] and Cholesky on
    # G G^T/n + I), FP32 factor + FP64 refinement; backward error of every
    # problem computed on the device in FP64
    import torch
    rng = np.random.default_rng(20000)
    cnt = 2000
    ns = np.exp(rng.uniform(np.log(8), np.log(512), cnt)).astype(int).clip(8, 512)
    kinds = rng.integers(0, 2, cnt)
    nrhs = rng.integers(1, 9, cnt)
    shapes = [(int(k), int(n), int(r)) for k, n, r in zip(kinds, ns, nrhs)]
    plist, na, nb, npv = lmx_gpu.batch_layout(shapes)
    g = torch.Generator(device="cuda").manual_seed(1)
    a = torch.empty(na, dtype=torch.float64, device="cuda")
    b = torch.rand(nb, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    for p in plist:
        n = p.n
        if p.kind == LU:
            a[p.a_offset:p.a_offset + n * n] = (torch.rand(n * n, dtype=torch.float64, device="cuda",
                                                           generator=g) * 2 - 1)
        else:
            G = torch.randn(n, n, dtype=torch.float64, device="cuda", generator=g)
            a[p.a_offset:p.a_offset + n * n] = (G @ G.T / n + torch.eye(n, dtype=torch.float64,
                                                                       device="cuda")).reshape(-1)
    a0, b0 = a.clone(), b.clone()
    info = torch.zeros(cnt, dtype=torch.int32, device="cuda")
    iters = torch.zeros(cnt, dtype=torch.int32, device="cuda")
    plan = lmx_gpu.BatchPlan(plist, na, nb, 0, lmx_gpu.PREC_F32, 0)
    try:
        plan.execute(a, b, info, None, iters)
        torch.cuda.synchronize()
    finally:
        plan.close()
    assert (info.cpu().numpy() == 0).all()
    assert (iters.cpu().numpy() >= 0).all()
    assert torch.equal(a, a0)  # refinement converged: A untouched
    worst = 0.0
    for p in plist:
        n, r = p.n, p.nrhs
        A = a0[p.a_offset:p.a_offset + n * n].view(n, n)
        X = b[p.b_offset:p.b_offset + n * r].view(n, r)
        B = b0[p.b_offset:p.b_offset + n * r].view(n, r)
        R = B - A @ X
        eta = (R.abs().max(0).values / (A.abs().sum(1).max() * X.abs().max(0).values + B.abs().max(0).values))
        worst = max(worst, eta.max().item() / (n * EPS))
    assert worst <= 10
