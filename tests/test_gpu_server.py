"""The solve server (lmx_server_*) through the C ABI: concurrent requests of
uneven sizes, resident factors reused for new right-hand sides, HBM-budget
admission; every result against the FP64 oracle (backward error
eta <= 10 n eps, scaledResidualLinear test/Test/Residuals.hs:69-78) and the
same bits alone as in the mix."""
import threading

import numpy as np
import pytest
import scipy.linalg as sl

pytestmark = pytest.mark.gpu
EPS = 2.220446049250313e-16
LU, CH = 0, 1


def spd(n, rng):
    G = rng.standard_normal((n, n))
    return G @ G.T / n + np.eye(n)


def make_requests(seed, count, nmax=1500):
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(count):
        kind = CH if rng.random() < 0.7 else LU
        n = int(np.exp(rng.uniform(np.log(16), np.log(nmax))))
        nrhs = int(rng.integers(1, 33))
        A = spd(n, rng) if kind == CH else rng.uniform(-1, 1, (n, n))
        B = rng.uniform(-1, 1, (n, nrhs))
        reqs.append((kind, A, B))
    return reqs


def check(orc, kind, A, B, X):
    n = A.shape[0]
    assert orc.scaled_residual_linear(A, X, B) <= 10 * n * EPS
    X64 = sl.solve(A, B)
    kappa = np.linalg.cond(A, np.inf)
    err = np.abs(X - X64).max() / np.abs(X64).max()
    assert err <= 10 * orc.condition_tolerance(kappa, n) + 1e-13


@pytest.mark.parametrize("small_n", [0, 256])
def test_server_mix_matches_alone(lmx_gpu, orc, small_n):
    reqs = make_requests(1, 30)
    srv = lmx_gpu.Server(slots=4, small_n=small_n)
    try:
        Xs = [np.zeros_like(B) for _, _, B in reqs]
        ids = [srv.submit(k, B, X, A=A) for (k, A, B), X in zip(reqs, Xs)]
        res = [srv.wait(i) for i in ids]
        st = srv.stats()
    finally:
        srv.close()
    assert all(r.rc == 0 and r.status == 1 for r in res)
    assert st["completed"] == len(reqs) and st["failed"] == 0
    assert 0 < st["busy_ms"] <= st["window_ms"]
    for (k, A, B), X in zip(reqs, Xs):
        check(orc, k, A, B, X)
    # each request alone, on a fresh one-slot server: the same bits
    alone = lmx_gpu.Server(slots=1, small_n=small_n)
    try:
        for (k, A, B), X in list(zip(reqs, Xs))[::3]:
            Y = np.zeros_like(B)
            r = alone.wait(alone.submit(k, B, Y, A=A))
            assert r.rc == 0
            np.testing.assert_array_equal(Y, X)
    finally:
        alone.close()


@pytest.mark.parametrize("n", [100, 700])
def test_server_resident_factor_reuse(lmx_gpu, orc, n):
    rng = np.random.default_rng(n)
    A = spd(n, rng)
    L = rng.uniform(-1, 1, (n, n))
    srv = lmx_gpu.Server(slots=2, small_n=256)
    try:
        B0 = rng.uniform(-1, 1, (n, 4))
        X0 = np.zeros_like(B0)
        r0 = srv.wait(srv.submit(CH, B0, X0, A=A, keep=True))
        assert r0.rc == 0 and r0.factor > 0
        C0 = rng.uniform(-1, 1, (n, 2))
        Y0 = np.zeros_like(C0)
        q0 = srv.wait(srv.submit(LU, C0, Y0, A=L, keep=True))
        assert q0.rc == 0 and q0.factor > 0
        check(orc, CH, A, B0, X0)
        st = srv.stats()
        assert st["resident_factors"] == 2
        ids, outs = [], []
        for j in range(5):
            B = rng.uniform(-1, 1, (n, 1 + 3 * j))
            X = np.zeros_like(B)
            ids.append(srv.submit(CH, B, X, factor=r0.factor))
            outs.append((CH, A, B, X))
            C = rng.uniform(-1, 1, (n, 2 + j))
            Y = np.zeros_like(C)
            ids.append(srv.submit(LU, C, Y, factor=q0.factor))
            outs.append((LU, L, C, Y))
        for i in ids:
            assert srv.wait(i).rc == 0
        for k, M, B, X in outs:
            check(orc, k, M, B, X)
        # a reused factor gives the bits of a fresh solve with the same matrix
        B = outs[2][2]
        Xf = np.zeros_like(B)
        assert srv.wait(srv.submit(CH, B, Xf, A=A)).rc == 0
        np.testing.assert_array_equal(Xf, outs[2][3])
        srv.release(r0.factor)
        srv.release(q0.factor)
        assert srv.stats()["resident_factors"] == 0
        with pytest.raises(lmx_gpu.LmxError):
            srv.submit(CH, B, Xf, factor=r0.factor)
    finally:
        srv.close()


def test_server_budget_admission(lmx_gpu, orc):
    # each n = 600 request needs ~10 MB of HBM: a 40 MB budget admits a few
    # at a time, the rest wait; a request that can never fit is refused
    budget = 40 << 20
    srv = lmx_gpu.Server(budget_bytes=budget, slots=8)
    rng = np.random.default_rng(2)
    try:
        reqs = [(LU, rng.uniform(-1, 1, (600, 600)), rng.uniform(-1, 1, (600, 3))) for _ in range(10)]
        Xs = [np.zeros_like(B) for _, _, B in reqs]
        ids = [srv.submit(k, B, X, A=A) for (k, A, B), X in zip(reqs, Xs)]
        res = [srv.wait(i) for i in ids]
        st = srv.stats()
        big = rng.uniform(-1, 1, (3000, 3000))
        with pytest.raises(lmx_gpu.LmxError) as e:
            srv.submit(LU, np.ones((3000, 1)), np.zeros((3000, 1)), A=big)
        assert e.value.code == lmx_gpu.LMX_ERR_BUDGET
    finally:
        srv.close()
    assert all(r.rc == 0 for r in res)
    assert 0 < st["peak_bytes"] <= budget
    assert max(r.queue_ms for r in res) > 0
    for (k, A, B), X in zip(reqs, Xs):
        check(orc, k, A, B, X)


def test_server_concurrent_clients(lmx_gpu, orc):
    # requests submitted from several threads while others run
    reqs = make_requests(9, 24, nmax=900)
    srv = lmx_gpu.Server(slots=3, small_n=128)
    Xs = [np.zeros_like(B) for _, _, B in reqs]
    errors = []

    def client(part):
        try:
            for i in part:
                k, A, B = reqs[i]
                r = srv.wait(srv.submit(k, B, Xs[i], A=A))
                assert r.rc == 0
        except Exception as ex:  # surfaced below
            errors.append(ex)

    ths = [threading.Thread(target=client, args=(list(range(t, len(reqs), 4)),)) for t in range(4)]
    try:
        for t in ths:
            t.start()
        for t in ths:
            t.join()
    finally:
        srv.close()
    assert not errors
    for (k, A, B), X in zip(reqs, Xs):
        check(orc, k, A, B, X)


def test_server_concurrent_large_lu(lmx_gpu, orc):
    # several n >= 4096 LU factorisations at once (8-workgroup panels on
    # concurrent streams, beside a Cholesky request): all complete, and each
    # gives the same bits as alone
    rng = np.random.default_rng(44)
    reqs = [(LU, rng.uniform(-1, 1, (n, n)), rng.uniform(-1, 1, (n, 2))) for n in (4096, 4100, 4500, 5000)]
    reqs.append((CH, spd(4096, rng), rng.uniform(-1, 1, (4096, 1))))
    srv = lmx_gpu.Server(slots=5)
    try:
        Xs = [np.zeros_like(B) for _, _, B in reqs]
        ids = [srv.submit(k, B, X, A=A) for (k, A, B), X in zip(reqs, Xs)]
        res = [srv.wait(i) for i in ids]
    finally:
        srv.close()
    assert all(r.rc == 0 for r in res), [r.rc for r in res]
    for (k, A, B), X in zip(reqs, Xs):
        assert orc.scaled_residual_linear(A, X, B) <= 10 * A.shape[0] * EPS
    alone = lmx_gpu.Server(slots=1)
    try:
        k, A, B = reqs[0]
        Y = np.zeros_like(B)
        assert alone.wait(alone.submit(k, B, Y, A=A)).rc == 0
        np.testing.assert_array_equal(Y, Xs[0])
    finally:
        alone.close()


def test_server_second_waiter_is_refused(lmx_gpu):
    # only one thread may wait for a request; a second one gets an error
    # instead of touching the request after the first has collected it
    rng = np.random.default_rng(5)
    n = 2000
    A, B = rng.uniform(-1, 1, (n, n)), rng.uniform(-1, 1, (n, 1))
    X = np.zeros_like(B)
    srv = lmx_gpu.Server(slots=1)
    try:
        rid = srv.submit(LU, B, X, A=A)
        out = {}

        def waiter(tag):
            try:
                out[tag] = srv.wait(rid).rc
            except lmx_gpu.LmxError as e:
                out[tag] = e.code

        t1 = threading.Thread(target=waiter, args=("a",))
        t1.start()
        import time
        time.sleep(0.05)
        t2 = threading.Thread(target=waiter, args=("b",))
        t2.start()
        t1.join()
        t2.join()
    finally:
        srv.close()
    assert sorted(out.values()) == [lmx_gpu.LMX_ERR_ARG, 0]
