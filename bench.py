#!/usr/bin/env python3
"""Benchmark of the dense-solve path on one MI355X.

Default workload (the configuration the metric is quoted on): luSolve of
A x = b, n = 16384, 1 right-hand side, A = U diag(sigma) V^T with Haar U, V
and geometric sigma (kappa = 1e6), factored in FP32 on MFMA and refined to
FP64 accuracy on the device.  One step = one FP64-accurate solve (factor +
refinement) with A, b already resident in HBM.  The same inputs are also
solved by the plain FP64 path (FP64 factor + solve) for the "vs plain FP64
GPU" part of the metric.

Prints ONE JSON line.  Other workloads (--workload): lu_fp64 / chol_fp64
(n = 8192, 16 RHS), gemm_fp64 (8192^3), chol_mixed, and the batched
workloads batch_mixed / batch_fp64: 20,000 LU (A iid U[-1,1]) plus 20,000
Cholesky (A = G G^T/n + I) independent solves, n log-uniform in [8, 512],
1-8 right-hand sides each, in one persistent launch (one step = the whole
batch); and serve: Poisson arrivals of solve requests (70% Cholesky, 30%
LU, n log-uniform in [16, 4096], 1-32 RHS, 30% reusing a resident factor)
through the solve server, FP32 factor + FP64 refinement, at rates swept
up to ~80% device utilisation; reports p50/p99 latency and throughput.
"""
import argparse
import json
import os
import platform
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "linear-massiv_amd"))

EPS = 2.220446049250313e-16
# TFLOP/s, MI355X dense MFMA (bgemm / bgemm3: the BF16 tiles, 2.5 PFLOP/s)
PEAK = {"gemm": 78.6, "sgemm": 157.3, "syrk": 78.6, "ssyrk": 157.3, "bgemm": 2500.0, "bgemm3": 2500.0}
KERNEL_FAMILIES = ["gemm", "sgemm", "gemm64", "sgemm64", "syrk", "ssyrk", "bgemm", "bgemm3", "getrf_panel", "trsm",
                   "laswp", "potrf_diag"]
# HBM bytes per launch of a workload's dominant kernel, from the rocprofv3
# FETCH_SIZE / WRITE_SIZE passes summarised in profiles/
# (workload, n, kernel family) -> file
TRAFFIC_PROFILES = {("lu_mixed", 16384, "sgemm"): "sgemm_lu_mixed_n16384_traffic.md",
                    ("chol_fp64", 8192, "syrk"): "dsyrk_chol_fp64_n8192_traffic.md",
                    ("lu_fp64", 8192, "gemm"): "dgemm_lu_fp64_n8192_traffic.md",
                    ("chol_mixed", 16384, "ssyrk"): "ssyrk_chol_mixed_n16384_traffic.md"}


def profiled_traffic(workload, n, family):
    name = TRAFFIC_PROFILES.get((workload, n, family))
    if not name:
        return None
    try:
        for line in open(os.path.join(ROOT, "profiles", name)):
            if line.startswith("traffic_bytes_per_launch:"):
                return int(line.split(":", 1)[1])
    except OSError:
        return None
    return None


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--workload", default="lu_mixed",
                   choices=["lu_mixed", "chol_mixed", "lu_fp64", "chol_fp64", "gemm_fp64", "gemm_bf16",
                            "gemm_bf16x3", "batch_mixed", "batch_fp64", "serve"])
    p.add_argument("--factor", default="f32", choices=["f32", "bf16"],
                   help="mixed workloads: the factor precision (bf16: BF16 tiles with split operands, GMRES-IR)")
    p.add_argument("--requests", type=int, default=400, help="serve: requests per rate point")
    p.add_argument("--count", type=int, default=20000, help="problems per family (batched workloads)")
    p.add_argument("--no-autotune", action="store_true",
                   help="batched workloads: keep the default kernel variant instead of timing all four")
    p.add_argument("--n", type=int, default=None)
    p.add_argument("--nrhs", type=int, default=None)
    p.add_argument("--kappa", type=float, default=1e6)
    p.add_argument("--seed", type=int, default=2026)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--no-fp64-compare", action="store_true")
    return p.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# ------------------------------------------------------------------ inputs

def haar(torch, n, g):
    G = torch.randn(n, n, dtype=torch.float64, device="cuda", generator=g)
    Q, R = torch.linalg.qr(G)
    del G
    Q *= torch.sign(torch.diagonal(R))[None, :]
    return Q


def randsvd(torch, n, kappa, g, spd):
    U = haar(torch, n, g)
    V = U if spd else haar(torch, n, g)
    sigma = kappa ** (-torch.arange(n, dtype=torch.float64, device="cuda") / max(n - 1, 1))
    A = (U * sigma[None, :]) @ V.T
    if spd:
        A = (A + A.T) * 0.5
    return A.contiguous()


def spd_gram(torch, n, g):
    G = torch.randn(n, n, dtype=torch.float64, device="cuda", generator=g)
    return (G @ G.T / n + torch.eye(n, dtype=torch.float64, device="cuda")).contiguous()


def uniform(torch, shape, g):
    return (torch.rand(*shape, dtype=torch.float64, device="cuda", generator=g) * 2 - 1).contiguous()


def backward_error(torch, A, X, B):
    R = B - A @ X
    num = R.abs().max(dim=0).values
    den = A.abs().sum(1).max() * X.abs().max(dim=0).values + B.abs().max(dim=0).values
    return (num / den).max().item()


# ------------------------------------------------------------------ CPU baseline

def cpu_baseline(workload, n, nrhs, batch=None):
    """The oracle (C restatement of the reference algorithm, one thread) on a
    bounded sample of the workload, scaled to the metric's unit."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    oracle.lib()
    rng = np.random.default_rng(7)
    if batch is not None:
        # a random sample of the batch's problems; the reference API solves
        # one right-hand side per call (and factors on every call)
        shapes, total_flops = batch
        idx = rng.choice(len(shapes), size=min(300, len(shapes)), replace=False)
        t_used, f_used = 0.0, 0.0
        for i in idx:
            kind, m, r = shapes[i]
            if kind == 0:
                A = rng.uniform(-1, 1, (m, m))
            else:
                G = rng.standard_normal((m, m))
                A = G @ G.T / m + np.eye(m)
            b = rng.uniform(-1, 1, m)
            t = time.perf_counter()
            for _ in range(r):
                (oracle.choleskySolveP if kind else oracle.luSolveP)(A, b)
            t_used += time.perf_counter() - t
            f_used += (m ** 3 / 3.0 if kind else 2.0 * m ** 3 / 3.0) + 2.0 * m * m * r
            if t_used > 20:
                break
        t_full = t_used * total_flops / f_used
        return {"value": len(shapes) / t_full, "unit": "solves/s", "cores": 1, "kind": "port",
                "sample": f"oracle luSolveP / choleskySolveP on {len(idx)} random problems of the batch, one call "
                          f"per RHS, 1 thread ({t_used:.1f} s), scaled by flops to the {len(shapes)} problems"}
    if workload.startswith("gemm"):
        # matMulPPar (BLAS/Level3.hs:213-294, gridDims 2D partition) on every
        # host core the box gives this job, and matMulP on one for reference
        cores = max(1, min(16, os.cpu_count() or 1))
        ns = 4096
        A = rng.uniform(-1, 1, (ns, ns))
        B = rng.uniform(-1, 1, (ns, ns))
        times = []
        for _ in range(3):
            t = time.perf_counter()
            oracle.matMulPPar(A, B, cores)
            times.append(time.perf_counter() - t)
        dt = sorted(times)[1]
        flops = 2.0 * ns ** 3
        t = time.perf_counter()
        oracle.matMulP(A[:2048, :2048].copy(), B[:2048, :2048].copy())
        dt1 = time.perf_counter() - t
        return {"value": flops / dt / 1e12, "unit": "TFLOP/s", "cores": cores, "kind": "port",
                "one_thread_tflops": 2.0 * 2048 ** 3 / dt1 / 1e12,
                "sample": f"oracle matMulPPar {ns}x{ns}x{ns} on {cores} threads, median of 3 runs "
                          f"({dt:.2f} s); matMulP 2048^3 on 1 thread ({dt1:.1f} s) beside it"}
    ns = 4096 if n >= 4096 else n
    A = rng.uniform(-1, 1, (ns, ns))
    b = rng.uniform(-1, 1, ns)
    spd = workload.startswith("chol")
    if spd:
        G = rng.standard_normal((ns, ns))
        A = G @ G.T / ns + np.eye(ns)
    reps = 3
    times = []
    for _ in range(reps):
        t = time.perf_counter()
        (oracle.choleskySolveP if spd else oracle.luSolveP)(A, b)
        times.append(time.perf_counter() - t)
    dt = sorted(times)[reps // 2]
    # The reference API takes one right-hand side per call and re-factors on
    # every call, so nrhs right-hand sides cost nrhs calls.
    calls = max(nrhs, 1)
    t_full = dt * (n / ns) ** 3 * calls
    name = "choleskySolveP" if spd else "luSolveP"
    return {"value": 1.0 / t_full, "unit": "solves/s", "cores": 1, "kind": "port",
            "sample": f"oracle {name} n={ns}, 1 RHS, median of {reps} runs, 1 thread ({dt:.1f} s), "
                      f"scaled by (n/{ns})^3 to n={n}"
                      + (f" x {calls} calls (one per RHS, as the reference API)" if calls > 1 else "")}


# ------------------------------------------------------------------ batched

def batch_inputs(torch, lmx, np, count, seed):
    """count LU and count Cholesky problems, n log-uniform in [8, 512], 1-8
    RHS; matrices generated on the device by size group."""
    rng = np.random.default_rng(seed)
    shapes = []
    for kind in (0, 1):
        ns = np.exp(rng.uniform(np.log(8), np.log(512), count)).round().astype(int).clip(8, 512)
        rs = rng.integers(1, 9, count)
        shapes += [(kind, int(m), int(r)) for m, r in zip(ns, rs)]
    shapes.sort(key=lambda x: (x[0], x[1]))  # size groups are contiguous in the pools
    probs, na, nb, npv = lmx.batch_layout(shapes)
    g = torch.Generator(device="cuda").manual_seed(seed)
    A = torch.empty(na, dtype=torch.float64, device="cuda")
    i = 0
    while i < len(shapes):
        kind, m = shapes[i][0], shapes[i][1]
        j = i
        while j < len(shapes) and shapes[j][0] == kind and shapes[j][1] == m:
            j += 1
        c = j - i
        off = probs[i].a_offset
        if kind == 0:
            blk = torch.rand(c, m, m, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
        else:
            G = torch.randn(c, m, m, dtype=torch.float64, device="cuda", generator=g)
            blk = torch.bmm(G, G.transpose(1, 2).contiguous()) / m + torch.eye(m, dtype=torch.float64,
                                                                                  device="cuda")
            del G
        A[off:off + c * m * m] = blk.reshape(-1)
        del blk
        i = j
    B = torch.rand(nb, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    return shapes, probs, na, nb, npv, A, B


def batch_backward_error(torch, probs, A, X, B):
    """max over problems of eta / (n eps), grouped by (n, nrhs) on the device."""
    groups = {}
    for p in probs:
        groups.setdefault((p.n, p.nrhs), []).append(p)
    worst = 0.0
    for (m, r), ps in groups.items():
        ia = torch.tensor([p.a_offset for p in ps], device="cuda")[:, None] + torch.arange(m * m, device="cuda")
        ib = torch.tensor([p.b_offset for p in ps], device="cuda")[:, None] + torch.arange(m * r, device="cuda")
        Am = A[ia].view(-1, m, m)
        kinds = torch.tensor([p.kind for p in ps], device="cuda")
        Al = torch.tril(Am)
        Asym = Al + torch.tril(Am, -1).transpose(1, 2)
        Am = torch.where((kinds == 1)[:, None, None], Asym, Am)
        Xm = X[ib].view(-1, m, r)
        Bm = B[ib].view(-1, m, r)
        R = Bm - torch.bmm(Am, Xm.contiguous())
        num = R.abs().amax(1)
        den = Am.abs().sum(2).amax(1)[:, None] * Xm.abs().amax(1) + Bm.abs().amax(1)
        worst = max(worst, (num / den).max().item() / (m * EPS))
    return worst


def bench_batch(a, np, torch, lmx):
    mixed = a.workload == "batch_mixed"
    t0 = time.perf_counter()
    shapes, probs, na, nb, npv, A, B = batch_inputs(torch, lmx, np, a.count, a.seed)
    torch.cuda.synchronize()
    log(f"[bench] {a.workload}: {len(shapes)} problems, A pool {na * 8 / 2**30:.2f} GiB, inputs ready in "
        f"{time.perf_counter() - t0:.1f} s")
    flags = lmx.BATCH_GRAPH | (0 if a.no_autotune else lmx.BATCH_AUTOTUNE)
    prec = lmx.PREC_F32 if mixed else lmx.PREC_F64
    plan = lmx.BatchPlan(probs, na, nb, 0, prec, flags)
    q = plan.query()
    log(f"[bench] plan: {q}")
    A0 = A.clone() if not mixed else A
    Bw = B.clone()
    info = torch.zeros(len(shapes), dtype=torch.int32, device="cuda")
    iters = torch.zeros(len(shapes), dtype=torch.int32, device="cuda")

    def reset():
        if not mixed:
            A.copy_(A0)
        Bw.copy_(B)

    for _ in range(a.warmup):
        reset()
        plan.execute(A, Bw, info, None, iters)
    torch.cuda.synchronize()
    log(f"[bench] warmup done ({a.warmup} steps)")
    # each step's inputs are restored outside the timed region (the FP64
    # path factors A in place); events bracket the batched launch only
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(a.steps)]
    lmx.kernel_timing_reset()
    lmx.kernel_timing(True)
    for i in range(a.steps):
        reset()
        ev[i][0].record()
        plan.execute(A, Bw, info, None, iters)
        ev[i][1].record()
    torch.cuda.synchronize()
    lmx.kernel_timing(False)
    lat = [e0.elapsed_time(e1) for e0, e1 in ev]
    cnt, kms, kwork = lmx.kernel_timing_get("batch")
    lmx.kernel_timing_reset()
    ms_per_step = sum(lat) / len(lat)
    flops = q["flops"]
    infos = info.cpu().numpy()
    its = iters.cpu().numpy()
    eta = batch_backward_error(torch, probs, A0 if not mixed else A, Bw, B)
    plan.close()
    peak = 157.3 if mixed else 78.6
    achieved = kwork / (kms * 1e-3) / 1e12 if kms > 0 else None
    sps = len(shapes) / (ms_per_step * 1e-3)
    result = {"metric": "FP64-accurate solves/sec & TFLOP/s vs plain FP64 GPU; p50/p99 solve latency",
              "value": round(sps, 1), "unit": "solves/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
              "ms_per_step": round(ms_per_step, 3), "higher_is_better": True, "scaling": "weak",
              "vs_baseline": None, "dtype": "f32+f64" if mixed else "f64", "data": "synthetic",
              "tflops": round(flops / (ms_per_step * 1e-3) / 1e12, 3),
              "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
              "max_backward_error_over_neps": eta, "info_nonzero": int((infos != 0).sum()),
              "config": {"workload": f"batched solves: {a.count} LU (A iid U[-1,1]) + {a.count} Cholesky "
                                     "(A = G G^T/n + I), n log-uniform in [8,512], 1-8 RHS, "
                                     + ("FP32 factor + FP64 refinement" if mixed else "FP64")
                                     + ", one persistent launch (hipGraph)",
                         "problems": len(shapes), "variant": q}}
    if mixed:
        result["refine"] = {"fell_back": int((its < 0).sum()), "mean_iterations": float(its[its >= 0].mean()),
                            "max_iterations": int(its.max())}
    result["roofline"] = {"bound": "mfma", "kernel": "batch_kernel", "achieved": round(achieved, 3) if achieved else None,
                          "peak": peak, "unit": "TFLOP/s", "frac": round(achieved / peak, 4) if achieved else None,
                          "traffic": None, "launches_per_step": cnt // max(a.steps, 1),
                          "avg_launch_us": round(kms * 1e3 / max(cnt, 1), 2)}
    if not a.no_cpu_baseline:
        log("[bench] cpu baseline (oracle) ...")
        try:
            result["cpu_baseline"] = cpu_baseline(a.workload, 0, 0, batch=(shapes, flops))
            result["cpu_baseline"]["host_cores"] = os.cpu_count()
        except Exception as e:
            result["cpu_baseline"] = {"error": str(e)}
    print(json.dumps(result), flush=True)


# ------------------------------------------------------------------ serving

def serve_library(torch, np, count, seed):
    """Matrices the clients send: 70% SPD (G G^T/n + I), 30% U[-1,1], n
    log-uniform in [16, 4096], generated on the device, kept in pinned host
    memory as a client's buffers would be."""
    rng = np.random.default_rng(seed)
    g = torch.Generator(device="cuda").manual_seed(seed)
    lib = []
    for i in range(count):
        kind = 1 if rng.random() < 0.7 else 0
        n = int(round(np.exp(rng.uniform(np.log(16), np.log(4096)))))
        if kind:
            G = torch.randn(n, n, dtype=torch.float64, device="cuda", generator=g)
            A = G @ G.T / n + torch.eye(n, dtype=torch.float64, device="cuda")
        else:
            A = torch.rand(n, n, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
        h = torch.empty((n, n), dtype=torch.float64, pin_memory=True)
        h.copy_(A)
        lib.append((kind, n, h.numpy(), h))
    torch.cuda.synchronize()
    return lib


def serve_run(lmx, np, srv, lib, nreq, rate, seed, rhs_pool):
    """Open loop (Poisson arrivals at `rate` req/s; rate <= 0: all at once).
    -> per-request records and the wall time."""
    import threading
    rng = np.random.default_rng(seed)
    live = []            # resident factors: (handle, kind, n)
    lock = threading.Lock()
    recs = []
    pending = []
    cv = threading.Condition()
    done_flag = [False]

    def collector():
        while True:
            with cv:
                while not pending and not done_flag[0]:
                    cv.wait()
                if not pending and done_flag[0]:
                    return
                rid, meta = pending.pop(0)
            r = srv.wait(rid)
            recs.append((meta, r.rc, r.latency_ms, r.queue_ms, r.iterations, r.fell_back))
            if r.factor:
                with lock:
                    live.append((r.factor, meta[1], meta[2]))
                    while len(live) > 12:
                        srv.release(live.pop(0)[0])

    th = threading.Thread(target=collector)
    th.start()
    t0 = time.perf_counter()
    t_next = t0
    for i in range(nreq):
        if rate > 0:
            t_next += rng.exponential(1.0 / rate)
            dt = t_next - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
        nrhs = int(rng.integers(1, 33))
        reuse = None
        if rng.random() < 0.3:
            with lock:
                if live:
                    reuse = live[int(rng.integers(0, len(live)))]
        if reuse:
            fh, kind, n = reuse
            B = np.ascontiguousarray(rhs_pool[:n * nrhs].reshape(n, nrhs))
            X = np.empty_like(B)
            rid = srv.submit(kind, B, X, factor=fh)
            meta = ("reuse", kind, n, nrhs)
        else:
            kind, n, A, _ = lib[int(rng.integers(0, len(lib)))]
            B = np.ascontiguousarray(rhs_pool[:n * nrhs].reshape(n, nrhs))
            X = np.empty_like(B)
            keep = rng.random() < 0.5
            rid = srv.submit(kind, B, X, A=A, keep=keep)
            meta = ("new", kind, n, nrhs)
        with cv:
            pending.append((rid, meta))
            cv.notify()
    with cv:
        done_flag[0] = True
        cv.notify()
    th.join()
    wall = time.perf_counter() - t0
    for fh, _, _ in live:
        srv.release(fh)
    return recs, wall


def serve_flops(meta):
    typ, kind, n, nrhs = meta
    solve = 2.0 * n * n * nrhs
    if typ == "reuse":
        return solve
    return (n ** 3 / 3.0 if kind == 1 else 2.0 * n ** 3 / 3.0) + solve


def bench_serve(a, np, torch, lmx):
    t0 = time.perf_counter()
    lib = serve_library(torch, np, 48, a.seed)
    rhs_pool = np.random.default_rng(a.seed + 1).uniform(-1, 1, 4096 * 32)
    log(f"[bench] serve: library of {len(lib)} matrices in {time.perf_counter() - t0:.1f} s")
    srv = lmx.Server(budget_bytes=32 << 30, slots=4, small_n=256)
    try:
        # warm-up, then the closed-loop capacity (all requests at once)
        serve_run(lmx, np, srv, lib, 40, 0, a.seed + 2, rhs_pool)
        srv.stats(reset=True)
        recs, wall = serve_run(lmx, np, srv, lib, a.requests, 0, a.seed + 3, rhs_pool)
        st = srv.stats(reset=True)
        capacity = len(recs) / wall
        log(f"[bench] capacity {capacity:.1f} req/s (busy {st['busy_ms'] / st['window_ms']:.2f})")
        points = []
        for frac in (0.5, 0.8, 0.9):
            rate = frac * capacity
            recs, wall = serve_run(lmx, np, srv, lib, a.requests, rate, a.seed + 4 + int(frac * 10), rhs_pool)
            st = srv.stats(reset=True)
            lat = np.array([r[2] for r in recs])
            flops = sum(serve_flops(r[0]) for r in recs)
            pt = {"offered_req_per_s": round(rate, 1), "req_per_s": round(len(recs) / wall, 1),
                  "tflops": round(flops / wall / 1e12, 3), "p50_ms": round(float(np.percentile(lat, 50)), 3),
                  "p99_ms": round(float(np.percentile(lat, 99)), 3),
                  "utilisation": round(st["busy_ms"] / st["window_ms"], 3), "failed": int(sum(r[1] != 0 for r in recs)),
                  "fell_back": int(sum(r[5] for r in recs)), "peak_hbm_gib": round(st["peak_bytes"] / 2**30, 2)}
            log(f"[bench] rate {rate:.1f}: {pt}")
            points.append(pt)
    finally:
        srv.close()
    best = min(points, key=lambda p: abs(p["utilisation"] - 0.8))
    result = {"metric": "FP64-accurate solves/sec & TFLOP/s vs plain FP64 GPU; p50/p99 solve latency",
              "value": best["req_per_s"], "unit": "solves/s", "n_gpus": 1, "steps": a.requests, "warmup": 40,
              "ms_per_step": round(1e3 / best["req_per_s"], 3), "higher_is_better": True, "scaling": "weak",
              "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
              "p50_ms": best["p50_ms"], "p99_ms": best["p99_ms"], "tflops": best["tflops"],
              "utilisation": best["utilisation"], "capacity_req_per_s": round(capacity, 1), "sweep": points,
              "config": {"workload": "serving: Poisson arrivals, 70% Cholesky / 30% LU, n log-uniform in [16,4096], "
                                     "1-32 RHS, 30% reuse a resident factor; FP32 factor + FP64 refinement; "
                                     "HBM budget 32 GiB, 4 slots, n <= 256 batched",
                         "requests_per_point": a.requests}}
    print(json.dumps(result), flush=True)


# ------------------------------------------------------------------ main

def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {a.gpus} is not supported "
              "(one GPU holds the whole problem; more GPUs would take independent solves each)",
              file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    import lmx

    lmx.init(0)
    wl = a.workload
    if wl.startswith("batch"):
        return bench_batch(a, np, torch, lmx)
    if wl == "serve":
        return bench_serve(a, np, torch, lmx)
    defaults = {"lu_mixed": (16384, 1), "chol_mixed": (16384, 1), "lu_fp64": (8192, 16),
                "chol_fp64": (8192, 16), "gemm_fp64": (8192, 0), "gemm_bf16": (8192, 0), "gemm_bf16x3": (8192, 0)}
    gemm_wl = wl.startswith("gemm")
    prec = lmx.PREC_BF16 if a.factor == "bf16" else lmx.PREC_F32
    n = a.n or defaults[wl][0]
    nrhs = a.nrhs if a.nrhs is not None else defaults[wl][1]
    g = torch.Generator(device="cuda").manual_seed(a.seed)
    spd = wl.startswith("chol")
    t0 = time.perf_counter()
    if gemm_wl:
        A = uniform(torch, (n, n), g)
        B = uniform(torch, (n, n), g)
        if wl == "gemm_bf16":
            A, B = A.to(torch.bfloat16), B.to(torch.bfloat16)
        elif wl == "gemm_bf16x3":
            A, B = A.float(), B.float()
        C = torch.empty(n, n, dtype=torch.float64 if wl == "gemm_fp64" else torch.float32, device="cuda")
    elif wl in ("lu_mixed", "chol_mixed"):
        A = randsvd(torch, n, a.kappa, g, spd)
        B = uniform(torch, (n, nrhs), g)
        X = torch.zeros_like(B)
    else:
        A = spd_gram(torch, n, g) if spd else uniform(torch, (n, n), g)
        B = uniform(torch, (n, nrhs), g)
    torch.cuda.synchronize()
    log(f"[bench] {wl} n={n} nrhs={nrhs}: inputs ready in {time.perf_counter() - t0:.1f} s")

    # FP64 path work buffers
    if not gemm_wl:
        LU = torch.empty_like(A)
        XB = torch.empty_like(B)
        ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")
        info = torch.zeros(1, dtype=torch.int32, device="cuda")

    last = {}

    def step():
        if wl == "gemm_fp64":
            lmx.dgemm_dev(A, B, C)
        elif wl == "gemm_bf16":
            lmx.gemm_bf16_dev(A, B, C)
        elif wl == "gemm_bf16x3":
            lmx.sgemm_bf16x3_dev(A, B, C)
        elif wl in ("lu_mixed", "chol_mixed"):
            ri = (lmx.posv_mixed_dev if spd else lmx.gesv_mixed_dev)(A, B, X, prec)
            last["ri"] = ri
        else:
            LU.copy_(A)
            XB.copy_(B)
            if spd:
                lmx.posv_dev(LU, XB, info)
            else:
                lmx.gesv_dev(LU, ipiv, XB, info)

    for _ in range(a.warmup):
        step()
    torch.cuda.synchronize()
    log(f"[bench] warmup done ({a.warmup} steps)")

    ev = [torch.cuda.Event(enable_timing=True) for _ in range(a.steps + 1)]
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    ev[0].record()
    for i in range(a.steps):
        step()
        ev[i + 1].record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    lat = [ev[i].elapsed_time(ev[i + 1]) for i in range(a.steps)]
    ms_per_step = wall * 1e3 / a.steps
    log(f"[bench] timed {a.steps} steps: {ms_per_step:.2f} ms/step")

    # per-family kernel times from one more step, outside the timed region
    # (the per-launch events would otherwise be part of the measurement)
    lmx.kernel_timing_reset()
    lmx.kernel_timing(True)
    step()
    torch.cuda.synchronize()
    lmx.kernel_timing(False)
    fam = {}
    for f in KERNEL_FAMILIES + ["getrf", "potrf", "getrs", "refine"]:
        cnt, ms, work = lmx.kernel_timing_get(f)
        if cnt:
            fam[f] = {"launches": cnt, "ms_per_step": ms, "work_per_step": work}
    lmx.kernel_timing_reset()

    # dominant MFMA kernel family and its roofline
    mf = [f for f in ("gemm", "sgemm", "syrk", "ssyrk", "bgemm", "bgemm3") if f in fam]
    dom = max(mf, key=lambda f: fam[f]["ms_per_step"]) if mf else None
    roofline = None
    if dom:
        cnt, ms, work = fam[dom]["launches"], fam[dom]["ms_per_step"], fam[dom]["work_per_step"]
        achieved = work / (ms * 1e-3) / 1e12
        roofline = {"bound": "mfma",
                    # the FP64 GEMMs of these workloads take the NN kernel (no
                    # transposes, aligned rows): 128 x 64 tiles for the plain
                    # GEMM, 128 x 128 beside the LU panels; the FP32 trailing
                    # GEMM of the mixed LU shares its CUs with the panels and
                    # keeps the first kernel
                    "kernel": {"gemm": "gemm_nn_kernel<double, 128, 64>" if wl == "gemm_fp64"
                               else "gemm_nn_kernel<double, 128, 128>",
                               "sgemm": "gemm_mfma_kernel<float, 128, 128, 16>",
                               "syrk": "gemm_nn_kernel<double, B^T> (lower tiles)",
                               "ssyrk": "gemm_nn_kernel<float, B^T> (lower tiles)",
                               "bgemm": "gemm_bf16_kernel<bf16, 128, 128>",
                               "bgemm3": "gemm_bf16_kernel<float, split, 128, 128> (3 BF16 MFMAs a product)"}[dom],
                    "achieved": round(achieved, 3), "peak": PEAK[dom], "unit": "TFLOP/s",
                    "frac": round(achieved / PEAK[dom], 4), "traffic": profiled_traffic(wl, n, dom),
                    "launches_per_step": cnt, "avg_launch_us": round(ms * 1e3 / cnt, 2)}

    result = {"metric": "FP64-accurate solves/sec & TFLOP/s vs plain FP64 GPU; p50/p99 solve latency",
              "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(ms_per_step, 3),
              "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "data": "synthetic"}
    lat_sorted = sorted(lat)
    p50 = float(np.percentile(lat_sorted, 50))
    p99 = float(np.percentile(lat_sorted, 99))

    if gemm_wl:
        flops = 2.0 * n ** 3
        kind = {"gemm_fp64": ("f64", "FP64 MFMA"), "gemm_bf16": ("bf16", "BF16 MFMA, FP32 accumulation"),
                "gemm_bf16x3": ("f32 (bf16x3)", "FP32 operands split in two BF16 parts, 3 BF16 MFMAs a product")}[wl]
        result.update({"value": round(flops / (ms_per_step * 1e-3) / 1e12, 3), "unit": "TFLOP/s", "dtype": kind[0],
                       "config": {"workload": f"GEMM C = A B, {n}x{n}x{n}, {kind[1]}, A, B iid U[-1,1]", "n": n}})
    else:
        if spd:
            flops = n ** 3 / 3.0 + 2.0 * n * n * nrhs
        else:
            flops = 2.0 * n ** 3 / 3.0 + 2.0 * n * n * nrhs
        sps = 1e3 / ms_per_step
        out = X if wl.endswith("mixed") else XB
        eta = backward_error(torch, A, out, B)
        name = "choleskySolve" if spd else "luSolve"
        result.update({"value": round(sps, 4), "unit": "solves/s", "tflops": round(flops * sps / 1e12, 3),
                       "p50_ms": round(p50, 3), "p99_ms": round(p99, 3), "backward_error": eta,
                       "backward_error_bound_10neps": 10 * n * EPS})
        if wl.endswith("mixed"):
            ri = last["ri"]
            bf = a.factor == "bf16"
            result["dtype"] = "bf16x3+f64" if bf else "f32+f64"
            result["refine"] = {"iterations": ri.iterations, "converged": ri.converged, "fell_back": ri.fell_back,
                                "method": "gmres-ir" if bf else "ir"}
            fac = "BF16-factor (BF16 tiles, split operands) + FP64 GMRES-IR" if bf else \
                "FP32-factor + FP64 refinement"
            result["config"] = {"workload": f"{name} {fac}, n={n}, {nrhs} RHS, "
                                            f"A = U diag(sigma) V^T (Haar, kappa={a.kappa:g})",
                                "n": n, "nrhs": nrhs, "kappa": a.kappa, "factor": a.factor}
            if not a.no_fp64_compare:
                k2 = max(2, a.steps // 2)
                for _ in range(1):
                    LU.copy_(A); XB.copy_(B)
                    (lmx.posv_dev(LU, XB, info) if spd else lmx.gesv_dev(LU, ipiv, XB, info))
                torch.cuda.synchronize()
                t1 = time.perf_counter()
                for _ in range(k2):
                    LU.copy_(A); XB.copy_(B)
                    (lmx.posv_dev(LU, XB, info) if spd else lmx.gesv_dev(LU, ipiv, XB, info))
                torch.cuda.synchronize()
                ms64 = (time.perf_counter() - t1) * 1e3 / k2
                eta64 = backward_error(torch, A, XB, B)
                result["fp64_path"] = {"solves_per_s": round(1e3 / ms64, 4), "ms_per_solve": round(ms64, 3),
                                       "tflops": round(flops / (ms64 * 1e-3) / 1e12, 3), "backward_error": eta64}
                result["speedup_vs_fp64"] = round(ms64 / ms_per_step, 3)
                fwd = ((out - XB).abs().max() / XB.abs().max()).item()
                result["rel_diff_vs_fp64_solution"] = fwd
        else:
            result["dtype"] = "f64"
            result["config"] = {"workload": f"{name} FP64, n={n}, {nrhs} RHS, "
                                            + ("A = G G^T/n + I" if spd else "A iid U[-1,1]"), "n": n, "nrhs": nrhs}
    result["roofline"] = roofline
    result["families"] = {k: {"launches_per_step": v["launches"], "ms_per_step": round(v["ms_per_step"], 3)}
                          for k, v in fam.items()}
    if not a.no_cpu_baseline:
        log("[bench] cpu baseline (oracle) ...")
        try:
            cpu = cpu_baseline(wl, n, nrhs)
            cpu["cpu_model"] = platform.processor() or ""
            try:
                model = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
                for line in model.splitlines():
                    if line.startswith("Model name"):
                        cpu["cpu_model"] = line.split(":", 1)[1].strip()
            except OSError:
                pass
            cpu["host_cores"] = os.cpu_count()
            result["cpu_baseline"] = cpu
        except Exception as e:  # the baseline is reported, never required
            result["cpu_baseline"] = {"error": str(e)}
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
