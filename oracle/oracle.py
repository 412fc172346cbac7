"""CPU oracle for the dense-solve path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline
leg import this module, and only as the checker or the timed CPU baseline.
The product library never imports it and never falls back to it.

Two parts:

* ctypes bindings to ``build/liblmx_oracle.so``, the C restatement of the
  reference's algorithms (``lmx_oracle.c`` documents which reference routine
  each entry restates, with file:line);
* numpy restatements of the reference's residual oracles
  (``test/Test/Residuals.hs``), used to state backward errors.

Parity status: every restated routine is pinned by the checks listed in the
header of ``lmx_oracle.c`` and exercised in ``tests/test_oracle.py``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "liblmx_oracle.so")
_lib = None

_ix = ctypes.c_int64
_pd = ctypes.POINTER(ctypes.c_double)
_pi = ctypes.POINTER(ctypes.c_int64)

# machineEps of test/Test/Residuals.hs:57
MACHINE_EPS = 2.220446049250313e-16


def build() -> str:
    """Compile the oracle with its Makefile (idempotent)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        sigs = {
            "orc_matmulp": [_ix, _ix, _ix, _pd, _pd, _pd],
            "orc_gemm_generic": [_ix, _ix, _ix, ctypes.c_double, _pd, _pd,
                                 ctypes.c_double, _pd, _pd],
            "orc_matmul_generic": [_ix, _ix, _ix, _pd, _pd, _pd],
            "orc_lu_panel_factor": [_ix, _pd, _ix, _pi],
            "orc_lu_unblocked_factor": [_ix, _pd, _pi],
            "orc_forward_unit_packed": [_ix, _pd, _pd],
            "orc_back_packed": [_ix, _pd, _pd],
            "orc_apply_swaps": [_ix, _pi, _pd],
            "orc_lu_solve_factored": [_ix, _pd, _ix, _pi, _pd],
            "orc_lusolvep": [_ix, _pd, _pd, _pd, _pd, _pi],
            "orc_lu_generic": [_ix, _pd, _pi],
            "orc_forward_sub": [_ix, _pd, _pd, _pd],
            "orc_back_sub": [_ix, _pd, _pd, _pd],
            "orc_forward_sub_unit": [_ix, _pd, _pd, _pd],
            "orc_back_sub_unit": [_ix, _pd, _pd, _pd],
            "orc_lusolve_generic": [_ix, _pd, _pd, _pd],
            "orc_cholesky_generic": [_ix, _pd, _pd],
            "orc_cholsolve_generic": [_ix, _pd, _pd, _pd],
            "orc_cholesky_p_factor": [_ix, _pd, _pd],
            "orc_forward_chol_packed": [_ix, _pd, _pd],
            "orc_back_cholT_packed": [_ix, _pd, _pd],
            "orc_cholsolvep": [_ix, _pd, _pd, _pd, _pd],
            "orc_chol_solve_factored": [_ix, _pd, _pd],
            "orc_matvecp": [_ix, _ix, _pd, _pd, _pd],
            "orc_gemv_generic": [_ix, _ix, ctypes.c_double, _pd, _pd, ctypes.c_double, _pd, _pd],
            "orc_inverse_iteration": [_ix, _pd, ctypes.c_double, _pd, _ix, ctypes.c_double, _pd, _pd],
            "orc_matmulp_par": [_ix, _ix, _ix, _pd, _pd, _pd, _ix],
        }
        for name, args in sigs.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = _ix if name.endswith("_factor") or name == "orc_inverse_iteration" else None
        for name, args in {"orc_raw_dot": [_ix, _pd, _pd], "orc_norm1": [_ix, _ix, _pd],
                           "orc_norm_inf": [_ix, _ix, _pd], "orc_norm_frob": [_ix, _ix, _pd],
                           "orc_det": [_ix, _pd]}.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = ctypes.c_double
        _lib = L
    return _lib


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_pd)


def _p(a):
    return a.ctypes.data_as(_pd)


# --------------------------------------------------------------------------
# GEMM
# --------------------------------------------------------------------------

def matMulP(A, B):
    """matMulP (BLAS/Level3.hs:183-207) -> C = A B with the reference's rounding."""
    A, pa = _d(A)
    B, pb = _d(B)
    m, k = A.shape
    k2, n = B.shape
    assert k == k2
    C = np.empty((m, n))
    lib().orc_matmulp(m, k, n, pa, pb, _p(C))
    return C


def gemm(alpha, A, B, beta, C):
    """Generic gemm (BLAS/Level3.hs:102-131): alpha A B + beta C."""
    A, pa = _d(A)
    B, pb = _d(B)
    C, pc = _d(C)
    m, k = A.shape
    n = B.shape[1]
    out = np.empty((m, n))
    lib().orc_gemm_generic(m, k, n, float(alpha), pa, pb, float(beta), pc, _p(out))
    return out


def matMulGeneric(A, B):
    """matMulGeneric (BLAS/Level3.hs:150-181)."""
    A, pa = _d(A)
    B, pb = _d(B)
    m, k = A.shape
    n = B.shape[1]
    C = np.empty((m, n))
    lib().orc_matmul_generic(m, k, n, pa, pb, _p(C))
    return C


# --------------------------------------------------------------------------
# LU
# --------------------------------------------------------------------------

def lu_p_factor(A, nb=32):
    """Factorisation inside luSolveP (Solve/LU.hs:276-284): the panel path
    (panelLUFactor, nb=32) for n >= 64, the unblocked path below.
    Returns (packed LU, swaps) with swaps[k] = pivot row chosen at step k
    (LAPACK-style ipiv, 0-based; steps beyond the last are identity)."""
    A = np.array(A, dtype=np.float64, order="C", copy=True)
    n = A.shape[0]
    sw = np.arange(max(n, 1), dtype=np.int64)
    if n >= 64:
        lib().orc_lu_panel_factor(n, _p(A), nb, sw.ctypes.data_as(_pi))
    else:
        lib().orc_lu_unblocked_factor(n, _p(A), sw.ctypes.data_as(_pi))
    return A, sw[:n]


def lu_panel_factor(A, nb):
    """panelLUFactor (Solve/LU.hs:311-357) with an explicit panel width."""
    A = np.array(A, dtype=np.float64, order="C", copy=True)
    n = A.shape[0]
    sw = np.arange(max(n, 1), dtype=np.int64)
    lib().orc_lu_panel_factor(n, _p(A), nb, sw.ctypes.data_as(_pi))
    return A, sw[:n]


def lu_unblocked_factor(A):
    A = np.array(A, dtype=np.float64, order="C", copy=True)
    n = A.shape[0]
    sw = np.arange(max(n, 1), dtype=np.int64)
    lib().orc_lu_unblocked_factor(n, _p(A), sw.ctypes.data_as(_pi))
    return A, sw[:n]


def lu_solve_factored(LU, swaps, b):
    """Solve phase of luSolveP (Solve/LU.hs:294-303) for one right-hand side."""
    LU, pl = _d(LU)
    n = LU.shape[0]
    sw = np.ascontiguousarray(swaps, dtype=np.int64)
    x = np.array(b, dtype=np.float64, copy=True)
    lib().orc_lu_solve_factored(n, pl, n, sw.ctypes.data_as(_pi), _p(x))
    return x


def lu_solve_factored_multi(LU, swaps, B):
    """Column-by-column lu_solve_factored for an n x nrhs block (the reference
    has one-RHS solves only; multi-RHS parity is per column)."""
    B = np.asarray(B, dtype=np.float64)
    if B.ndim == 1:
        return lu_solve_factored(LU, swaps, B)
    return np.stack([lu_solve_factored(LU, swaps, B[:, j]) for j in range(B.shape[1])], axis=1)


def luSolveP(A, b):
    """luSolveP (Solve/LU.hs:265-304)."""
    A, pa = _d(A)
    b, pb = _d(b)
    n = A.shape[0]
    x = np.empty(n)
    lib().orc_lusolvep(n, pa, pb, _p(x), None, None)
    return x


def lu(A):
    """Generic lu (Solve/LU.hs:121-167) -> (packed LU, permutation vector)."""
    A = np.array(A, dtype=np.float64, order="C", copy=True)
    n = A.shape[0]
    piv = np.zeros(max(n, 1), dtype=np.int64)
    lib().orc_lu_generic(n, _p(A), piv.ctypes.data_as(_pi))
    return A, piv[:n]


def luSolve(A, b):
    """Generic luSolve (Solve/LU.hs:240-259)."""
    A, pa = _d(A)
    b, pb = _d(b)
    n = A.shape[0]
    x = np.empty(n)
    lib().orc_lusolve_generic(n, pa, pb, _p(x))
    return x


def swaps_to_perm(swaps):
    """LAPACK-style swap list -> permutation vector (row i of PA = row perm[i] of A)."""
    n = len(swaps)
    perm = np.arange(n)
    for k, p in enumerate(swaps):
        if p != k:
            perm[[k, p]] = perm[[p, k]]
    return perm


# --------------------------------------------------------------------------
# Triangular (generic, Solve/Triangular.hs)
# --------------------------------------------------------------------------

def _tri(fn, T, b):
    T, pt = _d(T)
    b, pb = _d(b)
    n = T.shape[0]
    x = np.empty(n)
    getattr(lib(), fn)(n, pt, pb, _p(x))
    return x


def forwardSub(L, b):
    return _tri("orc_forward_sub", L, b)


def backSub(U, b):
    return _tri("orc_back_sub", U, b)


def forwardSubUnit(L, b):
    return _tri("orc_forward_sub_unit", L, b)


def backSubUnit(U, b):
    return _tri("orc_back_sub_unit", U, b)


# --------------------------------------------------------------------------
# Cholesky
# --------------------------------------------------------------------------

def cholesky(A):
    """Generic cholesky (Solve/Cholesky.hs:114-147) -> lower G."""
    A, pa = _d(A)
    n = A.shape[0]
    G = np.empty((n, n))
    lib().orc_cholesky_generic(n, pa, _p(G))
    return G


def cholesky_p_factor(A):
    """Factorisation inside choleskySolveP (Solve/Cholesky.hs:246-256)."""
    A, pa = _d(A)
    n = A.shape[0]
    G = np.empty((n, n))
    lib().orc_cholesky_p_factor(n, pa, _p(G))
    return G


def chol_solve_factored(G, b):
    G, pg = _d(G)
    x = np.array(b, dtype=np.float64, copy=True)
    lib().orc_chol_solve_factored(G.shape[0], pg, _p(x))
    return x


def chol_solve_factored_multi(G, B):
    B = np.asarray(B, dtype=np.float64)
    if B.ndim == 1:
        return chol_solve_factored(G, B)
    return np.stack([chol_solve_factored(G, B[:, j]) for j in range(B.shape[1])], axis=1)


def choleskySolveP(A, b):
    """choleskySolveP (Solve/Cholesky.hs:241-273)."""
    A, pa = _d(A)
    b, pb = _d(b)
    n = A.shape[0]
    x = np.empty(n)
    lib().orc_cholsolvep(n, pa, pb, _p(x), None)
    return x


def choleskySolve(A, b):
    """Generic choleskySolve (Solve/Cholesky.hs:229-235)."""
    A, pa = _d(A)
    b, pb = _d(b)
    n = A.shape[0]
    x = np.empty(n)
    lib().orc_cholsolve_generic(n, pa, pb, _p(x))
    return x


# --------------------------------------------------------------------------
# Residual oracles (test/Test/Residuals.hs)
# --------------------------------------------------------------------------

def scaled_residual_linear(A, x, b):
    """scaledResidualLinear (test/Test/Residuals.hs:69-78):
    ||Ax-b||_inf / (||A||_inf ||x||_inf + ||b||_inf + n eps), per column
    when x, b are n x nrhs (the max over columns is returned)."""
    A = np.asarray(A, dtype=np.float64)
    x = np.asarray(x, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if x.ndim == 1:
        x = x[:, None]
        b = b[:, None]
    n = A.shape[0]
    r = A @ x - b
    num = np.abs(r).max(axis=0) if n else np.zeros(x.shape[1])
    normA = np.abs(A).sum(axis=1).max() if n else 0.0
    den = normA * np.abs(x).max(axis=0) + np.abs(b).max(axis=0) + n * MACHINE_EPS
    return float((num / den).max()) if x.shape[1] else 0.0


def scaled_residual_lu(A, LU, perm):
    """scaledResidualLU (test/Test/Residuals.hs:167-189):
    ||PA - LU||_F / (||A||_F n eps + eps)."""
    A = np.asarray(A, dtype=np.float64)
    n = A.shape[0]
    L = np.tril(LU, -1) + np.eye(n)
    U = np.triu(LU)
    PA = A[np.asarray(perm)]
    return float(np.linalg.norm(PA - L @ U) / (np.linalg.norm(A) * n * MACHINE_EPS + MACHINE_EPS))


def scaled_residual_cholesky(A, G):
    """scaledResidualCholesky (test/Test/Residuals.hs:148-156):
    ||A - G G^T||_F / (||A||_F n eps + eps)."""
    A = np.asarray(A, dtype=np.float64)
    n = A.shape[0]
    G = np.asarray(G)
    return float(np.linalg.norm(A - G @ G.T) / (np.linalg.norm(A) * n * MACHINE_EPS + MACHINE_EPS))


def condition_tolerance(kappa, n):
    """conditionTolerance (test/Test/Residuals.hs:218-219): kappa n eps."""
    return kappa * n * MACHINE_EPS


# --------------------------------------------------------------------------
# Level 1/2, norms, SYRK, det, inverse iteration
# --------------------------------------------------------------------------

def rawDot(x, y):
    """rawDot# (Internal/Kernel.hs:89-110)."""
    x, px = _d(x)
    y, py = _d(y)
    return lib().orc_raw_dot(len(x), px, py)


def matvecP(A, x):
    """matvecP (BLAS/Level2.hs:132-140): y_i = rawDot#(row i, x)."""
    A, pa = _d(A)
    x, px = _d(x)
    y = np.empty(A.shape[0])
    lib().orc_matvecp(A.shape[0], A.shape[1], pa, px, _p(y))
    return y


def gemv(alpha, A, x, beta, y):
    """The generic gemv (BLAS/Level2.hs:92-100)."""
    A, pa = _d(A)
    x, px = _d(x)
    y, py = _d(y)
    out = np.empty(A.shape[0])
    lib().orc_gemv_generic(A.shape[0], A.shape[1], alpha, pa, px, beta, py, _p(out))
    return out


def norm1(A):
    A, pa = _d(np.atleast_2d(A))
    return lib().orc_norm1(A.shape[0], A.shape[1], pa)


def normInf(A):
    A, pa = _d(np.atleast_2d(A))
    return lib().orc_norm_inf(A.shape[0], A.shape[1], pa)


def normFrob(A):
    A, pa = _d(np.atleast_2d(A))
    return lib().orc_norm_frob(A.shape[0], A.shape[1], pa)


def matMulAtAP(A):
    """matMulAtAP (BLAS/Level3.hs:353-373): rawSyrkLowerKernel's lower
    triangle has, element by element, the rounding of rawGemmKernel on
    (A^T, A) (the same 64-tiles over k and columns, FMA chains, and the
    scalar column tail per k tile), mirrored into the upper triangle."""
    A = np.asarray(A, dtype=np.float64)
    C = matMulP(np.ascontiguousarray(A.T), A)
    L = np.tril(C)
    return L + np.tril(C, -1).T


def det(A):
    """det (Solve/LU.hs:534-558)."""
    A, pa = _d(A)
    return lib().orc_det(A.shape[0], pa)


def inverseIteration(A, mu, q0, max_iter, tol):
    """inverseIteration (Eigen/Power.hs:113-135) -> (lambda, q, steps)."""
    A, pa = _d(A)
    q0, pq = _d(q0)
    q = np.empty(A.shape[0])
    lam = ctypes.c_double(0.0)
    it = lib().orc_inverse_iteration(A.shape[0], pa, mu, pq, max_iter, tol, ctypes.pointer(lam), _p(q))
    return lam.value, q, it


def matMulPPar(A, B, threads):
    """matMulPPar (BLAS/Level3.hs:213-294) on `threads` host threads."""
    A, pa = _d(A)
    B, pb = _d(B)
    m, k = A.shape
    n = B.shape[1]
    C = np.empty((m, n))
    lib().orc_matmulp_par(m, k, n, pa, pb, _p(C), threads)
    return C
