"""Python host layer over liblmx.so (the C ABI of include/lmx.h).

Mirrors the public functions of Numeric.LinearAlgebra.Massiv on the
dense-solve path under their reference names and argument meaning
(row-major matrices, one-vector right-hand sides, permutation vectors from
`lu`), so tests read like the reference's own:

    matMul / matMulP / matMulPPar   src/.../BLAS/Level3.hs:144-147, 183-294
    gemm                            src/.../BLAS/Level3.hs:102-131
    lu, luSolve, luSolveP           src/.../Solve/LU.hs:121-167, 240-304
    cholesky, choleskySolve,
    choleskySolveP                  src/.../Solve/Cholesky.hs:114-147, 229-273
    forwardSub, backSub,
    forwardSubUnit, backSubUnit     src/.../Solve/Triangular.hs:90-249

plus the LAPACK-shaped entry points (getrf/getrs/gesv/potrf/potrs/posv/
trsm), the mixed-precision refined solvers and the device-pointer (`*_dev`)
calls that take torch tensors already resident in HBM.

Every call runs on the GPU through liblmx.so; there is no CPU fallback.  If
the library is missing the import fails.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblmx.so")
if not os.path.exists(LIB_PATH):
    raise ImportError(f"liblmx.so not built at {LIB_PATH} (run `make -C linear-massiv_amd`)")

# One HIP runtime per process: torch ships its own libamdhip64 (same SONAME
# libamdhip64.so.7).  Importing torch first lets liblmx.so bind to that copy;
# loading liblmx.so first would bring in a second runtime that torch cannot
# share the device with.
try:
    import torch  # noqa: F401
except ImportError:  # torch is optional plumbing
    torch = None

_lib = ctypes.CDLL(LIB_PATH)

LMX_OK = 0
LMX_ERR_ARG = -1
LMX_ERR_DEVICE = -2
LMX_ERR_NOMEM = -3
LMX_ERR_NOT_CONVERGED = -4
LMX_ERR_BUDGET = -5
LMX_ERR_INTERNAL = -6
PREC_F64, PREC_F32, PREC_BF16 = 0, 1, 2

EXPORTED = [
    "lmx_init", "lmx_finalize", "lmx_last_error", "lmx_version",
    "lmx_dgemm", "lmx_dgemm_dev", "lmx_sgemm", "lmx_sgemm_dev",
    "lmx_dgetrf", "lmx_dgetrf_dev", "lmx_sgetrf", "lmx_sgetrf_dev", "lmx_dgetrs", "lmx_dgetrs_dev", "lmx_dgesv", "lmx_dgesv_dev",
    "lmx_dpotrf", "lmx_dpotrf_dev", "lmx_dpotrs", "lmx_dpotrs_dev", "lmx_dposv", "lmx_dposv_dev",
    "lmx_dtrsm", "lmx_dtrsm_dev",
    "lmx_dgesv_mixed", "lmx_dgesv_mixed_dev", "lmx_dposv_mixed", "lmx_dposv_mixed_dev",
    "lmx_batch_plan_create", "lmx_batch_plan_execute", "lmx_batch_plan_autotune", "lmx_batch_plan_query",
    "lmx_batch_plan_destroy",
    "lmx_server_create", "lmx_server_submit", "lmx_server_wait", "lmx_server_poll", "lmx_server_factor_release",
    "lmx_server_stats", "lmx_server_destroy",
    "lmx_kernel_timing", "lmx_kernel_timing_reset", "lmx_kernel_timing_get",
]
BATCH_MAX_N, BATCH_MAX_NRHS = 512, 8
BATCH_LU, BATCH_CHOLESKY = 0, 1
BATCH_FACTORS, BATCH_GRAPH, BATCH_AUTOTUNE = 1, 2, 4
REQ_KEEP_FACTOR = 1


class LmxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"lmx error {code}: {msg}")
        self.code = code


class RefineInfo(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int32), ("converged", ctypes.c_int32),
                ("fell_back", ctypes.c_int32), ("info", ctypes.c_int32),
                ("backward_error", ctypes.c_double)]


class BatchProblem(ctypes.Structure):
    """lmx_batch_problem (include/lmx.h)."""
    _fields_ = [("kind", ctypes.c_int32), ("n", ctypes.c_int32), ("nrhs", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("a_offset", ctypes.c_int64), ("b_offset", ctypes.c_int64),
                ("ipiv_offset", ctypes.c_int64)]


class ServerConfig(ctypes.Structure):
    """lmx_server_config (include/lmx.h)."""
    _fields_ = [("budget_bytes", ctypes.c_int64), ("slots", ctypes.c_int32), ("small_n", ctypes.c_int32),
                ("small_max", ctypes.c_int32), ("reserved", ctypes.c_int32), ("aging_ms", ctypes.c_double)]


class RequestResult(ctypes.Structure):
    """lmx_request_result (include/lmx.h)."""
    _fields_ = [("status", ctypes.c_int32), ("rc", ctypes.c_int32), ("iterations", ctypes.c_int32),
                ("fell_back", ctypes.c_int32), ("backward_error", ctypes.c_double), ("factor", ctypes.c_int64),
                ("queue_ms", ctypes.c_double), ("run_ms", ctypes.c_double), ("latency_ms", ctypes.c_double)]


class ServerCounters(ctypes.Structure):
    """lmx_server_counters (include/lmx.h)."""
    _fields_ = [(f, ctypes.c_int64) for f in ("completed", "failed", "queued", "inflight", "bytes_in_use",
                                              "peak_bytes", "budget_bytes", "resident_factors")] + \
               [("busy_ms", ctypes.c_double), ("window_ms", ctypes.c_double)]


_i64, _dbl, _flt, _vp, _chr = ctypes.c_int64, ctypes.c_double, ctypes.c_float, ctypes.c_void_p, ctypes.c_char
_i32 = ctypes.c_int32
_pRI = ctypes.POINTER(RefineInfo)

_sigs = {
    "lmx_init": [ctypes.c_int], "lmx_finalize": [], "lmx_version": [],
    "lmx_dgemm": [_chr, _chr, _i64, _i64, _i64, _dbl, _vp, _i64, _vp, _i64, _dbl, _vp, _i64],
    "lmx_dgemm_dev": [_vp, _chr, _chr, _i64, _i64, _i64, _dbl, _vp, _i64, _vp, _i64, _dbl, _vp, _i64],
    "lmx_sgemm": [_chr, _chr, _i64, _i64, _i64, _flt, _vp, _i64, _vp, _i64, _flt, _vp, _i64],
    "lmx_sgemm_dev": [_vp, _chr, _chr, _i64, _i64, _i64, _flt, _vp, _i64, _vp, _i64, _flt, _vp, _i64],
    "lmx_dgetrf": [_i64, _vp, _i64, _vp],
    "lmx_dgetrf_dev": [_vp, _i64, _vp, _i64, _vp, _vp],
    "lmx_sgetrf": [_i64, _vp, _i64, _vp],
    "lmx_sgetrf_dev": [_vp, _i64, _vp, _i64, _vp, _vp],
    "lmx_dgetrs": [_i64, _i64, _vp, _i64, _vp, _vp, _i64],
    "lmx_dgetrs_dev": [_vp, _i64, _i64, _vp, _i64, _vp, _vp, _i64],
    "lmx_dgesv": [_i64, _i64, _vp, _i64, _vp, _vp, _i64],
    "lmx_dgesv_dev": [_vp, _i64, _i64, _vp, _i64, _vp, _vp, _i64, _vp],
    "lmx_dpotrf": [_i64, _vp, _i64],
    "lmx_dpotrf_dev": [_vp, _i64, _vp, _i64, _vp],
    "lmx_dpotrs": [_i64, _i64, _vp, _i64, _vp, _i64],
    "lmx_dpotrs_dev": [_vp, _i64, _i64, _vp, _i64, _vp, _i64],
    "lmx_dposv": [_i64, _i64, _vp, _i64, _vp, _i64],
    "lmx_dposv_dev": [_vp, _i64, _i64, _vp, _i64, _vp, _i64, _vp],
    "lmx_dtrsm": [_chr, _chr, _chr, _i64, _i64, _vp, _i64, _vp, _i64],
    "lmx_dtrsm_dev": [_vp, _chr, _chr, _chr, _i64, _i64, _vp, _i64, _vp, _i64],
    "lmx_dgesv_mixed": [ctypes.c_int, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64, _pRI],
    "lmx_dgesv_mixed_dev": [_vp, ctypes.c_int, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64, _pRI],
    "lmx_dposv_mixed": [ctypes.c_int, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64, _pRI],
    "lmx_dposv_mixed_dev": [_vp, ctypes.c_int, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64, _pRI],
    "lmx_batch_plan_create": [ctypes.POINTER(_vp), ctypes.POINTER(BatchProblem), _i32, _i64, _i64, _i64, _i32,
                              _i32],
    "lmx_batch_plan_execute": [_vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "lmx_batch_plan_autotune": [_vp, _vp],
    "lmx_batch_plan_query": [_vp, ctypes.POINTER(_i32), ctypes.POINTER(_i32), ctypes.POINTER(_i32),
                             ctypes.POINTER(_i32), ctypes.POINTER(_dbl), ctypes.POINTER(_dbl)],
    "lmx_batch_plan_destroy": [_vp],
    "lmx_server_create": [ctypes.POINTER(_vp), ctypes.POINTER(ServerConfig)],
    "lmx_server_submit": [_vp, _i32, _i64, _i64, _vp, _i64, _i64, _vp, _i64, _vp, _i64, _i32, ctypes.POINTER(_i64)],
    "lmx_server_wait": [_vp, _i64, ctypes.POINTER(RequestResult)],
    "lmx_server_poll": [_vp, _i64, ctypes.POINTER(RequestResult)],
    "lmx_server_factor_release": [_vp, _i64],
    "lmx_server_stats": [_vp, ctypes.POINTER(ServerCounters), _i32],
    "lmx_server_destroy": [_vp],
    "lmx_kernel_timing": [ctypes.c_int], "lmx_kernel_timing_reset": [],
    "lmx_kernel_timing_get": [ctypes.c_char_p, ctypes.POINTER(_i64), ctypes.POINTER(_dbl), ctypes.POINTER(_dbl)],
}
for _name, _args in _sigs.items():
    if not hasattr(_lib, _name):
        continue  # reported by exported_symbols() / the ABI test
    _f = getattr(_lib, _name)
    _f.argtypes = _args
    _f.restype = ctypes.c_int
_lib.lmx_last_error.argtypes = []
_lib.lmx_last_error.restype = ctypes.c_char_p


def missing_symbols():
    """Names declared by include/lmx.h that the loaded library does not export."""
    return [n for n in EXPORTED if not hasattr(_lib, n)]


def last_error() -> str:
    return _lib.lmx_last_error().decode()


def _check(rc, allow_info=False):
    if rc < 0:
        raise LmxError(rc, last_error())
    if rc > 0 and not allow_info:
        raise LmxError(rc, "unexpected info")
    return rc


def init(device: int = 0):
    _check(_lib.lmx_init(device))


def finalize():
    _check(_lib.lmx_finalize())


def version() -> int:
    return _lib.lmx_version()


def raw():
    """The ctypes handle of liblmx.so (for callers driving the C ABI directly)."""
    return _lib


def _f64(a, copy=True):
    return np.array(a, dtype=np.float64, order="C", copy=True) if copy else np.ascontiguousarray(a, np.float64)


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def _as_block(b):
    b = _f64(b)
    vec = b.ndim == 1
    if vec:
        b = b.reshape(-1, 1)
    return b, vec


# ---------------------------------------------------------------- host API

def dgemm(A, B, C=None, alpha=1.0, beta=0.0, transa="N", transb="N"):
    A = _f64(A, copy=False)
    B = _f64(B, copy=False)
    m = A.shape[1] if transa == "T" else A.shape[0]
    k = A.shape[0] if transa == "T" else A.shape[1]
    n = B.shape[0] if transb == "T" else B.shape[1]
    C = np.zeros((m, n)) if C is None else _f64(C)
    _check(_lib.lmx_dgemm(transa.encode(), transb.encode(), m, n, k, alpha, _ptr(A), max(A.shape[1], 1),
                          _ptr(B), max(B.shape[1], 1), beta, _ptr(C), max(n, 1)))
    return C


def sgemm(A, B, C=None, alpha=1.0, beta=0.0, transa="N", transb="N"):
    A = np.ascontiguousarray(A, np.float32)
    B = np.ascontiguousarray(B, np.float32)
    m = A.shape[1] if transa == "T" else A.shape[0]
    k = A.shape[0] if transa == "T" else A.shape[1]
    n = B.shape[0] if transb == "T" else B.shape[1]
    C = np.zeros((m, n), np.float32) if C is None else np.array(C, np.float32, order="C", copy=True)
    _check(_lib.lmx_sgemm(transa.encode(), transb.encode(), m, n, k, alpha, _ptr(A), max(A.shape[1], 1),
                          _ptr(B), max(B.shape[1], 1), beta, _ptr(C), max(n, 1)))
    return C


def getrf(A):
    """-> (packed LU, ipiv swap list (0-based), info)."""
    A = _f64(A)
    n = A.shape[0]
    ipiv = np.zeros(max(n, 1), np.int32)
    info = _check(_lib.lmx_dgetrf(n, _ptr(A), max(n, 1), _ptr(ipiv)), allow_info=True)
    return A, ipiv[:n], info


def sgetrf(A):
    """FP32 LU: -> (packed LU as float32, ipiv swap list (0-based), info)."""
    A = np.array(A, np.float32, order="C", copy=True)
    n = A.shape[0]
    ipiv = np.zeros(max(n, 1), np.int32)
    info = _check(_lib.lmx_sgetrf(n, _ptr(A), max(n, 1), _ptr(ipiv)), allow_info=True)
    return A, ipiv[:n], info


def getrs(LU, ipiv, B):
    LU = _f64(LU, copy=False)
    ipiv = np.ascontiguousarray(ipiv, np.int32)
    B, vec = _as_block(B)
    n, nrhs = B.shape
    _check(_lib.lmx_dgetrs(n, nrhs, _ptr(LU), max(n, 1), _ptr(ipiv), _ptr(B), max(nrhs, 1)))
    return B[:, 0] if vec else B


def gesv(A, B):
    """-> (X, packed LU, ipiv, info)."""
    A = _f64(A)
    B, vec = _as_block(B)
    n, nrhs = B.shape
    ipiv = np.zeros(max(n, 1), np.int32)
    info = _check(_lib.lmx_dgesv(n, nrhs, _ptr(A), max(n, 1), _ptr(ipiv), _ptr(B), max(nrhs, 1)),
                  allow_info=True)
    return (B[:, 0] if vec else B), A, ipiv[:n], info


def potrf(A):
    """-> (A with G in its lower triangle, info)."""
    A = _f64(A)
    n = A.shape[0]
    info = _check(_lib.lmx_dpotrf(n, _ptr(A), max(n, 1)), allow_info=True)
    return A, info


def potrs(G, B):
    G = _f64(G, copy=False)
    B, vec = _as_block(B)
    n, nrhs = B.shape
    _check(_lib.lmx_dpotrs(n, nrhs, _ptr(G), max(n, 1), _ptr(B), max(nrhs, 1)))
    return B[:, 0] if vec else B


def posv(A, B):
    A = _f64(A)
    B, vec = _as_block(B)
    n, nrhs = B.shape
    info = _check(_lib.lmx_dposv(n, nrhs, _ptr(A), max(n, 1), _ptr(B), max(nrhs, 1)), allow_info=True)
    return (B[:, 0] if vec else B), A, info


def trsm(T, B, uplo="L", trans="N", diag="N"):
    T = _f64(T, copy=False)
    B, vec = _as_block(B)
    n, nrhs = B.shape
    _check(_lib.lmx_dtrsm(uplo.encode(), trans.encode(), diag.encode(), n, nrhs, _ptr(T), max(n, 1),
                          _ptr(B), max(nrhs, 1)))
    return B[:, 0] if vec else B


def gesv_mixed(A, B, precision=PREC_F32):
    """-> (X, RefineInfo).  A and B are not modified."""
    A = _f64(A, copy=False)
    Bb, vec = _as_block(B)
    n, nrhs = Bb.shape
    X = np.zeros_like(Bb)
    ri = RefineInfo()
    _check(_lib.lmx_dgesv_mixed(precision, n, nrhs, _ptr(A), max(n, 1), _ptr(Bb), max(nrhs, 1), _ptr(X),
                                max(nrhs, 1), ctypes.byref(ri)), allow_info=True)
    return (X[:, 0] if vec else X), ri


def posv_mixed(A, B, precision=PREC_F32):
    A = _f64(A, copy=False)
    Bb, vec = _as_block(B)
    n, nrhs = Bb.shape
    X = np.zeros_like(Bb)
    ri = RefineInfo()
    _check(_lib.lmx_dposv_mixed(precision, n, nrhs, _ptr(A), max(n, 1), _ptr(Bb), max(nrhs, 1), _ptr(X),
                                max(nrhs, 1), ctypes.byref(ri)), allow_info=True)
    return (X[:, 0] if vec else X), ri


def swaps_to_perm(ipiv):
    """LAPACK swap list -> permutation vector p (row i of PA is row p[i] of A),
    the format of the reference's `lu` (Solve/LU.hs:121-167)."""
    n = len(ipiv)
    perm = np.arange(n)
    for k in range(n):
        p = int(ipiv[k])
        if p != k:
            perm[k], perm[p] = perm[p], perm[k]
    return perm


# ------------------------------------------------- reference-named surface

def matMul(a, b):
    """matMul (BLAS/Level3.hs:144-147; RULE to matMulP :296-300)."""
    return dgemm(a, b)


matMulP = matMul      # BLAS/Level3.hs:183-207
matMulPPar = matMul   # BLAS/Level3.hs:214-294: parallelism is the GPU grid


def gemm(alpha, a, b, beta, c):
    """gemm alpha a b beta c (BLAS/Level3.hs:102-131)."""
    return dgemm(a, b, c, alpha=alpha, beta=beta)


def lu(a):
    """lu (Solve/LU.hs:121-167) -> (packed LU, permutation vector)."""
    LU, ipiv, _ = getrf(a)
    return LU, swaps_to_perm(ipiv)


def luSolveP(a, b):
    """luSolveP (Solve/LU.hs:265-304)."""
    x, _, _, _ = gesv(a, b)
    return x


luSolve = luSolveP    # Solve/LU.hs:240-259


def cholesky(a):
    """cholesky (Solve/Cholesky.hs:114-147): lower G, upper triangle zero."""
    G, _ = potrf(a)
    return np.tril(G)


def choleskySolveP(a, b):
    """choleskySolveP (Solve/Cholesky.hs:241-273)."""
    x, _, _ = posv(a, b)
    return x


choleskySolve = choleskySolveP  # Solve/Cholesky.hs:229-235


def forwardSub(l, b):
    return trsm(l, b, "L", "N", "N")      # Solve/Triangular.hs:90-107


def backSub(u, b):
    return trsm(u, b, "U", "N", "N")      # Solve/Triangular.hs:136-153


def forwardSubUnit(l, b):
    return trsm(l, b, "L", "N", "U")      # Solve/Triangular.hs:187-201


def backSubUnit(u, b):
    return trsm(u, b, "U", "N", "U")      # Solve/Triangular.hs:235-249


# ---------------------------------------------------------------- device API (torch)

def _tp(t):
    return ctypes.c_void_p(t.data_ptr())


def _stream(stream):
    if stream is None:
        import torch
        return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    return ctypes.c_void_p(stream)


def dgemm_dev(A, B, C, alpha=1.0, beta=0.0, transa="N", transb="N", stream=None):
    m, n = C.shape
    k = A.shape[0] if transa == "T" else A.shape[1]
    _check(_lib.lmx_dgemm_dev(_stream(stream), transa.encode(), transb.encode(), m, n, k, alpha, _tp(A),
                              A.stride(0), _tp(B), B.stride(0), beta, _tp(C), C.stride(0)))


def sgemm_dev(A, B, C, alpha=1.0, beta=0.0, transa="N", transb="N", stream=None):
    m, n = C.shape
    k = A.shape[0] if transa == "T" else A.shape[1]
    _check(_lib.lmx_sgemm_dev(_stream(stream), transa.encode(), transb.encode(), m, n, k, alpha, _tp(A),
                              A.stride(0), _tp(B), B.stride(0), beta, _tp(C), C.stride(0)))


def getrf_dev(A, ipiv, info, stream=None):
    _check(_lib.lmx_dgetrf_dev(_stream(stream), A.shape[0], _tp(A), A.stride(0), _tp(ipiv), _tp(info)))


def sgetrf_dev(A, ipiv, info, stream=None):
    _check(_lib.lmx_sgetrf_dev(_stream(stream), A.shape[0], _tp(A), A.stride(0), _tp(ipiv), _tp(info)))


def getrs_dev(LU, ipiv, B, stream=None):
    _check(_lib.lmx_dgetrs_dev(_stream(stream), LU.shape[0], B.shape[1], _tp(LU), LU.stride(0), _tp(ipiv),
                               _tp(B), B.stride(0)))


def gesv_dev(A, ipiv, B, info, stream=None):
    _check(_lib.lmx_dgesv_dev(_stream(stream), A.shape[0], B.shape[1], _tp(A), A.stride(0), _tp(ipiv),
                              _tp(B), B.stride(0), _tp(info)))


def potrf_dev(A, info, stream=None):
    _check(_lib.lmx_dpotrf_dev(_stream(stream), A.shape[0], _tp(A), A.stride(0), _tp(info)))


def potrs_dev(G, B, stream=None):
    _check(_lib.lmx_dpotrs_dev(_stream(stream), G.shape[0], B.shape[1], _tp(G), G.stride(0), _tp(B),
                               B.stride(0)))


def posv_dev(A, B, info, stream=None):
    _check(_lib.lmx_dposv_dev(_stream(stream), A.shape[0], B.shape[1], _tp(A), A.stride(0), _tp(B),
                              B.stride(0), _tp(info)))


def trsm_dev(T, B, uplo="L", trans="N", diag="N", stream=None):
    _check(_lib.lmx_dtrsm_dev(_stream(stream), uplo.encode(), trans.encode(), diag.encode(), T.shape[0],
                              B.shape[1], _tp(T), T.stride(0), _tp(B), B.stride(0)))


def gesv_mixed_dev(A, B, X, precision=PREC_F32, stream=None):
    ri = RefineInfo()
    _check(_lib.lmx_dgesv_mixed_dev(_stream(stream), precision, A.shape[0], B.shape[1], _tp(A), A.stride(0),
                                    _tp(B), B.stride(0), _tp(X), X.stride(0), ctypes.byref(ri)), allow_info=True)
    return ri


def posv_mixed_dev(A, B, X, precision=PREC_F32, stream=None):
    ri = RefineInfo()
    _check(_lib.lmx_dposv_mixed_dev(_stream(stream), precision, A.shape[0], B.shape[1], _tp(A), A.stride(0),
                                    _tp(B), B.stride(0), _tp(X), X.stride(0), ctypes.byref(ri)), allow_info=True)
    return ri


# ---------------------------------------------------------------- batched solves

def batch_layout(shapes):
    """Packs problems of the given (kind, n, nrhs) shapes back to back:
    -> (list of BatchProblem, a_pool elements, b_pool elements, ipiv elements)."""
    probs, ao, bo, po = [], 0, 0, 0
    for kind, n, nrhs in shapes:
        probs.append(BatchProblem(kind, n, nrhs, 0, ao, bo, po))
        ao += n * n
        bo += n * nrhs
        po += n
    return probs, ao, bo, po


class BatchPlan:
    """A plan for one batch of independent LU / Cholesky solves
    (lmx_batch_plan_*); execute() runs them on torch tensors in HBM."""

    def __init__(self, problems, a_elems, b_elems, ipiv_elems=0, precision=PREC_F64, flags=0):
        arr = (BatchProblem * max(len(problems), 1))(*problems)
        h = _vp()
        _check(_lib.lmx_batch_plan_create(ctypes.byref(h), arr, len(problems), a_elems, b_elems, ipiv_elems,
                                          precision, flags))
        self._h = h
        self.count = len(problems)

    def execute(self, a_pool, b_pool, info, ipiv_pool=None, iterations=None, stream=None):
        _check(_lib.lmx_batch_plan_execute(self._h, _tp(a_pool), _tp(b_pool),
                                           _tp(ipiv_pool) if ipiv_pool is not None else None, _tp(info),
                                           _tp(iterations) if iterations is not None else None,
                                           _stream(stream)))

    def autotune(self, stream=None):
        _check(_lib.lmx_batch_plan_autotune(self._h, _stream(stream)))

    def query(self):
        v, nb, nt, wg = _i32(), _i32(), _i32(), _i32()
        fl = _dbl()
        ms = (_dbl * 4)()
        _check(_lib.lmx_batch_plan_query(self._h, ctypes.byref(v), ctypes.byref(nb), ctypes.byref(nt),
                                         ctypes.byref(wg), ctypes.byref(fl), ms))
        return {"variant": v.value, "block_cols": nb.value, "threads": nt.value, "workgroups": wg.value,
                "flops": fl.value, "tuned_ms": list(ms)}

    def close(self):
        if getattr(self, "_h", None):
            _lib.lmx_batch_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def solve_batched(problems, precision=PREC_F64, flags=0, return_factors=False):
    """Solves a list of (kind, A, B) numpy problems in one batched launch.
    -> (list of X, info array, iterations array[, list of (A_out, ipiv)])."""
    import torch
    shapes = [(k, A.shape[0], B.shape[1]) for k, A, B in problems]
    probs, na, nb, npv = batch_layout(shapes)
    a = np.concatenate([np.ascontiguousarray(A, dtype=np.float64).ravel() for _, A, _ in problems])
    b = np.concatenate([np.ascontiguousarray(B, dtype=np.float64).ravel() for _, _, B in problems])
    dA = torch.tensor(a, device="cuda")
    dB = torch.tensor(b, device="cuda")
    dP = torch.zeros(max(npv, 1), dtype=torch.int32, device="cuda")
    info = torch.full((len(problems),), -7, dtype=torch.int32, device="cuda")
    iters = torch.full((len(problems),), -7, dtype=torch.int32, device="cuda")
    plan = BatchPlan(probs, na, nb, npv, precision, flags)
    try:
        plan.execute(dA, dB, info, dP, iters)
        torch.cuda.synchronize()
    finally:
        plan.close()
    bh = dB.cpu().numpy()
    X = [bh[p.b_offset:p.b_offset + p.n * p.nrhs].reshape(p.n, p.nrhs) for p in probs]
    out = (X, info.cpu().numpy(), iters.cpu().numpy())
    if return_factors:
        ah, ph = dA.cpu().numpy(), dP.cpu().numpy()
        fac = [(ah[p.a_offset:p.a_offset + p.n * p.n].reshape(p.n, p.n), ph[p.ipiv_offset:p.ipiv_offset + p.n])
               for p in probs]
        out = out + (fac,)
    return out


# ---------------------------------------------------------------- solve server

class Server:
    """The request loop of lmx_server_* on host numpy buffers.  submit()
    keeps the buffers alive until the request's result is collected."""

    def __init__(self, budget_bytes=0, slots=4, small_n=0, small_max=0, aging_ms=0.0):
        cfg = ServerConfig(budget_bytes, slots, small_n, small_max, 0, aging_ms)
        h = _vp()
        _check(_lib.lmx_server_create(ctypes.byref(h), ctypes.byref(cfg)))
        self._h = h
        self._keep = {}

    def submit(self, kind, B, X, A=None, factor=0, keep=False):
        """A x = B for host arrays (row-major float64; X is written).  Either
        A or a resident factor handle of the same kind and order."""
        n, nrhs = B.shape
        for arr in (B, X) + ((A,) if A is not None else ()):
            if arr.dtype != np.float64 or not arr.flags["C_CONTIGUOUS"]:
                raise ValueError("server buffers must be C-contiguous float64")
        rid = _i64()
        _check(_lib.lmx_server_submit(self._h, kind, n, nrhs, _ptr(A) if A is not None else None,
                                      A.shape[1] if A is not None else 0, factor, _ptr(B), nrhs, _ptr(X), nrhs,
                                      REQ_KEEP_FACTOR if keep else 0, ctypes.byref(rid)))
        self._keep[rid.value] = (A, B, X)
        return rid.value

    def wait(self, rid):
        res = RequestResult()
        _check(_lib.lmx_server_wait(self._h, rid, ctypes.byref(res)))
        self._keep.pop(rid, None)
        return res

    def poll(self, rid):
        res = RequestResult()
        rc = _lib.lmx_server_poll(self._h, rid, ctypes.byref(res))
        if rc == 1:
            return None
        _check(rc)
        self._keep.pop(rid, None)
        return res

    def release(self, factor):
        _check(_lib.lmx_server_factor_release(self._h, factor))

    def stats(self, reset=False):
        st = ServerCounters()
        _check(_lib.lmx_server_stats(self._h, ctypes.byref(st), 1 if reset else 0))
        return {f: getattr(st, f) for f, _ in ServerCounters._fields_}

    def close(self):
        if getattr(self, "_h", None):
            _check(_lib.lmx_server_destroy(self._h))
            self._h = None
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- kernel timing

def kernel_timing(enable: bool):
    _check(_lib.lmx_kernel_timing(1 if enable else 0))


def kernel_timing_reset():
    _check(_lib.lmx_kernel_timing_reset())


def kernel_timing_get(family: str):
    """-> (launch count, total milliseconds, total algorithmic work)."""
    cnt = _i64(0)
    ms = _dbl(0)
    work = _dbl(0)
    _check(_lib.lmx_kernel_timing_get(family.encode(), ctypes.byref(cnt), ctypes.byref(ms), ctypes.byref(work)))
    return cnt.value, ms.value, work.value
