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

import atexit
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
    "lmx_dlu", "lmx_dlu_dev", "lmx_slu", "lmx_slu_dev",
    "lmx_gemm_bf16", "lmx_gemm_bf16_dev", "lmx_sgemm_bf16x3_dev", "lmx_gemm_tile_override",
    "lmx_dgemv", "lmx_dgemv_dev", "lmx_dnorm", "lmx_dnorm_dev", "lmx_dresidual_dev", "lmx_dscaled_residual",
    "lmx_dscaled_residual_dev", "lmx_dscaled_residual_noalloc_dev", "lmx_dsyrk", "lmx_dsyrk_dev", "lmx_ddet", "lmx_ddet_dev",
    "lmx_dinverse_iteration", "lmx_dinverse_iteration_dev",
    "lmx_factor_create", "lmx_factor_create_dev", "lmx_factor_solve", "lmx_factor_solve_dev", "lmx_factor_query",
    "lmx_factor_release",
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
    "lmx_gemm_bf16": [_chr, _chr, _i64, _i64, _i64, _flt, _vp, _i64, _vp, _i64, _flt, _vp, _i64],
    "lmx_gemm_bf16_dev": [_vp, _chr, _chr, _i64, _i64, _i64, _flt, _vp, _i64, _vp, _i64, _flt, _vp, _i64],
    "lmx_sgemm_bf16x3_dev": [_vp, _chr, _chr, _i64, _i64, _i64, _flt, _vp, _i64, _vp, _i64, _flt, _vp, _i64],
    "lmx_gemm_tile_override": [ctypes.c_int],
    "lmx_dgemv": [_i64, _i64, _dbl, _vp, _i64, _vp, _dbl, _vp],
    "lmx_dgemv_dev": [_vp, _i64, _i64, _dbl, _vp, _i64, _vp, _dbl, _vp],
    "lmx_dnorm": [_chr, _i64, _i64, _vp, _i64, ctypes.POINTER(_dbl)],
    "lmx_dnorm_dev": [_vp, _chr, _i64, _i64, _vp, _i64, _vp],
    "lmx_dresidual_dev": [_vp, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64, _vp, _i64],
    "lmx_dscaled_residual": [_i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64, _vp],
    "lmx_dscaled_residual_dev": [_vp, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64, _vp],
    "lmx_dsyrk": [_chr, _i64, _i64, _dbl, _vp, _i64, _dbl, _vp, _i64],
    "lmx_dsyrk_dev": [_vp, _chr, _i64, _i64, _dbl, _vp, _i64, _dbl, _vp, _i64],
    "lmx_ddet": [_i64, _vp, _i64, ctypes.POINTER(_dbl)],
    "lmx_ddet_dev": [_vp, _i64, _vp, _i64, _vp],
    "lmx_dinverse_iteration": [_i64, _vp, _i64, _dbl, _vp, _i32, _dbl, ctypes.POINTER(_dbl), _vp,
                               ctypes.POINTER(_i32)],
    "lmx_dinverse_iteration_dev": [_vp, _i64, _vp, _i64, _dbl, _vp, _i32, _dbl, ctypes.POINTER(_dbl), _vp,
                                   ctypes.POINTER(_i32)],
    "lmx_factor_create": [ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_int, _i64, _vp, _i64],
    "lmx_factor_create_dev": [ctypes.POINTER(_vp), _vp, ctypes.c_int, ctypes.c_int, _i64, _vp, _i64],
    "lmx_factor_solve": [_vp, _i64, _vp, _i64, _vp, _i64, _pRI],
    "lmx_factor_solve_dev": [_vp, _vp, _i64, _vp, _i64, _vp, _i64, _pRI],
    "lmx_factor_query": [_vp, ctypes.POINTER(_i32), ctypes.POINTER(_i64), ctypes.POINTER(_i32),
                         ctypes.POINTER(_i32), ctypes.POINTER(_i64)],
    "lmx_factor_release": [_vp],
    "lmx_dlu": [_i64, _vp, _i64, _vp],
    "lmx_dlu_dev": [_vp, _i64, _vp, _i64, _vp, _vp],
    "lmx_slu": [_i64, _vp, _i64, _vp],
    "lmx_slu_dev": [_vp, _i64, _vp, _i64, _vp, _vp],
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


def _finalize_at_exit():
    # The library's contexts (a stream, a memory pool that keeps freed
    # memory, staging buffers) are released here, while the HIP runtime and
    # torch are still whole; left to process exit they would be torn down
    # after (or while) the runtime and any profiler attached to it unload.
    try:
        _lib.lmx_finalize()
    except Exception:
        pass


atexit.register(_finalize_at_exit)  # registered after torch's: runs before them


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


def to_bf16_bits(x):
    """float32 -> BF16 bit patterns (uint16), rounding to nearest even."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    return (((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF).astype(np.uint16)


def from_bf16_bits(h):
    """BF16 bit patterns (uint16) -> float32."""
    return (np.ascontiguousarray(h, np.uint16).astype(np.uint32) << 16).view(np.float32)


def gemm_bf16(A, B, C=None, alpha=1.0, beta=0.0, transa="N", transb="N"):
    """C (float32) = alpha op(A) op(B) + beta C with A, B given as BF16 bit
    patterns (uint16 arrays), on BF16 matrix cores with FP32 accumulation."""
    A = np.ascontiguousarray(A, np.uint16)
    B = np.ascontiguousarray(B, np.uint16)
    m = A.shape[1] if transa == "T" else A.shape[0]
    k = A.shape[0] if transa == "T" else A.shape[1]
    n = B.shape[0] if transb == "T" else B.shape[1]
    C = np.zeros((m, n), np.float32) if C is None else np.array(C, np.float32, order="C", copy=True)
    _check(_lib.lmx_gemm_bf16(transa.encode(), transb.encode(), m, n, k, alpha, _ptr(A), max(A.shape[1], 1),
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
    """lu (Solve/LU.hs:121-167) -> (packed LU, permutation vector), with lu's
    own pivot rule: the last row of largest |a_ik| wins a tie (maximumBy,
    :128-135); the pivots of luSolveP (getrf) take the first."""
    A = _f64(a)
    n = A.shape[0]
    perm = np.zeros(max(n, 1), np.int32)
    _check(_lib.lmx_dlu(n, _ptr(A), max(n, 1), _ptr(perm)), allow_info=True)
    return A, perm[:n].astype(np.int64)


def lu_float(a):
    """lu at e = Float: -> (packed LU as float32, permutation vector)."""
    A = np.array(a, np.float32, order="C", copy=True)
    n = A.shape[0]
    perm = np.zeros(max(n, 1), np.int32)
    _check(_lib.lmx_slu(n, _ptr(A), max(n, 1), _ptr(perm)), allow_info=True)
    return A, perm[:n].astype(np.int64)


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


# ---------------------------------------------------------------- resident factors

FACTOR_LU, FACTOR_CHOLESKY = 0, 1


class Factor:
    """A factor resident in HBM (lmx_factor_*): factor once, solve many
    times.  precision PREC_F64 keeps the FP64 factor; PREC_F32 / PREC_BF16
    keep the low-precision factor and refine every solve to FP64 accuracy.
    `a` is a numpy matrix (host) or a CUDA float64 tensor (device)."""

    def __init__(self, a, kind=FACTOR_LU, precision=PREC_F64, stream=None):
        h = _vp()
        try:
            import torch
            dev = isinstance(a, torch.Tensor)
        except ImportError:
            dev = False
        if dev:
            n = _square(a, "A", "f64")
            rc = _lib.lmx_factor_create_dev(ctypes.byref(h), _stream(stream), kind, precision, n, _tp(a), _ld(a))
        else:
            A = _f64(a, copy=False)
            n = A.shape[0]
            rc = _lib.lmx_factor_create(ctypes.byref(h), kind, precision, n, _ptr(A), max(n, 1))
        self.info = _check(rc, allow_info=True)
        self._h = h
        self.n, self.kind, self.precision = n, kind, precision

    def solve(self, b):
        """Host solve: b is a vector or an n x nrhs block -> (x, RefineInfo)."""
        B, vec = _as_block(b)
        if B.shape[0] != self.n:
            raise ValueError(f"b has {B.shape[0]} rows, the factor is of order {self.n}")
        k = B.shape[1]
        X = np.zeros_like(B)
        ri = RefineInfo()
        _check(_lib.lmx_factor_solve(self._h, k, _ptr(B), max(k, 1), _ptr(X), max(k, 1), ctypes.byref(ri)),
               allow_info=True)
        return (X[:, 0] if vec else X), ri

    def solve_dev(self, B, X, stream=None):
        _dev_mat(B, "B", "f64", rows=self.n)
        _dev_mat(X, "X", "f64", rows=self.n, cols=B.shape[1])
        ri = RefineInfo()
        _check(_lib.lmx_factor_solve_dev(self._h, _stream(stream), B.shape[1], _tp(B), _ld(B), _tp(X), _ld(X),
                                         ctypes.byref(ri)), allow_info=True)
        return ri

    def query(self):
        info, n, kind, prec, nbytes = _i32(), _i64(), _i32(), _i32(), _i64()
        _check(_lib.lmx_factor_query(self._h, ctypes.byref(info), ctypes.byref(n), ctypes.byref(kind),
                                     ctypes.byref(prec), ctypes.byref(nbytes)))
        return {"info": info.value, "n": n.value, "kind": kind.value, "precision": prec.value,
                "bytes": nbytes.value}

    def close(self):
        if getattr(self, "_h", None):
            _check(_lib.lmx_factor_release(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def luFactor(a, precision=PREC_F64):
    """An LU factor kept on the device (no reference counterpart: lu,
    Solve/LU.hs:121, returns a factor nothing reuses)."""
    return Factor(a, FACTOR_LU, precision)


def luSolveWith(f, b):
    """x with A x = b for the factored A (a vector or an n x nrhs block)."""
    return f.solve(b)[0]


def choleskyFactor(a, precision=PREC_F64):
    return Factor(a, FACTOR_CHOLESKY, precision)


choleskySolveWith = luSolveWith


# ------------------------------------- level 2, norms, SYRK, det, inverse iteration

def gemv(alpha, a, x, beta, y):
    """gemv alpha a x beta y (BLAS/Level2.hs:92-100) -> alpha a x + beta y
    (beta = 0 does not read y)."""
    A = _f64(a, copy=False)
    x = _f64(x)
    y = _f64(y)
    m, n = A.shape
    _check(_lib.lmx_dgemv(m, n, alpha, _ptr(A), max(n, 1), _ptr(x), beta, _ptr(y)))
    return y


def matvec(a, x):
    """matvec / matvecP (BLAS/Level2.hs:123-144)."""
    A = _f64(a, copy=False)
    return gemv(1.0, A, x, 0.0, np.zeros(A.shape[0]))


matvecP = matvec


def norm(a, which):
    """A matrix norm: '1' (norm1), 'I' (normInf), 'F' (normFrob), 'M' (max |a_ij|)."""
    A = _f64(np.atleast_2d(a), copy=False)
    m, n = A.shape
    out = _dbl(0.0)
    _check(_lib.lmx_dnorm(which.encode(), m, n, _ptr(A), max(n, 1), ctypes.byref(out)))
    return out.value


def norm1(a):
    return norm(a, "1")          # Norms.hs:124-130


def normInf(a):
    return norm(a, "I")          # Norms.hs:142-148


def normFrob(a):
    return norm(a, "F")          # Norms.hs:110-112


def vnorm1(x):
    return norm(np.reshape(x, (-1, 1)), "1")     # Norms.hs:72-74


def vnorm2(x):
    return norm(np.reshape(x, (-1, 1)), "F")     # Norms.hs:83-85


def vnormInf(x):
    return norm(np.reshape(x, (-1, 1)), "I")     # Norms.hs:94-96


def scaled_residual_linear(a, x, b):
    """scaledResidualLinear (test/Test/Residuals.hs:69-78), per column of x / b."""
    A = _f64(a, copy=False)
    X, vec = _as_block(x)
    B, _ = _as_block(b)
    n, k = X.shape
    eta = np.zeros(max(k, 1))
    _check(_lib.lmx_dscaled_residual(n, k, _ptr(A), max(n, 1), _ptr(X), max(k, 1), _ptr(B), max(k, 1), _ptr(eta)))
    return float(eta[0]) if vec else eta[:k]


def syrk(a, trans="N", alpha=1.0, beta=0.0, c=None):
    """C = alpha op(a) op(a)^T + beta C, the whole symmetric matrix."""
    A = _f64(a, copy=False)
    n, k = (A.shape[1], A.shape[0]) if trans == "T" else A.shape
    C = np.zeros((n, n)) if c is None else _f64(c)
    _check(_lib.lmx_dsyrk(trans.encode(), n, k, alpha, _ptr(A), max(A.shape[1], 1), beta, _ptr(C), max(n, 1)))
    return C


def matMulAtAP(a):
    """matMulAtAP (BLAS/Level3.hs:353-373): A^T A."""
    return syrk(a, "T")


def det(a):
    """det (Solve/LU.hs:534-558)."""
    A = _f64(a, copy=False)
    out = _dbl(0.0)
    _check(_lib.lmx_ddet(A.shape[0], _ptr(A), max(A.shape[0], 1), ctypes.byref(out)))
    return out.value


def inverseIteration(a, mu, q0, max_iter, tol):
    """inverseIteration (Eigen/Power.hs:113-135) -> (lambda, q, steps)."""
    A = _f64(a, copy=False)
    q0 = _f64(q0)
    n = A.shape[0]
    q = np.zeros(max(n, 1))
    lam = _dbl(0.0)
    it = _i32(0)
    _check(_lib.lmx_dinverse_iteration(n, _ptr(A), max(n, 1), mu, _ptr(q0), max_iter, tol, ctypes.byref(lam),
                                       _ptr(q), ctypes.byref(it)))
    return lam.value, q[:n], it.value


# ---------------------------------------------------------------- device API (torch)

def _tp(t):
    return ctypes.c_void_p(t.data_ptr())


def _stream(stream):
    if stream is None:
        import torch
        return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    return ctypes.c_void_p(stream)


def _torch_dtype(name):
    import torch
    return {"f64": torch.float64, "f32": torch.float32, "i32": torch.int32, "bf16": torch.bfloat16}[name]


def _dev_mat(t, name, dtype, rows=None, cols=None):
    """Checks a device matrix argument before its pointer reaches the
    library: a CUDA tensor of the given dtype, 2-D, row-major with unit
    column stride and a row stride of at least its column count.  A wrong
    shape or a transposed view would otherwise make the kernels read or
    write outside the tensor."""
    import torch
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name}: expected a torch tensor")
    if not t.is_cuda:
        raise ValueError(f"{name}: must be a CUDA tensor")
    if dtype is not None and t.dtype != _torch_dtype(dtype):
        raise ValueError(f"{name}: dtype {t.dtype}, expected {_torch_dtype(dtype)}")
    if t.dim() != 2:
        raise ValueError(f"{name}: expected a 2-D matrix, got shape {tuple(t.shape)}")
    if t.shape[1] > 1 and t.stride(1) != 1:
        raise ValueError(f"{name}: columns must be contiguous (stride(1) == 1); pass .contiguous()")
    if t.shape[0] > 1 and t.stride(0) < t.shape[1]:
        raise ValueError(f"{name}: row stride {t.stride(0)} below the column count {t.shape[1]}")
    if rows is not None and t.shape[0] != rows:
        raise ValueError(f"{name}: {t.shape[0]} rows, expected {rows}")
    if cols is not None and t.shape[1] != cols:
        raise ValueError(f"{name}: {t.shape[1]} columns, expected {cols}")
    return t


def _ld(t):
    return max(t.stride(0), t.shape[1], 1) if t.shape[0] > 1 else max(t.shape[1], 1)


def _dev_vec(t, name, dtype, length):
    """Checks a device vector argument: a contiguous CUDA tensor of the
    given dtype with at least `length` elements."""
    import torch
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name}: expected a torch tensor")
    if not t.is_cuda:
        raise ValueError(f"{name}: must be a CUDA tensor")
    if t.dtype != _torch_dtype(dtype):
        raise ValueError(f"{name}: dtype {t.dtype}, expected {_torch_dtype(dtype)}")
    if not t.is_contiguous():
        raise ValueError(f"{name}: must be contiguous")
    if t.numel() < length:
        raise ValueError(f"{name}: {t.numel()} elements, need {length}")
    return t


def _square(A, name, dtype):
    _dev_mat(A, name, dtype)
    if A.shape[0] != A.shape[1]:
        raise ValueError(f"{name}: must be square, got {tuple(A.shape)}")
    return A.shape[0]


def _gemm_shapes(A, B, C, transa, transb, dtype):
    """Checks op(A) op(B) -> C; dtype None leaves A and B's dtype to the caller
    (C is then float32)."""
    for t, nm in ((A, "A"), (B, "B")):
        _dev_mat(t, nm, dtype)
    _dev_mat(C, "C", dtype or "f32")
    if transa not in ("N", "T") or transb not in ("N", "T"):
        raise ValueError("trans flags must be 'N' or 'T'")
    m, n = C.shape
    am, ak = (A.shape[1], A.shape[0]) if transa == "T" else tuple(A.shape)
    bk, bn = (B.shape[1], B.shape[0]) if transb == "T" else tuple(B.shape)
    if am != m or bn != n or ak != bk:
        raise ValueError(f"gemm: op(A) {am}x{ak}, op(B) {bk}x{bn}, C {m}x{n} do not conform")
    return m, n, ak


def dgemm_dev(A, B, C, alpha=1.0, beta=0.0, transa="N", transb="N", stream=None):
    m, n, k = _gemm_shapes(A, B, C, transa, transb, "f64")
    _check(_lib.lmx_dgemm_dev(_stream(stream), transa.encode(), transb.encode(), m, n, k, alpha, _tp(A),
                              _ld(A), _tp(B), _ld(B), beta, _tp(C), _ld(C)))


def sgemm_dev(A, B, C, alpha=1.0, beta=0.0, transa="N", transb="N", stream=None):
    m, n, k = _gemm_shapes(A, B, C, transa, transb, "f32")
    _check(_lib.lmx_sgemm_dev(_stream(stream), transa.encode(), transb.encode(), m, n, k, alpha, _tp(A),
                              _ld(A), _tp(B), _ld(B), beta, _tp(C), _ld(C)))


def gemm_bf16_dev(A, B, C, alpha=1.0, beta=0.0, transa="N", transb="N", stream=None):
    """C (float32) = alpha op(A) op(B) + beta C, A and B torch.bfloat16."""
    import torch
    for t, nm in ((A, "A"), (B, "B")):
        if not isinstance(t, torch.Tensor) or t.dtype != torch.bfloat16:
            raise ValueError(f"{nm}: expected a torch.bfloat16 tensor")
    m, n, k = _gemm_shapes(A, B, C, transa, transb, None)
    _check(_lib.lmx_gemm_bf16_dev(_stream(stream), transa.encode(), transb.encode(), m, n, k, alpha, _tp(A),
                                  _ld(A), _tp(B), _ld(B), beta, _tp(C), _ld(C)))


def sgemm_bf16x3_dev(A, B, C, alpha=1.0, beta=0.0, transa="N", transb="N", stream=None):
    """C = alpha op(A) op(B) + beta C for float32 tensors on BF16 matrix cores,
    each operand split in two BF16 parts (hi*hi + hi*lo + lo*hi)."""
    m, n, k = _gemm_shapes(A, B, C, transa, transb, "f32")
    _check(_lib.lmx_sgemm_bf16x3_dev(_stream(stream), transa.encode(), transb.encode(), m, n, k, alpha, _tp(A),
                                     _ld(A), _tp(B), _ld(B), beta, _tp(C), _ld(C)))


def gemm_tile_override(tile):
    """Forces the calling thread's plain GEMMs onto one tile shape (1 = 64x64,
    2 = 128x64, 3 = 128x128); 0 returns to the tuned table."""
    _check(_lib.lmx_gemm_tile_override(tile))


def getrf_dev(A, ipiv, info, stream=None):
    n = _square(A, "A", "f64")
    _dev_vec(ipiv, "ipiv", "i32", n)
    _dev_vec(info, "info", "i32", 1)
    _check(_lib.lmx_dgetrf_dev(_stream(stream), n, _tp(A), _ld(A), _tp(ipiv), _tp(info)))


def sgetrf_dev(A, ipiv, info, stream=None):
    n = _square(A, "A", "f32")
    _dev_vec(ipiv, "ipiv", "i32", n)
    _dev_vec(info, "info", "i32", 1)
    _check(_lib.lmx_sgetrf_dev(_stream(stream), n, _tp(A), _ld(A), _tp(ipiv), _tp(info)))


def lu_dev(A, perm, info, stream=None):
    """The generic lu on a device matrix: A <- packed factors, perm <- lu's
    permutation vector (int32), with lu's tie rule."""
    n = _square(A, "A", "f64")
    _dev_vec(perm, "perm", "i32", n)
    _dev_vec(info, "info", "i32", 1)
    _check(_lib.lmx_dlu_dev(_stream(stream), n, _tp(A), _ld(A), _tp(perm), _tp(info)))


def getrs_dev(LU, ipiv, B, stream=None):
    n = _square(LU, "LU", "f64")
    _dev_mat(B, "B", "f64", rows=n)
    _dev_vec(ipiv, "ipiv", "i32", n)
    _check(_lib.lmx_dgetrs_dev(_stream(stream), n, B.shape[1], _tp(LU), _ld(LU), _tp(ipiv), _tp(B), _ld(B)))


def gesv_dev(A, ipiv, B, info, stream=None):
    n = _square(A, "A", "f64")
    _dev_mat(B, "B", "f64", rows=n)
    _dev_vec(ipiv, "ipiv", "i32", n)
    _dev_vec(info, "info", "i32", 1)
    _check(_lib.lmx_dgesv_dev(_stream(stream), n, B.shape[1], _tp(A), _ld(A), _tp(ipiv), _tp(B), _ld(B),
                              _tp(info)))


def potrf_dev(A, info, stream=None):
    n = _square(A, "A", "f64")
    _dev_vec(info, "info", "i32", 1)
    _check(_lib.lmx_dpotrf_dev(_stream(stream), n, _tp(A), _ld(A), _tp(info)))


def potrs_dev(G, B, stream=None):
    n = _square(G, "G", "f64")
    _dev_mat(B, "B", "f64", rows=n)
    _check(_lib.lmx_dpotrs_dev(_stream(stream), n, B.shape[1], _tp(G), _ld(G), _tp(B), _ld(B)))


def posv_dev(A, B, info, stream=None):
    n = _square(A, "A", "f64")
    _dev_mat(B, "B", "f64", rows=n)
    _dev_vec(info, "info", "i32", 1)
    _check(_lib.lmx_dposv_dev(_stream(stream), n, B.shape[1], _tp(A), _ld(A), _tp(B), _ld(B), _tp(info)))


def trsm_dev(T, B, uplo="L", trans="N", diag="N", stream=None):
    n = _square(T, "T", "f64")
    _dev_mat(B, "B", "f64", rows=n)
    if uplo not in ("L", "U") or trans not in ("N", "T") or diag not in ("N", "U"):
        raise ValueError("trsm flags: uplo L/U, trans N/T, diag N/U")
    _check(_lib.lmx_dtrsm_dev(_stream(stream), uplo.encode(), trans.encode(), diag.encode(), n, B.shape[1],
                              _tp(T), _ld(T), _tp(B), _ld(B)))


def gemv_dev(A, x, y, alpha=1.0, beta=0.0, stream=None):
    _dev_mat(A, "A", "f64")
    m, n = A.shape
    _dev_vec(x, "x", "f64", n)
    _dev_vec(y, "y", "f64", m)
    _check(_lib.lmx_dgemv_dev(_stream(stream), m, n, alpha, _tp(A), _ld(A), _tp(x), beta, _tp(y)))


def norm_dev(A, which, out, stream=None):
    _dev_mat(A, "A", "f64")
    _dev_vec(out, "out", "f64", 1)
    _check(_lib.lmx_dnorm_dev(_stream(stream), which.encode(), A.shape[0], A.shape[1], _tp(A), _ld(A), _tp(out)))


def residual_dev(A, X, B, R, stream=None):
    n = _square(A, "A", "f64")
    _dev_mat(X, "X", "f64", rows=n)
    k = X.shape[1]
    _dev_mat(B, "B", "f64", rows=n, cols=k)
    _dev_mat(R, "R", "f64", rows=n, cols=k)
    _check(_lib.lmx_dresidual_dev(_stream(stream), n, k, _tp(A), _ld(A), _tp(X), _ld(X), _tp(B), _ld(B), _tp(R),
                                  _ld(R)))


def scaled_residual_dev(A, X, B, eta, stream=None):
    n = _square(A, "A", "f64")
    _dev_mat(X, "X", "f64", rows=n)
    k = X.shape[1]
    _dev_mat(B, "B", "f64", rows=n, cols=k)
    _dev_vec(eta, "eta", "f64", k)
    _check(_lib.lmx_dscaled_residual_dev(_stream(stream), n, k, _tp(A), _ld(A), _tp(X), _ld(X), _tp(B), _ld(B),
                                         _tp(eta)))


def syrk_dev(A, C, trans="N", alpha=1.0, beta=0.0, stream=None):
    _dev_mat(A, "A", "f64")
    n, k = (A.shape[1], A.shape[0]) if trans == "T" else tuple(A.shape)
    _dev_mat(C, "C", "f64", rows=n, cols=n)
    _check(_lib.lmx_dsyrk_dev(_stream(stream), trans.encode(), n, k, alpha, _tp(A), _ld(A), beta, _tp(C), _ld(C)))


def det_dev(A, out, stream=None):
    n = _square(A, "A", "f64")
    _dev_vec(out, "out", "f64", 1)
    _check(_lib.lmx_ddet_dev(_stream(stream), n, _tp(A), _ld(A), _tp(out)))


def _mixed_args(A, B, X):
    n = _square(A, "A", "f64")
    _dev_mat(B, "B", "f64", rows=n)
    _dev_mat(X, "X", "f64", rows=n, cols=B.shape[1])
    return n, B.shape[1]


def gesv_mixed_dev(A, B, X, precision=PREC_F32, stream=None):
    n, k = _mixed_args(A, B, X)
    ri = RefineInfo()
    _check(_lib.lmx_dgesv_mixed_dev(_stream(stream), precision, n, k, _tp(A), _ld(A), _tp(B), _ld(B), _tp(X),
                                    _ld(X), ctypes.byref(ri)), allow_info=True)
    return ri


def posv_mixed_dev(A, B, X, precision=PREC_F32, stream=None):
    n, k = _mixed_args(A, B, X)
    ri = RefineInfo()
    _check(_lib.lmx_dposv_mixed_dev(_stream(stream), precision, n, k, _tp(A), _ld(A), _tp(B), _ld(B), _tp(X),
                                    _ld(X), ctypes.byref(ri)), allow_info=True)
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
