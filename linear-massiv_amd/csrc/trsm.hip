// Triangular solves with many right-hand sides.
//
// Leaf kernel: one triangle of order <= 64 is staged in LDS (padded to a
// 64 x 64 identity); every thread owns one right-hand-side vector in 64
// registers and runs the substitution fully unrolled, reading the triangle
// by broadcast from LDS.  A vector is a column of B for a left-side solve
// (op(T) X = B) or a row of B for a right-side solve (X op(T) = B).
//
// Blocked driver: 64-row blocks in dependency order; each block is solved
// by the leaf kernel and the not-yet-solved rows are updated with the MFMA
// GEMM (right-looking).
#include "common.hpp"
#include "kernels.hpp"

namespace lmx {

namespace {

constexpr int LEAF = 64;

// Vector equation M x = b solved per thread, M = T (MT = false) or T^T.
// M is lower triangular when MLOWER.
template <class T, bool MT, bool MLOWER, bool UNIT, bool RIGHT>
__global__ __launch_bounds__(64) void trsm_leaf_kernel(int nb, int nvec, const T *__restrict__ A,
                                                       int64_t lda, T *__restrict__ B, int64_t ldb) {
    __shared__ T Ms[LEAF][LEAF + 1];
    const int tid = threadIdx.x;
    for (int e = tid; e < LEAF * LEAF; e += 64) {
        int r = e / LEAF, c = e % LEAF;
        T val = T(0);
        if (r < nb && c < nb) {
            bool in = MLOWER ? (c <= r) : (c >= r);
            if (in) val = MT ? A[(int64_t)c * lda + r] : A[(int64_t)r * lda + c];
            if (r == c && UNIT) val = T(1);
        } else if (r == c) {
            val = T(1);
        }
        Ms[r][c] = val;
    }
    __syncthreads();
    const int v = blockIdx.x * 64 + tid;
    if (v >= nvec) return;
    T x[LEAF];
#pragma unroll
    for (int r = 0; r < LEAF; ++r) {
        T b = T(0);
        if (r < nb) b = RIGHT ? B[(int64_t)v * ldb + r] : B[(int64_t)r * ldb + v];
        x[r] = b;
    }
    if (MLOWER) {
#pragma unroll
        for (int r = 0; r < LEAF; ++r) {
            T xr = UNIT ? x[r] : x[r] / Ms[r][r];
            x[r] = xr;
#pragma unroll
            for (int s = r + 1; s < LEAF; ++s) x[s] = fma(-Ms[s][r], xr, x[s]);
        }
    } else {
#pragma unroll
        for (int r = LEAF - 1; r >= 0; --r) {
            T xr = UNIT ? x[r] : x[r] / Ms[r][r];
            x[r] = xr;
#pragma unroll
            for (int s = 0; s < r; ++s) x[s] = fma(-Ms[s][r], xr, x[s]);
        }
    }
#pragma unroll
    for (int r = 0; r < LEAF; ++r) {
        if (r < nb) {
            if (RIGHT)
                B[(int64_t)v * ldb + r] = x[r];
            else
                B[(int64_t)r * ldb + v] = x[r];
        }
    }
}

template <class T, bool MT, bool MLOWER, bool UNIT, bool RIGHT>
void leaf(hipStream_t s, int nb, int64_t nvec, const T *A, int64_t lda, T *B, int64_t ldb) {
    if (nvec <= 0 || nb <= 0) return;
    if (nvec > INT32_MAX - 64) fail_arg("trsm: too many right-hand sides");
    hipLaunchKernelGGL((trsm_leaf_kernel<T, MT, MLOWER, UNIT, RIGHT>), dim3((unsigned)cdiv(nvec, 64)),
                       dim3(64), 0, s, nb, (int)nvec, A, lda, B, ldb);
    LMX_LAUNCH_CHECK();
}

template <class T, bool RIGHT>
void leaf_dispatch(hipStream_t s, bool mt, bool mlower, bool unit, int nb, int64_t nvec, const T *A,
                   int64_t lda, T *B, int64_t ldb) {
#define LMX_LEAF(MT_, ML_, U_)                                                      \
    if (mt == MT_ && mlower == ML_ && unit == U_) {                                 \
        leaf<T, MT_, ML_, U_, RIGHT>(s, nb, nvec, A, lda, B, ldb);                  \
        return;                                                                     \
    }
    LMX_LEAF(false, false, false)
    LMX_LEAF(false, false, true)
    LMX_LEAF(false, true, false)
    LMX_LEAF(false, true, true)
    LMX_LEAF(true, false, false)
    LMX_LEAF(true, false, true)
    LMX_LEAF(true, true, false)
    LMX_LEAF(true, true, true)
#undef LMX_LEAF
}

}  // namespace

// op(T) X = B, B n x nrhs.  M = op(T) is the matrix of the per-column
// equation; element M[r][c] is T[r][c] (no transpose) or T[c][r].
template <class T>
void trsm_left(hipStream_t s, bool lower, bool trans, bool unit, int64_t n, int64_t nrhs, const T *A,
               int64_t lda, T *B, int64_t ldb) {
    if (n <= 0 || nrhs <= 0) return;
    TimedRegion tr("trsm", s);
    const bool mlower = lower != trans;
    // Sub-block M[r0:r0+rn, c0:c0+cn] as a GEMM operand: T block, transposed when trans.
    auto mblock = [&](int64_t r0, int64_t c0) -> const T * {
        return trans ? A + c0 * lda + r0 : A + r0 * lda + c0;
    };
    const int64_t nblk = cdiv(n, LEAF);
    for (int64_t bi = 0; bi < nblk; ++bi) {
        // forward order for lower M, backward for upper
        int64_t b = mlower ? bi : nblk - 1 - bi;
        int64_t b0 = b * LEAF, bs = std::min<int64_t>(LEAF, n - b0);
        leaf_dispatch<T, false>(s, trans, mlower, unit, (int)bs, nrhs, A + b0 * lda + b0, lda,
                                B + b0 * ldb, ldb);
        if (mlower && b0 + bs < n) {
            int64_t r0 = b0 + bs, rn = n - r0;
            gemm(s, trans, false, rn, nrhs, bs, T(-1), mblock(r0, b0), lda, B + b0 * ldb, ldb, T(1),
                 B + r0 * ldb, ldb);
        } else if (!mlower && b0 > 0) {
            gemm(s, trans, false, b0, nrhs, bs, T(-1), mblock(0, b0), lda, B + b0 * ldb, ldb, T(1), B,
                 ldb);
        }
    }
}

// X op(T) = B with B nrows x n (row-vector equations M x = b, M = op(T)^T).
template <class T>
void trsm_right(hipStream_t s, bool lower, bool trans, bool unit, int64_t nrows, int64_t n, const T *A,
                int64_t lda, T *B, int64_t ldb) {
    if (n <= 0 || nrows <= 0) return;
    TimedRegion tr("trsm", s);
    const bool mt = !trans;           // M = op(T)^T
    const bool mlower = lower != mt;
    const int64_t nblk = cdiv(n, LEAF);
    for (int64_t bi = 0; bi < nblk; ++bi) {
        int64_t b = mlower ? bi : nblk - 1 - bi;
        int64_t b0 = b * LEAF, bs = std::min<int64_t>(LEAF, n - b0);
        leaf_dispatch<T, true>(s, mt, mlower, unit, (int)bs, nrows, A + b0 * lda + b0, lda, B + b0, ldb);
        // X[:, later] -= X[:, blk] * M[later, blk]^T ; M[r][c] = mt ? T[c][r] : T[r][c]
        // so M[later, blk]^T = T[later, blk]^T (mt = false) or T[blk, later] (mt = true).
        if (mlower && b0 + bs < n) {
            int64_t c0 = b0 + bs, cn = n - c0;
            const T *op = mt ? A + b0 * lda + c0 : A + c0 * lda + b0;
            gemm(s, false, !mt, nrows, cn, bs, T(-1), B + b0, ldb, op, lda, T(1), B + c0, ldb);
        } else if (!mlower && b0 > 0) {
            const T *op = mt ? A + b0 * lda : A + b0;
            gemm(s, false, !mt, nrows, b0, bs, T(-1), B + b0, ldb, op, lda, T(1), B, ldb);
        }
    }
}

template void trsm_left<double>(hipStream_t, bool, bool, bool, int64_t, int64_t, const double *, int64_t,
                                double *, int64_t);
template void trsm_left<float>(hipStream_t, bool, bool, bool, int64_t, int64_t, const float *, int64_t,
                               float *, int64_t);
template void trsm_right<double>(hipStream_t, bool, bool, bool, int64_t, int64_t, const double *, int64_t,
                                 double *, int64_t);
template void trsm_right<float>(hipStream_t, bool, bool, bool, int64_t, int64_t, const float *, int64_t,
                                float *, int64_t);

}  // namespace lmx

using namespace lmx;

namespace {

void parse_trsm(char uplo, char trans, char diag, bool &lower, bool &tr, bool &unit) {
    if (uplo == 'L' || uplo == 'l') lower = true;
    else if (uplo == 'U' || uplo == 'u') lower = false;
    else fail_arg("trsm: uplo must be L or U");
    tr = is_trans(trans);
    if (diag == 'U' || diag == 'u') unit = true;
    else if (diag == 'N' || diag == 'n') unit = false;
    else fail_arg("trsm: diag must be U or N");
}

}  // namespace

extern "C" {

int lmx_dtrsm_dev(void *stream, char uplo, char trans, char diag, int64_t n, int64_t nrhs, const double *dT,
                  int64_t ldt, double *dB, int64_t ldb) {
    return guarded([&] {
        bool lower, tr, unit;
        parse_trsm(uplo, trans, diag, lower, tr, unit);
        if (n < 0 || nrhs < 0) fail_arg("trsm: negative dimension");
        if (ldt < std::max<int64_t>(1, n) || ldb < std::max<int64_t>(1, nrhs)) fail_arg("trsm: bad stride");
        trsm_left<double>(resolve_stream(stream), lower, tr, unit, n, nrhs, dT, ldt, dB, ldb);
        return LMX_OK;
    });
}

int lmx_dtrsm(char uplo, char trans, char diag, int64_t n, int64_t nrhs, const double *T, int64_t ldt,
              double *B, int64_t ldb) {
    return guarded([&] {
        bool lower, tr, unit;
        parse_trsm(uplo, trans, diag, lower, tr, unit);
        if (n < 0 || nrhs < 0) fail_arg("trsm: negative dimension");
        if (ldt < std::max<int64_t>(1, n) || ldb < std::max<int64_t>(1, nrhs)) fail_arg("trsm: bad stride");
        if (n == 0 || nrhs == 0) return LMX_OK;
        Context &c = ctx();
        std::lock_guard<std::mutex> g(c.host_mu);
        size_t tb = (size_t)((n - 1) * ldt + n) * sizeof(double), bb = (size_t)((n - 1) * ldb + nrhs) * sizeof(double);
        double *dT = (double *)staging(c, 0, tb), *dB = (double *)staging(c, 1, bb);
        LMX_HIP(hipMemcpyAsync(dT, T, tb, hipMemcpyHostToDevice, c.stream));
        LMX_HIP(hipMemcpyAsync(dB, B, bb, hipMemcpyHostToDevice, c.stream));
        trsm_left<double>(c.stream, lower, tr, unit, n, nrhs, dT, ldt, dB, ldb);
        LMX_HIP(hipMemcpyAsync(B, dB, bb, hipMemcpyDeviceToHost, c.stream));
        LMX_HIP(hipStreamSynchronize(c.stream));
        return LMX_OK;
    });
}

}  // extern "C"
