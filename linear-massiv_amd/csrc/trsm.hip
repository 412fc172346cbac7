// Triangular solves with many right-hand sides.
//
// Every solve is written as M x = b for each right-hand-side vector x: a
// column of B for a left-side solve (M = op(T)), a row of B for a
// right-side solve (X op(T) = B, M = op(T)^T).  M is processed in 64 x 64
// blocks:
//   - many right-hand sides ("wide"): for each diagonal block in dependency
//     order, trsm_leaf_kernel solves the block (triangle in LDS, one vector
//     per thread) and the not-yet-solved part is updated with one MFMA GEMM;
//   - at most 64 right-hand sides and more than one block ("chain"):
//     trtri_blocks_kernel inverts every diagonal block in LDS (one launch),
//     then one cooperative launch with one workgroup per block row:
//     workgroup b streams the blocks M_bj through LDS (the next one in
//     flight while the current one is used), waits for block j's flag,
//     subtracts M_bj x_j on MFMA, and finally writes x_b = W_b acc and raises
//     its own flag -- no host round trip and no per-block launch.
#include "common.hpp"
#include "device.hpp"
#include "kernels.hpp"
#include "sync.hpp"

namespace lmx {

namespace {

constexpr int TB = 64;
constexpr long CHAIN_SPIN_LIMIT = 1L << 24;

// Element (r, c) of M = op(T) (MT = transpose), restricted to M's triangle.
template <class T, bool MT, bool MLOWER>
__device__ inline T m_elem(const T *A, int64_t lda, int64_t r, int64_t c) {
    bool in = MLOWER ? (c <= r) : (c >= r);
    if (!in) return T(0);
    return MT ? A[c * lda + r] : A[r * lda + c];
}

// W_b = inverse of diagonal block b of M (order min(64, n - 64 b), padded
// to a 64 x 64 identity), stored densely at W + 4096 b.
template <class T, bool MT, bool MLOWER, bool UNIT>
__global__ __launch_bounds__(256) void trtri_blocks_kernel(int n, const T *__restrict__ A, int64_t lda,
                                                           T *__restrict__ W) {
    __shared__ T Ms[TB][TB + 1];
    __shared__ T Xs[TB][TB + 1];
    __shared__ T red[4][TB];
    const int t = threadIdx.x, b = blockIdx.x;
    const int r0 = b * TB, nb = min(TB, n - r0);
    const T *Ab = A + (int64_t)r0 * lda + r0;
#pragma unroll
    for (int e = t; e < TB * TB; e += 256) {
        int r = e >> 6, c = e & 63;
        T v = T(0);
        if (r < nb && c < nb) {
            v = m_elem<T, MT, MLOWER>(Ab, lda, r, c);
            if (r == c && UNIT) v = T(1);
        } else if (r == c) {
            v = T(1);
        }
        Ms[r][c] = v;
        Xs[r][c] = T(0);
    }
    __syncthreads();
    const int c = t & 63, q = t >> 6;
    for (int step = 0; step < TB; ++step) {
        const int r = MLOWER ? step : TB - 1 - step;
        T s = T(0);
        if (MLOWER) {
            for (int k = c + q; k < r; k += 4) s = fma(Ms[r][k], Xs[k][c], s);
        } else {
            for (int k = r + 1 + q; k <= c; k += 4) s = fma(Ms[r][k], Xs[k][c], s);
        }
        red[q][c] = s;
        __syncthreads();
        if (q == 0) {
            bool in = MLOWER ? (r >= c) : (r <= c);
            T v = T(0);
            if (in) {
                T tot = (red[0][c] + red[1][c]) + (red[2][c] + red[3][c]);
                v = ((r == c ? T(1) : T(0)) - tot) / Ms[r][r];
            }
            Xs[r][c] = v;
        }
        __syncthreads();
    }
    T *Wb = W + (int64_t)b * TB * TB;
    for (int e = t; e < TB * TB; e += 256) Wb[e] = Xs[e >> 6][e & 63];
}

// ---------------------------------------------------------------- leaf kernel

// M x = b for the diagonal block (order nb <= 64) with one right-hand-side
// vector per thread: the block sits in LDS, each thread's vector in its own
// LDS column, so after the block is staged no barrier is needed.  Row form:
// x_r = (b_r - sum_k m_rk x_k) / m_rr (a division, as the reference's
// substitutions).  LEFT: vectors are columns of B; RIGHT: rows of B.
template <class T, bool MT, bool MLOWER, bool UNIT, bool RIGHT>
__global__ __launch_bounds__(64) void trsm_leaf_kernel(int nb, int nvec, const T *__restrict__ A, int64_t lda,
                                                       T *__restrict__ B, int64_t ldb) {
    constexpr int PAD = sizeof(T) == 8 ? 1 : 2;
    __shared__ T Ms[TB][TB + PAD];
    __shared__ T Xs[TB][64 + PAD];
    const int t = threadIdx.x;
    // unrolled so that the global loads of a batch are in flight together
#pragma unroll 16
    for (int e = t; e < TB * TB; e += 64) {
        int r = e >> 6, c = e & 63;
        T v = T(0);
        if (r < nb && c < nb) {
            v = m_elem<T, MT, MLOWER>(A, lda, r, c);
            if (r == c && UNIT) v = T(1);
        }
        Ms[r][c] = v;
    }
    const int v = blockIdx.x * 64 + t;
    const bool live = v < nvec;
    if (RIGHT) {
#pragma unroll 16
        for (int r = 0; r < nb; ++r) Xs[r][t] = live ? B[(int64_t)v * ldb + r] : T(0);
    } else {
#pragma unroll 16
        for (int r = 0; r < nb; ++r) Xs[r][t] = live ? B[(int64_t)r * ldb + v] : T(0);
    }
    __syncthreads();
    if (MLOWER) {
        for (int r = 0; r < nb; ++r) {
            T s = Xs[r][t];
            for (int k = 0; k < r; ++k) s = fma(-Ms[r][k], Xs[k][t], s);
            Xs[r][t] = UNIT ? s : s / Ms[r][r];
        }
    } else {
        for (int r = nb - 1; r >= 0; --r) {
            T s = Xs[r][t];
            for (int k = r + 1; k < nb; ++k) s = fma(-Ms[r][k], Xs[k][t], s);
            Xs[r][t] = UNIT ? s : s / Ms[r][r];
        }
    }
    if (!live) return;
    if (RIGHT) {
        for (int r = 0; r < nb; ++r) B[(int64_t)v * ldb + r] = Xs[r][t];
    } else {
        for (int r = 0; r < nb; ++r) B[(int64_t)r * ldb + v] = Xs[r][t];
    }
}

// ---------------------------------------------------------------- chain kernel

// The 64 x 64 block of M at (r0, c0) (zero outside the matrix), fetched
// into 16 registers per thread so all loads are in flight together, and
// stored to LDS separately.
template <class T, bool MT>
__device__ inline void fetch_mblock(T (&reg)[16], const T *A, int64_t lda, int n, int r0, int c0, int t) {
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        int e = t + 256 * i, r = e >> 6, c = e & 63;
        int64_t gr = r0 + r, gc = c0 + c;
        reg[i] = (gr < n && gc < n) ? (MT ? A[gc * lda + gr] : A[gr * lda + gc]) : T(0);
    }
}
template <class T, int LD>
__device__ inline void store_mblock(T (*Ms)[LD], const T (&reg)[16], int t) {
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        int e = t + 256 * i;
        Ms[e >> 6][e & 63] = reg[i];
    }
}

// Visibility between workgroups: a block's solution is written to B with
// device-coherent stores (xstore), every wave drains them, and after a
// barrier thread 0 raises the flag; the reader's thread 0 spins on the flag,
// and after the barrier the block is read with device-coherent loads.
template <class T, bool MT, bool MLOWER>
__global__ __launch_bounds__(256) void trsm_chain_kernel(int n, int nrhs, const T *__restrict__ A, int64_t lda,
                                                         const T *__restrict__ W, T *B, int64_t ldb,
                                                         int *flags, int *err) {
    constexpr int PAD = sizeof(T) == 8 ? 1 : 2;
    constexpr int LD = TB + PAD;
    __shared__ T Ms[2][TB][LD];
    __shared__ T Ws[TB][LD];
    __shared__ T Xs[TB][LD];
    typedef typename V4<T>::type V;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6, lr = lane & 15, lk = lane >> 4;
    const int nblk = gridDim.x;
    const int b = MLOWER ? (int)blockIdx.x : nblk - 1 - (int)blockIdx.x;
    const int r0 = b * TB;
    const int ncb = (nrhs + 15) >> 4;  // 16-column blocks (<= 4)
    const int ndep = MLOWER ? b : nblk - 1 - b;

    // W_b and the first off-diagonal block are fetched before any waiting.
    const T *Wb = W + (int64_t)b * TB * TB;
#pragma unroll
    for (int e = t; e < TB * TB; e += 256) Ws[e >> 6][e & 63] = Wb[e];
    T nxt[16];
    if (ndep > 0) {
        fetch_mblock<T, MT>(nxt, A, lda, n, r0, (MLOWER ? 0 : nblk - 1) * TB, t);
        store_mblock<T, LD>(Ms[0], nxt, t);
    }

    V acc[4];
    // acc <- B_b (rows r0 + 16 wave + acc_row, cols 16 cb + lr)
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            int gi = r0 + wave * 16 + acc_row<T>(lk, r), gc = cb * 16 + lr;
            acc[cb][r] = (cb < ncb && gi < n && gc < nrhs) ? B[(int64_t)gi * ldb + gc] : T(0);
        }
    }
    bool failed = false;
    for (int d = 0; d < ndep; ++d) {
        const int j = MLOWER ? d : nblk - 1 - d;
        const int c0 = j * TB;
        const int cur = d & 1;
        if (t == 0 && !failed) {
            long spins = 0;
            while (xload(&flags[j]) == 0) {
                __builtin_amdgcn_s_sleep(1);
                if (++spins > CHAIN_SPIN_LIMIT) {
                    xstore(err, 1);
                    failed = true;
                    break;
                }
            }
        }
        __syncthreads();  // Ms[cur] loaded, flag j seen, Xs free
#pragma unroll
        for (int i = 0; i < TB * TB / 256; ++i) {
            const int e = t + 256 * i;
            if (e < TB * nrhs) {
                int r = e / nrhs, c = e - r * nrhs;
                int64_t gr = c0 + r;
                Xs[r][c] = gr < n ? xload(&B[gr * ldb + c]) : T(0);
            }
        }
        // the next block's loads fly while this block is used; they land in
        // the other LDS buffer after the MFMAs
        const bool more = d + 1 < ndep;
        if (more) fetch_mblock<T, MT>(nxt, A, lda, n, r0, (MLOWER ? j + 1 : j - 1) * TB, t);
        __syncthreads();
#pragma unroll 4
        for (int kk = 0; kk < TB; kk += 4) {
            T a = -Ms[cur][wave * 16 + lr][kk + lk];
#pragma unroll
            for (int cb = 0; cb < 4; ++cb)
                if (cb < ncb) acc[cb] = mfma(a, Xs[kk + lk][cb * 16 + lr], acc[cb]);
        }
        if (more) store_mblock<T, LD>(Ms[cur ^ 1], nxt, t);
    }
    // x_b = W_b acc
    __syncthreads();
#pragma unroll
    for (int cb = 0; cb < 4; ++cb)
#pragma unroll
        for (int r = 0; r < 4; ++r) Xs[wave * 16 + acc_row<T>(lk, r)][cb * 16 + lr] = acc[cb][r];
    __syncthreads();
    V out[4];
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) out[cb] = V{0, 0, 0, 0};
#pragma unroll 4
    for (int kk = 0; kk < TB; kk += 4) {
        T a = Ws[wave * 16 + lr][kk + lk];
#pragma unroll
        for (int cb = 0; cb < 4; ++cb)
            if (cb < ncb) out[cb] = mfma(a, Xs[kk + lk][cb * 16 + lr], out[cb]);
    }
#pragma unroll
    for (int cb = 0; cb < 4; ++cb)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            int gi = r0 + wave * 16 + acc_row<T>(lk, r), gc = cb * 16 + lr;
            if (cb < ncb && gi < n && gc < nrhs) xstore(&B[(int64_t)gi * ldb + gc], out[cb][r]);
        }
    xdrain();
    __syncthreads();
    if (t == 0) xstore(&flags[b], 1);
}

__global__ void zero_i32_kernel(int *p, int n) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 0;
}

// ---------------------------------------------------------------- dispatch

template <class T>
void trtri_blocks(hipStream_t s, bool mt, bool mlower, bool unit, int n, const T *A, int64_t lda, T *W) {
    dim3 g((unsigned)cdiv(n, TB)), blk(256);
#define LMX_TRTRI(MT_, ML_, U_)                                                                 \
    if (mt == MT_ && mlower == ML_ && unit == U_) {                                            \
        hipLaunchKernelGGL((trtri_blocks_kernel<T, MT_, ML_, U_>), g, blk, 0, s, n, A, lda, W); \
        LMX_LAUNCH_CHECK();                                                                     \
        return;                                                                                 \
    }
    LMX_TRTRI(false, false, false)
    LMX_TRTRI(false, false, true)
    LMX_TRTRI(false, true, false)
    LMX_TRTRI(false, true, true)
    LMX_TRTRI(true, false, false)
    LMX_TRTRI(true, false, true)
    LMX_TRTRI(true, true, false)
    LMX_TRTRI(true, true, true)
#undef LMX_TRTRI
}

template <class T>
int chain_max_blocks() {
    static int cached = -1;
    if (cached < 0) {
        int dev = 0, per_cu = 0;
        LMX_HIP(hipGetDevice(&dev));
        LMX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, trsm_chain_kernel<T, false, true>, 256, 0));
        cached = std::max(per_cu, 0) * ctx().cu_count;
    }
    return cached;
}

// Chain solve for nrhs <= 64; returns false when the grid cannot be co-resident.
template <class T>
bool chain_solve(hipStream_t s, bool mt, bool mlower, int n, int nrhs, const T *A, int64_t lda, const T *W, T *B,
                 int64_t ldb) {
    int nblk = (int)cdiv(n, TB);
    if (nblk > chain_max_blocks<T>()) return false;
    Scratch fl((size_t)(nblk + 64) * sizeof(int), s);
    int *flags = fl.as<int>(), *err = flags + nblk;
    hipLaunchKernelGGL(zero_i32_kernel, dim3((unsigned)cdiv(nblk + 1, 256)), dim3(256), 0, s, flags, nblk + 1);
    LMX_LAUNCH_CHECK();
    void *args[] = {&n, &nrhs, &A, &lda, &W, &B, &ldb, &flags, &err};
    const void *fn = nullptr;
    if (!mt && mlower) fn = (const void *)trsm_chain_kernel<T, false, true>;
    if (!mt && !mlower) fn = (const void *)trsm_chain_kernel<T, false, false>;
    if (mt && mlower) fn = (const void *)trsm_chain_kernel<T, true, true>;
    if (mt && !mlower) fn = (const void *)trsm_chain_kernel<T, true, false>;
    LMX_HIP(hipLaunchCooperativeKernel(fn, dim3(nblk), dim3(256), args, 0, s));
    return true;
}

// Leaf solve of one diagonal block (order bs) for nvec vectors.
template <class T, bool RIGHT>
void leaf(hipStream_t s, bool mt, bool mlower, bool unit, int bs, int64_t nvec, const T *A, int64_t lda, T *B,
          int64_t ldb) {
    dim3 g((unsigned)cdiv(nvec, 64)), blk(64);
#define LMX_LEAF(MT_, ML_, U_)                                                                                  \
    if (mt == MT_ && mlower == ML_ && unit == U_) {                                                            \
        hipLaunchKernelGGL((trsm_leaf_kernel<T, MT_, ML_, U_, RIGHT>), g, blk, 0, s, bs, (int)nvec, A, lda, B, \
                           ldb);                                                                               \
        LMX_LAUNCH_CHECK();                                                                                     \
        return;                                                                                                 \
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

// op(T) X = B, B n x nrhs.
template <class T>
void trsm_left(hipStream_t s, bool lower, bool trans, bool unit, int64_t n, int64_t nrhs, const T *A, int64_t lda,
               T *B, int64_t ldb) {
    if (n <= 0 || nrhs <= 0) return;
    if (n > INT32_MAX / 2 || nrhs > INT32_MAX / 2) fail_arg("trsm: dimension too large");
    TimedRegion tr("trsm", s, (double)n * (double)n * (double)nrhs);
    const bool mt = trans, mlower = lower != trans;
    const int64_t nblk = cdiv(n, TB);
    if (nrhs <= 64 && nblk > 1) {
        Scratch sw((size_t)nblk * TB * TB * sizeof(T), s);
        T *W = sw.as<T>();
        trtri_blocks<T>(s, mt, mlower, unit, (int)n, A, lda, W);
        if (chain_solve<T>(s, mt, mlower, (int)n, (int)nrhs, A, lda, W, B, ldb)) return;
    }
    // M[r0.., c0..] as a GEMM operand: the T block, transposed when trans.
    auto mblock = [&](int64_t r0, int64_t c0) -> const T * { return trans ? A + c0 * lda + r0 : A + r0 * lda + c0; };
    for (int64_t bi = 0; bi < nblk; ++bi) {
        int64_t b = mlower ? bi : nblk - 1 - bi;
        int64_t b0 = b * TB, bs = std::min<int64_t>(TB, n - b0);
        leaf<T, false>(s, mt, mlower, unit, (int)bs, nrhs, A + b0 * lda + b0, lda, B + b0 * ldb, ldb);
        if (mlower && b0 + bs < n) {
            int64_t r0 = b0 + bs, rn = n - r0;
            gemm(s, trans, false, rn, nrhs, bs, T(-1), mblock(r0, b0), lda, B + b0 * ldb, ldb, T(1), B + r0 * ldb,
                 ldb);
        } else if (!mlower && b0 > 0) {
            gemm(s, trans, false, b0, nrhs, bs, T(-1), mblock(0, b0), lda, B + b0 * ldb, ldb, T(1), B, ldb);
        }
    }
}

// X op(T) = B with B nrows x n (row-vector equations M x = b, M = op(T)^T).
template <class T>
void trsm_right(hipStream_t s, bool lower, bool trans, bool unit, int64_t nrows, int64_t n, const T *A, int64_t lda,
                T *B, int64_t ldb) {
    if (n <= 0 || nrows <= 0) return;
    if (n > INT32_MAX / 2 || nrows > INT32_MAX / 2) fail_arg("trsm: dimension too large");
    TimedRegion tr("trsm", s, (double)n * (double)n * (double)nrows);
    const bool mt = !trans, mlower = lower != mt;
    const int64_t nblk = cdiv(n, TB);
    for (int64_t bi = 0; bi < nblk; ++bi) {
        int64_t b = mlower ? bi : nblk - 1 - bi;
        int64_t b0 = b * TB, bs = std::min<int64_t>(TB, n - b0);
        leaf<T, true>(s, mt, mlower, unit, (int)bs, nrows, A + b0 * lda + b0, lda, B + b0, ldb);
        // X[:, later] -= X[:, blk] M[later, blk]^T ; M[r][c] = mt ? T[c][r] : T[r][c]
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

template void trsm_left<double>(hipStream_t, bool, bool, bool, int64_t, int64_t, const double *, int64_t, double *,
                                int64_t);
template void trsm_left<float>(hipStream_t, bool, bool, bool, int64_t, int64_t, const float *, int64_t, float *,
                               int64_t);
template void trsm_right<double>(hipStream_t, bool, bool, bool, int64_t, int64_t, const double *, int64_t, double *,
                                 int64_t);
template void trsm_right<float>(hipStream_t, bool, bool, bool, int64_t, int64_t, const float *, int64_t, float *,
                                int64_t);

}  // namespace lmx

using namespace lmx;

namespace {

void parse_trsm(char uplo, char trans, char diag, bool &lower, bool &tr, bool &unit) {
    if (uplo == 'L' || uplo == 'l')
        lower = true;
    else if (uplo == 'U' || uplo == 'u')
        lower = false;
    else
        fail_arg("trsm: uplo must be L or U");
    tr = is_trans(trans);
    if (diag == 'U' || diag == 'u')
        unit = true;
    else if (diag == 'N' || diag == 'n')
        unit = false;
    else
        fail_arg("trsm: diag must be U or N");
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

int lmx_dtrsm(char uplo, char trans, char diag, int64_t n, int64_t nrhs, const double *T, int64_t ldt, double *B,
              int64_t ldb) {
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
