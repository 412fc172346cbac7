// Lower Cholesky factorisation (right-looking, blocked) and solve, for
// row-major matrices.  Only the lower triangle is read or written.
//
// potrf_blocked(nb): for each diagonal block of width nb:
//   - factor the block (recursively with nb = 64, then the one-workgroup
//     LDS kernel potrf_diag_kernel for blocks of order <= 64);
//   - L21 = A21 L11^{-T} (right-side blocked TRSM);
//   - A22 -= L21 L21^T on the lower-triangle-only MFMA SYRK.
// The reference's left-looking panelCholFactor (src/.../Solve/Cholesky.hs:
// 279-321) computes the same factor; the right-looking order keeps the
// large update a single wide MFMA launch per block column.
// A pivot that is not > 0 (including NaN) sets info to its 1-based column
// the first time; the factorisation continues and propagates NaN, as the
// reference's sqrt of a negative number does.
#include "common.hpp"
#include "device.hpp"
#include "kernels.hpp"

namespace lmx {

namespace {

constexpr int DIAG = 64;

// In-place lower Cholesky of an nb x nb block (nb <= 64): one wave, lane i
// holding row i in registers.  Step j takes the diagonal and each l_cj
// from the lane that holds it (v_readlane into a scalar register), so a
// step has no barrier: l_jj = sqrt(a_jj), l_ij = a_ij / l_jj (a division,
// as rawCholColumnSIMD), then a_ic -= l_ij l_cj for c > j, in ascending j
// for every element.  Lanes also update their row's columns above the
// diagonal; those registers are never read or stored.
template <class T>
__global__ __launch_bounds__(64) void potrf_diag_kernel(int nb, T *__restrict__ A, int64_t lda, int col0,
                                                        int32_t *__restrict__ info) {
    __shared__ T S[DIAG][DIAG + 1];
    const int i = threadIdx.x;
    // coalesced loads through LDS (lane i reads column i of each row), then
    // lane i takes row i.  The loads are unconditional, at addresses clamped
    // into the block, and issued 16 at a time: a load under a per-lane
    // condition waits for memory before the next one issues.
    const T *Ai = A + min(i, nb - 1);
    for (int r0 = 0; r0 < DIAG; r0 += 16) {
        T v[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) v[u] = Ai[(int64_t)min(r0 + u, nb - 1) * lda];
#pragma unroll
        for (int u = 0; u < 16; ++u) asm volatile("" : "+v"(v[u]));
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int r = r0 + u;
            S[r][i] = (r < nb && i <= r) ? v[u] : T(0);
        }
    }
    __syncthreads();
    T a[DIAG];
#pragma unroll
    for (int c = 0; c < DIAG; ++c) a[c] = S[i][c];
    // All 64 steps run (compile-time indices keep a[] in registers); rows
    // and columns from nb on start at zero, may turn to NaN, and are never
    // stored.
    int bad = 0;
#pragma unroll
    for (int j = 0; j < DIAG; ++j) {
        const T d = lane_value(a[j], j);
        const T sjj = sqrt(d);
        if (j < nb && !(d > T(0)) && bad == 0) bad = col0 + j + 1;
        a[j] = i == j ? sjj : a[j] / sjj;  // rows above j: unused registers
#pragma unroll
        for (int c = j + 1; c < DIAG; ++c) a[c] = fma(-a[j], lane_value(a[j], c), a[c]);
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < DIAG; ++c) S[i][c] = a[c];
    __syncthreads();
    for (int r = 0; r < nb; ++r)
        if (i <= r) A[(int64_t)r * lda + i] = S[r][i];
    if (i == 0 && bad != 0 && *info == 0) *info = bad;
}

// C -= A B^T (lower tiles only with `lower`): FP64 / FP32 MFMA, or the BF16
// tiles with split operands for the BF16 factor.
template <class T>
void chol_update(hipStream_t s, bool bf16, bool lower, int64_t m, int64_t n, int64_t k, const T *A, int64_t lda,
                 const T *B, T *C) {
    if constexpr (sizeof(T) == 4) {
        if (bf16) {
            gemm_bf16x3(s, false, true, m, n, k, -1.0f, A, lda, B, lda, 1.0f, C, lda, lower);
            return;
        }
    }
    if (lower)
        syrk_lower(s, m, k, T(-1), A, lda, T(1), C, lda);
    else
        gemm(s, false, true, m, n, k, T(-1), A, lda, B, lda, T(1), C, lda);
}

template <class T>
void potrf_blocked(hipStream_t s, int64_t n, T *A, int64_t lda, int64_t col0, int32_t *d_info, int64_t nb,
                   bool bf16) {
    for (int64_t k0 = 0; k0 < n; k0 += nb) {
        const int64_t kb = std::min(nb, n - k0);
        T *A11 = A + k0 * lda + k0;
        if (kb <= DIAG) {
            TimedRegion tr("potrf_diag", s);
            hipLaunchKernelGGL(potrf_diag_kernel<T>, dim3(1), dim3(64), 0, s, (int)kb, A11, lda,
                               (int)(col0 + k0), d_info);
            LMX_LAUNCH_CHECK();
        } else {
            potrf_blocked<T>(s, kb, A11, lda, col0 + k0, d_info, DIAG, bf16);
        }
        const int64_t rest = n - k0 - kb;
        if (rest > 0) {
            T *A21 = A11 + kb * lda;
            trsm_right<T>(s, true, true, false, rest, kb, A11, lda, A21, lda);
            chol_update<T>(s, bf16, true, rest, rest, kb, A21, lda, A21, A21 + kb);
        }
    }
}

// The top level with lookahead: the caller's stream factors block column
// b + 1 (diagonal block, then the column below it) while an auxiliary
// stream applies block b's SYRK to the columns beyond b + 1.
//   stream s:   panel(b); wait done(b-1); b's update of block column b + 1
//               (SYRK of its diagonal block, GEMM below it); record ready(b)
//   stream aux: wait ready(b); SYRK of the trailing square beyond b + 1;
//               record done(b)
// While they overlap the streams write disjoint columns (b + 1's against
// those beyond it), and the auxiliary stream only reads block column b.
template <class T>
void potrf_lookahead(hipStream_t s, int64_t n, T *A, int64_t lda, int32_t *d_info, int64_t nb, bool bf16) {
    AuxStream la;
    bool pending = false;
    for (int64_t k0 = 0; k0 < n; k0 += nb) {
        const int64_t kb = std::min(nb, n - k0);
        T *A11 = A + k0 * lda + k0;
        if (kb <= DIAG) {
            TimedRegion tr("potrf_diag", s);
            hipLaunchKernelGGL(potrf_diag_kernel<T>, dim3(1), dim3(64), 0, s, (int)kb, A11, lda, (int)k0, d_info);
            LMX_LAUNCH_CHECK();
        } else {
            potrf_blocked<T>(s, kb, A11, lda, k0, d_info, DIAG, bf16);
        }
        const int64_t rest = n - k0 - kb;
        T *A21 = A11 + kb * lda;
        if (rest > 0) trsm_right<T>(s, true, true, false, rest, kb, A11, lda, A21, lda);
        if (pending) LMX_HIP(hipStreamWaitEvent(s, la.done, 0));
        const int64_t narrow = std::min(nb, rest);
        if (rest > 0) {
            chol_update<T>(s, bf16, true, narrow, narrow, kb, A21, lda, A21, A21 + kb);
            if (rest > narrow)
                chol_update<T>(s, bf16, false, rest - narrow, narrow, kb, A21 + narrow * lda, lda, A21,
                               A21 + narrow * lda + kb);
        }
        LMX_HIP(hipEventRecord(la.ready, s));
        LMX_HIP(hipStreamWaitEvent(la.aux, la.ready, 0));
        if (rest > narrow)
            chol_update<T>(la.aux, bf16, true, rest - narrow, rest - narrow, kb, A21 + narrow * lda, lda,
                           A21 + narrow * lda, A21 + narrow * lda + kb + narrow);
        LMX_HIP(hipEventRecord(la.done, la.aux));
        pending = true;
    }
    if (pending) LMX_HIP(hipStreamWaitEvent(s, la.done, 0));
}

}  // namespace

template <class T>
void potrf(hipStream_t s, int64_t n, T *A, int64_t lda, int32_t *d_info, bool bf16) {
    if (n < 0 || lda < std::max<int64_t>(1, n)) fail_arg("potrf: bad dimension or stride");
    if (n > INT32_MAX / 2) fail_arg("potrf: n too large");
    clear_info(s, d_info);
    if (n == 0) return;
    TimedRegion trall("potrf", s, (double)n * (double)n * (double)n / 3.0);
    if (bf16 && sizeof(T) != 4) fail_arg("potrf: BF16 updates are for the FP32-stored factor");
    potrf_lookahead<T>(s, n, A, lda, d_info, 128, bf16);
}

template <class T>
void potrs(hipStream_t s, int64_t n, int64_t nrhs, const T *L, int64_t lda, T *B, int64_t ldb, const T *WL,
           const T *WU) {
    if (n == 0 || nrhs == 0) return;
    if (WL && WU) {
        tri_inverse_solve<T>(s, true, false, n, nrhs, L, lda, WL, B, ldb);
        tri_inverse_solve<T>(s, true, true, n, nrhs, L, lda, WU, B, ldb);
        return;
    }
    trsm_left<T>(s, true, false, false, n, nrhs, L, lda, B, ldb);
    trsm_left<T>(s, true, true, false, n, nrhs, L, lda, B, ldb);
}

template void potrf<double>(hipStream_t, int64_t, double *, int64_t, int32_t *, bool);
template void potrf<float>(hipStream_t, int64_t, float *, int64_t, int32_t *, bool);
template void potrs<double>(hipStream_t, int64_t, int64_t, const double *, int64_t, double *, int64_t,
                            const double *, const double *);
template void potrs<float>(hipStream_t, int64_t, int64_t, const float *, int64_t, float *, int64_t, const float *,
                           const float *);

}  // namespace lmx

using namespace lmx;

namespace {

size_t extent(int64_t rows, int64_t cols, int64_t ld) {
    return (rows > 0 && cols > 0) ? (size_t)((rows - 1) * ld + cols) * sizeof(double) : 0;
}

void check(int64_t n, int64_t lda, const char *who) {
    if (n < 0) fail_arg(std::string(who) + ": negative dimension");
    if (lda < std::max<int64_t>(1, n)) fail_arg(std::string(who) + ": lda too small");
}

void check_rhs(int64_t nrhs, int64_t ldb, const char *who) {
    if (nrhs < 0) fail_arg(std::string(who) + ": negative nrhs");
    if (ldb < std::max<int64_t>(1, nrhs)) fail_arg(std::string(who) + ": ldb too small");
}

int read_info(hipStream_t s, const int32_t *d_info) {
    int32_t info = 0;
    LMX_HIP(hipMemcpyAsync(&info, d_info, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    LMX_HIP(hipStreamSynchronize(s));
    return info;
}

}  // namespace

extern "C" {

int lmx_dpotrf_dev(void *stream, int64_t n, double *dA, int64_t lda, int32_t *d_info) {
    return guarded([&] {
        check(n, lda, "potrf");
        potrf<double>(resolve_stream(stream), n, dA, lda, d_info);
        return LMX_OK;
    });
}

int lmx_dpotrs_dev(void *stream, int64_t n, int64_t nrhs, const double *dG, int64_t lda, double *dB,
                   int64_t ldb) {
    return guarded([&] {
        check(n, lda, "potrs");
        check_rhs(nrhs, ldb, "potrs");
        potrs<double>(resolve_stream(stream), n, nrhs, dG, lda, dB, ldb);
        return LMX_OK;
    });
}

int lmx_dposv_dev(void *stream, int64_t n, int64_t nrhs, double *dA, int64_t lda, double *dB, int64_t ldb,
                  int32_t *d_info) {
    return guarded([&] {
        check(n, lda, "posv");
        check_rhs(nrhs, ldb, "posv");
        hipStream_t s = resolve_stream(stream);
        potrf<double>(s, n, dA, lda, d_info);
        potrs<double>(s, n, nrhs, dA, lda, dB, ldb);
        return LMX_OK;
    });
}

int lmx_dpotrf(int64_t n, double *A, int64_t lda) {
    return guarded([&] {
        check(n, lda, "potrf");
        if (n == 0) return LMX_OK;
        Context &c = ctx();
        std::lock_guard<std::mutex> g(c.host_mu);
        size_t ab = extent(n, n, lda);
        double *dA = (double *)staging(c, 0, ab);
        int32_t *di = (int32_t *)staging(c, 1, 64);
        LMX_HIP(hipMemcpyAsync(dA, A, ab, hipMemcpyHostToDevice, c.stream));
        potrf<double>(c.stream, n, dA, lda, di);
        LMX_HIP(hipMemcpyAsync(A, dA, ab, hipMemcpyDeviceToHost, c.stream));
        return read_info(c.stream, di);
    });
}

int lmx_dpotrs(int64_t n, int64_t nrhs, const double *G, int64_t lda, double *B, int64_t ldb) {
    return guarded([&] {
        check(n, lda, "potrs");
        check_rhs(nrhs, ldb, "potrs");
        if (n == 0 || nrhs == 0) return LMX_OK;
        Context &c = ctx();
        std::lock_guard<std::mutex> g(c.host_mu);
        size_t ab = extent(n, n, lda), bb = extent(n, nrhs, ldb);
        double *dA = (double *)staging(c, 0, ab), *dB = (double *)staging(c, 1, bb);
        LMX_HIP(hipMemcpyAsync(dA, G, ab, hipMemcpyHostToDevice, c.stream));
        LMX_HIP(hipMemcpyAsync(dB, B, bb, hipMemcpyHostToDevice, c.stream));
        potrs<double>(c.stream, n, nrhs, dA, lda, dB, ldb);
        LMX_HIP(hipMemcpyAsync(B, dB, bb, hipMemcpyDeviceToHost, c.stream));
        LMX_HIP(hipStreamSynchronize(c.stream));
        return LMX_OK;
    });
}

int lmx_dposv(int64_t n, int64_t nrhs, double *A, int64_t lda, double *B, int64_t ldb) {
    return guarded([&] {
        check(n, lda, "posv");
        check_rhs(nrhs, ldb, "posv");
        if (n == 0) return LMX_OK;
        Context &c = ctx();
        std::lock_guard<std::mutex> g(c.host_mu);
        size_t ab = extent(n, n, lda), bb = extent(n, nrhs, ldb);
        double *dA = (double *)staging(c, 0, ab), *dB = (double *)staging(c, 1, bb);
        int32_t *di = (int32_t *)staging(c, 2, 64);
        LMX_HIP(hipMemcpyAsync(dA, A, ab, hipMemcpyHostToDevice, c.stream));
        if (bb) LMX_HIP(hipMemcpyAsync(dB, B, bb, hipMemcpyHostToDevice, c.stream));
        potrf<double>(c.stream, n, dA, lda, di);
        potrs<double>(c.stream, n, nrhs, dA, lda, dB, ldb);
        LMX_HIP(hipMemcpyAsync(A, dA, ab, hipMemcpyDeviceToHost, c.stream));
        if (bb) LMX_HIP(hipMemcpyAsync(B, dB, bb, hipMemcpyDeviceToHost, c.stream));
        return read_info(c.stream, di);
    });
}

}  // extern "C"
