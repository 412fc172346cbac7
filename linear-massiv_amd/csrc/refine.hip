// Mixed-precision solves with FP64 iterative refinement.
//
// A (FP64, unchanged) is rounded to the factor precision and factored there
// (LU with partial pivoting, or Cholesky) on the MFMA kernels.  Then, in
// FP64 on the device:
//   X_0 = 0, R_0 = B
//   repeat: D = solve_lowprec(R);  X += D;  R = B - A X  (FP64 MFMA GEMM);
//           eta = max_c ||r_c||_inf / (||A||_inf ||x_c||_inf + ||b_c||_inf)
// until every column meets LAPACK dsgesv's test
//   ||r_c||_inf <= sqrt(n) eps ||A||_inf ||x_c||_inf,
// or eta stops shrinking (less than 10% per step), or 30 steps.  A solve that does not converge is
// redone with the FP64 factorisation (fell_back = 1).  Only eta (one double)
// crosses to the host per step.
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "common.hpp"
#include "device.hpp"
#include "kernels.hpp"

namespace lmx {

namespace {

constexpr int ITMAX = 30;
constexpr double EPS64 = 2.220446049250313e-16;

__global__ void d2f_kernel(const double *__restrict__ in, int64_t ldi, float *__restrict__ out, int64_t ldo,
                           int64_t rows, int64_t cols) {
    int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= rows * cols) return;
    int64_t i = e / cols, c = e % cols;
    out[i * ldo + c] = (float)in[i * ldi + c];
}

// X += (double) D ; optionally X = D when `assign`.
__global__ void f2d_acc_kernel(const float *__restrict__ D, int64_t ldd, double *__restrict__ X, int64_t ldx,
                               int64_t rows, int64_t cols, int assign) {
    int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= rows * cols) return;
    int64_t i = e / cols, c = e % cols;
    double d = (double)D[i * ldd + c];
    X[i * ldx + c] = assign ? d : X[i * ldx + c] + d;
}

__global__ void copy_d_kernel(const double *__restrict__ in, int64_t ldi, double *__restrict__ out, int64_t ldo,
                              int64_t rows, int64_t cols) {
    int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= rows * cols) return;
    int64_t i = e / cols, c = e % cols;
    out[i * ldo + c] = in[i * ldi + c];
}

__global__ void zero_d_kernel(double *p, int n) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

// Non-negative doubles order like their bit patterns: max via integer atomics.
__device__ inline void atomic_max_nonneg(double *addr, double v) {
    if (!(v >= 0.0)) v = __builtin_huge_val();  // NaN poisons the norm
    atomicMax(reinterpret_cast<unsigned long long *>(addr), (unsigned long long)__double_as_longlong(v));
}

// ||A||_inf = max_i sum_j |a_ij|: one wave per row at a time (coalesced
// row reads, four partial sums per lane), rows strided over the grid, one
// atomic per workgroup.
__global__ __launch_bounds__(256) void norm_inf_rows_kernel(const double *__restrict__ A, int64_t lda,
                                                            int64_t n, double *__restrict__ out) {
    __shared__ double part[4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t nw = (int64_t)gridDim.x * 4;
    double best = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 4 + wave; i < n; i += nw) {
        const double *row = A + i * lda;
        double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
        int64_t j = lane;
        for (; j + 192 < n; j += 256) {
            s0 += fabs(row[j]);
            s1 += fabs(row[j + 64]);
            s2 += fabs(row[j + 128]);
            s3 += fabs(row[j + 192]);
        }
        for (; j < n; j += 64) s0 += fabs(row[j]);
        const double sum = wave_sum((s0 + s1) + (s2 + s3));
        best = (sum > best || isnan(sum)) ? sum : best;
    }
    if (lane == 0) part[wave] = best;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = part[0];
        for (int x = 1; x < 4; ++x) m = (part[x] > m || isnan(part[x])) ? part[x] : m;
        atomic_max_nonneg(out, m);
    }
}

// out[c] = max_i |X[i][c]| for an n x ncols block: a workgroup covers
// COLMAX_ROWS rows and 64 columns (blockIdx.y), reduces in LDS, and issues
// one atomic per column.
constexpr int COLMAX_ROWS = 1024;
__global__ __launch_bounds__(256) void colmax_kernel(const double *__restrict__ X, int64_t ldx, int64_t n,
                                                     int ncols, double *__restrict__ out) {
    __shared__ double part[4][64];
    const int t = threadIdx.x, cl = t & 63, q = t >> 6;
    const int c = blockIdx.y * 64 + cl;
    const int64_t r0 = (int64_t)blockIdx.x * COLMAX_ROWS;
    double m = 0.0;
    if (c < ncols)
        for (int64_t i = r0 + q; i < n && i < r0 + COLMAX_ROWS; i += 4) {
            double v = fabs(X[i * ldx + c]);
            if (!(v <= m)) m = v;  // keeps NaN
        }
    part[q][cl] = m;
    __syncthreads();
    if (q == 0 && c < ncols) {
        double r = part[0][cl];
        for (int x = 1; x < 4; ++x)
            if (!(part[x][cl] <= r)) r = part[x][cl];
        atomic_max_nonneg(&out[c], r);
    }
}

// The same for at most 8 columns: one thread per row (16 rows per thread,
// coalesced), wave and workgroup reductions, one atomic per column and
// workgroup.  NaN becomes +inf first (as the atomic would make it).
constexpr int COLMAX_FEW = 8;
__global__ __launch_bounds__(256) void colmax_few_kernel(const double *__restrict__ X, int64_t ldx, int64_t n,
                                                         int ncols, double *__restrict__ out) {
    __shared__ double part[4][COLMAX_FEW];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    double m[COLMAX_FEW];
#pragma unroll
    for (int c = 0; c < COLMAX_FEW; ++c) m[c] = 0.0;
    const int64_t base = (int64_t)blockIdx.x * 4096;
#pragma unroll 4
    for (int k = 0; k < 16; ++k) {
        const int64_t i = base + k * 256 + t;
        if (i < n) {
#pragma unroll
            for (int c = 0; c < COLMAX_FEW; ++c)
                if (c < ncols) {
                    double v = fabs(X[i * ldx + c]);
                    if (isnan(v)) v = __builtin_huge_val();
                    m[c] = fmax(m[c], v);
                }
        }
    }
#pragma unroll
    for (int c = 0; c < COLMAX_FEW; ++c) {
        const double w = wave_max(m[c]);
        if (lane == 0) part[wave][c] = w;
    }
    __syncthreads();
    if (t < ncols) atomic_max_nonneg(&out[t], fmax(fmax(part[0][t], part[1][t]), fmax(part[2][t], part[3][t])));
}

// R = B - A X in FP64 for nrhs <= NR: each wave takes RESID_RPW rows of A
// and walks them together, so one load of x[k] serves all of them (coalesced
// row reads).  The columns go in pairs of chunks of RESID_CH * 64, pipelined
// one chunk ahead in two register buffers used in turn: a chunk's products
// wait only for that chunk while the next one loads.  The loop has no exit
// between the two halves of a pair (an exit there lets the compiler move the
// prefetch down to its use); columns past the last full pair go through a
// plain tail loop.  The prefetch chunk is clamped to the last full one, so
// every load is in bounds.  A wave's rows past n repeat
// row n - 1 (in bounds) and are not stored.  Wave-shuffle reduction.
constexpr int RESID_RPW = 4;
constexpr int RESID_CH = 2;

template <int NR>
struct ResidChunk {
    double a[RESID_CH][RESID_RPW], x[RESID_CH][NR];
};

template <int NR>
__device__ __forceinline__ void resid_load(ResidChunk<NR> &b, const double *const (&a)[RESID_RPW],
                                           const double *__restrict__ X, int64_t ldx, int nrhs, int64_t k0,
                                           int lane) {
#pragma unroll
    for (int u = 0; u < RESID_CH; ++u) {
        const int64_t k = k0 + 64 * u + lane;
#pragma unroll
        for (int r = 0; r < RESID_RPW; ++r) b.a[u][r] = a[r][k];
#pragma unroll
        for (int c = 0; c < NR; ++c) b.x[u][c] = X[k * ldx + (c < nrhs ? c : 0)];
    }
}

template <int NR>
__device__ __forceinline__ void resid_fma(const ResidChunk<NR> &b, double (&acc)[RESID_RPW][NR]) {
#pragma unroll
    for (int u = 0; u < RESID_CH; ++u)
#pragma unroll
        for (int r = 0; r < RESID_RPW; ++r)
#pragma unroll
            for (int c = 0; c < NR; ++c) acc[r][c] = fma(b.a[u][r], b.x[u][c], acc[r][c]);
}

template <int NR>
__global__ __launch_bounds__(256) void resid_kernel(int64_t n, int nrhs, const double *__restrict__ A, int64_t lda,
                                                    const double *__restrict__ X, int64_t ldx,
                                                    const double *__restrict__ B, int64_t ldb,
                                                    double *__restrict__ R, int64_t ldr) {
    constexpr int RW = RESID_RPW, CW = RESID_CH * 64;
    const int lane = threadIdx.x & 63;
    const int64_t i0 = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * RW;
    if (i0 >= n) return;  // wave-uniform
    const double *a[RW];
#pragma unroll
    for (int r = 0; r < RW; ++r) a[r] = A + (i0 + r < n ? i0 + r : n - 1) * lda;
    double acc[RW][NR];
#pragma unroll
    for (int r = 0; r < RW; ++r)
#pragma unroll
        for (int c = 0; c < NR; ++c) acc[r][c] = 0.0;
    const int64_t nfull = n / (2 * CW) * (2 * CW), klast = nfull - CW;
    if (nfull > 0) {
        ResidChunk<NR> b0, b1;
        resid_load<NR>(b0, a, X, ldx, nrhs, 0, lane);
        for (int64_t k0 = 0; k0 < nfull; k0 += 2 * CW) {
            resid_load<NR>(b1, a, X, ldx, nrhs, k0 + CW, lane);
            __builtin_amdgcn_sched_barrier(0);  // keep the prefetch ahead of the products
            resid_fma<NR>(b0, acc);
            resid_load<NR>(b0, a, X, ldx, nrhs, k0 + 2 * CW < klast ? k0 + 2 * CW : klast, lane);
            __builtin_amdgcn_sched_barrier(0);
            resid_fma<NR>(b1, acc);
        }
    }
    for (int64_t k = nfull + lane; k < n; k += 64) {
#pragma unroll
        for (int r = 0; r < RW; ++r) {
            const double av = a[r][k];
#pragma unroll
            for (int c = 0; c < NR; ++c) acc[r][c] = fma(av, X[k * ldx + (c < nrhs ? c : 0)], acc[r][c]);
        }
    }
    double out = 0.0;
#pragma unroll
    for (int r = 0; r < RW; ++r)
#pragma unroll
        for (int c = 0; c < NR; ++c) {
            double v = acc[r][c];
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
            out = lane == r * NR + c ? v : out;
        }
    const int r = lane / NR, c = lane % NR;
    if (r < RW && i0 + r < n && c < nrhs) R[(i0 + r) * ldr + c] = B[(i0 + r) * ldb + c] - out;
}

void colmax(hipStream_t s, const double *X, int64_t ldx, int64_t n, int ncols, double *out) {
    if (ncols <= COLMAX_FEW) {
        hipLaunchKernelGGL(colmax_few_kernel, dim3((unsigned)cdiv(n, 4096)), dim3(256), 0, s, X, ldx, n, ncols, out);
        LMX_LAUNCH_CHECK();
        return;
    }
    dim3 g((unsigned)cdiv(n, COLMAX_ROWS), (unsigned)cdiv(ncols, 64));
    hipLaunchKernelGGL(colmax_kernel, g, dim3(256), 0, s, X, ldx, n, ncols, out);
    LMX_LAUNCH_CHECK();
}

// R = B - A X (FP64).
void residual(hipStream_t s, int64_t n, int64_t k, const double *A, int64_t lda, const double *X, int64_t ldx,
              const double *B, int64_t ldb, double *R, int64_t ldr) {
    if (k <= 4) {
        const dim3 g((unsigned)cdiv(n, 4 * RESID_RPW));
        if (k == 1)
            hipLaunchKernelGGL(resid_kernel<1>, g, dim3(256), 0, s, n, 1, A, lda, X, ldx, B, ldb, R, ldr);
        else
            hipLaunchKernelGGL(resid_kernel<4>, g, dim3(256), 0, s, n, (int)k, A, lda, X, ldx, B, ldb, R, ldr);
        LMX_LAUNCH_CHECK();
        return;
    }
    hipLaunchKernelGGL(copy_d_kernel, dim3((unsigned)cdiv(n * k, 256)), dim3(256), 0, s, B, ldb, R, ldr, n, k);
    LMX_LAUNCH_CHECK();
    gemm(s, false, false, n, k, n, -1.0, A, lda, X, ldx, 1.0, R, ldr);
}

// norms layout: [0] ||A||, [1..1+k) ||b_c||, [1+k..1+2k) ||x_c||, [1+2k..1+3k) ||r_c||,
// [1+3k] eta, [2+3k] all-columns-pass flag (1.0 / 0.0)
__global__ void eta_kernel(double *norms, int k, double crit) {
    if (threadIdx.x != 0) return;
    double nA = norms[0], eta = 0.0, pass = 1.0;
    for (int c = 0; c < k; ++c) {
        double nb = norms[1 + c], nx = norms[1 + k + c], nr = norms[1 + 2 * k + c];
        double den = nA * nx + nb;
        double e = den > 0.0 ? nr / den : (nr > 0.0 ? __builtin_huge_val() : 0.0);
        if (!(e <= eta)) eta = e;  // keeps NaN
        if (!(nr <= crit * nA * nx)) pass = 0.0;
    }
    norms[1 + 3 * k] = eta;
    norms[2 + 3 * k] = pass;
}

template <class F>
void launch1d(hipStream_t s, int64_t elems, F launch) {
    if (elems > 0) launch(dim3((unsigned)cdiv(elems, 256)));
}

}  // namespace

namespace {

constexpr int64_t INVERSE_MIN_N = 256;  // smaller factors solve by substitution

size_t round256(size_t b) { return (b + 255) / 256 * 256; }
size_t inverse_bytes(int64_t n) {
    return n > INVERSE_MIN_N ? round256((size_t)tri_inverse_elems(n) * sizeof(float)) : 0;
}

}  // namespace

// Layout: A32 | anorm (256 B) | ipiv, perm, info | WL | WU.
size_t low_factor_bytes(int64_t n) {
    return round256((size_t)n * (size_t)n * sizeof(float)) + 256 + round256((size_t)(2 * n + 8) * sizeof(int32_t)) +
           2 * inverse_bytes(n);
}

void low_factor_bind(LowFactor &f, void *mem) {
    char *p = static_cast<char *>(mem);
    const size_t a32 = round256((size_t)f.n * (size_t)f.n * sizeof(float));
    const size_t ints = round256((size_t)(2 * f.n + 8) * sizeof(int32_t));
    const size_t w = inverse_bytes(f.n);
    f.A32 = reinterpret_cast<float *>(p);
    f.anorm = reinterpret_cast<double *>(p + a32);
    f.ipiv = reinterpret_cast<int32_t *>(p + a32 + 256);
    f.perm = f.ipiv + f.n;
    f.dinfo = f.perm + f.n;
    f.WL = w ? reinterpret_cast<float *>(p + a32 + 256 + ints) : nullptr;
    f.WU = w ? reinterpret_cast<float *>(p + a32 + 256 + ints + w) : nullptr;
}

// LMX_LU_PIVOTING=tournament selects tournament pivoting for the FP32 LU
// factor (read once per process).
bool tournament_pivoting() {
    static const bool on = [] {
        const char *e = std::getenv("LMX_LU_PIVOTING");
        return e != nullptr && std::strcmp(e, "tournament") == 0;
    }();
    return on;
}

void low_factor(hipStream_t s, LowFactor &f) {
    const int64_t n = f.n;
    launch1d(s, n * n, [&](dim3 g) { hipLaunchKernelGGL(d2f_kernel, g, dim3(256), 0, s, f.A, f.lda, f.A32, n, n, n); });
    LMX_LAUNCH_CHECK();
    if (f.spd)
        potrf<float>(s, n, f.A32, n, f.dinfo);
    else
        getrf_impl<float>(s, n, f.A32, n, f.ipiv, f.perm, f.dinfo, tournament_pivoting());
    // inverses of the factor's diagonal blocks for the refinement solves (a
    // failed factor leaves them unused)
    if (f.WL) {
        if (f.spd) {
            tri_inverse_build<float>(s, true, false, false, n, f.A32, n, f.WL);
            tri_inverse_build<float>(s, true, true, false, n, f.A32, n, f.WU);
        } else {
            tri_inverse_build<float>(s, true, false, true, n, f.A32, n, f.WL);
            tri_inverse_build<float>(s, false, false, false, n, f.A32, n, f.WU);
        }
    }
    hipLaunchKernelGGL(zero_d_kernel, dim3(1), dim3(256), 0, s, f.anorm, 1);
    hipLaunchKernelGGL(norm_inf_rows_kernel, dim3((unsigned)std::min<int64_t>(cdiv(n, 4), 1024)), dim3(256), 0, s, f.A,
                       f.lda, n, f.anorm);
    LMX_LAUNCH_CHECK();
    int32_t hinfo = 0;
    LMX_HIP(hipMemcpyAsync(&hinfo, f.dinfo, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    LMX_HIP(hipStreamSynchronize(s));
    f.info = hinfo;
}

RefineOut low_refine(hipStream_t s, const LowFactor &f, int64_t k, const double *B, int64_t ldb, double *X,
                     int64_t ldx) {
    RefineOut out;
    const int64_t n = f.n;
    out.info = f.info;
    if (n == 0 || k == 0) {
        out.converged = true;
        return out;
    }
    const size_t nk = (size_t)n * (size_t)k;
    Scratch sR(nk * sizeof(double), s), sR32(nk * sizeof(float), s), sN((size_t)(3 * k + 8) * sizeof(double), s),
        sI((size_t)(2 * n + 8) * sizeof(int32_t), s);
    float *R32 = sR32.as<float>();
    double *R = sR.as<double>(), *norms = sN.as<double>();
    const double *A = f.A;
    const int64_t lda = f.lda;
    const int K = (int)k;
    const double crit = std::sqrt((double)n) * EPS64;
    if (f.info == 0) {
        hipLaunchKernelGGL(zero_d_kernel, dim3((unsigned)cdiv(3 * K + 8, 256)), dim3(256), 0, s, norms, 3 * K + 8);
        LMX_LAUNCH_CHECK();
        hipLaunchKernelGGL(copy_d_kernel, dim3(1), dim3(256), 0, s, f.anorm, 1, norms, 1, 1, 1);
        colmax(s, B, ldb, n, K, norms + 1);
        LMX_LAUNCH_CHECK();
        // R = B
        launch1d(s, (int64_t)nk, [&](dim3 g) { hipLaunchKernelGGL(copy_d_kernel, g, dim3(256), 0, s, B, ldb, R, k, n, k); });
        double prev = INFINITY;
        for (int it = 0; it < ITMAX; ++it) {
            launch1d(s, (int64_t)nk, [&](dim3 g) { hipLaunchKernelGGL(d2f_kernel, g, dim3(256), 0, s, R, k, R32, k, n, k); });
            if (f.spd)
                potrs<float>(s, n, k, f.A32, n, R32, k, f.WL, f.WU);
            else
                getrs_perm<float>(s, n, k, f.A32, n, f.perm, R32, k, f.WL, f.WU);
            launch1d(s, (int64_t)nk, [&](dim3 g) {
                hipLaunchKernelGGL(f2d_acc_kernel, g, dim3(256), 0, s, R32, k, X, ldx, n, k, it == 0 ? 1 : 0);
            });
            // R = B - A X
            residual(s, n, k, A, lda, X, ldx, B, ldb, R, k);
            hipLaunchKernelGGL(zero_d_kernel, dim3((unsigned)cdiv(2 * K + 2, 256)), dim3(256), 0, s, norms + 1 + K, 2 * K + 2);
            colmax(s, X, ldx, n, K, norms + 1 + K);
            colmax(s, R, k, n, K, norms + 1 + 2 * K);
            hipLaunchKernelGGL(eta_kernel, dim3(1), dim3(64), 0, s, norms, K, crit);
            LMX_LAUNCH_CHECK();
            double h[2];
            LMX_HIP(hipMemcpyAsync(h, norms + 1 + 3 * K, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
            LMX_HIP(hipStreamSynchronize(s));
            out.iterations = it + 1;
            out.eta = h[0];
            if (h[1] == 1.0) {
                out.converged = true;
                break;
            }
            if (!(h[0] < 0.9 * prev)) break;  // stagnation (or NaN)
            prev = h[0];
        }
    }
    if (!out.converged) {
        // FP64 factorisation of a copy of A.
        out.fell_back = true;
        const size_t nn = (size_t)n * (size_t)n;
        Scratch sA64(nn * sizeof(double), s);
        double *A64 = sA64.as<double>();
        int32_t *ipiv = sI.as<int32_t>(), *perm = ipiv + n, *dinfo = perm + n;
        launch1d(s, (int64_t)nn, [&](dim3 g) { hipLaunchKernelGGL(copy_d_kernel, g, dim3(256), 0, s, A, lda, A64, n, n, n); });
        launch1d(s, (int64_t)nk, [&](dim3 g) { hipLaunchKernelGGL(copy_d_kernel, g, dim3(256), 0, s, B, ldb, X, ldx, n, k); });
        if (f.spd) {
            potrf<double>(s, n, A64, n, dinfo);
            potrs<double>(s, n, k, A64, n, X, ldx);
        } else {
            getrf_impl<double>(s, n, A64, n, ipiv, perm, dinfo);
            getrs_perm<double>(s, n, k, A64, n, perm, X, ldx);
        }
        // report the backward error of the FP64 solve
        residual(s, n, k, A, lda, X, ldx, B, ldb, R, k);
        hipLaunchKernelGGL(zero_d_kernel, dim3((unsigned)cdiv(3 * K + 8, 256)), dim3(256), 0, s, norms, 3 * K + 8);
        hipLaunchKernelGGL(norm_inf_rows_kernel, dim3((unsigned)std::min<int64_t>(cdiv(n, 4), 1024)), dim3(256), 0, s,
                           A, lda, n, norms);
        colmax(s, B, ldb, n, K, norms + 1);
        colmax(s, X, ldx, n, K, norms + 1 + K);
        colmax(s, R, k, n, K, norms + 1 + 2 * K);
        hipLaunchKernelGGL(eta_kernel, dim3(1), dim3(64), 0, s, norms, K, crit);
        LMX_LAUNCH_CHECK();
        double h[2];
        int32_t i64 = 0;
        LMX_HIP(hipMemcpyAsync(h, norms + 1 + 3 * K, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
        LMX_HIP(hipMemcpyAsync(&i64, dinfo, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        LMX_HIP(hipStreamSynchronize(s));
        out.eta = h[0];
        out.info = i64;
        out.converged = (h[1] == 1.0);
    }
    return out;
}

// precision: LMX_PREC_F32 (factor in FP32).
RefineOut refine_solve(hipStream_t s, bool spd, int precision, int64_t n, int64_t k, const double *A,
                       int64_t lda, const double *B, int64_t ldb, double *X, int64_t ldx) {
    RefineOut out;
    if (precision != LMX_PREC_F32) fail_arg("mixed solve: factor precision must be LMX_PREC_F32");
    if (n == 0 || k == 0) {
        out.converged = true;
        return out;
    }
    TimedRegion trall("refine", s);
    Scratch mem(low_factor_bytes(n), s);
    LowFactor f;
    f.spd = spd;
    f.n = n;
    f.A = A;
    f.lda = lda;
    low_factor_bind(f, mem.p);
    low_factor(s, f);
    return low_refine(s, f, k, B, ldb, X, ldx);
}

namespace {

void fill(lmx_refine_info *info, const RefineOut &o) {
    if (!info) return;
    info->iterations = o.iterations;
    info->converged = o.converged ? 1 : 0;
    info->fell_back = o.fell_back ? 1 : 0;
    info->info = o.info;
    info->backward_error = o.eta;
}

void check_args(int64_t n, int64_t k, int64_t lda, int64_t ldb, int64_t ldx) {
    if (n < 0 || k < 0) fail_arg("mixed solve: negative dimension");
    if (lda < std::max<int64_t>(1, n) || ldb < std::max<int64_t>(1, k) || ldx < std::max<int64_t>(1, k))
        fail_arg("mixed solve: bad stride");
}

int host_mixed(bool spd, int precision, int64_t n, int64_t k, const double *A, int64_t lda, const double *B,
               int64_t ldb, double *X, int64_t ldx, lmx_refine_info *info) {
    check_args(n, k, lda, ldb, ldx);
    if (n == 0 || k == 0) {
        fill(info, RefineOut{0, true, false, 0, 0.0});
        return LMX_OK;
    }
    Context &c = ctx();
    std::lock_guard<std::mutex> g(c.host_mu);
    auto ext = [](int64_t r, int64_t cc, int64_t ld) { return (size_t)((r - 1) * ld + cc) * sizeof(double); };
    size_t ab = ext(n, n, lda), bb = ext(n, k, ldb), xb = ext(n, k, ldx);
    double *dA = (double *)staging(c, 0, ab), *dB = (double *)staging(c, 1, bb), *dX = (double *)staging(c, 2, xb);
    LMX_HIP(hipMemcpyAsync(dA, A, ab, hipMemcpyHostToDevice, c.stream));
    LMX_HIP(hipMemcpyAsync(dB, B, bb, hipMemcpyHostToDevice, c.stream));
    if (ldx > k) LMX_HIP(hipMemcpyAsync(dX, X, xb, hipMemcpyHostToDevice, c.stream));
    RefineOut o = refine_solve(c.stream, spd, precision, n, k, dA, lda, dB, ldb, dX, ldx);
    LMX_HIP(hipMemcpyAsync(X, dX, xb, hipMemcpyDeviceToHost, c.stream));
    LMX_HIP(hipStreamSynchronize(c.stream));
    fill(info, o);
    if (o.info < 0) throw Error(LMX_ERR_INTERNAL, "panel synchronisation timed out");
    return o.converged ? (o.fell_back ? o.info : LMX_OK) : LMX_ERR_NOT_CONVERGED;
}

}  // namespace

}  // namespace lmx

using namespace lmx;

extern "C" {

int lmx_dgesv_mixed_dev(void *stream, int precision, int64_t n, int64_t nrhs, const double *dA, int64_t lda,
                        const double *dB, int64_t ldb, double *dX, int64_t ldx, lmx_refine_info *info) {
    return guarded([&] {
        check_args(n, nrhs, lda, ldb, ldx);
        RefineOut o = refine_solve(resolve_stream(stream), false, precision, n, nrhs, dA, lda, dB, ldb, dX, ldx);
        fill(info, o);
        return o.converged ? (o.fell_back ? o.info : LMX_OK) : LMX_ERR_NOT_CONVERGED;
    });
}

int lmx_dposv_mixed_dev(void *stream, int precision, int64_t n, int64_t nrhs, const double *dA, int64_t lda,
                        const double *dB, int64_t ldb, double *dX, int64_t ldx, lmx_refine_info *info) {
    return guarded([&] {
        check_args(n, nrhs, lda, ldb, ldx);
        RefineOut o = refine_solve(resolve_stream(stream), true, precision, n, nrhs, dA, lda, dB, ldb, dX, ldx);
        fill(info, o);
        return o.converged ? (o.fell_back ? o.info : LMX_OK) : LMX_ERR_NOT_CONVERGED;
    });
}

int lmx_dgesv_mixed(int precision, int64_t n, int64_t nrhs, const double *A, int64_t lda, const double *B,
                    int64_t ldb, double *X, int64_t ldx, lmx_refine_info *info) {
    return guarded([&] { return host_mixed(false, precision, n, nrhs, A, lda, B, ldb, X, ldx, info); });
}

int lmx_dposv_mixed(int precision, int64_t n, int64_t nrhs, const double *A, int64_t lda, const double *B,
                    int64_t ldb, double *X, int64_t ldx, lmx_refine_info *info) {
    return guarded([&] { return host_mixed(true, precision, n, nrhs, A, lda, B, ldb, X, ldx, info); });
}

}  // extern "C"
