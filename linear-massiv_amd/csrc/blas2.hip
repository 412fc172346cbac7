// Level-2 work on the device: conversions and copies, the FP64 residual
// R = B - A X, matrix-vector products, norms and the normwise backward
// error, plus the consumers of an LU factor that need nothing else (the
// determinant and inverse iteration).
//
// The residual / GEMV kernel (resid_kernel) is HBM-bound: every element of
// A is read once per call (m n 8 bytes), four rows per wave share each load
// of x, and the column loop is pipelined one chunk ahead.
#include <cmath>
#include <vector>

#include "common.hpp"
#include "device.hpp"
#include "kernels.hpp"

namespace lmx {

namespace {

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
                                                            int64_t m, int64_t n, double *__restrict__ out) {
    __shared__ double part[4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t nw = (int64_t)gridDim.x * 4;
    double best = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 4 + wave; i < m; i += nw) {
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

// What resid_kernel does with out = (A X)[i][c]:
//   RESID_STORE: R[i][c] = B[i][c] - out;
//   RESID_GEMV:  R[i][c] = alpha out + beta R[i][c] (beta = 0 does not read R);
//   RESID_MAX:   colmax[c] = max(colmax[c], |B[i][c] - out|), R not written.
enum { RESID_STORE = 0, RESID_GEMV = 1, RESID_MAX = 2 };

template <int NR, int MODE>
__global__ __launch_bounds__(256) void resid_kernel(int64_t m, int64_t n, int nrhs, const double *__restrict__ A,
                                                    int64_t lda, const double *__restrict__ X, int64_t ldx,
                                                    const double *__restrict__ B, int64_t ldb,
                                                    double *__restrict__ R, int64_t ldr, double alpha, double beta,
                                                    double *__restrict__ colmax_out) {
    constexpr int RW = RESID_RPW, CW = RESID_CH * 64;
    __shared__ double wmax[4][NR];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t i0 = ((int64_t)blockIdx.x * 4 + wave) * RW;
    const bool active = i0 < m;  // wave-uniform
    double out = 0.0;
    if (active) {
        const double *a[RW];
#pragma unroll
        for (int r = 0; r < RW; ++r) a[r] = A + (i0 + r < m ? i0 + r : m - 1) * lda;
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
#pragma unroll
        for (int r = 0; r < RW; ++r)
#pragma unroll
            for (int c = 0; c < NR; ++c) {
                double v = acc[r][c];
#pragma unroll
                for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
                out = lane == r * NR + c ? v : out;
            }
    }
    const int r = lane / NR, c = lane % NR;
    const bool mine = active && r < RW && i0 + r < m && c < nrhs;
    if constexpr (MODE == RESID_STORE) {
        if (mine) R[(i0 + r) * ldr + c] = B[(i0 + r) * ldb + c] - out;
    } else if constexpr (MODE == RESID_GEMV) {
        if (mine) {
            double y = alpha * out;
            if (beta != 0.0) y += beta * R[(i0 + r) * ldr + c];
            R[(i0 + r) * ldr + c] = y;
        }
    } else {
        // |r| of this wave's RW rows, reduced over the lanes of one column
        // (lanes c, NR + c, 2 NR + c, 3 NR + c), then over the workgroup's
        // waves in LDS; one atomic per column and workgroup.  NaN becomes
        // +inf, as the atomic would make it.
        double v = mine ? fabs(B[(i0 + r) * ldb + c] - out) : 0.0;
        if (isnan(v)) v = __builtin_huge_val();
#pragma unroll
        for (int off = NR; off < RW * NR; off <<= 1) v = fmax(v, __shfl_xor(v, off));
        if (lane < NR) wmax[wave][lane] = v;
        __syncthreads();
        if (threadIdx.x < NR && (int)threadIdx.x < nrhs)
            atomic_max_nonneg(&colmax_out[threadIdx.x],
                              fmax(fmax(wmax[0][threadIdx.x], wmax[1][threadIdx.x]),
                                   fmax(wmax[2][threadIdx.x], wmax[3][threadIdx.x])));
    }
}


__global__ void scale_d_kernel(double *y, int64_t m, double beta) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < m) y[i] = beta != 0.0 ? beta * y[i] : 0.0;
}

// max_j sum_i |a_ij| (the 1-norm): a workgroup owns 64 columns; wave w sums
// rows w, w + 4, ... of each (lane = column, coalesced row segments), the
// four partial sums are added in a fixed order, then one atomic max per
// column block.  Every column sum is formed in the same order on every run.
__global__ __launch_bounds__(256) void norm1_cols_kernel(const double *__restrict__ A, int64_t lda, int64_t m,
                                                         int64_t n, double *__restrict__ out) {
    __shared__ double part[4][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t c = (int64_t)blockIdx.x * 64 + lane;
    double sum = 0.0;
    if (c < n)
        for (int64_t i = wave; i < m; i += 4) sum += fabs(A[i * lda + c]);
    part[wave][lane] = sum;
    __syncthreads();
    if (wave == 0) {
        double v = c < n ? (part[0][lane] + part[1][lane]) + (part[2][lane] + part[3][lane]) : 0.0;
        if (isnan(v)) v = __builtin_huge_val();
        v = wave_max(v);
        if (lane == 0) atomic_max_nonneg(out, v);
    }
}

// max_ij |a_ij|: rows strided over the grid, one atomic per workgroup.
__global__ __launch_bounds__(256) void absmax_kernel(const double *__restrict__ A, int64_t lda, int64_t m, int64_t n,
                                                     double *__restrict__ out) {
    __shared__ double part[4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double best = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 4 + wave; i < m; i += (int64_t)gridDim.x * 4)
        for (int64_t j = lane; j < n; j += 64) {
            double v = fabs(A[i * lda + j]);
            if (isnan(v)) v = __builtin_huge_val();
            best = fmax(best, v);
        }
    best = wave_max(best);
    if (lane == 0) part[wave] = best;
    __syncthreads();
    if (threadIdx.x == 0) atomic_max_nonneg(out, fmax(fmax(part[0], part[1]), fmax(part[2], part[3])));
}

// Sums of squares per workgroup (rows strided over a fixed grid, a fixed
// order inside it) into partial[blockIdx.x]; frob_finish adds the partials
// in order and takes the square root, so the result does not vary by run.
constexpr int FROB_GRID = 1024;
__global__ __launch_bounds__(256) void frob_partial_kernel(const double *__restrict__ A, int64_t lda, int64_t m,
                                                           int64_t n, double *__restrict__ partial) {
    __shared__ double part[4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double sum = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 4 + wave; i < m; i += (int64_t)gridDim.x * 4)
        for (int64_t j = lane; j < n; j += 64) {
            const double v = A[i * lda + j];
            sum = fma(v, v, sum);
        }
    sum = wave_sum(sum);
    if (lane == 0) part[wave] = sum;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = (part[0] + part[1]) + (part[2] + part[3]);
}

__global__ void frob_finish_kernel(const double *__restrict__ partial, int count, double *__restrict__ out) {
    __shared__ double part[4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double sum = 0.0;
    for (int i = threadIdx.x; i < count; i += 256) sum += partial[i];
    sum = wave_sum(sum);
    if (lane == 0) part[wave] = sum;
    __syncthreads();
    if (threadIdx.x == 0) *out = sqrt((part[0] + part[1]) + (part[2] + part[3]));
}

// eta_c = ||r_c|| / (||A|| ||x_c|| + ||b_c|| + n eps), the normwise backward
// error of scaledResidualLinear (test/Test/Residuals.hs:69-78).
// acc layout: [0] ||A||_inf, [1, 1 + k) ||b_c||, [1 + k, 1 + 2k) ||x_c||,
// [1 + 2k, 1 + 3k) ||r_c||.
__global__ void scaled_residual_finish_kernel(const double *__restrict__ acc, int k, double neps,
                                              double *__restrict__ eta) {
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= k) return;
    const double na = acc[0], nb = acc[1 + c], nx = acc[1 + k + c], nr = acc[1 + 2 * k + c];
    eta[c] = nr / (na * nx + nb + neps);
}

// Dot product in one workgroup: each thread sums its strided elements in
// order, then a fixed tree; the result does not vary by run.  `root` takes
// the square root (nrm2).
__global__ __launch_bounds__(1024) void dot_kernel(const double *__restrict__ x, const double *__restrict__ y,
                                                   int64_t n, double *__restrict__ out, int root) {
    __shared__ double part[16];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double sum = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += 1024) sum = fma(x[i], y[i], sum);
    sum = wave_sum(sum);
    if (lane == 0) part[wave] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < 16; ++w) t += part[w];
        *out = root ? sqrt(t) : t;
    }
}

// y = x * (1 / *nrm) (scal (1 / znorm) z: the reciprocal, then products).
__global__ void scale_by_recip_kernel(const double *__restrict__ x, const double *__restrict__ nrm,
                                      double *__restrict__ y, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) y[i] = x[i] * (1.0 / *nrm);
}

// A - mu I into S (n x n, stride n).
__global__ void shift_copy_kernel(const double *__restrict__ A, int64_t lda, double mu, double *__restrict__ S,
                                  int64_t n) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= n * n) return;
    const int64_t i = e / n, j = e % n;
    const double a = A[i * lda + j];
    S[i * n + j] = i == j ? a - mu : a;
}

// det = sign * prod_i u_ii with the product formed left to right, as det's
// foldl does (Solve/LU.hs:534-545), and the sign from the parity of the swap
// list (the parity of lu's permutation, which countSwaps measures).  One
// wave: the next 64 diagonal entries load while lane 0 multiplies the
// current ones in order; the swaps are counted by ballots.
__global__ __launch_bounds__(64) void det_kernel(const double *__restrict__ LU, int64_t lda,
                                                 const int32_t *__restrict__ ipiv, int64_t n,
                                                 double *__restrict__ det) {
    const int lane = threadIdx.x;
    int swaps = 0;
    for (int64_t k0 = 0; k0 < n; k0 += 64) {
        const int64_t k = k0 + lane;
        swaps += __popcll(__ballot(k < n && ipiv[k] != (int32_t)k));
    }
    double prod = 1.0;
    double cur = lane < n ? LU[(int64_t)lane * lda + lane] : 1.0;
    for (int64_t k0 = 0; k0 < n; k0 += 64) {
        const int64_t kn = k0 + 64 + lane;
        const double next = kn < n ? LU[kn * lda + kn] : 1.0;
        const int cnt = (int)(n - k0 < 64 ? n - k0 : 64);
        for (int q = 0; q < cnt; ++q) prod *= lane_value(cur, q);
        cur = next;
    }
    if (lane == 0) *det = (swaps & 1) ? -prod : prod;
}

}  // namespace

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


// R = B - A X (FP64), A n x n, X and B n x k.
void residual(hipStream_t s, int64_t n, int64_t k, const double *A, int64_t lda, const double *X, int64_t ldx,
              const double *B, int64_t ldb, double *R, int64_t ldr) {
    if (n == 0 || k == 0) return;
    if (k <= 4) {
        const dim3 g((unsigned)cdiv(n, 4 * RESID_RPW));
        if (k == 1)
            hipLaunchKernelGGL((resid_kernel<1, RESID_STORE>), g, dim3(256), 0, s, n, n, 1, A, lda, X, ldx, B, ldb, R,
                               ldr, 0.0, 0.0, nullptr);
        else
            hipLaunchKernelGGL((resid_kernel<4, RESID_STORE>), g, dim3(256), 0, s, n, n, (int)k, A, lda, X, ldx, B,
                               ldb, R, ldr, 0.0, 0.0, nullptr);
        LMX_LAUNCH_CHECK();
        return;
    }
    copy_d(s, B, ldb, R, ldr, n, k);
    gemm(s, false, false, n, k, n, -1.0, A, lda, X, ldx, 1.0, R, ldr);
}

// colmax_out[c] = max(colmax_out[c], max_i |(B - A X)[i][c]|) without storing
// the residual; A is m x n, columns in groups of four.
void residual_colmax(hipStream_t s, int64_t m, int64_t n, int64_t k, const double *A, int64_t lda, const double *X,
                     int64_t ldx, const double *B, int64_t ldb, double *colmax_out) {
    if (m == 0 || k == 0) return;
    const dim3 g((unsigned)cdiv(m, 4 * RESID_RPW));
    for (int64_t c0 = 0; c0 < k; c0 += 4) {
        const int kc = (int)std::min<int64_t>(4, k - c0);
        if (kc == 1)
            hipLaunchKernelGGL((resid_kernel<1, RESID_MAX>), g, dim3(256), 0, s, m, n, 1, A, lda, X + c0, ldx, B + c0,
                               ldb, nullptr, 0, 0.0, 0.0, colmax_out + c0);
        else
            hipLaunchKernelGGL((resid_kernel<4, RESID_MAX>), g, dim3(256), 0, s, m, n, kc, A, lda, X + c0, ldx,
                               B + c0, ldb, nullptr, 0, 0.0, 0.0, colmax_out + c0);
        LMX_LAUNCH_CHECK();
    }
}

// y = alpha A x + beta y, A m x n (x, y unit stride).
void gemv(hipStream_t s, int64_t m, int64_t n, double alpha, const double *A, int64_t lda, const double *x,
          double beta, double *y) {
    if (m == 0) return;
    if (n == 0) {
        // alpha A x is empty: y = beta y
        hipLaunchKernelGGL(scale_d_kernel, dim3((unsigned)cdiv(m, 256)), dim3(256), 0, s, y, m, beta);
        LMX_LAUNCH_CHECK();
        return;
    }
    TimedRegion tr("gemv", s, 2.0 * (double)m * (double)n);
    hipLaunchKernelGGL((resid_kernel<1, RESID_GEMV>), dim3((unsigned)cdiv(m, 4 * RESID_RPW)), dim3(256), 0, s, m, n,
                       1, A, lda, x, 1, nullptr, 0, y, 1, alpha, beta, nullptr);
    LMX_LAUNCH_CHECK();
}

void matmul_block(hipStream_t s, int64_t m, int64_t n, int64_t k, const double *A, int64_t lda, const double *X,
                  int64_t ldx, double *Y, int64_t ldy) {
    if (m == 0 || k == 0) return;
    if (k > 4 || n == 0) {
        gemm(s, false, false, m, k, n, 1.0, A, lda, X, ldx, 0.0, Y, ldy);
        return;
    }
    const dim3 g((unsigned)cdiv(m, 4 * RESID_RPW));
    if (k == 1)
        hipLaunchKernelGGL((resid_kernel<1, RESID_GEMV>), g, dim3(256), 0, s, m, n, 1, A, lda, X, ldx, nullptr, 0, Y,
                           ldy, 1.0, 0.0, nullptr);
    else
        hipLaunchKernelGGL((resid_kernel<4, RESID_GEMV>), g, dim3(256), 0, s, m, n, (int)k, A, lda, X, ldx, nullptr,
                           0, Y, ldy, 1.0, 0.0, nullptr);
    LMX_LAUNCH_CHECK();
}

void d2f(hipStream_t s, const double *in, int64_t ldi, float *out, int64_t ldo, int64_t rows, int64_t cols) {
    if (rows * cols == 0) return;
    hipLaunchKernelGGL(d2f_kernel, dim3((unsigned)cdiv(rows * cols, 256)), dim3(256), 0, s, in, ldi, out, ldo, rows,
                       cols);
    LMX_LAUNCH_CHECK();
}

void f2d_acc(hipStream_t s, const float *D, int64_t ldd, double *X, int64_t ldx, int64_t rows, int64_t cols,
             bool assign) {
    if (rows * cols == 0) return;
    hipLaunchKernelGGL(f2d_acc_kernel, dim3((unsigned)cdiv(rows * cols, 256)), dim3(256), 0, s, D, ldd, X, ldx, rows,
                       cols, assign ? 1 : 0);
    LMX_LAUNCH_CHECK();
}

void copy_d(hipStream_t s, const double *in, int64_t ldi, double *out, int64_t ldo, int64_t rows, int64_t cols) {
    if (rows * cols == 0) return;
    hipLaunchKernelGGL(copy_d_kernel, dim3((unsigned)cdiv(rows * cols, 256)), dim3(256), 0, s, in, ldi, out, ldo,
                       rows, cols);
    LMX_LAUNCH_CHECK();
}

void zero_d(hipStream_t s, double *p, int64_t n) {
    if (n <= 0) return;
    hipLaunchKernelGGL(zero_d_kernel, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, s, p, (int)n);
    LMX_LAUNCH_CHECK();
}

void norm_inf_acc(hipStream_t s, const double *A, int64_t lda, int64_t m, int64_t n, double *out) {
    if (m == 0) return;
    hipLaunchKernelGGL(norm_inf_rows_kernel, dim3((unsigned)std::min<int64_t>(cdiv(m, 4), 1024)), dim3(256), 0, s, A,
                       lda, m, n, out);
    LMX_LAUNCH_CHECK();
}

namespace {

__global__ void mirror_lower_kernel(double *C, int64_t ldc, int64_t n) {
    // C[i][j] = C[j][i] for j > i; tiles of 32 x 32 through LDS so both the
    // reads and the writes run along rows
    __shared__ double t[32][33];
    const int64_t bi = (int64_t)blockIdx.y * 32, bj = (int64_t)blockIdx.x * 32;
    if (bj < bi) return;  // only tiles touching the upper triangle
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8 threads
    for (int r = ty; r < 32; r += 8) {
        const int64_t i = bj + r, j = bi + tx;  // source: lower element (i, j) with i >= j
        t[r][tx] = (i < n && j < n) ? C[i * ldc + j] : 0.0;
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int64_t i = bi + r, j = bj + tx;  // destination (i, j), j > i
        if (i < n && j < n && j > i) C[i * ldc + j] = t[tx][r];
    }
}

void check_mat(int64_t m, int64_t n, int64_t ld, const char *who) {
    if (m < 0 || n < 0) fail_arg(std::string(who) + ": negative dimension");
    if (ld < std::max<int64_t>(1, n)) fail_arg(std::string(who) + ": leading dimension too small");
}

size_t bytes_of(int64_t rows, int64_t cols, int64_t ld) {
    return (rows > 0 && cols > 0) ? (size_t)((rows - 1) * ld + cols) * sizeof(double) : 0;
}

}  // namespace

void matrix_norm(hipStream_t s, char which, int64_t m, int64_t n, const double *A, int64_t lda, double *d_out) {
    zero_d(s, d_out, 1);
    if (m == 0 || n == 0) return;
    switch (which) {
    case 'I':
    case 'i':
        norm_inf_acc(s, A, lda, m, n, d_out);
        return;
    case '1':
    case 'O':
    case 'o':
        hipLaunchKernelGGL(norm1_cols_kernel, dim3((unsigned)cdiv(n, 64)), dim3(256), 0, s, A, lda, m, n, d_out);
        LMX_LAUNCH_CHECK();
        return;
    case 'M':
    case 'm':
        hipLaunchKernelGGL(absmax_kernel, dim3((unsigned)std::min<int64_t>(cdiv(m, 4), 1024)), dim3(256), 0, s, A,
                           lda, m, n, d_out);
        LMX_LAUNCH_CHECK();
        return;
    case 'F':
    case 'f':
    case 'E':
    case 'e': {
        Scratch part(FROB_GRID * sizeof(double), s);
        const int grid = (int)std::min<int64_t>(cdiv(m, 4), FROB_GRID);
        hipLaunchKernelGGL(frob_partial_kernel, dim3(grid), dim3(256), 0, s, A, lda, m, n, part.as<double>());
        hipLaunchKernelGGL(frob_finish_kernel, dim3(1), dim3(256), 0, s, part.as<double>(), grid, d_out);
        LMX_LAUNCH_CHECK();
        return;
    }
    default:
        fail_arg("norm: `which` must be '1', 'I', 'F' or 'M'");
    }
}

// The normwise backward error of each column, accumulating into acc
// (1 + 3k doubles of device memory, see scaled_residual_finish_kernel) and R
// (n x k doubles, used only for k > 16; may be null otherwise).
void scaled_residual(hipStream_t s, int64_t n, int64_t k, const double *A, int64_t lda, const double *X, int64_t ldx,
                     const double *B, int64_t ldb, double *d_eta, double *acc, double *R) {
    if (k == 0) return;
    zero_d(s, acc, 1 + 3 * k);
    norm_inf_acc(s, A, lda, n, n, acc);
    colmax(s, B, ldb, n, (int)k, acc + 1);
    colmax(s, X, ldx, n, (int)k, acc + 1 + k);
    if (k <= SCALED_RESIDUAL_FUSED_MAX) {
        residual_colmax(s, n, n, k, A, lda, X, ldx, B, ldb, acc + 1 + 2 * k);
    } else {
        residual(s, n, k, A, lda, X, ldx, B, ldb, R, k);
        colmax(s, R, k, n, (int)k, acc + 1 + 2 * k);
    }
    hipLaunchKernelGGL(scaled_residual_finish_kernel, dim3((unsigned)cdiv(k, 256)), dim3(256), 0, s, acc, (int)k,
                       (double)n * 2.220446049250313e-16, d_eta);
    LMX_LAUNCH_CHECK();
}

void syrk_full(hipStream_t s, bool trans, int64_t n, int64_t k, double alpha, const double *A, int64_t lda,
               double beta, double *C, int64_t ldc) {
    if (n == 0) return;
    syrk_lower_op(s, trans, n, k, alpha, A, lda, beta, C, ldc);
    const dim3 g((unsigned)cdiv(n, 32), (unsigned)cdiv(n, 32));
    hipLaunchKernelGGL(mirror_lower_kernel, g, dim3(256), 0, s, C, ldc, n);
    LMX_LAUNCH_CHECK();
}

// det of A (not modified) into d_det.
void determinant(hipStream_t s, int64_t n, const double *A, int64_t lda, double *d_det) {
    const size_t nn = (size_t)std::max<int64_t>(n, 1);
    Scratch lu(nn * nn * sizeof(double), s), piv((2 * nn + 8) * sizeof(int32_t), s);
    int32_t *ipiv = piv.as<int32_t>(), *info = ipiv + nn;
    if (n > 0) copy_d(s, A, lda, lu.as<double>(), n, n, n);
    getrf_impl<double>(s, n, lu.as<double>(), std::max<int64_t>(n, 1), ipiv, nullptr, info, false, true);
    hipLaunchKernelGGL(det_kernel, dim3(1), dim3(64), 0, s, lu.as<double>(), std::max<int64_t>(n, 1), ipiv, n, d_det);
    LMX_LAUNCH_CHECK();
}

}  // namespace lmx

using namespace lmx;

extern "C" {

int lmx_dgemv_dev(void *stream, int64_t m, int64_t n, double alpha, const double *dA, int64_t lda, const double *dx,
                  double beta, double *dy) {
    return guarded([&] {
        check_mat(m, n, lda, "gemv");
        gemv(resolve_stream(stream), m, n, alpha, dA, lda, dx, beta, dy);
        return LMX_OK;
    });
}

int lmx_dgemv(int64_t m, int64_t n, double alpha, const double *A, int64_t lda, const double *x, double beta,
              double *y) {
    return guarded([&] {
        check_mat(m, n, lda, "gemv");
        if (m == 0) return LMX_OK;
        Context &c = ctx();
        std::lock_guard<std::mutex> g(c.host_mu);
        const size_t ab = bytes_of(m, n, lda);
        double *dA = (double *)staging(c, 0, ab), *dx = (double *)staging(c, 1, (size_t)std::max<int64_t>(n, 1) * 8),
               *dy = (double *)staging(c, 2, (size_t)m * 8);
        if (ab) LMX_HIP(hipMemcpyAsync(dA, A, ab, hipMemcpyHostToDevice, c.stream));
        if (n) LMX_HIP(hipMemcpyAsync(dx, x, (size_t)n * 8, hipMemcpyHostToDevice, c.stream));
        if (beta != 0.0) LMX_HIP(hipMemcpyAsync(dy, y, (size_t)m * 8, hipMemcpyHostToDevice, c.stream));
        gemv(c.stream, m, n, alpha, dA, lda, dx, beta, dy);
        LMX_HIP(hipMemcpyAsync(y, dy, (size_t)m * 8, hipMemcpyDeviceToHost, c.stream));
        LMX_HIP(hipStreamSynchronize(c.stream));
        return LMX_OK;
    });
}

int lmx_dnorm_dev(void *stream, char which, int64_t m, int64_t n, const double *dA, int64_t lda, double *d_out) {
    return guarded([&] {
        check_mat(m, n, lda, "norm");
        matrix_norm(resolve_stream(stream), which, m, n, dA, lda, d_out);
        return LMX_OK;
    });
}

int lmx_dnorm(char which, int64_t m, int64_t n, const double *A, int64_t lda, double *out) {
    return guarded([&] {
        check_mat(m, n, lda, "norm");
        Context &c = ctx();
        std::lock_guard<std::mutex> g(c.host_mu);
        const size_t ab = bytes_of(m, n, lda);
        double *dA = (double *)staging(c, 0, std::max<size_t>(ab, 8)), *dout = (double *)staging(c, 1, 8);
        if (ab) LMX_HIP(hipMemcpyAsync(dA, A, ab, hipMemcpyHostToDevice, c.stream));
        matrix_norm(c.stream, which, m, n, dA, lda, dout);
        LMX_HIP(hipMemcpyAsync(out, dout, 8, hipMemcpyDeviceToHost, c.stream));
        LMX_HIP(hipStreamSynchronize(c.stream));
        return LMX_OK;
    });
}

int lmx_dresidual_dev(void *stream, int64_t n, int64_t nrhs, const double *dA, int64_t lda, const double *dX,
                      int64_t ldx, const double *dB, int64_t ldb, double *dR, int64_t ldr) {
    return guarded([&] {
        check_mat(n, n, lda, "residual");
        check_mat(n, nrhs, ldx, "residual");
        check_mat(n, nrhs, ldb, "residual");
        check_mat(n, nrhs, ldr, "residual");
        residual(resolve_stream(stream), n, nrhs, dA, lda, dX, ldx, dB, ldb, dR, ldr);
        return LMX_OK;
    });
}

int lmx_dscaled_residual_dev(void *stream, int64_t n, int64_t nrhs, const double *dA, int64_t lda, const double *dX,
                             int64_t ldx, const double *dB, int64_t ldb, double *d_eta) {
    return guarded([&] {
        check_mat(n, n, lda, "scaled_residual");
        check_mat(n, nrhs, ldx, "scaled_residual");
        check_mat(n, nrhs, ldb, "scaled_residual");
        hipStream_t s = resolve_stream(stream);
        if (nrhs == 0) return LMX_OK;
        Scratch acc((size_t)(1 + 3 * nrhs) * sizeof(double), s);
        Scratch R(nrhs > SCALED_RESIDUAL_FUSED_MAX ? (size_t)(n * nrhs) * sizeof(double) : 0, s);
        scaled_residual(s, n, nrhs, dA, lda, dX, ldx, dB, ldb, d_eta, acc.as<double>(), R.as<double>());
        return LMX_OK;
    });
}

int lmx_dscaled_residual(int64_t n, int64_t nrhs, const double *A, int64_t lda, const double *X, int64_t ldx,
                         const double *B, int64_t ldb, double *eta) {
    return guarded([&] {
        check_mat(n, n, lda, "scaled_residual");
        check_mat(n, nrhs, ldx, "scaled_residual");
        check_mat(n, nrhs, ldb, "scaled_residual");
        if (nrhs == 0) return LMX_OK;
        Context &c = ctx();
        std::lock_guard<std::mutex> g(c.host_mu);
        const size_t ab = bytes_of(n, n, lda), xb = bytes_of(n, nrhs, ldx), bb = bytes_of(n, nrhs, ldb);
        double *dA = (double *)staging(c, 0, std::max<size_t>(ab, 8)), *dX = (double *)staging(c, 1, std::max<size_t>(xb, 8)),
               *dB = (double *)staging(c, 2, std::max<size_t>(bb, 8));
        if (ab) LMX_HIP(hipMemcpyAsync(dA, A, ab, hipMemcpyHostToDevice, c.stream));
        if (xb) LMX_HIP(hipMemcpyAsync(dX, X, xb, hipMemcpyHostToDevice, c.stream));
        if (bb) LMX_HIP(hipMemcpyAsync(dB, B, bb, hipMemcpyHostToDevice, c.stream));
        Scratch acc((size_t)(1 + 3 * nrhs) * sizeof(double), c.stream), deta((size_t)nrhs * sizeof(double), c.stream);
        Scratch R(nrhs > SCALED_RESIDUAL_FUSED_MAX ? (size_t)(n * nrhs) * sizeof(double) : 0, c.stream);
        scaled_residual(c.stream, n, nrhs, dA, lda, dX, ldx, dB, ldb, deta.as<double>(), acc.as<double>(), R.as<double>());
        LMX_HIP(hipMemcpyAsync(eta, deta.p, (size_t)nrhs * sizeof(double), hipMemcpyDeviceToHost, c.stream));
        LMX_HIP(hipStreamSynchronize(c.stream));
        return LMX_OK;
    });
}

int lmx_dsyrk_dev(void *stream, char trans, int64_t n, int64_t k, double alpha, const double *dA, int64_t lda,
                  double beta, double *dC, int64_t ldc) {
    return guarded([&] {
        const bool t = is_trans(trans);
        check_mat(t ? k : n, t ? n : k, lda, "syrk");
        check_mat(n, n, ldc, "syrk");
        syrk_full(resolve_stream(stream), t, n, k, alpha, dA, lda, beta, dC, ldc);
        return LMX_OK;
    });
}

int lmx_dsyrk(char trans, int64_t n, int64_t k, double alpha, const double *A, int64_t lda, double beta, double *C,
              int64_t ldc) {
    return guarded([&] {
        const bool t = is_trans(trans);
        check_mat(t ? k : n, t ? n : k, lda, "syrk");
        check_mat(n, n, ldc, "syrk");
        if (n == 0) return LMX_OK;
        Context &c = ctx();
        std::lock_guard<std::mutex> g(c.host_mu);
        const size_t ab = bytes_of(t ? k : n, t ? n : k, lda), cb = bytes_of(n, n, ldc);
        double *dA = (double *)staging(c, 0, std::max<size_t>(ab, 8)), *dC = (double *)staging(c, 1, cb);
        if (ab) LMX_HIP(hipMemcpyAsync(dA, A, ab, hipMemcpyHostToDevice, c.stream));
        if (beta != 0.0 || ldc > n) LMX_HIP(hipMemcpyAsync(dC, C, cb, hipMemcpyHostToDevice, c.stream));
        syrk_full(c.stream, t, n, k, alpha, dA, lda, beta, dC, ldc);
        LMX_HIP(hipMemcpyAsync(C, dC, cb, hipMemcpyDeviceToHost, c.stream));
        LMX_HIP(hipStreamSynchronize(c.stream));
        return LMX_OK;
    });
}

int lmx_ddet_dev(void *stream, int64_t n, const double *dA, int64_t lda, double *d_det) {
    return guarded([&] {
        check_mat(n, n, lda, "det");
        determinant(resolve_stream(stream), n, dA, lda, d_det);
        return LMX_OK;
    });
}

int lmx_ddet(int64_t n, const double *A, int64_t lda, double *det) {
    return guarded([&] {
        check_mat(n, n, lda, "det");
        Context &c = ctx();
        std::lock_guard<std::mutex> g(c.host_mu);
        const size_t ab = bytes_of(n, n, lda);
        double *dA = (double *)staging(c, 0, std::max<size_t>(ab, 8)), *dd = (double *)staging(c, 1, 8);
        if (ab) LMX_HIP(hipMemcpyAsync(dA, A, ab, hipMemcpyHostToDevice, c.stream));
        determinant(c.stream, n, dA, lda, dd);
        LMX_HIP(hipMemcpyAsync(det, dd, 8, hipMemcpyDeviceToHost, c.stream));
        LMX_HIP(hipStreamSynchronize(c.stream));
        return LMX_OK;
    });
}

int lmx_dinverse_iteration_dev(void *stream, int64_t n, const double *dA, int64_t lda, double mu, const double *dq0,
                               int32_t max_iter, double tol, double *lambda, double *dq, int32_t *iterations) {
    return guarded([&] {
        check_mat(n, n, lda, "inverse_iteration");
        if (max_iter < 0) fail_arg("inverse_iteration: negative max_iter");
        hipStream_t s = resolve_stream(stream);
        const size_t nn = (size_t)std::max<int64_t>(n, 1);
        // (A - mu I), factored once with lu's pivots: luSolve of the same
        // matrix at every step gives the same factors
        Scratch S(nn * nn * sizeof(double), s), piv((3 * nn + 8) * sizeof(int32_t), s),
            vec((3 * nn + 8) * sizeof(double), s);
        int32_t *ipiv = piv.as<int32_t>(), *perm = ipiv + nn, *info = perm + nn;
        double *z = vec.as<double>(), *t = z + nn, *sc = t + nn;  // sc: [0] ||z||, [1] lambda
        if (n > 0) {
            hipLaunchKernelGGL(shift_copy_kernel, dim3((unsigned)cdiv(n * n, 256)), dim3(256), 0, s, dA, lda, mu,
                               S.as<double>(), n);
            LMX_LAUNCH_CHECK();
        }
        getrf_impl<double>(s, n, S.as<double>(), (int64_t)nn, ipiv, perm, info, false, true);
        copy_d(s, dq0, 1, dq, 1, n, 1);
        double prev = 0.0;
        int it = 0;
        for (; it < max_iter; ++it) {
            // z = (A - mu I)^{-1} q ; znorm = nrm2 z ; qNew = scal (1 / znorm) z
            copy_d(s, dq, 1, z, 1, n, 1);
            getrs_perm<double>(s, n, 1, S.as<double>(), (int64_t)nn, perm, z, 1);
            hipLaunchKernelGGL(dot_kernel, dim3(1), dim3(1024), 0, s, z, z, n, sc, 1);
            hipLaunchKernelGGL(scale_by_recip_kernel, dim3((unsigned)cdiv(std::max<int64_t>(n, 1), 256)), dim3(256),
                               0, s, z, sc, z, n);
            // lambda = qNew . (A qNew), the Rayleigh quotient with the original A
            gemv(s, n, n, 1.0, dA, lda, z, 0.0, t);
            hipLaunchKernelGGL(dot_kernel, dim3(1), dim3(1024), 0, s, z, t, n, sc + 1, 0);
            LMX_LAUNCH_CHECK();
            double lam = 0.0;
            LMX_HIP(hipMemcpyAsync(&lam, sc + 1, sizeof(double), hipMemcpyDeviceToHost, s));
            LMX_HIP(hipStreamSynchronize(s));
            copy_d(s, z, 1, dq, 1, n, 1);
            if (std::fabs(lam - prev) < tol) {
                prev = lam;
                ++it;
                break;
            }
            prev = lam;
        }
        LMX_HIP(hipStreamSynchronize(s));
        *lambda = prev;
        if (iterations) *iterations = it;
        return LMX_OK;
    });
}

int lmx_dinverse_iteration(int64_t n, const double *A, int64_t lda, double mu, const double *q0, int32_t max_iter,
                           double tol, double *lambda, double *q, int32_t *iterations) {
    return guarded([&] {
        check_mat(n, n, lda, "inverse_iteration");
        Context &c = ctx();
        double *dA = nullptr, *dq0 = nullptr, *dq = nullptr;
        {
            std::lock_guard<std::mutex> g(c.host_mu);
            const size_t ab = bytes_of(n, n, lda), vb = (size_t)std::max<int64_t>(n, 1) * 8;
            dA = (double *)staging(c, 0, std::max<size_t>(ab, 8));
            dq0 = (double *)staging(c, 1, vb);
            dq = (double *)staging(c, 2, vb);
            if (ab) LMX_HIP(hipMemcpyAsync(dA, A, ab, hipMemcpyHostToDevice, c.stream));
            if (n) LMX_HIP(hipMemcpyAsync(dq0, q0, (size_t)n * 8, hipMemcpyHostToDevice, c.stream));
            const int rc = lmx_dinverse_iteration_dev(c.stream, n, dA, lda, mu, dq0, max_iter, tol, lambda, dq,
                                                      iterations);
            if (rc != LMX_OK) return rc;
            if (n) LMX_HIP(hipMemcpyAsync(q, dq, (size_t)n * 8, hipMemcpyDeviceToHost, c.stream));
            LMX_HIP(hipStreamSynchronize(c.stream));
        }
        return LMX_OK;
    });
}

}  // extern "C"

// ---------------------------------------------------------------- without allocation
//
// lmx_dscaled_residual_noalloc_dev keeps its accumulators (||A||_inf and,
// per column, ||b||, ||x||, ||r||) in a file-scope device array instead of
// memory from the pool, for callers that cannot allocate.  Each GPU holds
// its own copy of the array for the life of the process; its address is
// looked up once per device.  Every call clears what it uses first; calls
// that overlap on one GPU share the array and race.
namespace lmx {
namespace {

constexpr int NOALLOC_MAX_NRHS = 16;  // the fused residual path: no residual block
constexpr int NOALLOC_DOUBLES = 64;   // >= 1 + 3 NOALLOC_MAX_NRHS
static_assert(NOALLOC_DOUBLES >= 1 + 3 * NOALLOC_MAX_NRHS, "scratch too small");
static_assert(NOALLOC_MAX_NRHS <= SCALED_RESIDUAL_FUSED_MAX, "the no-allocation path stores no residual");

__device__ __align__(256) double scratch[NOALLOC_DOUBLES];

std::mutex g_scratch_mu;
double *g_scratch_ptr[64] = {};

// The device address of `scratch` on the current device.
double *scratch_on_device() {
    int dev = 0;
    LMX_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) throw Error(LMX_ERR_DEVICE, "device id out of range");
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        LMX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[dev] = static_cast<double *>(p);
    }
    return g_scratch_ptr[dev];
}

}  // namespace
}  // namespace lmx

extern "C" int lmx_dscaled_residual_noalloc_dev(void *stream, int64_t n, int64_t nrhs, const double *dA,
                                                int64_t lda, const double *dX, int64_t ldx, const double *dB,
                                                int64_t ldb, double *d_eta) {
    return guarded([&] {
        check_mat(n, n, lda, "scaled_residual_noalloc");
        check_mat(n, nrhs, ldx, "scaled_residual_noalloc");
        check_mat(n, nrhs, ldb, "scaled_residual_noalloc");
        if (nrhs > NOALLOC_MAX_NRHS)
            fail_arg("scaled_residual_noalloc: nrhs above 16, the columns its device scratch holds");
        if (nrhs == 0) return LMX_OK;
        double *acc = scratch_on_device();
        scaled_residual(resolve_stream(stream), n, nrhs, dA, lda, dX, ldx, dB, ldb, d_eta, acc, nullptr);
        return LMX_OK;
    });
}
