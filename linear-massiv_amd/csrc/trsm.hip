// Triangular solves with many right-hand sides.
//
// Every solve is written as M x = b for each right-hand-side vector x: a
// column of B for a left-side solve (M = op(T)), a row of B for a
// right-side solve (X op(T) = B, M = op(T)^T).
//   - many right-hand sides ("wide"): M is processed in 64 x 64 blocks; for
//     each diagonal block in dependency order, trsm_leaf_kernel solves the
//     block (triangle in LDS, one vector per thread) and the not-yet-solved
//     part is updated with one MFMA GEMM;
//   - at most 64 right-hand sides ("block inverse"): the diagonal blocks of
//     M, up to 1024 x 1024, are inverted explicitly (64 x 64 inverses in
//     LDS, then doubling with batched MFMA GEMMs), and the solve is a
//     sequence of n / 1024 block steps, each a product with the block's
//     inverse and one update of all unsolved rows -- every launch spans
//     the chip, and the work is one read of the triangle.  A factor that is
//     solved with many times keeps its inverses (tri_inverse_build once,
//     tri_inverse_solve per solve).
#include "common.hpp"
#include "device.hpp"
#include "kernels.hpp"

namespace lmx {

namespace {

constexpr int TB = 64;

// A diagonal block of M = op(T) in LDS as a LOWER triangle L of order S,
// stored by columns (Ms[c][r] = L[r][c]): M restricted to its triangle and
// padded with the identity beyond nb (nb <= 0: the identity); an upper M
// is staged reversed, L[r][c] = M[S-1-r][S-1-c] (J M J with J the reversal
// is lower), so one forward solve serves both.  Thread t reads column t of
// each stored row i -- coalesced -- which is M[i][t], or M[t][i] for a
// transposed operand.  Rows of Ms are 16-byte aligned so the solve reads
// four coefficients per access.  The loads are unconditional, at addresses
// clamped into the nb x nb block, and issued 16 at a time before the
// triangle is selected with no branch: a load left under a per-lane
// condition waits for memory before the next one issues.
constexpr int ldm_of(int bytes) { return TB + 16 / bytes; }
template <class T>
__device__ inline void stage_block(T (*Ms)[ldm_of(sizeof(T))], const T *A, int64_t lda, int nb, int t, bool mt,
                                   bool mlower, bool unit, int S) {
    if (nb <= 0) {
#pragma unroll 16
        for (int i = 0; i < TB; ++i) {
            if (i >= S || t >= S) continue;
            Ms[i][t] = i == t ? T(1) : T(0);
        }
        return;
    }
    const T *At = A + min(t, nb - 1);
    for (int i0 = 0; i0 < TB; i0 += 16) {
        T a[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) a[u] = At[(int64_t)min(i0 + u, nb - 1) * lda];
        // keeps the compiler from sinking each load into the branch that
        // uses it: all 16 are in flight before the first wait
#pragma unroll
        for (int u = 0; u < 16; ++u) asm volatile("" : "+v"(a[u]));
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int i = i0 + u;
            const int r = mt ? t : i, c = mt ? i : t;
            const bool inblk = r < nb && c < nb, in = mlower ? c <= r : c >= r;
            T v = inblk && in ? a[u] : T(0);
            v = r == c && (unit || !inblk) ? T(1) : v;
            T *dst = mlower ? &Ms[c][r] : &Ms[S - 1 - c][S - 1 - r];
            if (r < S && c < S) *dst = v;
        }
    }
}

// L x = b in place on the registers x[0..S), L lower in LDS by columns:
// x_r = (b_r - sum_{k<r} l_rk x_k) / l_rr with the terms subtracted in
// ascending k (for a reversed upper block: the reference back substitution
// with its sum taken from the far end) -- evaluated column by column, so
// that the updates of one step are independent and the dependent path is S
// steps, not S^2/2.  A unit diagonal is staged as 1; dividing by 1 is exact.
template <class T, int S>
__device__ __forceinline__ void solve_regs(const T (*Ms)[ldm_of(sizeof(T))], T (&x)[TB]) {
#pragma unroll
    for (int k = 0; k < S; ++k) {
        x[k] = x[k] / Ms[k][k];
#pragma unroll
        for (int r = k + 1; r < S; ++r) x[r] = fma(-Ms[k][r], x[k], x[r]);
    }
}

// M x = b for one diagonal block (order nb <= S <= 64) and the lane's
// right-hand-side vector, element r at p[r rs]: the vector is held in
// registers (reversed for an upper M), the block in LDS.  A lane that is
// not live points p at a valid vector and stores nothing; `ident` takes
// e_t, the lane's unit vector, in place of the loads.  The results go back
// through the lane's own LDS column: a per-row guard on stores from
// registers costs the compiler its register allocation.
template <class T, int S>
__device__ __forceinline__ void leaf_body(int nb, bool live, const T *__restrict__ A, int64_t lda, T *p, int64_t rs,
                                          bool mt, bool mlower, bool unit, bool ident) {
    __shared__ __attribute__((aligned(16))) T Ms[TB][ldm_of(sizeof(T))];
    __shared__ T Xs[TB][TB + 1];
    const int t = threadIdx.x;
    T x[TB];
    if (ident) {
#pragma unroll
        for (int i = 0; i < TB; ++i) {
            const int r = mlower ? i : S - 1 - i;  // register i holds element r
            x[i] = (i < S && r < nb && r == t) ? T(1) : T(0);
        }
    } else {
        const int nbc = max(nb, 1);
#pragma unroll
        for (int i = 0; i < TB; ++i) {
            const int r = mlower ? i : S - 1 - i;
            x[i] = T(0);
            if (i < S) {
                const T a = p[(int64_t)min(r, nbc - 1) * rs];
                x[i] = r < nb ? a : T(0);
            }
        }
    }
    stage_block<T>(Ms, A, lda, nb, t, mt, mlower, unit, S);
    __syncthreads();
    solve_regs<T, S>(Ms, x);
#pragma unroll
    for (int i = 0; i < S; ++i) Xs[i][t] = x[i];
    if (!live) return;
    for (int r = 0; r < nb; ++r) p[r * rs] = Xs[mlower ? r : S - 1 - r][t];
}

// M x = b for the diagonal block (order nb <= S) with one right-hand-side
// vector per thread.  right = 0: vectors are columns of B; right = 1: rows.
template <class T, int S>
__global__ __launch_bounds__(64) void trsm_leaf_kernel(int nb, int nvec, const T *__restrict__ A, int64_t lda,
                                                       T *__restrict__ B, int64_t ldb, int mt, int mlower, int unit,
                                                       int right) {
    const int v = blockIdx.x * 64 + threadIdx.x, vc = min(v, nvec - 1);
    T *p = right ? B + (int64_t)vc * ldb : B + vc;
    leaf_body<T, S>(nb, v < nvec, A, lda, p, right ? 1 : ldb, mt != 0, mlower != 0, unit != 0, false);
}

// W_b = inverse of diagonal block b of M (order min(64, n - 64 b), padded
// to a 64 x 64 identity; the identity when the block starts at or beyond
// n), solved in place over the identity the diagonal copy left there: lane
// t solves M x = e_t, giving column t.  W holds dense wld x wld blocks (row
// stride wld), and the 64-blocks lie along their diagonals: block b sits
// in wld-block b / (wld / 64) at diagonal offset 64 (b % (wld / 64)).
template <class T>
__global__ __launch_bounds__(64) void trtri_blocks_kernel(int n, const T *__restrict__ A, int64_t lda,
                                                          T *__restrict__ W, int wld, int mt, int mlower, int unit) {
    const int b = blockIdx.x, r0 = b * TB;
    const int per = wld / TB, zk = b / per, i0 = b % per;
    T *Wb = W + (int64_t)zk * wld * wld + (int64_t)i0 * TB * wld + i0 * TB;
    leaf_body<T, TB>(min(TB, n - r0), true, A + (int64_t)r0 * lda + r0, lda, Wb + threadIdx.x, wld, mt != 0,
                     mlower != 0, unit != 0, true);
}

// ---------------------------------------------------------------- block inverses

// Largest diagonal block inverted explicitly, and the longest row the GEMV
// kernel takes.
constexpr int IBW = 1024;

// W (nbw dense blocks of bw x bw, row stride bw) = the diagonal bw-blocks of
// M = op(T) restricted to M's strict triangle, zero elsewhere and beyond n,
// with the identity in the diagonal 64 x 64 tiles (the 64-block inverses
// are solved in place over it).
// One workgroup per 64 x 64 tile (blockIdx.y = diagonal block); a
// transposed operand goes through LDS so both sides stay coalesced.
template <class T, bool MT, bool MLOWER>
__global__ __launch_bounds__(256) void tri_diag_copy_kernel(int n, int bw, const T *__restrict__ A, int64_t lda,
                                                            T *__restrict__ W) {
    __shared__ T tile[TB][TB + 1];
    const int tpb = bw / TB, t = threadIdx.x;
    const int ti = (int)blockIdx.x / tpb, tj = (int)blockIdx.x % tpb;
    const int64_t r0 = (int64_t)blockIdx.y * bw + ti * TB, c0 = (int64_t)blockIdx.y * bw + tj * TB;
    T *Wt = W + (int64_t)blockIdx.y * bw * bw + (int64_t)ti * TB * bw + tj * TB;
    const bool live = MLOWER ? ti >= tj : ti <= tj;  // uniform per workgroup
    if (live) {
#pragma unroll 4
        for (int e = t; e < TB * TB; e += 256) {
            const int a = e >> 6, b = e & 63;
            // (a, b) in the stored operand's memory order
            const int64_t gr = MT ? c0 + a : r0 + a, gc = MT ? r0 + b : c0 + b;
            const T v = (gr < n && gc < n) ? A[gr * lda + gc] : T(0);
            if (MT)
                tile[b][a] = v;
            else
                tile[a][b] = v;
        }
        __syncthreads();
    }
#pragma unroll 4
    for (int e = t; e < TB * TB; e += 256) {
        const int r = e >> 6, c = e & 63;
        const int64_t gr = r0 + r, gc = c0 + c;
        const bool in = live && (MLOWER ? gc < gr : gc > gr) && ti != tj;
        Wt[(int64_t)r * bw + c] = in ? tile[r][c] : (ti == tj && r == c ? T(1) : T(0));
    }
}

// out[r][q] = alpha sum_{c<k} M[r][c] x[c][q] (added to out[r][q] when
// `accumulate`), r < rows, q < nrhs <= NR, k <= IBW.  One wave per row, rpw
// rows per wave (rpw * NR <= 64); x is staged in LDS column by column
// (consecutive lanes read consecutive banks).  The loads are pipelined one
// row ahead in two register buffers used in turn (no copy between them): the
// first row is in flight while x is staged, and a row's products wait only
// for that row while the next one loads.  Lane rr * NR + q keeps row rr's
// result for column q and the wave stores its rows once at the end, so no
// store inside the loop orders the loads behind it.  The prefetch row index
// is clamped to rows - 1, so every load is in bounds.  FULL (k == IBW) drops
// the column mask, so the loads carry no branches.
template <class T, bool FULL>
__device__ __forceinline__ void gemv_load_row(const T *__restrict__ M, int64_t ldm, int64_t r, int k, int lane,
                                              T (&mv)[IBW / 64]) {
    const T *mr = M + r * ldm;
#pragma unroll
    for (int i = 0; i < IBW / 64; ++i) {
        const int c = lane + 64 * i;
        mv[i] = FULL || c < k ? mr[c] : T(0);
    }
}

template <class T, int NR>
__device__ __forceinline__ void gemv_finish_row(const T (&mv)[IBW / 64], const T (&xs)[NR][IBW], int lane, int rr,
                                                T &res) {
    T acc[NR];
#pragma unroll
    for (int q = 0; q < NR; ++q) acc[q] = T(0);
#pragma unroll
    for (int i = 0; i < IBW / 64; ++i)
#pragma unroll
        for (int q = 0; q < NR; ++q) acc[q] = fma(mv[i], xs[q][lane + 64 * i], acc[q]);
#pragma unroll
    for (int q = 0; q < NR; ++q) {
        const T v = wave_sum(acc[q]);
        res = lane == rr * NR + q ? v : res;
    }
}

template <class T, int NR, bool FULL>
__global__ __launch_bounds__(256) void gemv_rows_kernel(int rows, int k, int nrhs, int rpw, T alpha,
                                                        const T *__restrict__ M, int64_t ldm,
                                                        const T *__restrict__ x, int64_t ldx, int accumulate,
                                                        T *__restrict__ out, int64_t ldo) {
    __shared__ T xs[NR][IBW];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int64_t rbeg = ((int64_t)blockIdx.x * 4 + wave) * rpw, last = rows - 1;
    // this lane's output element: row rbeg + lane / NR, column lane % NR
    const int64_t orow = rbeg + lane / NR;
    const int oq = lane % NR;
    const bool mine = lane / NR < rpw && orow < rows && oq < nrhs;
    T old = T(0), res = T(0);
    if (accumulate && mine) old = out[orow * ldo + oq];
    T b0[IBW / 64], b1[IBW / 64];
    gemv_load_row<T, FULL>(M, ldm, rbeg < last ? rbeg : last, k, lane, b0);
    for (int e = t; e < NR * IBW; e += 256) {
        const int c = e / NR, q = e % NR;
        xs[q][c] = (c < k && q < nrhs) ? x[(int64_t)c * ldx + q] : T(0);
    }
    __syncthreads();
    // pairs of rows with no exit between their halves (an exit there lets
    // the compiler move the prefetch down to its use), then an odd last row
    int rr = 0;
    for (; rr + 1 < rpw && rbeg + rr + 1 < rows; rr += 2) {
        const int64_t r = rbeg + rr;
        gemv_load_row<T, FULL>(M, ldm, r + 1, k, lane, b1);
        __builtin_amdgcn_sched_barrier(0);  // keep the prefetch ahead of the products
        gemv_finish_row<T, NR>(b0, xs, lane, rr, res);
        gemv_load_row<T, FULL>(M, ldm, r + 2 < last ? r + 2 : last, k, lane, b0);
        __builtin_amdgcn_sched_barrier(0);
        gemv_finish_row<T, NR>(b1, xs, lane, rr + 1, res);
    }
    if (rr < rpw && rbeg + rr < rows) gemv_finish_row<T, NR>(b0, xs, lane, rr, res);
    if (mine) out[orow * ldo + oq] = accumulate ? fma(alpha, res, old) : alpha * res;
}

template <class T>
__global__ void copy_rows_kernel(const T *__restrict__ in, int64_t ldi, T *__restrict__ out, int64_t ldo,
                                 int64_t rows, int cols) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= rows * cols) return;
    const int64_t i = e / cols;
    const int c = (int)(e - i * cols);
    out[i * ldo + c] = in[i * ldi + c];
}

// ---------------------------------------------------------------- dispatch

template <class T>
void trtri_blocks(hipStream_t s, bool mt, bool mlower, bool unit, int n, int nblocks, const T *A, int64_t lda, T *W,
                  int wld) {
    hipLaunchKernelGGL(trtri_blocks_kernel<T>, dim3((unsigned)nblocks), dim3(64), 0, s, n, A, lda, W, wld, mt ? 1 : 0,
                       mlower ? 1 : 0, unit ? 1 : 0);
    LMX_LAUNCH_CHECK();
}

template <class T>
void diag_copy(hipStream_t s, bool mt, bool mlower, int n, int bw, int nbw, const T *A, int64_t lda, T *W) {
    const int tpb = bw / TB;
    dim3 g((unsigned)(tpb * tpb), (unsigned)nbw), blk(256);
#define LMX_DCOPY(MT_, ML_)                                                                             \
    if (mt == MT_ && mlower == ML_) {                                                                  \
        hipLaunchKernelGGL((tri_diag_copy_kernel<T, MT_, ML_>), g, blk, 0, s, n, bw, A, lda, W);        \
        LMX_LAUNCH_CHECK();                                                                             \
        return;                                                                                         \
    }
    LMX_DCOPY(false, false)
    LMX_DCOPY(false, true)
    LMX_DCOPY(true, false)
    LMX_DCOPY(true, true)
#undef LMX_DCOPY
}

// out = alpha op(M) x (+ out when `accumulate`), op(M) rows x k, x k x nrhs:
// the GEMV kernel for at most 8 right-hand sides of a row-major M, the MFMA
// GEMM otherwise.
template <class T>
void apply_block(hipStream_t s, int64_t rows, int64_t k, int64_t nrhs, T alpha, const T *M, int64_t ldm, bool mtrans,
                 const T *x, int64_t ldx, bool accumulate, T *out, int64_t ldo) {
    if (rows <= 0) return;
    if (mtrans || nrhs > 8) {
        gemm(s, mtrans, false, rows, nrhs, k, alpha, M, ldm, x, ldx, accumulate ? T(1) : T(0), out, ldo);
        return;
    }
    // rows per wave: at most 8, so rpw * NR <= 64 result lanes for NR <= 8
    const int rpw = (int)std::min<int64_t>(8, std::max<int64_t>(1, rows / 4096));
    dim3 g((unsigned)cdiv(rows, 4 * rpw)), blk(256);
#define LMX_GEMV(NR_)                                                                                          \
    hipLaunchKernelGGL((k == IBW ? gemv_rows_kernel<T, NR_, true> : gemv_rows_kernel<T, NR_, false>), g, blk, 0, \
                       s, (int)rows, (int)k, (int)nrhs, rpw, alpha, M, ldm, x, ldx, accumulate ? 1 : 0, out, ldo)
    if (nrhs == 1)
        LMX_GEMV(1);
    else if (nrhs == 2)
        LMX_GEMV(2);
    else if (nrhs <= 4)
        LMX_GEMV(4);
    else
        LMX_GEMV(8);
#undef LMX_GEMV
    LMX_LAUNCH_CHECK();
}

// Leaf solve of one diagonal block (order bs <= 64) for nvec vectors.
template <class T, bool RIGHT>
void leaf(hipStream_t s, bool mt, bool mlower, bool unit, int bs, int64_t nvec, const T *A, int64_t lda, T *B,
          int64_t ldb) {
    dim3 g((unsigned)cdiv(nvec, 64)), blk(64);
    const int imt = mt ? 1 : 0, iml = mlower ? 1 : 0, iu = unit ? 1 : 0, ir = RIGHT ? 1 : 0;
    if (bs <= 32)
        hipLaunchKernelGGL((trsm_leaf_kernel<T, 32>), g, blk, 0, s, bs, (int)nvec, A, lda, B, ldb, imt, iml, iu, ir);
    else
        hipLaunchKernelGGL((trsm_leaf_kernel<T, 64>), g, blk, 0, s, bs, (int)nvec, A, lda, B, ldb, imt, iml, iu, ir);
    LMX_LAUNCH_CHECK();
}

}  // namespace

int tri_block(int64_t n) {
    int b = TB;
    while (b < IBW && b < n) b *= 2;
    return b;
}

int64_t tri_inverse_elems(int64_t n) {
    const int bw = tri_block(n);
    return cdiv(std::max<int64_t>(n, 1), bw) * (int64_t)bw * bw;
}

// Inverses of M's diagonal bw-blocks: masked copy, the 64 x 64 inverses in
// place on its diagonal, then doubling: for each pair of inverted h-blocks
// of a lower M = [[A, 0], [C, B]] the off-diagonal block of the inverse is
// -B^-1 (C A^-1); of an upper M = [[A, C], [0, B]] it is -A^-1 (C B^-1).
// Each level is two batched GEMMs over every pair of every diagonal block.
template <class T>
void tri_inverse_build(hipStream_t s, bool lower, bool trans, bool unit, int64_t n, const T *A, int64_t lda, T *W) {
    if (n <= 0) return;
    if (n > INT32_MAX / 2) fail_arg("trsm: dimension too large");
    TimedRegion tr("trsm", s);
    const bool mt = trans, mlower = lower != trans;
    const int bw = tri_block(n), nbw = (int)cdiv(n, bw), tpb = bw / TB;
    diag_copy<T>(s, mt, mlower, (int)n, bw, nbw, A, lda, W);
    trtri_blocks<T>(s, mt, mlower, unit, (int)n, nbw * tpb, A, lda, W, bw);
    if (bw == TB) return;
    Scratch sT((size_t)nbw * bw * (bw / 4) * sizeof(T), s);
    T *Tb = sT.as<T>();
    const int64_t bw2 = (int64_t)bw * bw;
    for (int h = TB; h < bw; h *= 2) {
        const int np = bw / (2 * h);
        const int64_t dstep = 2 * (int64_t)h * bw + 2 * h, hh = (int64_t)h * h;
        T *Cp = mlower ? W + (int64_t)h * bw : W + h;        // off-diagonal block of each pair
        const T *Rinv = mlower ? W : W + (int64_t)h * bw + h;  // inverse applied on C's right
        const T *Linv = mlower ? W + (int64_t)h * bw + h : W;  // inverse applied on the left
        gemm_batched<T>(s, h, h, h, T(1), Cp, bw, Rinv, bw, T(0), Tb, h, np, nbw, dstep, dstep, hh, bw2, bw2,
                        np * hh);
        gemm_batched<T>(s, h, h, h, T(-1), Linv, bw, Tb, h, T(0), Cp, bw, np, nbw, dstep, hh, dstep, bw2, np * hh,
                        bw2);
    }
}

// op(T) X = B with the inverses of tri_inverse_build: for each diagonal
// block in dependency order, X_b = W_b B_b, then the rows still unsolved
// take B_i -= M_ib X_b -- two launches per block, each a full-chip pass.
template <class T>
void tri_inverse_solve(hipStream_t s, bool lower, bool trans, int64_t n, int64_t nrhs, const T *A, int64_t lda,
                       const T *W, T *B, int64_t ldb) {
    if (n <= 0 || nrhs <= 0) return;
    if (n > INT32_MAX / 2 || nrhs > INT32_MAX / 2) fail_arg("trsm: dimension too large");
    TimedRegion tr("trsm", s, (double)n * (double)n * (double)nrhs);
    const bool mt = trans, mlower = lower != trans;
    const int bw = tri_block(n), nbw = (int)cdiv(n, bw);
    Scratch sX((size_t)n * nrhs * sizeof(T), s);
    T *X = sX.as<T>();
    const int64_t ldx = nrhs;
    for (int bi = 0; bi < nbw; ++bi) {
        const int b = mlower ? bi : nbw - 1 - bi;
        const int64_t j0 = (int64_t)b * bw, nj = std::min<int64_t>(bw, n - j0);
        apply_block<T>(s, nj, nj, nrhs, T(1), W + (int64_t)b * bw * bw, bw, false, B + j0 * ldb, ldb, false,
                       X + j0 * ldx, ldx);
        const int64_t i0 = mlower ? j0 + nj : 0, ni = mlower ? n - (j0 + nj) : j0;
        if (ni > 0) {
            const T *Mb = mt ? A + j0 * lda + i0 : A + i0 * lda + j0;
            apply_block<T>(s, ni, nj, nrhs, T(-1), Mb, lda, mt, X + j0 * ldx, ldx, true, B + i0 * ldb, ldb);
        }
    }
    hipLaunchKernelGGL(copy_rows_kernel<T>, dim3((unsigned)cdiv(n * nrhs, 256)), dim3(256), 0, s, X, ldx, B, ldb, n,
                       (int)nrhs);
    LMX_LAUNCH_CHECK();
}

// op(T) X = B, B n x nrhs: one leaf launch for a single 64-block, the
// block-inverse path otherwise.
template <class T>
void trsm_left(hipStream_t s, bool lower, bool trans, bool unit, int64_t n, int64_t nrhs, const T *A, int64_t lda,
               T *B, int64_t ldb) {
    if (n <= 0 || nrhs <= 0) return;
    if (n > INT32_MAX / 2 || nrhs > INT32_MAX / 2) fail_arg("trsm: dimension too large");
    TimedRegion tr("trsm", s, (double)n * (double)n * (double)nrhs);
    if (n <= TB) {
        leaf<T, false>(s, trans, lower != trans, unit, (int)n, nrhs, A, lda, B, ldb);
        return;
    }
    Scratch sw((size_t)tri_inverse_elems(n) * sizeof(T), s);
    tri_inverse_build<T>(s, lower, trans, unit, n, A, lda, sw.as<T>());
    tri_inverse_solve<T>(s, lower, trans, n, nrhs, A, lda, sw.as<T>(), B, ldb);
}

// X op(T) = B with B nrows x n.  A single 64-block is one leaf launch
// (row-vector equations M x = b, M = op(T)^T).  Otherwise, with R = op(T)
// and W_j the inverses of R's diagonal blocks, in dependency order (last
// block first for a lower R): X_j = B_j W_j, then the unsolved columns take
// B_i -= X_j R_ji -- MFMA GEMMs over all rows at once.
template <class T>
void trsm_right(hipStream_t s, bool lower, bool trans, bool unit, int64_t nrows, int64_t n, const T *A, int64_t lda,
                T *B, int64_t ldb) {
    if (n <= 0 || nrows <= 0) return;
    if (n > INT32_MAX / 2 || nrows > INT32_MAX / 2) fail_arg("trsm: dimension too large");
    TimedRegion tr("trsm", s, (double)n * (double)n * (double)nrows);
    if (n <= TB) {
        const bool mt = !trans;
        leaf<T, true>(s, mt, lower != mt, unit, (int)n, nrows, A, lda, B, ldb);
        return;
    }
    const bool rlower = lower != trans;
    const int bw = tri_block(n), nbw = (int)cdiv(n, bw);
    Scratch sw((size_t)tri_inverse_elems(n) * sizeof(T), s), sX((size_t)nrows * n * sizeof(T), s);
    T *W = sw.as<T>(), *X = sX.as<T>();
    tri_inverse_build<T>(s, lower, trans, unit, n, A, lda, W);
    // R[r0.., c0..] as the right GEMM operand (transposed when trans)
    auto rblock = [&](int64_t r0, int64_t c0) -> const T * { return trans ? A + c0 * lda + r0 : A + r0 * lda + c0; };
    for (int bi = 0; bi < nbw; ++bi) {
        const int b = rlower ? nbw - 1 - bi : bi;
        const int64_t j0 = (int64_t)b * bw, nj = std::min<int64_t>(bw, n - j0);
        gemm(s, false, false, nrows, nj, nj, T(1), B + j0, ldb, W + (int64_t)b * bw * bw, bw, T(0), X + j0, n);
        const int64_t i0 = rlower ? 0 : j0 + nj, ni = rlower ? j0 : n - (j0 + nj);
        if (ni > 0)
            gemm(s, false, trans, nrows, ni, nj, T(-1), X + j0, n, rblock(j0, i0), lda, T(1), B + i0, ldb);
    }
    hipLaunchKernelGGL(copy_rows_kernel<T>, dim3((unsigned)cdiv(nrows * n, 256)), dim3(256), 0, s, X, n, B, ldb,
                       nrows, (int)n);
    LMX_LAUNCH_CHECK();
}

template void trsm_left<double>(hipStream_t, bool, bool, bool, int64_t, int64_t, const double *, int64_t, double *,
                                int64_t);
template void trsm_left<float>(hipStream_t, bool, bool, bool, int64_t, int64_t, const float *, int64_t, float *,
                               int64_t);
template void tri_inverse_build<double>(hipStream_t, bool, bool, bool, int64_t, const double *, int64_t, double *);
template void tri_inverse_build<float>(hipStream_t, bool, bool, bool, int64_t, const float *, int64_t, float *);
template void tri_inverse_solve<double>(hipStream_t, bool, bool, int64_t, int64_t, const double *, int64_t,
                                        const double *, double *, int64_t);
template void tri_inverse_solve<float>(hipStream_t, bool, bool, int64_t, int64_t, const float *, int64_t,
                                       const float *, float *, int64_t);
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
