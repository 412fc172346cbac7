// Lower Cholesky factorisation (right-looking, blocked, with lookahead) and
// solve, for row-major matrices.  Only the lower triangle is written, and
// the results depend on it alone (the GEMM tiles on the diagonal load the
// upper elements they cover and never use them).
//
// For each block column of width 256:
//   - the diagonal block and its inverse (diag_block), by halves down to
//     64 x 64, which potrf_diag_kernel factors and inverts in one pass;
//     between two halves, L10 = A10 L00^{-T} and A11 -= L10 L10^T are MFMA
//     products, and the inverse's off-diagonal block -L11^{-1} L10 L00^{-1}
//     two more;
//   - L21 = A21 L11^{-T}: one GEMM with the block's inverse;
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

// Rows [0, rows) of a 64-column block at G (row stride ldg) into S[r][c]
// (an LDS array of row stride LD), zeros below them, in two parts: load_rows issues the loads (wave w rows
// 16w .. 16w+15, lane c column c, unconditional at clamped addresses) and
// put_rows stores them, so that other loads can be in flight in between.
template <class T>
__device__ inline void load_rows(T (&v)[16], const T *G, int64_t ldg, int rows, int cols, int w, int c) {
    const T *Gc = G + min(c, cols - 1);
#pragma unroll
    for (int u = 0; u < 16; ++u) v[u] = Gc[(int64_t)min(16 * w + u, rows - 1) * ldg];
}

template <class T, int LD>
__device__ inline void put_rows(T (*S)[LD], const T (&v)[16], int rows, int cols, int w, int c) {
#pragma unroll
    for (int u = 0; u < 16; ++u) S[16 * w + u][c] = (16 * w + u < rows && c < cols) ? v[u] : T(0);
}

// D += A B over a 64 x 64 x 64 product from LDS, a(r, k) = Ap[r ar + k ak],
// b(k, c) = Bp[k bk + c bc]: wave wv takes rows 16wv .. 16wv+15 and the four
// 16-column tiles, acc[ct] holding tile ct (MFMA 16 x 16 x 4; lane l gives
// a(16wv + l % 16, k + l / 16) and b(k + l / 16, 16ct + l % 16)).
template <class T>
__device__ inline void tile_product(const T *Ap, int ar, int ak, const T *Bp, int bk, int bc, int wv, int lane,
                                    typename V4<T>::type (&acc)[4]) {
    const int lr = lane & 15, lk = lane >> 4;
#pragma unroll 4
    for (int k0 = 0; k0 < DIAG; k0 += 4) {
        const T a = Ap[(16 * wv + lr) * ar + (k0 + lk) * ak];
        T b[4];
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) b[ct] = Bp[(k0 + lk) * bk + (16 * ct + lr) * bc];
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) acc[ct] = mfma(a, b[ct], acc[ct]);
    }
}

// Rows 16q .. 16q+15 of X = L^{-1} in wave q (lane i holds column i, x[u]
// row 16q + u), once rows 0 .. 16q-1 of X are in R and columns 0 .. 16q+15
// of L in C (C[j][r] = l_rj): x_r = (e_r - sum_{k<r} l_rk x_k) / l_rr, the
// terms in ascending k, as the 64 x 64 inverse kernel's forward
// substitution.  inverse_terms applies the rows before the group,
// inverse_finish the group's own and writes each finished row to R.
template <class T>
__device__ inline void inverse_terms(const T (*C)[DIAG + 2], const T (*R)[DIAG + 1], T (&x)[16], int q, int i) {
    for (int k = 0; k < 16 * q; ++k) {
        const T xk = R[k][i];
        T ck[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) ck[u] = C[k][16 * q + u];
#pragma unroll
        for (int u = 0; u < 16; ++u) x[u] = fma(-ck[u], xk, x[u]);
    }
}

template <class T>
__device__ inline void inverse_finish(const T (*C)[DIAG + 2], T (*R)[DIAG + 1], T (&x)[16], int q, int i) {
#pragma unroll
    for (int kk = 0; kk < 16; ++kk) {
        const int k = 16 * q + kk;
        x[kk] = x[kk] / C[k][k];
        R[k][i] = x[kk];
        T ck[16];
#pragma unroll
        for (int u = kk + 1; u < 16; ++u) ck[u] = C[k][16 * q + u];
#pragma unroll
        for (int u = kk + 1; u < 16; ++u) x[u] = fma(-ck[u], x[kk], x[u]);
    }
}

// In-place lower Cholesky of an nb x nb block (nb <= 64) by four waves:
// lane i holds row i, wave w its columns 16w .. 16w+15.  Column j: the wave
// that holds it takes l_jj = sqrt(a_jj) from lane j (v_readlane), divides
// its column by it (l_ij = a_ij / l_jj, a division, as rawCholColumnSIMD),
// publishes the column in LDS and updates its own columns after j, all 16
// of its columns without a barrier; after one barrier the waves holding
// later columns apply the 16 columns to all of theirs (a_ic -= l_ij l_cj).
// Every element takes its terms in ascending j.  Lanes also update their
// row's columns above the diagonal; those values are never stored.
//
// The inverse X = L^{-1} (the identity beyond nb) is built beside it: lane
// i holds column i of X, wave q rows 16q .. 16q+15 (inverse_rows), which it
// finishes while the next wave factors its columns.  X^T, upper, goes to WT
// (row stride ldw), so that B L^{-T} is a plain product B WT.
//
// The first half of a 128-column diagonal block (MODE 1) goes on with the
// rows below it, k1 of them: L10 = A10 L^{-T} and A11 -= L10 L10^T (lower),
// both in place.  The second half (MODE 2) goes on with the off-diagonal
// block of the whole block's inverse: WT01 = -WT00 (L10^T WT11), with L10
// the DIAG columns to the left of A and WT00 / WT01 the blocks DIAG rows up
// in WT.  The products run on MFMA from LDS (tile_product).
template <class T, int MODE>
__global__ __launch_bounds__(256, 2) void potrf_diag_kernel(int nb, T *__restrict__ A, int64_t lda, int col0,
                                                         int32_t *__restrict__ info, T *__restrict__ WT,
                                                         int64_t ldw, int k1) {
    // C[j][i] = l_ij; afterwards the products' operand / result block
    __shared__ __attribute__((aligned(16))) T C[DIAG][DIAG + 2];
    // R[k][t] = x_kt; before that the block as loaded, afterwards a
    // product's operand (no third array: with 67 KB of LDS the kernel shares
    // a CU with two workgroups of the lookahead SYRK)
    __shared__ __attribute__((aligned(16))) T R[DIAG][DIAG + 1];
    __shared__ int sbad[4];
    // the wave index, uniform in a scalar register
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), i = threadIdx.x & 63;
    const int lr = i & 15, lk = i >> 4;
    // coalesced loads through LDS (lane i reads column i of the wave's 16
    // rows), unconditional at addresses clamped into the block: a load under
    // a per-lane condition waits for memory before the next one issues.  The
    // products' operands after the factorisation are loaded here too (they
    // do not depend on it), so all of the kernel's loads wait once:
    //   MODE 1: p0 = A10's rows, p1 = A11's old values in the product's
    //           accumulator layout;
    //   MODE 2: p0 = L10's rows, p1 = WT00's rows.
    T p0[16], p1[16];
    {
        const T *Ai = A + min(i, nb - 1);
        T v[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) v[u] = Ai[(int64_t)min(16 * w + u, nb - 1) * lda];
        if constexpr (MODE == 1) {
            const T *A10 = A + (int64_t)DIAG * lda, *A11 = A10 + DIAG;
            load_rows<T>(p0, A10, lda, k1, DIAG, w, i);
#pragma unroll
            for (int ct = 0; ct < 4; ++ct)
#pragma unroll
                for (int r4 = 0; r4 < 4; ++r4) {
                    const int row = 16 * w + acc_row<T>(lk, r4), col = 16 * ct + lr;
                    p1[4 * ct + r4] = A11[(int64_t)min(row, k1 - 1) * lda + min(col, k1 - 1)];
                }
        } else if constexpr (MODE == 2) {
            load_rows<T>(p0, A - DIAG, lda, nb, DIAG, w, i);
            load_rows<T>(p1, WT - DIAG * ldw - DIAG, ldw, DIAG, DIAG, w, i);
        }
#pragma unroll
        for (int u = 0; u < 16; ++u) asm volatile("" : "+v"(v[u]));
        if constexpr (MODE != 0) {
#pragma unroll
            for (int u = 0; u < 16; ++u) asm volatile("" : "+v"(p0[u]), "+v"(p1[u]));
        }
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int r = 16 * w + u;
            R[r][i] = (r < nb && i <= r) ? v[u] : T(0);  // R stages the block
        }
    }
    __syncthreads();
    T a[16], x[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
        a[u] = R[i][16 * w + u];
        x[u] = 16 * w + u == i ? T(1) : T(0);
    }
    // Four phases and four barriers.  Phase p: the waves from p on apply
    // block p - 1's 16 columns to theirs (a rank-16 update, the terms in
    // ascending j); wave p then factors its 16 columns with no barrier (it
    // publishes each column in LDS and updates its own later columns);
    // wave p - 1, idle from then on, finishes rows 16(p-1) .. 16p-1 of X.
    // All 64 steps run; rows and columns from nb on start at zero, may turn
    // to NaN, and are never stored.  A published column is the identity's
    // beyond nb (no real entry depends on those), which gives the inverse
    // its identity padding.
    int bad = 0;
    for (int p = 0; p < 4; ++p) {
        if (p > 0 && w >= p) {
            for (int jj = 0; jj < 16; ++jj) {
                const int j = 16 * (p - 1) + jj;
                const T li = C[j][i];
                T cj[16];
#pragma unroll
                for (int u = 0; u < 16; ++u) cj[u] = C[j][16 * w + u];
#pragma unroll
                for (int u = 0; u < 16; ++u) a[u] = fma(-li, cj[u], a[u]);
            }
        }
        if (w == p) {
            // Column j's multiplier for column j + 1 comes by v_readlane,
            // and its terms for the columns after j + 1 are applied one
            // column later, so that no LDS round trip lies between one
            // column's square root and the next one's.
            T lp = T(0), cp[16];
#pragma unroll
            for (int jj = 0; jj < 16; ++jj) {
                const int j = 16 * p + jj;
                const T d = lane_value(a[jj], j);
                const T sjj = sqrt(d);
                if (j < nb && !(d > T(0)) && bad == 0) bad = col0 + j + 1;
                // column j - 1's terms, deferred
#pragma unroll
                for (int u = jj + 1; u < 16; ++u)
                    if (jj > 0) a[u] = fma(-lp, cp[u], a[u]);
                a[jj] = i == j ? sjj : a[jj] / sjj;
                const T li = (i >= j && i < nb && j < nb) ? a[jj] : (i == j ? T(1) : T(0));
                C[j][i] = li;
                // (this wave's own LDS writes are visible to its reads
                // without a barrier)
#pragma unroll
                for (int u = jj + 2; u < 16; ++u) cp[u] = C[j][16 * p + u];
                if (jj < 15) a[jj + 1] = fma(-li, lane_value(li, j + 1), a[jj + 1]);
                lp = li;
            }
        }
        if (p > 0 && w == p - 1) {
            inverse_terms<T>(C, R, x, w, i);
            inverse_finish<T>(C, R, x, w, i);
        }
        __syncthreads();
    }
    // L is final in C (C[c][r] = l_rc on and below the diagonal): its
    // stores go now, beside the last rows of X.
#pragma unroll 4
    for (int u = 0; u < 16; ++u) {
        const int r = 16 * w + u;
        if (r < nb && i <= r) A[(int64_t)r * lda + i] = C[i][r];
    }
    // Rows 48 .. 63 of X: the terms of rows 0 .. 47, four rows a wave (in
    // ascending k, as inverse_terms), then the group's own in wave 3.
    {
        T y[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) y[t] = 48 + 4 * w + t == i ? T(1) : T(0);
#pragma unroll 4
        for (int k = 0; k < 48; ++k) {
            const T xk = R[k][i];
            T ck[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) ck[t] = C[k][48 + 4 * w + t];
#pragma unroll
            for (int t = 0; t < 4; ++t) y[t] = fma(-ck[t], xk, y[t]);
        }
#pragma unroll
        for (int t = 0; t < 4; ++t) R[48 + 4 * w + t][i] = y[t];
    }
    __syncthreads();
    if (w == 3) {
#pragma unroll
        for (int u = 0; u < 16; ++u) x[u] = R[48 + u][i];
        inverse_finish<T>(C, R, x, 3, i);
    }
    if (i == 0) sbad[w] = bad;
    __syncthreads();
#pragma unroll 4
    for (int u = 0; u < 16; ++u) {
        const int r = 16 * w + u;
        WT[(int64_t)r * ldw + i] = R[i][r];  // WT[r][i] = x_ir
    }
    if (threadIdx.x == 0) {
        int first = 0;
        for (int q = 0; q < 4; ++q)
            if (sbad[q] != 0 && (first == 0 || sbad[q] < first)) first = sbad[q];
        if (first != 0 && *info == 0) *info = first;
    }
    if constexpr (MODE == 0) return;
    typedef typename V4<T>::type V;
    __syncthreads();
    if constexpr (MODE == 1) {
        // A10 (k1 rows) into C
        T *A10 = A + (int64_t)DIAG * lda, *A11 = A10 + DIAG;
        put_rows<T, DIAG + 2>(C, p0, k1, DIAG, w, i);
        __syncthreads();
        // L10 = A10 X^T (X^T[k][c] = R[c][k]), to A10 and, once every wave
        // has read A10, to C
        V acc[4];
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) acc[ct] = V{};
        tile_product<T>(&C[0][0], DIAG + 2, 1, &R[0][0], 1, DIAG + 1, w, i, acc);
        __syncthreads();
#pragma unroll
        for (int ct = 0; ct < 4; ++ct)
#pragma unroll
            for (int r4 = 0; r4 < 4; ++r4) {
                const int row = 16 * w + acc_row<T>(lk, r4), col = 16 * ct + lr;
                C[row][col] = acc[ct][r4];
                if (row < k1) A10[(int64_t)row * lda + col] = acc[ct][r4];
            }
        __syncthreads();
        // A11 -= L10 L10^T on the lower triangle of its k1 x k1 block
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) acc[ct] = V{};
        tile_product<T>(&C[0][0], DIAG + 2, 1, &C[0][0], 1, DIAG + 2, w, i, acc);
#pragma unroll
        for (int ct = 0; ct < 4; ++ct)
#pragma unroll
            for (int r4 = 0; r4 < 4; ++r4) {
                const int row = 16 * w + acc_row<T>(lk, r4), col = 16 * ct + lr;
                if (row < k1 && col <= row) A11[(int64_t)row * lda + col] = p1[4 * ct + r4] - acc[ct][r4];
            }
    } else {
        // L10 (nb rows, the DIAG columns left of A) into C
        put_rows<T, DIAG + 2>(C, p0, nb, DIAG, w, i);
        __syncthreads();
        // T1 = L10^T WT11 (T1[k][c] = sum_r L10[r][k] X[c][r]) into R, once
        // every wave has read X
        V acc[4];
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) acc[ct] = V{};
        tile_product<T>(&C[0][0], 1, DIAG + 2, &R[0][0], 1, DIAG + 1, w, i, acc);
        __syncthreads();
#pragma unroll
        for (int ct = 0; ct < 4; ++ct)
#pragma unroll
            for (int r4 = 0; r4 < 4; ++r4) R[16 * w + acc_row<T>(lk, r4)][16 * ct + lr] = acc[ct][r4];
        // WT00 into C; WT01 = -WT00 T1 for the nb columns
        T *WT01 = WT - DIAG * ldw;
        put_rows<T, DIAG + 2>(C, p1, DIAG, DIAG, w, i);
        __syncthreads();
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) acc[ct] = V{};
        tile_product<T>(&C[0][0], DIAG + 2, 1, &R[0][0], DIAG + 1, 1, w, i, acc);
#pragma unroll
        for (int ct = 0; ct < 4; ++ct)
#pragma unroll
            for (int r4 = 0; r4 < 4; ++r4) {
                const int row = 16 * w + acc_row<T>(lk, r4), col = 16 * ct + lr;
                if (col < nb) WT01[(int64_t)row * ldw + col] = -acc[ct][r4];
            }
    }
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
__global__ __launch_bounds__(256) void copy_block_kernel(const T *__restrict__ in, int64_t ldi, T *__restrict__ out,
                                                         int64_t ldo, int64_t rows, int cols) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= rows * cols) return;
    const int64_t r = e / cols;
    const int c = (int)(e - r * cols);
    out[r * ldo + c] = in[r * ldi + c];
}

template <class T>
void copy_block(hipStream_t s, const T *in, int64_t ldi, T *out, int64_t ldo, int64_t rows, int64_t cols) {
    if (rows <= 0 || cols <= 0) return;
    hipLaunchKernelGGL(copy_block_kernel<T>, dim3((unsigned)cdiv(rows * cols, 256)), dim3(256), 0, s, in, ldi, out,
                       ldo, rows, (int)cols);
    LMX_LAUNCH_CHECK();
}

// Factors a 64 x 64 (or smaller) diagonal block and writes its L^{-T} to WT
// (MODE as potrf_diag_kernel).
template <class T, int MODE = 0>
void diag_inv(hipStream_t s, int64_t nb, T *A, int64_t lda, int64_t col0, int32_t *d_info, T *WT, int64_t ldw,
              int64_t k1 = 0) {
    TimedRegion tr("potrf_diag", s);
    hipLaunchKernelGGL((potrf_diag_kernel<T, MODE>), dim3(1), dim3(256), 0, s, (int)nb, A, lda, (int)col0, d_info,
                       WT, ldw, (int)k1);
    LMX_LAUNCH_CHECK();
}

template <class T>
__global__ __launch_bounds__(256) void fill_zero_kernel(T *__restrict__ p, int64_t n) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e < n) p[e] = T(0);
}

constexpr int BLK = 4 * DIAG;  // block column width (64 times a power of two); WT's row stride

// The diagonal block A11 (order kb <= BLK) in place, and L11^{-T} (upper,
// kb x kb) in WT, whose part below the diagonal blocks the caller has set
// to zero.  Order > 64: halves of h (64 times a power of two, h < kb <=
// 2h) and kb - h; with L = [[L00, 0], [L10, L11]], L^{-T} = [[L00^{-T},
// -L00^{-T} L10^T L11^{-T}], [0, L11^{-T}]], the off-diagonal block as the
// inverse's doubling step forms it.  tmp holds (BLK / 2)^2.
template <class T>
void diag_block(hipStream_t s, bool bf16, int64_t kb, T *A11, int64_t lda, int64_t col0, int32_t *d_info, T *WT,
                T *tmp) {
    if (kb <= DIAG) {
        diag_inv<T>(s, kb, A11, lda, col0, d_info, WT, BLK);
        return;
    }
    if (kb <= 2 * DIAG && !bf16) {
        // two kernels: the first half with L10 and the second half's update,
        // the second half with the inverse's off-diagonal block
        const int64_t k1 = kb - DIAG;
        diag_inv<T, 1>(s, DIAG, A11, lda, col0, d_info, WT, BLK, k1);
        diag_inv<T, 2>(s, k1, A11 + DIAG * lda + DIAG, lda, col0 + DIAG, d_info, WT + DIAG * BLK + DIAG, BLK);
        return;
    }
    int64_t h = DIAG;
    while (2 * h < kb) h *= 2;
    const int64_t k1 = kb - h;
    diag_block<T>(s, bf16, h, A11, lda, col0, d_info, WT, tmp);
    T *A10 = A11 + h * lda, *A11b = A10 + h, *WT11 = WT + h * BLK + h;
    gemm(s, false, false, k1, h, h, T(1), A10, lda, WT, BLK, T(0), tmp, h);  // L10 = A10 L00^{-T}
    copy_block(s, tmp, h, A10, lda, k1, h);
    chol_update<T>(s, bf16, true, k1, k1, h, A10, lda, A10, A11b);
    diag_block<T>(s, bf16, k1, A11b, lda, col0 + h, d_info, WT11, tmp);
    gemm(s, true, false, h, k1, k1, T(1), A10, lda, WT11, BLK, T(0), tmp, k1);
    gemm(s, false, false, h, k1, h, T(-1), WT, BLK, tmp, k1, T(0), WT + h, BLK);
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
// The panel below a diagonal block is L21 = A21 L11^{-T}: one GEMM into
// scratch and a copy back.
template <class T>
void potrf_lookahead(hipStream_t s, int64_t n, T *A, int64_t lda, int32_t *d_info, bool bf16) {
    constexpr int64_t nb = BLK;
    AuxStream la;
    bool pending = false;
    Scratch sw((size_t)BLK * BLK * sizeof(T), s), st((size_t)(BLK / 2) * (BLK / 2) * sizeof(T), s);
    Scratch sx((size_t)std::max<int64_t>(n - nb, 1) * nb * sizeof(T), s);
    T *WT = sw.as<T>(), *tmp = st.as<T>(), *X = sx.as<T>();
    // the inverse's blocks below its diagonal blocks stay zero: nothing
    // writes there
    hipLaunchKernelGGL(fill_zero_kernel<T>, dim3((unsigned)cdiv(BLK * BLK, 256)), dim3(256), 0, s, WT,
                       (int64_t)BLK * BLK);
    LMX_LAUNCH_CHECK();
    for (int64_t k0 = 0; k0 < n; k0 += nb) {
        const int64_t kb = std::min(nb, n - k0);
        T *A11 = A + k0 * lda + k0;
        const int64_t rest = n - k0 - kb;
        T *A21 = A11 + kb * lda;
        diag_block<T>(s, bf16, kb, A11, lda, k0, d_info, WT, tmp);
        if (rest > 0) {
            gemm(s, false, false, rest, kb, kb, T(1), A21, lda, WT, BLK, T(0), X, kb);
            copy_block(s, X, kb, A21, lda, rest, kb);
        }
        // The trailing SYRK may start once the panel is in place: it and
        // the next block column's update below touch different elements.
        // Recording `ready` here keeps the event markers out of the gap
        // between that update and the next diagonal kernel.
        if (pending) LMX_HIP(hipStreamWaitEvent(s, la.done, 0));
        const int64_t narrow = std::min(nb, rest);
        LMX_HIP(hipEventRecord(la.ready, s));
        LMX_HIP(hipStreamWaitEvent(la.aux, la.ready, 0));
        if (rest > narrow) {
            if (!bf16) {
                // Few enough SYRK workgroups a CU that when one ends, the
                // diagonal kernel fits in its place: FP64 two a CU (80 KB
                // of LDS each; the kernel takes 67 KB and 4 x 224
                // registers), FP32 five (30 KB each; 34 KB and 4 x 128).
                // As many as fit, they left it waiting for the SYRK to
                // drain, up to 300 us a block (FP64 n = 8192: 12.7 against
                // 13.1 ms; FP32 n = 16384: 32.0 against 32.9 ms).
                const size_t fp = (size_t)(sizeof(T) == 8 ? 80 : 30) * 1024;
                syrk_lower_sharing<T>(la.aux, rest - narrow, kb, T(-1), A21 + narrow * lda, lda, T(1),
                                      A21 + narrow * lda + kb + narrow, lda, fp);
            } else {
                chol_update<T>(la.aux, bf16, true, rest - narrow, rest - narrow, kb, A21 + narrow * lda, lda,
                               A21 + narrow * lda, A21 + narrow * lda + kb + narrow);
            }
        }
        LMX_HIP(hipEventRecord(la.done, la.aux));
        pending = true;
        if (rest > 0 && !bf16) {
            // the next block column, on and below its diagonal, in one launch
            syrk_trap<T>(s, rest, narrow, kb, T(-1), A21, lda, T(1), A21 + kb, lda);
        } else if (rest > 0) {
            chol_update<T>(s, bf16, true, narrow, narrow, kb, A21, lda, A21, A21 + kb);
            if (rest > narrow)
                chol_update<T>(s, bf16, false, rest - narrow, narrow, kb, A21 + narrow * lda, lda, A21,
                               A21 + narrow * lda + kb);
        }
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
    potrf_lookahead<T>(s, n, A, lda, d_info, bf16);
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
