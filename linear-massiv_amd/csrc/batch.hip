// Batched solves of many independent small systems (order n <= 512, at
// most 8 right-hand sides each), LU with partial pivoting and Cholesky
// mixed freely, in one persistent launch.
//
// One workgroup solves one problem at a time and takes the next one from an
// atomic work counter; the problems are queued largest first, so the tail of
// the launch is made of small problems.  A problem's matrix stays in HBM
// (where the caller put it); the workgroup's LDS holds one 32-column block
// column at a time.
//
// LU (lu_factor) is the left-looking (Crout) order of partial-pivoting LU.
// At block column j0 (width NB):
//   C  = A[j0:n, j0:j0+NB] - L[j0:n, 0:j0] U[0:j0, j0:j0+NB]   (MFMA)
//   pivoted panel factorisation of C in LDS (rawPivotSearch's rule:
//   largest |a_ik|, first row on ties, a NaN never chosen below the
//   diagonal; multipliers by true division)
//   R  = L11^{-1} (A[piv rows, j0+NB:n] - L[piv rows, 0:j0] U[0:j0, j0+NB:n])
// Rows are never moved in memory: logical row r of the matrix lives in
// storage row r for the whole factorisation, and the permutation perm
// (position -> storage row, in LDS) is the only thing a row interchange
// changes.  Both updates therefore read every operand straight from its
// final place, and the swap list (LAPACK ipiv) is recorded as it happens.
// The inverses of the diagonal blocks of L and U are kept in the
// workgroup's slot of workspace, so the triangular solves are block
// products.
//
// Cholesky (chol_factor) is the left-looking lower factorisation of the
// reference's panelCholFactor (src/.../Solve/Cholesky.hs:279-321), reading
// only the lower triangle of A.
//
// Mixed precision: the factorisation runs in FP32 on a copy of A in the
// slot, and refinement in FP64 (LAPACK dsgesv: x = solve(b); repeat
// r = b - A x, stop when ||r||inf <= ||x||inf ||A||inf eps sqrt(n) for every
// right-hand side, else x += solve(r); at most 30 corrections), and falls
// back to the FP64 factorisation of A when it does not converge.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <numeric>
#include <string>
#include <vector>

#include "common.hpp"
#include "device.hpp"

namespace lmx {

namespace {

constexpr int BN = LMX_BATCH_MAX_N;       // largest order
constexpr int BR = LMX_BATCH_MAX_NRHS;    // most right-hand sides
constexpr int ITMAX = 30;

static_assert(sizeof(lmx_batch_problem) == 40, "lmx_batch_problem layout");

struct BatchArgs {
    const lmx_batch_problem *probs;  // sorted, `reserved` = index in the caller's array
    int count;
    unsigned *counter;
    double *a_pool;
    double *b_pool;
    int32_t *ipiv_pool;
    int32_t *info;
    int32_t *iters;
    unsigned char *ws;
    int64_t slot_bytes;
    int64_t off_inv;     // FP64-sized inverse blocks (L then U)
    int64_t off_inv32;   // FP32 inverse blocks (mixed)
    int64_t off_b;       // FP64 right-hand sides (mixed)
    int64_t off_g32;     // FP32 matrix copy (mixed)
    int64_t off_gather;  // FP64 gather buffer (factors in LAPACK layout)
    int nmax_pad;        // largest order rounded up to the block width
    int flags;
};

// LDS layout of a kernel variant (bytes).  The small arrays come first, so
// their offsets do not depend on the size of the region after them; the
// region is shared by the phases: the block column C, then the block row R,
// then the solve and refinement vectors (offsets relative to reg_off).
template <int NB, int NT, bool MIXED = true>
struct Lay {
    static constexpr int NW = NT / 64;
    static constexpr int LDC = NB + 1;
    static constexpr int LDR = BN + 1;
    static constexpr int NS = NT / NB;  // k-slices of a column-oriented block dot
    static constexpr size_t li_off = 0;
    static constexpr size_t ui_off = li_off + (size_t)NB * LDC * 8;
    static constexpr size_t perm_off = ui_off + (size_t)NB * LDC * 8;
    static constexpr size_t piv_off = perm_off + (size_t)BN * 4;
    static constexpr size_t red_off = piv_off + (size_t)BN * 4;
    static constexpr size_t reg_off = red_off + 1024;
    static constexpr size_t c_bytes = (size_t)BN * LDC * 8;
    static constexpr size_t r_bytes = (size_t)NB * LDR * 8;
    // solve (element type up to 8 bytes): X, S, P
    static constexpr size_t x_off = 0;
    static constexpr size_t s_off = x_off + (size_t)BN * BR * 8;
    static constexpr size_t p_off = s_off + (size_t)NB * BR * 8;
    static constexpr size_t solve_bytes = p_off + (size_t)NS * NB * BR * 8;
    // refinement: FP64 x and r, then an FP32 solve area
    static constexpr size_t ir_x = 0;
    static constexpr size_t ir_r = ir_x + (size_t)BN * BR * 8;
    static constexpr size_t ir_solve = ir_r + (size_t)BN * BR * 8;
    static constexpr size_t ir_bytes = MIXED ? ir_solve + solve_bytes : 0;
    static constexpr size_t m1 = c_bytes > r_bytes ? c_bytes : r_bytes;
    static constexpr size_t m2 = m1 > solve_bytes ? m1 : solve_bytes;
    static constexpr size_t region = m2 > ir_bytes ? m2 : ir_bytes;
    static constexpr size_t total = reg_off + region;
};

__device__ inline int64_t ix(int r, int c, int n) { return (int64_t)r * n + c; }

// ------------------------------------------------------------- small pieces

// Li (unit lower inverse) and Ui (upper inverse) of the w x w diagonal block
// held in C rows 0..w-1 (LU), padded to NB x NB with the identity; written
// to LDS and to the slot (row-major NB x NB).  Lane c of wave 0 builds
// column c of Li, lane c of wave 1 column c of Ui (separate waves, so the
// two run side by side); each element is a dot product whose loads are
// unrolled by 8 (one LDS round trip for eight terms, not one each).
template <class F, int NB, int NT>
__device__ void lu_tri_inverses_lds(const F *C, int w, F *Li, F *Ui, F *gL, F *gU) {
    static_assert(NT >= 128, "the two inverses run in waves 0 and 1");
    constexpr int LDC = NB + 1;
    const int t = threadIdx.x;
    if (t < NB) {
        const int c = t;
        for (int i = 0; i < NB; ++i) {
            F v;
            if (i < c) v = F(0);
            else if (i == c) v = F(1);
            else if (i < w && c < w) {
                F s = F(0);
#pragma unroll 8
                for (int k = c; k < i; ++k) s = fma(C[i * LDC + k], Li[k * LDC + c], s);
                v = -s;
            } else v = F(0);
            Li[i * LDC + c] = v;
        }
    } else if (t >= 64 && t < 64 + NB) {
        const int c = t - 64;
        for (int i = NB - 1; i >= 0; --i) {
            F v;
            if (i > c) v = F(0);
            else if (c >= w) v = (i == c) ? F(1) : F(0);
            else {
                F s = (i == c) ? F(1) : F(0);
#pragma unroll 8
                for (int k = i + 1; k <= c; ++k) s = fma(-C[i * LDC + k], Ui[k * LDC + c], s);
                v = s / C[i * LDC + i];
            }
            Ui[i * LDC + c] = v;
        }
    }
    __syncthreads();
    for (int e = t; e < NB * NB; e += NT) {
        const int i = e / NB, c = e % NB;
        gL[e] = Li[i * LDC + c];
        gU[e] = Ui[i * LDC + c];
    }
}

// Li = inverse of the non-unit lower w x w block in C (Cholesky); the dot
// products as in lu_tri_inverses.
template <class F, int NB, int NT>
__device__ void chol_tri_inverse_lds(const F *C, int w, F *Li, F *gL) {
    constexpr int LDC = NB + 1;
    const int t = threadIdx.x;
    if (t < NB) {
        const int c = t;
        for (int i = 0; i < NB; ++i) {
            F v;
            if (i < c) v = F(0);
            else if (c >= w) v = (i == c) ? F(1) : F(0);
            else if (i == c) v = F(1) / C[c * LDC + c];
            else if (i < w) {
                F s = F(0);
#pragma unroll 8
                for (int k = c; k < i; ++k) s = fma(C[i * LDC + k], Li[k * LDC + c], s);
                v = -s / C[i * LDC + i];
            } else v = F(0);
            Li[i * LDC + c] = v;
        }
    }
    __syncthreads();
    for (int e = t; e < NB * NB; e += NT) gL[e] = Li[(e / NB) * LDC + e % NB];
}

// The same inverses with each column held in registers by its lane (lane c
// of wave 0 for Li, of wave 1 for Ui), every element a dot product in
// ascending k over the column's triangle (terms outside it are skipped);
// the coefficients are broadcast LDS reads that do not depend on the
// column, so they issue together.  Used for FP32 factors: with FP64 the
// unrolled columns and their loads cost the kernel its registers.
template <class F, int NB, int NT>
__device__ void lu_tri_inverses_regs(const F *C, int w, F *Li, F *Ui, F *gL, F *gU) {
    static_assert(NT >= 128, "the two inverses run in waves 0 and 1");
    constexpr int LDC = NB + 1;
    const int t = threadIdx.x;
    if (t < NB) {
        const int c = t;
        F x[NB];
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            // Li[i][c] = -sum_{k=c}^{i-1} L[i][k] Li[k][c]  (i > c, both < w)
            F sum = F(0);
#pragma unroll
            for (int k = 0; k < i; ++k) {
                const F a = C[i * LDC + k];
                sum = k >= c ? fma(a, x[k], sum) : sum;
            }
            x[i] = i == c ? F(1) : ((i > c && i < w && c < w) ? -sum : F(0));
        }
#pragma unroll
        for (int i = 0; i < NB; ++i) Li[i * LDC + c] = x[i];
    } else if (t >= 64 && t < 64 + NB) {
        const int c = t - 64;
        F y[NB];
#pragma unroll
        for (int i = NB - 1; i >= 0; --i) {
            // Ui[i][c] = ((i == c) - sum_{k=i+1}^{c} U[i][k] Ui[k][c]) / U[i][i]
            F sum = i == c ? F(1) : F(0);
#pragma unroll
            for (int k = i + 1; k < NB; ++k) {
                const F a = C[i * LDC + k];
                sum = k <= c ? fma(-a, y[k], sum) : sum;
            }
            const F q = sum / C[i * LDC + i];
            y[i] = i > c ? F(0) : (c >= w ? (i == c ? F(1) : F(0)) : q);
        }
#pragma unroll
        for (int i = 0; i < NB; ++i) Ui[i * LDC + c] = y[i];
    }
    __syncthreads();
    for (int e = t; e < NB * NB; e += NT) {
        const int i = e / NB, c = e % NB;
        gL[e] = Li[i * LDC + c];
        gU[e] = Ui[i * LDC + c];
    }
}

template <class F, int NB, int NT>
__device__ void chol_tri_inverse_regs(const F *C, int w, F *Li, F *gL) {
    constexpr int LDC = NB + 1;
    const int t = threadIdx.x;
    if (t < NB) {
        const int c = t;
        F x[NB];
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            // Li[i][c] = -sum_{k=c}^{i-1} L[i][k] Li[k][c] / L[i][i] (i > c),
            // 1 / L[c][c] on the diagonal
            F sum = F(0);
#pragma unroll
            for (int k = 0; k < i; ++k) {
                const F a = C[i * LDC + k];
                sum = k >= c ? fma(a, x[k], sum) : sum;
            }
            const F dv = C[i * LDC + i];
            F v;
            if (i < c)
                v = F(0);
            else if (c >= w)
                v = i == c ? F(1) : F(0);
            else if (i == c)
                v = F(1) / dv;
            else if (i < w)
                v = -sum / dv;
            else
                v = F(0);
            x[i] = v;
        }
#pragma unroll
        for (int i = 0; i < NB; ++i) Li[i * LDC + c] = x[i];
    }
    __syncthreads();
    for (int e = t; e < NB * NB; e += NT) gL[e] = Li[(e / NB) * LDC + e % NB];
}

template <class F, int NB, int NT>
__device__ void lu_tri_inverses(const F *C, int w, F *Li, F *Ui, F *gL, F *gU) {
    if constexpr (sizeof(F) == 4)
        lu_tri_inverses_regs<F, NB, NT>(C, w, Li, Ui, gL, gU);
    else
        lu_tri_inverses_lds<F, NB, NT>(C, w, Li, Ui, gL, gU);
}

template <class F, int NB, int NT>
__device__ void chol_tri_inverse(const F *C, int w, F *Li, F *gL) {
    if constexpr (sizeof(F) == 4)
        chol_tri_inverse_regs<F, NB, NT>(C, w, Li, gL);
    else
        chol_tri_inverse_lds<F, NB, NT>(C, w, Li, gL);
}

// cr[cc] -= l * cp[cc * S] for cc in [c1, c2): one row's rank-1 update in
// LDS, against a row (S = 1) or a column (S = the row stride) that shares
// no element with the written ones.  The restrict qualifiers let the
// compiler issue an unrolled group's reads before its writes; without them
// every iteration waits for the previous write.
template <class F, int S>
__device__ __forceinline__ void row_axpy(F *__restrict__ cr, const F *__restrict__ cp, F l, int c1, int c2) {
#pragma unroll 8
    for (int cc = c1; cc < c2; ++cc) cr[cc] = fma(-l, cp[cc * S], cr[cc]);
}

// ------------------------------------------------------------------- LU

// Pivoted factorisation of the m x w block column C (LDS), positions
// relative to the block column.  perm and piv are offset to the block
// column's first position.  Returns the first zero pivot (1-based) or 0.
template <class F, int NB, int NT>
__device__ int lu_panel(F *C, int m, int w, int *perm, int *piv, F *red_v, int *red_r) {
    constexpr int LDC = NB + 1, NW = NT / 64;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    int info = 0;
    for (int c = 0; c < w; ++c) {
        F best = F(-1);
        int brow = INT_MAX;
        for (int r = c + t; r < m; r += NT) {
            const F v = fabs(C[r * LDC + c]);
            if (v > best) {
                best = v;
                brow = r;
            }
        }
        const F wb = wave_max(best);
        const int wr = wave_min(best == wb ? brow : INT_MAX);
        if (lane == 0) {
            red_v[wave] = wb;
            red_r[wave] = wr;
        }
        __syncthreads();
        F gb = red_v[0];
#pragma unroll
        for (int x = 1; x < NW; ++x) gb = fmax(gb, red_v[x]);
        int p = INT_MAX;
#pragma unroll
        for (int x = 0; x < NW; ++x)
            if (red_v[x] == gb) p = min(p, red_r[x]);
        if (p == INT_MAX || isnan(C[c * LDC + c])) p = c;
        if (p != c) {
            if (t < w) {
                const F a = C[c * LDC + t];
                C[c * LDC + t] = C[p * LDC + t];
                C[p * LDC + t] = a;
            }
            if (t == NT - 1) {
                const int a = perm[c];
                perm[c] = perm[p];
                perm[p] = a;
            }
        }
        if (t == NT - 2) piv[c] = p;
        __syncthreads();
        const F pv = C[c * LDC + c];
        if (pv != F(0)) {
            for (int r = c + 1 + t; r < m; r += NT) {
                const F l = C[r * LDC + c] / pv;
                C[r * LDC + c] = l;
                row_axpy<F, 1>(C + r * LDC, C + c * LDC, l, c + 1, w);
            }
        } else if (info == 0) {
            info = c + 1;
        }
        // No barrier here: the next column's search reads, in each thread,
        // the rows this thread just updated (rows c + 1 + t + k NT both
        // times); its interchange touches other threads' rows only after the
        // search's barrier.
    }
    __syncthreads();
    return info;
}

// LU of the n x n matrix G (row-major, storage rows never move).  On return
// perm[i] is the storage row of position i and piv[] the swap list; the
// inverses of the diagonal blocks are at gLinv / gUinv (NB x NB each).
template <class F, int NB, int NT>
__device__ int lu_factor(F *G, int n, unsigned char *sm, F *gLinv, F *gUinv) {
    typedef Lay<NB, NT> L;
    typedef typename V4<F>::type V;
    constexpr int LDC = L::LDC, NW = L::NW, CT = NB / 16;
    constexpr int LDR = L::LDR;
    constexpr int MRT = (BN / 16 + NW - 1) / NW;  // row tiles per wave
    F *C = (F *)(sm + L::reg_off);
    F *Rs = C;
    F *Li = (F *)(sm + L::li_off);
    F *Ui = (F *)(sm + L::ui_off);
    int *perm = (int *)(sm + L::perm_off);
    int *piv = (int *)(sm + L::piv_off);
    F *red_v = (F *)(sm + L::red_off);
    int *red_r = (int *)(sm + L::red_off + 64 * 8);
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6, lr = lane & 15, lk = lane >> 4;

    for (int i = t; i < n; i += NT) perm[i] = i;
    __syncthreads();
    int info = 0;
    for (int j0 = 0; j0 < n; j0 += NB) {
        const int w = min(NB, n - j0), m = n - j0;
        const int RT = (m + 15) >> 4;
        // C = A[j0:n, j0:j0+w] - L[j0:n, 0:j0] U[0:j0, j0:j0+w]
        {
            V acc[MRT][CT];
#pragma unroll
            for (int q = 0; q < MRT; ++q) {
                const int rt = wave + q * NW;
#pragma unroll
                for (int ct = 0; ct < CT; ++ct)
#pragma unroll
                    for (int r4 = 0; r4 < 4; ++r4) {
                        const int row = j0 + rt * 16 + acc_row<F>(lk, r4), col = ct * 16 + lr;
                        acc[q][ct][r4] = (rt < RT && row < n && col < w) ? G[ix(perm[row], j0 + col, n)] : F(0);
                    }
            }
            // The operands of a k step are loaded unconditionally (addresses
            // clamped into the matrix, out-of-range values zeroed after) and
            // all before the MFMAs: loads under a per-tile branch wait for
            // memory one tile after another.  The A rows' storage indices do
            // not change over the k loop.
            int arow[MRT];
#pragma unroll
            for (int q = 0; q < MRT; ++q) arow[q] = perm[min(j0 + (wave + q * NW) * 16 + lr, n - 1)];
            for (int k0 = 0; k0 < j0; k0 += 16) {
                int brow[4];
#pragma unroll
                for (int s = 0; s < 4; ++s) brow[s] = perm[k0 + lk * 4 + s];
                F bv[CT][4], av[MRT][4];
#pragma unroll
                for (int ct = 0; ct < CT; ++ct)
#pragma unroll
                    for (int s = 0; s < 4; ++s) {
                        const int col = ct * 16 + lr;
                        const F v = G[ix(brow[s], j0 + min(col, w - 1), n)];
                        bv[ct][s] = col < w ? v : F(0);
                    }
#pragma unroll
                for (int q = 0; q < MRT; ++q)
#pragma unroll
                    for (int s = 0; s < 4; ++s) av[q][s] = G[ix(arow[q], k0 + lk * 4 + s, n)];
#pragma unroll
                for (int q = 0; q < MRT; ++q) {
                    const int rt = wave + q * NW;
                    if (rt < RT) {
                        const bool live = j0 + rt * 16 + lr < n;
#pragma unroll
                        for (int s = 0; s < 4; ++s)
#pragma unroll
                            for (int ct = 0; ct < CT; ++ct)
                                acc[q][ct] = mfma(live ? -av[q][s] : F(0), bv[ct][s], acc[q][ct]);
                    }
                }
            }
#pragma unroll
            for (int q = 0; q < MRT; ++q) {
                const int rt = wave + q * NW;
#pragma unroll
                for (int ct = 0; ct < CT; ++ct)
#pragma unroll
                    for (int r4 = 0; r4 < 4; ++r4) {
                        const int rl = rt * 16 + acc_row<F>(lk, r4), col = ct * 16 + lr;
                        if (rt < RT && rl < m && col < w) C[rl * LDC + col] = acc[q][ct][r4];
                    }
            }
        }
        __syncthreads();
        const int pinfo = lu_panel<F, NB, NT>(C, m, w, perm + j0, piv + j0, red_v, red_r);
        if (info == 0 && pinfo != 0) info = j0 + pinfo;
        if (w == NB) {
            for (int e = t; e < m * NB; e += NT) {
                const int r = e / NB, c = e % NB;
                G[ix(perm[j0 + r], j0 + c, n)] = C[r * LDC + c];
            }
        } else {
            for (int e = t; e < m * w; e += NT) {
                const int r = e / w, c = e - r * w;
                G[ix(perm[j0 + r], j0 + c, n)] = C[r * LDC + c];
            }
        }
        lu_tri_inverses<F, NB, NT>(C, w, Li, Ui, gLinv + (int64_t)(j0 / NB) * NB * NB,
                                   gUinv + (int64_t)(j0 / NB) * NB * NB);
        __syncthreads();
        if (j0 + w >= n) break;
        // R = Li (A[piv rows, c0:n] - L[piv rows, 0:j0] U[0:j0, c0:n]), w == NB here
        {
            const int c0 = j0 + NB, ncols = n - c0;
            const int NCT = (ncols + 15) >> 4;
            constexpr int MCT = (BN / 16 + NW - 1) / NW;
            V acc[MCT][CT];
#pragma unroll
            for (int q = 0; q < MCT; ++q) {
                const int ct = wave + q * NW;
#pragma unroll
                for (int rt = 0; rt < CT; ++rt)
#pragma unroll
                    for (int r4 = 0; r4 < 4; ++r4) {
                        const int row = j0 + rt * 16 + acc_row<F>(lk, r4), col = c0 + ct * 16 + lr;
                        acc[q][rt][r4] = (ct < NCT && col < n) ? G[ix(perm[row], col, n)] : F(0);
                    }
            }
            // loads as in the block-column update: all of a k step's
            // operands in flight before its MFMAs
            int arow[CT];
#pragma unroll
            for (int rt = 0; rt < CT; ++rt) arow[rt] = perm[j0 + rt * 16 + lr];
            for (int k0 = 0; k0 < j0; k0 += 16) {
                F av[CT][4];
#pragma unroll
                for (int rt = 0; rt < CT; ++rt)
#pragma unroll
                    for (int s = 0; s < 4; ++s) av[rt][s] = -G[ix(arow[rt], k0 + lk * 4 + s, n)];
                int srow[4];
#pragma unroll
                for (int s = 0; s < 4; ++s) srow[s] = perm[k0 + lk * 4 + s];
                F bv[MCT][4];
#pragma unroll
                for (int q = 0; q < MCT; ++q) {
                    const int col = c0 + (wave + q * NW) * 16 + lr;
#pragma unroll
                    for (int s = 0; s < 4; ++s) {
                        const F v = G[ix(srow[s], min(col, n - 1), n)];
                        bv[q][s] = col < n ? v : F(0);
                    }
                }
#pragma unroll
                for (int q = 0; q < MCT; ++q) {
                    const int ct = wave + q * NW;
                    if (ct < NCT) {
#pragma unroll
                        for (int s = 0; s < 4; ++s)
#pragma unroll
                            for (int rt = 0; rt < CT; ++rt) acc[q][rt] = mfma(av[rt][s], bv[q][s], acc[q][rt]);
                    }
                }
            }
#pragma unroll
            for (int q = 0; q < MCT; ++q) {
                const int ct = wave + q * NW;
#pragma unroll
                for (int rt = 0; rt < CT; ++rt)
#pragma unroll
                    for (int r4 = 0; r4 < 4; ++r4) {
                        const int rl = rt * 16 + acc_row<F>(lk, r4), cl = ct * 16 + lr;
                        if (ct < NCT) Rs[rl * LDR + cl] = acc[q][rt][r4];
                    }
            }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < MCT; ++q) {
                const int ct = wave + q * NW;
                if (ct < NCT) {
#pragma unroll
                    for (int rt = 0; rt < CT; ++rt) {
                        V o = {0, 0, 0, 0};
#pragma unroll
                        for (int kk = 0; kk < NB; kk += 4)
                            o = mfma(Li[(rt * 16 + lr) * LDC + kk + lk], Rs[(kk + lk) * LDR + ct * 16 + lr], o);
#pragma unroll
                        for (int r4 = 0; r4 < 4; ++r4) {
                            const int row = j0 + rt * 16 + acc_row<F>(lk, r4), col = c0 + ct * 16 + lr;
                            if (col < n) G[ix(perm[row], col, n)] = o[r4];
                        }
                    }
                }
            }
            __syncthreads();
        }
    }
    return info;
}

// ------------------------------------------------------------- Cholesky

// Lower Cholesky of G (reads and writes only the lower triangle).  Returns
// the first column whose pivot is not positive (1-based) or 0.
template <class F, int NB, int NT>
__device__ int chol_factor(F *G, int n, unsigned char *sm, F *gLinv) {
    typedef Lay<NB, NT> L;
    typedef typename V4<F>::type V;
    constexpr int LDC = L::LDC, NW = L::NW, CT = NB / 16;
    constexpr int MRT = (BN / 16 + NW - 1) / NW;
    F *C = (F *)(sm + L::reg_off);
    F *Li = (F *)(sm + L::li_off);
    F *red_v = (F *)(sm + L::red_off);
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6, lr = lane & 15, lk = lane >> 4;
    for (int j0 = 0; j0 < n; j0 += NB) {
        const int w = min(NB, n - j0), m = n - j0;
        const int RT = (m + 15) >> 4;
        {
            V acc[MRT][CT];
#pragma unroll
            for (int q = 0; q < MRT; ++q) {
                const int rt = wave + q * NW;
#pragma unroll
                for (int ct = 0; ct < CT; ++ct)
#pragma unroll
                    for (int r4 = 0; r4 < 4; ++r4) {
                        const int rl = rt * 16 + acc_row<F>(lk, r4), col = ct * 16 + lr;
                        acc[q][ct][r4] = (rt < RT && rl < m && col < w && col <= rl) ? G[ix(j0 + rl, j0 + col, n)] : F(0);
                    }
            }
            // all of a k step's operands in flight before its MFMAs (as in
            // lu_factor)
            for (int k0 = 0; k0 < j0; k0 += 16) {
                F bv[CT][4], av[MRT][4];
#pragma unroll
                for (int ct = 0; ct < CT; ++ct)
#pragma unroll
                    for (int s = 0; s < 4; ++s) {
                        const int col = ct * 16 + lr;
                        const F v = G[ix(j0 + min(col, w - 1), k0 + lk * 4 + s, n)];
                        bv[ct][s] = col < w ? v : F(0);
                    }
#pragma unroll
                for (int q = 0; q < MRT; ++q) {
                    const int row = min(j0 + (wave + q * NW) * 16 + lr, n - 1);
#pragma unroll
                    for (int s = 0; s < 4; ++s) av[q][s] = G[ix(row, k0 + lk * 4 + s, n)];
                }
#pragma unroll
                for (int q = 0; q < MRT; ++q) {
                    const int rt = wave + q * NW;
                    if (rt < RT) {
                        const bool live = j0 + rt * 16 + lr < n;
#pragma unroll
                        for (int s = 0; s < 4; ++s)
#pragma unroll
                            for (int ct = 0; ct < CT; ++ct)
                                acc[q][ct] = mfma(live ? -av[q][s] : F(0), bv[ct][s], acc[q][ct]);
                    }
                }
            }
#pragma unroll
            for (int q = 0; q < MRT; ++q) {
                const int rt = wave + q * NW;
#pragma unroll
                for (int ct = 0; ct < CT; ++ct)
#pragma unroll
                    for (int r4 = 0; r4 < 4; ++r4) {
                        const int rl = rt * 16 + acc_row<F>(lk, r4), col = ct * 16 + lr;
                        if (rt < RT && rl < m && col < w) C[rl * LDC + col] = col <= rl ? acc[q][ct][r4] : F(0);
                    }
            }
        }
        __syncthreads();
        // column by column: sqrt of the pivot, true division below it, then
        // the rank-1 update of the block column's remaining lower part.  The
        // square roots wait in dg and enter the diagonal after the loop, so
        // every thread's read of the pivot needs no barrier against its
        // overwrite (the updates read only below the diagonal).
        F *dg = red_v;
        for (int c = 0; c < w; ++c) {
            const F d = C[c * LDC + c];
            if (!(d > F(0))) return j0 + c + 1;
            const F sd = sqrt(d);
            if (t == 0) dg[c] = sd;
            for (int r = c + 1 + t; r < m; r += NT) C[r * LDC + c] = C[r * LDC + c] / sd;
            __syncthreads();
            for (int r = c + 1 + t; r < m; r += NT) {
                const F l = C[r * LDC + c];
                row_axpy<F, LDC>(C + r * LDC, C + c, l, c + 1, min(r, w - 1) + 1);
            }
            __syncthreads();
        }
        for (int c = t; c < w; c += NT) C[c * LDC + c] = dg[c];
        __syncthreads();
        if (w == NB) {
            for (int e = t; e < m * NB; e += NT) {
                const int r = e / NB, c = e % NB;
                if (c <= r) G[ix(j0 + r, j0 + c, n)] = C[r * LDC + c];
            }
        } else {
            for (int e = t; e < m * w; e += NT) {
                const int r = e / w, c = e - r * w;
                if (c <= r) G[ix(j0 + r, j0 + c, n)] = C[r * LDC + c];
            }
        }
        chol_tri_inverse<F, NB, NT>(C, w, Li, gLinv + (int64_t)(j0 / NB) * NB * NB);
        __syncthreads();
    }
    return 0;
}

// ---------------------------------------------------------------- solves

// S[r][j] = X[kb + r][j] - sum_{k in [k1, k2)} M[srow(kb + r)][k] X[k][j],
// r < w (row-oriented: 16 lanes per row, stride 16 over k).
template <class F, int NB, int NT>
__device__ void row_dots(const F *G, int n, const int *perm, int kb, int w, int k1, int k2, const F *X, int nrhs,
                         F *S) {
    const int t = threadIdx.x, sub = t & 15;
    for (int r = t >> 4; r < NB; r += NT / 16) {
        F part[BR];
#pragma unroll
        for (int j = 0; j < BR; ++j) part[j] = F(0);
        if (r < w) {
            const int srow = perm ? perm[kb + r] : kb + r;
            const F *Mr = G + (int64_t)srow * n;
            for (int k = k1 + sub; k < k2; k += 16) {
                const F a = Mr[k];
#pragma unroll
                for (int j = 0; j < BR; ++j)
                    if (j < nrhs) part[j] = fma(a, X[k * BR + j], part[j]);
            }
        }
#pragma unroll
        for (int j = 0; j < BR; ++j) part[j] = row16_sum(part[j]);
        if (sub == 0 && r < w) {
#pragma unroll
            for (int j = 0; j < BR; ++j)
                if (j < nrhs) S[r * BR + j] = X[(kb + r) * BR + j] - part[j];
        }
    }
}

// S[r][j] = X[kb + r][j] - sum_{k in [k1, n)} M[k][kb + r] X[k][j]
// (column-oriented, for the transposed factor).
template <class F, int NB, int NT>
__device__ void col_dots(const F *G, int n, int kb, int w, int k1, const F *X, int nrhs, F *S, F *P) {
    constexpr int NS = NT / NB;
    const int t = threadIdx.x, r = t % NB, ks = t / NB;
    F part[BR];
#pragma unroll
    for (int j = 0; j < BR; ++j) part[j] = F(0);
    if (r < w)
        for (int k = k1 + ks; k < n; k += NS) {
            const F a = G[ix(k, kb + r, n)];
#pragma unroll
            for (int j = 0; j < BR; ++j)
                if (j < nrhs) part[j] = fma(a, X[k * BR + j], part[j]);
        }
#pragma unroll
    for (int j = 0; j < BR; ++j) P[(ks * NB + r) * BR + j] = part[j];
    __syncthreads();
    for (int e = t; e < NB * BR; e += NT) {
        const int rr = e / BR, j = e % BR;
        if (rr < w && j < nrhs) {
            F s = F(0);
            for (int x = 0; x < NS; ++x) s += P[(x * NB + rr) * BR + j];
            S[rr * BR + j] = X[(kb + rr) * BR + j] - s;
        }
    }
}

// X[kb + r][j] = sum_c Inv[r][c] S[c][j]  (TR: Inv^T), r < w.
template <class F, int NB, int NT, bool TR>
__device__ void apply_inv(const F *Inv, int kb, int w, const F *S, F *X, int nrhs) {
    for (int e = threadIdx.x; e < NB * BR; e += NT) {
        const int r = e / BR, j = e % BR;
        if (r < w && j < nrhs) {
            F s = F(0);
            for (int c = 0; c < w; ++c) s = fma(TR ? Inv[c * NB + r] : Inv[r * NB + c], S[c * BR + j], s);
            X[(kb + r) * BR + j] = s;
        }
    }
}

// Solves with the factors in place: X (LDS, BN x BR) holds the right-hand
// sides in position order (LU: row i is b[perm[i]]) and receives x.
template <class F, int NB, int NT>
__device__ void lu_solve(const F *G, int n, const int *perm, const F *gLinv, const F *gUinv, F *X, int nrhs,
                         F *S) {
    for (int kb = 0; kb < n; kb += NB) {
        const int w = min(NB, n - kb);
        row_dots<F, NB, NT>(G, n, perm, kb, w, 0, kb, X, nrhs, S);
        __syncthreads();
        apply_inv<F, NB, NT, false>(gLinv + (int64_t)(kb / NB) * NB * NB, kb, w, S, X, nrhs);
        __syncthreads();
    }
    for (int kb = ((n - 1) / NB) * NB; kb >= 0; kb -= NB) {
        const int w = min(NB, n - kb);
        row_dots<F, NB, NT>(G, n, perm, kb, w, kb + w, n, X, nrhs, S);
        __syncthreads();
        apply_inv<F, NB, NT, false>(gUinv + (int64_t)(kb / NB) * NB * NB, kb, w, S, X, nrhs);
        __syncthreads();
    }
}

template <class F, int NB, int NT>
__device__ void chol_solve(const F *G, int n, const F *gLinv, F *X, int nrhs, F *S, F *P) {
    for (int kb = 0; kb < n; kb += NB) {
        const int w = min(NB, n - kb);
        row_dots<F, NB, NT>(G, n, nullptr, kb, w, 0, kb, X, nrhs, S);
        __syncthreads();
        apply_inv<F, NB, NT, false>(gLinv + (int64_t)(kb / NB) * NB * NB, kb, w, S, X, nrhs);
        __syncthreads();
    }
    for (int kb = ((n - 1) / NB) * NB; kb >= 0; kb -= NB) {
        const int w = min(NB, n - kb);
        col_dots<F, NB, NT>(G, n, kb, w, kb + w, X, nrhs, S, P);
        __syncthreads();
        apply_inv<F, NB, NT, true>(gLinv + (int64_t)(kb / NB) * NB * NB, kb, w, S, X, nrhs);
        __syncthreads();
    }
}

// ------------------------------------------------------------- refinement

// R = B - A X in FP64 (X, R in LDS, BN x BR; B row-major n x nrhs in HBM)
// and, into *anorm, ||A||inf.  Cholesky reads only the lower triangle of A
// (A = tril + tril^T - diag): a row pass over k <= i, then a column pass
// over k > i in strips of NB columns.  scratch: BN + (NS + 1) NB (BR + 1)
// doubles of LDS.
template <int NB, int NT>
__device__ void residual(bool chol, const double *A, int n, const double *B, int nrhs, const double *X, double *R,
                         double *scratch, double *anorm, double *red) {
    constexpr int NS = NT / NB;
    double *rowsum = scratch;           // BN
    double *Q = scratch + BN;           // NS x NB x BR
    double *Qa = Q + NS * NB * BR;      // NS x NB
    const int t = threadIdx.x, sub = t & 15;
    for (int i = t >> 4; i < n; i += NT / 16) {
        double part[BR];
#pragma unroll
        for (int j = 0; j < BR; ++j) part[j] = 0.0;
        double asum = 0.0;
        const double *Ar = A + (int64_t)i * n;
        const int kend = chol ? i + 1 : n;
        for (int k = sub; k < kend; k += 16) {
            const double a = Ar[k];
            asum += fabs(a);
#pragma unroll
            for (int j = 0; j < BR; ++j)
                if (j < nrhs) part[j] = fma(a, X[k * BR + j], part[j]);
        }
#pragma unroll
        for (int j = 0; j < BR; ++j) part[j] = row16_sum(part[j]);
        asum = row16_sum(asum);
        if (sub == 0) {
#pragma unroll
            for (int j = 0; j < BR; ++j)
                if (j < nrhs) R[i * BR + j] = B[(int64_t)i * nrhs + j] - part[j];
            rowsum[i] = asum;
        }
    }
    __syncthreads();
    if (chol) {
        for (int c0 = 0; c0 < n; c0 += NB) {
            const int r = t % NB, ks = t / NB, i = c0 + r;
            double part[BR];
#pragma unroll
            for (int j = 0; j < BR; ++j) part[j] = 0.0;
            double asum = 0.0;
            if (i < n)
                for (int k = i + 1 + ks; k < n; k += NS) {
                    const double a = A[ix(k, i, n)];
                    asum += fabs(a);
#pragma unroll
                    for (int j = 0; j < BR; ++j)
                        if (j < nrhs) part[j] = fma(a, X[k * BR + j], part[j]);
                }
#pragma unroll
            for (int j = 0; j < BR; ++j) Q[(ks * NB + r) * BR + j] = part[j];
            Qa[ks * NB + r] = asum;
            __syncthreads();
            for (int e = t; e < NB * (BR + 1); e += NT) {
                const int rr = e / (BR + 1), j = e % (BR + 1), ii = c0 + rr;
                if (ii >= n) continue;
                double acc = 0.0;
                if (j < BR) {
                    if (j >= nrhs) continue;
                    for (int x = 0; x < NS; ++x) acc += Q[(x * NB + rr) * BR + j];
                    R[ii * BR + j] -= acc;
                } else {
                    for (int x = 0; x < NS; ++x) acc += Qa[x * NB + rr];
                    rowsum[ii] += acc;
                }
            }
            __syncthreads();
        }
    }
    double amax = 0.0;
    for (int i = t; i < n; i += NT) amax = fmax(amax, rowsum[i]);
    const double wm = wave_max(amax);
    if ((t & 63) == 0) red[t >> 6] = wm;
    __syncthreads();
    double v = 0.0;
#pragma unroll
    for (int x = 0; x < NT / 64; ++x) v = fmax(v, red[x]);
    *anorm = v;
    __syncthreads();
}

// out[j] = max_i |V[i][j]| (V in LDS, n x BR); a NaN counts as +inf.
template <int NT>
__device__ void col_absmax(const double *V, int n, int nrhs, double *out, double *red) {
    const int t = threadIdx.x;
    for (int j = 0; j < nrhs; ++j) {
        double m = 0.0;
        for (int i = t; i < n; i += NT) {
            double a = fabs(V[i * BR + j]);
            if (a != a) a = __builtin_inf();
            m = fmax(m, a);
        }
        const double wm = wave_max(m);
        if ((t & 63) == 0) red[t >> 6] = wm;
        __syncthreads();
        if (t == 0) {
            double v = 0.0;
            for (int x = 0; x < NT / 64; ++x) v = fmax(v, red[x]);
            out[j] = v;
        }
        __syncthreads();
    }
}

// ----------------------------------------------------------------- kernel

template <class F, int NB, int NT>
struct SolveAreas {
    F *X, *S, *P;
};
template <class F, int NB, int NT>
__device__ inline SolveAreas<F, NB, NT> solve_areas(unsigned char *base) {
    typedef Lay<NB, NT> L;
    return {(F *)(base + L::x_off), (F *)(base + L::s_off), (F *)(base + L::p_off)};
}

// The FP64 factorisation and solve of one problem in place on A and B.
template <int NB, int NT>
__device__ int solve_fp64(const lmx_batch_problem &p, double *A, double *B, unsigned char *sm, double *gL,
                          double *gU) {
    typedef Lay<NB, NT> L;
    const int n = p.n, nrhs = p.nrhs, t = threadIdx.x;
    const bool chol = p.kind == LMX_BATCH_CHOLESKY;
    int info = chol ? chol_factor<double, NB, NT>(A, n, sm, gL) : lu_factor<double, NB, NT>(A, n, sm, gL, gU);
    __syncthreads();
    if (info != 0) return info;
    const int *perm = chol ? nullptr : (const int *)(sm + L::perm_off);
    SolveAreas<double, NB, NT> a = solve_areas<double, NB, NT>(sm + L::reg_off);
    for (int e = t; e < n * BR; e += NT) {
        const int i = e / BR, j = e % BR;
        a.X[e] = j < nrhs ? B[(int64_t)(perm ? perm[i] : i) * nrhs + j] : 0.0;
    }
    __syncthreads();
    if (chol) chol_solve<double, NB, NT>(A, n, gL, a.X, nrhs, a.S, a.P);
    else lu_solve<double, NB, NT>(A, n, perm, gL, gU, a.X, nrhs, a.S);
    for (int e = t; e < n * nrhs; e += NT) B[e] = a.X[(e / nrhs) * BR + e % nrhs];
    __syncthreads();
    return 0;
}

template <int NB, int NT, bool MIXED>
__global__ __launch_bounds__(NT) void batch_kernel(BatchArgs g) {
    typedef Lay<NB, NT> L;
    __shared__ __align__(16) unsigned char sm[Lay<NB, NT, MIXED>::total];
    __shared__ int s_idx;
    __shared__ int s_flag;
    const int t = threadIdx.x;
    unsigned char *slot = g.ws + (int64_t)blockIdx.x * g.slot_bytes;
    double *gL = (double *)(slot + g.off_inv);
    double *gU = gL + (int64_t)g.nmax_pad * NB;
    for (;;) {
        if (t == 0) s_idx = (int)atomicAdd(g.counter, 1u);
        __syncthreads();
        const int idx = s_idx;
        __syncthreads();
        if (idx >= g.count) break;
        const lmx_batch_problem p = g.probs[idx];
        const int n = p.n, nrhs = p.nrhs;
        const bool chol = p.kind == LMX_BATCH_CHOLESKY;
        double *A = g.a_pool + p.a_offset;
        double *B = g.b_pool + p.b_offset;
        int info = 0, iters = 0;
        bool fp64 = !MIXED;
        if (MIXED) {
            float *G = (float *)(slot + g.off_g32);
            float *fL = (float *)(slot + g.off_inv32);
            float *fU = fL + (int64_t)g.nmax_pad * NB;
            double *bsave = (double *)(slot + g.off_b);
            for (int e = t; e < n * n; e += NT) {
                const int i = e / n, k = e - i * n;
                if (!chol || k <= i) G[e] = (float)A[e];
            }
            for (int e = t; e < n * nrhs; e += NT) bsave[e] = B[e];
            __syncthreads();
            const int info32 = chol ? chol_factor<float, NB, NT>(G, n, sm, fL)
                                    : lu_factor<float, NB, NT>(G, n, sm, fL, fU);
            __syncthreads();
            if (info32 != 0) {
                fp64 = true;
            } else {
                const int *perm = chol ? nullptr : (const int *)(sm + L::perm_off);
                unsigned char *reg = sm + L::reg_off;
                double *Xd = (double *)(reg + L::ir_x);
                double *Rd = (double *)(reg + L::ir_r);
                SolveAreas<float, NB, NT> a = solve_areas<float, NB, NT>(reg + L::ir_solve);
                double *red = (double *)(sm + L::red_off);
                double *colmax = red + 40;  // BR values
                double *xmax = red + 48;    // BR values
                double anorm = 0.0;
                // x0 = solve(b)
                for (int e = t; e < n * BR; e += NT) {
                    const int i = e / BR, j = e % BR;
                    a.X[e] = j < nrhs ? (float)bsave[(int64_t)(perm ? perm[i] : i) * nrhs + j] : 0.f;
                }
                __syncthreads();
                if (chol) chol_solve<float, NB, NT>(G, n, fL, a.X, nrhs, a.S, a.P);
                else lu_solve<float, NB, NT>(G, n, perm, fL, fU, a.X, nrhs, a.S);
                for (int e = t; e < n * BR; e += NT) Xd[e] = (double)a.X[e];
                __syncthreads();
                bool converged = false;
                for (int it = 0; it <= ITMAX; ++it) {
                    // the residual pass keeps its partial sums in the FP32
                    // solve area, which is free here
                    double an = 0.0;
                    residual<NB, NT>(chol, A, n, bsave, nrhs, Xd, Rd, (double *)(reg + L::ir_solve), &an, red);
                    if (it == 0) anorm = an;
                    __syncthreads();
                    col_absmax<NT>(Rd, n, nrhs, colmax, red);
                    col_absmax<NT>(Xd, n, nrhs, xmax, red);
                    if (t == 0) {
                        const double cte = anorm * 2.220446049250313e-16 * sqrt((double)n);
                        bool ok = true;
                        for (int j = 0; j < nrhs; ++j)
                            ok = ok && colmax[j] <= xmax[j] * cte && colmax[j] < __builtin_inf() &&
                                 xmax[j] < __builtin_inf();
                        s_flag = ok ? 1 : 0;
                    }
                    __syncthreads();
                    if (s_flag) {
                        converged = true;
                        iters = it;
                        break;
                    }
                    if (it == ITMAX) break;
                    for (int e = t; e < n * BR; e += NT) {
                        const int i = e / BR, j = e % BR;
                        a.X[e] = j < nrhs ? (float)Rd[(perm ? perm[i] : i) * BR + j] : 0.f;
                    }
                    __syncthreads();
                    if (chol) chol_solve<float, NB, NT>(G, n, fL, a.X, nrhs, a.S, a.P);
                    else lu_solve<float, NB, NT>(G, n, perm, fL, fU, a.X, nrhs, a.S);
                    for (int e = t; e < n * BR; e += NT) Xd[e] += (double)a.X[e];
                    __syncthreads();
                }
                if (converged) {
                    for (int e = t; e < n * nrhs; e += NT) B[e] = Xd[(e / nrhs) * BR + e % nrhs];
                } else {
                    fp64 = true;
                    for (int e = t; e < n * nrhs; e += NT) B[e] = bsave[e];
                }
                __syncthreads();
            }
            if (fp64) iters = -1;
        }
        if (fp64) info = solve_fp64<NB, NT>(p, A, B, sm, gL, gU);
        if (!chol && fp64 && g.ipiv_pool) {
            const int *piv = (const int *)(sm + L::piv_off);
            for (int k = t; k < n; k += NT) {
                // positions are relative to their block column
                g.ipiv_pool[p.ipiv_offset + k] = (k / NB) * NB + piv[k];
            }
        }
        if (!chol && fp64 && (g.flags & LMX_BATCH_FACTORS)) {
            // storage row perm[i] -> row i (LAPACK layout) through the slot
            const int *perm = (const int *)(sm + L::perm_off);
            double *tmp = (double *)(slot + g.off_gather);
            for (int e = t; e < n * n; e += NT) {
                const int i = e / n, k = e - i * n;
                tmp[e] = A[ix(perm[i], k, n)];
            }
            __syncthreads();
            for (int e = t; e < n * n; e += NT) A[e] = tmp[e];
        }
        if (t == 0) {
            g.info[p.reserved] = info;
            if (g.iters) g.iters[p.reserved] = iters;
        }
        __syncthreads();
    }
}

__global__ void batch_reset_kernel(unsigned *counter) {
    if (threadIdx.x == 0) *counter = 0u;
}

// ------------------------------------------------------------------- host

struct Variant {
    int nb, nt;
    const char *name;
};
constexpr Variant VARIANTS[] = {{32, 512, "nb32_t512"}, {32, 256, "nb32_t256"}, {16, 256, "nb16_t256"},
                                {16, 512, "nb16_t512"}};
constexpr int NVARIANTS = 4;

const void *kernel_of(int v, bool mixed) {
    switch (v) {
        case 0: return mixed ? (const void *)batch_kernel<32, 512, true> : (const void *)batch_kernel<32, 512, false>;
        case 1: return mixed ? (const void *)batch_kernel<32, 256, true> : (const void *)batch_kernel<32, 256, false>;
        case 2: return mixed ? (const void *)batch_kernel<16, 256, true> : (const void *)batch_kernel<16, 256, false>;
        default: return mixed ? (const void *)batch_kernel<16, 512, true> : (const void *)batch_kernel<16, 512, false>;
    }
}

double problem_work(const lmx_batch_problem &p) {
    const double n = p.n;
    return (p.kind == LMX_BATCH_CHOLESKY ? n * n * n / 3.0 : 2.0 * n * n * n / 3.0) + 2.0 * n * n * p.nrhs;
}

}  // namespace

}  // namespace lmx

struct lmx_batch_plan {
    int device = 0;
    int count = 0;
    int precision = LMX_PREC_F64;
    int flags = 0;
    int nmax = 0;
    int variant = 0;
    int grid = 0;
    int64_t a_elems = 0, b_elems = 0, ipiv_elems = 0;
    double flops = 0;  // FP64-equivalent LAPACK count of the batch
    lmx_batch_problem *d_probs = nullptr;
    unsigned *d_counter = nullptr;
    unsigned char *d_ws = nullptr;
    size_t ws_bytes = 0;
    lmx::BatchArgs args{};
    std::vector<lmx_batch_problem> sorted;
    // captured graph, keyed by the launch arguments it was captured with
    hipGraphExec_t graph = nullptr;
    hipStream_t graph_stream = nullptr;
    double *g_a = nullptr, *g_b = nullptr;
    int32_t *g_ipiv = nullptr, *g_info = nullptr, *g_iters = nullptr;
    int g_variant = -1;
    double tuned_ms[lmx::NVARIANTS] = {0, 0, 0, 0};
    size_t probs_cap = 0;
};

namespace lmx {
namespace {

void configure(lmx_batch_plan *pl, int v) {
    const Variant &var = VARIANTS[v];
    const bool mixed = pl->precision == LMX_PREC_F32;
    const int nb = var.nb;
    const int nmax_pad = (int)cdiv(std::max(pl->nmax, 1), nb) * nb;
    int64_t off = 0;
    auto take = [&](int64_t bytes) {
        int64_t o = off;
        off += (bytes + 255) / 256 * 256;
        return o;
    };
    BatchArgs &a = pl->args;
    a.off_inv = take(2LL * nmax_pad * nb * 8);
    a.off_inv32 = mixed ? take(2LL * nmax_pad * nb * 4) : 0;
    a.off_b = mixed ? take((int64_t)pl->nmax * BR * 8) : 0;
    a.off_g32 = mixed ? take((int64_t)pl->nmax * pl->nmax * 4) : 0;
    a.off_gather = (pl->flags & LMX_BATCH_FACTORS) ? take((int64_t)pl->nmax * pl->nmax * 8) : 0;
    a.slot_bytes = std::max<int64_t>(off, 256);
    a.nmax_pad = nmax_pad;
    a.flags = pl->flags;
    int per_cu = 0;
    LMX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel_of(v, mixed), var.nt, 0));
    if (per_cu < 1) throw Error(LMX_ERR_INTERNAL, std::string("batch kernel ") + var.name + " cannot be resident");
    pl->grid = std::max(1, std::min(per_cu * ctx().cu_count, pl->count));
    const size_t need = (size_t)pl->grid * (size_t)a.slot_bytes;
    if (need > pl->ws_bytes) {
        if (pl->d_ws) LMX_HIP(hipFree(pl->d_ws));
        pl->d_ws = nullptr;
        pl->ws_bytes = 0;
        hipError_t e = hipMalloc(&pl->d_ws, need);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            throw Error(LMX_ERR_NOMEM, "batch workspace of " + std::to_string(need) + " bytes");
        }
        pl->ws_bytes = need;
    }
    a.ws = pl->d_ws;
    pl->variant = v;
}

void launch(lmx_batch_plan *pl, hipStream_t s) {
    const bool mixed = pl->precision == LMX_PREC_F32;
    hipLaunchKernelGGL(batch_reset_kernel, dim3(1), dim3(64), 0, s, pl->d_counter);
    LMX_LAUNCH_CHECK();
    BatchArgs a = pl->args;
    void *kargs[] = {&a};
    LMX_HIP(hipLaunchKernel(kernel_of(pl->variant, mixed), dim3(pl->grid), dim3(VARIANTS[pl->variant].nt), kargs, 0,
                            s));
}

void drop_graph(lmx_batch_plan *pl) {
    if (pl->graph) (void)hipGraphExecDestroy(pl->graph);
    pl->graph = nullptr;
    pl->g_variant = -1;
}

void execute(lmx_batch_plan *pl, double *A, double *B, int32_t *ipiv, int32_t *info, int32_t *iters,
             hipStream_t s) {
    BatchArgs &a = pl->args;
    a.a_pool = A;
    a.b_pool = B;
    a.ipiv_pool = ipiv;
    a.info = info;
    a.iters = iters;
    TimedRegion tr("batch", s, pl->flops);
    if (!(pl->flags & LMX_BATCH_GRAPH)) {
        launch(pl, s);
        return;
    }
    const bool same = pl->graph && pl->g_a == A && pl->g_b == B && pl->g_ipiv == ipiv && pl->g_info == info &&
                      pl->g_iters == iters && pl->g_variant == pl->variant;
    if (!same) {
        drop_graph(pl);
        hipStream_t cap;
        LMX_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
        hipGraph_t gr = nullptr;
        LMX_HIP(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
        try {
            launch(pl, cap);
        } catch (...) {
            (void)hipStreamEndCapture(cap, &gr);
            if (gr) (void)hipGraphDestroy(gr);
            (void)hipStreamDestroy(cap);
            throw;
        }
        LMX_HIP(hipStreamEndCapture(cap, &gr));
        hipError_t e = hipGraphInstantiate(&pl->graph, gr, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gr);
        (void)hipStreamDestroy(cap);
        LMX_HIP(e);
        pl->g_a = A;
        pl->g_b = B;
        pl->g_ipiv = ipiv;
        pl->g_info = info;
        pl->g_iters = iters;
        pl->g_variant = pl->variant;
    }
    LMX_HIP(hipGraphLaunch(pl->graph, s));
}

// Synthetic problems of the plan's shapes for the autotuner: LU matrices
// with a dominant diagonal, SPD matrices for Cholesky.
__global__ void tune_fill_kernel(const lmx_batch_problem *probs, int count, double *A, double *B) {
    const int idx = blockIdx.x;
    if (idx >= count) return;
    const lmx_batch_problem p = probs[idx];
    const int n = p.n;
    double *a = A + p.a_offset;
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
        const int i = e / n, k = e % n;
        unsigned h = (unsigned)(e * 2654435761u) ^ (unsigned)(idx * 40503u + 17u);
        h ^= h >> 13;
        h *= 0x5bd1e995u;
        h ^= h >> 15;
        const double u = (double)(h & 0xFFFFFF) / 16777216.0 - 0.5;
        a[e] = (i == k) ? (double)n : (p.kind == LMX_BATCH_CHOLESKY ? 0.5 / (1.0 + fabs((double)(i - k))) : u);
    }
    double *b = B + p.b_offset;
    for (int e = threadIdx.x; e < n * p.nrhs; e += blockDim.x) b[e] = 1.0;
}

void autotune(lmx_batch_plan *pl, hipStream_t s) {
    Scratch sa((size_t)std::max<int64_t>(pl->a_elems, 1) * 8, s);
    Scratch sb((size_t)std::max<int64_t>(pl->b_elems, 1) * 8, s);
    Scratch si((size_t)std::max(pl->count, 1) * 4 * 2, s);
    Scratch sp((size_t)std::max<int64_t>(pl->ipiv_elems, 1) * 4, s);
    hipEvent_t e0, e1;
    LMX_HIP(hipEventCreate(&e0));
    LMX_HIP(hipEventCreate(&e1));
    const int saved_flags = pl->flags;
    pl->flags &= ~LMX_BATCH_GRAPH;
    int best = 0;
    double best_ms = 1e300;
    try {
        for (int v = 0; v < NVARIANTS; ++v) {
            configure(pl, v);
            float ms_min = 1e30f;
            for (int rep = 0; rep < 2; ++rep) {
                hipLaunchKernelGGL(tune_fill_kernel, dim3((unsigned)pl->count), dim3(256), 0, s, pl->d_probs,
                                   pl->count, sa.as<double>(), sb.as<double>());
                LMX_LAUNCH_CHECK();
                LMX_HIP(hipEventRecord(e0, s));
                execute(pl, sa.as<double>(), sb.as<double>(), pl->ipiv_elems > 0 ? sp.as<int32_t>() : nullptr,
                        si.as<int32_t>(), si.as<int32_t>() + pl->count, s);
                LMX_HIP(hipEventRecord(e1, s));
                LMX_HIP(hipEventSynchronize(e1));
                float ms = 0;
                LMX_HIP(hipEventElapsedTime(&ms, e0, e1));
                ms_min = std::min(ms_min, ms);
            }
            pl->tuned_ms[v] = ms_min;
            if (ms_min < best_ms) {
                best_ms = ms_min;
                best = v;
            }
        }
    } catch (...) {
        pl->flags = saved_flags;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    pl->flags = saved_flags;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    configure(pl, best);
    drop_graph(pl);
}

}  // namespace

void batch_replan(lmx_batch_plan *pl, const lmx_batch_problem *problems, int32_t count, int64_t a_pool_elems,
                  int64_t b_pool_elems, int64_t ipiv_pool_elems, hipStream_t s) {
    if (count < 0) fail_arg("count < 0");
    if (count > 0 && !problems) fail_arg("problems is NULL");
    std::vector<lmx_batch_problem> sorted(problems, problems + count);
    int nmax = 0;
    double flops = 0;
    for (int i = 0; i < count; ++i) {
        lmx_batch_problem &p = sorted[i];
        const std::string at = "problem " + std::to_string(i) + ": ";
        if (p.kind != LMX_BATCH_LU && p.kind != LMX_BATCH_CHOLESKY) fail_arg(at + "kind must be LU or Cholesky");
        if (p.n < 1 || p.n > BN) fail_arg(at + "n must be in [1, " + std::to_string(BN) + "]");
        if (p.nrhs < 1 || p.nrhs > BR) fail_arg(at + "nrhs must be in [1, " + std::to_string(BR) + "]");
        const int64_t n = p.n;
        if (p.a_offset < 0 || p.a_offset + n * n > a_pool_elems) fail_arg(at + "A lies outside the A pool");
        if (p.b_offset < 0 || p.b_offset + n * p.nrhs > b_pool_elems) fail_arg(at + "B lies outside the B pool");
        if (p.kind == LMX_BATCH_LU && ipiv_pool_elems > 0 &&
            (p.ipiv_offset < 0 || p.ipiv_offset + n > ipiv_pool_elems))
            fail_arg(at + "ipiv lies outside the ipiv pool");
        p.reserved = i;
        nmax = std::max(nmax, p.n);
        flops += problem_work(p);
    }
    // largest first (longest-processing-time order)
    std::stable_sort(sorted.begin(), sorted.end(), [](const lmx_batch_problem &x, const lmx_batch_problem &y) {
        return problem_work(x) > problem_work(y);
    });
    ctx();
    pl->count = count;
    pl->a_elems = a_pool_elems;
    pl->b_elems = b_pool_elems;
    pl->ipiv_elems = ipiv_pool_elems;
    pl->nmax = nmax;
    pl->flops = flops;
    pl->sorted.swap(sorted);
    if ((size_t)count > pl->probs_cap) {
        if (pl->d_probs) LMX_HIP(hipFree(pl->d_probs));
        pl->d_probs = nullptr;
        pl->probs_cap = 0;
        LMX_HIP(hipMalloc(&pl->d_probs, sizeof(lmx_batch_problem) * (size_t)count));
        pl->probs_cap = (size_t)count;
    }
    if (count > 0) {
        if (s) {
            LMX_HIP(hipMemcpyAsync(pl->d_probs, pl->sorted.data(), sizeof(lmx_batch_problem) * (size_t)count,
                                   hipMemcpyHostToDevice, s));
        } else {
            LMX_HIP(hipMemcpy(pl->d_probs, pl->sorted.data(), sizeof(lmx_batch_problem) * (size_t)count,
                              hipMemcpyHostToDevice));
        }
    }
    if (!pl->d_counter) LMX_HIP(hipMalloc(&pl->d_counter, 256));
    pl->args.probs = pl->d_probs;
    pl->args.count = count;
    pl->args.counter = pl->d_counter;
    drop_graph(pl);
    if (count > 0) configure(pl, pl->variant);
}

size_t batch_slot_bytes(int nmax, int precision, int flags) {
    const int nb = VARIANTS[0].nb;
    const int64_t nmax_pad = cdiv(std::max(nmax, 1), nb) * nb;
    const bool mixed = precision == LMX_PREC_F32;
    auto r = [](int64_t b) { return (b + 255) / 256 * 256; };
    int64_t b = r(2LL * nmax_pad * 32 * 8) + (mixed ? r(2LL * nmax_pad * 32 * 4) + r((int64_t)nmax * BR * 8) +
                                                          r((int64_t)nmax * nmax * 4)
                                                    : 0) +
                ((flags & LMX_BATCH_FACTORS) ? r((int64_t)nmax * nmax * 8) : 0);
    return (size_t)b;
}

}  // namespace lmx

using namespace lmx;

extern "C" {

int lmx_batch_plan_create(lmx_batch_plan **plan, const lmx_batch_problem *problems, int32_t count,
                          int64_t a_pool_elems, int64_t b_pool_elems, int64_t ipiv_pool_elems, int32_t precision,
                          int32_t flags) {
    return guarded([&]() -> int {
        if (!plan) fail_arg("plan is NULL");
        *plan = nullptr;
        if (precision != LMX_PREC_F64 && precision != LMX_PREC_F32)
            fail_arg("batched solves support LMX_PREC_F64 and LMX_PREC_F32 (FP32 factor, FP64 refinement)");
        if (flags & ~(LMX_BATCH_FACTORS | LMX_BATCH_GRAPH | LMX_BATCH_AUTOTUNE)) fail_arg("unknown flag bits");
        std::unique_ptr<lmx_batch_plan> pl(new lmx_batch_plan());
        pl->precision = precision;
        pl->flags = flags;
        batch_replan(pl.get(), problems, count, a_pool_elems, b_pool_elems, ipiv_pool_elems, nullptr);
        if (count > 0 && (flags & LMX_BATCH_AUTOTUNE)) {
            hipStream_t s = ctx().stream;
            autotune(pl.get(), s);
            LMX_HIP(hipStreamSynchronize(s));
        }
        *plan = pl.release();
        return LMX_OK;
    });
}

int lmx_batch_plan_execute(lmx_batch_plan *plan, double *a_pool, double *b_pool, int32_t *ipiv_pool, int32_t *info,
                           int32_t *iterations, void *stream) {
    return guarded([&]() -> int {
        if (!plan) fail_arg("plan is NULL");
        if (plan->count == 0) return LMX_OK;
        if (!a_pool || !b_pool || !info) fail_arg("a_pool, b_pool and info must be device pointers");
        if (plan->ipiv_elems > 0 && !ipiv_pool) fail_arg("the plan has an ipiv pool but ipiv_pool is NULL");
        execute(plan, a_pool, b_pool, plan->ipiv_elems > 0 ? ipiv_pool : nullptr, info, iterations,
                resolve_stream(stream));
        return LMX_OK;
    });
}

int lmx_batch_plan_autotune(lmx_batch_plan *plan, void *stream) {
    return guarded([&]() -> int {
        if (!plan) fail_arg("plan is NULL");
        if (plan->count == 0) return LMX_OK;
        hipStream_t s = resolve_stream(stream);
        autotune(plan, s);
        LMX_HIP(hipStreamSynchronize(s));
        return LMX_OK;
    });
}

int lmx_batch_plan_query(const lmx_batch_plan *plan, int32_t *variant, int32_t *block_cols, int32_t *threads,
                         int32_t *workgroups, double *flops, double *tuned_ms) {
    return guarded([&]() -> int {
        if (!plan) fail_arg("plan is NULL");
        if (variant) *variant = plan->variant;
        if (block_cols) *block_cols = VARIANTS[plan->variant].nb;
        if (threads) *threads = VARIANTS[plan->variant].nt;
        if (workgroups) *workgroups = plan->grid;
        if (flops) *flops = plan->flops;
        if (tuned_ms)
            for (int v = 0; v < NVARIANTS; ++v) tuned_ms[v] = plan->tuned_ms[v];
        return LMX_OK;
    });
}

int lmx_batch_plan_destroy(lmx_batch_plan *plan) {
    return guarded([&]() -> int {
        if (!plan) return LMX_OK;
        drop_graph(plan);
        if (plan->d_ws) (void)hipFree(plan->d_ws);
        if (plan->d_probs) (void)hipFree(plan->d_probs);
        if (plan->d_counter) (void)hipFree(plan->d_counter);
        delete plan;
        return LMX_OK;
    });
}

}  // extern "C"
