/*
 * lmx_oracle.c -- CPU restatement of linear-massiv's dense-solve algorithms.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle: it is loaded by
 * tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and
 * nowhere else.  The product library (liblmx.so) never links, loads or calls
 * it, and nothing in the product falls back to it.
 *
 * Every routine restates one reference routine, in the reference's own loop
 * order and rounding order, so that the oracle reproduces the reference's
 * floating-point result:
 *   - an FMA in the reference (GHC's fmaddDoubleX4#, lane-wise) is fma();
 *   - a separate multiply and add in the reference (`acc +## a *## b`) is a
 *     multiply and an add here; the file is compiled with -ffp-contract=off
 *     so the C compiler never fuses them;
 *   - 4-lane SIMD accumulators are restated as 4-element arrays and reduced
 *     in the reference's order ((a + b) + c) + d.
 * All matrices are row-major and dense (leading dimension = column count),
 * as massiv's P arrays are (src/Numeric/LinearAlgebra/Massiv/Types.hs:90).
 *
 * Pinning (what checks each function; see tests/test_oracle.py):
 *   orc_matmulp        KATs test/Test/BLAS.hs:132-144, exact integer
 *                      products, brute-force FMA-chain emulation in exact
 *                      rational arithmetic, numpy within gamma_k|A||B|.
 *   orc_gemm_generic   KAT test/Test/BLAS.hs:156-172, exact rational
 *                      emulation of the mul+add chain.
 *   orc_lu_generic     KAT test/Test/Solve.hs:107-119, LAPACK getrf factors
 *                      (scipy), tie-breaking closed form, PA = LU residual.
 *   orc_lu_panel_factor / orc_lusolvep
 *                      LAPACK getrf pivots and factors (idamax == first max),
 *                      permutation-matrix inputs, Hilbert-5 residual,
 *                      mpmath 50-digit solutions, unblocked == blocked path.
 *   orc_cholesky_*     LAPACK potrf (scipy), diagonal and tridiagonal closed
 *                      forms, A - GG^T residual.
 *   orc_*sub*          scipy solve_triangular, triangular systems with known
 *                      integer solutions.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_API __attribute__((visibility("default")))

typedef int64_t ix;

static ix min_ix(ix a, ix b) { return a < b ? a : b; }

/* ------------------------------------------------------------------------ */
/* GEMM                                                                     */
/* ------------------------------------------------------------------------ */

/*
 * C += A * B with the rounding of rawGemmKernel
 * (src/Numeric/LinearAlgebra/Massiv/Internal/Kernel.hs:145-806).
 *
 * Both of its paths (unpacked rawGemmBIBJSlice :163-406 and packed
 * rawGemmBISlicePackedBK :412-805) tile k in steps of 64 and give every
 * element of C the same arithmetic:
 *   - columns j < n - n%4 (the DoubleX4# 4x8, 4x4, 1x8 and 1x4 blocks,
 *     :232-279, :286-301, :339-348, :361-366, :624-671): the accumulator is
 *     loaded from C, receives fma(a_ik, b_kj, acc) for ascending k inside a
 *     k tile and is stored back, so over all tiles C is one ascending FMA
 *     chain;
 *   - columns j >= n - n%4 (goJScalar4/goJScalar1, :305-331, :370-382):
 *     per 64-wide k tile a partial sum starting at 0.0 accumulates
 *     acc + a_ik*b_kj with a rounded product, and C += partial after the tile.
 * lda, ldb, ldc are row strides.  The loops below are ordered for the host
 * cache (column blocks of 512, k blocks of 256, rows in fours); the order in
 * which one element sees its k terms is the reference's.
 */
static void gemm_acc_rawkernel(ix m, ix k, ix n,
                               const double *A, ix lda,
                               const double *B, ix ldb,
                               double *C, ix ldc)
{
    const ix n4 = n - (n % 4);
    for (ix bj = 0; bj < n4; bj += 512) {
        ix je = min_ix(bj + 512, n4);
        for (ix bk = 0; bk < k; bk += 256) {
            ix ke = min_ix(bk + 256, k);
            ix i = 0;
            for (; i + 4 <= m; i += 4) {
                double *c0 = C + i * ldc, *c1 = c0 + ldc, *c2 = c1 + ldc,
                       *c3 = c2 + ldc;
                const double *a0 = A + i * lda, *a1 = a0 + lda,
                             *a2 = a1 + lda, *a3 = a2 + lda;
                for (ix kk = bk; kk < ke; ++kk) {
                    const double *b = B + kk * ldb;
                    double x0 = a0[kk], x1 = a1[kk], x2 = a2[kk], x3 = a3[kk];
                    for (ix j = bj; j < je; ++j) {
                        double bv = b[j];
                        c0[j] = fma(x0, bv, c0[j]);
                        c1[j] = fma(x1, bv, c1[j]);
                        c2[j] = fma(x2, bv, c2[j]);
                        c3[j] = fma(x3, bv, c3[j]);
                    }
                }
            }
            for (; i < m; ++i) {
                double *c = C + i * ldc;
                const double *a = A + i * lda;
                for (ix kk = bk; kk < ke; ++kk) {
                    const double *b = B + kk * ldb;
                    double x = a[kk];
                    for (ix j = bj; j < je; ++j)
                        c[j] = fma(x, b[j], c[j]);
                }
            }
        }
    }
    for (ix i = 0; i < m; ++i) {
        const double *a = A + i * lda;
        double *c = C + i * ldc;
        for (ix j = n4; j < n; ++j) {
            for (ix bk = 0; bk < k; bk += 64) {
                ix kend = min_ix(bk + 64, k);
                double part = 0.0;
                for (ix kk = bk; kk < kend; ++kk) {
                    double prod = a[kk] * B[kk * ldb + j];
                    part = part + prod;
                }
                c[j] = c[j] + part;
            }
        }
    }
}

/* matMulP (src/Numeric/LinearAlgebra/Massiv/BLAS/Level3.hs:183-207):
 * zero C, then rawGemmKernel.  A is m x k, B is k x n, C is m x n. */
ORC_API void orc_matmulp(ix m, ix k, ix n, const double *A, const double *B,
                         double *C)
{
    memset(C, 0, (size_t)(m * n) * sizeof(double));
    gemm_acc_rawkernel(m, k, n, A, k, B, n, C, n);
}

/* Generic gemm (src/Numeric/LinearAlgebra/Massiv/BLAS/Level3.hs:102-131):
 * C = beta*C0, then for every (i, l) in ascending l order
 * C[i,j] = C[i,j] + (alpha*a_il)*b_lj with separate roundings.  The 32-wide
 * blocking of that loop does not change the per-element order (l ascends
 * through the bk blocks).  Cout may alias C0. */
ORC_API void orc_gemm_generic(ix m, ix k, ix n, double alpha,
                              const double *A, const double *B, double beta,
                              const double *C0, double *Cout)
{
    for (ix i = 0; i < m * n; ++i)
        Cout[i] = beta * C0[i];
    for (ix i = 0; i < m; ++i)
        for (ix l = 0; l < k; ++l) {
            double aik = alpha * A[i * k + l];
            for (ix j = 0; j < n; ++j) {
                double prod = aik * B[l * n + j];
                Cout[i * n + j] = Cout[i * n + j] + prod;
            }
        }
}

/* matMulGeneric (src/Numeric/LinearAlgebra/Massiv/BLAS/Level3.hs:150-181):
 * the path `matMul` takes for element types other than P Double. */
ORC_API void orc_matmul_generic(ix m, ix k, ix n, const double *A,
                                const double *B, double *C)
{
    memset(C, 0, (size_t)(m * n) * sizeof(double));
    for (ix i = 0; i < m; ++i)
        for (ix l = 0; l < k; ++l) {
            double aik = A[i * k + l];
            for (ix j = 0; j < n; ++j) {
                double prod = aik * B[l * n + j];
                C[i * n + j] = C[i * n + j] + prod;
            }
        }
}

/* ------------------------------------------------------------------------ */
/* LU (P Double fast path)                                                  */
/* ------------------------------------------------------------------------ */

/* rawPivotSearch (Internal/Kernel.hs:1541-1556): argmax |A[i,k]| over
 * i >= fromRow; strict '>' keeps the first maximum; a NaN is chosen only
 * when it sits at fromRow. */
static ix pivot_search(const double *A, ix n, ix k, ix fromRow)
{
    double v0 = A[fromRow * n + k];
    double best = (v0 >= 0.0) ? v0 : -v0;
    ix bi = fromRow;
    for (ix i = fromRow + 1; i < n; ++i) {
        double v = A[i * n + k];
        double av = (v >= 0.0) ? v : -v;
        if (av > best) {
            best = av;
            bi = i;
        }
    }
    return bi;
}

/* rawSwapRows (Internal/Kernel.hs:1507-1537) over columns [fromCol, n). */
static void swap_rows(double *A, ix n, ix r1, ix r2, ix fromCol)
{
    for (ix j = fromCol; j < n; ++j) {
        double t = A[r1 * n + j];
        A[r1 * n + j] = A[r2 * n + j];
        A[r2 * n + j] = t;
    }
}

/* rawLUEliminateColumnTo (Internal/Kernel.hs:1464-1503); with colEnd = n it
 * is rawLUEliminateColumn (:1414-1458).  mult = a_ik / a_kk (a division),
 * stored at A[i,k] for every row i > k; columns k+1 .. colEnd-1 of row i get
 * fma(-mult, a_kj, a_ij) in the SIMD span and a_ij - mult*a_kj (rounded
 * product) in the scalar tail of (colEnd-k-1) % 4 columns. */
static void eliminate_column_to(double *A, ix n, ix k, ix colEnd)
{
    const double akk = A[k * n + k];
    const ix jspan = colEnd - k - 1;
    const ix j4end = (k + 1) + (jspan - jspan % 4);
    const double *rk = A + k * n;
    for (ix i = k + 1; i < n; ++i) {
        double *ri = A + i * n;
        double mult = ri[k] / akk;
        ri[k] = mult;
        for (ix j = k + 1; j < j4end; ++j)
            ri[j] = fma(-mult, rk[j], ri[j]);
        for (ix j = j4end; j < colEnd; ++j) {
            double prod = mult * rk[j];
            ri[j] = ri[j] - prod;
        }
    }
}

/* rawTriSolvePanelTrail (Solve/LU.hs:364-407): U12 <- L11^{-1} A12, row by
 * row, for k in [k0, panelEnd), i in (k, panelEnd):
 *   a_ij = fma(-l_ik, a_kj, a_ij) over the 4-aligned span of [panelEnd, n),
 *   a_ij = a_ij + (-l_ik)*a_kj (rounded product) over the tail. */
static void trisolve_panel_trail(double *A, ix n, ix k0, ix panelEnd)
{
    const ix jspan = n - panelEnd;
    const ix j4end = panelEnd + (jspan - jspan % 4);
    for (ix k = k0; k < panelEnd; ++k) {
        const double *rk = A + k * n;
        for (ix i = k + 1; i < panelEnd; ++i) {
            double *ri = A + i * n;
            double neg = -ri[k];
            for (ix j = panelEnd; j < j4end; ++j)
                ri[j] = fma(neg, rk[j], ri[j]);
            for (ix j = j4end; j < n; ++j) {
                double prod = neg * rk[j];
                ri[j] = ri[j] + prod;
            }
        }
    }
}

/*
 * panelLUFactor (Solve/LU.hs:311-357), right-looking with panel width nb:
 * per panel column k: pivot search, full-row swap, elimination restricted to
 * the panel; then rawTriSolvePanelTrail, then A22 -= L21*U12 where the
 * product is formed by rawGemmKernel into a zeroed dense buffer (:336-351)
 * and subtracted element-wise (rawSubtractFromStrided :444-476; its SIMD
 * `aij + (-cij)` and scalar `aij - cij` round identically).
 * swaps[k] receives the pivot row chosen at step k (the (k, pivRow) list the
 * reference returns, in ascending k).  Returns the number of steps taken.
 */
ORC_API ix orc_lu_panel_factor(ix n, double *A, ix nb, ix *swaps)
{
    ix steps = 0;
    double *L = NULL, *U = NULL, *C = NULL;
    for (ix k0 = 0; k0 < n - 1;) {
        ix panelEnd = min_ix(k0 + nb, n);
        ix w = panelEnd - k0;
        for (ix k = k0; k < panelEnd; ++k) {
            ix p = pivot_search(A, n, k, k);
            if (p != k)
                swap_rows(A, n, k, p, 0);
            eliminate_column_to(A, n, k, panelEnd);
            swaps[k] = p;
            ++steps;
        }
        if (panelEnd < n) {
            trisolve_panel_trail(A, n, k0, panelEnd);
            ix mt = n - panelEnd, nt = n - panelEnd;
            L = (double *)realloc(L, (size_t)(mt * w) * sizeof(double));
            U = (double *)realloc(U, (size_t)(w * nt) * sizeof(double));
            C = (double *)realloc(C, (size_t)(mt * nt) * sizeof(double));
            for (ix i = 0; i < mt; ++i)
                for (ix j = 0; j < w; ++j)
                    L[i * w + j] = A[(panelEnd + i) * n + k0 + j];
            for (ix i = 0; i < w; ++i)
                for (ix j = 0; j < nt; ++j)
                    U[i * nt + j] = A[(k0 + i) * n + panelEnd + j];
            memset(C, 0, (size_t)(mt * nt) * sizeof(double));
            gemm_acc_rawkernel(mt, w, nt, L, w, U, nt, C, nt);
            for (ix i = 0; i < mt; ++i)
                for (ix j = 0; j < nt; ++j) {
                    double *a = &A[(panelEnd + i) * n + panelEnd + j];
                    *a = *a - C[i * nt + j];
                }
        }
        k0 = panelEnd;
    }
    free(L);
    free(U);
    free(C);
    return steps;
}

/* The unblocked path of luSolveP for n < 64 (Solve/LU.hs:278-284):
 * k = 0 .. n-2, full-width elimination (rawLUEliminateColumn). */
ORC_API ix orc_lu_unblocked_factor(ix n, double *A, ix *swaps)
{
    ix steps = 0;
    for (ix k = 0; k + 1 < n; ++k) {
        ix p = pivot_search(A, n, k, k);
        if (p != k)
            swap_rows(A, n, k, p, 0);
        eliminate_column_to(A, n, k, n);
        swaps[k] = p;
        ++steps;
    }
    return steps;
}

/* Dot product with the accumulator scheme of the packed substitutions
 * (Internal/Kernel.hs:2205-2256 and :2261-2314): two DoubleX4# FMA
 * accumulators over 8-wide steps, one more 4-wide step into the first,
 * lane-wise sum of the two, ((a+b)+c)+d, then a mul+add scalar tail. */
static double dot_sub_simd(const double *row, const double *x, ix start,
                           ix len)
{
    ix d8 = start + (len - len % 8);
    ix d4 = start + (len - len % 4);
    double a0[4] = {0, 0, 0, 0}, a1[4] = {0, 0, 0, 0};
    ix k = start;
    for (; k < d8; k += 8)
        for (int l = 0; l < 4; ++l) {
            a0[l] = fma(row[k + l], x[k + l], a0[l]);
            a1[l] = fma(row[k + 4 + l], x[k + 4 + l], a1[l]);
        }
    for (; k < d4; k += 4)
        for (int l = 0; l < 4; ++l)
            a0[l] = fma(row[k + l], x[k + l], a0[l]);
    double c[4];
    for (int l = 0; l < 4; ++l)
        c[l] = a0[l] + a1[l];
    double s = ((c[0] + c[1]) + c[2]) + c[3];
    for (k = d4; k < start + len; ++k) {
        double prod = row[k] * x[k];
        s = s + prod;
    }
    return s;
}

/* rawForwardSubUnitPackedSIMD (Internal/Kernel.hs:2205-2256): L y = x with
 * unit L in the strict lower triangle of LU; x overwritten. */
ORC_API void orc_forward_unit_packed(ix n, const double *LU, double *x)
{
    for (ix i = 0; i < n; ++i) {
        double d = dot_sub_simd(LU + i * n, x, 0, i);
        x[i] = x[i] - d;
    }
}

/* rawBackSubPackedSIMD (Internal/Kernel.hs:2261-2314): U x = y, i from n-1
 * down to 0, x_i = (x_i - dot) / u_ii. */
ORC_API void orc_back_packed(ix n, const double *LU, double *x)
{
    for (ix i = n - 1; i >= 0; --i) {
        double d = dot_sub_simd(LU + i * n, x, i + 1, n - i - 1);
        x[i] = (x[i] - d) / LU[i * n + i];
    }
}

/* applyPivotsForward (Solve/LU.hs:493-507): swap x[k], x[p] for the swap
 * list in ascending k. */
ORC_API void orc_apply_swaps(ix nsteps, const ix *swaps, double *x)
{
    for (ix k = 0; k < nsteps; ++k) {
        ix p = swaps[k];
        if (p != k) {
            double t = x[k];
            x[k] = x[p];
            x[p] = t;
        }
    }
}

/* Solve with an existing luSolveP factorisation (Solve/LU.hs:294-303). */
ORC_API void orc_lu_solve_factored(ix n, const double *LU, ix nsteps,
                                   const ix *swaps, double *x)
{
    orc_apply_swaps(nsteps, swaps, x);
    orc_forward_unit_packed(n, LU, x);
    orc_back_packed(n, LU, x);
}

/* luSolveP (Solve/LU.hs:265-304): factor a copy of A (panel path, nb = 32,
 * for n >= 64; unblocked below), then solve for b.  LU and swaps may be NULL;
 * when given they receive the packed factors and the swap list. */
ORC_API void orc_lusolvep(ix n, const double *A, const double *b, double *x,
                          double *LUout, ix *swapsOut)
{
    double *LU = (double *)malloc((size_t)(n * n) * sizeof(double));
    ix *sw = (ix *)calloc((size_t)(n > 0 ? n : 1), sizeof(ix));
    memcpy(LU, A, (size_t)(n * n) * sizeof(double));
    ix steps = (n >= 64) ? orc_lu_panel_factor(n, LU, 32, sw)
                         : orc_lu_unblocked_factor(n, LU, sw);
    memcpy(x, b, (size_t)n * sizeof(double));
    orc_lu_solve_factored(n, LU, steps, sw, x);
    if (LUout)
        memcpy(LUout, LU, (size_t)(n * n) * sizeof(double));
    if (swapsOut) {
        for (ix k = 0; k < n; ++k)
            swapsOut[k] = (k < steps) ? sw[k] : k;
    }
    free(LU);
    free(sw);
}

/* ------------------------------------------------------------------------ */
/* LU (generic path)                                                        */
/* ------------------------------------------------------------------------ */

/* Generic lu (Solve/LU.hs:121-167).  Pivot row = maximumBy (comparing abs)
 * over rows k..n-1: Data.Foldable.maximumBy replaces the running best x by
 * the next y unless compare x y == GT, so ties (and NaN comparisons that are
 * not GT) pick the later row.  piv is the permutation vector: row i of PA is
 * row piv[i] of A.  Elimination is skipped when a_kk == 0 (:152). */
ORC_API void orc_lu_generic(ix n, double *A, ix *piv)
{
    for (ix i = 0; i < n; ++i)
        piv[i] = i;
    for (ix k = 0; k + 1 < n; ++k) {
        ix bi = k;
        double bv = fabs(A[k * n + k]);
        for (ix i = k + 1; i < n; ++i) {
            double v = fabs(A[i * n + k]);
            /* compare bv v: LT if bv < v, EQ if bv == v, else GT */
            int gt = !(bv < v) && !(bv == v);
            if (!gt) {
                bi = i;
                bv = v;
            }
        }
        if (bi != k) {
            ix t = piv[k];
            piv[k] = piv[bi];
            piv[bi] = t;
            swap_rows(A, n, k, bi, 0);
        }
        double akk = A[k * n + k];
        if (akk != 0.0) {
            for (ix i = k + 1; i < n; ++i) {
                double mult = A[i * n + k] / akk;
                A[i * n + k] = mult;
                for (ix j = k + 1; j < n; ++j) {
                    double prod = mult * A[k * n + j];
                    A[i * n + j] = A[i * n + j] - prod;
                }
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Generic triangular solves (Solve/Triangular.hs), full n x n matrices     */
/* ------------------------------------------------------------------------ */

/* forwardSub (Triangular.hs:90-107): column-oriented, divide then update. */
ORC_API void orc_forward_sub(ix n, const double *L, const double *b, double *x)
{
    memcpy(x, b, (size_t)n * sizeof(double));
    for (ix j = 0; j < n; ++j) {
        double xj = x[j] / L[j * n + j];
        x[j] = xj;
        for (ix i = j + 1; i < n; ++i) {
            double prod = L[i * n + j] * xj;
            x[i] = x[i] - prod;
        }
    }
}

/* backSub (Triangular.hs:136-153). */
ORC_API void orc_back_sub(ix n, const double *U, const double *b, double *x)
{
    memcpy(x, b, (size_t)n * sizeof(double));
    for (ix j = n - 1; j >= 0; --j) {
        double xj = x[j] / U[j * n + j];
        x[j] = xj;
        for (ix i = 0; i < j; ++i) {
            double prod = U[i * n + j] * xj;
            x[i] = x[i] - prod;
        }
    }
}

/* forwardSubUnit (Triangular.hs:187-201): the diagonal is not read. */
ORC_API void orc_forward_sub_unit(ix n, const double *L, const double *b,
                                  double *x)
{
    memcpy(x, b, (size_t)n * sizeof(double));
    for (ix j = 0; j < n; ++j) {
        double xj = x[j];
        for (ix i = j + 1; i < n; ++i) {
            double prod = L[i * n + j] * xj;
            x[i] = x[i] - prod;
        }
    }
}

/* backSubUnit (Triangular.hs:235-249). */
ORC_API void orc_back_sub_unit(ix n, const double *U, const double *b,
                               double *x)
{
    memcpy(x, b, (size_t)n * sizeof(double));
    for (ix j = n - 1; j >= 0; --j) {
        double xj = x[j];
        for (ix i = 0; i < j; ++i) {
            double prod = U[i * n + j] * xj;
            x[i] = x[i] - prod;
        }
    }
}

/* Generic luSolve (Solve/LU.hs:240-259): lu, build unit L and U, pb[i] =
 * b[piv[i]], forwardSubUnit, backSub. */
ORC_API void orc_lusolve_generic(ix n, const double *A, const double *b,
                                 double *x)
{
    size_t nn = (size_t)(n * n);
    double *LU = (double *)malloc(nn * sizeof(double));
    double *L = (double *)calloc(nn ? nn : 1, sizeof(double));
    double *U = (double *)calloc(nn ? nn : 1, sizeof(double));
    double *pb = (double *)malloc((size_t)(n ? n : 1) * sizeof(double));
    double *y = (double *)malloc((size_t)(n ? n : 1) * sizeof(double));
    ix *piv = (ix *)malloc((size_t)(n ? n : 1) * sizeof(ix));
    memcpy(LU, A, nn * sizeof(double));
    orc_lu_generic(n, LU, piv);
    for (ix i = 0; i < n; ++i)
        for (ix j = 0; j < n; ++j) {
            if (i == j)
                L[i * n + j] = 1.0;
            else if (i > j)
                L[i * n + j] = LU[i * n + j];
            if (i <= j)
                U[i * n + j] = LU[i * n + j];
        }
    for (ix i = 0; i < n; ++i)
        pb[i] = b[piv[i]];
    orc_forward_sub_unit(n, L, pb, y);
    orc_back_sub(n, U, y, x);
    free(LU);
    free(L);
    free(U);
    free(pb);
    free(y);
    free(piv);
}

/* ------------------------------------------------------------------------ */
/* Cholesky                                                                 */
/* ------------------------------------------------------------------------ */

/* Generic cholesky (Solve/Cholesky.hs:114-147): G starts as the lower
 * triangle of A (upper zero); for each column j, subtract g_ik*g_jk for
 * k < j over rows i >= j (mul+add), sqrt the diagonal, divide below it. */
ORC_API void orc_cholesky_generic(ix n, const double *A, double *G)
{
    memset(G, 0, (size_t)(n * n) * sizeof(double));
    for (ix i = 0; i < n; ++i)
        for (ix j = 0; j <= i; ++j)
            G[i * n + j] = A[i * n + j];
    for (ix j = 0; j < n; ++j) {
        for (ix k = 0; k < j; ++k) {
            double gjk = G[j * n + k];
            for (ix i = j; i < n; ++i) {
                double prod = G[i * n + k] * gjk;
                G[i * n + j] = G[i * n + j] - prod;
            }
        }
        double sjj = sqrt(G[j * n + j]);
        G[j * n + j] = sjj;
        for (ix i = j + 1; i < n; ++i)
            G[i * n + j] = G[i * n + j] / sjj;
    }
}

/* Generic choleskySolve (Solve/Cholesky.hs:229-235): forwardSub with G then
 * backSub with G^T. */
ORC_API void orc_cholsolve_generic(ix n, const double *A, const double *b,
                                   double *x)
{
    size_t nn = (size_t)(n * n);
    double *G = (double *)malloc((nn ? nn : 1) * sizeof(double));
    double *Gt = (double *)malloc((nn ? nn : 1) * sizeof(double));
    double *y = (double *)malloc((size_t)(n ? n : 1) * sizeof(double));
    orc_cholesky_generic(n, A, G);
    for (ix i = 0; i < n; ++i)
        for (ix j = 0; j < n; ++j)
            Gt[j * n + i] = G[i * n + j];
    orc_forward_sub(n, G, b, y);
    orc_back_sub(n, Gt, y, x);
    free(G);
    free(Gt);
    free(y);
}

/* rawCholColumnSIMDFrom (Internal/Kernel.hs:2137-2195); fromCol = 0 is
 * rawCholColumnSIMD (:2071-2131).  For rows i = j .. n-1:
 * g_ij -= dot(G[i, fromCol:j], G[j, fromCol:j]) with one DoubleX4# FMA
 * accumulator over the 4-aligned span, ((a+b)+c)+d, then a mul+add tail;
 * then g_jj = sqrt(g_jj) and g_ij /= g_jj for i > j. */
static void chol_column_from(double *G, ix n, ix j, ix fromCol)
{
    const double *rj = G + j * n;
    ix len = j - fromCol;
    ix k4 = fromCol + (len - len % 4);
    for (ix i = j; i < n; ++i) {
        double *ri = G + i * n;
        double s;
        if (fromCol < k4) {
            double a[4] = {0, 0, 0, 0};
            for (ix k = fromCol; k < k4; k += 4)
                for (int l = 0; l < 4; ++l)
                    a[l] = fma(ri[k + l], rj[k + l], a[l]);
            s = ((a[0] + a[1]) + a[2]) + a[3];
            for (ix k = k4; k < j; ++k) {
                double prod = ri[k] * rj[k];
                s = s + prod;
            }
        } else {
            s = 0.0;
            for (ix k = fromCol; k < j; ++k) {
                double prod = ri[k] * rj[k];
                s = s + prod;
            }
        }
        ri[j] = ri[j] - s;
    }
    double sjj = sqrt(G[j * n + j]);
    G[j * n + j] = sjj;
    for (ix i = j + 1; i < n; ++i)
        G[i * n + j] = G[i * n + j] / sjj;
}

/* panelCholFactor (Solve/Cholesky.hs:279-321), left-looking, width nb:
 * for each panel [j, j+jb): C = G[j:n, 0:j] * G[j:j+jb, 0:j]^T formed by
 * rawGemmKernel into a zeroed buffer from dense copies (:296-312), G[j:n,
 * j:j+jb] -= C (:385-417), then rawCholColumnSIMDFrom c j for each panel
 * column c. */
static void chol_panel_factor(double *G, ix n, ix nb)
{
    double *Ab = NULL, *Bt = NULL, *C = NULL;
    for (ix j = 0; j < n;) {
        ix jb = min_ix(nb, n - j);
        ix nbelow = n - j;
        if (j > 0) {
            Ab = (double *)realloc(Ab, (size_t)(nbelow * j) * sizeof(double));
            Bt = (double *)realloc(Bt, (size_t)(j * jb) * sizeof(double));
            C = (double *)realloc(C, (size_t)(nbelow * jb) * sizeof(double));
            for (ix i = 0; i < nbelow; ++i)
                for (ix c = 0; c < j; ++c)
                    Ab[i * j + c] = G[(j + i) * n + c];
            for (ix i = 0; i < jb; ++i)
                for (ix c = 0; c < j; ++c)
                    Bt[c * jb + i] = G[(j + i) * n + c];
            memset(C, 0, (size_t)(nbelow * jb) * sizeof(double));
            gemm_acc_rawkernel(nbelow, j, jb, Ab, j, Bt, jb, C, jb);
            for (ix i = 0; i < nbelow; ++i)
                for (ix c = 0; c < jb; ++c) {
                    double *g = &G[(j + i) * n + j + c];
                    *g = *g - C[i * jb + c];
                }
        }
        for (ix c = j; c < j + jb; ++c)
            chol_column_from(G, n, c, j);
        j += jb;
    }
    free(Ab);
    free(Bt);
    free(C);
}

/* The factorisation inside choleskySolveP (Solve/Cholesky.hs:246-256):
 * zeroed G, lower triangle of A copied (copyLowerTriangle :420-434), panel
 * factor for n >= 64, column kernel per column below. */
ORC_API void orc_cholesky_p_factor(ix n, const double *A, double *G)
{
    memset(G, 0, (size_t)(n * n) * sizeof(double));
    for (ix i = 0; i < n; ++i)
        for (ix j = 0; j <= i; ++j)
            G[i * n + j] = A[i * n + j];
    if (n >= 64)
        chol_panel_factor(G, n, 32);
    else
        for (ix j = 0; j < n; ++j)
            chol_column_from(G, n, j, 0);
}

/* rawForwardSubCholPackedSIMD (Internal/Kernel.hs:2319-2371). */
ORC_API void orc_forward_chol_packed(ix n, const double *G, double *x)
{
    for (ix i = 0; i < n; ++i) {
        double d = dot_sub_simd(G + i * n, x, 0, i);
        x[i] = (x[i] - d) / G[i * n + i];
    }
}

/* rawBackSubCholTPackedSIMD (Internal/Kernel.hs:2378-2430): G^T x = y in
 * saxpy form: x_j /= g_jj, then x_i = fma(-x_j, g_ji, x_i) over the 8- and
 * 4-aligned spans of [0, j) and x_i - g_ji*x_j over the tail. */
ORC_API void orc_back_cholT_packed(ix n, const double *G, double *x)
{
    for (ix j = n - 1; j >= 0; --j) {
        double xj = x[j] / G[j * n + j];
        x[j] = xj;
        const double *rj = G + j * n;
        ix u4 = j - j % 4;
        for (ix i = 0; i < u4; ++i)
            x[i] = fma(-xj, rj[i], x[i]);
        for (ix i = u4; i < j; ++i) {
            double prod = rj[i] * xj;
            x[i] = x[i] - prod;
        }
    }
}

/* choleskySolveP (Solve/Cholesky.hs:241-273).  Gout may be NULL. */
ORC_API void orc_cholsolvep(ix n, const double *A, const double *b, double *x,
                            double *Gout)
{
    double *G = (double *)malloc((size_t)(n > 0 ? n * n : 1) * sizeof(double));
    orc_cholesky_p_factor(n, A, G);
    memcpy(x, b, (size_t)n * sizeof(double));
    orc_forward_chol_packed(n, G, x);
    orc_back_cholT_packed(n, G, x);
    if (Gout)
        memcpy(Gout, G, (size_t)(n * n) * sizeof(double));
    free(G);
}

/* Solve with an existing choleskySolveP factor (Solve/Cholesky.hs:258-272). */
ORC_API void orc_chol_solve_factored(ix n, const double *G, double *x)
{
    orc_forward_chol_packed(n, G, x);
    orc_back_cholT_packed(n, G, x);
}

/* ------------------------------------------------------------------------ */
/* Level 1/2, norms, det, inverse iteration                                 */
/* ------------------------------------------------------------------------ */

/* rawDot# (Internal/Kernel.hs:89-110): four FMA lanes over i = 0, 4, 8, ...,
 * reduced ((a + b) + c) + d, then the tail i >= 4 floor(n/4) with
 * acc + x * y (a separate multiply and add). */
ORC_API double orc_raw_dot(ix n, const double *x, const double *y)
{
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    ix n4 = n - n % 4, i;
    for (i = 0; i < n4; i += 4)
        for (int l = 0; l < 4; ++l)
            acc[l] = fma(x[i + l], y[i + l], acc[l]);
    double s = ((acc[0] + acc[1]) + acc[2]) + acc[3];
    for (i = n4; i < n; ++i) {
        double p = x[i] * y[i];
        s = s + p;
    }
    return s;
}

/* matvecP (BLAS/Level2.hs:132-140) through rawGemv (Internal/Kernel.hs:
 * 119-134): y_i = rawDot#(row i of A, x). */
ORC_API void orc_matvecp(ix m, ix n, const double *A, const double *x, double *y)
{
    for (ix i = 0; i < m; ++i)
        y[i] = orc_raw_dot(n, A + i * n, x);
}

/* The generic gemv (BLAS/Level2.hs:92-100): a_i . x as foldl (acc + a x)
 * from 0 in ascending j, then alpha * axi + beta * y_i. */
ORC_API void orc_gemv_generic(ix m, ix n, double alpha, const double *A, const double *x, double beta,
                              const double *y, double *out)
{
    for (ix i = 0; i < m; ++i) {
        double acc = 0.0;
        for (ix j = 0; j < n; ++j) {
            double p = A[i * n + j] * x[j];
            acc = acc + p;
        }
        double a = alpha * acc, b = beta * y[i];
        out[i] = a + b;
    }
}

/* norm1 (Norms.hs:124-130): column sums foldl (acc + |a_ij|) over i, the
 * maximum over columns; normInf (:142-148) the same over rows; normFrob
 * (:110-112) sqrt of foldl (acc + x * x) over the array in row-major order.
 * `maximum` of a list keeps its running maximum with max (which keeps the
 * first argument unless it is <= the second); NaN inputs are not restated. */
ORC_API double orc_norm1(ix m, ix n, const double *A)
{
    double best = 0.0;
    for (ix j = 0; j < n; ++j) {
        double s = 0.0;
        for (ix i = 0; i < m; ++i) s = s + fabs(A[i * n + j]);
        if (j == 0 || best <= s) best = s;
    }
    return best;
}

ORC_API double orc_norm_inf(ix m, ix n, const double *A)
{
    double best = 0.0;
    for (ix i = 0; i < m; ++i) {
        double s = 0.0;
        for (ix j = 0; j < n; ++j) s = s + fabs(A[i * n + j]);
        if (i == 0 || best <= s) best = s;
    }
    return best;
}

ORC_API double orc_norm_frob(ix m, ix n, const double *A)
{
    double s = 0.0;
    for (ix e = 0; e < m * n; ++e) {
        double p = A[e] * A[e];
        s = s + p;
    }
    return sqrt(s);
}

/* det (Solve/LU.hs:534-558): the generic lu, the product of U's diagonal as
 * foldl' (acc * u_ii) from 1, and (-1)^countSwaps of lu's permutation;
 * countSwaps moves each p into place by one swap, so its count is
 * n - (number of cycles), whose parity is the permutation's. */
ORC_API double orc_det(ix n, const double *A)
{
    double *LU = malloc(sizeof(double) * (size_t)(n > 0 ? n * n : 1));
    ix *piv = malloc(sizeof(ix) * (size_t)(n > 0 ? n : 1));
    memcpy(LU, A, sizeof(double) * (size_t)(n * n));
    orc_lu_generic(n, LU, piv);
    double prod = 1.0;
    for (ix i = 0; i < n; ++i) prod = prod * LU[i * n + i];
    /* countSwaps restated: walk the list, swapping p into place */
    ix nswaps = 0;
    for (ix i = 0; i < n; ++i) {
        if (piv[i] == i) continue;
        ix p = piv[i];
        for (ix t = i + 1; t < n; ++t)
            if (piv[t] == i) piv[t] = p;
        ++nswaps;
    }
    free(LU);
    free(piv);
    return (nswaps % 2 == 0) ? prod : -prod;
}

/* nrm2 (BLAS/Level1.hs:179-181): sqrt of foldl (acc + x * x). */
static double nrm2_seq(ix n, const double *x)
{
    double s = 0.0;
    for (ix i = 0; i < n; ++i) {
        double p = x[i] * x[i];
        s = s + p;
    }
    return sqrt(s);
}

/* inverseIteration (Eigen/Power.hs:113-135) at P Double: aShifted = A - mu I;
 * z = luSolve aShifted q (the generic lu, forwardSubUnit, backSub),
 * qNew = scal (1 / nrm2 z) z, lambda = dot qNew (matvec A qNew) with dot and
 * matvec at their P Double specialisations (rawDot#), stopping when
 * |lambda - prevLambda| < tol, or after max_iter steps with (prevLambda, q).
 * Returns the steps taken. */
ORC_API ix orc_inverse_iteration(ix n, const double *A, double mu, const double *q0, ix max_iter, double tol,
                                 double *lambda, double *q)
{
    double *S = malloc(sizeof(double) * (size_t)(n > 0 ? n * n : 1));
    double *z = malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    double *t = malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    for (ix i = 0; i < n; ++i)
        for (ix j = 0; j < n; ++j) S[i * n + j] = i == j ? A[i * n + j] - mu : A[i * n + j];
    memcpy(q, q0, sizeof(double) * (size_t)n);
    double prev = 0.0;
    ix it = 0;
    for (; it < max_iter; ++it) {
        orc_lusolve_generic(n, S, q, z);
        double zn = nrm2_seq(n, z), r = 1.0 / zn;
        for (ix i = 0; i < n; ++i) z[i] = z[i] * r;
        orc_matvecp(n, n, A, z, t);
        double lam = orc_raw_dot(n, z, t);
        memcpy(q, z, sizeof(double) * (size_t)n);
        if (fabs(lam - prev) < tol) {
            prev = lam;
            ++it;
            break;
        }
        prev = lam;
    }
    *lambda = prev;
    free(S);
    free(z);
    free(t);
    return it;
}

/* ------------------------------------------------------------------------ */
/* matMulPPar: the all-cores GEMM                                           */
/* ------------------------------------------------------------------------ */
#include <pthread.h>

/* gridDims (BLAS/Level3.hs:276-294): the factorisation pr x pc = p whose
 * aspect pr/pc is nearest m/n, the first best among i = 1..floor(sqrt p),
 * then its transpose if that scores strictly better. */
static void grid_dims(ix p, ix m, ix n, ix *pr_out, ix *pc_out)
{
    ix sq = (ix)floor(sqrt((double)p));
    double target = (double)m / (double)(n > 1 ? n : 1);
    ix bpr = 1, bpc = p;
    double bscore = -1.0;
    for (ix i = 1; i <= sq; ++i) {
        if (p % i) continue;
        ix pr = i, pc = p / i;
        double sc = fabs((double)pr / (double)(pc > 1 ? pc : 1) - target);
        if (bscore < 0.0 || bscore > sc) {
            bscore = sc;
            bpr = pr;
            bpc = pc;
        }
    }
    double tscore = fabs((double)bpc / (double)(bpr > 1 ? bpr : 1) - target);
    if (tscore < bscore) {
        ix t = bpr;
        bpr = bpc;
        bpc = t;
    }
    *pr_out = bpr;
    *pc_out = bpc;
}

struct par_slice {
    ix m, k, n, i0, i1, j0, j1;
    const double *A, *B;
    double *C;
};

static void *par_worker(void *arg)
{
    struct par_slice *s = arg;
    if (s->i1 > s->i0 && s->j1 > s->j0)
        gemm_acc_rawkernel(s->i1 - s->i0, s->k, s->j1 - s->j0, s->A + s->i0 * s->k, s->k, s->B + s->j0, s->n,
                           s->C + s->i0 * s->n + s->j0, s->n);
    return NULL;
}

/* matMulPPar (BLAS/Level3.hs:213-272) with `caps` threads: at least 16 rows
 * a thread; a pr x pc grid of (row, column) slices (gridDims) when there are
 * >= 4 threads and m, n >= 128, else row slices; each slice is the
 * reference's slice kernel, whose k tiles and column tails start at the
 * slice's own first column. */
ORC_API void orc_matmulp_par(ix m, ix k, ix n, const double *A, const double *B, double *C, ix caps)
{
    ix maxt = m / 16 > 1 ? m / 16 : 1;
    ix nt = caps < m ? caps : m;
    if (nt > maxt) nt = maxt;
    if (nt <= 1) {
        orc_matmulp(m, k, n, A, B, C);
        return;
    }
    memset(C, 0, (size_t)(m * n) * sizeof(double));
    ix pr = nt, pc = 1;
    if (nt >= 4 && m >= 128 && n >= 128) grid_dims(nt, m, n, &pr, &pc);
    ix rch = (m + pr - 1) / pr, cch = (n + pc - 1) / pc;
    struct par_slice *sl = malloc(sizeof(struct par_slice) * (size_t)(pr * pc));
    pthread_t *th = malloc(sizeof(pthread_t) * (size_t)(pr * pc));
    for (ix tr = 0; tr < pr; ++tr)
        for (ix tc = 0; tc < pc; ++tc) {
            struct par_slice *s = &sl[tr * pc + tc];
            s->m = m;
            s->k = k;
            s->n = n;
            s->i0 = tr * rch;
            s->i1 = min_ix(s->i0 + rch, m);
            s->j0 = tc * cch;
            s->j1 = min_ix(s->j0 + cch, n);
            s->A = A;
            s->B = B;
            s->C = C;
            pthread_create(&th[tr * pc + tc], NULL, par_worker, s);
        }
    for (ix t = 0; t < pr * pc; ++t) pthread_join(th[t], NULL);
    free(sl);
    free(th);
}
