/*
 * lmx.h -- C ABI of the MI355X (gfx950) dense-solve library.
 *
 * This is the drop-in boundary for the dense-solve path of linear-massiv
 * (Numeric.LinearAlgebra.Massiv): a Haskell `foreign import ccall` module
 * binds these symbols and keeps the library's public signatures unchanged.
 * Each entry point names the reference function it replaces (paths relative
 * to the reference repository).
 *
 * Conventions
 *   - Matrices are ROW-MAJOR, as massiv's P arrays are
 *     (src/Numeric/LinearAlgebra/Massiv/Types.hs:90); `ld*` is the row
 *     stride in elements (>= the column count).
 *   - Right-hand sides are an n x nrhs row-major block; nrhs = 1 with
 *     ldb = 1 is a massiv Vector.
 *   - Pivots use the LAPACK swap-list convention, 0-based: at step k rows k
 *     and ipiv[k] were exchanged.  This is the (k, pivRow) list that
 *     luSolveP builds (src/.../Solve/LU.hs:276-284, applied at :493-507);
 *     the permutation vector of the generic `lu` (src/.../Solve/LU.hs:121)
 *     follows by applying the swaps to 0..n-1 in ascending k.
 *   - Return value: LMX_OK (0); > 0 is a LAPACK-style `info` (1-based index
 *     of the first exactly-zero pivot, or of the first Cholesky column whose
 *     pivot is not positive) -- the computation still runs to the end and
 *     its Inf/NaN output is returned, as the reference does not stop either;
 *     < 0 is an error code (below).  lmx_last_error() describes the most
 *     recent error of the calling thread.
 *   - Host-pointer functions (no suffix) are synchronous: they copy to HBM,
 *     compute and copy back.  `_dev` functions take device pointers and a
 *     hipStream_t (passed as void*, NULL = HIP's default stream) and are
 *     stream-ordered; their `info` goes to a device int32 so no host
 *     synchronisation is needed.  A device `info` of -1 means the LU
 *     panel's cross-workgroup exchange did not complete within its bounded
 *     wait (its workgroups could not all be resident at once): the factors
 *     are then not valid.  Host functions report it as LMX_ERR_INTERNAL.
 *   - Every function is thread-safe.
 */
#ifndef LMX_H
#define LMX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LMX_OK 0
#define LMX_ERR_ARG (-1)          /* invalid argument (size, stride, flag) */
#define LMX_ERR_DEVICE (-2)       /* HIP runtime / device failure */
#define LMX_ERR_NOMEM (-3)        /* device memory exhausted */
#define LMX_ERR_NOT_CONVERGED (-4) /* refinement did not reach the target */
#define LMX_ERR_BUDGET (-5)       /* request larger than the HBM budget */
#define LMX_ERR_INTERNAL (-6)

/* Factor precisions for the mixed-precision solvers. */
#define LMX_PREC_F64 0
#define LMX_PREC_F32 1
#define LMX_PREC_BF16 2

/* ---- context ---------------------------------------------------------- */

/* Select and initialise a device (idempotent).  Replaces nothing in the
 * reference, which has no device. */
int lmx_init(int device);
/* Release device resources held by the library: its streams, its memory
 * pool and the freed scratch blocks it keeps per stream for reuse (up to
 * 8 GiB per device). */
int lmx_finalize(void);
/* Thread-local description of the last error (never NULL). */
const char *lmx_last_error(void);
/* Library version, major*10000 + minor*100 + patch. */
int lmx_version(void);

/* ---- GEMM --------------------------------------------------------------- */

/* C = alpha op(A) op(B) + beta C, op(X) = X ('N') or X^T ('T'); op(A) is
 * m x k, op(B) is k x n.  beta = 0 does not read C.
 * Replaces gemm (src/.../BLAS/Level3.hs:102-131) and, with alpha = 1,
 * beta = 0, matMul / matMulP / matMulPPar (:144-147, :183-207, :214-294). */
int lmx_dgemm(char transa, char transb, int64_t m, int64_t n, int64_t k,
              double alpha, const double *A, int64_t lda, const double *B,
              int64_t ldb, double beta, double *C, int64_t ldc);
int lmx_dgemm_dev(void *stream, char transa, char transb, int64_t m,
                  int64_t n, int64_t k, double alpha, const double *dA,
                  int64_t lda, const double *dB, int64_t ldb, double beta,
                  double *dC, int64_t ldc);
/* Single precision on FP32 MFMA; matMul at e = Float (the generic path
 * src/.../BLAS/Level3.hs:150-181). */
int lmx_sgemm_dev(void *stream, char transa, char transb, int64_t m,
                  int64_t n, int64_t k, float alpha, const float *dA,
                  int64_t lda, const float *dB, int64_t ldb, float beta,
                  float *dC, int64_t ldc);
int lmx_sgemm(char transa, char transb, int64_t m, int64_t n, int64_t k,
              float alpha, const float *A, int64_t lda, const float *B,
              int64_t ldb, float beta, float *C, int64_t ldc);

/* C (FP32) = alpha op(A) op(B) + beta C on the BF16 matrix cores
 * (v_mfma_f32_16x16x32_bf16) with FP32 accumulation; A and B hold BF16 bit
 * patterns (the upper half of an IEEE single).  matMul at a BF16 element
 * type: the reference multiplies any Num e through the generic path
 * (src/.../BLAS/Level3.hs:150-181); this is the BF16 tile the survey's GEMM
 * row asks for beside FP64 and FP32. */
int lmx_gemm_bf16(char transa, char transb, int64_t m, int64_t n, int64_t k,
                  float alpha, const uint16_t *A, int64_t lda,
                  const uint16_t *B, int64_t ldb, float beta, float *C,
                  int64_t ldc);
int lmx_gemm_bf16_dev(void *stream, char transa, char transb, int64_t m,
                      int64_t n, int64_t k, float alpha, const uint16_t *dA,
                      int64_t lda, const uint16_t *dB, int64_t ldb,
                      float beta, float *dC, int64_t ldc);
/* FP32 operands on the BF16 matrix cores: each element is split as
 * x = hi + lo (hi = bf16(x), lo = bf16(x - hi)) and every product is
 * hi*hi + hi*lo + lo*hi, accumulated in FP32 (relative error about 2^-16
 * per product).  This is the trailing-update GEMM of the BF16 factor of
 * the mixed-precision solvers (LMX_PREC_BF16). */
int lmx_sgemm_bf16x3_dev(void *stream, char transa, char transb, int64_t m,
                         int64_t n, int64_t k, float alpha, const float *dA,
                         int64_t lda, const float *dB, int64_t ldb,
                         float beta, float *dC, int64_t ldc);

/* The tile shape of the plain GEMMs issued by the calling thread: 0 = the
 * autotuned table (csrc/gemm_tiles.inc, written by tune_gemm.py; shapes it
 * does not cover take the fixed rule), 1 = 64 x 64, 2 = 128 x 64,
 * 3 = 128 x 128.  For the autotuner, which times each shape under every
 * tile; no reference counterpart (rawGemmKernel's dispatch,
 * src/.../Internal/Kernel.hs:145-159, picks its path by size alone). */
int lmx_gemm_tile_override(int tile);

/* ---- LU ----------------------------------------------------------------- */

/* In-place LU with partial pivoting of the n x n matrix A: unit L below the
 * diagonal, U on and above it; ipiv[k] = pivot row of step k (n entries).
 * The pivot of each column is its first row of largest magnitude, the rule
 * of rawPivotSearch (src/.../Internal/Kernel.hs:1541-1556).  Replaces the
 * factorisation of luSolveP (panelLUFactor, src/.../Solve/LU.hs:276-284,
 * :311-357). */
int lmx_dgetrf(int64_t n, double *A, int64_t lda, int32_t *ipiv);
int lmx_dgetrf_dev(void *stream, int64_t n, double *dA, int64_t lda,
                   int32_t *d_ipiv, int32_t *d_info);
/* The same factorisation in FP32: lu at e = Float (the reference's lu is
 * generic over Fractional e, src/.../Solve/LU.hs:121).  This is the factor
 * the mixed-precision solvers refine with. */
int lmx_sgetrf(int64_t n, float *A, int64_t lda, int32_t *ipiv);
int lmx_sgetrf_dev(void *stream, int64_t n, float *dA, int64_t lda,
                   int32_t *d_ipiv, int32_t *d_info);
/* The generic lu (src/.../Solve/LU.hs:121-167): A is overwritten by the
 * packed factors and perm (n entries) receives lu's permutation vector: row
 * i of PA is row perm[i] of A.  Pivots follow lu's maximumBy (:128-135): the
 * LAST row of largest |a_ik| wins a tie, a NaN is chosen only where it sits
 * on the diagonal, and a zero pivot leaves its column unscaled (:151-152).
 * The return value is the 1-based column of the first zero pivot, or 0.
 * lmx_slu is lu at e = Float. */
int lmx_dlu(int64_t n, double *A, int64_t lda, int32_t *perm);
int lmx_dlu_dev(void *stream, int64_t n, double *dA, int64_t lda,
                int32_t *d_perm, int32_t *d_info);
int lmx_slu(int64_t n, float *A, int64_t lda, int32_t *perm);
int lmx_slu_dev(void *stream, int64_t n, float *dA, int64_t lda,
                int32_t *d_perm, int32_t *d_info);
/* Solve A X = B with the factors of lmx_dgetrf; B (n x nrhs) is
 * overwritten by X.  Replaces the solve phase of luSolveP
 * (src/.../Solve/LU.hs:294-303: applyPivotsForward,
 * rawForwardSubUnitPackedSIMD, rawBackSubPackedSIMD) for nrhs columns. */
int lmx_dgetrs(int64_t n, int64_t nrhs, const double *LU, int64_t lda,
               const int32_t *ipiv, double *B, int64_t ldb);
int lmx_dgetrs_dev(void *stream, int64_t n, int64_t nrhs, const double *dLU,
                   int64_t lda, const int32_t *d_ipiv, double *dB,
                   int64_t ldb);
/* Factor and solve: A is overwritten by its factors, B by X.
 * Replaces luSolveP / luSolve (src/.../Solve/LU.hs:265-304, :240-259). */
int lmx_dgesv(int64_t n, int64_t nrhs, double *A, int64_t lda, int32_t *ipiv,
              double *B, int64_t ldb);
int lmx_dgesv_dev(void *stream, int64_t n, int64_t nrhs, double *dA,
                  int64_t lda, int32_t *d_ipiv, double *dB, int64_t ldb,
                  int32_t *d_info);

/* ---- Cholesky ----------------------------------------------------------- */

/* In-place lower Cholesky A = G G^T from the lower triangle only (GEMM
 * tiles on the diagonal may load the upper elements they cover, and never
 * use them); the strict upper triangle is left as given.  Replaces the factorisation of
 * choleskySolveP (src/.../Solve/Cholesky.hs:246-256, panelCholFactor
 * :279-321) and cholesky (:114-147, whose upper triangle is zero: the
 * binding zeroes it). */
int lmx_dpotrf(int64_t n, double *A, int64_t lda);
int lmx_dpotrf_dev(void *stream, int64_t n, double *dA, int64_t lda,
                   int32_t *d_info);
/* Solve G G^T X = B with the lower factor.  Replaces the solve phase of
 * choleskySolveP (src/.../Solve/Cholesky.hs:258-272). */
int lmx_dpotrs(int64_t n, int64_t nrhs, const double *G, int64_t lda,
               double *B, int64_t ldb);
int lmx_dpotrs_dev(void *stream, int64_t n, int64_t nrhs, const double *dG,
                   int64_t lda, double *dB, int64_t ldb);
/* Factor and solve.  Replaces choleskySolveP / choleskySolve
 * (src/.../Solve/Cholesky.hs:241-273, :229-235). */
int lmx_dposv(int64_t n, int64_t nrhs, double *A, int64_t lda, double *B,
              int64_t ldb);
int lmx_dposv_dev(void *stream, int64_t n, int64_t nrhs, double *dA,
                  int64_t lda, double *dB, int64_t ldb, int32_t *d_info);

/* ---- triangular solves -------------------------------------------------- */

/* Solve op(T) X = B in place (B is n x nrhs).  uplo 'L'/'U' selects the
 * triangle of T that is read, trans 'N'/'T' selects op, diag 'U' treats the
 * diagonal as ones without reading it, 'N' divides by it.
 * Replaces forwardSub (uplo L, diag N), backSub (U, N), forwardSubUnit
 * (L, U) and backSubUnit (U, U) of src/.../Solve/Triangular.hs:90-107,
 * :136-153, :187-201, :235-249, and the packed substitutions
 * src/.../Internal/Kernel.hs:2205-2430. */
int lmx_dtrsm(char uplo, char trans, char diag, int64_t n, int64_t nrhs,
              const double *T, int64_t ldt, double *B, int64_t ldb);
int lmx_dtrsm_dev(void *stream, char uplo, char trans, char diag, int64_t n,
                  int64_t nrhs, const double *dT, int64_t ldt, double *dB,
                  int64_t ldb);

/* ---- level-2 utilities, norms, SYRK, determinant, inverse iteration ---- */

/* y = alpha A x + beta y for an m x n row-major A (x and y unit stride).
 * Replaces gemv (src/.../BLAS/Level2.hs:92-100), matvec / matvecP
 * (:123-144, alpha = 1, beta = 0; rawGemv, Internal/Kernel.hs:119-134).
 * beta = 0 does not read y (BLAS convention; the reference's gemv forms
 * beta * y_i, so a NaN in y reaches its result only there). */
int lmx_dgemv(int64_t m, int64_t n, double alpha, const double *A, int64_t lda,
              const double *x, double beta, double *y);
int lmx_dgemv_dev(void *stream, int64_t m, int64_t n, double alpha,
                  const double *dA, int64_t lda, const double *dx, double beta,
                  double *dy);
/* A matrix norm of the m x n A into *out: which = '1' (max column sum,
 * norm1), 'I' (max row sum, normInf), 'F' (Frobenius, normFrob) or 'M'
 * (max |a_ij|).  Replaces src/.../Norms.hs norm1 (:124-130), normInf
 * (:142-148) and normFrob (:110-112); a vector is an n x 1 matrix (vnorm1
 * = '1', :72-74; vnorm2 = 'F', :83-85; vnormInf = 'I', :94-96).  A NaN
 * entry gives +Inf for '1', 'I' and 'M'. */
int lmx_dnorm(char which, int64_t m, int64_t n, const double *A, int64_t lda,
              double *out);
int lmx_dnorm_dev(void *stream, char which, int64_t m, int64_t n,
                  const double *dA, int64_t lda, double *d_out);
/* R = B - A X in FP64 (A n x n, X, B, R n x nrhs); the residual the
 * refinement loop and test/Test/Residuals.hs:69-78 form. */
int lmx_dresidual_dev(void *stream, int64_t n, int64_t nrhs, const double *dA,
                      int64_t lda, const double *dX, int64_t ldx,
                      const double *dB, int64_t ldb, double *dR, int64_t ldr);
/* eta[c] = ||A x_c - b_c||_inf / (||A||_inf ||x_c||_inf + ||b_c||_inf + n eps)
 * for each of the nrhs columns: scaledResidualLinear
 * (test/Test/Residuals.hs:69-78), the backward error the refinement stops
 * on.  nrhs <= 16 is formed without storing the residual. */
int lmx_dscaled_residual(int64_t n, int64_t nrhs, const double *A, int64_t lda,
                         const double *X, int64_t ldx, const double *B,
                         int64_t ldb, double *eta);
int lmx_dscaled_residual_dev(void *stream, int64_t n, int64_t nrhs,
                             const double *dA, int64_t lda, const double *dX,
                             int64_t ldx, const double *dB, int64_t ldb,
                             double *d_eta);
/* lmx_dscaled_residual_dev without any allocation: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops).  Its accumulators live in a
 * fixed device array of the library (64 doubles per GPU, kept until the
 * process ends), so nrhs is limited to 16 (larger values return
 * LMX_ERR_ARG).  Calls that overlap on one GPU share that array and race. */
int lmx_dscaled_residual_noalloc_dev(void *stream, int64_t n, int64_t nrhs,
                                     const double *dA, int64_t lda,
                                     const double *dX, int64_t ldx,
                                     const double *dB, int64_t ldb,
                                     double *d_eta);
/* C = alpha op(A) op(A)^T + beta C, the whole symmetric n x n C (its lower
 * triangle computed on the FP64 MFMA tiles, then mirrored).  trans 'N':
 * op(A) = A (n x k); 'T': op(A) = A^T with A stored k x n, so alpha = 1,
 * beta = 0 is matMulAtAP (src/.../BLAS/Level3.hs:353-373,
 * rawSyrkLowerKernel Internal/Kernel.hs:812-1058). */
int lmx_dsyrk(char trans, int64_t n, int64_t k, double alpha, const double *A,
              int64_t lda, double beta, double *C, int64_t ldc);
int lmx_dsyrk_dev(void *stream, char trans, int64_t n, int64_t k,
                  double alpha, const double *dA, int64_t lda, double beta,
                  double *dC, int64_t ldc);
/* det A (A not modified): lu's factors (its tie rule), the product of U's
 * diagonal formed left to right, and the sign of the permutation.  Replaces
 * det (src/.../Solve/LU.hs:534-558). */
int lmx_ddet(int64_t n, const double *A, int64_t lda, double *det);
int lmx_ddet_dev(void *stream, int64_t n, const double *dA, int64_t lda,
                 double *d_det);
/* Inverse iteration (src/.../Eigen/Power.hs:113-135): with q = q0 and
 * lambda_prev = 0, repeat z = (A - mu I)^{-1} q, q' = z / ||z||_2,
 * lambda = q'^T A q' until |lambda - lambda_prev| < tol (returning lambda,
 * q') or max_iter steps (returning the last lambda and q).  A - mu I is
 * factored once with lu's pivots; each step is a solve with the resident
 * factor, a norm, a GEMV and a dot product on the device, and lambda is read
 * back for the test.  *iterations receives the number of steps taken. */
int lmx_dinverse_iteration(int64_t n, const double *A, int64_t lda, double mu,
                           const double *q0, int32_t max_iter, double tol,
                           double *lambda, double *q, int32_t *iterations);
int lmx_dinverse_iteration_dev(void *stream, int64_t n, const double *dA,
                               int64_t lda, double mu, const double *dq0,
                               int32_t max_iter, double tol, double *lambda,
                               double *dq, int32_t *iterations);

/* ---- mixed precision with FP64 iterative refinement -------------------- */

/* Outcome of a refined solve. */
typedef struct lmx_refine_info {
    int32_t iterations;      /* refinement steps taken */
    int32_t converged;       /* 1 if the FP64 backward-error target was met */
    int32_t fell_back;       /* 1 if the solve fell back to an FP64 factor */
    int32_t info;            /* factorisation info of the final factor */
    double backward_error;   /* max over columns of ||b-Ax||_inf /
                                (||A||_inf ||x||_inf + ||b||_inf) */
} lmx_refine_info;

/* Solve A X = B to FP64 accuracy: factor a copy of A at `precision`
 * (LMX_PREC_F32 or LMX_PREC_BF16) on MFMA with FP32 accumulation, then
 * refine in FP64 (residual, correction, convergence test on the device)
 * until the normwise backward error is <= 10 n eps or stagnates; falls back
 * to the FP64 factorisation if it does not converge.  A and B are not
 * modified.  No reference counterpart: pinned by the FP64 oracle. */
int lmx_dgesv_mixed(int precision, int64_t n, int64_t nrhs, const double *A,
                    int64_t lda, const double *B, int64_t ldb, double *X,
                    int64_t ldx, lmx_refine_info *info);
int lmx_dgesv_mixed_dev(void *stream, int precision, int64_t n, int64_t nrhs,
                        const double *dA, int64_t lda, const double *dB,
                        int64_t ldb, double *dX, int64_t ldx,
                        lmx_refine_info *info);
int lmx_dposv_mixed(int precision, int64_t n, int64_t nrhs, const double *A,
                    int64_t lda, const double *B, int64_t ldb, double *X,
                    int64_t ldx, lmx_refine_info *info);
int lmx_dposv_mixed_dev(void *stream, int precision, int64_t n, int64_t nrhs,
                        const double *dA, int64_t lda, const double *dB,
                        int64_t ldb, double *dX, int64_t ldx,
                        lmx_refine_info *info);

/* ---- resident factors ---------------------------------------------------- */

/* A factor kept in HBM for repeated solves.  The reference's lu and
 * cholesky (src/.../Solve/LU.hs:121, Solve/Cholesky.hs:114) return factors
 * its solvers never reuse (luSolveP and choleskySolveP refactor on every
 * call, LU.hs:265-304, Cholesky.hs:241-273); a handle factors once.
 *   kind: LMX_FACTOR_LU (partial pivoting) or LMX_FACTOR_CHOLESKY (lower).
 *   precision: LMX_PREC_F64 (the FP64 factor; solves are the permutation and
 *     two triangular solves through the kept diagonal-block inverses) or
 *     LMX_PREC_F32 / LMX_PREC_BF16 (the low-precision factor plus an FP64
 *     copy of A; every solve is refined to FP64 accuracy as
 *     lmx_dgesv_mixed does, with the same fallback).
 * create returns the factorisation's info (> 0: singular / not positive
 * definite; the handle is still created) or an error; A is not modified.
 * Solves only read the handle: several may run at once (on different
 * streams); release waits for the device and must not overlap them. */
typedef struct lmx_factor_s *lmx_factor;
#define LMX_FACTOR_LU 0
#define LMX_FACTOR_CHOLESKY 1
int lmx_factor_create(lmx_factor *h, int kind, int precision, int64_t n,
                      const double *A, int64_t lda);
int lmx_factor_create_dev(lmx_factor *h, void *stream, int kind,
                          int precision, int64_t n, const double *dA,
                          int64_t lda);
/* X (n x nrhs) = A^{-1} B for the handle's A.  Returns 0, the factor's info
 * (FP64 handles), LMX_ERR_NOT_CONVERGED, or an error; `info` (may be NULL)
 * receives the refinement outcome (iterations 0 for FP64 handles). */
int lmx_factor_solve(lmx_factor h, int64_t nrhs, const double *B, int64_t ldb,
                     double *X, int64_t ldx, lmx_refine_info *info);
int lmx_factor_solve_dev(lmx_factor h, void *stream, int64_t nrhs,
                         const double *dB, int64_t ldb, double *dX,
                         int64_t ldx, lmx_refine_info *info);
int lmx_factor_query(lmx_factor h, int32_t *info, int64_t *n, int32_t *kind,
                     int32_t *precision, int64_t *bytes);
int lmx_factor_release(lmx_factor h);

/* ---- batched solves of many small systems ------------------------------- */

/* No reference counterpart (the reference solves one system per call, with
 * luSolveP / choleskySolveP, src/.../Solve/LU.hs:265-304 and
 * Solve/Cholesky.hs:241-273): these entries run thousands of independent
 * solves of uneven sizes in one persistent launch, each to the reference's
 * result within the stated tolerance. */
#define LMX_BATCH_MAX_N 512
#define LMX_BATCH_MAX_NRHS 8
#define LMX_BATCH_LU 0
#define LMX_BATCH_CHOLESKY 1
/* plan flags */
#define LMX_BATCH_FACTORS 1   /* FP64 LU: leave the factors in LAPACK layout
                                 in A (else A holds them with rows in their
                                 original order) */
#define LMX_BATCH_GRAPH 2     /* capture the launch in a hipGraph and replay
                                 it while the pool pointers are unchanged */
#define LMX_BATCH_AUTOTUNE 4  /* time the kernel variants on synthetic
                                 problems of the plan's shapes at creation
                                 and keep the fastest */

/* One problem of a batch.  A is n x n row-major (lda = n) at
 * a_pool + a_offset; B is n x nrhs row-major (ldb = nrhs) at
 * b_pool + b_offset and receives X; an LU problem's swap list (0-based,
 * LAPACK order) goes to ipiv_pool + ipiv_offset when an ipiv pool is given.
 * Cholesky reads only the lower triangle of A. */
typedef struct lmx_batch_problem {
    int32_t kind;       /* LMX_BATCH_LU or LMX_BATCH_CHOLESKY */
    int32_t n;          /* 1 .. LMX_BATCH_MAX_N */
    int32_t nrhs;       /* 1 .. LMX_BATCH_MAX_NRHS */
    int32_t reserved;   /* ignored */
    int64_t a_offset;   /* elements */
    int64_t b_offset;   /* elements */
    int64_t ipiv_offset;
} lmx_batch_problem;

typedef struct lmx_batch_plan lmx_batch_plan;

/* Validates the problems against the pool sizes (in elements; the ipiv pool
 * may be 0 = no pivots returned), orders them largest first, sizes the
 * per-workgroup workspace and picks the kernel variant.  precision:
 * LMX_PREC_F64 (FP64 factorisations), or LMX_PREC_F32 (FP32 factorisation
 * on MFMA, FP64 refinement as LAPACK dsgesv, FP64 fallback; A is left
 * unchanged unless a problem falls back). */
int lmx_batch_plan_create(lmx_batch_plan **plan, const lmx_batch_problem *problems,
                          int32_t count, int64_t a_pool_elems,
                          int64_t b_pool_elems, int64_t ipiv_pool_elems,
                          int32_t precision, int32_t flags);
/* Runs every problem of the plan on device pools (stream-ordered).  info
 * (count entries, device) gets 0, or the LAPACK info of a singular / not
 * positive definite problem (its B is then unchanged); iterations (count
 * entries, device, may be NULL) gets the refinement steps of a mixed
 * solve, or -1 where it fell back to FP64. */
int lmx_batch_plan_execute(lmx_batch_plan *plan, double *a_pool, double *b_pool,
                           int32_t *ipiv_pool, int32_t *info,
                           int32_t *iterations, void *stream);
/* Re-runs the autotuner (synchronous). */
int lmx_batch_plan_autotune(lmx_batch_plan *plan, void *stream);
/* The chosen variant: block-column width, threads per workgroup, number of
 * persistent workgroups, the batch's FP64-equivalent flops, and (4
 * entries) the autotuner's time per variant in ms (0 if not tuned).
 * Pointers may be NULL. */
int lmx_batch_plan_query(const lmx_batch_plan *plan, int32_t *variant,
                         int32_t *block_cols, int32_t *threads,
                         int32_t *workgroups, double *flops, double *tuned_ms);
int lmx_batch_plan_destroy(lmx_batch_plan *plan);

/* ---- solve server ------------------------------------------------------- */

/* No reference counterpart: a request loop over the mixed-precision solvers
 * for callers that issue many independent solves (the reference's
 * luSolveP / choleskySolveP, one call each).  Requests are admitted as
 * worker slots and HBM (within the budget) free up, run concurrently on
 * separate streams, small ones grouped into one batched launch; a factor
 * can be kept resident and reused for new right-hand sides. */
typedef struct lmx_server_config {
    int64_t budget_bytes; /* HBM the server may use; 0 = 80% of free memory */
    int32_t slots;        /* concurrent large requests (1..16; 0 = 4) */
    int32_t small_n;      /* requests with n <= small_n (<= 512) and at most
                             8 right-hand sides go through the batched kernel
                             in groups; 0 = never */
    int32_t small_max;    /* most requests in one batched launch (0 = 256) */
    int32_t reserved;
    double aging_ms;      /* a queued request older than this is not
                             overtaken by smaller later ones (0 = 100 ms) */
} lmx_server_config;

#define LMX_REQ_KEEP_FACTOR 1 /* keep the factor resident; its handle is
                                 returned in lmx_request_result.factor */

typedef struct lmx_request_result {
    int32_t status;        /* 1 = done, 0 = still queued or running */
    int32_t rc;            /* LMX_OK, LAPACK info > 0, or an error < 0 */
    int32_t iterations;    /* refinement steps */
    int32_t fell_back;     /* 1 if it was solved by an FP64 factorisation */
    double backward_error; /* from the refinement (large requests) */
    int64_t factor;        /* resident factor handle (KEEP_FACTOR), else 0 */
    double queue_ms;       /* submit -> admitted */
    double run_ms;         /* admitted -> done */
    double latency_ms;     /* submit -> done */
} lmx_request_result;

typedef struct lmx_server_counters {
    int64_t completed, failed, queued, inflight;
    int64_t bytes_in_use, peak_bytes, budget_bytes, resident_factors;
    double busy_ms;   /* union of the requests' device intervals */
    double window_ms; /* wall time of the same window */
} lmx_server_counters;

typedef struct lmx_server lmx_server;

int lmx_server_create(lmx_server **server, const lmx_server_config *config);
/* Queues one solve A X = B (kind LMX_BATCH_LU or LMX_BATCH_CHOLESKY, host
 * row-major buffers that must stay valid until the request is done).
 * Either A (with lda) or a resident factor handle of the same kind and n is
 * given.  Fails with LMX_ERR_BUDGET if the request can never fit. */
int lmx_server_submit(lmx_server *server, int32_t kind, int64_t n, int64_t nrhs,
                      const double *A, int64_t lda, int64_t factor,
                      const double *B, int64_t ldb, double *X, int64_t ldx,
                      int32_t flags, int64_t *request_id);
/* Blocks until the request is done, fills the result and forgets it. */
int lmx_server_wait(lmx_server *server, int64_t request_id, lmx_request_result *result);
/* LMX_OK and the result if done (then forgets it), 1 if not yet. */
int lmx_server_poll(lmx_server *server, int64_t request_id, lmx_request_result *result);
/* Frees a resident factor once the requests using it are done. */
int lmx_server_factor_release(lmx_server *server, int64_t factor);
/* Counters and the device-busy fraction (busy_ms / window_ms) since the
 * last reset; reset != 0 starts a new window. */
int lmx_server_stats(lmx_server *server, lmx_server_counters *stats, int32_t reset);
/* Finishes running requests, fails queued ones, frees everything. */
int lmx_server_destroy(lmx_server *server);

/* ---- measurement -------------------------------------------------------- */

/* When enabled, the library brackets every launch of the named kernel
 * family with HIP events on the launch stream and accumulates the elapsed
 * time and the algorithmic work (flops) of each launch.  Kernel families:
 * "gemm" / "sgemm" (FP64 / FP32 MFMA GEMM, 128 x 128 tiles), "gemm64" /
 * "sgemm64" (the same with 64 x 64 tiles, and the batched GEMMs), "syrk" /
 * "ssyrk" (FP64 / FP32),
 * "getrf_panel", "trsm", "laswp", "potrf_diag", "batch"; whole-call
 * families: "getrf", "potrf", "getrs", "refine".  No reference
 * counterpart. */
int lmx_kernel_timing(int enable);
int lmx_kernel_timing_reset(void);
/* Count of timed launches, their total milliseconds and total work for one
 * family (synchronises on the recorded events).  Pointers may be NULL. */
int lmx_kernel_timing_get(const char *family, int64_t *count,
                          double *total_ms, double *total_work);

#ifdef __cplusplus
}
#endif

#endif /* LMX_H */
