// Internal (C++) entry points of the kernels, shared between the solver
// drivers.  All matrices are row-major with an explicit row stride; all
// launches are asynchronous on the given stream.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "lmx.h"

namespace lmx {

// C = alpha op(A) op(B) + beta C on MFMA (FP64: v_mfma_f64_16x16x4_f64,
// FP32: v_mfma_f32_16x16x4_f32).  beta == 0 never reads C.
void gemm(hipStream_t s, bool ta, bool tb, int64_t m, int64_t n, int64_t k, double alpha,
          const double *A, int64_t lda, const double *B, int64_t ldb, double beta, double *C,
          int64_t ldc);
void gemm(hipStream_t s, bool ta, bool tb, int64_t m, int64_t n, int64_t k, float alpha,
          const float *A, int64_t lda, const float *B, int64_t ldb, float beta, float *C,
          int64_t ldc);

// C = alpha op(A) op(B) + beta C for FP32 C on BF16 matrix cores
// (v_mfma_f32_16x16x32_bf16, FP32 accumulation):
//   gemm_bf16x3: FP32 operands split into hi = bf16(x) and lo = bf16(x - hi),
//     products hi*hi + hi*lo + lo*hi (about 16 significant bits); with
//     `lower`, C is square and only its lower triangle's tiles are computed;
//   gemm_bf16: BF16 operands (bit patterns) as stored;
//   gemm_bf16_rounded: FP32 operands rounded once to BF16.
void gemm_bf16x3(hipStream_t s, bool ta, bool tb, int64_t m, int64_t n, int64_t k, float alpha, const float *A,
                 int64_t lda, const float *B, int64_t ldb, float beta, float *C, int64_t ldc, bool lower = false,
                 size_t lds_footprint = 0);
void gemm_bf16(hipStream_t s, bool ta, bool tb, int64_t m, int64_t n, int64_t k, float alpha, const uint16_t *A,
               int64_t lda, const uint16_t *B, int64_t ldb, float beta, float *C, int64_t ldc);
void gemm_bf16_rounded(hipStream_t s, bool ta, bool tb, int64_t m, int64_t n, int64_t k, float alpha, const float *A,
                       int64_t lda, const float *B, int64_t ldb, float beta, float *C, int64_t ldc);

// C = alpha A B + beta C with each workgroup holding lds_footprint bytes of
// LDS in all (its tiles plus an unused reserve): fewer GEMM workgroups share
// a CU, which leaves room for the workgroups of a kernel running beside it
// on another stream.
template <class T>
void gemm_sharing(hipStream_t s, int64_t m, int64_t n, int64_t k, T alpha, const T *A, int64_t lda, const T *B,
                  int64_t ldb, T beta, T *C, int64_t ldc, size_t lds_footprint);

// Batched C_z = alpha A_z B_z + beta C_z (no transposes) for z = zk nb1 + zp,
// zk < nb2, zp < nb1; entry z's matrices start at zk x2 + zp x1 elements
// from A, B, C.
template <class T>
void gemm_batched(hipStream_t s, int64_t m, int64_t n, int64_t k, T alpha, const T *A, int64_t lda, const T *B,
                  int64_t ldb, T beta, T *C, int64_t ldc, int nb1, int nb2, int64_t a1, int64_t b1, int64_t c1,
                  int64_t a2, int64_t b2, int64_t c2);

// Lower triangle of C = alpha A A^T + beta C (A n x k); tiles strictly above
// the diagonal are not launched and elements above it are not written.
void syrk_lower(hipStream_t s, int64_t n, int64_t k, double alpha, const double *A, int64_t lda,
                double beta, double *C, int64_t ldc);
void syrk_lower(hipStream_t s, int64_t n, int64_t k, float alpha, const float *A, int64_t lda, float beta,
                float *C, int64_t ldc);
// C (m x n, m >= n) = alpha A A[0:n]^T + beta C on the elements with row >=
// column (the lower trapezoid), one launch.
// syrk_lower with each workgroup of the 64 x 64 tiles holding lds_footprint
// bytes of LDS (its tiles plus a reserve), so that fewer share a CU.
template <class T>
void syrk_lower_sharing(hipStream_t s, int64_t n, int64_t k, T alpha, const T *A, int64_t lda, T beta, T *C,
                        int64_t ldc, size_t lds_footprint);
template <class T>
void syrk_trap(hipStream_t s, int64_t m, int64_t n, int64_t k, T alpha, const T *A, int64_t lda, T beta, T *C,
               int64_t ldc);

// Lower triangle of C = alpha op(A) op(A)^T + beta C: op(A) = A (n x k) or,
// with `trans`, A^T (A stored k x n, so C = A^T A).
void syrk_lower_op(hipStream_t s, bool trans, int64_t n, int64_t k, double alpha, const double *A, int64_t lda,
                   double beta, double *C, int64_t ldc);
// The same, then the lower triangle mirrored into the upper one (a full
// symmetric C, as matMulAtAP returns).
void syrk_full(hipStream_t s, bool trans, int64_t n, int64_t k, double alpha, const double *A, int64_t lda,
               double beta, double *C, int64_t ldc);

// ---- level-2 work (blas2.hip)
// R = B - A X in FP64 (A n x n, X and B n x k).
void residual(hipStream_t s, int64_t n, int64_t k, const double *A, int64_t lda, const double *X, int64_t ldx,
              const double *B, int64_t ldb, double *R, int64_t ldr);
// colmax_out[c] = max(colmax_out[c], max_i |(B - A X)[i][c]|), A m x n,
// without storing the residual.
void residual_colmax(hipStream_t s, int64_t m, int64_t n, int64_t k, const double *A, int64_t lda, const double *X,
                     int64_t ldx, const double *B, int64_t ldb, double *colmax_out);
// Y = A X in FP64 (A m x n, X n x k, Y m x k): the GEMV kernel for k <= 4,
// the MFMA GEMM otherwise.
void matmul_block(hipStream_t s, int64_t m, int64_t n, int64_t k, const double *A, int64_t lda, const double *X,
                  int64_t ldx, double *Y, int64_t ldy);
// y = alpha A x + beta y (A m x n; beta = 0 does not read y).
void gemv(hipStream_t s, int64_t m, int64_t n, double alpha, const double *A, int64_t lda, const double *x,
          double beta, double *y);
// out[c] = max(out[c], max_i |X[i][c]|) (NaN counts as +inf).
void colmax(hipStream_t s, const double *X, int64_t ldx, int64_t n, int ncols, double *out);
// *out = max(*out, ||A||_inf) for A m x n.
void norm_inf_acc(hipStream_t s, const double *A, int64_t lda, int64_t m, int64_t n, double *out);
// *d_out = the norm `which` of A: '1' (max column sum), 'I' (max row sum),
// 'F' (Frobenius), 'M' (max |a_ij|).
void matrix_norm(hipStream_t s, char which, int64_t m, int64_t n, const double *A, int64_t lda, double *d_out);
// d_eta[c] = scaledResidualLinear of column c; acc holds 1 + 3k doubles of
// device memory and R n x k doubles (read only when k > SCALED_RESIDUAL_FUSED_MAX).
constexpr int64_t SCALED_RESIDUAL_FUSED_MAX = 16;
void scaled_residual(hipStream_t s, int64_t n, int64_t k, const double *A, int64_t lda, const double *X, int64_t ldx,
                     const double *B, int64_t ldb, double *d_eta, double *acc, double *R);
// *d_det = det A from lu's factors (A is not modified).
void determinant(hipStream_t s, int64_t n, const double *A, int64_t lda, double *d_det);
void d2f(hipStream_t s, const double *in, int64_t ldi, float *out, int64_t ldo, int64_t rows, int64_t cols);
// X += D (or X = D with `assign`), D in FP32.
void f2d_acc(hipStream_t s, const float *D, int64_t ldd, double *X, int64_t ldx, int64_t rows, int64_t cols,
             bool assign);
void copy_d(hipStream_t s, const double *in, int64_t ldi, double *out, int64_t ldo, int64_t rows, int64_t cols);
void zero_d(hipStream_t s, double *p, int64_t n);

// Triangular solve with many right-hand sides, in place on B (n x nrhs):
// op(T) X = B, T read from the `lower` or upper triangle, `unit` diagonal
// not read.
template <class T>
void trsm_left(hipStream_t s, bool lower, bool trans, bool unit, int64_t n, int64_t nrhs,
               const T *A, int64_t lda, T *B, int64_t ldb);
// X op(T) = B in place on B (nrows x n).
template <class T>
void trsm_right(hipStream_t s, bool lower, bool trans, bool unit, int64_t nrows, int64_t n,
                const T *A, int64_t lda, T *B, int64_t ldb);

// Explicit inverses of the diagonal blocks of op(T) (blocks of order
// tri_block(n), stored densely in tri_inverse_elems(n) elements of W), for
// repeated solves with tri_inverse_solve (same lower / trans flags; B n x
// nrhs, in place).
int tri_block(int64_t n);
int64_t tri_inverse_elems(int64_t n);
template <class T>
void tri_inverse_build(hipStream_t s, bool lower, bool trans, bool unit, int64_t n, const T *A, int64_t lda, T *W);
template <class T>
void tri_inverse_solve(hipStream_t s, bool lower, bool trans, int64_t n, int64_t nrhs, const T *A, int64_t lda,
                       const T *W, T *B, int64_t ldb);

// LU with partial pivoting, in place; writes d_ipiv[0..n) and *d_info.
template <class T>
void getrf(hipStream_t s, int64_t n, T *A, int64_t lda, int32_t *d_ipiv, int32_t *d_info);
// getrf that also writes the permutation vector (row i of PA is row
// d_perm[i] of A) when d_perm is not null.  `tournament` chooses each
// panel's pivots by a tournament over the panel's row blocks (CALU) instead
// of column-by-column partial pivoting: one exchange per panel.
// `tie_last` breaks ties between equal pivot magnitudes towards the highest
// row (the generic lu's maximumBy) instead of the lowest (rawPivotSearch).
// `bf16` (FP32 factors only) runs the trailing updates on the BF16 tiles with
// split operands (gemm_bf16x3).
template <class T>
void getrf_impl(hipStream_t s, int64_t n, T *A, int64_t lda, int32_t *d_ipiv, int32_t *d_perm,
                int32_t *d_info, bool tournament = false, bool tie_last = false, bool bf16 = false);
// Solve with getrf's factors and its permutation vector; WL / WU, when not
// null, are tri_inverse_build's inverses of L (unit lower) and U.
template <class T>
void getrs_perm(hipStream_t s, int64_t n, int64_t nrhs, const T *LU, int64_t lda,
                const int32_t *d_perm, T *B, int64_t ldb, const T *WL = nullptr, const T *WU = nullptr);
// Solve with getrf's factors and swap list.
template <class T>
void getrs(hipStream_t s, int64_t n, int64_t nrhs, const T *LU, int64_t lda,
           const int32_t *d_ipiv, T *B, int64_t ldb);

// Lower Cholesky, in place; writes *d_info.  `bf16` (FP32 only): the SYRK
// and GEMM updates on the BF16 tiles with split operands.
template <class T>
void potrf(hipStream_t s, int64_t n, T *A, int64_t lda, int32_t *d_info, bool bf16 = false);
// WL / WU, when not null, are tri_inverse_build's inverses of L and L^T.
template <class T>
void potrs(hipStream_t s, int64_t n, int64_t nrhs, const T *L, int64_t lda, T *B, int64_t ldb,
           const T *WL = nullptr, const T *WU = nullptr);

// Set *d_info = 0.
void clear_info(hipStream_t s, int32_t *d_info);

// ---- mixed precision: an FP32 factor of an FP64 matrix, and FP64
// refinement with it (refine.hip).
struct RefineOut {
    int iterations = 0;
    bool converged = false;
    bool fell_back = false;
    int info = 0;
    double eta = 0.0;
};
struct LowFactor {
    int precision = LMX_PREC_F32;  // LMX_PREC_F32 or LMX_PREC_BF16 (FP32 storage, BF16 updates)
    bool spd = false;
    bool tournament = false;       // LU: tournament pivoting for this factor
    int64_t n = 0;
    const double *A = nullptr;  // the FP64 matrix (residuals); must outlive the factor's use
    int64_t lda = 0;
    float *A32 = nullptr;       // n x n FP32 factor
    float *WL = nullptr, *WU = nullptr;  // diagonal-block inverses of the two triangles (n > 256)
    double *anorm = nullptr;    // ||A||_inf (device)
    int32_t *ipiv = nullptr, *perm = nullptr, *dinfo = nullptr;
    int info = 0;
};
// Device bytes of a factor of order n (low_factor_bind carves them).
size_t low_factor_bytes(int64_t n);
void low_factor_bind(LowFactor &f, void *mem);
// Factors f.A into f's buffers (synchronises the stream for info).
void low_factor(hipStream_t s, LowFactor &f);
// Refines X for B (n x k) with the factor; falls back to an FP64
// factorisation of f.A when refinement does not converge.
RefineOut low_refine(hipStream_t s, const LowFactor &f, int64_t k, const double *B, int64_t ldb, double *X,
                     int64_t ldx);
RefineOut refine_solve(hipStream_t s, bool spd, int precision, int64_t n, int64_t k, const double *A,
                       int64_t lda, const double *B, int64_t ldb, double *X, int64_t ldx);

// ---- batched solves (batch.hip): re-targets a plan at a new list of
// problems, reusing its device buffers (problem table copied on stream s,
// or synchronously when s is null).
void batch_replan(lmx_batch_plan *pl, const lmx_batch_problem *problems, int32_t count, int64_t a_pool_elems,
                  int64_t b_pool_elems, int64_t ipiv_pool_elems, hipStream_t s);
// Workspace bytes one persistent workgroup needs for problems up to nmax.
size_t batch_slot_bytes(int nmax, int precision, int flags);

}  // namespace lmx
