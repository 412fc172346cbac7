// Resident factors: factor a matrix once, then solve any number of
// right-hand-side blocks against the factor kept in HBM.
//
// The reference's lu and cholesky (src/.../Solve/LU.hs:121, Cholesky.hs:114)
// return factors that its solvers never reuse: every luSolveP /
// choleskySolveP call refactors (LU.hs:265-304, Cholesky.hs:241-273).  A
// handle keeps, per precision:
//   LMX_PREC_F64:  the FP64 factor, its permutation and the inverses of its
//                  triangles' diagonal blocks, so a solve is the permutation
//                  and two block-inverse triangular solves;
//   LMX_PREC_F32 / LMX_PREC_BF16:  the low-precision factor with its block
//                  inverses (LowFactor) and an FP64 copy of A for the
//                  residuals of the refinement, so a solve is refinement to
//                  FP64 accuracy (falling back to an FP64 factorisation of
//                  the kept copy when it does not converge).
// Solves only read the handle, so several may run at once on different
// streams; release must not overlap them.
#include <algorithm>

#include "common.hpp"
#include "kernels.hpp"

struct lmx_factor_s {
    int kind = LMX_FACTOR_LU;
    int precision = LMX_PREC_F64;
    int device = 0;
    int64_t n = 0;
    int info = 0;
    size_t bytes = 0;
    void *mem = nullptr;
    // FP64 factor
    double *F = nullptr, *WL = nullptr, *WU = nullptr;
    int32_t *ipiv = nullptr, *perm = nullptr, *dinfo = nullptr;
    // low-precision factor
    double *A64 = nullptr;
    lmx::LowFactor low;
};

namespace lmx {
namespace {

constexpr int64_t F64_INVERSE_MIN_N = 64;  // smaller factors solve by substitution

size_t r256(size_t b) { return (b + 255) / 256 * 256; }

// Factors dA (n x n, stride lda, device) into a new handle on stream s.
lmx_factor_s *factor_create(hipStream_t s, int kind, int precision, int64_t n, const double *dA, int64_t lda) {
    if (kind != LMX_FACTOR_LU && kind != LMX_FACTOR_CHOLESKY) fail_arg("factor: kind must be LU or Cholesky");
    if (precision != LMX_PREC_F64 && precision != LMX_PREC_F32 && precision != LMX_PREC_BF16)
        fail_arg("factor: precision must be LMX_PREC_F64, LMX_PREC_F32 or LMX_PREC_BF16");
    if (n < 0 || lda < std::max<int64_t>(1, n)) fail_arg("factor: bad dimension or stride");
    auto *h = new lmx_factor_s();
    h->kind = kind;
    h->precision = precision;
    h->n = n;
    LMX_HIP(hipGetDevice(&h->device));
    const bool spd = kind == LMX_FACTOR_CHOLESKY;
    const size_t nn = (size_t)std::max<int64_t>(n, 1) * (size_t)std::max<int64_t>(n, 1);
    try {
        if (precision == LMX_PREC_F64) {
            const bool inv = n > F64_INVERSE_MIN_N;
            const size_t wb = inv ? r256((size_t)tri_inverse_elems(n) * sizeof(double)) : 0;
            const size_t fb = r256(nn * sizeof(double)), ib = r256((size_t)(2 * std::max<int64_t>(n, 1) + 8) * 4);
            h->bytes = fb + 2 * wb + ib;
            LMX_HIP(hipMalloc(&h->mem, h->bytes));
            char *p = static_cast<char *>(h->mem);
            h->F = reinterpret_cast<double *>(p);
            h->WL = inv ? reinterpret_cast<double *>(p + fb) : nullptr;
            h->WU = inv ? reinterpret_cast<double *>(p + fb + wb) : nullptr;
            h->ipiv = reinterpret_cast<int32_t *>(p + fb + 2 * wb);
            h->perm = h->ipiv + std::max<int64_t>(n, 1);
            h->dinfo = h->perm + std::max<int64_t>(n, 1);
            if (n > 0) copy_d(s, dA, lda, h->F, n, n, n);
            if (spd) {
                potrf<double>(s, n, h->F, std::max<int64_t>(n, 1), h->dinfo);
                if (inv) {
                    tri_inverse_build<double>(s, true, false, false, n, h->F, n, h->WL);
                    tri_inverse_build<double>(s, true, true, false, n, h->F, n, h->WU);
                }
            } else {
                getrf_impl<double>(s, n, h->F, std::max<int64_t>(n, 1), h->ipiv, h->perm, h->dinfo);
                if (inv) {
                    tri_inverse_build<double>(s, true, false, true, n, h->F, n, h->WL);
                    tri_inverse_build<double>(s, false, false, false, n, h->F, n, h->WU);
                }
            }
            int32_t info = 0;
            LMX_HIP(hipMemcpyAsync(&info, h->dinfo, sizeof(int32_t), hipMemcpyDeviceToHost, s));
            LMX_HIP(hipStreamSynchronize(s));
            if (info < 0) throw Error(LMX_ERR_INTERNAL, "panel synchronisation timed out");
            h->info = info;
        } else {
            const size_t ab = r256(nn * sizeof(double)), lb = low_factor_bytes(std::max<int64_t>(n, 1));
            h->bytes = ab + lb;
            LMX_HIP(hipMalloc(&h->mem, h->bytes));
            char *p = static_cast<char *>(h->mem);
            h->A64 = reinterpret_cast<double *>(p);
            if (n > 0) copy_d(s, dA, lda, h->A64, n, n, n);
            LowFactor &f = h->low;
            f.precision = precision;
            f.spd = spd;
            f.n = n;
            f.A = h->A64;
            f.lda = std::max<int64_t>(n, 1);
            low_factor_bind(f, p + ab);
            if (n > 0) low_factor(s, f);
            if (f.info < 0) throw Error(LMX_ERR_INTERNAL, "panel synchronisation timed out");
            h->info = f.info;
        }
    } catch (...) {
        if (h->mem) (void)hipFree(h->mem);
        delete h;
        throw;
    }
    return h;
}

RefineOut factor_solve(hipStream_t s, const lmx_factor_s *h, int64_t k, const double *dB, int64_t ldb, double *dX,
                       int64_t ldx) {
    if (k < 0 || ldb < std::max<int64_t>(1, k) || ldx < std::max<int64_t>(1, k)) fail_arg("factor_solve: bad nrhs or stride");
    const int64_t n = h->n;
    RefineOut out;
    out.info = h->info;
    if (n == 0 || k == 0) {
        out.converged = true;
        return out;
    }
    if (h->precision == LMX_PREC_F64) {
        copy_d(s, dB, ldb, dX, ldx, n, k);
        if (h->kind == LMX_FACTOR_CHOLESKY)
            potrs<double>(s, n, k, h->F, n, dX, ldx, h->WL, h->WU);
        else
            getrs_perm<double>(s, n, k, h->F, n, h->perm, dX, ldx, h->WL, h->WU);
        out.converged = true;
        return out;
    }
    return low_refine(s, h->low, k, dB, ldb, dX, ldx);
}

void check_handle(const lmx_factor_s *h) {
    if (!h) fail_arg("factor: null handle");
    int dev = 0;
    LMX_HIP(hipGetDevice(&dev));
    if (dev != h->device) fail_arg("factor: the handle belongs to another device");
}

void fill_info(lmx_refine_info *info, const RefineOut &o) {
    if (!info) return;
    info->iterations = o.iterations;
    info->converged = o.converged ? 1 : 0;
    info->fell_back = o.fell_back ? 1 : 0;
    info->info = o.info;
    info->backward_error = o.eta;
}

int status_of(const lmx_factor_s *h, const RefineOut &o) {
    if (h->precision == LMX_PREC_F64) return h->info;
    return o.converged ? (o.fell_back ? o.info : LMX_OK) : LMX_ERR_NOT_CONVERGED;
}

}  // namespace
}  // namespace lmx

using namespace lmx;

extern "C" {

int lmx_factor_create_dev(lmx_factor *out, void *stream, int kind, int precision, int64_t n, const double *dA,
                          int64_t lda) {
    return guarded([&] {
        if (!out) fail_arg("factor_create: null output");
        *out = nullptr;
        lmx_factor_s *h = factor_create(resolve_stream(stream), kind, precision, n, dA, lda);
        *out = h;
        return h->info;
    });
}

int lmx_factor_create(lmx_factor *out, int kind, int precision, int64_t n, const double *A, int64_t lda) {
    return guarded([&] {
        if (!out) fail_arg("factor_create: null output");
        *out = nullptr;
        if (n < 0 || lda < std::max<int64_t>(1, n)) fail_arg("factor_create: bad dimension or stride");
        Context &c = ctx();
        std::lock_guard<std::mutex> g(c.host_mu);
        const size_t ab = n > 0 ? (size_t)((n - 1) * lda + n) * sizeof(double) : 8;
        auto *dA = (double *)staging(c, 0, ab);
        if (n > 0) LMX_HIP(hipMemcpyAsync(dA, A, ab, hipMemcpyHostToDevice, c.stream));
        lmx_factor_s *h = factor_create(c.stream, kind, precision, n, dA, lda);
        *out = h;
        return h->info;
    });
}

int lmx_factor_solve_dev(lmx_factor h, void *stream, int64_t nrhs, const double *dB, int64_t ldb, double *dX,
                         int64_t ldx, lmx_refine_info *info) {
    return guarded([&] {
        check_handle(h);
        RefineOut o = factor_solve(resolve_stream(stream), h, nrhs, dB, ldb, dX, ldx);
        fill_info(info, o);
        if (o.info < 0) throw Error(LMX_ERR_INTERNAL, "panel synchronisation timed out");
        return status_of(h, o);
    });
}

int lmx_factor_solve(lmx_factor h, int64_t nrhs, const double *B, int64_t ldb, double *X, int64_t ldx,
                     lmx_refine_info *info) {
    return guarded([&] {
        check_handle(h);
        if (nrhs < 0 || ldb < std::max<int64_t>(1, nrhs) || ldx < std::max<int64_t>(1, nrhs))
            fail_arg("factor_solve: bad nrhs or stride");
        const int64_t n = h->n;
        if (n == 0 || nrhs == 0) {
            fill_info(info, RefineOut{0, true, false, h->info, 0.0});
            return h->precision == LMX_PREC_F64 ? h->info : LMX_OK;
        }
        Context &c = ctx();
        std::lock_guard<std::mutex> g(c.host_mu);
        const size_t bb = (size_t)((n - 1) * ldb + nrhs) * sizeof(double),
                     xb = (size_t)((n - 1) * ldx + nrhs) * sizeof(double);
        auto *dB = (double *)staging(c, 0, bb), *dX = (double *)staging(c, 1, xb);
        LMX_HIP(hipMemcpyAsync(dB, B, bb, hipMemcpyHostToDevice, c.stream));
        if (ldx > nrhs) LMX_HIP(hipMemcpyAsync(dX, X, xb, hipMemcpyHostToDevice, c.stream));
        RefineOut o = factor_solve(c.stream, h, nrhs, dB, ldb, dX, ldx);
        LMX_HIP(hipMemcpyAsync(X, dX, xb, hipMemcpyDeviceToHost, c.stream));
        LMX_HIP(hipStreamSynchronize(c.stream));
        fill_info(info, o);
        return status_of(h, o);
    });
}

int lmx_factor_query(lmx_factor h, int32_t *info, int64_t *n, int32_t *kind, int32_t *precision, int64_t *bytes) {
    return guarded([&] {
        if (!h) fail_arg("factor_query: null handle");
        if (info) *info = h->info;
        if (n) *n = h->n;
        if (kind) *kind = h->kind;
        if (precision) *precision = h->precision;
        if (bytes) *bytes = (int64_t)h->bytes;
        return LMX_OK;
    });
}

int lmx_factor_release(lmx_factor h) {
    return guarded([&] {
        if (!h) return LMX_OK;
        int cur = 0;
        LMX_HIP(hipGetDevice(&cur));
        LMX_HIP(hipSetDevice(h->device));
        // solves are stream-ordered on the callers' streams: wait for them
        LMX_HIP(hipDeviceSynchronize());
        if (h->mem) LMX_HIP(hipFree(h->mem));
        delete h;
        LMX_HIP(hipSetDevice(cur));
        return LMX_OK;
    });
}

}  // extern "C"
