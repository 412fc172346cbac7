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
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "device.hpp"
#include "kernels.hpp"

namespace lmx {

namespace {

constexpr int ITMAX = 30;
constexpr double EPS64 = 2.220446049250313e-16;

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



// ---------------------------------------------------------------- GMRES-IR
//
// Refinement with GMRES on the left-preconditioned system M^{-1} A d =
// M^{-1} r (M = the low-precision factor), for factors too coarse for plain
// refinement to contract (the BF16 factor at large kappa): Carson & Higham,
// SIAM J. Sci. Comput. 40(2), 2018.  Every right-hand-side column has its
// own Krylov space; the columns advance together, so each step is one FP64
// product A V_j (n x k), one factor solve of the n x k block and classical
// Gram-Schmidt with reorthogonalisation, all on the device.  The Hessenberg
// columns (2 (j + 1) k + k doubles per step) come back to the host for the
// Givens rotations.  Outer steps recompute r = b - A x in FP64 and test
// convergence with the same rule as plain refinement.
constexpr int GM_COLS = 64;      // right-hand sides per Krylov group
constexpr int GM_RESTART = 50;   // basis vectors per outer step
constexpr int GM_OUTER = 12;     // outer steps
constexpr double GM_INNER_TOL = 1e-8;
constexpr int GM_CHUNK = 1024;   // rows per partial dot product

// part[(i * nchunk + chunk) * k + c] = sum over the chunk's rows of
// V_i[r][c] W[r][c] (V_i = V + i vstride, row stride k); rows in a fixed
// order, then a fixed tree: the sums do not vary by run.
__global__ __launch_bounds__(256) void coldots_partial_kernel(const double *__restrict__ V, int64_t vstride,
                                                              const double *__restrict__ W, int64_t n, int k,
                                                              double *__restrict__ part) {
    __shared__ double red[4][64];
    const int c = threadIdx.x & 63, rsub = threadIdx.x >> 6;
    const int i = blockIdx.x, chunk = blockIdx.y;
    const double *v = V + (int64_t)i * vstride;
    const int64_t r0 = (int64_t)chunk * GM_CHUNK;
    double s = 0.0;
    if (c < k)
        for (int64_t r = r0 + rsub; r < n && r < r0 + GM_CHUNK; r += 4) s = fma(v[r * k + c], W[r * k + c], s);
    red[rsub][c] = s;
    __syncthreads();
    if (rsub == 0 && c < k)
        part[((int64_t)i * gridDim.y + chunk) * k + c] = (red[0][c] + red[1][c]) + (red[2][c] + red[3][c]);
}

// out[i * k + c] = sum over chunks (in order) of the partials.
__global__ void coldots_finish_kernel(const double *__restrict__ part, int nvec, int nchunk, int k,
                                      double *__restrict__ out) {
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= nvec * k) return;
    const int i = e / k, c = e % k;
    double s = 0.0;
    for (int q = 0; q < nchunk; ++q) s += part[((int64_t)i * nchunk + q) * k + c];
    out[e] = s;
}

// W[r][c] += sign * sum_i h[i][c] V_i[r][c] (W row stride ldw).
__global__ void block_axpy_kernel(double *__restrict__ W, int64_t ldw, const double *__restrict__ V,
                                  int64_t vstride, const double *__restrict__ h, int nvec, int64_t n, int k,
                                  double sign) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= n * k) return;
    const int64_t r = e / k;
    const int c = (int)(e % k);
    double acc = 0.0;
    for (int i = 0; i < nvec; ++i) acc = fma(h[i * k + c], V[i * vstride + e], acc);
    W[r * ldw + c] += sign * acc;
}

// dst[r][c] = src[r][c] / sqrt(nrm2sq[c]) (0 when the norm is 0).
__global__ void col_normalise_kernel(double *__restrict__ dst, const double *__restrict__ src,
                                     const double *__restrict__ nrm2sq, int64_t n, int k) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= n * k) return;
    const double d = sqrt(nrm2sq[e % k]);
    dst[e] = d > 0.0 ? src[e] / d : 0.0;
}

struct GmresBuffers {
    double *V, *W, *R, *part, *dots;
    float *R32;
};

void coldots(hipStream_t s, const double *V, int64_t vstride, int nvec, const double *W, int64_t n, int k,
             double *part, double *out) {
    const int nchunk = (int)cdiv(n, GM_CHUNK);
    hipLaunchKernelGGL(coldots_partial_kernel, dim3(nvec, nchunk), dim3(256), 0, s, V, vstride, W, n, k, part);
    hipLaunchKernelGGL(coldots_finish_kernel, dim3((unsigned)cdiv((int64_t)nvec * k, 256)), dim3(256), 0, s, part,
                       nvec, nchunk, k, out);
    LMX_LAUNCH_CHECK();
}

// Y (n x k, stride k) = M^{-1} Y with the low-precision factor, through the
// FP32 staging block.
void precondition(hipStream_t s, const LowFactor &f, int64_t k, double *Y, float *Y32) {
    d2f(s, Y, k, Y32, k, f.n, k);
    if (f.spd)
        potrs<float>(s, f.n, k, f.A32, f.n, Y32, k, f.WL, f.WU);
    else
        getrs_perm<float>(s, f.n, k, f.A32, f.n, f.perm, Y32, k, f.WL, f.WU);
    f2d_acc(s, Y32, k, Y, k, f.n, k, true);
}

// One group of k <= GM_COLS columns.  Returns the outer steps taken and sets
// *conv when every column met the criterion.
int gmres_group(hipStream_t s, const LowFactor &f, int k, const double *B, int64_t ldb, double *X, int64_t ldx,
                const GmresBuffers &g, double *norms, double crit, double *eta_out, bool *conv, int *inner_total) {
    const int64_t n = f.n, nk = n * k;
    const int K = k, M = GM_RESTART;
    std::vector<double> H((size_t)K * (M + 1) * M), cs((size_t)K * M), sn((size_t)K * M), gv((size_t)K * (M + 1)),
        beta(K), y((size_t)M), hbuf((size_t)(2 * (M + 1) + 1) * K);
    auto Hc = [&](int c, int i, int j) -> double & { return H[((size_t)c * (M + 1) + i) * M + j]; };
    zero_d(s, norms, 3 * K + 8);
    copy_d(s, f.anorm, 1, norms, 1, 1, 1);
    colmax(s, B, ldb, n, K, norms + 1);
    // X = 0
    zero_d(s, g.R, nk);
    copy_d(s, g.R, k, X, ldx, n, k);
    *conv = false;
    double prev = INFINITY;
    int outer = 0;
    for (; outer < GM_OUTER; ++outer) {
        residual(s, n, k, f.A, f.lda, X, ldx, B, ldb, g.R, k);
        if (outer > 0) {
            zero_d(s, norms + 1 + K, 2 * K + 2);
            colmax(s, X, ldx, n, K, norms + 1 + K);
            colmax(s, g.R, k, n, K, norms + 1 + 2 * K);
            hipLaunchKernelGGL(eta_kernel, dim3(1), dim3(64), 0, s, norms, K, crit);
            LMX_LAUNCH_CHECK();
            double h[2];
            LMX_HIP(hipMemcpyAsync(h, norms + 1 + 3 * K, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
            LMX_HIP(hipStreamSynchronize(s));
            *eta_out = h[0];
            if (h[1] == 1.0) {
                *conv = true;
                return outer;
            }
            if (!(h[0] < 0.9 * prev)) return outer;  // stagnation (or NaN)
            prev = h[0];
        }
        // v_0 = M^{-1} r / ||M^{-1} r||
        copy_d(s, g.R, k, g.V, k, n, k);
        precondition(s, f, k, g.V, g.R32);
        coldots(s, g.V, nk, 1, g.V, n, k, g.part, g.dots);
        LMX_HIP(hipMemcpyAsync(beta.data(), g.dots, K * sizeof(double), hipMemcpyDeviceToHost, s));
        hipLaunchKernelGGL(col_normalise_kernel, dim3((unsigned)cdiv(nk, 256)), dim3(256), 0, s, g.V, g.V, g.dots, n,
                           k);
        LMX_LAUNCH_CHECK();
        LMX_HIP(hipStreamSynchronize(s));
        for (int c = 0; c < K; ++c) {
            beta[c] = std::sqrt(beta[c]);
            std::fill(&gv[(size_t)c * (M + 1)], &gv[(size_t)c * (M + 1)] + M + 1, 0.0);
            gv[(size_t)c * (M + 1)] = beta[c];
        }
        int jdone = 0;
        for (int j = 0; j < M; ++j) {
            double *Vj = g.V + (int64_t)j * nk, *Vn = g.V + (int64_t)(j + 1) * nk;
            // W = M^{-1} A v_j
            matmul_block(s, n, n, k, f.A, f.lda, Vj, k, g.W, k);
            precondition(s, f, k, g.W, g.R32);
            // classical Gram-Schmidt, twice
            double *h1 = g.dots, *h2 = g.dots + (int64_t)(j + 1) * K, *hn = h2 + (int64_t)(j + 1) * K;
            coldots(s, g.V, nk, j + 1, g.W, n, k, g.part, h1);
            hipLaunchKernelGGL(block_axpy_kernel, dim3((unsigned)cdiv(nk, 256)), dim3(256), 0, s, g.W, k, g.V, nk, h1,
                               j + 1, n, k, -1.0);
            coldots(s, g.V, nk, j + 1, g.W, n, k, g.part, h2);
            hipLaunchKernelGGL(block_axpy_kernel, dim3((unsigned)cdiv(nk, 256)), dim3(256), 0, s, g.W, k, g.V, nk, h2,
                               j + 1, n, k, -1.0);
            coldots(s, g.W, nk, 1, g.W, n, k, g.part, hn);
            hipLaunchKernelGGL(col_normalise_kernel, dim3((unsigned)cdiv(nk, 256)), dim3(256), 0, s, Vn, g.W, hn, n,
                               k);
            LMX_LAUNCH_CHECK();
            const size_t cnt = (size_t)(2 * (j + 1) + 1) * K;
            LMX_HIP(hipMemcpyAsync(hbuf.data(), g.dots, cnt * sizeof(double), hipMemcpyDeviceToHost, s));
            LMX_HIP(hipStreamSynchronize(s));
            ++*inner_total;
            bool all = true;
            for (int c = 0; c < K; ++c) {
                for (int i = 0; i <= j; ++i)
                    Hc(c, i, j) = hbuf[(size_t)i * K + c] + hbuf[(size_t)(j + 1 + i) * K + c];
                Hc(c, j + 1, j) = std::sqrt(hbuf[(size_t)(2 * (j + 1)) * K + c]);
                double *csc = &cs[(size_t)c * M], *snc = &sn[(size_t)c * M], *gc = &gv[(size_t)c * (M + 1)];
                for (int i = 0; i < j; ++i) {
                    const double t = csc[i] * Hc(c, i, j) + snc[i] * Hc(c, i + 1, j);
                    Hc(c, i + 1, j) = -snc[i] * Hc(c, i, j) + csc[i] * Hc(c, i + 1, j);
                    Hc(c, i, j) = t;
                }
                const double d = std::hypot(Hc(c, j, j), Hc(c, j + 1, j));
                csc[j] = d > 0.0 ? Hc(c, j, j) / d : 1.0;
                snc[j] = d > 0.0 ? Hc(c, j + 1, j) / d : 0.0;
                Hc(c, j, j) = d;
                Hc(c, j + 1, j) = 0.0;
                gc[j + 1] = -snc[j] * gc[j];
                gc[j] = csc[j] * gc[j];
                if (!(std::fabs(gc[j + 1]) <= GM_INNER_TOL * beta[c])) all = false;
            }
            jdone = j + 1;
            if (all) break;
        }
        // x += V y, y = H^{-1} g per column (upper triangular, jdone x jdone)
        std::vector<double> ycol((size_t)jdone * K);
        for (int c = 0; c < K; ++c) {
            const double *gc = &gv[(size_t)c * (M + 1)];
            for (int i = jdone - 1; i >= 0; --i) {
                double t = gc[i];
                for (int l = i + 1; l < jdone; ++l) t -= Hc(c, i, l) * y[l];
                y[i] = Hc(c, i, i) != 0.0 ? t / Hc(c, i, i) : 0.0;
            }
            for (int i = 0; i < jdone; ++i) ycol[(size_t)i * K + c] = y[i];
        }
        LMX_HIP(hipMemcpyAsync(g.dots, ycol.data(), ycol.size() * sizeof(double), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(block_axpy_kernel, dim3((unsigned)cdiv(nk, 256)), dim3(256), 0, s, X, ldx, g.V, nk, g.dots,
                           jdone, n, k, 1.0);
        LMX_LAUNCH_CHECK();
        LMX_HIP(hipStreamSynchronize(s));  // ycol is a host buffer
    }
    // the last update's residual
    residual(s, n, k, f.A, f.lda, X, ldx, B, ldb, g.R, k);
    zero_d(s, norms + 1 + K, 2 * K + 2);
    colmax(s, X, ldx, n, K, norms + 1 + K);
    colmax(s, g.R, k, n, K, norms + 1 + 2 * K);
    hipLaunchKernelGGL(eta_kernel, dim3(1), dim3(64), 0, s, norms, K, crit);
    LMX_LAUNCH_CHECK();
    double h[2];
    LMX_HIP(hipMemcpyAsync(h, norms + 1 + 3 * K, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
    LMX_HIP(hipStreamSynchronize(s));
    *eta_out = h[0];
    *conv = h[1] == 1.0;
    return outer;
}

// GMRES-IR over the right-hand sides in groups of GM_COLS columns.
RefineOut gmres_refine(hipStream_t s, const LowFactor &f, int64_t k, const double *B, int64_t ldb, double *X,
                       int64_t ldx) {
    RefineOut out;
    out.info = f.info;
    const int64_t n = f.n;
    const int kg = (int)std::min<int64_t>(k, GM_COLS);
    const int64_t nk = n * kg, nchunk = cdiv(n, GM_CHUNK);
    Scratch sV((size_t)(GM_RESTART + 1) * nk * sizeof(double), s), sW((size_t)nk * sizeof(double), s),
        sR((size_t)nk * sizeof(double), s), sR32((size_t)nk * sizeof(float), s),
        sP((size_t)(GM_RESTART + 1) * nchunk * kg * sizeof(double), s),
        sD((size_t)(2 * (GM_RESTART + 1) + 1) * kg * sizeof(double), s),
        sN((size_t)(3 * kg + 8) * sizeof(double), s);
    GmresBuffers g{sV.as<double>(), sW.as<double>(), sR.as<double>(), sP.as<double>(), sD.as<double>(),
                   sR32.as<float>()};
    const double crit = std::sqrt((double)n) * EPS64;
    bool all = true;
    double eta = 0.0;
    int inner = 0, outer = 0;
    for (int64_t c0 = 0; c0 < k; c0 += GM_COLS) {
        const int kc = (int)std::min<int64_t>(GM_COLS, k - c0);
        bool conv = false;
        double e = 0.0;
        outer = std::max(outer, gmres_group(s, f, kc, B + c0, ldb, X + c0, ldx, g, sN.as<double>(), crit, &e, &conv,
                                            &inner));
        all = all && conv;
        if (!(e <= eta)) eta = e;
    }
    out.iterations = inner;
    out.eta = eta;
    out.converged = all;
    return out;
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

// LMX_REFINE=gmres refines FP32 factors with GMRES-IR too (BF16 factors
// always use it; read once per process).
bool gmres_requested() {
    static const bool on = [] {
        const char *e = std::getenv("LMX_REFINE");
        return e != nullptr && std::strcmp(e, "gmres") == 0;
    }();
    return on;
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
    d2f(s, f.A, f.lda, f.A32, n, n, n);
    const bool bf16 = f.precision == LMX_PREC_BF16;
    if (f.spd)
        potrf<float>(s, n, f.A32, n, f.dinfo, bf16);
    else
        getrf_impl<float>(s, n, f.A32, n, f.ipiv, f.perm, f.dinfo, (tournament_pivoting() || f.tournament) && !bf16,
                          false, bf16);
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
    zero_d(s, f.anorm, 1);
    norm_inf_acc(s, f.A, f.lda, n, n, f.anorm);
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
    if (f.info == 0 && (f.precision == LMX_PREC_BF16 || gmres_requested())) {
        RefineOut g = gmres_refine(s, f, k, B, ldb, X, ldx);
        out.iterations = g.iterations;
        out.eta = g.eta;
        out.converged = g.converged;
    } else if (f.info == 0) {
        zero_d(s, norms, 3 * K + 8);
        copy_d(s, f.anorm, 1, norms, 1, 1, 1);
        colmax(s, B, ldb, n, K, norms + 1);
        // R = B
        copy_d(s, B, ldb, R, k, n, k);
        double prev = INFINITY;
        for (int it = 0; it < ITMAX; ++it) {
            d2f(s, R, k, R32, k, n, k);
            if (f.spd)
                potrs<float>(s, n, k, f.A32, n, R32, k, f.WL, f.WU);
            else
                getrs_perm<float>(s, n, k, f.A32, n, f.perm, R32, k, f.WL, f.WU);
            f2d_acc(s, R32, k, X, ldx, n, k, it == 0);
            // R = B - A X
            residual(s, n, k, A, lda, X, ldx, B, ldb, R, k);
            zero_d(s, norms + 1 + K, 2 * K + 2);
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
        copy_d(s, A, lda, A64, n, n, n);
        copy_d(s, B, ldb, X, ldx, n, k);
        if (f.spd) {
            potrf<double>(s, n, A64, n, dinfo);
            potrs<double>(s, n, k, A64, n, X, ldx);
        } else {
            getrf_impl<double>(s, n, A64, n, ipiv, perm, dinfo);
            getrs_perm<double>(s, n, k, A64, n, perm, X, ldx);
        }
        // report the backward error of the FP64 solve
        residual(s, n, k, A, lda, X, ldx, B, ldb, R, k);
        zero_d(s, norms, 3 * K + 8);
        norm_inf_acc(s, A, lda, n, n, norms);
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

// precision: LMX_PREC_F32 (factor in FP32, plain refinement) or
// LMX_PREC_BF16 (FP32 storage, trailing updates on BF16 tiles with split
// operands, GMRES-IR).
RefineOut refine_solve(hipStream_t s, bool spd, int precision, int64_t n, int64_t k, const double *A,
                       int64_t lda, const double *B, int64_t ldb, double *X, int64_t ldx) {
    RefineOut out;
    if (precision != LMX_PREC_F32 && precision != LMX_PREC_BF16)
        fail_arg("mixed solve: factor precision must be LMX_PREC_F32 or LMX_PREC_BF16");
    if (n == 0 || k == 0) {
        out.converged = true;
        return out;
    }
    TimedRegion trall("refine", s);
    Scratch mem(low_factor_bytes(n), s);
    LowFactor f;
    f.precision = precision;
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
