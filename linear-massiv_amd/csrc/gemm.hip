// Row-major GEMM on CDNA4 matrix cores.
//
// C = alpha op(A) op(B) + beta C.  One workgroup of 256 threads (4 wave64s
// in a 2 x 2 grid) computes a BM x BN tile of C; each wave owns a
// (BM/2) x (BN/2) sub-tile made of 16 x 16 MFMA blocks held in accumulator
// registers.  K advances in steps of BK through two LDS buffers: the next
// step's global loads are issued into registers before the current step's
// MFMAs and written to the other buffer after them, one barrier per step.
//
// The NN case (both operands as stored) with 16-byte aligned rows has its
// own kernel, gemm_nn_kernel, described there.
//
// MFMA 16x16x4 operands (f64 and f32 alike): lane l supplies
// A[i = l%16][kk = l/16] and B[kk = l/16][j = l%16], so both operand tiles
// live in LDS k-major (As[kk][i], Bs[kk][j]) and the 16 lanes of a
// quarter-wave read 16 consecutive elements.  The four accumulator values of
// lane l sit in column l%16 at rows 4*(l/16) + r for f32 but at rows
// l/16 + 4*r for f64 (measured on gfx950); acc_row() encodes both.
#include "common.hpp"
#include "kernels.hpp"

namespace lmx {

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef double d2 __attribute__((ext_vector_type(2)));

// 16-byte vectors of T (the width of one global or LDS access)
template <class T>
struct V16;
template <>
struct V16<float> {
    typedef f4 type;
    static constexpr int N = 4;
};
template <>
struct V16<double> {
    typedef d2 type;
    static constexpr int N = 2;
};

template <class T>
struct Acc;
template <>
struct Acc<double> {
    typedef d4 type;
};
template <>
struct Acc<float> {
    typedef f4 type;
};

__device__ inline d4 mfma16x16x4(double a, double b, d4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ inline f4 mfma16x16x4(float a, float b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// Maps the linear workgroup id to a C tile.  Workgroups are dispatched
// round-robin over the 8 XCDs; remapping gives each XCD a contiguous range
// of ids, and the grouped order inside a range keeps GROUP_M tile-rows
// sharing B panels in that XCD's L2.
// With `lower` the grid enumerates only the tiles on or below the diagonal
// of a lower-trapezoidal tile grid (tn <= tm, tiles_n <= tiles_m): the
// triangle of the first tiles_n tile rows (in groups of 8 tile rows), then
// the full rows below it.
__host__ __device__ inline int lower_tiles(int tiles_m, int tiles_n) {
    return tiles_n * (tiles_n + 1) / 2 + (tiles_m - tiles_n) * tiles_n;
}

__device__ inline void tile_of(int tiles_m, int tiles_n, bool lower, int &tm, int &tn) {
    const int nwg = lower ? lower_tiles(tiles_m, tiles_n) : tiles_m * tiles_n;
    // workgroup b runs on XCD b mod 8: XCD x takes the x-th contiguous range
    // of the enumeration (ranges of q + 1 tiles for the first nwg mod 8 XCDs,
    // of q for the others)
    const int q = nwg >> 3, r = nwg & 7, x = (int)blockIdx.x & 7, j = (int)blockIdx.x >> 3;
    const int bid = x < r ? x * (q + 1) + j : r * (q + 1) + (x - r) * q + j;
    if (lower) {
        const int tri = tiles_n * (tiles_n + 1) / 2;
        if (bid >= tri) {
            tm = tiles_n + (bid - tri) / tiles_n;
            tn = (bid - tri) % tiles_n;
            return;
        }
        // the triangle in groups of 8 tile rows (the groups in row order, a
        // group column by column), so that the workgroups of an XCD share
        // their A and B panels in its L2 as the full grid's GROUP_M order
        // does; row by row, an FP64 SYRK fetched 2.5 times its bytes
        int r = (int)((sqrtf(8.0f * (float)bid + 1.0f) - 1.0f) * 0.5f);
        while (r * (r + 1) / 2 > bid) --r;
        while ((r + 1) * (r + 2) / 2 <= bid) ++r;
        const int g0 = r & ~7, rows = min(8, tiles_n - g0);
        int l = bid - g0 * (g0 + 1) / 2;
        if (l < g0 * rows) {
            tm = g0 + l % rows;
            tn = l / rows;
            return;
        }
        l -= g0 * rows;
        int c = 0;
        while (l >= rows - c) {
            l -= rows - c;
            ++c;
        }
        tm = g0 + c + l;
        tn = g0 + c;
        return;
    }
    const int GROUP_M = 8;
    int per_group = GROUP_M * tiles_n;
    int g = bid / per_group;
    int first_m = g * GROUP_M;
    int gsize = min(tiles_m - first_m, GROUP_M);
    int in_g = bid - g * per_group;
    tm = first_m + in_g % gsize;
    tn = in_g / gsize;
}

// Pointer offsets of a batched launch: entry (zk, zp) starts at
// zk * x2 + zp * x1 elements from the base pointers.
struct BatchStrides {
    int nb1;
    int64_t a1, b1, c1, a2, b2, c2;
};

// Row inside a 16 x 16 accumulator block of value r held by lane group lk.
template <class T>
__device__ inline int acc_row(int lk, int r) {
    return sizeof(T) == 8 ? lk + 4 * r : 4 * lk + r;
}

template <class T, int BM, int BN, int BK, bool TA, bool TB>
__global__ __launch_bounds__(256) void gemm_mfma_kernel(int m, int n, int k, T alpha,
                                                        const T *__restrict__ A, int64_t lda,
                                                        const T *__restrict__ B, int64_t ldb,
                                                        T beta, T *__restrict__ C, int64_t ldc,
                                                        int tiles_m, int tiles_n, int lower, BatchStrides bs) {
    constexpr int PAD = (sizeof(T) == 8) ? 1 : 2;
    constexpr int WM = BM / 2, WN = BN / 2;
    constexpr int MT = WM / 16, NT = WN / 16;
    constexpr int LA = BM * BK / 256;  // A elements loaded per thread per step
    constexpr int LB = BN * BK / 256;
    static_assert(BM * BK % 256 == 0 && BN * BK % 256 == 0 && BK % 4 == 0, "tile shape");
    typedef typename Acc<T>::type V;

    __shared__ T As[2][BK][BM + PAD];
    __shared__ T Bs[2][BK][BN + PAD];

    if (gridDim.y > 1) {
        // batch entry blockIdx.y = zk * nb1 + zp
        const int zk = (int)blockIdx.y / bs.nb1, zp = (int)blockIdx.y - zk * bs.nb1;
        A += zk * bs.a2 + zp * bs.a1;
        B += zk * bs.b2 + zp * bs.b1;
        C += zk * bs.c2 + zp * bs.c1;
    }
    int tm, tn;
    tile_of(tiles_m, tiles_n, lower != 0, tm, tn);
    const int m0 = tm * BM, n0 = tn * BN;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm0 = (wave >> 1) * WM, wn0 = (wave & 1) * WN;
    const int lr = lane & 15, lk = lane >> 4;

    V acc[MT][NT];
#pragma unroll
    for (int a = 0; a < MT; ++a)
#pragma unroll
        for (int b = 0; b < NT; ++b) acc[a][b] = V{0, 0, 0, 0};

    T ra[LA], rb[LB];

    auto load = [&](int k0) {
#pragma unroll
        for (int it = 0; it < LA; ++it) {
            int e = it * 256 + tid;
            int i, kk;
            if (TA) {
                i = e % BM;
                kk = e / BM;
            } else {
                kk = e % BK;
                i = e / BK;
            }
            int gi = m0 + i, gk = k0 + kk;
            T v = T(0);
            if (gi < m && gk < k) v = TA ? A[(int64_t)gk * lda + gi] : A[(int64_t)gi * lda + gk];
            ra[it] = v;
        }
#pragma unroll
        for (int it = 0; it < LB; ++it) {
            int e = it * 256 + tid;
            int j, kk;
            if (TB) {
                kk = e % BK;
                j = e / BK;
            } else {
                j = e % BN;
                kk = e / BN;
            }
            int gj = n0 + j, gk = k0 + kk;
            T v = T(0);
            if (gj < n && gk < k) v = TB ? B[(int64_t)gj * ldb + gk] : B[(int64_t)gk * ldb + gj];
            rb[it] = v;
        }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int it = 0; it < LA; ++it) {
            int e = it * 256 + tid;
            int i, kk;
            if (TA) {
                i = e % BM;
                kk = e / BM;
            } else {
                kk = e % BK;
                i = e / BK;
            }
            As[buf][kk][i] = ra[it];
        }
#pragma unroll
        for (int it = 0; it < LB; ++it) {
            int e = it * 256 + tid;
            int j, kk;
            if (TB) {
                kk = e % BK;
                j = e / BK;
            } else {
                j = e % BN;
                kk = e / BN;
            }
            Bs[buf][kk][j] = rb[it];
        }
    };

    // C values read before the k loop (beta != 0), as in the NN kernel: all
    // of a 64 x 64 tile's, the first half of the rows of an FP32 128 x 128
    // tile's (32 more registers keep it within two waves a SIMD)
    constexpr int PH = (BM == 64 && BN == 64) ? MT : ((sizeof(T) == 4 && BM == 128 && BN == 128) ? MT / 2 : 0);
    T cpre[PH > 0 ? PH : 1][NT][4];
    if constexpr (PH > 0) {
        if (beta != T(0)) {
#pragma unroll
            for (int a = 0; a < PH; ++a)
#pragma unroll
                for (int b = 0; b < NT; ++b)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int gi = min(m0 + wm0 + a * 16 + acc_row<T>(lk, r), m - 1);
                        const int gj = min(n0 + wn0 + b * 16 + lr, n - 1);
                        cpre[a][b][r] = C[(int64_t)gi * ldc + gj];
                    }
        }
    }

    const int KT = (k + BK - 1) / BK;
    if (KT > 0) {
        load(0);
        store(0);
        __syncthreads();
    }
    for (int kt = 0; kt < KT; ++kt) {
        const int cur = kt & 1;
        const bool more = kt + 1 < KT;
        if (more) load((kt + 1) * BK);
#pragma unroll
        for (int ks = 0; ks < BK; ks += 4) {
            T fa[MT], fb[NT];
#pragma unroll
            for (int a = 0; a < MT; ++a) fa[a] = As[cur][ks + lk][wm0 + a * 16 + lr];
#pragma unroll
            for (int b = 0; b < NT; ++b) fb[b] = Bs[cur][ks + lk][wn0 + b * 16 + lr];
#pragma unroll
            for (int a = 0; a < MT; ++a)
#pragma unroll
                for (int b = 0; b < NT; ++b) acc[a][b] = mfma16x16x4(fa[a], fb[b], acc[a][b]);
        }
        if (more) store(cur ^ 1);
        __syncthreads();
    }

    // Epilogue: lane holds C[acc_row(lk, r)][lr] of each 16 x 16 block.
#pragma unroll
    for (int a = 0; a < MT; ++a)
#pragma unroll
        for (int b = 0; b < NT; ++b) {
            int gj = n0 + wn0 + b * 16 + lr;
            if (gj >= n) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                int gi = m0 + wm0 + a * 16 + acc_row<T>(lk, r);
                if (gi < m && (!lower || gj <= gi)) {
                    T *c = C + (int64_t)gi * ldc + gj;
                    T v = alpha * acc[a][b][r];
                    if (beta != T(0)) v += beta * (a < PH ? cpre[a < PH ? a : 0][b][r] : *c);
                    *c = v;
                }
            }
        }
}

// C = alpha A B + beta C with A (m x k) and B (k x n) as stored, every row
// 16-byte aligned and k, n multiples of the 16-byte vector width.
//   Global loads are 16-byte vectors.  A tiles go to LDS as they are
//   (As[i][k]); B tiles are transposed on the way in (Bs[j][k]), the lanes of
//   a wave on consecutive k so that the transposing writes do not collide.
//   MFMA step s of a 16-k chunk takes k = 4 (lane / 16) + s from lane group
//   lane / 16 (the same products as steps of consecutive k, summed in
//   another order), so a lane's operands for the whole chunk are four
//   consecutive elements of one LDS row: one 128-bit read per 16 x 16 block
//   for FP32 (two for FP64) instead of four scalar reads.  The row stride
//   (20 floats, 18 doubles) puts the 128-bit reads of 8 consecutive rows on
//   distinct banks.
//   With TB, B is stored transposed (n x k, so op(B) = B^T, as in A A^T):
//   its k runs along its rows too, and its tiles load and store like A's.
//   With `lower`, C is lower-trapezoidal (square for a SYRK) and only the
//   tiles on or below its diagonal are launched, and only elements on or
//   below it are written.
template <class T, int BM, int BN, bool TB>
__global__ __launch_bounds__(256) void gemm_nn_kernel(int m, int n, int k, T alpha, const T *__restrict__ A,
                                                      int64_t lda, const T *__restrict__ B, int64_t ldb, T beta,
                                                      T *__restrict__ C, int64_t ldc, int tiles_m, int tiles_n,
                                                      BatchStrides bs, int lower) {
    constexpr int BK = 16;
    constexpr int SK = sizeof(T) == 4 ? BK + 4 : BK + 2;
    typedef typename V16<T>::type VT;
    constexpr int E = V16<T>::N;
    constexpr int WM = BM / 2, WN = BN / 2;
    constexpr int MT = WM / 16, NT = WN / 16;
    constexpr int LA = BM * BK / (E * 256);  // A vectors loaded per thread per step
    constexpr int LB = BN * BK / (E * 256);
    static_assert(LA >= 1 && LB >= 1 && BM * BK % (E * 256) == 0 && BN * BK % (E * 256) == 0, "tile shape");
    typedef typename Acc<T>::type V;

    __shared__ __attribute__((aligned(16))) T As[2][BM][SK];
    __shared__ __attribute__((aligned(16))) T Bs[2][BN][SK];

    if (gridDim.y > 1) {
        const int zk = (int)blockIdx.y / bs.nb1, zp = (int)blockIdx.y - zk * bs.nb1;
        A += zk * bs.a2 + zp * bs.a1;
        B += zk * bs.b2 + zp * bs.b1;
        C += zk * bs.c2 + zp * bs.c1;
    }
    int tm, tn;
    tile_of(tiles_m, tiles_n, lower != 0, tm, tn);
    const int m0 = tm * BM, n0 = tn * BN;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm0 = (wave >> 1) * WM, wn0 = (wave & 1) * WN;
    const int lr = lane & 15, lk = lane >> 4;

    V acc[MT][NT];
#pragma unroll
    for (int a = 0; a < MT; ++a)
#pragma unroll
        for (int b = 0; b < NT; ++b) acc[a][b] = V{0, 0, 0, 0};

    VT ra[LA], rb[LB];
    auto load = [&](int k0) {
#pragma unroll
        for (int it = 0; it < LA; ++it) {
            const int e = it * 256 + tid;
            const int i = e / (BK / E), kq = e % (BK / E);
            const int gi = m0 + i, gk = k0 + kq * E;
            ra[it] = (gi < m && gk < k) ? *reinterpret_cast<const VT *>(A + (int64_t)gi * lda + gk) : VT{};
        }
#pragma unroll
        for (int it = 0; it < LB; ++it) {
            const int e = it * 256 + tid;
            if (TB) {
                const int j = e / (BK / E), kq = e % (BK / E);
                const int gj = n0 + j, gk = k0 + kq * E;
                rb[it] = (gj < n && gk < k) ? *reinterpret_cast<const VT *>(B + (int64_t)gj * ldb + gk) : VT{};
            } else {
                const int kk = e % BK, jq = e / BK;
                const int gj = n0 + jq * E, gk = k0 + kk;
                rb[it] = (gk < k && gj < n) ? *reinterpret_cast<const VT *>(B + (int64_t)gk * ldb + gj) : VT{};
            }
        }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int it = 0; it < LA; ++it) {
            const int e = it * 256 + tid;
            const int i = e / (BK / E), kq = e % (BK / E);
            *reinterpret_cast<VT *>(&As[buf][i][kq * E]) = ra[it];
        }
#pragma unroll
        for (int it = 0; it < LB; ++it) {
            const int e = it * 256 + tid;
            if (TB) {
                const int j = e / (BK / E), kq = e % (BK / E);
                *reinterpret_cast<VT *>(&Bs[buf][j][kq * E]) = rb[it];
            } else {
                const int kk = e % BK, jq = e / BK;
#pragma unroll
                for (int q = 0; q < E; ++q) Bs[buf][jq * E + q][kk] = rb[it][q];
            }
        }
    };

    // The 64 x 64 tile reads its C values (beta != 0) before the k loop, at
    // addresses clamped into C: with a short k (the factorisations' updates
    // have k = 32 or 256) the epilogue's reads otherwise wait on memory after
    // the last MFMA.  16 values a lane.
    constexpr bool PREC = BM == 64 && BN == 64;
    T cpre[PREC ? MT : 1][PREC ? NT : 1][4];
    if constexpr (PREC) {
        if (beta != T(0)) {
#pragma unroll
            for (int a = 0; a < MT; ++a)
#pragma unroll
                for (int b = 0; b < NT; ++b)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int gi = min(m0 + wm0 + a * 16 + acc_row<T>(lk, r), m - 1);
                        const int gj = min(n0 + wn0 + b * 16 + lr, n - 1);
                        cpre[a][b][r] = C[(int64_t)gi * ldc + gj];
                    }
        }
    }

    const int KT = (k + BK - 1) / BK;
    if (KT > 0) {
        load(0);
        store(0);
        __syncthreads();
    }
    for (int kt = 0; kt < KT; ++kt) {
        const int cur = kt & 1;
        const bool more = kt + 1 < KT;
        if (more) load((kt + 1) * BK);
        T fa[MT][4], fb[NT][4];
#pragma unroll
        for (int a = 0; a < MT; ++a) {
            const T *p = &As[cur][wm0 + a * 16 + lr][4 * lk];
#pragma unroll
            for (int v = 0; v < 4 / E; ++v) {
                const VT x = *reinterpret_cast<const VT *>(p + v * E);
#pragma unroll
                for (int q = 0; q < E; ++q) fa[a][v * E + q] = x[q];
            }
        }
#pragma unroll
        for (int b = 0; b < NT; ++b) {
            const T *p = &Bs[cur][wn0 + b * 16 + lr][4 * lk];
#pragma unroll
            for (int v = 0; v < 4 / E; ++v) {
                const VT x = *reinterpret_cast<const VT *>(p + v * E);
#pragma unroll
                for (int q = 0; q < E; ++q) fb[b][v * E + q] = x[q];
            }
        }
#pragma unroll
        for (int st = 0; st < 4; ++st)
#pragma unroll
            for (int a = 0; a < MT; ++a)
#pragma unroll
                for (int b = 0; b < NT; ++b) acc[a][b] = mfma16x16x4(fa[a][st], fb[b][st], acc[a][b]);
        if (more) store(cur ^ 1);
        __syncthreads();
    }

#pragma unroll
    for (int a = 0; a < MT; ++a)
#pragma unroll
        for (int b = 0; b < NT; ++b) {
            const int gj = n0 + wn0 + b * 16 + lr;
            if (gj >= n) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gi = m0 + wm0 + a * 16 + acc_row<T>(lk, r);
                if (gi < m && (!lower || gj <= gi)) {
                    T *c = C + (int64_t)gi * ldc + gj;
                    T v = alpha * acc[a][b][r];
                    if constexpr (PREC) {
                        if (beta != T(0)) v += beta * cpre[a][b][r];
                    } else {
                        if (beta != T(0)) v += beta * *c;
                    }
                    *c = v;
                }
            }
        }
}

// ---------------------------------------------------------------- BF16 tiles
//
// C (FP32) = alpha op(A) op(B) + beta C on v_mfma_f32_16x16x32_bf16 (FP32
// accumulation).  The operands are either BF16 in memory (S = uint16_t, the
// bit pattern) or FP32 rounded to BF16 on the way into LDS (S = float).
// With SPLIT (FP32 operands only) each element x becomes hi = bf16(x) and
// lo = bf16(x - hi), and every block product is hi*hi + hi*lo + lo*hi:
// three BF16 MFMAs recover about 16 significant bits of each product (the
// dropped lo*lo term is below 2^-16 of it).
//   Tiles are BM x BN x 32 (one MFMA k-chunk per LDS stage), 256 threads in a
//   2 x 2 grid of waves, double-buffered LDS with one barrier per stage.  In
//   LDS both operands are k-contiguous (As[i][k], Bs[j][k]) so that a lane's
//   MFMA operand (8 consecutive k of one row: A[i = l%16][k = 8(l/16)+0..7],
//   B[k = 8(l/16)+0..7][j = l%16]) is one 128-bit read; the row stride of 40
//   BF16 (80 bytes) keeps those reads off each other's banks.
//   An operand whose k runs along its stored rows (A as stored, B transposed)
//   is read in VE-element vectors along k and written as VE-element runs; an
//   operand whose rows run along memory (B as stored, A transposed) is read
//   as two VE-element vectors at k and k + 1 and written as VE packed BF16
//   pairs, the lanes of a wave on consecutive k pairs so the writes fall on
//   distinct banks.  VE = 4 needs every row 16-byte (FP32) or 8-byte (BF16)
//   aligned and the vector dimension a multiple of 4; VE = 1 takes any shape.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

__device__ inline unsigned short bf16_bits(float x) { return __builtin_bit_cast(unsigned short, (__bf16)x); }
__device__ inline float bf16_value(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

template <class S, bool SPLIT>
__device__ inline void bf16_parts(S x, unsigned short &hi, unsigned short &lo) {
    if constexpr (sizeof(S) == 2) {
        hi = x;
        lo = 0;
    } else {
        hi = bf16_bits(x);
        lo = SPLIT ? bf16_bits(x - bf16_value(hi)) : (unsigned short)0;
    }
}

template <class S, int VE>
struct alignas(sizeof(S) * VE) SVec {
    S v[VE];
};

constexpr int BF_BK = 32;
constexpr int BF_SK = 40;

// One operand tile of R rows (of op(X)) x BF_BK k, KC = k contiguous in memory.
template <class S, bool KC, int R, int VE, bool SPLIT>
struct BfTile {
    static constexpr int TASKS = KC ? R * BF_BK / VE / 256 : R * BF_BK / (2 * VE) / 256;
    static_assert(TASKS >= 1, "tile too small for the vector width");
    SVec<S, VE> r[TASKS][KC ? 1 : 2];

    // op(X)[row][kk] is X[row * ldx + kk] (KC) or X[kk * ldx + row]
    __device__ void load(const S *__restrict__ X, int64_t ldx, int r0, int rows, int k0, int kdim) {
        const int tid = threadIdx.x;
#pragma unroll
        for (int it = 0; it < TASKS; ++it) {
            const int e = it * 256 + tid;
            if (KC) {
                const int row = e / (BF_BK / VE), kq = e % (BF_BK / VE);
                const int gr = r0 + row, gk = k0 + kq * VE;
                if (gr < rows && gk < kdim)
                    r[it][0] = *reinterpret_cast<const SVec<S, VE> *>(X + (int64_t)gr * ldx + gk);
                else
#pragma unroll
                    for (int q = 0; q < VE; ++q) r[it][0].v[q] = S(0);
            } else {
                const int kp = e % (BF_BK / 2), rq = e / (BF_BK / 2);
                const int gr = r0 + rq * VE;
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int gk = k0 + 2 * kp + h;
                    if (gr < rows && gk < kdim)
                        r[it][h] = *reinterpret_cast<const SVec<S, VE> *>(X + (int64_t)gk * ldx + gr);
                    else
#pragma unroll
                        for (int q = 0; q < VE; ++q) r[it][h].v[q] = S(0);
                }
            }
        }
    }
    __device__ void store(unsigned short (*hi)[BF_SK], unsigned short (*lo)[BF_SK]) const {
        const int tid = threadIdx.x;
#pragma unroll
        for (int it = 0; it < TASKS; ++it) {
            const int e = it * 256 + tid;
            if (KC) {
                const int row = e / (BF_BK / VE), kq = e % (BF_BK / VE);
                unsigned short h[VE], l[VE];
#pragma unroll
                for (int q = 0; q < VE; ++q) bf16_parts<S, SPLIT>(r[it][0].v[q], h[q], l[q]);
#pragma unroll
                for (int q = 0; q < VE; ++q) {
                    hi[row][kq * VE + q] = h[q];
                    if constexpr (SPLIT) lo[row][kq * VE + q] = l[q];
                }
            } else {
                const int kp = e % (BF_BK / 2), rq = e / (BF_BK / 2);
#pragma unroll
                for (int q = 0; q < VE; ++q) {
                    unsigned short h0, l0, h1, l1;
                    bf16_parts<S, SPLIT>(r[it][0].v[q], h0, l0);
                    bf16_parts<S, SPLIT>(r[it][1].v[q], h1, l1);
                    *reinterpret_cast<unsigned *>(&hi[rq * VE + q][2 * kp]) = (unsigned)h0 | ((unsigned)h1 << 16);
                    if constexpr (SPLIT)
                        *reinterpret_cast<unsigned *>(&lo[rq * VE + q][2 * kp]) = (unsigned)l0 | ((unsigned)l1 << 16);
                }
            }
        }
    }
};

template <class S, bool TA, bool TB, bool SPLIT, int BM, int BN, int VE>
__global__ __launch_bounds__(256) void gemm_bf16_kernel(int m, int n, int k, float alpha, const S *__restrict__ A,
                                                        int64_t lda, const S *__restrict__ B, int64_t ldb,
                                                        float beta, float *__restrict__ C, int64_t ldc, int tiles_m,
                                                        int tiles_n, int lower) {
    constexpr int NP = SPLIT ? 2 : 1;
    constexpr int WM = BM / 2, WN = BN / 2, MT = WM / 16, NT = WN / 16;
    __shared__ __attribute__((aligned(16))) unsigned short As[2][NP][BM][BF_SK];
    __shared__ __attribute__((aligned(16))) unsigned short Bs[2][NP][BN][BF_SK];
    int tm, tn;
    tile_of(tiles_m, tiles_n, lower != 0, tm, tn);
    const int m0 = tm * BM, n0 = tn * BN;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm0 = (wave >> 1) * WM, wn0 = (wave & 1) * WN;
    const int lr = lane & 15, lk = lane >> 4;

    f4 acc[MT][NT];
#pragma unroll
    for (int a = 0; a < MT; ++a)
#pragma unroll
        for (int b = 0; b < NT; ++b) acc[a][b] = f4{0, 0, 0, 0};

    BfTile<S, !TA, BM, VE, SPLIT> ta;
    BfTile<S, TB, BN, VE, SPLIT> tb;
    const int KT = (k + BF_BK - 1) / BF_BK;
    if (KT > 0) {
        ta.load(A, lda, m0, m, 0, k);
        tb.load(B, ldb, n0, n, 0, k);
        ta.store(As[0][0], As[0][NP - 1]);
        tb.store(Bs[0][0], Bs[0][NP - 1]);
        __syncthreads();
    }
    for (int kt = 0; kt < KT; ++kt) {
        const int cur = kt & 1;
        const bool more = kt + 1 < KT;
        if (more) {
            ta.load(A, lda, m0, m, (kt + 1) * BF_BK, k);
            tb.load(B, ldb, n0, n, (kt + 1) * BF_BK, k);
        }
        bf16x8 fa[NP][MT], fb[NP][NT];
#pragma unroll
        for (int p = 0; p < NP; ++p) {
#pragma unroll
            for (int a = 0; a < MT; ++a)
                fa[p][a] = *reinterpret_cast<const bf16x8 *>(&As[cur][p][wm0 + a * 16 + lr][8 * lk]);
#pragma unroll
            for (int b = 0; b < NT; ++b)
                fb[p][b] = *reinterpret_cast<const bf16x8 *>(&Bs[cur][p][wn0 + b * 16 + lr][8 * lk]);
        }
#pragma unroll
        for (int a = 0; a < MT; ++a)
#pragma unroll
            for (int b = 0; b < NT; ++b) {
                if constexpr (SPLIT) {
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[1][a], fb[0][b], acc[a][b], 0, 0, 0);
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[0][a], fb[1][b], acc[a][b], 0, 0, 0);
                }
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[0][a], fb[0][b], acc[a][b], 0, 0, 0);
            }
        if (more) {
            ta.store(As[cur ^ 1][0], As[cur ^ 1][NP - 1]);
            tb.store(Bs[cur ^ 1][0], Bs[cur ^ 1][NP - 1]);
        }
        __syncthreads();
    }

#pragma unroll
    for (int a = 0; a < MT; ++a)
#pragma unroll
        for (int b = 0; b < NT; ++b) {
            const int gj = n0 + wn0 + b * 16 + lr;
            if (gj >= n) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gi = m0 + wm0 + a * 16 + acc_row<float>(lk, r);
                if (gi < m && (!lower || gj <= gi)) {
                    float *c = C + (int64_t)gi * ldc + gj;
                    float v = alpha * acc[a][b][r];
                    if (beta != 0.0f) v += beta * *c;
                    *c = v;
                }
            }
        }
}

template <class S, bool SPLIT, int BM, int BN, int VE>
void launch_bf16_tile(hipStream_t s, bool ta, bool tb, int m, int n, int k, float alpha, const S *A, int64_t lda,
                      const S *B, int64_t ldb, float beta, float *C, int64_t ldc, int lower, size_t lds_footprint) {
    const int tiles_m = (int)cdiv(m, BM), tiles_n = (int)cdiv(n, BN);
    const dim3 grid(lower ? lower_tiles(tiles_m, tiles_n) : tiles_m * tiles_n);
    const size_t st = 2 * (size_t)(SPLIT ? 2 : 1) * (BM + BN) * BF_SK * sizeof(unsigned short);
    const size_t reserve = lds_footprint > st ? lds_footprint - st : 0;
#define LMX_BF16_CASE(TA_, TB_)                                                                                   \
    hipLaunchKernelGGL((gemm_bf16_kernel<S, TA_, TB_, SPLIT, BM, BN, VE>), grid, dim3(256), reserve, s, m, n, k, \
                       alpha, A, lda, B, ldb, beta, C, ldc, tiles_m, tiles_n, lower)
    if (!ta && !tb)
        LMX_BF16_CASE(false, false);
    else if (!ta && tb)
        LMX_BF16_CASE(false, true);
    else if (ta && !tb)
        LMX_BF16_CASE(true, false);
    else
        LMX_BF16_CASE(true, true);
#undef LMX_BF16_CASE
    LMX_LAUNCH_CHECK();
}

// VE = 4 applies when every vector is wholly inside or outside the operand
// and aligned: rows aligned to 4 elements (16 bytes FP32, 8 bytes BF16) and
// the vector dimension (k for a k-contiguous operand, its rows otherwise) a
// multiple of 4.
template <class S>
bool bf16_vec_ok(bool ta, bool tb, int64_t m, int64_t n, int64_t k, const S *A, int64_t lda, const S *B,
                 int64_t ldb) {
    auto al = [](const void *p) { return ((uintptr_t)p & (4 * sizeof(S) - 1)) == 0; };
    const bool a_ok = ta ? (m % 4 == 0) : (k % 4 == 0);
    const bool b_ok = tb ? (k % 4 == 0) : (n % 4 == 0);
    return al(A) && al(B) && lda % 4 == 0 && ldb % 4 == 0 && a_ok && b_ok;
}

template <class S, bool SPLIT>
void gemm_bf16_impl(hipStream_t s, bool ta, bool tb, int64_t m, int64_t n, int64_t k, float alpha, const S *A,
                    int64_t lda, const S *B, int64_t ldb, float beta, float *C, int64_t ldc, int lower,
                    size_t lds_footprint) {
    if (m < 0 || n < 0 || k < 0) fail_arg("gemm_bf16: negative dimension");
    if (m == 0 || n == 0) return;
    if (m > INT32_MAX / 2 || n > INT32_MAX / 2 || k > INT32_MAX / 2) fail_arg("gemm_bf16: dimension too large");
    if (lower && m != n) fail_arg("gemm_bf16: lower tiles need a square C");
    const double flops = (lower ? 1.0 : 2.0) * (double)m * (double)n * (double)k;
    TimedRegion tr(SPLIT ? "bgemm3" : "bgemm", s, flops);
    const bool vec = bf16_vec_ok<S>(ta, tb, m, n, k, A, lda, B, ldb);
    const int64_t big = lower ? cdiv(m, 128) * (cdiv(m, 128) + 1) / 2 : cdiv(m, 128) * cdiv(n, 128);
    const int M = (int)m, N = (int)n, K = (int)k;
    if (big >= 512) {
        if (vec)
            launch_bf16_tile<S, SPLIT, 128, 128, 4>(s, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, lower,
                                                    lds_footprint);
        else
            launch_bf16_tile<S, SPLIT, 128, 128, 1>(s, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, lower,
                                                    lds_footprint);
    } else {
        if (vec)
            launch_bf16_tile<S, SPLIT, 64, 64, 4>(s, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, lower,
                                                  lds_footprint);
        else
            launch_bf16_tile<S, SPLIT, 64, 64, 1>(s, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, lower,
                                                  lds_footprint);
    }
}

// LDS of one workgroup of the two kernels (static arrays)
template <class T, int BM, int BN, int BK>
constexpr size_t gemm_lds() {
    return 2 * (size_t)BK * (BM + BN + 2 * ((sizeof(T) == 8) ? 1 : 2)) * sizeof(T);
}
template <class T, int BM, int BN>
constexpr size_t gemm_nn_lds() {
    return 2 * (size_t)(BM + BN) * (sizeof(T) == 4 ? 20 : 18) * sizeof(T);
}

// Whether gemm_nn_kernel applies: A as stored, B as stored or transposed,
// 16-byte aligned rows, and the dimensions its vectors run along (k, and n
// for a B as stored) multiples of the vector width, so that a vector is
// wholly inside or outside the operand.
template <class T>
bool nn_ok(bool ta, bool tb, int64_t n, int64_t k, const T *A, int64_t lda, const T *B, int64_t ldb,
           const BatchStrides &bs) {
    constexpr int E = V16<T>::N;
    auto al = [](const void *p) { return ((uintptr_t)p & 15) == 0; };
    return !ta && al(A) && al(B) && lda % E == 0 && ldb % E == 0 && k % E == 0 && (tb || n % E == 0) &&
           bs.a1 % E == 0 && bs.a2 % E == 0 && bs.b1 % E == 0 && bs.b2 % E == 0;
}

template <class T, int BM, int BN, int BK>
void launch_tile(hipStream_t s, bool ta, bool tb, int m, int n, int k, T alpha, const T *A,
                 int64_t lda, const T *B, int64_t ldb, T beta, T *C, int64_t ldc, int lower = 0,
                 int batch = 1, BatchStrides bs = BatchStrides{1, 0, 0, 0, 0, 0, 0}, size_t lds_footprint = 0,
                 bool nn_sharing = false) {
    int tiles_m = (int)cdiv(m, BM), tiles_n = (int)cdiv(n, BN);
    dim3 grid(lower ? lower_tiles(tiles_m, tiles_n) : tiles_m * tiles_n, (unsigned)batch), block(256);
    // lds_footprint: the LDS one workgroup should hold in all (its tiles plus
    // an unused reserve), so that fewer workgroups share a CU; 0 = no reserve.
    // A GEMM that shares its CUs with FP32 panel workgroups keeps the first
    // kernel: the NN kernel is faster on its own but slows those panels more
    // (FP32 LU at n=16384: 155 against 151 ms).
    const bool sharing = lds_footprint > 0 && sizeof(T) == 4 && !nn_sharing;
    if (!sharing && nn_ok<T>(ta, tb, n, k, A, lda, B, ldb, bs)) {
        const size_t st = gemm_nn_lds<T, BM, BN>(), reserve = lds_footprint > st ? lds_footprint - st : 0;
        if (tb)
            hipLaunchKernelGGL((gemm_nn_kernel<T, BM, BN, true>), grid, block, reserve, s, m, n, k, alpha, A, lda, B,
                               ldb, beta, C, ldc, tiles_m, tiles_n, bs, lower);
        else
            hipLaunchKernelGGL((gemm_nn_kernel<T, BM, BN, false>), grid, block, reserve, s, m, n, k, alpha, A, lda, B,
                               ldb, beta, C, ldc, tiles_m, tiles_n, bs, lower);
        LMX_LAUNCH_CHECK();
        return;
    }
    const size_t st = gemm_lds<T, BM, BN, BK>(), reserve_lds = lds_footprint > st ? lds_footprint - st : 0;
#define LMX_GEMM_CASE(TA_, TB_)                                                              \
    hipLaunchKernelGGL((gemm_mfma_kernel<T, BM, BN, BK, TA_, TB_>), grid, block, reserve_lds, s, m, n, k, \
                       alpha, A, lda, B, ldb, beta, C, ldc, tiles_m, tiles_n, lower, bs)
    if (!ta && !tb)
        LMX_GEMM_CASE(false, false);
    else if (!ta && tb)
        LMX_GEMM_CASE(false, true);
    else if (ta && !tb)
        LMX_GEMM_CASE(true, false);
    else
        LMX_GEMM_CASE(true, true);
#undef LMX_GEMM_CASE
    LMX_LAUNCH_CHECK();
}

// Tile shapes of the plain (non-batched, non-SYRK) GEMM dispatch.
enum GemmTile { TILE_DEFAULT = 0, TILE_64x64 = 1, TILE_128x64 = 2, TILE_128x128 = 3 };

// The autotuned table (gemm_tiles.inc, written by tune_gemm.py): for an
// element size and ranges of ceil(log2) of m, n and k, the tile that timed
// fastest on MI355X.  Shapes no rule covers take the fixed rule below.
struct GemmTileRule {
    int bytes;
    int m_lo, m_hi, n_lo, n_hi, k_lo, k_hi;  // inclusive ranges of ceil(log2(dim))
    int tile;
};
#include "gemm_tiles.inc"

int ceil_log2(int64_t v) {
    int b = 0;
    while (((int64_t)1 << b) < v) ++b;
    return b;
}

thread_local int t_tile_override = TILE_DEFAULT;

int table_tile(int bytes, int64_t m, int64_t n, int64_t k) {
    if (t_tile_override != TILE_DEFAULT) return t_tile_override;
    const int lm = ceil_log2(m), ln = ceil_log2(n), lk = ceil_log2(std::max<int64_t>(k, 1));
    for (const GemmTileRule &r : kGemmTileRules)
        if (r.bytes == bytes && lm >= r.m_lo && lm <= r.m_hi && ln >= r.n_lo && ln <= r.n_hi && lk >= r.k_lo &&
            lk <= r.k_hi)
            return r.tile;
    return TILE_DEFAULT;
}

template <class T>
void gemm_impl(hipStream_t s, bool ta, bool tb, int64_t m, int64_t n, int64_t k, T alpha,
               const T *A, int64_t lda, const T *B, int64_t ldb, T beta, T *C, int64_t ldc, size_t lds_footprint = 0,
               bool sharing = false) {
    if (m < 0 || n < 0 || k < 0) fail_arg("gemm: negative dimension");
    if (m == 0 || n == 0) return;
    if (m > INT32_MAX / 2 || n > INT32_MAX / 2 || k > INT32_MAX / 2) fail_arg("gemm: dimension too large");
    const double flops = 2.0 * (double)m * (double)n * (double)k;
    // A GEMM sharing its CUs with the LU panels keeps its measured tile (the
    // table is timed on an idle device); the others take the tuned table's
    // tile where it has one.
    const int tile = sharing ? TILE_DEFAULT : table_tile((int)sizeof(T), m, n, k);
    if (tile != TILE_DEFAULT) {
        const BatchStrides nob{1, 0, 0, 0, 0, 0, 0};
        const bool big = tile != TILE_64x64;
        TimedRegion tr(sizeof(T) == 8 ? (big ? "gemm" : "gemm64") : (big ? "sgemm" : "sgemm64"), s, flops);
        if (tile == TILE_64x64)
            launch_tile<T, 64, 64, 16>(s, ta, tb, (int)m, (int)n, (int)k, alpha, A, lda, B, ldb, beta, C, ldc, 0, 1,
                                       nob, lds_footprint);
        else if (tile == TILE_128x64)
            launch_tile<T, 128, 64, 16>(s, ta, tb, (int)m, (int)n, (int)k, alpha, A, lda, B, ldb, beta, C, ldc, 0, 1,
                                        nob, lds_footprint);
        else
            launch_tile<T, 128, 128, 16>(s, ta, tb, (int)m, (int)n, (int)k, alpha, A, lda, B, ldb, beta, C, ldc, 0,
                                         1, nob, lds_footprint);
        return;
    }
    // The fixed rule: large outputs take 128 x 128 tiles, smaller ones
    // 64 x 64 so the launch still spreads over the 256 CUs.  Each tile size
    // is its own timing family.
    int64_t big_tiles = cdiv(m, 128) * cdiv(n, 128);
    if (big_tiles >= 512) {
        TimedRegion tr(sizeof(T) == 8 ? "gemm" : "sgemm", s, flops);
        const BatchStrides nob{1, 0, 0, 0, 0, 0, 0};
        // FP64 NN: 128 x 64 tiles (224 registers a wave against 392 at
        // 128 x 128, so two waves share a SIMD): dgemm 8192^3 46 -> 53 TF.  A
        // GEMM beside the LU panels keeps 128 x 128: faster, it slowed the
        // panels more (FP64 LU at n=8192: 75.2 against 72.4 ms).
        if (sizeof(T) == 8 && !sharing && nn_ok<T>(ta, tb, n, k, A, lda, B, ldb, nob))
            launch_tile<T, 128, 64, 16>(s, ta, tb, (int)m, (int)n, (int)k, alpha, A, lda, B, ldb, beta, C, ldc, 0, 1,
                                        nob, lds_footprint);
        else
            launch_tile<T, 128, 128, 16>(s, ta, tb, (int)m, (int)n, (int)k, alpha, A, lda, B, ldb, beta,
                                         C, ldc, 0, 1, nob, lds_footprint);
    } else {
        TimedRegion tr(sizeof(T) == 8 ? "gemm64" : "sgemm64", s, flops);
        launch_tile<T, 64, 64, 16>(s, ta, tb, (int)m, (int)n, (int)k, alpha, A, lda, B, ldb, beta,
                                   C, ldc, 0, 1, BatchStrides{1, 0, 0, 0, 0, 0, 0}, lds_footprint);
    }
}

template <class T>
void gemm_batched_impl(hipStream_t s, int64_t m, int64_t n, int64_t k, T alpha, const T *A, int64_t lda,
                       const T *B, int64_t ldb, T beta, T *C, int64_t ldc, int nb1, int nb2, int64_t a1,
                       int64_t b1, int64_t c1, int64_t a2, int64_t b2, int64_t c2) {
    if (m <= 0 || n <= 0 || nb1 <= 0 || nb2 <= 0) return;
    if ((int64_t)nb1 * nb2 > 65535) fail_arg("gemm_batched: too many batch entries");
    TimedRegion tr(sizeof(T) == 8 ? "gemm64" : "sgemm64", s, 2.0 * (double)m * (double)n * (double)k * nb1 * nb2);
    BatchStrides bs{nb1, a1, b1, c1, a2, b2, c2};
    launch_tile<T, 64, 64, 16>(s, false, false, (int)m, (int)n, (int)k, alpha, A, lda, B, ldb, beta, C, ldc, 0,
                               nb1 * nb2, bs);
}

// Lower tiles of C = alpha op(A) op(A)^T + beta C: op(A) = A (A A^T, the
// Cholesky update) or, with `trans`, A^T (A^T A of matMulAtAP).
template <class T>
void syrk_impl(hipStream_t s, int64_t n, int64_t k, T alpha, const T *A, int64_t lda, T beta, T *C,
               int64_t ldc, bool trans = false, size_t lds_footprint = 0) {
    if (n <= 0) return;
    if (n > INT32_MAX / 2 || k > INT32_MAX / 2) fail_arg("syrk: dimension too large");
    TimedRegion tr(sizeof(T) == 8 ? "syrk" : "ssyrk", s, (double)n * (double)(n + 1) * (double)k);
    // 64 x 64 tiles at every size: the 128 x 128 FP64 tile holds one wave a
    // SIMD and ran at 20.8 TF alone (n = 7936, k = 256) against 29.4 TF
    // (40.5 against 28.8 at k = 512); FP32's lookahead SYRK beside the mixed
    // Cholesky's diagonal kernels: 76.6 against 64.8 TF, 32.8 against 33.1 ms.
    launch_tile<T, 64, 64, 16>(s, trans, !trans, (int)n, (int)n, (int)k, alpha, A, lda, A, lda, beta, C, ldc, 1,
                               1, BatchStrides{1, 0, 0, 0, 0, 0, 0}, lds_footprint, true);
}

}  // namespace

void gemm_bf16x3(hipStream_t s, bool ta, bool tb, int64_t m, int64_t n, int64_t k, float alpha, const float *A,
                 int64_t lda, const float *B, int64_t ldb, float beta, float *C, int64_t ldc, bool lower,
                 size_t lds_footprint) {
    gemm_bf16_impl<float, true>(s, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, lower ? 1 : 0,
                                lds_footprint);
}

void gemm_bf16(hipStream_t s, bool ta, bool tb, int64_t m, int64_t n, int64_t k, float alpha, const uint16_t *A,
               int64_t lda, const uint16_t *B, int64_t ldb, float beta, float *C, int64_t ldc) {
    gemm_bf16_impl<uint16_t, false>(s, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, 0, 0);
}

void gemm_bf16_rounded(hipStream_t s, bool ta, bool tb, int64_t m, int64_t n, int64_t k, float alpha, const float *A,
                       int64_t lda, const float *B, int64_t ldb, float beta, float *C, int64_t ldc) {
    gemm_bf16_impl<float, false>(s, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, 0, 0);
}

void syrk_lower(hipStream_t s, int64_t n, int64_t k, double alpha, const double *A, int64_t lda, double beta,
                double *C, int64_t ldc) {
    syrk_impl<double>(s, n, k, alpha, A, lda, beta, C, ldc);
}

template <class T>
void syrk_lower_sharing(hipStream_t s, int64_t n, int64_t k, T alpha, const T *A, int64_t lda, T beta, T *C,
                        int64_t ldc, size_t lds_footprint) {
    syrk_impl<T>(s, n, k, alpha, A, lda, beta, C, ldc, false, lds_footprint);
}
template void syrk_lower_sharing<double>(hipStream_t, int64_t, int64_t, double, const double *, int64_t, double,
                                         double *, int64_t, size_t);
template void syrk_lower_sharing<float>(hipStream_t, int64_t, int64_t, float, const float *, int64_t, float,
                                        float *, int64_t, size_t);

void syrk_lower_op(hipStream_t s, bool trans, int64_t n, int64_t k, double alpha, const double *A, int64_t lda,
                   double beta, double *C, int64_t ldc) {
    syrk_impl<double>(s, n, k, alpha, A, lda, beta, C, ldc, trans);
}

void syrk_lower(hipStream_t s, int64_t n, int64_t k, float alpha, const float *A, int64_t lda, float beta,
                float *C, int64_t ldc) {
    syrk_impl<float>(s, n, k, alpha, A, lda, beta, C, ldc);
}

// C (m x n, m >= n) = alpha A A[0:n]^T + beta C on the elements with row >=
// column: the update of a block column below and on the diagonal, in one
// launch of 64 x 64 tiles.
template <class T>
void syrk_trap(hipStream_t s, int64_t m, int64_t n, int64_t k, T alpha, const T *A, int64_t lda, T beta, T *C,
               int64_t ldc) {
    if (m <= 0 || n <= 0) return;
    if (n > m) fail_arg("syrk_trap: more columns than rows");
    if (m > INT32_MAX / 2 || k > INT32_MAX / 2) fail_arg("syrk_trap: dimension too large");
    const double elems = (double)n * (double)(n + 1) / 2 + (double)(m - n) * (double)n;
    TimedRegion tr(sizeof(T) == 8 ? "gemm64" : "sgemm64", s, 2.0 * elems * (double)k);
    launch_tile<T, 64, 64, 16>(s, false, true, (int)m, (int)n, (int)k, alpha, A, lda, A, lda, beta, C, ldc, 1);
}
template void syrk_trap<double>(hipStream_t, int64_t, int64_t, int64_t, double, const double *, int64_t, double,
                                double *, int64_t);
template void syrk_trap<float>(hipStream_t, int64_t, int64_t, int64_t, float, const float *, int64_t, float, float *,
                               int64_t);

void gemm(hipStream_t s, bool ta, bool tb, int64_t m, int64_t n, int64_t k, double alpha,
          const double *A, int64_t lda, const double *B, int64_t ldb, double beta, double *C,
          int64_t ldc) {
    gemm_impl<double>(s, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

void gemm(hipStream_t s, bool ta, bool tb, int64_t m, int64_t n, int64_t k, float alpha,
          const float *A, int64_t lda, const float *B, int64_t ldb, float beta, float *C,
          int64_t ldc) {
    gemm_impl<float>(s, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

template <class T>
void gemm_sharing(hipStream_t s, int64_t m, int64_t n, int64_t k, T alpha, const T *A, int64_t lda, const T *B,
                  int64_t ldb, T beta, T *C, int64_t ldc, size_t lds_footprint) {
    gemm_impl<T>(s, false, false, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, lds_footprint, true);
}
template void gemm_sharing<double>(hipStream_t, int64_t, int64_t, int64_t, double, const double *, int64_t,
                                   const double *, int64_t, double, double *, int64_t, size_t);
template void gemm_sharing<float>(hipStream_t, int64_t, int64_t, int64_t, float, const float *, int64_t,
                                  const float *, int64_t, float, float *, int64_t, size_t);

template <class T>
void gemm_batched(hipStream_t s, int64_t m, int64_t n, int64_t k, T alpha, const T *A, int64_t lda, const T *B,
                  int64_t ldb, T beta, T *C, int64_t ldc, int nb1, int nb2, int64_t a1, int64_t b1, int64_t c1,
                  int64_t a2, int64_t b2, int64_t c2) {
    gemm_batched_impl<T>(s, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, nb1, nb2, a1, b1, c1, a2, b2, c2);
}
template void gemm_batched<double>(hipStream_t, int64_t, int64_t, int64_t, double, const double *, int64_t,
                                   const double *, int64_t, double, double *, int64_t, int, int, int64_t, int64_t,
                                   int64_t, int64_t, int64_t, int64_t);
template void gemm_batched<float>(hipStream_t, int64_t, int64_t, int64_t, float, const float *, int64_t,
                                  const float *, int64_t, float, float *, int64_t, int, int, int64_t, int64_t,
                                  int64_t, int64_t, int64_t, int64_t);

}  // namespace lmx

using namespace lmx;

namespace {

void check_gemm_args(bool ta, bool tb, int64_t m, int64_t n, int64_t k, int64_t lda, int64_t ldb,
                     int64_t ldc) {
    if (m < 0 || n < 0 || k < 0) fail_arg("gemm: negative dimension");
    int64_t acols = ta ? m : k, bcols = tb ? k : n;
    if (lda < (acols > 1 ? acols : 1)) fail_arg("gemm: lda too small");
    if (ldb < (bcols > 1 ? bcols : 1)) fail_arg("gemm: ldb too small");
    if (ldc < (n > 1 ? n : 1)) fail_arg("gemm: ldc too small");
}

template <class T>
int gemm_host(char transa, char transb, int64_t m, int64_t n, int64_t k, T alpha, const T *A,
              int64_t lda, const T *B, int64_t ldb, T beta, T *C, int64_t ldc) {
    return guarded([&] {
        bool ta = is_trans(transa), tb = is_trans(transb);
        check_gemm_args(ta, tb, m, n, k, lda, ldb, ldc);
        if (m == 0 || n == 0) return LMX_OK;
        Context &c = ctx();
        std::lock_guard<std::mutex> g(c.host_mu);
        int64_t arows = ta ? k : m, brows = tb ? n : k;
        int64_t acols = ta ? m : k, bcols = tb ? k : n;
        // exact extents: the last row ends at its last column, not at the stride
        auto extent = [](int64_t rows, int64_t cols, int64_t ld) -> size_t {
            return (rows > 0 && cols > 0) ? (size_t)((rows - 1) * ld + cols) * sizeof(T) : 0;
        };
        size_t abytes = extent(arows, acols, lda), bbytes = extent(brows, bcols, ldb),
               cbytes = extent(m, n, ldc);
        T *dA = (T *)staging(c, 0, abytes), *dB = (T *)staging(c, 1, bbytes),
          *dC = (T *)staging(c, 2, cbytes);
        if (abytes) LMX_HIP(hipMemcpyAsync(dA, A, abytes, hipMemcpyHostToDevice, c.stream));
        if (bbytes) LMX_HIP(hipMemcpyAsync(dB, B, bbytes, hipMemcpyHostToDevice, c.stream));
        if (beta != T(0)) LMX_HIP(hipMemcpyAsync(dC, C, cbytes, hipMemcpyHostToDevice, c.stream));
        else if (ldc > n)  // keep the untouched padding of C intact on the way back
            LMX_HIP(hipMemcpyAsync(dC, C, cbytes, hipMemcpyHostToDevice, c.stream));
        gemm(c.stream, ta, tb, m, n, k, alpha, dA, lda, dB, ldb, beta, dC, ldc);
        LMX_HIP(hipMemcpyAsync(C, dC, cbytes, hipMemcpyDeviceToHost, c.stream));
        LMX_HIP(hipStreamSynchronize(c.stream));
        return LMX_OK;
    });
}

}  // namespace

extern "C" {

int lmx_dgemm(char transa, char transb, int64_t m, int64_t n, int64_t k, double alpha,
              const double *A, int64_t lda, const double *B, int64_t ldb, double beta, double *C,
              int64_t ldc) {
    return gemm_host<double>(transa, transb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

int lmx_sgemm(char transa, char transb, int64_t m, int64_t n, int64_t k, float alpha,
              const float *A, int64_t lda, const float *B, int64_t ldb, float beta, float *C,
              int64_t ldc) {
    return gemm_host<float>(transa, transb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

int lmx_dgemm_dev(void *stream, char transa, char transb, int64_t m, int64_t n, int64_t k,
                  double alpha, const double *dA, int64_t lda, const double *dB, int64_t ldb,
                  double beta, double *dC, int64_t ldc) {
    return guarded([&] {
        bool ta = is_trans(transa), tb = is_trans(transb);
        check_gemm_args(ta, tb, m, n, k, lda, ldb, ldc);
        gemm(resolve_stream(stream), ta, tb, m, n, k, alpha, dA, lda, dB, ldb, beta, dC, ldc);
        return LMX_OK;
    });
}

int lmx_sgemm_dev(void *stream, char transa, char transb, int64_t m, int64_t n, int64_t k,
                  float alpha, const float *dA, int64_t lda, const float *dB, int64_t ldb,
                  float beta, float *dC, int64_t ldc) {
    return guarded([&] {
        bool ta = is_trans(transa), tb = is_trans(transb);
        check_gemm_args(ta, tb, m, n, k, lda, ldb, ldc);
        gemm(resolve_stream(stream), ta, tb, m, n, k, alpha, dA, lda, dB, ldb, beta, dC, ldc);
        return LMX_OK;
    });
}

int lmx_gemm_tile_override(int tile) {
    return guarded([&] {
        if (tile < TILE_DEFAULT || tile > TILE_128x128) fail_arg("gemm_tile_override: tile must be 0..3");
        t_tile_override = tile;
        return LMX_OK;
    });
}

int lmx_gemm_bf16_dev(void *stream, char transa, char transb, int64_t m, int64_t n, int64_t k, float alpha,
                      const uint16_t *dA, int64_t lda, const uint16_t *dB, int64_t ldb, float beta, float *dC,
                      int64_t ldc) {
    return guarded([&] {
        bool ta = is_trans(transa), tb = is_trans(transb);
        check_gemm_args(ta, tb, m, n, k, lda, ldb, ldc);
        gemm_bf16(resolve_stream(stream), ta, tb, m, n, k, alpha, dA, lda, dB, ldb, beta, dC, ldc);
        return LMX_OK;
    });
}

int lmx_sgemm_bf16x3_dev(void *stream, char transa, char transb, int64_t m, int64_t n, int64_t k, float alpha,
                         const float *dA, int64_t lda, const float *dB, int64_t ldb, float beta, float *dC,
                         int64_t ldc) {
    return guarded([&] {
        bool ta = is_trans(transa), tb = is_trans(transb);
        check_gemm_args(ta, tb, m, n, k, lda, ldb, ldc);
        gemm_bf16x3(resolve_stream(stream), ta, tb, m, n, k, alpha, dA, lda, dB, ldb, beta, dC, ldc, false, 0);
        return LMX_OK;
    });
}

int lmx_gemm_bf16(char transa, char transb, int64_t m, int64_t n, int64_t k, float alpha, const uint16_t *A,
                  int64_t lda, const uint16_t *B, int64_t ldb, float beta, float *C, int64_t ldc) {
    return guarded([&] {
        bool ta = is_trans(transa), tb = is_trans(transb);
        check_gemm_args(ta, tb, m, n, k, lda, ldb, ldc);
        if (m == 0 || n == 0) return LMX_OK;
        Context &c = ctx();
        std::lock_guard<std::mutex> g(c.host_mu);
        const int64_t arows = ta ? k : m, brows = tb ? n : k, acols = ta ? m : k, bcols = tb ? k : n;
        auto ext = [](int64_t rows, int64_t cols, int64_t ld, size_t es) -> size_t {
            return (rows > 0 && cols > 0) ? (size_t)((rows - 1) * ld + cols) * es : 0;
        };
        const size_t ab = ext(arows, acols, lda, 2), bb = ext(brows, bcols, ldb, 2), cb = ext(m, n, ldc, 4);
        auto *dA = (uint16_t *)staging(c, 0, ab), *dB = (uint16_t *)staging(c, 1, bb);
        auto *dC = (float *)staging(c, 2, cb);
        if (ab) LMX_HIP(hipMemcpyAsync(dA, A, ab, hipMemcpyHostToDevice, c.stream));
        if (bb) LMX_HIP(hipMemcpyAsync(dB, B, bb, hipMemcpyHostToDevice, c.stream));
        if (beta != 0.0f || ldc > n) LMX_HIP(hipMemcpyAsync(dC, C, cb, hipMemcpyHostToDevice, c.stream));
        gemm_bf16(c.stream, ta, tb, m, n, k, alpha, dA, lda, dB, ldb, beta, dC, ldc);
        LMX_HIP(hipMemcpyAsync(C, dC, cb, hipMemcpyDeviceToHost, c.stream));
        LMX_HIP(hipStreamSynchronize(c.stream));
        return LMX_OK;
    });
}

}  // extern "C"
