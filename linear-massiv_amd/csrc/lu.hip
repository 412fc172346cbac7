// LU factorisation with partial pivoting (right-looking, blocked) and the
// LU solve, for row-major matrices.
//
// Panel kernel (getf2_panel_kernel): the m x kb panel (kb <= 32) is spread
// over P concurrent workgroups of 512 rows each (FP32: 4 waves, two rows a
// thread; FP64: 8 waves, one row a thread), rows held in registers.  Per
// column j:
//   1. each workgroup finds its candidate max |a_ij| over rows i >= j (ties:
//      lowest row, NaN never chosen -- the rule of rawPivotSearch,
//      src/.../Internal/Kernel.hs:1541-1556; or, for the generic lu, the
//      highest row among equal magnitudes -- maximumBy keeps the later of
//      two equal elements, src/.../Solve/LU.hs:128-135): DPP wave
//      reductions and ballots, then one LDS exchange between the waves;
//   2. it publishes the candidate (value, row, full row data) in slot j%2,
//      and workgroup 0 publishes row j, each word tagged with the column;
//   3. every workgroup polls all P records, reduces them to the pivot p,
//      swaps rows j and p, and all rows i > j get l = a_ij / a_pj (a
//      division, as the reference) and the rank-1 update of the remaining
//      panel columns.
// The current column always sits in register 0..3: after every 4 columns
// the row's registers rotate left by 4, so the column loop needs no dynamic
// register indexing.  At the end every workgroup derives the panel's
// row-move list from the pivots and applies it to the other columns of the
// outer panel (and the permutation vector) with one parallel gather, and
// solves U12 = L11^{-1} B12 for the columns to the right of the panel.
//
// Driver (getrf_impl): outer panels of NB = 256 columns, each factored as
// eight 32-column sub-panels (panel kernel, then A22 -= L21 U12 inside the
// outer panel on the MFMA GEMM), then U12 = L11^{-1} A12 (TRSM through the
// inverse of L11) and the trailing update A22 -= L21 U12 on the MFMA GEMM.
#include <climits>
#include <vector>

#include "common.hpp"
#include "device.hpp"
#include "kernels.hpp"
#include "sync.hpp"

namespace lmx {

namespace {

constexpr int KB = 32;        // panel kernel width
constexpr int RPT = 1;        // rows per thread in the panel kernel
constexpr int PT = 512;       // threads per panel workgroup (8 waves)
constexpr int PW = PT / 64;   // waves per panel workgroup
constexpr int PANEL_ROWS = PT * RPT;
constexpr int MAXP = 64;      // max panel workgroups (panel height <= 32768)
constexpr int NB_OUTER = 256; // outer panel width (k of the trailing GEMM)
constexpr long SPIN_LIMIT = 1L << 22;

// Cross-workgroup records of the panel kernel.  Every word is 64 bits: a
// 32-bit payload and, in the upper half, the tag base + j + 1 of the column
// it belongs to, written with one device-coherent store; a reader polls a
// word until its tag matches, so the poll is the load and no flag is needed.
// Tags grow over the whole factorisation (the records are zeroed once per
// getrf) and records are double-buffered by column parity.
//   workgroup record: value (WPV words), row, pad, row data
//   row-j record:     pad, pad, row data
template <class T>
struct PanelRec {
    static constexpr int WPV = sizeof(T) / 4;
    static constexpr int DATA = WPV + 2;  // data offset (even for doubles)
    static constexpr int W = DATA + KB * WPV;                  // words per workgroup record
    static constexpr int DDATA = 2;
    static constexpr int DW = DDATA + KB * WPV;                // words of the row-j record
    static constexpr int TOTAL = MAXP * W + DW;
};
template <class T>
struct PanelSync {
    // the P workgroup records, then the row-j record (at P * W)
    unsigned long long rec[2][PanelRec<T>::TOTAL];
    unsigned error;
    unsigned pad[63];
};
template <class T>
constexpr int sync_clear_words() {
    return (int)(sizeof(PanelSync<T>) / 4);
}

__device__ inline unsigned long long tagword(unsigned payload, unsigned tag) {
    return (unsigned long long)payload | ((unsigned long long)tag << 32);
}
template <class T>
__device__ inline unsigned word_of(T v, int h) {
    if constexpr (sizeof(T) == 4)
        return __builtin_bit_cast(unsigned, v);
    else
        return h == 0 ? (unsigned)__builtin_bit_cast(unsigned long long, v)
                      : (unsigned)(__builtin_bit_cast(unsigned long long, v) >> 32);
}

constexpr int PG = 4;  // columns per register rotation

// Shape of the partial-pivoting panel kernel's workgroup (PANEL_ROWS rows):
// 8 waves of one row a thread.  A wave's rows are consecutive: row block q
// of a wave is rows 64 q .. 64 q + 63 of its 64 RPT.  (Two rows a thread in
// 4 waves, and a one-row panel held to 85 registers, both let a panel
// workgroup share its CU with two workgroups of the FP32 lookahead GEMM;
// both measured slower beside that GEMM.)
template <class T>
struct PanelShape {
    static constexpr int RPT = 1;               // rows per thread
    static constexpr int PT = PANEL_ROWS / RPT;  // threads
    static constexpr int PW = PT / 64;           // waves
    static constexpr int MIN_WAVES_PER_SIMD = 2;
};

// Per-thread state of the panel kernel: R rows of the panel in registers,
// in the rotated frame (register c holds column j0 + c of the current
// group).
template <class T, int R>
struct PanelRows {
    T v[R][KB];
    int row[R];
};

// Columns a pass of the panel kernel's closing row gather: the U12 block
// of a pass (KB x GCOLS) shares LDS with the poll payloads.
template <class T>
constexpr int gather_cols() {
    return 64;
}

// LDS of the panel kernel.  What is written before a column's first
// barrier (wave slots, row j) is double-buffered by column parity; the poll
// payloads are written after it and read before the next column's, so one
// buffer serves.
template <class T>
struct PanelLds {
    static constexpr int NW = PanelShape<T>::PW;
    alignas(16) T slot_data[2][NW][KB];  // each wave's best row
    T slot_val[2][NW];
    int slot_row[2][NW];
    T drow[2][KB];                 // row j (single workgroup)
    int pivs[KB];
    T l11[KB][KB];                 // the panel's unit-lower L11: row j recorded at column j
    union {
        alignas(8) unsigned pay[PanelRec<T>::TOTAL];  // polled record payloads (several workgroups)
        alignas(16) T b12[KB][gather_cols<T>()];      // closing gather: B12, then U12
    };
    int mv_dst[2 * KB];  // the rows that move (panel-relative) ...
    int mv_src[2 * KB];  // ... and the original row each one receives
    int mv_cnt;
};

// Tie rules of the pivot search: the lowest row among equal magnitudes
// (TL = false, rawPivotSearch) or the highest (TL = true, lu's maximumBy).
// A row id of "none" loses to every real row under either rule.
template <bool TL>
struct TieRule {
    static constexpr int NONE = TL ? -1 : INT_MAX;
    __device__ static int pick(int a, int b) { return TL ? max(a, b) : min(a, b); }
    __device__ static bool better(int a, int b) { return TL ? a > b : a < b; }
    __device__ static int wave_pick(int v) { return TL ? wave_max_int(v) : wave_min(v); }
};

// The workgroup's best of the PW wave slots: largest value, the tie rule's
// id on ties.  All slots are loaded before any comparison: compared as they
// are read, every step waits for its own LDS round trip.
template <class T, bool TL = false>
__device__ __forceinline__ void wg_select(const T *vals, const int *ids, T &best, int &bid, int &bw) {
    T v[PW];
    int id[PW];
#pragma unroll
    for (int x = 0; x < PW; ++x) {
        v[x] = vals[x];
        id[x] = ids[x];
    }
    best = v[0];
    bid = id[0];
    bw = 0;
#pragma unroll
    for (int x = 1; x < PW; ++x) {
        const bool take = v[x] > best || (v[x] == best && TieRule<TL>::better(id[x], bid));
        best = take ? v[x] : best;
        bid = take ? id[x] : bid;
        bw = take ? x : bw;
    }
}

// The tie rule's bit among the set bits of a wave-uniform mask whose bit
// order is row order: the lowest (rawPivotSearch) or the highest (lu's
// maximumBy); -1 for an empty mask.
template <bool TL>
__device__ __forceinline__ int tie_index(unsigned long long mask) {
    if (mask == 0) return -1;
    return TL ? 63 - __builtin_clzll(mask) : __builtin_ctzll(mask);
}

// One column (j = j0 + U) of the panel: pivot search, interchange, scaling
// and rank-1 update of the rest of the panel.
template <class T, int U, bool TL>
__device__ __forceinline__ void panel_column(PanelRows<T, PanelShape<T>::RPT> &st, PanelLds<T> &L, int j0, int kb,
                                             int m, int P, PanelSync<T> *__restrict__ sy, unsigned base, int k0,
                                             int32_t *__restrict__ ipiv, int &local_info, bool &failed) {
    typedef PanelShape<T> S;
    constexpr int NR = S::RPT;
    typedef PanelRec<T> R;
    typedef TieRule<TL> Tie;
    const int j = j0 + U;
    if (j >= kb) return;
    const int w = blockIdx.x, t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int par = j & 1;
    // 1. this wave's candidate: max |a_ij| over its rows i >= j (fmax skips
    //    NaN), then the tie rule's row among those attaining it (row block
    //    q of a wave precedes block q + 1, lanes ascend within a block);
    //    the owning lane writes that row to the wave slot
    bool live[NR];
    T a[NR];
    T tmax = T(-1);
#pragma unroll
    for (int q = 0; q < NR; ++q) {
        live[q] = st.row[q] >= j && st.row[q] < m;
        a[q] = live[q] ? fabs(st.v[q][U]) : T(-1);
        tmax = fmax(tmax, a[q]);
    }
    const T wmax = wave_max(tmax);
    int wl = -1, wq = 0;
#pragma unroll
    for (int qq = 0; qq < NR; ++qq) {
        const int q = TL ? NR - 1 - qq : qq;
        const unsigned long long mk = __ballot(a[q] == wmax && a[q] >= T(0));
        if (wl < 0 && mk != 0) {
            wl = tie_index<TL>(mk);
            wq = q;
        }
    }
    if (lane == wl) {
#pragma unroll
        for (int c = 0; c < KB; ++c) {
            T x = st.v[0][c];
#pragma unroll
            for (int q = 1; q < NR; ++q) x = wq == q ? st.v[q][c] : x;
            L.slot_data[par][wave][c] = x;
        }
    }
#pragma unroll
    for (int q = 0; q < NR; ++q)
        if (st.row[q] == j) {
#pragma unroll
            for (int c = 0; c < KB; ++c) L.drow[par][c] = st.v[q][c];
        }
    if (lane == 0) {
        // a wave with no candidate offers -1, so no slot holds a NaN
        L.slot_val[par][wave] = wl < 0 ? T(-1) : wmax;
        L.slot_row[par][wave] = wl < 0 ? Tie::NONE : w * PANEL_ROWS + wave * 64 * NR + wq * 64 + wl;
    }
    __syncthreads();
    // 2. this workgroup's candidate: lane x reads slot x; waves hold
    //    ascending rows, so the tie rule over rows is the one over waves
    const T sv = lane < S::PW ? L.slot_val[par][lane] : T(-1);
    const T best = wave_max(sv);
    const int bw = tie_index<TL>(__ballot(lane < S::PW && sv == best && best >= T(0)));
    const int brow = bw < 0 ? Tie::NONE : L.slot_row[par][bw];
    int p;
    const T *prow;
    const T *drow;
    if (P == 1) {
        drow = L.drow[par];
        // a NaN at row j keeps row j, as rawPivotSearch does
        if (bw < 0 || isnan(drow[U])) {
            p = j;
            prow = drow;
        } else {
            p = brow;
            prow = L.slot_data[par][bw];
        }
    } else {
        const unsigned tg = base + (unsigned)j + 1u;
        // The records' base, opaque to the compiler: every address below is
        // formed here, in the column, rather than hoisted out of the column
        // loop into registers held across all of it.
        unsigned long long *rbase = sy->rec[par];
        asm volatile("" : "+v"(rbase));
        // 3. publish this workgroup's record (and row j from workgroup 0):
        //    lane e of wave 0 writes word e
        if (wave == 0) {
            unsigned long long *rec = rbase + w * R::W;
            const unsigned *data = reinterpret_cast<const unsigned *>(L.slot_data[par][bw < 0 ? 0 : bw]);
            for (int e = lane; e < R::W; e += 64) {
                const unsigned dv = data[max(e - R::DATA, 0)];
                const unsigned hv = e < R::WPV ? word_of(best, e) : (e == R::WPV ? (unsigned)brow : 0u);
                const unsigned pl = e >= R::DATA ? (bw < 0 ? 0u : dv) : hv;
                xstore(rec + e, tagword(pl, tg));
            }
            if (w == 0) {
                const unsigned *dd = reinterpret_cast<const unsigned *>(L.drow[par]);
                for (int e = lane; e < R::DW; e += 64) {
                    const unsigned dv = dd[max(e - R::DDATA, 0)];
                    xstore(rbase + P * R::W + e, tagword(e >= R::DDATA ? dv : 0u, tg));
                }
            }
        }
        // 4. poll every record: all loads in flight first, then the
        //    stragglers one by one
        constexpr int NPOLL = (R::TOTAL + S::PT - 1) / S::PT;
        const int total = P * R::W + R::DW;
        unsigned long long wv[NPOLL];
#pragma unroll
        for (int i = 0; i < NPOLL; ++i) {
            const int e = t + i * S::PT;
            if (e < total) {
                wv[i] = xload(rbase + e);
            }
        }
#pragma unroll
        for (int i = 0; i < NPOLL; ++i) {
            const int e = t + i * S::PT;
            if (e < total) {
                const unsigned long long *wp = rbase + e;
                long spins = 0;
                while ((unsigned)(wv[i] >> 32) != tg && !failed) {
                    __builtin_amdgcn_s_sleep(1);
                    if ((++spins & 255) == 0 && xload(&sy->error) != 0u) failed = true;
                    if (spins > SPIN_LIMIT) {
                        xstore(&sy->error, 1u);
                        failed = true;
                    }
                    wv[i] = xload(wp);
                }
                L.pay[e] = (unsigned)wv[i];
            }
        }
        __syncthreads();
        // 5. the pivot: largest value over the workgroups; workgroups hold
        //    ascending rows, so the tie rule over rows is the one over
        //    workgroups.  Every wave reduces the P records itself (lane x
        //    reads record x).  Record values are -1 or a magnitude, never NaN.
        const unsigned *pay = L.pay;
        const T xv = lane < P ? *reinterpret_cast<const T *>(pay + lane * R::W) : T(-1);
        const T gb = wave_max(xv);
        const int gw = tie_index<TL>(__ballot(lane < P && xv == gb && gb >= T(0)));
        drow = reinterpret_cast<const T *>(pay + P * R::W + R::DDATA);
        if (gw < 0 || isnan(drow[U])) {
            p = j;
            prow = drow;
        } else {
            p = (int)pay[gw * R::W + R::WPV];
            prow = reinterpret_cast<const T *>(pay + gw * R::W + R::DATA);
        }
    }
    // row j of L11: the pivot row's multipliers of columns k < j (final
    // since column k; register (k - j0) mod KB in the rotated frame)
    if (t < j) L.l11[j][t] = prow[(t - j0) & (KB - 1)];
    // 6. interchange rows j and p
    if (p != j) {
#pragma unroll
        for (int q = 0; q < NR; ++q) {
            if (st.row[q] == p) {
#pragma unroll
                for (int c = 0; c < KB; ++c) st.v[q][c] = drow[c];
            } else if (st.row[q] == j) {
#pragma unroll
                for (int c = 0; c < KB; ++c) st.v[q][c] = prow[c];
            }
        }
    }
    if (t == 0) L.pivs[j] = p;
    if (w == 0 && t == 0) ipiv[k0 + j] = k0 + p;
    // 7. multipliers (a division, as the reference) and the rank-1 update
    //    of the panel columns after j.  The pivot row is read from LDS in
    //    one batch first.  Registers past KB - j0 hold the finished columns
    //    (rotated to the end) and are skipped in blocks of PG; registers of
    //    columns past kb (a narrow last panel) are padding that is updated
    //    but never stored.
    T pr[KB];
#pragma unroll
    for (int c = U; c < KB; ++c) pr[c] = prow[c];
    const T piv = pr[U];
    if (piv != T(0)) {
#pragma unroll
        for (int q = 0; q < NR; ++q)
            if (live[q] && st.row[q] > j) {
                const T l = st.v[q][U] / piv;
                st.v[q][U] = l;
#pragma unroll
                for (int b = (U + 1) / PG * PG; b < KB; b += PG)
                    if (b < KB - j0) {
#pragma unroll
                        for (int e = 0; e < PG; ++e)
                            if (b + e > U) st.v[q][b + e] = fma(-l, pr[b + e], st.v[q][b + e]);
                    }
            }
    } else if (local_info == 0) {
        local_info = k0 + j + 1;
    }
}

// 16-byte vectors of panel row elements.
template <class T>
struct RowVec;
template <>
struct RowVec<float> {
    typedef f4 type;
};
template <>
struct RowVec<double> {
    typedef d2 type;
};

// The row interchanges of a panel applied to other columns of the matrix
// (rows relative to the panel's first row): columns [c0, c1) and [c2, c3)
// of G, and the permutation vector perm when not null; pairs_out receives
// the move list (rows absolute) for the columns left for later.
//   With u12 set, rows 0 .. KB - 1 of columns [c2, c3) (B12 of a full
// panel) receive U12 = L11^{-1} B12 instead of the gathered B12.
template <class T>
struct PanelGather {
    T *G;
    int64_t ldg;
    int c0, c1, c2, c3;
    int32_t *perm;
    int32_t *pairs_out;
    int u12;
};

// At most 128 VGPRs (four waves a SIMD), so that GEMM waves of the other
// stream still find registers on the panel's CUs.
template <class T, bool TL>
__global__ __launch_bounds__(PanelShape<T>::PT)
__attribute__((amdgpu_waves_per_eu(PanelShape<T>::MIN_WAVES_PER_SIMD))) void getf2_panel_kernel(int m, int kb, T *__restrict__ A, int64_t lda,
                                                          int k0, int32_t *__restrict__ ipiv,
                                                          int32_t *__restrict__ info,
                                                          PanelSync<T> *__restrict__ sy, int P,
                                                          unsigned base, PanelGather<T> gx) {
    critical_path_priority();
    typedef PanelShape<T> S;
    const int w = blockIdx.x, t = threadIdx.x;
    __shared__ PanelLds<T> L;
    PanelRows<T, S::RPT> st;
    // A full-width panel whose rows start on 16-byte boundaries moves each
    // row in 16-byte vectors (a quarter of the address work of 4-byte ones)
    typedef typename RowVec<T>::type V;
    constexpr int VE = 16 / sizeof(T);
    const bool vec = kb == KB && (reinterpret_cast<uintptr_t>(A) & 15) == 0 && ((lda * (int64_t)sizeof(T)) & 15) == 0;
#pragma unroll
    for (int q = 0; q < S::RPT; ++q) {
        st.row[q] = w * PANEL_ROWS + (t >> 6) * 64 * S::RPT + q * 64 + (t & 63);
        if (vec) {
            const T *src = A + (int64_t)st.row[q] * lda;
#pragma unroll
            for (int c = 0; c < KB; c += VE) {
                V x = {};
                if (st.row[q] < m) x = *reinterpret_cast<const V *>(src + c);
#pragma unroll
                for (int e = 0; e < VE; ++e) st.v[q][c + e] = x[e];
            }
        } else {
#pragma unroll
            for (int c = 0; c < KB; ++c)
                st.v[q][c] = (st.row[q] < m && c < kb) ? A[(int64_t)st.row[q] * lda + c] : T(0);
        }
    }
    int local_info = 0;
    bool failed = false;
    int groups = 0;
    for (int j0 = 0; j0 < kb; j0 += PG) {
        panel_column<T, 0, TL>(st, L, j0, kb, m, P, sy, base, k0, ipiv, local_info, failed);
        panel_column<T, 1, TL>(st, L, j0, kb, m, P, sy, base, k0, ipiv, local_info, failed);
        panel_column<T, 2, TL>(st, L, j0, kb, m, P, sy, base, k0, ipiv, local_info, failed);
        panel_column<T, 3, TL>(st, L, j0, kb, m, P, sy, base, k0, ipiv, local_info, failed);
        // rotate: the next group's first column moves to register 0
#pragma unroll
        for (int q = 0; q < S::RPT; ++q) {
            T keep[PG];
#pragma unroll
            for (int c = 0; c < PG; ++c) keep[c] = st.v[q][c];
#pragma unroll
            for (int c = 0; c < KB - PG; ++c) st.v[q][c] = st.v[q][c + PG];
#pragma unroll
            for (int c = 0; c < PG; ++c) st.v[q][KB - PG + c] = keep[c];
        }
        ++groups;
    }
    __syncthreads();

    // After the rotations register s holds column (s + PG*groups) % KB
    // (column s itself for a full-width panel: KB / PG rotations).
    const int shift = (PG * groups) % KB;
#pragma unroll
    for (int q = 0; q < S::RPT; ++q)
        if (st.row[q] < m) {
            T *dst = A + (int64_t)st.row[q] * lda;
            if (vec) {
#pragma unroll
                for (int c = 0; c < KB; c += VE) {
                    V x;
#pragma unroll
                    for (int e = 0; e < VE; ++e) x[e] = st.v[q][c + e];
                    *reinterpret_cast<V *>(dst + c) = x;
                }
            } else {
#pragma unroll
                for (int s2 = 0; s2 < KB; ++s2) {
                    int c = (s2 + shift) % KB;
                    if (c < kb) dst[c] = st.v[q][s2];
                }
            }
        }
    // The row interchanges on the other columns.  Every workgroup knows
    // every pivot (L.pivs), so each one derives the move list itself:
    //  a. the rows that move: panel rows 0 .. min(kb, m) - 1, then each pivot
    //     row below the panel's diagonal block once, in the order of the
    //     columns that first chose it (lane j of wave 0 takes column j's
    //     pivot; a ballot gives its place in the list);
    //  b. the original row each of them receives: lane e holds the tag of
    //     the row at position mv_dst[e] and the kb interchanges are
    //     replayed on those tags;
    //  c. the gather itself, gather_cols() columns a pass, each column read
    //     by one workgroup before it is written; for B12's columns one
    //     thread a column then solves U12 = L11^{-1} B12 (unit lower,
    //     x_i -= l_ik x_k with k ascending for every i) between the reads
    //     and the writes.
    const int nd = min(kb, m);
    if (t < 64) {
        const int pj = t < kb ? L.pivs[t] : -1;
        bool keep = t < kb && pj >= kb;
#pragma unroll
        for (int x = 0; x < KB; ++x)
            if (x < kb) keep = keep && !(x < t && L.pivs[x] == pj);
        const unsigned long long mask = __ballot(keep);
        if (t < nd) L.mv_dst[t] = t;
        if (keep) L.mv_dst[nd + __popcll(mask & ((1ull << t) - 1ull))] = pj;
        if (t == 0) L.mv_cnt = nd + __popcll(mask);
    }
    __syncthreads();
    const int cnt = L.mv_cnt;
    if (t < 64) {
        const int pos = t < cnt ? L.mv_dst[t] : -1;
        const int pv = t < kb ? L.pivs[t] : 0;
        int tag = pos;
        for (int jj = 0; jj < kb; ++jj) {
            const int p = __builtin_amdgcn_readlane(pv, jj);
            if (p == jj) continue;
            // the lane holding position p (positions below kb are lanes)
            const int lp = p < nd ? p : __builtin_ctzll(__ballot(pos == p));
            const int tj = __builtin_amdgcn_readlane(tag, jj);
            const int tp = __builtin_amdgcn_readlane(tag, lp);
            tag = t == jj ? tp : (t == lp ? tj : tag);
        }
        L.mv_src[t] = t < cnt ? tag : 0;
        if (t >= cnt) L.mv_dst[t] = 0;
        if (w == 0 && t == 0 && local_info != 0 && *info == 0) *info = local_info;
    }
    __syncthreads();
    if (w == 0 && gx.pairs_out != nullptr && t < 64) {
        if (t == 0) gx.pairs_out[0] = cnt;
        gx.pairs_out[1 + t] = k0 + L.mv_dst[t];
        gx.pairs_out[1 + 2 * KB + t] = k0 + L.mv_src[t];
    }
    const int n1 = gx.c1 - gx.c0, ncols = n1 + (gx.c3 - gx.c2);
    constexpr int GCOLS = gather_cols<T>(), GPER = 2 * KB / (S::PT / GCOLS);  // columns a pass; moves per thread
    const bool u12 = gx.u12 != 0 && kb == KB;
    const int lane = t & 63, wave = t >> 6;
    // for the U12 solve: lane i holds row i of L11 (columns k < i)
    T lrow[KB];
#pragma unroll
    for (int k = 0; k < KB; ++k) lrow[k] = (u12 && lane < KB && k < lane) ? L.l11[lane][k] : T(0);
    // workgroup w takes a contiguous chunk of the columns
    const int chunk = (ncols + P - 1) / P;
    const int cbeg = w * chunk, cend = min(ncols, cbeg + chunk);
    for (int cb = cbeg; cb < cend; cb += GCOLS) {
        const int tc = t % GCOLS, cc = cb + tc, grp = t / GCOLS;
        const int col = cc >= cend ? -1 : (cc < n1 ? gx.c0 + cc : gx.c2 + (cc - n1));
        const bool ucol = u12 && cc >= n1 && col >= 0;  // a column of B12
        // a wave's threads share grp, so the rows are wave-uniform
        T val[GPER];
#pragma unroll
        for (int i = 0; i < GPER; ++i) {
            const int e = grp * GPER + i;
            const T *src = gx.G + (int64_t)__builtin_amdgcn_readfirstlane(L.mv_src[e]) * gx.ldg;
            val[i] = (col >= 0 && e < cnt) ? src[col] : T(0);
            // positions 0 .. KB - 1 are the first KB entries of the list
            if (ucol && e < KB) L.b12[e][tc] = val[i];
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < GPER; ++i) {
            const int e = grp * GPER + i;
            T *dst = gx.G + (int64_t)__builtin_amdgcn_readfirstlane(L.mv_dst[e]) * gx.ldg;
            if (col >= 0 && e < cnt && !(ucol && e < KB)) dst[col] = val[i];
        }
        if (u12) {
            // U12 = L11^{-1} B12, a column per wave: lane i holds x_i, and
            // step k broadcasts x_k (by v_readlane) to x_i -= l_ik x_k for
            // the lanes i > k -- every x_i takes its terms in ascending k.
            // Rows 0 .. KB - 1 are positions 0 .. KB - 1.
            for (int q = wave; q < GCOLS; q += S::PW) {
                const int cq = cb + q;
                if (cq >= cend || cq < n1) continue;
                T x = lane < KB ? L.b12[lane][q] : T(0);
#pragma unroll
                for (int k = 0; k < KB - 1; ++k) {
                    const T xk = bcast_lane(x, k);
                    x = lane > k ? fma(-lrow[k], xk, x) : x;
                }
                if (lane < KB) (gx.G + (int64_t)lane * gx.ldg)[gx.c2 + (cq - n1)] = x;
            }
            __syncthreads();  // the next pass rewrites b12
        }
    }
    if (w == 0 && gx.perm != nullptr) {
        const int pv = t < cnt ? gx.perm[L.mv_src[t]] : 0;
        __syncthreads();
        if (t < cnt) gx.perm[L.mv_dst[t]] = pv;
    }
}


// ---------------------------------------------------------------- tournament panel
//
// Panel with one exchange per panel instead of one per column (for the FP32
// factor of the mixed solver, which only has to precondition the FP64
// refinement):
//   tslu_local_kernel: every workgroup runs GEPP on its own rows only (no
//     row moves; chosen rows are marked) and writes the kb rows it chose,
//     with their original values, as candidates;
//   tslu_final_kernel: every workgroup runs GEPP on the P*kb candidates --
//     identical work in each, so no cross-workgroup wait -- which gives the
//     panel's pivot rows in order with their factored values (L11 \ U11),
//     turns the pivot order into LAPACK's sequential swaps, and computes the
//     multipliers of its own rows by the update sequence GEPP applies to a
//     row that is never chosen: l_c = (a_c - sum_{k<c} l_k u_kc) / u_cc.
// The result is the LU factorisation of the panel with the tournament's
// pivots: GEPP itself when the panel fits one workgroup, in general the
// row choice of CALU with a flat reduction tree (Grigori, Demmel and Xiang,
// SIAM J. Matrix Anal. Appl. 32(4), 2011).  Ties go to the lowest original
// row, NaN is chosen only when nothing else is left.

template <class T, int NR>
struct GeppRows {
    T v[NR][KB];
    int id[NR];   // panel-relative row, -1 = none
    int col[NR];  // column at which the row became a pivot, -1 = not yet
};

template <class T>
struct GeppLds {
    alignas(16) T slot_data[2][PW][KB];
    T slot_val[2][PW];
    int slot_id[2][PW];
    int piv_id[KB];
};

// Column J of GEPP over the rows the workgroup holds.
template <class T, int NR, int J>
__device__ __forceinline__ void gepp_column(GeppRows<T, NR> &st, GeppLds<T> &L) {
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6, par = J & 1;
    T a[NR];
    T tmax = T(-1);
#pragma unroll
    for (int q = 0; q < NR; ++q) {
        const T x = fabs(st.v[q][J]);
        a[q] = (st.id[q] >= 0 && st.col[q] < 0) ? (isnan(x) ? T(-0.5) : x) : T(-1);
        tmax = fmax(tmax, a[q]);
    }
    const T wmax = wave_max(tmax);
    int tid = INT_MAX;
#pragma unroll
    for (int q = 0; q < NR; ++q)
        if (a[q] == wmax && a[q] > T(-1)) tid = min(tid, st.id[q]);
    const int wid = wave_min(tid);
#pragma unroll
    for (int q = 0; q < NR; ++q)
        if (st.id[q] == wid && st.col[q] < 0) {
#pragma unroll
            for (int c = 0; c < KB; ++c) L.slot_data[par][wave][c] = st.v[q][c];  // whole row: vector stores
        }
    if (lane == 0) {
        L.slot_val[par][wave] = wmax;
        L.slot_id[par][wave] = wid;
    }
    __syncthreads();
    T best;
    int bid, bw;
    wg_select<T>(L.slot_val[par], L.slot_id[par], best, bid, bw);
    if (!(best > T(-1))) {  // no row left (uniform)
        if (t == 0) L.piv_id[J] = -1;
        return;
    }
    T pr[KB];
#pragma unroll
    for (int c = 0; c < KB; ++c) pr[c] = L.slot_data[par][bw][c];  // whole row: vector loads
#pragma unroll
    for (int q = 0; q < NR; ++q)
        if (st.id[q] == bid && st.col[q] < 0) st.col[q] = J;
    if (t == 0) L.piv_id[J] = bid;
    const T piv = pr[J];
    if (piv != T(0)) {
#pragma unroll
        for (int q = 0; q < NR; ++q)
            if (st.id[q] >= 0 && st.col[q] < 0) {
                const T l = st.v[q][J] / piv;
                st.v[q][J] = l;
#pragma unroll
                for (int c = J + 1; c < KB; ++c) st.v[q][c] = fma(-l, pr[c], st.v[q][c]);
            }
    }
}

template <class T, int NR, int J>
__device__ __forceinline__ void gepp_from(GeppRows<T, NR> &st, GeppLds<T> &L, int kb) {
    if constexpr (J < KB) {
        if (J < kb) {
            gepp_column<T, NR, J>(st, L);
            gepp_from<T, NR, J + 1>(st, L, kb);
        }
    }
}

// Scratch of the tournament panel: candidates of up to MAXP workgroups and
// the panel's first kb rows (original values).
template <class T>
struct TsluScratch {
    int32_t cand[MAXP * KB];
    alignas(16) T candv[MAXP * KB * KB];
    alignas(16) T top[KB * KB];
};

// A block of PANEL_ROWS panel rows staged in LDS: the global side is read
// and written row-contiguously (consecutive lanes, consecutive columns),
// the thread side row per thread (odd row stride: no bank conflicts).
constexpr int SLD = KB + 1;
template <class T>
__device__ inline void stage_in(T (*S)[SLD], const T *__restrict__ A, int64_t lda, int r0, int m, int kb) {
#pragma unroll 16
    for (int e = threadIdx.x; e < PANEL_ROWS * KB; e += PT) {
        const int row = e / KB, c = e % KB, gr = r0 + row;
        S[row][c] = (gr < m && c < kb) ? A[(int64_t)gr * lda + c] : T(0);
    }
}
template <class T>
__device__ inline void stage_out(const T (*S)[SLD], T *__restrict__ A, int64_t lda, int r0, int m, int kb) {
#pragma unroll 16
    for (int e = threadIdx.x; e < PANEL_ROWS * KB; e += PT) {
        const int row = e / KB, c = e % KB, gr = r0 + row;
        if (gr < m && c < kb) A[(int64_t)gr * lda + c] = S[row][c];
    }
}

template <class T>
__global__ __launch_bounds__(PT) void tslu_local_kernel(int m, int kb, const T *__restrict__ A, int64_t lda,
                                                         TsluScratch<T> *__restrict__ ts) {
    const int w = blockIdx.x, t = threadIdx.x;
    __shared__ GeppLds<T> L;
    __shared__ T S[PANEL_ROWS][SLD];
    GeppRows<T, RPT> st;
    stage_in<T>(S, A, lda, w * PANEL_ROWS, m, kb);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
        const int r = w * PANEL_ROWS + q * PT + t;
        st.id[q] = r < m ? r : -1;
        st.col[q] = -1;
#pragma unroll
        for (int c = 0; c < KB; ++c) st.v[q][c] = S[q * PT + t][c];
    }
    gepp_from<T, RPT, 0>(st, L, kb);
    __syncthreads();
    // candidates with their original values (still in S)
    for (int e = t; e < KB * KB; e += PT) {
        const int i = e / KB, c = e % KB;
        const int r = i < kb ? L.piv_id[i] : -1;
        if (c == 0) ts->cand[w * KB + i] = r;
        ts->candv[(w * KB + i) * KB + c] = r >= 0 ? S[r - w * PANEL_ROWS][c] : T(0);
    }
    if (w == 0)
        for (int e = t; e < KB * KB; e += PT) ts->top[e] = S[e / KB][e % KB];
}

template <class T, int NC>
__global__ __launch_bounds__(PT) void tslu_final_kernel(int m, int kb, T *__restrict__ A, int64_t lda, int k0,
                                                         int P, const TsluScratch<T> *__restrict__ ts,
                                                         int32_t *__restrict__ ipiv, int32_t *__restrict__ srcmap,
                                                         int32_t *__restrict__ moves, int32_t *__restrict__ info) {
    static_assert(sizeof(T) == 4, "the tournament panel is for FP32 factors");
    const int w = blockIdx.x, t = threadIdx.x, lane = t & 63, wave = t >> 6;
    __shared__ GeppLds<T> L;
    __shared__ __attribute__((aligned(16))) T F[KB][KB];  // the pivot rows, factored: L11 \ U11
    __shared__ int ent_pos[2 * KB], ent_row[2 * KB];
    __shared__ int s_nent;
    __shared__ T S[PANEL_ROWS][SLD];
    GeppRows<T, NC> st;
    for (int e = t; e < KB * KB; e += PT) (&F[0][0])[e] = T(0);
    const int ncand = P * KB;
#pragma unroll
    for (int q = 0; q < NC; ++q) {
        const int ci = q * PT + t;
        const int r = ci < ncand ? ts->cand[ci] : -1;
        st.id[q] = r;
        st.col[q] = -1;
        const float4 *src = reinterpret_cast<const float4 *>(ts->candv + (int64_t)(ci < ncand ? ci : 0) * KB);
#pragma unroll
        for (int c4 = 0; c4 < KB / 4; ++c4) {
            const float4 v4 = src[c4];
            st.v[q][4 * c4] = r >= 0 ? v4.x : T(0);
            st.v[q][4 * c4 + 1] = r >= 0 ? v4.y : T(0);
            st.v[q][4 * c4 + 2] = r >= 0 ? v4.z : T(0);
            st.v[q][4 * c4 + 3] = r >= 0 ? v4.w : T(0);
        }
    }
    // own rows, staged while the tournament runs
    stage_in<T>(S, A, lda, w * PANEL_ROWS, m, kb);
    gepp_from<T, NC, 0>(st, L, kb);
#pragma unroll
    for (int q = 0; q < NC; ++q)
        if (st.col[q] >= 0) {
#pragma unroll
            for (int c = 0; c < KB; ++c) F[st.col[q]][c] = st.v[q][c];
        }
    __syncthreads();
    // pivot order -> sequential swaps of positions j and cur(p_j); lane e of
    // wave 0 holds entry e of the (position, row) list of moved positions,
    // lane j holds pivot j
    if (wave == 0) {
        const int pj = lane < kb ? L.piv_id[lane] : -1;
        int epos = -1, erow = -1, nent = 0;
        for (int j = 0; j < kb; ++j) {
            const int r = __builtin_amdgcn_readlane(pj, j);
            const unsigned long long m1 = __ballot(erow == r);
            const int cur = m1 ? __builtin_amdgcn_readlane(epos, __ffsll((long long)m1) - 1) : r;
            const unsigned long long m2 = __ballot(epos == j);
            const int l2 = __ffsll((long long)m2) - 1;
            const int rj = m2 ? __builtin_amdgcn_readlane(erow, l2) : j;
            if (m2) {
                if (lane == l2) erow = r;
            } else {
                if (lane == nent) {
                    epos = j;
                    erow = r;
                }
                ++nent;
            }
            if (cur != j) {
                const unsigned long long m3 = __ballot(epos == cur);
                const int l3 = __ffsll((long long)m3) - 1;
                if (m3) {
                    if (lane == l3) erow = rj;
                } else {
                    if (lane == nent) {
                        epos = cur;
                        erow = rj;
                    }
                    ++nent;
                }
            }
            if (w == 0 && lane == 0) ipiv[k0 + j] = k0 + cur;
        }
        ent_pos[lane] = epos;
        ent_row[lane] = erow;
        if (lane == 0) s_nent = nent;
        if (w == 0) {
            if (epos >= 0) {
                srcmap[epos] = erow;
                moves[1 + lane] = epos;
            }
            if (lane == 0) moves[0] = nent;
            if (lane == 0) {
                int inf = 0;
                for (int j = 0; j < kb; ++j)
                    if (F[j][j] == T(0)) {
                        inf = k0 + j + 1;
                        break;
                    }
                if (inf != 0 && *info == 0) *info = inf;
            }
        }
    }
    __syncthreads();
    const int nent = s_nent;
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
        const int row = q * PT + t, pos = w * PANEL_ROWS + row;
        if (pos >= m) continue;
        if (pos < kb) {
#pragma unroll
            for (int c = 0; c < KB; ++c) S[row][c] = F[pos][c];
            continue;
        }
        int src = pos;  // a displaced top row, or the row itself
        for (int e = 0; e < nent; ++e)
            if (ent_pos[e] == pos) src = ent_row[e];
        const bool own = src == pos || src < 0 || src >= KB;  // displaced rows come from the first kb
        T x[KB];
#pragma unroll
        for (int c = 0; c < KB; ++c) x[c] = own ? S[row][c] : ts->top[src * KB + c];
        // GEPP's updates of a row never chosen, branch-free: l_c = a_c / u_cc,
        // skipped (no division, no update) when u_cc = 0 as in the panel
#pragma unroll
        for (int c = 0; c < KB; ++c) {
            const T u = F[c][c];
            const bool ok = c < kb && u != T(0);
            const T l = x[c] / (ok ? u : T(1));
            x[c] = ok ? l : x[c];
            const T use = ok ? l : T(0);
#pragma unroll
            for (int k = c + 1; k < KB; ++k) x[k] = fma(-use, F[c][k], x[k]);
        }
#pragma unroll
        for (int c = 0; c < KB; ++c) S[row][c] = x[c];
    }
    __syncthreads();
    stage_out<T>(S, A, lda, w * PANEL_ROWS, m, kb);
}

// ---------------------------------------------------------------- inner update

// The update inside an outer panel after a 32-column sub-panel, fused into
// one launch: with L11 the sub-panel's unit-lower 32 x 32 block at P0 and
// B12 the 32 x rest block to its right, U12 = L11^{-1} B12, then the
// workgroup's 32-row tile of A22 (rows 32 + 32 b ..) takes A22 -= L21 U12.
// Thread c owns column c: it solves U12's column c in registers (every
// workgroup solves all of U12 for itself, so no workgroup waits on
// another) and then updates column c of its tile.  L11 (transposed, so the
// coefficients of one step are consecutive) and the tile of L21 sit in LDS
// and are read as broadcasts.  U12 goes to `ust` (row stride ldst), not
// into B12: other workgroups may still be reading B12; the driver copies it
// into place once the outer panel is done (nothing reads it before).
template <class T>
constexpr int iu_rows() {
    return 32;
}
constexpr int IU_MAXREST = NB_OUTER - KB;
template <class T>
__global__ __launch_bounds__(256) void inner_update_kernel(int rest, int mrows, T *__restrict__ P0, int64_t lda,
                                                           T *__restrict__ ust, int64_t ldst) {
    critical_path_priority();
    static_assert(IU_MAXREST <= 256, "one thread per column");
    constexpr int ROWS = iu_rows<T>();
    __shared__ __attribute__((aligned(16))) T LsT[KB][KB];   // LsT[k][i] = l_ik
    __shared__ __attribute__((aligned(16))) T L2[ROWS][KB];  // the tile of L21
    const int t = threadIdx.x;
    const int64_t rt = KB + (int64_t)blockIdx.x * ROWS;  // first tile row, panel-relative
    for (int e = t; e < KB * KB; e += 256) {
        const int i = e / KB, k = e % KB;
        LsT[k][i] = P0[(int64_t)i * lda + k];
    }
    for (int e = t; e < ROWS * KB; e += 256) {
        const int i = e / KB, k = e % KB;
        L2[i][k] = rt + i < KB + mrows ? P0[(rt + i) * lda + k] : T(0);
    }
    const bool act = t < rest;
    T x[KB], pre[ROWS];
#pragma unroll
    for (int i = 0; i < KB; ++i) x[i] = act ? P0[(int64_t)i * lda + KB + t] : T(0);
    // this thread's column of the A22 tile, loaded now so its latency
    // overlaps the solve
#pragma unroll
    for (int r = 0; r < ROWS; ++r) pre[r] = act && rt + r < KB + mrows ? P0[(rt + r) * lda + KB + t] : T(0);
    __syncthreads();
    if (!act) return;
    // forward substitution with the unit lower L11, a column at a time:
    // x_i -= l_ik x_k, every x_i taking its terms in ascending k
#pragma unroll
    for (int k = 0; k < KB - 1; ++k)
#pragma unroll
        for (int i = k + 1; i < KB; ++i) x[i] = fma(-LsT[k][i], x[k], x[i]);
    if (blockIdx.x == 0)
#pragma unroll
        for (int i = 0; i < KB; ++i) ust[(int64_t)i * ldst + t] = x[i];
    // A22 tile -= L21 U12, each product summed in ascending k
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
        T acc = T(0);
#pragma unroll
        for (int k = 0; k < KB; ++k) acc = fma(L2[r][k], x[k], acc);
        if (rt + r < KB + mrows) P0[(rt + r) * lda + KB + t] = pre[r] - acc;
    }
}

// U12 blocks of an outer panel from the stash into place: element (r, c) of
// the nbw x nbw diagonal block with c at or beyond the end of r's sub-panel.
template <class T>
__global__ void inner_flush_kernel(int nbw, const T *__restrict__ ust, int64_t ldst, T *__restrict__ A11,
                                   int64_t lda) {
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= nbw * nbw) return;
    const int r = e / nbw, c = e % nbw;
    if (c >= (r / KB + 1) * KB) A11[(int64_t)r * lda + c] = ust[(int64_t)r * ldst + c];
}

// X[d] <- X[src[d]] for every d of the move list, for columns [c0, c1) and
// [c2, c3) of X (rows relative to X), and the same gather on the
// permutation vector when perm is not null (workgroup 0).  The move list
// is read once into LDS; all reads of a column finish before any write (one
// workgroup per 64 columns, barrier in between).
template <class T>
__global__ __launch_bounds__(256) void row_gather_kernel(T *__restrict__ X, int64_t ldx, int c0, int c1,
                                                         int c2, int c3, const int32_t *__restrict__ moves,
                                                         const int32_t *__restrict__ srcmap,
                                                         int32_t *__restrict__ perm, int32_t *__restrict__ pairs_out,
                                                         int row_off) {
    critical_path_priority();
    constexpr int MAXM = 2 * KB;
    constexpr int PER = MAXM / 4;  // moves per thread
    __shared__ int s_dst[MAXM], s_src[MAXM], s_cnt;
    const int t = threadIdx.x, cl = t & 63, grp = t >> 6;
    if (t < 64) {
        const int cnt = min(moves[0], MAXM);
        const int d = t < cnt ? moves[1 + t] : 0;
        s_dst[t] = d;
        s_src[t] = t < cnt ? srcmap[d] : 0;
        if (t == 0) s_cnt = cnt;
        if (pairs_out != nullptr && blockIdx.x == 0) {  // the move list, rows absolute, for deferred gathers
            if (t == 0) pairs_out[0] = cnt;
            pairs_out[1 + t] = row_off + d;
            pairs_out[1 + MAXM + t] = row_off + s_src[t];
        }
    }
    __syncthreads();
    const int cnt = s_cnt;
    const int n1 = c1 - c0;
    const int cc = blockIdx.x * 64 + cl;
    int col = -1;
    if (cc < n1)
        col = c0 + cc;
    else if (cc - n1 < c3 - c2)
        col = c2 + (cc - n1);
    T val[PER];
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int e = grp * PER + i;
        val[i] = T(0);
        if (col >= 0 && e < cnt) val[i] = X[(int64_t)s_src[e] * ldx + col];
    }
    int pv = 0;
    const bool pm = perm != nullptr && blockIdx.x == 0 && t < cnt;
    if (pm) pv = perm[s_src[t]];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int e = grp * PER + i;
        if (col >= 0 && e < cnt) X[(int64_t)s_dst[e] * ldx + col] = val[i];
    }
    if (pm) perm[s_dst[t]] = pv;
}

// The move lists of npanels sub-panels (pairs: cnt, dst[2 KB], src[2 KB],
// rows absolute, PAIR_STRIDE ints each), applied in order to columns
// [c0, c1) of X: the row interchanges an outer panel deferred for columns
// that another stream was still using.
constexpr int PAIR_STRIDE = 4 * KB + 4;
constexpr int PANELS_PER_BLOCK = NB_OUTER / KB;
// LDS footprint of a workgroup of the lookahead's trailing GEMM: 64000
// bytes keep the FP32 GEMM at two workgroups per CU, leaving a CU room for
// an FP32 panel workgroup (21 KB) or inner-update workgroup (33 KB); the
// FP64 GEMM tiles already hold that much, and a reserve would cost them
// half their occupancy (0: none).
template <class T>
constexpr size_t aux_gemm_lds_footprint() {
    return sizeof(T) == 4 ? 64000 : 0;
}
// The BF16 lookahead GEMM's workgroup (80 KB of tiles) held to one per CU.
constexpr size_t BF16_SHARING_LDS = 96 * 1024;
template <class T>
__global__ __launch_bounds__(256) void row_gather_multi_kernel(T *__restrict__ X, int64_t ldx, int c0, int c1,
                                                               const int32_t *__restrict__ pairs, int npanels) {
    critical_path_priority();
    constexpr int MAXM = 2 * KB;
    constexpr int PER = MAXM / 4;
    __shared__ int s_dst[MAXM], s_src[MAXM], s_cnt;
    const int t = threadIdx.x, cl = t & 63, grp = t >> 6;
    const int cc = c0 + blockIdx.x * 64 + cl;
    const int col = cc < c1 ? cc : -1;
    for (int p = 0; p < npanels; ++p) {
        const int32_t *pp = pairs + (int64_t)p * PAIR_STRIDE;
        if (t < 64) {
            const int cnt = min(pp[0], MAXM);
            s_dst[t] = t < cnt ? pp[1 + t] : 0;
            s_src[t] = t < cnt ? pp[1 + MAXM + t] : 0;
            if (t == 0) s_cnt = cnt;
        }
        __syncthreads();
        const int cnt = s_cnt;
        T val[PER];
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int e = grp * PER + i;
            val[i] = (col >= 0 && e < cnt) ? X[(int64_t)s_src[e] * ldx + col] : T(0);
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int e = grp * PER + i;
            if (col >= 0 && e < cnt) X[(int64_t)s_dst[e] * ldx + col] = val[i];
        }
        __syncthreads();
    }
}

// The same moves for rows [r0, r0 + nr), nr <= GATHER_ROWS, with the
// sub-panels' permutations composed first: M (LDS) maps each row to the
// row its final value comes from, built move list by move list (all of a
// list's sources read before its destinations are written, as the moves
// themselves); then every destination row takes its source's value in one
// load pass and one store pass, instead of a load, store and barrier round
// trip per sub-panel.  A workgroup takes 16 columns; thread t column t % 16
// and destinations t / 16, t / 16 + 16, ....  A row listed as destination
// in several lists is written several times with the same value.
constexpr int GATHER_ROWS = 16384;
constexpr int GATHER_COLS = 16;
template <class T>
__global__ __launch_bounds__(256) void row_gather_composed_kernel(T *__restrict__ X, int64_t ldx, int c0, int c1,
                                                                  int r0, int nr, const int32_t *__restrict__ pairs,
                                                                  int npanels) {
    critical_path_priority();
    constexpr int MAXM = 2 * KB, MAXD = PANELS_PER_BLOCK * MAXM, PER = MAXD / (256 / GATHER_COLS);
    __shared__ int M[GATHER_ROWS];
    __shared__ int sd[MAXD];
    const int t = threadIdx.x;
    for (int r = t; r < nr; r += 256) M[r] = r0 + r;
    __syncthreads();
    int tot = 0;
    for (int p = 0; p < npanels; ++p) {
        const int32_t *pp = pairs + (int64_t)p * PAIR_STRIDE;
        const int cnt = min(pp[0], MAXM);
        int v = 0, d = 0;
        if (t < cnt) {
            d = pp[1 + t];
            v = M[pp[1 + MAXM + t] - r0];
        }
        __syncthreads();
        if (t < cnt) {
            M[d - r0] = v;
            sd[tot + t] = d;
        }
        tot += cnt;
        __syncthreads();
    }
    const int cl = t % GATHER_COLS, g = t / GATHER_COLS;
    const int col = c0 + (int)blockIdx.x * GATHER_COLS + cl;
    const int cc = min(col, c1 - 1);
    T val[PER];
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int e = g + i * (256 / GATHER_COLS);
        const int src = e < tot ? M[sd[e] - r0] : r0;
        val[i] = X[(int64_t)src * ldx + cc];
    }
    // every load has returned before any thread stores (a destination can
    // be another entry's source)
#pragma unroll
    for (int i = 0; i < PER; ++i) asm volatile("" : "+v"(val[i]));
    __syncthreads();
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int e = g + i * (256 / GATHER_COLS);
        if (col < c1 && e < tot) X[(int64_t)sd[e] * ldx + col] = val[i];
    }
}

__global__ void iota_kernel(int32_t *p, int n) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = i;
}

__global__ void zero_u32_kernel(unsigned *p, int n) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 0u;
}

__global__ void finish_kernel(const unsigned *err, int32_t *info) {
    if (threadIdx.x == 0 && *err != 0u) *info = -1;
}

// Swap list -> permutation vector: perm[i] = original row at position i.
// The swaps are sequential (step k reads what step k' < k left at ipiv[k]),
// so one lane applies them to the vector held in LDS (n <= PERM_MAX_N);
// wave 0 loads the swap list 64 entries at a time and hands each entry to
// that lane by v_readlane, so no step waits for a global load.
constexpr int PERM_MAX_N = 32768;
__global__ __launch_bounds__(256) void ipiv_to_perm_kernel(const int32_t *__restrict__ ipiv, int32_t *__restrict__ perm,
                                                           int n) {
    __shared__ int32_t sp[PERM_MAX_N];
    for (int i = threadIdx.x; i < n; i += blockDim.x) sp[i] = i;
    __syncthreads();
    if (threadIdx.x < 64) {
        const int lane = threadIdx.x;
        for (int k0 = 0; k0 < n; k0 += 64) {
            const int chunk = ipiv[k0 + lane < n ? k0 + lane : n - 1];
            const int cnt = min(64, n - k0);
            for (int kk = 0; kk < cnt; ++kk) {
                const int k = k0 + kk;
                const int p = __builtin_amdgcn_readlane(chunk, kk);
                if (lane == 0 && p != k && p >= 0 && p < n) {
                    const int a = sp[k], b = sp[p];
                    sp[k] = b;
                    sp[p] = a;
                }
            }
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) perm[i] = sp[i];
}

// out[i] = in[perm[i]] for an n x ncols block; then copied back by the caller.
template <class T>
__global__ void row_permute_kernel(const T *__restrict__ in, int64_t ldi, T *__restrict__ out, int64_t ldo,
                                   const int32_t *__restrict__ perm, int n, int ncols) {
    int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (int64_t)n * ncols) return;
    int i = (int)(e / ncols), c = (int)(e % ncols);
    out[(int64_t)i * ldo + c] = in[(int64_t)perm[i] * ldi + c];
}

template <class T>
__global__ void copy2d_kernel(const T *__restrict__ in, int64_t ldi, T *__restrict__ out, int64_t ldo, int n,
                              int ncols) {
    int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (int64_t)n * ncols) return;
    int i = (int)(e / ncols), c = (int)(e % ncols);
    out[(int64_t)i * ldo + c] = in[(int64_t)i * ldi + c];
}

__global__ void set_i32_kernel(int32_t *p, int32_t v) {
    if (threadIdx.x == 0) *p = v;
}

template <class T>
void launch_panel(hipStream_t s, int m, int kb, T *A, int64_t lda, int k0, int32_t *ipiv, int32_t *info,
                  PanelSync<T> *sy, unsigned &base, bool tie_last, const PanelGather<T> &gx) {
    int P = (int)cdiv(m, PANEL_ROWS);
    if (P > MAXP) fail_arg("getrf: matrix too tall for the panel kernel");
    TimedRegion tr("getrf_panel", s);
    // An ordinary launch on the caller's stream.  The P <= 64 workgroups
    // (one CU each, of 256) wait on each other, which needs them all
    // resident at once: a workgroup that cannot start yet (its CU busy with
    // GEMM tiles of the other stream) starts when those tiles retire, and
    // every wait is bounded (an error, not a hang, if one never arrives).
    // A cooperative launch adds nothing to that here, and it runs on a
    // hardware queue of its own: each hand-over between it and the stream's
    // other kernels cost about 11 us (about 10 ms per FP32 factorisation at
    // n = 16384).
    if (tie_last)
        hipLaunchKernelGGL((getf2_panel_kernel<T, true>), dim3(P), dim3(PanelShape<T>::PT), 0, s, m, kb, A, lda, k0, ipiv, info,
                           sy, P, base, gx);
    else
        hipLaunchKernelGGL((getf2_panel_kernel<T, false>), dim3(P), dim3(PanelShape<T>::PT), 0, s, m, kb, A, lda, k0, ipiv, info,
                           sy, P, base, gx);
    LMX_LAUNCH_CHECK();
    base += (unsigned)kb;  // sequence numbers stay monotonic across launches
}

// The tournament panel of an m x kb panel (same outputs as launch_panel).
template <class T>
void launch_tslu(hipStream_t s, int m, int kb, T *A, int64_t lda, int k0, int32_t *ipiv, int32_t *srcmap,
                 int32_t *moves, int32_t *info, TsluScratch<T> *ts) {
    const int P = (int)cdiv(m, PANEL_ROWS);
    if (P > MAXP) fail_arg("getrf: matrix too tall for the panel kernel");
    TimedRegion tr("getrf_panel", s);
    hipLaunchKernelGGL(tslu_local_kernel<T>, dim3(P), dim3(PT), 0, s, m, kb, A, lda, ts);
    LMX_LAUNCH_CHECK();
    const int nc = (int)cdiv((int64_t)P * KB, PT);
#define LMX_TSLU(NC_)                                                                                              \
    hipLaunchKernelGGL((tslu_final_kernel<T, NC_>), dim3(P), dim3(PT), 0, s, m, kb, A, lda, k0, P, ts, ipiv, srcmap, \
                       moves, info)
    if (nc <= 1)
        LMX_TSLU(1);
    else if (nc <= 2)
        LMX_TSLU(2);
    else
        LMX_TSLU(4);
#undef LMX_TSLU
    LMX_LAUNCH_CHECK();
}

template <class T>
void gather_rows(hipStream_t s, T *X, int64_t ldx, int c0, int c1, int c2, int c3, const int32_t *moves,
                 const int32_t *srcmap, int32_t *perm, int32_t *pairs_out, int row_off) {
    const int ncols = (c1 - c0) + (c3 - c2);
    if (ncols <= 0 && !perm && !pairs_out) return;
    TimedRegion tr("laswp", s);
    hipLaunchKernelGGL(row_gather_kernel<T>, dim3((unsigned)std::max<int64_t>(1, cdiv(ncols, 64))), dim3(256), 0, s,
                       X, ldx, c0, c1, c2, c3, moves, srcmap, perm, pairs_out, row_off);
    LMX_LAUNCH_CHECK();
}

// The deferred moves of npanels sub-panels whose rows lie in [r0, r1);
// `composed` (on the critical path) takes the composed-permutation kernel,
// whose 66 KB of LDS a workgroup would otherwise crowd the panels' CUs.
template <class T>
void gather_rows_deferred(hipStream_t s, T *A, int64_t lda, int c0, int c1, int r0, int r1, const int32_t *pairs,
                          int npanels, bool composed) {
    if (c1 <= c0 || npanels <= 0) return;
    TimedRegion tr("laswp", s);
    if (composed && r1 - r0 <= GATHER_ROWS && npanels <= PANELS_PER_BLOCK) {
        hipLaunchKernelGGL(row_gather_composed_kernel<T>, dim3((unsigned)cdiv(c1 - c0, GATHER_COLS)), dim3(256), 0, s,
                           A, lda, c0, c1, r0, r1 - r0, pairs, npanels);
    } else {
        hipLaunchKernelGGL(row_gather_multi_kernel<T>, dim3((unsigned)cdiv(c1 - c0, 64)), dim3(256), 0, s, A, lda, c0,
                           c1, pairs, npanels);
    }
    LMX_LAUNCH_CHECK();
}

}  // namespace

void clear_info(hipStream_t s, int32_t *d_info) {
    hipLaunchKernelGGL(set_i32_kernel, dim3(1), dim3(64), 0, s, d_info, 0);
    LMX_LAUNCH_CHECK();
}

// getrf with optional permutation-vector output (d_perm may be null).
// A22 -= L21 U12 for the trailing updates: FP32 or FP64 MFMA tiles, or, for
// the BF16 factor (bf16 = true, FP32 storage), the BF16 tiles with each
// operand split in two BF16 parts.  `sharing`: the lookahead GEMM that runs
// beside the panels, with the tiles and LDS footprint measured for that
// (its workgroups per CU held down so that panel workgroups find room).
template <class T>
void trailing_update(hipStream_t s, bool bf16, int64_t m, int64_t n, int64_t k, const T *L21, int64_t lda,
                     const T *U12, T *A22, bool sharing) {
    if constexpr (sizeof(T) == 4) {
        if (bf16) {
            gemm_bf16x3(s, false, false, m, n, k, -1.0f, L21, lda, U12, lda, 1.0f, A22, lda, false,
                        sharing ? BF16_SHARING_LDS : 0);
            return;
        }
    }
    if (sharing)
        gemm_sharing<T>(s, m, n, k, T(-1), L21, lda, U12, lda, T(1), A22, lda, aux_gemm_lds_footprint<T>());
    else
        gemm(s, false, false, m, n, k, T(-1), L21, lda, U12, lda, T(1), A22, lda);
}

template <class T>
void getrf_impl(hipStream_t s, int64_t n, T *A, int64_t lda, int32_t *d_ipiv, int32_t *d_perm,
                int32_t *d_info, bool tournament, bool tie_last, bool bf16) {
    if (n < 0 || lda < std::max<int64_t>(1, n)) fail_arg("getrf: bad dimension or stride");
    if (n > INT32_MAX / 2) fail_arg("getrf: n too large");
    if (tournament && sizeof(T) != 4) fail_arg("getrf: tournament pivoting is for the FP32 factor");
    if (tournament && tie_last) fail_arg("getrf: the tie rule of lu needs column-by-column pivoting");
    if (bf16 && sizeof(T) != 4) fail_arg("getrf: BF16 updates are for the FP32-stored factor");
    clear_info(s, d_info);
    if (n == 0) return;
    TimedRegion trall("getrf", s, 2.0 / 3.0 * (double)n * (double)n * (double)n);
    const int N = (int)n;
    size_t sync_bytes = (sizeof(PanelSync<T>) + 255) & ~size_t(255);
    const size_t map_bytes = (sizeof(int32_t) * ((size_t)N + 2 * KB + 8) + 255) & ~size_t(255);
    const size_t stash_bytes = (sizeof(T) * NB_OUTER * NB_OUTER + 255) & ~size_t(255);
    const size_t inv_bytes = 2 * stash_bytes;  // L11^-1 of two outer panels (double-buffered)
    const size_t pair_bytes = (sizeof(int32_t) * 2 * PANELS_PER_BLOCK * PAIR_STRIDE + 255) & ~size_t(255);
    Scratch scr(sync_bytes + map_bytes + stash_bytes + inv_bytes + pair_bytes +
                    (tournament ? sizeof(TsluScratch<T>) : 0) + 256,
                s);
    auto *sy = scr.as<PanelSync<T>>();
    int32_t *srcmap = reinterpret_cast<int32_t *>(reinterpret_cast<char *>(scr.p) + sync_bytes);
    int32_t *moves = srcmap + N;
    char *cur = reinterpret_cast<char *>(scr.p) + sync_bytes + map_bytes;
    T *stash = reinterpret_cast<T *>(cur);
    T *Wall = reinterpret_cast<T *>(cur + stash_bytes);
    int32_t *pairs_all = reinterpret_cast<int32_t *>(cur + stash_bytes + inv_bytes);
    auto *ts = tournament ? reinterpret_cast<TsluScratch<T> *>(cur + stash_bytes + inv_bytes + pair_bytes) : nullptr;
    hipLaunchKernelGGL(zero_u32_kernel, dim3((unsigned)cdiv(sync_clear_words<T>(), 256)), dim3(256), 0, s,
                       (unsigned *)sy, sync_clear_words<T>());
    LMX_LAUNCH_CHECK();
    if (d_perm) {
        hipLaunchKernelGGL(iota_kernel, dim3((unsigned)cdiv(N, 256)), dim3(256), 0, s, d_perm, N);
        LMX_LAUNCH_CHECK();
    }
    // Lookahead: the caller's stream factors outer panel b + 1 while an
    // auxiliary stream applies panel b's update to the columns beyond
    // b + 1.  Row interchanges of a panel reach, on the caller's stream, its
    // own outer panel at once and the trailing columns once the auxiliary
    // stream is done with them; the columns of earlier panels (L) get them
    // on the auxiliary stream, after the GEMM that reads them.
    //   stream s:   panels(b); W = L11^-1; wait F(b-1); swaps(b) on [k0+nbw, N);
    //               record E(b); U12 and A22 -= L21 U12 on panel b+1's columns
    //   stream aux: wait E(b); swaps(b) on [0, k0); U12 and A22 -= L21 U12 on
    //               the columns beyond panel b+1; record F(b)
    AuxStream la;
    hipStream_t aux = la.aux;
    hipEvent_t evE = la.ready, evF = la.done;
    bool pending = false;  // evF marks the auxiliary stream's last block
    unsigned base = 0;
    for (int b = 0, k0 = 0; k0 < N; ++b, k0 += NB_OUTER) {
        const int nbw = std::min(NB_OUTER, N - k0);
        int32_t *pairs = pairs_all + (int64_t)(b & 1) * PANELS_PER_BLOCK * PAIR_STRIDE;
        int np = 0;
        for (int s0 = k0; s0 < k0 + nbw; s0 += KB, ++np) {
            const int sb = std::min(KB, k0 + nbw - s0);
            const int m = N - s0;
            T *P0 = A + (int64_t)s0 * lda + s0;
            // the row interchanges reach this outer panel's other columns
            // at once, the rest later (move list kept)
            int32_t *perm_s = d_perm ? d_perm + s0 : nullptr;
            int32_t *pairs_s = pairs + (int64_t)np * PAIR_STRIDE;
            bool tslu_done = false;
            if constexpr (sizeof(T) == 4) {
                if (tournament) {
                    launch_tslu<T>(s, m, sb, P0, lda, s0, d_ipiv, srcmap, moves, d_info, ts);
                    gather_rows<T>(s, A + (int64_t)s0 * lda, lda, k0, s0, s0 + sb, k0 + nbw, moves, srcmap, perm_s,
                                   pairs_s, s0);
                    tslu_done = true;
                }
            }
            if (!tslu_done) {
                const PanelGather<T> gx{A + (int64_t)s0 * lda, lda, k0, s0, s0 + sb, k0 + nbw, perm_s, pairs_s, 1};
                launch_panel<T>(s, m, sb, P0, lda, s0, d_ipiv, d_info, sy, base, tie_last, gx);
            }
            const int rest = k0 + nbw - (s0 + sb);
            if (rest > 0) {
                if (sb != KB) fail_arg("getrf: inner update needs a full sub-panel");  // rest > 0 implies sb = KB
                if (tslu_done) {
                    // U12 into the stash at (s0 - k0, s0 - k0 + sb) of the outer block
                    T *ust = stash + (int64_t)(s0 - k0) * NB_OUTER + (s0 - k0) + sb;
                    hipLaunchKernelGGL(inner_update_kernel<T>,
                                       dim3((unsigned)std::max<int64_t>(1, cdiv(m - sb, iu_rows<T>()))), dim3(256), 0,
                                       s, rest, m - sb, P0, lda, ust, (int64_t)NB_OUTER);
                    LMX_LAUNCH_CHECK();
                } else if (m > sb) {
                    // the panel kernel left U12 in place: A22 -= L21 U12
                    gemm(s, false, false, m - sb, rest, sb, T(-1), P0 + (int64_t)sb * lda, lda, P0 + sb, lda, T(1),
                            P0 + (int64_t)sb * lda + sb, lda);
                }
            }
        }
        if (nbw > KB && tournament) {
            hipLaunchKernelGGL(inner_flush_kernel<T>, dim3((unsigned)cdiv((int64_t)nbw * nbw, 256)), dim3(256), 0, s,
                               nbw, stash, (int64_t)NB_OUTER, A + (int64_t)k0 * lda + k0, lda);
            LMX_LAUNCH_CHECK();
        }
        const int trail = N - (k0 + nbw);
        T *A11 = A + (int64_t)k0 * lda + k0;
        T *A12 = A11 + nbw;
        T *W = Wall + (int64_t)(b & 1) * NB_OUTER * NB_OUTER;
        const int narrow = std::min(NB_OUTER, trail);
        // L11^-1 before the wait: it reads only this outer panel's L11 (final
        // once its sub-panels are done; the auxiliary stream never touches
        // these columns) and writes W of this panel's parity (last read by the
        // auxiliary stream two panels ago), so it fills the time the caller's
        // stream would otherwise spend waiting for the previous GEMM
        if (trail > 0) tri_inverse_build<T>(s, true, false, true, nbw, A11, lda, W);
        if (pending) LMX_HIP(hipStreamWaitEvent(s, evF, 0));
        // The auxiliary stream's part needs the row moves on the trailing
        // columns and W, not the next columns' solve and update, which touch
        // other columns: E goes before those, so that no event marker sits
        // between them and the next panel.
        if (trail > 0) gather_rows_deferred<T>(s, A, lda, k0 + nbw, N, k0, N, pairs, np, true);
        LMX_HIP(hipEventRecord(evE, s));
        LMX_HIP(hipStreamWaitEvent(aux, evE, 0));
        if (trail > 0) {
            tri_inverse_solve<T>(s, true, false, nbw, narrow, A11, lda, W, A12, lda);
            trailing_update<T>(s, bf16, trail, narrow, nbw, A11 + (int64_t)nbw * lda, lda, A12,
                               A12 + (int64_t)nbw * lda, false);
        }
        gather_rows_deferred<T>(aux, A, lda, 0, k0, k0, N, pairs, np, false);
        if (trail > narrow) {
            tri_inverse_solve<T>(aux, true, false, nbw, trail - narrow, A11, lda, W, A12 + narrow, lda);
            // at most two GEMM workgroups per CU, so every CU keeps room for
            // a panel workgroup of the caller's stream
            trailing_update<T>(aux, bf16, trail, trail - narrow, nbw, A11 + (int64_t)nbw * lda, lda, A12 + narrow,
                               A12 + (int64_t)nbw * lda + narrow, true);
        }
        LMX_HIP(hipEventRecord(evF, aux));
        pending = true;
    }
    if (pending) LMX_HIP(hipStreamWaitEvent(s, evF, 0));
    hipLaunchKernelGGL(finish_kernel, dim3(1), dim3(64), 0, s, &sy->error, d_info);
    LMX_LAUNCH_CHECK();
}

template <class T>
void getrf(hipStream_t s, int64_t n, T *A, int64_t lda, int32_t *d_ipiv, int32_t *d_info) {
    getrf_impl<T>(s, n, A, lda, d_ipiv, nullptr, d_info);
}

// Solve with a permutation vector: B <- U^{-1} L^{-1} P B.
template <class T>
void getrs_perm(hipStream_t s, int64_t n, int64_t nrhs, const T *LU, int64_t lda, const int32_t *d_perm, T *B,
                int64_t ldb, const T *WL, const T *WU) {
    if (n == 0 || nrhs == 0) return;
    TimedRegion tr("getrs", s);
    Scratch tmp((size_t)(n * nrhs) * sizeof(T), s);
    unsigned g = (unsigned)cdiv(n * nrhs, 256);
    hipLaunchKernelGGL(row_permute_kernel<T>, dim3(g), dim3(256), 0, s, B, ldb, tmp.as<T>(), nrhs, d_perm,
                       (int)n, (int)nrhs);
    LMX_LAUNCH_CHECK();
    hipLaunchKernelGGL(copy2d_kernel<T>, dim3(g), dim3(256), 0, s, tmp.as<T>(), nrhs, B, ldb, (int)n,
                       (int)nrhs);
    LMX_LAUNCH_CHECK();
    if (WL && WU) {
        tri_inverse_solve<T>(s, true, false, n, nrhs, LU, lda, WL, B, ldb);
        tri_inverse_solve<T>(s, false, false, n, nrhs, LU, lda, WU, B, ldb);
        return;
    }
    trsm_left<T>(s, true, false, true, n, nrhs, LU, lda, B, ldb);
    trsm_left<T>(s, false, false, false, n, nrhs, LU, lda, B, ldb);
}

template <class T>
void getrs(hipStream_t s, int64_t n, int64_t nrhs, const T *LU, int64_t lda, const int32_t *d_ipiv, T *B,
           int64_t ldb) {
    if (n == 0 || nrhs == 0) return;
    if (n > PERM_MAX_N) fail_arg("getrs_dev: n above 32768 (the largest order getrf factors)");
    Scratch perm((size_t)n * sizeof(int32_t), s);
    hipLaunchKernelGGL(ipiv_to_perm_kernel, dim3(1), dim3(256), 0, s, d_ipiv, perm.as<int32_t>(), (int)n);
    LMX_LAUNCH_CHECK();
    getrs_perm<T>(s, n, nrhs, LU, lda, perm.as<int32_t>(), B, ldb);
}

template void getrf<double>(hipStream_t, int64_t, double *, int64_t, int32_t *, int32_t *);
template void getrf<float>(hipStream_t, int64_t, float *, int64_t, int32_t *, int32_t *);
template void getrf_impl<double>(hipStream_t, int64_t, double *, int64_t, int32_t *, int32_t *, int32_t *, bool,
                                 bool, bool);
template void getrf_impl<float>(hipStream_t, int64_t, float *, int64_t, int32_t *, int32_t *, int32_t *, bool,
                                bool, bool);
template void getrs<double>(hipStream_t, int64_t, int64_t, const double *, int64_t, const int32_t *, double *,
                            int64_t);
template void getrs<float>(hipStream_t, int64_t, int64_t, const float *, int64_t, const int32_t *, float *,
                           int64_t);
template void getrs_perm<double>(hipStream_t, int64_t, int64_t, const double *, int64_t, const int32_t *,
                                 double *, int64_t, const double *, const double *);
template void getrs_perm<float>(hipStream_t, int64_t, int64_t, const float *, int64_t, const int32_t *, float *,
                                int64_t, const float *, const float *);

}  // namespace lmx

using namespace lmx;

namespace {

void check_square(int64_t n, int64_t lda, const char *who) {
    if (n < 0) fail_arg(std::string(who) + ": negative dimension");
    if (lda < std::max<int64_t>(1, n)) fail_arg(std::string(who) + ": lda too small");
}

void check_rhs(int64_t nrhs, int64_t ldb, const char *who) {
    if (nrhs < 0) fail_arg(std::string(who) + ": negative nrhs");
    if (ldb < std::max<int64_t>(1, nrhs)) fail_arg(std::string(who) + ": ldb too small");
}

size_t extent(int64_t rows, int64_t cols, int64_t ld) {
    return (rows > 0 && cols > 0) ? (size_t)((rows - 1) * ld + cols) * sizeof(double) : 0;
}

int read_info(hipStream_t s, const int32_t *d_info) {
    int32_t info = 0;
    LMX_HIP(hipMemcpyAsync(&info, d_info, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    LMX_HIP(hipStreamSynchronize(s));
    if (info < 0) throw Error(LMX_ERR_INTERNAL, "panel synchronisation timed out");
    return info;
}

template <class T>
int getrf_host(int64_t n, T *A, int64_t lda, int32_t *ipiv) {
    return guarded([&] {
        check_square(n, lda, "getrf");
        if (n == 0) return LMX_OK;
        Context &c = ctx();
        std::lock_guard<std::mutex> g(c.host_mu);
        size_t ab = extent(n, n, lda) / sizeof(double) * sizeof(T);
        T *dA = (T *)staging(c, 0, ab);
        int32_t *dp = (int32_t *)staging(c, 1, (size_t)n * 4 + 64);
        int32_t *di = dp + n;
        LMX_HIP(hipMemcpyAsync(dA, A, ab, hipMemcpyHostToDevice, c.stream));
        getrf<T>(c.stream, n, dA, lda, dp, di);
        LMX_HIP(hipMemcpyAsync(A, dA, ab, hipMemcpyDeviceToHost, c.stream));
        LMX_HIP(hipMemcpyAsync(ipiv, dp, (size_t)n * 4, hipMemcpyDeviceToHost, c.stream));
        return read_info(c.stream, di);
    });
}

// The generic lu: factors with the tie rule of maximumBy and returns the
// permutation vector instead of the swap list.
template <class T>
void lu_dev(hipStream_t s, int64_t n, T *dA, int64_t lda, int32_t *d_perm, int32_t *d_info) {
    Scratch ipiv((size_t)std::max<int64_t>(n, 1) * sizeof(int32_t), s);
    getrf_impl<T>(s, n, dA, lda, ipiv.as<int32_t>(), d_perm, d_info, false, true);
}

template <class T>
int lu_host(int64_t n, T *A, int64_t lda, int32_t *perm) {
    return guarded([&] {
        check_square(n, lda, "lu");
        if (n == 0) return LMX_OK;
        Context &c = ctx();
        std::lock_guard<std::mutex> g(c.host_mu);
        size_t ab = extent(n, n, lda) / sizeof(double) * sizeof(T);
        T *dA = (T *)staging(c, 0, ab);
        int32_t *dp = (int32_t *)staging(c, 1, (size_t)n * 4 + 64);
        int32_t *di = dp + n;
        LMX_HIP(hipMemcpyAsync(dA, A, ab, hipMemcpyHostToDevice, c.stream));
        lu_dev<T>(c.stream, n, dA, lda, dp, di);
        LMX_HIP(hipMemcpyAsync(A, dA, ab, hipMemcpyDeviceToHost, c.stream));
        LMX_HIP(hipMemcpyAsync(perm, dp, (size_t)n * 4, hipMemcpyDeviceToHost, c.stream));
        return read_info(c.stream, di);
    });
}

}  // namespace

extern "C" {

int lmx_dlu(int64_t n, double *A, int64_t lda, int32_t *perm) { return lu_host<double>(n, A, lda, perm); }

int lmx_slu(int64_t n, float *A, int64_t lda, int32_t *perm) { return lu_host<float>(n, A, lda, perm); }

int lmx_dlu_dev(void *stream, int64_t n, double *dA, int64_t lda, int32_t *d_perm, int32_t *d_info) {
    return guarded([&] {
        check_square(n, lda, "lu");
        lu_dev<double>(resolve_stream(stream), n, dA, lda, d_perm, d_info);
        return LMX_OK;
    });
}

int lmx_slu_dev(void *stream, int64_t n, float *dA, int64_t lda, int32_t *d_perm, int32_t *d_info) {
    return guarded([&] {
        check_square(n, lda, "lu");
        lu_dev<float>(resolve_stream(stream), n, dA, lda, d_perm, d_info);
        return LMX_OK;
    });
}

int lmx_dgetrf_dev(void *stream, int64_t n, double *dA, int64_t lda, int32_t *d_ipiv, int32_t *d_info) {
    return guarded([&] {
        check_square(n, lda, "getrf");
        getrf<double>(resolve_stream(stream), n, dA, lda, d_ipiv, d_info);
        return LMX_OK;
    });
}

int lmx_dgetrs_dev(void *stream, int64_t n, int64_t nrhs, const double *dLU, int64_t lda, const int32_t *d_ipiv,
                   double *dB, int64_t ldb) {
    return guarded([&] {
        check_square(n, lda, "getrs");
        check_rhs(nrhs, ldb, "getrs");
        getrs<double>(resolve_stream(stream), n, nrhs, dLU, lda, d_ipiv, dB, ldb);
        return LMX_OK;
    });
}

int lmx_dgesv_dev(void *stream, int64_t n, int64_t nrhs, double *dA, int64_t lda, int32_t *d_ipiv, double *dB,
                  int64_t ldb, int32_t *d_info) {
    return guarded([&] {
        check_square(n, lda, "gesv");
        check_rhs(nrhs, ldb, "gesv");
        hipStream_t s = resolve_stream(stream);
        Scratch perm((size_t)std::max<int64_t>(n, 1) * sizeof(int32_t), s);
        getrf_impl<double>(s, n, dA, lda, d_ipiv, perm.as<int32_t>(), d_info);
        getrs_perm<double>(s, n, nrhs, dA, lda, perm.as<int32_t>(), dB, ldb);
        return LMX_OK;
    });
}

int lmx_dgetrf(int64_t n, double *A, int64_t lda, int32_t *ipiv) { return getrf_host<double>(n, A, lda, ipiv); }

int lmx_sgetrf(int64_t n, float *A, int64_t lda, int32_t *ipiv) { return getrf_host<float>(n, A, lda, ipiv); }

int lmx_sgetrf_dev(void *stream, int64_t n, float *dA, int64_t lda, int32_t *d_ipiv, int32_t *d_info) {
    return guarded([&] {
        check_square(n, lda, "getrf");
        getrf<float>(resolve_stream(stream), n, dA, lda, d_ipiv, d_info);
        return LMX_OK;
    });
}

int lmx_dgetrs(int64_t n, int64_t nrhs, const double *LU, int64_t lda, const int32_t *ipiv, double *B,
               int64_t ldb) {
    return guarded([&] {
        check_square(n, lda, "getrs");
        check_rhs(nrhs, ldb, "getrs");
        if (n == 0 || nrhs == 0) return LMX_OK;
        for (int64_t k = 0; k < n; ++k)
            if (ipiv[k] < k || ipiv[k] >= n) fail_arg("getrs: ipiv entry out of range");
        // swap list -> permutation on the host: O(n)
        std::vector<int32_t> perm((size_t)n);
        for (int64_t i = 0; i < n; ++i) perm[(size_t)i] = (int32_t)i;
        for (int64_t k = 0; k < n; ++k) std::swap(perm[(size_t)k], perm[(size_t)ipiv[k]]);
        Context &c = ctx();
        std::lock_guard<std::mutex> g(c.host_mu);
        size_t ab = extent(n, n, lda), bb = extent(n, nrhs, ldb);
        double *dA = (double *)staging(c, 0, ab), *dB = (double *)staging(c, 1, bb);
        int32_t *dp = (int32_t *)staging(c, 2, (size_t)n * 4);
        LMX_HIP(hipMemcpyAsync(dA, LU, ab, hipMemcpyHostToDevice, c.stream));
        LMX_HIP(hipMemcpyAsync(dB, B, bb, hipMemcpyHostToDevice, c.stream));
        LMX_HIP(hipMemcpyAsync(dp, perm.data(), (size_t)n * 4, hipMemcpyHostToDevice, c.stream));
        getrs_perm<double>(c.stream, n, nrhs, dA, lda, dp, dB, ldb);
        LMX_HIP(hipMemcpyAsync(B, dB, bb, hipMemcpyDeviceToHost, c.stream));
        LMX_HIP(hipStreamSynchronize(c.stream));
        return LMX_OK;
    });
}

int lmx_dgesv(int64_t n, int64_t nrhs, double *A, int64_t lda, int32_t *ipiv, double *B, int64_t ldb) {
    return guarded([&] {
        check_square(n, lda, "gesv");
        check_rhs(nrhs, ldb, "gesv");
        if (n == 0) return LMX_OK;
        Context &c = ctx();
        std::lock_guard<std::mutex> g(c.host_mu);
        size_t ab = extent(n, n, lda), bb = extent(n, nrhs, ldb);
        double *dA = (double *)staging(c, 0, ab), *dB = (double *)staging(c, 1, bb);
        int32_t *dp = (int32_t *)staging(c, 2, (size_t)n * 8 + 64);
        int32_t *dperm = dp + n, *di = dperm + n;
        LMX_HIP(hipMemcpyAsync(dA, A, ab, hipMemcpyHostToDevice, c.stream));
        if (bb) LMX_HIP(hipMemcpyAsync(dB, B, bb, hipMemcpyHostToDevice, c.stream));
        getrf_impl<double>(c.stream, n, dA, lda, dp, dperm, di);
        getrs_perm<double>(c.stream, n, nrhs, dA, lda, dperm, dB, ldb);
        LMX_HIP(hipMemcpyAsync(A, dA, ab, hipMemcpyDeviceToHost, c.stream));
        if (bb) LMX_HIP(hipMemcpyAsync(B, dB, bb, hipMemcpyDeviceToHost, c.stream));
        LMX_HIP(hipMemcpyAsync(ipiv, dp, (size_t)n * 4, hipMemcpyDeviceToHost, c.stream));
        return read_info(c.stream, di);
    });
}

}  // extern "C"
