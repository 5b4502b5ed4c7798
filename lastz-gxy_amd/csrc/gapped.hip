// gapped.hip — anchor selection and banded 3-state affine y-drop DP
// (src/anchor.rs:25-69, src/gapped_extend.rs:78-463).
//
// DP kernel: one block of four wavefronts (256 lanes) per side-problem, rows
// swept in order.  Pruning in the reference is row-synchronous (the alive band
// of row i depends on the best score after the whole row,
// src/gapped_extend.rs:388-412), so the exact parallel form is row-parallel:
// the 256 lanes own contiguous segments of <= 5 cells of the row's window
// [jl, jh] (<= 1208 cells with HOXD70), the previous row lives in LDS, X and M
// come straight from it, and the in-row Y recurrence
//     Y[j] = max(M[j-1] - (O+E), Y[j-1] - E)
// becomes, with Z[j] = Y[j] + (j-jl)*E (offsets relative to the row window
// start jl, so they stay below 1280*E), an inclusive prefix max
//     Z[j] = max(Z[j-1], M[j-1] - (O+E) + (j-jl)*E),  Z[j0-1] = NEG + (j0-1-jl)*E,
// computed lane-locally, carried across lanes with a shuffle scan and across
// wavefronts through LDS.  max and constant shifts commute exactly on int32
// (no value comes near overflow), so every M/X/Y value — including the
// NEG-derived ones in dead cells — equals the reference's, and so do the
// from-states, best cell and alive band.  Four wavefronts per problem keep the
// per-row latency short (the sweep is sequential in rows) and give the SIMDs
// enough wavefronts to hide it.
//
// Traceback: each row stores 4 bits per cell (M-from: 2 bits, X-from: 1 bit,
// Y-from: 1 bit), packed in window order (640 bytes for the widest row), plus
// the row window (jl, jh).  A second kernel walks back from the best cell, one
// wavefront per problem, and emits run-length edit runs tip-to-root.
#include <algorithm>

#include <thread>

#include "internal.h"

using namespace gxy;

namespace {

constexpr int kMaxW = 1280;                 // widest row window
constexpr int kDpThreads = 256;             // lanes per side-problem
constexpr int kDpWaves = kDpThreads / 64;
constexpr int kCMax = kMaxW / kDpThreads;   // cells per lane
// Traceback rows hold 4-bit from-states, one per cell, in one of two layouts
// (DpRes::fmt): k_dp_forward packs them in window order, cell idx = j - jl at
// bits [4 idx, 4 idx + 4) (lane t's five cells are bits [20 t, 20 t + 20)),
// 256 lanes x 20 bits = 640 bytes; k_dp_fwd1 stores them lane-major, 96 bits
// per lane (see there), 768 bytes.  Rows sit kRowBytes apart in the pool.
constexpr int kRowWordsW4 = kDpThreads * kCMax * 4 / 32;
static_assert(kRowWordsW4 % kDpWaves == 0 && kRowWordsW4 / kDpWaves * 32 == 64 * kCMax * 4,
              "a wavefront's lanes fill whole words of the packed row");
constexpr int kRowWords = 64 * 3;           // lane-major layout: three words per lane
constexpr int kRowBytes = kRowWords * 4;
static_assert(kRowWordsW4 <= kRowWords, "both layouts fit one row");
constexpr int kRecs = kMaxW + 2;            // previous-row records in LDS: [0] NEG, [1, 1+W) the row, NEG after
constexpr int kQCache = 4096;               // query columns cached in LDS, one byte each
constexpr int kQMargin = 512;               // cached columns kept left of the window on a refill
constexpr int kChunkRows = 64;              // traceback rows per pool chunk
constexpr uint64_t kChunkBytes = (uint64_t)kChunkRows * kRowBytes;
constexpr uint32_t kNoChunk = 0xFFFFFFFFu;
constexpr int kTileRows = 8;                // traceback rows staged in LDS by the walk
constexpr int kRowLoads = kRowBytes / 16;    // lanes that load a packed row, 16 bytes each

enum : uint32_t { kStM = 0, kStX = 1, kStY = 2 };

struct DpProb {
    const Blk *T;
    const Blk *Q;
    uint64_t t0, q0;   // position of T[0] / Q[0]
    int32_t tdir, qdir;
    uint32_t n, m;
};

struct DpRes {
    int32_t score;
    uint32_t bi, bj, bstate;
    uint32_t rows;       // rows swept
    uint32_t truncated;  // 1 when the traceback pool ran out before the sweep ended
    uint32_t last_chunk; // pool chunk holding the last swept row
    uint32_t fmt;        // traceback row layout: 0 window order, 1 + log2(K) lane-major (k_dp_fwd1<K>)
    uint64_t cells;
};

// Traceback store shared by the blocks of one launch: 64-row chunks handed out
// in order by an atomic counter; each chunk records the one before it in its
// side's sweep, so a walk from the last row back to row 0 follows `prev`.
struct DpPool {
    uint8_t *tb;       // kChunkBytes per chunk: kRowBytes of packed from-states per row
    uint2 *win;        // (jl, jh) per row, kChunkRows per chunk
    uint32_t *prev;    // previous chunk of the same side (kNoChunk for the first)
    uint32_t *next;    // allocation counter
    uint32_t nchunks;
};

// Inclusive prefix max over the 64 lanes of a wavefront: shifts of 1, 2, 4, 8
// within each 16-lane row, then the row-15 and row-31 broadcasts.  Called with
// every lane active.
__device__ __forceinline__ int32_t wave_scan_max(int32_t v) {
    v = max(v, __builtin_amdgcn_update_dpp(INT32_MIN, v, 0x111, 0xF, 0xF, false));  // row_shr:1
    v = max(v, __builtin_amdgcn_update_dpp(INT32_MIN, v, 0x112, 0xF, 0xF, false));  // row_shr:2
    v = max(v, __builtin_amdgcn_update_dpp(INT32_MIN, v, 0x114, 0xF, 0xF, false));  // row_shr:4
    v = max(v, __builtin_amdgcn_update_dpp(INT32_MIN, v, 0x118, 0xF, 0xF, false));  // row_shr:8
    v = max(v, __builtin_amdgcn_update_dpp(INT32_MIN, v, 0x142, 0xA, 0xF, false));  // row_bcast:15
    v = max(v, __builtin_amdgcn_update_dpp(INT32_MIN, v, 0x143, 0xC, 0xF, false));  // row_bcast:31
    return v;
}

// The same in place: each step is one DPP max whose missing source lanes
// keep their own value (v_max_i32_dpp with destination = both sources), with
// the two wait states a DPP read of a just-written VGPR needs.  Every lane
// must be active.
__device__ __forceinline__ int32_t wave_scan_max_self(int32_t v) {
    asm volatile(
        "s_nop 1\n"
        "v_max_i32_dpp %0, %0, %0 row_shr:1 row_mask:0xf bank_mask:0xf\n"
        "s_nop 1\n"
        "v_max_i32_dpp %0, %0, %0 row_shr:2 row_mask:0xf bank_mask:0xf\n"
        "s_nop 1\n"
        "v_max_i32_dpp %0, %0, %0 row_shr:4 row_mask:0xf bank_mask:0xf\n"
        "s_nop 1\n"
        "v_max_i32_dpp %0, %0, %0 row_shr:8 row_mask:0xf bank_mask:0xf\n"
        "s_nop 1\n"
        "v_max_i32_dpp %0, %0, %0 row_bcast:15 row_mask:0xa bank_mask:0xf\n"
        "s_nop 1\n"
        "v_max_i32_dpp %0, %0, %0 row_bcast:31 row_mask:0xc bank_mask:0xf\n"
        "s_nop 1\n"
        : "+v"(v));
    return v;
}

// Value of the previous lane (lane 0 gets `fill`).
__device__ __forceinline__ int32_t wave_shr1(int32_t v, int32_t fill) {
    return __builtin_amdgcn_update_dpp(fill, v, 0x138, 0xF, 0xF, false);  // wave_shr:1
}

// 32 bases at positions p0, p0+dir, ..., p0+31*dir in slots 0..31; for
// dir < 0 the slots past position 0 hold garbage.
__device__ __forceinline__ Win dir_win(const Blk *__restrict__ b, uint64_t p0, int32_t dir) {
    if (dir > 0) return load_win(b, p0);
    Win w;
    if (p0 >= 31) {
        const Win x = load_win(b, p0 - 31);
        w.codes = rev2(x.codes);
        w.valid = __builtin_bitreverse32(x.valid);
    } else {
        const Win x = load_win(b, 0);
        const uint32_t sh = 31 - (uint32_t)p0;
        w.codes = rev2(x.codes) >> (2 * sh);
        w.valid = __builtin_bitreverse32(x.valid) >> sh;
    }
    w.mask = 0;
    return w;
}

// Scores of one target base against a cached query byte qb = 8 * (code |
// 4 * invalid) (src/gapped_extend.rs:339-340 via pair_score).
struct RowScoreI8 {
    uint64_t tab;  // byte b: the score of query byte 8b
    __device__ __forceinline__ int32_t operator()(uint32_t qb) const { return (int32_t)(int8_t)(uint8_t)(tab >> qb); }
};
__device__ __forceinline__ RowScoreI8 row_scores(const ScorerI8 &s, uint32_t tc, uint32_t tv) {
    const uint32_t lo = (tc & 1u) ? s.r1 : s.r0;
    const uint32_t hi = (tc & 1u) ? s.r3 : s.r2;
    const uint32_t row = (tc & 2u) ? hi : lo;
    const uint32_t amb = (uint32_t)(s.ambig & 0xFF) * 0x01010101u;
    return {((uint64_t)amb << 32) | (tv ? row : amb)};
}
struct RowScoreI32 {
    int32_t s0, s1, s2, s3, amb;
    __device__ __forceinline__ int32_t operator()(uint32_t qb) const {
        const int32_t a = (qb & 8u) ? s1 : s0;
        const int32_t b = (qb & 8u) ? s3 : s2;
        const int32_t x = (qb & 16u) ? b : a;
        return (qb & 32u) ? amb : x;
    }
};
__device__ __forceinline__ RowScoreI32 row_scores(const ScorerI32 &s, uint32_t tc, uint32_t tv) {
    RowScoreI32 r;
    r.amb = s.ambig;
    const int32_t *row = s.s + 4 * tc;  // tc is uniform: no dynamic register indexing per lane
    int32_t v[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int32_t a = (tc & 1u) ? s.s[4 + q] : s.s[q];
        const int32_t b = (tc & 1u) ? s.s[12 + q] : s.s[8 + q];
        v[q] = tv ? ((tc & 2u) ? b : a) : s.ambig;
    }
    (void)row;
    r.s0 = v[0];
    r.s1 = v[1];
    r.s2 = v[2];
    r.s3 = v[3];
    return r;
}

// One side-problem per block of kDpThreads lanes (four wavefronts); lane t
// owns the contiguous cells [jl + 5t, jl + 5t + cnt) of the row window
// (C = kCMax = 5 cells per lane in every row, so the row loop has no uniform
// branches on the window width).  Per row:
//   * the previous row comes from LDS as one {M, X, Y} record per cell,
//     indexed by j - jl(prev) + 1 and clamped into [0, kRecs): record 0 and
//     the min(br, m) + 1 records after the previous row hold NEG, which is
//     what the reference reads outside the previous row's window;
//   * query bases come from a 4096-column LDS cache (refilled when the window
//     leaves it), target bases from a 32-row register window;
//   * Y by the max-plus prefix scan: each lane's M values feed the next
//     column, so lane totals need no neighbour; DPP scan inside a wavefront,
//     wavefront totals across them through LDS;
//   * row best (first strict maximum in j order) and alive band reduced
//     through LDS; the new row overwrites the old records.
// Three barriers per row.
template <class S>
__global__ __launch_bounds__(kDpThreads) __attribute__((amdgpu_waves_per_eu(5))) void k_dp_forward(const DpProb *__restrict__ probs,
                                                           const uint32_t *__restrict__ list, uint32_t nprob, S sc,
                                                           int32_t oe, int32_t e, int32_t ydrop, DpPool pool,
                                                           DpRes *__restrict__ res) {
    __shared__ int4 rec[kRecs];                       // .x = M, .y = X, .z = Y of the previous row
    __shared__ __align__(16) uint32_t qcache[kQCache / 4];  // query byte of column qc0 + x at byte x
    __shared__ int32_t w_tot[kDpWaves], w_alast[kDpWaves];
    __shared__ int32_t w_max[kDpWaves];
    __shared__ uint32_t w_j[kDpWaves], w_st[kDpWaves], w_lo[kDpWaves], w_hi[kDpWaves];
    __shared__ uint32_t s_chunk;
    const uint32_t p = blockIdx.x;
    if (p >= nprob) return;  // uniform per block
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const DpProb pr = probs[list[p]];
    const uint32_t n = pr.n, m = pr.m;
    DpRes out = {0, 0, 0, kStM, 0, 0, kNoChunk, 0, 0};
    if (n == 0 || m == 0) {  // src/gapped_extend.rs:188-195
        if (t == 0) res[p] = out;
        return;
    }
    // band radius / width (src/gapped_extend.rs:214-233)
    uint32_t br;
    if (e > 0) {
        const int32_t r = sat_sub(ydrop, oe) / e;
        br = (uint32_t)(max(r, 0) + 2);
    } else {
        br = m + 1;
    }
    const uint64_t bw64 = min((uint64_t)4 * br + 4, (uint64_t)m + 1);
    const uint32_t bw = (uint32_t)bw64;  // host guarantees bw <= kMaxW
    // a row window reaches at most min(br, m) columns right of the previous one
    const uint32_t padr = min(br, m) + 1;
    const int4 negr = make_int4(kNegInf, kNegInf, kNegInf, 0);
    for (int x = t; x < kRecs; x += kDpThreads) rec[x] = negr;
    // Y-from of the window's first cell: its left neighbour is outside the
    // window, so both candidates are NEG-based (src/gapped_extend.rs:318-334)
    const uint32_t ybit_lead = (kNegInf - oe >= kNegInf - e) ? 0u : 1u;
    const int32_t lead_y = max(kNegInf - oe, kNegInf - e);

    int32_t best_score = 0;
    uint32_t bi = 0, bj = 0, bst = kStM;
    uint32_t lo = 0, hi = 0;
    uint32_t pjl = 0;
    uint32_t qc0 = 0xFFFFFFFFu;  // first cached query column (none yet)
    Win tw = {0, 0, 0};          // target bases of rows 32g+1 .. 32g+32
    uint64_t cells = 0;
    uint32_t rows_done = 0;
    uint32_t chunk = kNoChunk;
    const uint8_t *qbytes = (const uint8_t *)qcache;

    uint32_t truncated = 0;
    for (uint32_t i = 0; i <= n; i++) {
        // a fresh traceback chunk every kChunkRows rows; read back after barrier (1)
        const bool newc = (i % kChunkRows) == 0;
        if (newc && t == 0) {
            const uint32_t c = atomicAdd(pool.next, 1u);
            if (c < pool.nchunks) pool.prev[c] = chunk;
            s_chunk = c;
        }
        const uint32_t jl = lo >= br ? lo - br : 0;
        const uint32_t jh = (uint32_t)min(min((uint64_t)hi + br, (uint64_t)m), (uint64_t)jl + bw - 1);
        const uint32_t W = jh - jl + 1;
        // every lane takes kCMax cells (lanes past the window compute padding
        // cells that are masked out): no per-row uniform branches on C
        constexpr uint32_t C = kCMax;
        const uint32_t c0 = (uint32_t)t * C;
        const uint32_t cnt = c0 < W ? min(C, W - c0) : 0;
        const uint32_t js = jl + c0;  // first cell of this lane
        cells += W - ((i == 0 && jl == 0) ? 1 : 0);

        // query columns [max(jl,1)-1, jh-1] must be cached (uniform)
        if (jh >= 1) {
            const uint32_t need = (jl >= 1 ? jl : 1) - 1;
            if (qc0 == 0xFFFFFFFFu || need < qc0 || jh - 1 >= qc0 + kQCache) {
                qc0 = need > (uint32_t)kQMargin ? need - kQMargin : 0;
                const uint32_t c = qc0 + 16u * (uint32_t)t;
                if (c < m) {
                    const Win w = dir_win(pr.Q, pr.qdir > 0 ? pr.q0 + c : pr.q0 - c, pr.qdir);
                    uint32_t wd[4];
#pragma unroll
                    for (int x = 0; x < 4; x++) {
                        uint32_t v = 0;
#pragma unroll
                        for (int y = 0; y < 4; y++) {
                            const int k = 4 * x + y;
                            const uint32_t code = (uint32_t)(w.codes >> (2 * k)) & 3u;
                            const uint32_t inv = ((w.valid >> k) & 1u) ^ 1u;
                            v |= (8u * (code | (inv << 2))) << (8 * y);
                        }
                        wd[x] = v;
                    }
                    *(uint4 *)&qcache[4 * t] = make_uint4(wd[0], wd[1], wd[2], wd[3]);
                }
                __syncthreads();
            }
        }
        // target base of row i (T[i-1]); uniform
        uint32_t tc = 0, tv = 0;
        if (i >= 1) {
            const uint32_t r = (i - 1) & 31u;
            if (r == 0) tw = dir_win(pr.T, pr.tdir > 0 ? pr.t0 + (i - 1) : pr.t0 - (i - 1), pr.tdir);
            tc = (uint32_t)(tw.codes >> (2 * r)) & 3u;
            tv = (tw.valid >> r) & 1u;
        }
        const auto rs = row_scores(sc, tc, tv);

        // ---- X and M from the previous row ------------------------------------
        int32_t Mv[kCMax], Xv[kCMax], Yv[kCMax];
        uint32_t nib[kCMax];
        if (i >= 1) {
            int4 P[kCMax + 1];  // previous row at j = js - 1 + k
            const int32_t r0 = (int32_t)js - (int32_t)pjl;  // record index of j = js - 1
#pragma unroll
            for (int k = 0; k <= kCMax; k++)
                P[k] = rec[min(max(r0 + k, 0), kRecs - 1)];
#pragma unroll
            for (int k = 0; k < kCMax; k++) {
                if ((uint32_t)k < C) {
                    const uint32_t j = js + k;
                    const uint32_t jq = min(j, jh);  // cells past the window: any cached column
                    const uint32_t qb = qbytes[jq >= 1 ? jq - 1 - qc0 : 0];
                    const int4 a = P[k + 1], d = P[k];
                    // X (src/gapped_extend.rs:293-312)
                    const int32_t xm = a.x - oe, xx = a.y - e;
                    Xv[k] = max(xm, xx);
                    uint32_t nb = (xm >= xx ? 0u : 1u) << 2;
                    // M (src/gapped_extend.rs:339-363)
                    const int32_t dmax = max(max(d.x, d.y), d.z);
                    const uint32_t arg = d.x == dmax ? kStM : (d.y == dmax ? kStX : kStY);
                    int32_t mv = dmax + rs(qb);
                    if (k == 0 && js == 0) {  // column 0 has no M
                        mv = kNegInf;
                    } else {
                        nb |= arg;
                    }
                    Mv[k] = mv;
                    nib[k] = nb;
                }
            }
        } else {
#pragma unroll
            for (int k = 0; k < kCMax; k++) {
                Xv[k] = kNegInf;  // row 0: no X; M only at the seed cell (src/gapped_extend.rs:245)
                Mv[k] = (k == 0 && js == 0) ? 0 : kNegInf;
                nib[k] = 0;
            }
        }

        // ---- Y as a max-plus prefix scan --------------------------------------
        // Z[j] = Y[j] + (j-jl)*E = max(Z[j-1], a[j]),  a[j] = M[j-1] - (O+E) + (j-jl)*E;
        // the lane's cell k contributes a[js+k+1].
        const int32_t cterm = (int32_t)(js + 1 - jl) * e - oe;
        int32_t tot = INT32_MIN, alast = INT32_MIN;
#pragma unroll
        for (int k = 0; k < kCMax; k++) {
            if ((uint32_t)k < C && (uint32_t)k < cnt) {
                const int32_t a = Mv[k] + cterm + k * e;
                tot = max(tot, a);
                alast = a;
            }
        }
        const int32_t incl = wave_scan_max(tot);
        int32_t excl = wave_shr1(incl, INT32_MIN);
        int32_t aprev = wave_shr1(alast, INT32_MIN);
        if (lane == 63) {
            w_tot[wv] = incl;
            w_alast[wv] = alast;
        }
        __syncthreads();  // (1) wavefront totals of the Y scan; the row's chunk
        if (newc) {
            chunk = s_chunk;
            if (chunk >= pool.nchunks) {  // pool exhausted: the side is swept again later
                truncated = 1;
                break;
            }
        }
        for (int v = 0; v < wv; v++) excl = max(excl, w_tot[v]);
        if (lane == 0 && wv > 0) aprev = w_alast[wv - 1];
        // Z at the window's first cell: for jl >= 1 both of Y[jl]'s candidates
        // come from outside the window; for jl = 0, Y[0] is not computed
        const int32_t zinit = jl >= 1 ? lead_y : kNegInf;
        {
            int32_t z = max(excl, zinit);
#pragma unroll
            for (int k = 0; k < kCMax; k++) {
                if ((uint32_t)k < C) {
                    const uint32_t j = js + k;
                    uint32_t yb;
                    if (k == 0) {
                        yb = t == 0 ? ybit_lead : (z != aprev ? 1u : 0u);
                    } else {
                        const int32_t an = Mv[k - 1] + cterm + (k - 1) * e;
                        yb = an < z ? 1u : 0u;
                        z = max(z, an);
                    }
                    int32_t y = z - (int32_t)(j - jl) * e;
                    if (j == 0) {  // column 0 has no Y
                        y = kNegInf;
                        yb = 0;
                    }
                    Yv[k] = y;
                    nib[k] |= yb << 3;
                }
            }
        }

        // ---- row best (first strict maximum in j order) -------------------------
        int32_t rmax = INT32_MIN;
        int rk = 0;
#pragma unroll
        for (int k = 0; k < kCMax; k++) {
            if ((uint32_t)k < C) {
                const int32_t cb = max(max(Mv[k], Xv[k]), Yv[k]);
                if ((uint32_t)k < cnt && cb > rmax) {
                    rmax = cb;
                    rk = k;
                }
            }
        }
        {
            const int32_t wmax = __builtin_amdgcn_readlane(wave_scan_max(rmax), 63);
            const unsigned long long hasmax = __ballot(cnt > 0 && rmax == wmax);
            if (hasmax) {
                const int src = __ffsll((long long)hasmax) - 1;
                if (lane == src) {
                    uint32_t st = kStY;
#pragma unroll
                    for (int k = kCMax - 1; k >= 0; k--) {
                        if ((uint32_t)k < C && k == rk) st = Mv[k] == rmax ? kStM : (Xv[k] == rmax ? kStX : kStY);
                    }
                    w_max[wv] = wmax;
                    w_j[wv] = js + rk;
                    w_st[wv] = st;
                }
            } else if (lane == 0) {
                w_max[wv] = INT32_MIN;
            }
        }
        __syncthreads();  // (2) row best per wavefront
        {
            int32_t gmax = INT32_MIN;
            uint32_t gw = 0;
#pragma unroll
            for (int v = 0; v < kDpWaves; v++) {
                const int32_t x = w_max[v];
                if (x > gmax) {  // strict: the earliest wavefront holding the maximum
                    gmax = x;
                    gw = v;
                }
            }
            if (gmax > best_score) {  // src/gapped_extend.rs:365-380
                best_score = gmax;
                bi = i;
                bj = w_j[gw];
                bst = w_st[gw];
            }
        }
        const int32_t thr = sat_sub(best_score, ydrop);

        // ---- alive band, the new row into LDS, traceback -----------------------
        uint32_t alo = 0xFFFFFFFFu, ahi = 0;
        uint32_t word = 0;
#pragma unroll
        for (int k = 0; k < kCMax; k++) {
            if ((uint32_t)k < C && (uint32_t)k < cnt) {
                const uint32_t j = js + k;
                int4 r = make_int4(Mv[k], Xv[k], Yv[k], 0);
                if (max(max(Mv[k], Xv[k]), Yv[k]) >= thr) {  // src/gapped_extend.rs:388-406
                    alo = min(alo, j);
                    ahi = max(ahi, j);
                } else {
                    r = negr;
                }
                rec[1 + c0 + k] = r;
                word |= nib[k] << (4 * k);
            }
        }
        {
            const uint32_t clear_end = min((uint32_t)kRecs, 1 + W + padr);
            for (uint32_t x = 1 + W + (uint32_t)t; x < clear_end; x += kDpThreads) rec[x] = negr;
        }
        {
            const unsigned long long anyalive = __ballot(alo != 0xFFFFFFFFu);
            const int fl = anyalive ? __ffsll((long long)anyalive) - 1 : 0;
            const int ll = anyalive ? 63 - __clzll((long long)anyalive) : 0;
            const uint32_t wl = __builtin_amdgcn_readlane(alo, fl);
            const uint32_t wh = __builtin_amdgcn_readlane(ahi, ll);
            if (lane == 0) {
                w_lo[wv] = anyalive ? wl : 0xFFFFFFFFu;
                w_hi[wv] = wh;
            }
        }
        const uint32_t ir = i % kChunkRows;
        {
            // The packed row: lane t's 20 bits at bit 20 t, so a wavefront's 64
            // lanes fill exactly 40 words.  Lane w < 40 builds word w of its
            // wavefront from the lanes a = 32 w / 20, a + 1 and a + 2 that it
            // spans (empty lanes contribute 0).
            const uint32_t a = (32u * (uint32_t)lane) / 20u, o = 32u * (uint32_t)lane - 20u * a;
            const uint32_t v0 = (uint32_t)__shfl((int)word, (int)min(a, 63u));
            const uint32_t v1 = (uint32_t)__shfl((int)word, (int)min(a + 1u, 63u));
            const uint32_t v2 = (uint32_t)__shfl((int)word, (int)min(a + 2u, 63u));
            uint32_t pw = (v0 >> o) | (v1 << (20u - o));
            if (o > 8u) pw |= v2 << (40u - o);
            if (lane < kRowWordsW4 / kDpWaves)
                *(uint32_t *)(pool.tb + chunk * kChunkBytes + ir * kRowBytes +
                              ((uint32_t)wv * (kRowWordsW4 / kDpWaves) + (uint32_t)lane) * 4) = pw;
        }
        if (t == 0) pool.win[(uint64_t)chunk * kChunkRows + ir] = make_uint2(jl, jh);
        rows_done = i + 1;
        __syncthreads();  // (3) alive band; the new row is in LDS
        uint32_t nlo = 0xFFFFFFFFu, nhi = 0;
        bool alive = false;
#pragma unroll
        for (int v = 0; v < kDpWaves; v++) {
            const uint32_t l = w_lo[v];
            if (l != 0xFFFFFFFFu) {
                if (!alive) nlo = l;
                nhi = w_hi[v];
                alive = true;
            }
        }
        if (!alive) break;  // whole row dead (src/gapped_extend.rs:407-409)
        lo = nlo;
        hi = nhi;
        pjl = jl;
    }
    if (t == 0) {
        out.score = best_score;
        out.bi = bi;
        out.bj = bj;
        out.bstate = bst;
        out.rows = rows_done;
        out.truncated = truncated;
        out.last_chunk = chunk;
        out.cells = cells;
        res[p] = out;
    }
}

// ---- one wavefront per side, rows in registers --------------------------------
//
// k_dp_fwd1<S, K>: the same row recurrence as k_dp_forward, for one side per
// wavefront, with the row held in registers instead of LDS.  Columns map to
// lanes by absolute position, so a cell of the previous row sits in the same
// lane and register as the cell of this row below it: column j belongs to
// slot k = j / SW - B (SW = 64 K columns per slot, B = jl / SW the row's base
// block), lane (j % SW) / K and cell j % K of that lane.  NS slots cover any
// window of at most kMaxW columns wherever it starts.  Per lane and cell the
// registers carry, from the previous row:
//   dd = max(M, X, Y)                  (what M of the column to the right adds to)
//   xn = max(M - (O+E), X - E)         (X of the same column in this row)
//   sb = arg of dd (2 bits) | from-bit of xn << 2
// with cells outside the previous row's window or dead at its end holding the
// values the reference reads there (NEG for M, X and Y).  The window moves
// right as the alignment proceeds: when jl crosses a slot boundary the slots
// shift down by one (rarely up, when the band grows left), once per SW rows.
//
// Y is the max-plus prefix scan of k_dp_forward with offsets relative to the
// row's base column B * SW: each cell contributes c = M - (O+E) + (j + 1 -
// B SW) E to every cell to its right, Z[j] = max(zinit, c over [jl, j)), Y[j]
// = Z[j] - (j - B SW) E; inside a lane over its K cells, across lanes with a
// DPP scan, across slots through a carry.  The Y from-bit of cell j is
// Z[j] != c(j - 1) (ties go to M), and the window's first cell takes the
// reference's from-state of a cell whose left neighbour lies outside.
//
// Traceback rows are stored lane-major: lane l's 96 bits at bit 96 l, cell
// (k, p) of the lane at bits [4 (k K + p), + 4).
template <int K>
struct Fwd1Shape {
    static constexpr int SW = 64 * K;                     // columns per slot
    static constexpr int NS = (kMaxW + SW - 1) / SW + 1;  // slots: any window of <= kMaxW columns
    static_assert(NS * K * 4 <= 96, "a lane's traceback cells fit 96 bits");
};

// Query bytes 8 * (code | 4 * invalid) (as in the LDS cache of k_dp_forward)
// of query indices xs .. xs + 31 into out[0..8) (byte t of the 32 at byte t),
// zero for indices outside [0, m).
__device__ __forceinline__ void fwd1_qchunk(const DpProb &pr, int64_t xs, uint32_t *out) {
    uint32_t v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (xs + 32 > 0 && xs < (int64_t)pr.m) {
        const int64_t lx = xs < 0 ? 0 : xs;  // first index read
        const uint32_t sh = (uint32_t)(lx - xs);
        const Win w = dir_win(pr.Q, pr.qdir > 0 ? pr.q0 + (uint64_t)lx : pr.q0 - (uint64_t)lx, pr.qdir);
#pragma unroll
        for (int t = 0; t < 32; t++) {
            const int64_t x = xs + t;
            if ((uint32_t)t >= sh && x < (int64_t)pr.m) {
                const uint32_t u = (uint32_t)t - sh;
                const uint32_t code = (uint32_t)(w.codes >> (2 * u)) & 3u;
                const uint32_t inv = ((w.valid >> u) & 1u) ^ 1u;
                v[t >> 2] |= (8u * (code | (inv << 2))) << (8 * (t & 3));
            }
        }
    }
#pragma unroll
    for (int t = 0; t < 8; t++) out[t] = v[t];
}

// Where a one-wavefront sweep takes the traceback chunk of its rows
// 64 x .. 64 x + 63 (take(x, prev): the chunk, linked to the side's previous
// chunk `prev`, or kNoChunk when none is left).
// The pass kernels share the pool: chunks in arrival order from one counter.
struct PoolChunks {
    DpPool pool;
    __device__ __forceinline__ uint32_t take(uint32_t, uint32_t prev) const {
        uint32_t c = 0;
        if (threadIdx.x == 0) {
            c = atomicAdd(pool.next, 1u);
            if (c < pool.nchunks) pool.prev[c] = prev;
        }
        c = __builtin_amdgcn_readfirstlane(c);
        return c < pool.nchunks ? c : kNoChunk;
    }
};
// The persistent kernel: each block owns `count` consecutive chunks from
// `base` and reuses them for every side it sweeps.
struct RegionChunks {
    DpPool pool;
    uint32_t base, count;
    __device__ __forceinline__ uint32_t take(uint32_t x, uint32_t prev) const {
        if (x >= count) return kNoChunk;
        const uint32_t c = base + x;
        if (threadIdx.x == 0) pool.prev[c] = prev;
        return c;
    }
};

// query bytes a one-wavefront sweep caches in LDS
constexpr int kQC1 = 4096;

// One side swept by one wavefront (the body of k_dp_fwd1 and k_dp_persist);
// every lane returns the side's result.  qcache: kQC1 bytes of LDS.
template <class S, int K, class Chunks>
__device__ __forceinline__ DpRes fwd1_sweep(const DpProb &pr, S sc, int32_t oe, int32_t e, int32_t ydrop,
                                            DpPool pool, const Chunks &chunks, uint32_t *__restrict__ qcache) {
    constexpr int SW = Fwd1Shape<K>::SW, NS = Fwd1Shape<K>::NS;
    constexpr uint32_t kFmt = K == 1 ? 1u : (K == 2 ? 2u : 3u);
    const int lane = threadIdx.x;
    const uint32_t n = pr.n, m = pr.m;
    DpRes out = {0, 0, 0, kStM, 0, 0, kNoChunk, kFmt, 0};
    if (n == 0 || m == 0) return out;  // src/gapped_extend.rs:188-195
    // band radius / width (src/gapped_extend.rs:214-233)
    uint32_t br;
    if (e > 0) {
        const int32_t r = sat_sub(ydrop, oe) / e;
        br = (uint32_t)(max(r, 0) + 2);
    } else {
        br = m + 1;
    }
    const uint32_t bw = (uint32_t)min((uint64_t)4 * br + 4, (uint64_t)m + 1);  // host guarantees bw <= kMaxW
    // what a cell outside the previous row's window (or dead) yields
    const int32_t negx = max(kNegInf - oe, kNegInf - e);
    const uint32_t xb_neg = (kNegInf - oe >= kNegInf - e) ? 0u : 1u;
    uint32_t sb_negw = 0;  // arg M, X from-bit of NEG, in every cell field
#pragma unroll
    for (int p = 0; p < K; p++) sb_negw |= (xb_neg << 2) << (4 * p);
    const int32_t lead_y = negx;  // Y of a cell whose left neighbour is outside
    const uint32_t ybit_lead = xb_neg;

    // query bytes of query indices [qc0, qc0 + kQC1): column j reads index j - 1
    int64_t qc0 = INT64_MIN / 2;  // nothing cached yet

    // Row 0 from the same code: every column scores `ambig` in row 0 (no
    // target base), so D = NEG - ambig gives M = NEG exactly, X starts at NEG
    // and row 0 has no M or X from-states (src/gapped_extend.rs:245).
    int32_t dd[NS][K], xn[NS][K];
    uint32_t sb[NS];
#pragma unroll
    for (int k = 0; k < NS; k++) {
#pragma unroll
        for (int p = 0; p < K; p++) {
            dd[k][p] = kNegInf - sc.ambig;
            xn[k][p] = kNegInf;
        }
        sb[k] = 0;
    }
    auto reset_slot = [&](int k) {
#pragma unroll
        for (int p = 0; p < K; p++) {
            dd[k][p] = kNegInf;
            xn[k][p] = negx;
        }
        sb[k] = sb_negw;
    };
    uint32_t B = 0;  // base block: slot k holds columns [(B + k) SW, (B + k + 1) SW)

    int32_t best_score = 0;
    uint32_t bi = 0, bj = 0, bst = kStM;
    uint32_t lo = 0, hi = 0;
    Win tw = {0, 0, 0};
    uint64_t cells = 0;
    uint32_t rows_done = 0, chunk = kNoChunk, truncated = 0;
    for (uint32_t i = 0; i <= n; i++) {
        // a fresh traceback chunk every kChunkRows rows
        if (i % kChunkRows == 0) {
            const uint32_t c = chunks.take(i / kChunkRows, chunk);
            if (c == kNoChunk) {  // no chunk left: the side is swept again later
                truncated = 1;
                break;
            }
            chunk = c;
        }
        const uint32_t jl = lo >= br ? lo - br : 0;
        const uint32_t jh = (uint32_t)min(min((uint64_t)hi + br, (uint64_t)m), (uint64_t)jl + bw - 1);
        cells += (jh - jl + 1) - ((i == 0 && jl == 0) ? 1 : 0);

        // move the slots with the window; the column left of the new base
        // keeps its previous-row value (it may lie inside the previous window)
        const uint32_t Bn = jl / SW;
        int32_t carry_dd = kNegInf;  // D and sb word of the column left of slot 0 (lane 63 of a slot)
        uint32_t carry_sb = sb_negw;
        if (Bn > B) {
            const uint32_t d = Bn - B;
#pragma unroll
            for (int k = 0; k < NS; k++) {
                if ((uint32_t)k + 1 == d) {
                    carry_dd = __builtin_amdgcn_readlane(dd[k][K - 1], 63);
                    carry_sb = __builtin_amdgcn_readlane(sb[k], 63);
                }
            }
            const uint32_t steps = min(d, (uint32_t)NS);
            for (uint32_t x = 0; x < steps; x++) {
#pragma unroll
                for (int k = 0; k + 1 < NS; k++) {
#pragma unroll
                    for (int p = 0; p < K; p++) {
                        dd[k][p] = dd[k + 1][p];
                        xn[k][p] = xn[k + 1][p];
                    }
                    sb[k] = sb[k + 1];
                }
                reset_slot(NS - 1);
            }
            B = Bn;
        } else if (Bn < B) {
            const uint32_t steps = min(B - Bn, (uint32_t)NS);
            for (uint32_t x = 0; x < steps; x++) {
#pragma unroll
                for (int k = NS - 1; k > 0; k--) {
#pragma unroll
                    for (int p = 0; p < K; p++) {
                        dd[k][p] = dd[k - 1][p];
                        xn[k][p] = xn[k - 1][p];
                    }
                    sb[k] = sb[k - 1];
                }
                reset_slot(0);
            }
            B = Bn;
        }
        const uint32_t kmax = jh / SW - B;  // last slot the window reaches (< NS)
        {
            // the active slots read query indices [B SW - 1, (B + kmax + 1) SW - 1);
            // a refill starts the cache at B SW - 1, so a slot's K bytes stay aligned
            const int64_t xlo = (int64_t)B * SW - 1, xhi = (int64_t)(B + kmax + 1) * SW - 1;
            if (xlo < qc0 || xhi > qc0 + kQC1) {
                qc0 = xlo;
#pragma unroll 1
                for (int h = 0; h < kQC1 / 64 / 32; h++) {
                    const int c = lane * (kQC1 / 64 / 32) + h;  // 32-index chunk of this lane
                    uint32_t w[8];
                    fwd1_qchunk(pr, qc0 + 32 * c, w);
                    *(uint4 *)&qcache[8 * c] = make_uint4(w[0], w[1], w[2], w[3]);
                    *(uint4 *)&qcache[8 * c + 4] = make_uint4(w[4], w[5], w[6], w[7]);
                }
                __syncthreads();
            }
        }

        // target base of row i (T[i-1]); uniform
        uint32_t tc = 0, tv = 0;
        if (i >= 1) {
            const uint32_t r = (i - 1) & 31u;
            if (r == 0) tw = dir_win(pr.T, pr.tdir > 0 ? pr.t0 + (i - 1) : pr.t0 - (i - 1), pr.tdir);
            tc = (uint32_t)(tw.codes >> (2 * r)) & 3u;
            tv = (tw.valid >> r) & 1u;
        }
        const auto rs = row_scores(sc, tc, tv);
        // Per-row copies of the gap costs and the lane id: the column offsets
        // below are computed per row instead of being hoisted out of the row
        // loop into dozens of live registers.
        int32_t er = e, oer = oe;
        int32_t lr = (int32_t)__builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));  // lane id
        asm volatile("" : "+s"(er), "+s"(oer), "+v"(lr));
        const int32_t lc = K * lr;  // the lane's first column within a slot
        const int32_t lane_e = lc * er;

        const int32_t base_col = (int32_t)(B * SW);
        // Z at the window's first cell: for jl >= 1 both of Y[jl]'s candidates
        // lie outside the window; Y[0] is not computed
        int32_t zc = (jl >= 1 ? lead_y : kNegInf) + ((int32_t)jl - base_col) * er;
        // lane 0 of these holds the previous-row D / sb word, and the Y-scan
        // term, of the column left of the current slot (lane 63 of the slot
        // before, moved round with a DPP rotate)
        int32_t ldd = carry_dd, lsb = (int32_t)carry_sb, lcv = INT32_MIN;
        int32_t rmax = INT32_MIN;
        uint32_t nw0 = 0, nw1 = 0, nw2 = 0;  // the lane's traceback cells of this row
#pragma unroll
        for (int k = 0; k < NS; k++) {
            if ((uint32_t)k <= kmax) {
                const int32_t scol = base_col + k * SW;  // first column of the slot
                const int32_t jl_rel = (int32_t)jl - scol, jh_rel = (int32_t)jh - scol;
                // previous-row D and sb of column j - 1 for cell 0: lane - 1, or the slot before
                const int32_t dl = __builtin_amdgcn_update_dpp(ldd, dd[k][K - 1], 0x138, 0xF, 0xF, false);
                const uint32_t al = (uint32_t)__builtin_amdgcn_update_dpp(lsb, (int)sb[k], 0x138, 0xF, 0xF, false);
                ldd = __builtin_amdgcn_update_dpp(0, dd[k][K - 1], 0x13C, 0xF, 0xF, false);  // wave_ror:1
                lsb = __builtin_amdgcn_update_dpp(0, (int)sb[k], 0x13C, 0xF, 0xF, false);
                // this lane's K query bytes: index j - 1 - qc0 = (B + k) SW + lc - 1 - qc0, a multiple of K
                const uint32_t qoff = (uint32_t)((int64_t)scol - 1 - qc0) + (uint32_t)lc;
                uint32_t qk;
                if (K == 1)
                    qk = ((const uint8_t *)qcache)[qoff];
                else if (K == 2)
                    qk = ((const uint16_t *)qcache)[qoff / 2];
                else
                    qk = qcache[qoff / 4];
                int32_t Mv[K], cv[K];
#pragma unroll
                for (int p = 0; p < K; p++) {
                    const int32_t dprev = p == 0 ? dl : dd[k][p - 1];
                    Mv[p] = dprev + rs((qk >> (8 * p)) & 0xFFu);  // src/gapped_extend.rs:339-363
                    cv[p] = Mv[p] + lane_e + ((k * SW + p + 1) * er - oer);
                }
                uint32_t mfix = 0;  // column 0 (no M, no Y): lane 0 cell 0 of the slot at column 0
                if (k == 0 && scol == 0) {
                    if (lr == 0) {
                        Mv[0] = i == 0 ? 0 : kNegInf;
                        cv[0] = Mv[0] + ((1 * er) - oer);
                    }
                    mfix = lr == 0 ? 1u : 0u;
                }
                if (k == 0) {  // columns left of the window add nothing to the Y scan
#pragma unroll
                    for (int p = 0; p < K; p++)
                        if (lc + p < jl_rel) cv[p] = INT32_MIN;
                }
                // Y: lane-local exclusive prefix, DPP scan across lanes, carry across slots
                int32_t E[K];
                E[0] = INT32_MIN;
#pragma unroll
                for (int p = 1; p < K; p++) E[p] = max(E[p - 1], cv[p - 1]);
                const int32_t incl = wave_scan_max_self(max(E[K - 1], cv[K - 1]));
                const int32_t excl = max(__builtin_amdgcn_update_dpp(zc, incl, 0x138, 0xF, 0xF, false), zc);
                const int32_t cl0 = __builtin_amdgcn_update_dpp(lcv, cv[K - 1], 0x138, 0xF, 0xF, false);
                lcv = __builtin_amdgcn_update_dpp(0, cv[K - 1], 0x13C, 0xF, 0xF, false);
                zc = max(zc, __builtin_amdgcn_readlane(incl, 63));
                const bool full = jl_rel <= 0 && jh_rel >= SW - 1;
                uint32_t sbn = 0;
#pragma unroll
                for (int p = 0; p < K; p++) {
                    const int32_t z = max(excl, E[p]);
                    int32_t yv = z - lane_e - (k * SW + p) * er;
                    uint32_t yb = z != (p == 0 ? cl0 : cv[p - 1]) ? 8u : 0u;
                    // from-states of the previous row: M from column j - 1, X from column j
                    const uint32_t sprev = p == 0 ? (al >> (4 * (K - 1))) : (sb[k] >> (4 * (p - 1)));
                    uint32_t nib = __builtin_amdgcn_ubfe(sprev, 0, 2) | (sb[k] >> (4 * p) & 4u);
                    if (k == 0) {
                        if (lc + p == jl_rel) yb = ybit_lead << 3;
                        if (p == 0 && mfix) {
                            yv = kNegInf;
                            yb = 0;
                            nib &= 4u;
                        }
                    }
                    nib |= yb;
                    const int cidx = k * K + p;
                    if (cidx < 8)
                        nw0 |= nib << (4 * cidx);
                    else if (cidx < 16)
                        nw1 |= nib << (4 * (cidx - 8));
                    else
                        nw2 |= nib << (4 * (cidx - 16));
                    const int32_t xv = xn[k][p];
                    const int32_t cb = max(max(Mv[p], xv), yv);
                    const uint32_t st = Mv[p] != cb ? (xv != cb ? kStY : kStX) : kStM;
                    const int32_t tm = Mv[p] - oer, tx = xv - er;
                    xn[k][p] = max(tm, tx);
                    dd[k][p] = cb;
                    sbn |= (st | (tm >= tx ? 0u : 4u)) << (4 * p);
                    if (full)
                        rmax = max(rmax, cb);
                    else
                        rmax = max(rmax, (uint32_t)(lc + p - jl_rel) <= (uint32_t)(jh_rel - jl_rel) ? cb : INT32_MIN);
                }
                sb[k] = sbn;
            } else {
                reset_slot(k);  // outside the window: NEG for the next row
            }
            // keep the scheduler from interleaving slots: it would hold every
            // slot's temporaries at once
            __builtin_amdgcn_sched_barrier(0);
        }

        // ---- row best (first strict maximum in j order, src/gapped_extend.rs:365-380)
        const int32_t wmax = __builtin_amdgcn_readlane(wave_scan_max_self(rmax), 63);
        if (wmax > best_score) {
            bool found = false;
#pragma unroll
            for (int k = 0; k < NS; k++) {
                if ((uint32_t)k <= kmax && !found) {
                    const int32_t scol = base_col + k * SW;
                    const int32_t jl_rel = (int32_t)jl - scol, jh_rel = (int32_t)jh - scol;
                    uint64_t hit[K], any = 0;
#pragma unroll
                    for (int p = 0; p < K; p++) {
                        const bool in = (uint32_t)(lc + p - jl_rel) <= (uint32_t)(jh_rel - jl_rel);
                        hit[p] = __ballot(in && dd[k][p] == wmax);
                        any |= hit[p];
                    }
                    if (any) {  // the first lane holding the maximum, then its first cell
                        const int l = __ffsll((long long)any) - 1;
                        int pf = K - 1;
#pragma unroll
                        for (int p = K - 1; p >= 0; p--)
                            if ((hit[p] >> l) & 1ull) pf = p;
                        best_score = wmax;
                        bi = i;
                        bj = (uint32_t)(scol + K * l + pf);
                        bst = ((uint32_t)__builtin_amdgcn_readlane((int)sb[k], l) >> (4 * pf)) & 3u;
                        found = true;
                    }
                }
            }
        }
        const int32_t thr = sat_sub(best_score, ydrop);

        // ---- alive band; dead cells become NEG for the next row (src/gapped_extend.rs:388-412)
        uint32_t alo = 0xFFFFFFFFu, ahi = 0;
#pragma unroll
        for (int k = 0; k < NS; k++) {
            if ((uint32_t)k <= kmax) {
                const int32_t scol = base_col + k * SW;
                const int32_t jl_rel = (int32_t)jl - scol, jh_rel = (int32_t)jh - scol;
                uint64_t al[K], any = 0;
                uint32_t sbk = sb[k];
#pragma unroll
                for (int p = 0; p < K; p++) {
                    const bool in = (uint32_t)(lc + p - jl_rel) <= (uint32_t)(jh_rel - jl_rel);
                    const bool alive = in && dd[k][p] >= thr;
                    al[p] = __ballot(alive);
                    any |= al[p];
                    dd[k][p] = alive ? dd[k][p] : kNegInf;
                    xn[k][p] = alive ? xn[k][p] : negx;
                    const uint32_t f = 0xFu << (4 * p);
                    sbk = alive ? sbk : ((sbk & ~f) | (sb_negw & f));
                }
                sb[k] = sbk;
                if (any) {
                    const int lf = __ffsll((long long)any) - 1, ll = 63 - __clzll((long long)any);
                    int pf = K - 1, pl = 0;
#pragma unroll
                    for (int p = K - 1; p >= 0; p--)
                        if ((al[p] >> lf) & 1ull) pf = p;
#pragma unroll
                    for (int p = 0; p < K; p++)
                        if ((al[p] >> ll) & 1ull) pl = p;
                    alo = min(alo, (uint32_t)(scol + K * lf + pf));
                    ahi = (uint32_t)(scol + K * ll + pl);
                }
            }
        }
        const uint32_t ir = i % kChunkRows;
        {
            uint32_t *row = (uint32_t *)(pool.tb + (uint64_t)chunk * kChunkBytes + ir * kRowBytes) + 3 * lr;
            row[0] = nw0;
            if (NS * K > 8) row[1] = nw1;
            if (NS * K > 16) row[2] = nw2;
        }
        if (lane == 0) pool.win[(uint64_t)chunk * kChunkRows + ir] = make_uint2(jl, jh);
        rows_done = i + 1;
        if (alo == 0xFFFFFFFFu) break;  // whole row dead (src/gapped_extend.rs:407-409)
        lo = alo;
        hi = ahi;
    }
    out.score = best_score;
    out.bi = bi;
    out.bj = bj;
    out.bstate = bst;
    out.rows = rows_done;
    out.truncated = truncated;
    out.last_chunk = chunk;
    out.cells = cells;
    return out;
}

template <class S, int K, int WPE>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(WPE))) void k_dp_fwd1(
    const DpProb *__restrict__ probs, const uint32_t *__restrict__ list, uint32_t nprob, S sc, int32_t oe, int32_t e,
    int32_t ydrop, DpPool pool, DpRes *__restrict__ res) {
    __shared__ uint32_t qcache[kQC1 / 4];
    const uint32_t side = blockIdx.x;
    if (side >= nprob) return;
    const DpProb pr = probs[list[side]];
    const DpRes out = fwd1_sweep<S, K>(pr, sc, oe, e, ydrop, pool, PoolChunks{pool}, qcache);
    if (threadIdx.x == 0) res[side] = out;
}

// One lane's words of kTileRows consecutive traceback rows.  Named members
// (not an array) keep them in registers while they are carried round the
// walk loop.
static_assert(kTileRows == 8, "TbRows holds eight rows");
struct TbRows {
    uint4 r0, r1, r2, r3, r4, r5, r6, r7;
};

__device__ __forceinline__ TbRows tb_load(const uint8_t *p) {
    TbRows t;
    t.r0 = *(const uint4 *)(p + 0 * kRowBytes);
    t.r1 = *(const uint4 *)(p + 1 * kRowBytes);
    t.r2 = *(const uint4 *)(p + 2 * kRowBytes);
    t.r3 = *(const uint4 *)(p + 3 * kRowBytes);
    t.r4 = *(const uint4 *)(p + 4 * kRowBytes);
    t.r5 = *(const uint4 *)(p + 5 * kRowBytes);
    t.r6 = *(const uint4 *)(p + 6 * kRowBytes);
    t.r7 = *(const uint4 *)(p + 7 * kRowBytes);
    return t;
}

__device__ __forceinline__ void tb_store(uint32_t (*tile)[kRowWords], int lane, const TbRows &t) {
    *(uint4 *)&tile[0][lane * 4] = t.r0;
    *(uint4 *)&tile[1][lane * 4] = t.r1;
    *(uint4 *)&tile[2][lane * 4] = t.r2;
    *(uint4 *)&tile[3][lane * 4] = t.r3;
    *(uint4 *)&tile[4][lane * 4] = t.r4;
    *(uint4 *)&tile[5][lane * 4] = t.r5;
    *(uint4 *)&tile[6][lane * 4] = t.r6;
    *(uint4 *)&tile[7][lane * 4] = t.r7;
}

// Walk back from the best cell (src/gapped_extend.rs:415-463), one wavefront
// per side: the wavefront stages kTileRows packed traceback rows (plus their
// windows) in LDS, and each iteration takes a whole
// run of one op at once (its lanes read the run's next 64 steps in parallel
// and a ballot finds where it ends), so a run costs one pair of LDS reads
// instead of one dependent read per step.  Runs are emitted tip-to-root,
// merged while the op repeats, through out.put(k, run) for k = 0, 1, ... (at
// most bi + bj of them); the walk returns their count.
struct TbLds {
    uint32_t tile[kTileRows][kRowWords];
    uint2 twin[kTileRows];
    uint32_t tlink;  // chunk below the staged tile's chunk
};
// Run sinks of the walk: lane 0 writes run k.
struct RunsOut {  // a side's reserved range in device memory
    gxy_edit_run *p;
    __device__ __forceinline__ void put(uint32_t k, gxy_edit_run r) const {
        if (threadIdx.x == 0) p[k] = r;
    }
};
template <class Out>
__device__ __forceinline__ uint32_t tb_walk(const DpRes &r, DpPool pool, TbLds &L, const Out &out) {
    uint32_t(*tile)[kRowWords] = L.tile;
    uint2 *twin = L.twin;
    uint32_t &tlink = L.tlink;
    const int lane = threadIdx.x;
    if (r.truncated || (r.bi == 0 && r.bj == 0) || r.rows == 0) return 0;
    uint32_t i = r.bi, j = r.bj, st = r.bstate;
    // chunk holding row bi
    uint32_t c = r.last_chunk;
    for (uint32_t d = (r.rows - 1) / kChunkRows; d > r.bi / kChunkRows; d--) c = pool.prev[c];
    uint32_t nr = 0;
    uint32_t cur_op = 0xFFFFFFFFu, cur_n = 0;
    const uint32_t cap = r.bi + r.bj;  // runs reserved for this side
    // The walk moves down at most one row a step, so tiles are visited in
    // order: the tile below the staged one is loaded into registers (pt, pw)
    // while the walk crosses the staged one, hiding the global round trip.
    // Lane kTileRows carries the link to the chunk below the loaded tile's.
    int64_t tl = -1;                   // first row of the staged tile
    int64_t ptl = -1;                  // first row of the loaded tile
    uint32_t pc = c;                   // chunk of the loaded tile
    uint32_t cbelow = 0;               // chunk below c
    TbRows pt;
    uint2 pw = {0u, 0u};
    uint32_t plink = 0;
    while (i > 0 || j > 0) {
        // A consistent traceback never leaves the swept cells; these guards
        // stop a walk rather than read outside the traceback store.
        if (i >= r.rows || (st == kStM && (i == 0 || j == 0)) || (st == kStX && i == 0) ||
            (st == kStY && j == 0) || st > kStY || nr >= cap)
            break;
        if ((int64_t)i < tl || tl < 0) {
            const uint32_t ntl = i / kTileRows * kTileRows;
            if ((int64_t)ntl != ptl) {
                // first tile of the walk: load it now
                const uint32_t r0 = ntl % kChunkRows;
                pc = c;
                if (lane < kRowLoads) pt = tb_load(pool.tb + (uint64_t)pc * kChunkBytes + r0 * kRowBytes + lane * 16);
                if (lane < kTileRows) pw = pool.win[(uint64_t)pc * kChunkRows + r0 + lane];
                if (lane == kTileRows) plink = pool.prev[pc];
            }
            __syncthreads();
            if (lane < kRowLoads) tb_store(tile, lane, pt);
            if (lane < kTileRows) twin[lane] = pw;
            if (lane == kTileRows) tlink = plink;
            __syncthreads();
            c = pc;
            cbelow = tlink;
            tl = ntl;
            if (ntl >= kTileRows) {
                ptl = ntl - kTileRows;
                pc = (ntl % kChunkRows == 0) ? cbelow : c;
                const uint32_t r0 = (uint32_t)ptl % kChunkRows;
                if (lane < kRowLoads) pt = tb_load(pool.tb + (uint64_t)pc * kChunkBytes + r0 * kRowBytes + lane * 16);
                if (lane < kTileRows) pw = pool.win[(uint64_t)pc * kChunkRows + r0 + lane];
                if (lane == kTileRows) plink = pool.prev[pc];
            } else {
                ptl = -1;
            }
        }
        // Lane k takes step k of a run of the current op from (i, j): the
        // cell (i - k, j - k) for M, (i - k, j) for X, (i, j - k) for Y.  A
        // step is valid where the one-step walk's guards pass (cell off row
        // or column 0 as the op needs, j inside the row's window) and
        // continues the run where its from-state is the run's own state.
        // The first step that does not continue ends the run: it is taken
        // (one more op, then its from-state) if valid, else the walk stops
        // before it.  Rows below the staged tile are left to the next tile.
        const uint32_t dk = (uint32_t)lane;
        const uint32_t di = st == kStY ? 0u : dk, dj = st == kStX ? 0u : dk;
        const uint32_t lim = st == kStY ? 64u : min(64u, i - (uint32_t)tl + 1u);  // steps inside the tile
        bool valid = false;
        uint32_t next = kStY + 1u;
        if (dk < lim && dj <= j) {
            const uint32_t ci = i - di, cj = j - dj;
            const bool edge_ok = st == kStM ? (ci > 0 && cj > 0) : st == kStX ? ci > 0 : cj > 0;
            const uint2 win = twin[ci - (uint32_t)tl];
            if (edge_ok && cj >= win.x && cj <= win.y) {
                valid = true;
                uint32_t word, sh;
                if (r.fmt == 0) {  // window order: cell idx = j - jl at bits [4 idx, 4 idx + 4)
                    const uint32_t idx = cj - win.x;
                    word = idx >> 3;
                    sh = 4 * (idx & 7u);
                } else {  // lane-major (k_dp_fwd1<K>, K = 2^(fmt-1)): slot, lane, cell of column cj
                    const uint32_t ks = r.fmt - 1u, sws = 6u + ks;
                    const uint32_t c = (((cj >> sws) - (win.x >> sws)) << ks) + (cj & ((1u << ks) - 1u));
                    word = 3u * ((cj >> ks) & 63u) + (c >> 3);
                    sh = 4 * (c & 7u);
                }
                const uint32_t nb = (tile[ci - (uint32_t)tl][word] >> sh) & 0xFu;
                next = st == kStM ? (nb & 3u)
                                  : st == kStX ? (((nb >> 2) & 1u) ? kStX : kStM) : (((nb >> 3) & 1u) ? kStY : kStM);
            }
        }
        const uint64_t ends = __ballot(dk < lim && (!valid || next != st));
        uint32_t n;               // ops of this run taken now
        bool stop = false;        // the walk ends before step n
        uint32_t nst = st;
        if (ends == 0) {
            n = lim;
        } else {
            const uint32_t kf = (uint32_t)__ffsll((long long)ends) - 1u;
            const bool vf = (__ballot(valid) >> kf) & 1u;
            if (vf) {
                n = kf + 1u;
                nst = (uint32_t)__shfl((int)next, (int)kf);
            } else {
                n = kf;
                stop = true;
            }
        }
        const uint32_t op = st == kStM ? GXY_OP_MATCH : st == kStX ? GXY_OP_DELETE : GXY_OP_INSERT;
        if (n) {
            if (op == cur_op) {
                cur_n += n;
            } else {
                if (cur_n) {
                    if (nr >= cap) break;
                    out.put(nr, gxy_edit_run{cur_op, cur_n});
                    nr++;
                }
                cur_op = op;
                cur_n = n;
            }
            if (st != kStY) i -= n;
            if (st != kStX) j -= n;
        }
        if (stop) break;
        st = nst;
    }
    if (cur_n && nr < cap) {
        out.put(nr, gxy_edit_run{cur_op, cur_n});
        nr++;
    }
    return nr;
}

// The walk of side p into runs[run_off[p] ..].
__global__ __launch_bounds__(64) void k_dp_traceback(const DpRes *__restrict__ res, uint32_t nprob, DpPool pool,
                                                     const uint64_t *__restrict__ run_off,
                                                     gxy_edit_run *__restrict__ runs, uint32_t *__restrict__ nruns) {
    __shared__ __align__(16) TbLds L;
    const uint32_t p = blockIdx.x;
    if (p >= nprob) return;
    const DpRes r = res[p];
    const uint32_t nr = tb_walk(r, pool, L, RunsOut{runs + run_off[p]});
    if (threadIdx.x == 0) nruns[p] = nr;
}

// ---- persistent form of a batch -------------------------------------------
//
// One launch of as many one-wavefront blocks as the GPU holds at once.  Each
// block takes sides from a device queue in list order (the host sorts the
// list longest first, so the long sweeps start early and the short ones fill
// the tail), sweeps a side forward into its own run of pool chunks, walks it
// back at once and appends the side's runs to one array, then takes the next
// side: no pass boundaries and no host work between sides.  A side whose
// sweep outgrows the block's chunks is left truncated, like a pass kernel's
// side that finds the pool empty, and the host sweeps it again in a pass.
constexpr uint32_t kRunStage = 256;  // runs a block stages in LDS per side
struct PersistArgs {
    uint32_t *head;                   // next list index to take
    uint32_t region;                  // pool chunks per block
    unsigned long long *run_ctr;      // runs appended so far
    gxy_edit_run *runs;
    uint64_t run_cap;
    uint64_t *run_off;                // per list index: first run
    uint32_t *nruns;                  //                 run count
};
struct StageOut {  // the first kRunStage runs into LDS; the walk counts the rest
    gxy_edit_run *stage;
    __device__ __forceinline__ void put(uint32_t k, gxy_edit_run r) const {
        if (threadIdx.x == 0 && k < kRunStage) stage[k] = r;
    }
};

template <class S, int K, int WPE>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(WPE))) void k_dp_persist(
    const DpProb *__restrict__ probs, const uint32_t *__restrict__ list, uint32_t nlist, S sc, int32_t oe, int32_t e,
    int32_t ydrop, DpPool pool, DpRes *__restrict__ res, PersistArgs pa) {
    // the forward sweep's query cache and the walk's tiles share the LDS
    union Lds {
        uint32_t qcache[kQC1 / 4];
        struct {
            TbLds tb;
            gxy_edit_run stage[kRunStage];
        } w;
    };
    __shared__ __align__(16) Lds L;
    const RegionChunks chunks{pool, blockIdx.x * pa.region, pa.region};
    for (;;) {
        uint32_t x = 0;
        if (threadIdx.x == 0) x = atomicAdd(pa.head, 1u);
        x = __builtin_amdgcn_readfirstlane(x);
        if (x >= nlist) break;
        const DpProb pr = probs[list[x]];
        DpRes r = fwd1_sweep<S, K>(pr, sc, oe, e, ydrop, pool, chunks, L.qcache);
        // the rows, windows and links written by the sweep, for the walk's loads
        __threadfence();
        __syncthreads();  // the query cache's LDS becomes the walk's
        uint32_t nr = tb_walk(r, pool, L.w.tb, StageOut{L.w.stage});
        uint64_t off = 0;
        if (nr) {
            unsigned long long o = 0;
            if (threadIdx.x == 0) o = atomicAdd(pa.run_ctr, (unsigned long long)nr);
            off = ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(o >> 32)) << 32) |
                  (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)o);
            if (off + nr > pa.run_cap) {  // no room for the runs: swept again in a pass
                r.truncated = 1;
                nr = 0;
                off = 0;
            } else if (nr <= kRunStage) {
                __syncthreads();  // lane 0's staged runs
                for (uint32_t k = threadIdx.x; k < nr; k += 64) pa.runs[off + k] = L.w.stage[k];
            } else {  // more runs than the stage holds: walk again into the reserved range
                tb_walk(r, pool, L.w.tb, RunsOut{pa.runs + off});
            }
        }
        if (threadIdx.x == 0) {
            res[x] = r;
            pa.run_off[x] = off;
            pa.nruns[x] = nr;
        }
        __syncthreads();  // the walk's LDS becomes the next sweep's query cache
    }
}

// Copy each side's runs from its reserved range to a packed range.
__global__ __launch_bounds__(64) void k_pack_runs(const uint64_t *__restrict__ run_off,
                                                  const uint64_t *__restrict__ pack_off,
                                                  const uint32_t *__restrict__ nruns, uint32_t nprob,
                                                  const gxy_edit_run *__restrict__ src, gxy_edit_run *__restrict__ dst) {
    const uint32_t p = blockIdx.x;
    if (p >= nprob) return;
    const uint32_t nr = nruns[p];
    const gxy_edit_run *a = src + run_off[p];
    gxy_edit_run *b = dst + pack_off[p];
    for (uint32_t x = threadIdx.x; x < nr; x += 64) b[x] = a[x];
}

// choose_anchor (src/anchor.rs:25-69), one lane per HSP.
// ---- sides wider than kMaxW ---------------------------------------------------
//
// k_dp_band<S>: the same row recurrence for sides whose band is wider than the
// kernels above hold (y_drop / gap_extend large, or gap_extend = 0, where the
// band is the whole query; src/gapped_extend.rs:214-233 sets no limit).  One
// block of four wavefronts per side; the row lives in global memory (two row
// buffers of the side's band width, int2 {D = max(M, X, Y), X-term =
// max(M - (O+E), X - E)} plus a byte of from-states per cell, as the
// register kernels keep them), lane t owning C (even) consecutive cells of
// the window.  Dead cells are not rewritten: the next row reads a cell as NEG
// when it lies outside the previous window or below the previous row's
// threshold.  Per row three passes over a lane's cells: (1) M and X from the
// previous row, the Y-scan terms; block scan of the lane totals; (2) Y,
// from-states, the new row and the traceback nibbles; block reduction of the
// row best; (3) the alive band under the new threshold.
// Traceback rows go to the side's own range (row i at i * rb bytes, nibble
// of column j at index j - jl, low nibble first) with the row windows beside
// them; a side that outruns its reserved rows stops (truncated) and is swept
// again with room for every row.
struct BandSide {
    uint64_t tb;        // byte offset of the side's traceback rows
    uint64_t win;       // index of its first row window
    uint64_t rowbuf;    // index of its first row-buffer cell (two rows of bw cells)
    uint32_t cap_rows;  // traceback rows reserved
    uint32_t rb;        // traceback bytes per row
};
constexpr uint32_t kFmtBand = 16;  // DpRes::fmt of k_dp_band's sides

// query byte 8 * (code | 4 * invalid) of base u of a window
__device__ __forceinline__ uint32_t win_qbyte(const Win &w, uint32_t u) {
    const uint32_t code = (uint32_t)(w.codes >> (2 * u)) & 3u;
    const uint32_t inv = ((w.valid >> u) & 1u) ^ 1u;
    return 8u * (code | (inv << 2));
}

template <class S>
__global__ __launch_bounds__(kDpThreads) void k_dp_band(const DpProb *__restrict__ probs,
                                                        const uint32_t *__restrict__ list,
                                                        const BandSide *__restrict__ bsides, uint32_t nside, S sc,
                                                        int32_t oe, int32_t e, int32_t ydrop, uint8_t *__restrict__ tb,
                                                        uint2 *__restrict__ wins, int2 *__restrict__ rows,
                                                        uint8_t *__restrict__ rows_sb, DpRes *__restrict__ res) {
    __shared__ int32_t w_tot[kDpWaves], w_alast[kDpWaves], w_max[kDpWaves];
    __shared__ uint32_t w_j[kDpWaves], w_st[kDpWaves], w_lo[kDpWaves], w_hi[kDpWaves];
    const uint32_t sidx = blockIdx.x;
    if (sidx >= nside) return;  // uniform per block
    const int t = threadIdx.x, lane = t & 63;
    const int wv = __builtin_amdgcn_readfirstlane(t >> 6);
    const DpProb pr = probs[list[sidx]];
    const BandSide bs = bsides[sidx];
    const uint32_t n = pr.n, m = pr.m;
    DpRes out = {0, 0, 0, kStM, 0, 0, kNoChunk, kFmtBand, 0};
    if (n == 0 || m == 0) {  // src/gapped_extend.rs:188-195
        if (t == 0) res[sidx] = out;
        return;
    }
    // band radius / width (src/gapped_extend.rs:214-233)
    uint32_t br;
    if (e > 0) {
        const int32_t r = sat_sub(ydrop, oe) / e;
        br = (uint32_t)(max(r, 0) + 2);
    } else {
        br = m + 1;
    }
    const uint32_t bw = (uint32_t)min((uint64_t)4 * br + 4, (uint64_t)m + 1);
    const int32_t negx = max(kNegInf - oe, kNegInf - e);
    const uint32_t xb_neg = (kNegInf - oe >= kNegInf - e) ? 0u : 1u;
    const int32_t lead_y = negx;  // Y of the window's first cell (src/gapped_extend.rs:318-334)
    const uint32_t ybit_lead = xb_neg;
    int2 *buf[2] = {rows + bs.rowbuf, rows + bs.rowbuf + bw};
    uint8_t *sbuf[2] = {rows_sb + bs.rowbuf, rows_sb + bs.rowbuf + bw};
    uint8_t *const tb0 = tb + bs.tb;
    uint2 *const win0 = wins + bs.win;

    int32_t best_score = 0;
    uint32_t bi = 0, bj = 0, bst = kStM;
    uint32_t lo = 0, hi = 0, pjl = 0, pjh = 0;
    int32_t thr_prev = 0;
    Win tw = {0, 0, 0};
    uint64_t cells = 0;
    uint32_t rows_done = 0, truncated = 0, cur = 0;
    for (uint32_t i = 0; i <= n; i++) {
        if (i >= bs.cap_rows) {  // out of reserved rows: swept again with room for all
            truncated = 1;
            break;
        }
        const uint32_t jl = lo >= br ? lo - br : 0;
        const uint32_t jh = (uint32_t)min(min((uint64_t)hi + br, (uint64_t)m), (uint64_t)jl + bw - 1);
        const uint32_t W = jh - jl + 1;
        cells += W - ((i == 0 && jl == 0) ? 1 : 0);
        // lane t owns cells [c0, c0 + cnt) of the window, C even
        const uint32_t C = 2u * ((W + 2u * kDpThreads - 1) / (2u * kDpThreads));
        const uint32_t c0 = (uint32_t)t * C;
        const uint32_t cnt = c0 < W ? min(C, W - c0) : 0;
        // target base of row i (T[i-1]); uniform
        uint32_t tc = 0, tv = 0;
        if (i >= 1) {
            const uint32_t r = (i - 1) & 31u;
            if (r == 0) tw = dir_win(pr.T, pr.tdir > 0 ? pr.t0 + (i - 1) : pr.t0 - (i - 1), pr.tdir);
            tc = (uint32_t)(tw.codes >> (2 * r)) & 3u;
            tv = (tw.valid >> r) & 1u;
        }
        const auto rs = row_scores(sc, tc, tv);
        int2 *const pb = buf[cur ^ 1u];
        const uint8_t *const psb = sbuf[cur ^ 1u];
        int2 *const cb_ = buf[cur];
        uint8_t *const csb = sbuf[cur];
        // the previous row's cell at column c as the reference reads it:
        // NEG outside its window or below its threshold
        auto prev = [&](uint32_t c, int32_t &d, int32_t &xv, uint32_t &s) {
            d = kNegInf;
            xv = negx;
            s = xb_neg << 2;
            if (i >= 1 && c >= pjl && c <= pjh) {
                const int2 v = pb[c - pjl];
                if (v.x >= thr_prev) {
                    d = v.x;
                    xv = v.y;
                    s = psb[c - pjl];
                }
            }
        };

        // ---- pass 1: M, X, their from-states; the Y-scan terms ----------------
        // Z[j] = Y[j] + (j - jl) E = max(Z[j-1], M[j-1] - (O+E) + (j - jl) E)
        int32_t tot = INT32_MIN, alast = INT32_MIN;
        {
            Win qw = {0, 0, 0};
            int64_t qbase = -64;  // query index of qw's base 0
            int32_t dl = 0, xl = 0;
            uint32_t sl = 0;
            if (cnt) {
                const uint32_t j0 = jl + c0;
                if (j0 >= 1) prev(j0 - 1, dl, xl, sl);
            }
            for (uint32_t k = 0; k < cnt; k++) {
                const uint32_t j = jl + c0 + k;
                int32_t dj, xj;
                uint32_t sj;
                prev(j, dj, xj, sj);
                int32_t mv;
                uint32_t nib;
                if (i == 0) {  // row 0: M only at the anchor cell, no X (src/gapped_extend.rs:245)
                    mv = j == 0 ? 0 : kNegInf;
                    xj = kNegInf;
                    nib = 0;
                } else {
                    nib = sj & 4u;  // X from M or X (src/gapped_extend.rs:293-312)
                    if (j == 0) {   // column 0 has no M
                        mv = kNegInf;
                    } else {        // src/gapped_extend.rs:339-363
                        const int64_t qi = (int64_t)j - 1;
                        if (qi - qbase >= 32 || qi < qbase) {
                            qbase = qi;
                            qw = dir_win(pr.Q, pr.qdir > 0 ? pr.q0 + (uint64_t)qi : pr.q0 - (uint64_t)qi, pr.qdir);
                        }
                        mv = dl + rs(win_qbyte(qw, (uint32_t)(qi - qbase)));
                        nib |= sl & 3u;
                    }
                }
                cb_[c0 + k] = make_int2(mv, xj);
                csb[c0 + k] = (uint8_t)nib;
                const int32_t a = mv - oe + (int32_t)(j - jl + 1) * e;
                tot = max(tot, a);
                alast = a;
                dl = dj;
                sl = sj;
            }
        }
        const int32_t incl = wave_scan_max(tot);
        int32_t excl = wave_shr1(incl, INT32_MIN);
        int32_t aprev = wave_shr1(alast, INT32_MIN);
        if (lane == 63) {
            w_tot[wv] = incl;
            w_alast[wv] = alast;
        }
        __syncthreads();  // (1) wavefront totals of the Y scan
        for (int v = 0; v < wv; v++) excl = max(excl, w_tot[v]);
        if (lane == 0 && wv > 0) aprev = w_alast[wv - 1];

        // ---- pass 2: Y, the new row, traceback nibbles, the lane's row best ----
        int32_t rmax = INT32_MIN;
        uint32_t rj = 0, rst = kStY;
        {
            int32_t z = max(excl, jl >= 1 ? lead_y : kNegInf);
            int32_t an = aprev;  // Y-scan term of the cell left of the current one
            uint32_t byte = 0;
            uint8_t *const trow = tb0 + (uint64_t)i * bs.rb;
            for (uint32_t k = 0; k < cnt; k++) {
                const uint32_t j = jl + c0 + k;
                const int2 mx = cb_[c0 + k];
                uint32_t nib = csb[c0 + k];
                uint32_t yb;
                if (k == 0) {
                    yb = t == 0 ? ybit_lead : (z != an ? 1u : 0u);
                } else {
                    yb = an < z ? 1u : 0u;
                    z = max(z, an);
                }
                int32_t y = z - (int32_t)(j - jl) * e;
                if (j == 0) {  // column 0 has no Y
                    y = kNegInf;
                    yb = 0;
                    nib &= 4u;
                }
                nib |= yb << 3;
                const int32_t mv = mx.x, xv = mx.y;
                const int32_t cb = max(max(mv, xv), y);
                const uint32_t st = mv == cb ? kStM : (xv == cb ? kStX : kStY);
                const int32_t tm = mv - oe, tx = xv - e;
                cb_[c0 + k] = make_int2(cb, max(tm, tx));
                csb[c0 + k] = (uint8_t)(st | (tm >= tx ? 0u : 4u));
                if (cb > rmax) {  // first strict maximum
                    rmax = cb;
                    rj = j;
                    rst = st;
                }
                an = mv - oe + (int32_t)(j - jl + 1) * e;
                // nibbles in window order, low nibble first; C is even, so a
                // lane's bytes are its own
                byte |= nib << (4 * (k & 1u));
                if ((k & 1u) || k + 1 == cnt) {
                    trow[(c0 + k) >> 1] = (uint8_t)byte;
                    byte = 0;
                }
            }
        }
        {
            const int32_t wmax = __builtin_amdgcn_readlane(wave_scan_max(rmax), 63);
            const unsigned long long hasmax = __ballot(cnt > 0 && rmax == wmax);
            if (hasmax) {
                const int src = __ffsll((long long)hasmax) - 1;
                if (lane == src) {
                    w_max[wv] = wmax;
                    w_j[wv] = rj;
                    w_st[wv] = rst;
                }
            } else if (lane == 0) {
                w_max[wv] = INT32_MIN;
            }
        }
        __syncthreads();  // (2) row best per wavefront
        {
            int32_t gmax = INT32_MIN;
            uint32_t gw = 0;
#pragma unroll
            for (int v = 0; v < kDpWaves; v++) {
                const int32_t x = w_max[v];
                if (x > gmax) {  // strict: the earliest wavefront holding the maximum
                    gmax = x;
                    gw = v;
                }
            }
            if (gmax > best_score) {  // src/gapped_extend.rs:365-380
                best_score = gmax;
                bi = i;
                bj = w_j[gw];
                bst = w_st[gw];
            }
        }
        const int32_t thr = sat_sub(best_score, ydrop);

        // ---- pass 3: the alive band (src/gapped_extend.rs:388-412) ------------
        uint32_t alo = 0xFFFFFFFFu, ahi = 0;
        for (uint32_t k = 0; k < cnt; k++) {
            if (cb_[c0 + k].x >= thr) {
                const uint32_t j = jl + c0 + k;
                alo = min(alo, j);
                ahi = j;
            }
        }
        {
            const unsigned long long anyalive = __ballot(alo != 0xFFFFFFFFu);
            const int fl = anyalive ? __ffsll((long long)anyalive) - 1 : 0;
            const int ll = anyalive ? 63 - __clzll((long long)anyalive) : 0;
            const uint32_t wl = __builtin_amdgcn_readlane(alo, fl);
            const uint32_t wh = __builtin_amdgcn_readlane(ahi, ll);
            if (lane == 0) {
                w_lo[wv] = anyalive ? wl : 0xFFFFFFFFu;
                w_hi[wv] = wh;
            }
        }
        if (t == 0) win0[i] = make_uint2(jl, jh);
        rows_done = i + 1;
        __syncthreads();  // (3) alive band; the new row is written
        uint32_t nlo = 0xFFFFFFFFu, nhi = 0;
        bool alive = false;
#pragma unroll
        for (int v = 0; v < kDpWaves; v++) {
            const uint32_t l = w_lo[v];
            if (l != 0xFFFFFFFFu) {
                if (!alive) nlo = l;
                nhi = w_hi[v];
                alive = true;
            }
        }
        if (!alive) break;  // whole row dead (src/gapped_extend.rs:407-409)
        lo = nlo;
        hi = nhi;
        pjl = jl;
        pjh = jh;
        thr_prev = thr;
        cur ^= 1u;  // every read of the row now overwritten was before (1)
    }
    if (t == 0) {
        out.score = best_score;
        out.bi = bi;
        out.bj = bj;
        out.bstate = bst;
        out.rows = rows_done;
        out.truncated = truncated;
        out.cells = cells;
        res[sidx] = out;
    }
}

// The walk of a k_dp_band side (tb_walk's runs, one wavefront, reading the
// side's traceback rows and windows straight from device memory).
__global__ __launch_bounds__(64) void k_dp_traceback_band(const DpRes *__restrict__ res,
                                                          const BandSide *__restrict__ bsides, uint32_t nside,
                                                          const uint8_t *__restrict__ tb, const uint2 *__restrict__ wins,
                                                          const uint64_t *__restrict__ run_off,
                                                          gxy_edit_run *__restrict__ runs, uint32_t *__restrict__ nruns) {
    const uint32_t p = blockIdx.x;
    if (p >= nside) return;
    const DpRes r = res[p];
    const BandSide bs = bsides[p];
    const int lane = threadIdx.x;
    gxy_edit_run *const out = runs + run_off[p];
    uint32_t nr = 0;
    if (!r.truncated && !(r.bi == 0 && r.bj == 0) && r.rows != 0) {
        const uint8_t *const tb0 = tb + bs.tb;
        const uint2 *const win0 = wins + bs.win;
        uint32_t i = r.bi, j = r.bj, st = r.bstate;
        uint32_t cur_op = 0xFFFFFFFFu, cur_n = 0;
        const uint32_t cap = r.bi + r.bj;
        while (i > 0 || j > 0) {
            // a consistent traceback never leaves the swept cells
            if (i >= r.rows || (st == kStM && (i == 0 || j == 0)) || (st == kStX && i == 0) ||
                (st == kStY && j == 0) || st > kStY || nr >= cap)
                break;
            // lane k takes step k of a run of the current op (as tb_walk)
            const uint32_t dk = (uint32_t)lane;
            const uint32_t di = st == kStY ? 0u : dk, dj = st == kStX ? 0u : dk;
            bool valid = false;
            uint32_t next = kStY + 1u;
            if (di <= i && dj <= j) {
                const uint32_t ci = i - di, cj = j - dj;
                const bool edge_ok = st == kStM ? (ci > 0 && cj > 0) : st == kStX ? ci > 0 : cj > 0;
                const uint2 w = win0[ci];
                if (edge_ok && cj >= w.x && cj <= w.y) {
                    valid = true;
                    const uint32_t idx = cj - w.x;
                    const uint32_t nb = (tb0[(uint64_t)ci * bs.rb + (idx >> 1)] >> (4 * (idx & 1u))) & 0xFu;
                    next = st == kStM ? (nb & 3u)
                                      : st == kStX ? (((nb >> 2) & 1u) ? kStX : kStM)
                                                   : (((nb >> 3) & 1u) ? kStY : kStM);
                }
            }
            const uint64_t ends = __ballot(!valid || next != st);
            uint32_t nstep;
            bool stop = false;
            uint32_t nst = st;
            if (ends == 0) {
                nstep = 64;
            } else {
                const uint32_t kf = (uint32_t)__ffsll((long long)ends) - 1u;
                const bool vf = (__ballot(valid) >> kf) & 1u;
                if (vf) {
                    nstep = kf + 1u;
                    nst = (uint32_t)__shfl((int)next, (int)kf);
                } else {
                    nstep = kf;
                    stop = true;
                }
            }
            const uint32_t op = st == kStM ? GXY_OP_MATCH : st == kStX ? GXY_OP_DELETE : GXY_OP_INSERT;
            if (nstep) {
                if (op == cur_op) {
                    cur_n += nstep;
                } else {
                    if (cur_n) {
                        if (nr >= cap) break;
                        if (lane == 0) out[nr] = gxy_edit_run{cur_op, cur_n};
                        nr++;
                    }
                    cur_op = op;
                    cur_n = nstep;
                }
                if (st != kStY) i -= nstep;
                if (st != kStX) j -= nstep;
            }
            if (stop) break;
            st = nst;
        }
        if (cur_n && nr < cap) {
            if (lane == 0) out[nr] = gxy_edit_run{cur_op, cur_n};
            nr++;
        }
    }
    if (lane == 0) nruns[p] = nr;
}

template <class S>
__global__ void k_choose_anchor(const gxy_hsp *__restrict__ hsps, uint64_t n, const Blk *const *__restrict__ Ts,
                                const Blk *const *__restrict__ Qs, S sc, uint32_t window,
                                gxy_anchor *__restrict__ out) {
    const uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= n) return;
    const Blk *T = Ts[x];
    const Blk *Q = Qs[x];
    const gxy_hsp h = hsps[x];
    const uint32_t len = h.length;
    const uint32_t w = min(max(window, 1u), len);
    auto col = [&](uint32_t k) {
        const uint64_t ti = (uint64_t)h.t_start + k, qi = (uint64_t)h.q_start + k;
        return sc(base_code(T, ti), base_code(Q, qi), base_valid(T, ti) & base_valid(Q, qi));
    };
    int32_t best = 0;
    for (uint32_t k = 0; k < w; k++) best += col(k);
    int32_t win = best;
    uint32_t best_start = 0;
    for (uint32_t s = 1; s + w <= len; s++) {
        win = win - col(s - 1) + col(s + w - 1);
        if (win > best) {
            best = win;
            best_start = s;
        }
    }
    const uint32_t mid = best_start + w / 2;
    out[x] = {h.t_start + mid, h.q_start + mid, h};
}

template <class S>
__global__ void k_anchor_cols(const Blk *const *__restrict__ Ts, const Blk *const *__restrict__ Qs,
                              const gxy_anchor *__restrict__ a, uint64_t n, S sc, int32_t *__restrict__ out) {
    const uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= n) return;
    const Blk *T = Ts[x];
    const Blk *Q = Qs[x];
    const uint64_t t = a[x].t_pos, q = a[x].q_pos;
    out[x] = sc(base_code(T, t), base_code(Q, q), base_valid(T, t) & base_valid(Q, q));
}

inline unsigned grid_for(uint64_t n, unsigned th) { return (unsigned)((n + th - 1) / th); }

// DP forward kernels (GXY_DP_* in gxy.h): four wavefronts per side with the
// row in LDS, or one wavefront per side with the row in registers, 1/2/4
// columns per lane per slot, compiled for 3 or 4 wavefronts per SIMD.  Every
// variant gives the same results.
// Autotuner candidates, and the pass-size classes they are tuned for.
constexpr int kDpCand[gxy_ctx::kDpCands] = {GXY_DP_WAVE_K2_W4, GXY_DP_WAVE_K4_W4, GXY_DP_WAVE_K2, GXY_DP_WAVE_K4,
                                           GXY_DP_BLOCK4};
inline int dp_class(uint64_t sides) { return sides <= 8 ? 0 : sides <= 512 ? 1 : sides <= 8192 ? 2 : 3; }
// sides a probe pass of a big pass takes: enough to fill the device
constexpr uint64_t kDpProbeSides = 8192;
template <class S>
void launch_dp_forward(int variant, unsigned npp, hipStream_t st, const DpProb *probs, const uint32_t *list, S sc,
                       int32_t oe, int32_t e, int32_t ydrop, DpPool pool, DpRes *res) {
    switch (variant) {
    case GXY_DP_BLOCK4:
        k_dp_forward<S><<<npp, kDpThreads, 0, st>>>(probs, list, npp, sc, oe, e, ydrop, pool, res);
        break;
    case GXY_DP_WAVE_K1:
        k_dp_fwd1<S, 1, 3><<<npp, 64, 0, st>>>(probs, list, npp, sc, oe, e, ydrop, pool, res);
        break;
    case GXY_DP_WAVE_K4:
        k_dp_fwd1<S, 4, 3><<<npp, 64, 0, st>>>(probs, list, npp, sc, oe, e, ydrop, pool, res);
        break;
    case GXY_DP_WAVE_K2_W4:
        k_dp_fwd1<S, 2, 4><<<npp, 64, 0, st>>>(probs, list, npp, sc, oe, e, ydrop, pool, res);
        break;
    case GXY_DP_WAVE_K4_W4:
        k_dp_fwd1<S, 4, 4><<<npp, 64, 0, st>>>(probs, list, npp, sc, oe, e, ydrop, pool, res);
        break;
    default:
        k_dp_fwd1<S, 2, 3><<<npp, 64, 0, st>>>(probs, list, npp, sc, oe, e, ydrop, pool, res);
        break;
    }
}

// One-wavefront variants the persistent kernel runs, and their waves per SIMD
// (0: not run persistently).
inline int persist_wpe(int variant) {
    switch (variant) {
    case GXY_DP_WAVE_K2:
    case GXY_DP_WAVE_K4:
        return 3;
    case GXY_DP_WAVE_K2_W4:
    case GXY_DP_WAVE_K4_W4:
        return 4;
    default:
        return 0;
    }
}
template <class S>
void launch_dp_persist(int variant, unsigned grid, hipStream_t st, const DpProb *probs, const uint32_t *list,
                       uint32_t nlist, S sc, int32_t oe, int32_t e, int32_t ydrop, DpPool pool, DpRes *res,
                       PersistArgs pa) {
    switch (variant) {
    case GXY_DP_WAVE_K2:
        k_dp_persist<S, 2, 3><<<grid, 64, 0, st>>>(probs, list, nlist, sc, oe, e, ydrop, pool, res, pa);
        break;
    case GXY_DP_WAVE_K4:
        k_dp_persist<S, 4, 3><<<grid, 64, 0, st>>>(probs, list, nlist, sc, oe, e, ydrop, pool, res, pa);
        break;
    case GXY_DP_WAVE_K2_W4:
        k_dp_persist<S, 2, 4><<<grid, 64, 0, st>>>(probs, list, nlist, sc, oe, e, ydrop, pool, res, pa);
        break;
    default:
        k_dp_persist<S, 4, 4><<<grid, 64, 0, st>>>(probs, list, nlist, sc, oe, e, ydrop, pool, res, pa);
        break;
    }
}

// The DP's per-row score table packs the target row and the ambiguity score
// as signed bytes; other matrices use the int32 row table.
bool dp_scores_i8(const gxy_scoring &m) { return scoring_is_i8(m) && m.ambig >= -128 && m.ambig <= 127; }

uint32_t band_width_host(int32_t oe, int32_t e, int32_t ydrop, uint64_t m) {
    uint64_t br;
    if (e > 0) {
        int64_t d = (int64_t)ydrop - (int64_t)oe;
        if (d < INT32_MIN) d = INT32_MIN;
        const int32_t r = (int32_t)d / e;
        br = (uint64_t)(std::max(r, 0) + 2);
    } else {
        br = m + 1;
    }
    return (uint32_t)std::min<uint64_t>(4 * br + 4, m + 1);
}


// The sides whose band is wider than kMaxW, through k_dp_band in groups that
// fit a memory budget: each side reserves traceback rows for its expected
// sweep (est_v) and, if it outruns them, is swept again with room for all of
// its rows.  Results, runs and cells as the passes of gapped_batch_dev give
// them.
int band_sides_dev(gxy_ctx *ctx, const std::vector<DpProb> &probs, const std::vector<uint32_t> &wide,
                   const std::vector<uint32_t> &est_v, const gxy_scoring &sc, int32_t oe, int32_t e, int32_t ydrop,
                   std::vector<DpRes> *res, std::vector<std::vector<gxy_edit_run>> *run_store,
                   std::vector<const gxy_edit_run *> *side_ptr, std::vector<uint32_t> *side_n, uint64_t *cells) {
    const uint64_t np = probs.size();
    auto *dprob = (DpProb *)scratch(ctx, kSlotDpProb, np * sizeof(DpProb));
    if (!dprob) return set_err(ctx, GXY_ENOMEM, "DP scratch allocation failed");
    GXY_HIP(ctx, hipMemcpyAsync(dprob, probs.data(), np * sizeof(DpProb), hipMemcpyHostToDevice, ctx->stream));
    size_t free_b = 0, total_b = 0;
    GXY_HIP(ctx, hipMemGetInfo(&free_b, &total_b));
    const uint64_t budget = std::max<uint64_t>(256ull << 20, std::min<uint64_t>(16ull << 30, free_b / 4));
    std::vector<uint32_t> cap(np, 0);
    for (uint32_t x : wide) cap[x] = (uint32_t)std::min<uint64_t>((uint64_t)probs[x].n + 1, (uint64_t)est_v[x] + 64);
    auto bw_of = [&](uint32_t x) { return band_width_host(oe, e, ydrop, probs[x].m); };
    auto rb_of = [&](uint32_t x) { return (uint32_t)(((uint64_t)bw_of(x) + 1) / 2 + 15) & ~15u; };
    auto bytes_of = [&](uint32_t x) {
        return (uint64_t)cap[x] * (rb_of(x) + sizeof(uint2)) + 2ull * bw_of(x) * (sizeof(int2) + 1);
    };
    std::vector<uint32_t> todo = wide;
    while (!todo.empty()) {
        // a group within the budget (at least one side)
        size_t ng = 0;
        uint64_t used = 0;
        while (ng < todo.size() && (ng == 0 || used + bytes_of(todo[ng]) <= budget)) used += bytes_of(todo[ng++]);
        std::vector<BandSide> bs(ng);
        uint64_t tb_b = 0, nwin = 0, ncell = 0;
        for (size_t y = 0; y < ng; y++) {
            const uint32_t x = todo[y];
            bs[y].tb = tb_b;
            bs[y].win = nwin;
            bs[y].rowbuf = ncell;
            bs[y].cap_rows = cap[x];
            bs[y].rb = rb_of(x);
            tb_b += (uint64_t)cap[x] * bs[y].rb;
            nwin += cap[x];
            ncell += 2ull * bw_of(x);
        }
        // side descriptors | list | results | run offsets | run counts
        const size_t o_bs = 0, o_list = o_bs + ng * sizeof(BandSide), o_res = (o_list + ng * 4 + 15) & ~15ull,
                     o_roff = o_res + ng * sizeof(DpRes), o_nr = o_roff + ng * 8, o_end = o_nr + ng * 4;
        auto *work = (uint8_t *)scratch(ctx, kSlotDpOut, o_end);
        auto *dtb = (uint8_t *)scratch(ctx, kSlotDpTb, std::max<uint64_t>(tb_b, 16));
        auto *dwin = (uint2 *)scratch(ctx, kSlotDpRows, nwin * sizeof(uint2));
        auto *drows = (uint8_t *)scratch(ctx, kSlotDpPersist, ncell * (sizeof(int2) + 1));
        if (!work || !dtb || !dwin || !drows)
            return set_err(ctx, GXY_ENOMEM, "gapped extension of a %u-cell band: device allocation failed",
                           bw_of(todo[0]));
        auto *dbs = (BandSide *)(work + o_bs);
        auto *dlist = (uint32_t *)(work + o_list);
        auto *dres = (DpRes *)(work + o_res);
        auto *droff = (uint64_t *)(work + o_roff);
        auto *dnr = (uint32_t *)(work + o_nr);
        GXY_HIP(ctx, hipMemcpyAsync(dbs, bs.data(), ng * sizeof(BandSide), hipMemcpyHostToDevice, ctx->stream));
        GXY_HIP(ctx, hipMemcpyAsync(dlist, todo.data(), ng * 4, hipMemcpyHostToDevice, ctx->stream));
        {
            KernelTimer kt_(ctx, "dp_forward");
            if (dp_scores_i8(sc))
                k_dp_band<ScorerI8><<<(unsigned)ng, kDpThreads, 0, ctx->stream>>>(
                    dprob, dlist, dbs, (uint32_t)ng, make_scorer_i8(sc), oe, e, ydrop, dtb, dwin, (int2 *)drows,
                    drows + ncell * sizeof(int2), dres);
            else
                k_dp_band<ScorerI32><<<(unsigned)ng, kDpThreads, 0, ctx->stream>>>(
                    dprob, dlist, dbs, (uint32_t)ng, make_scorer_i32(sc), oe, e, ydrop, dtb, dwin, (int2 *)drows,
                    drows + ncell * sizeof(int2), dres);
        }
        GXY_LAUNCH_CHECK(ctx);
        std::vector<DpRes> pr(ng);
        GXY_HIP(ctx, hipMemcpyAsync(pr.data(), dres, ng * sizeof(DpRes), hipMemcpyDeviceToHost, ctx->stream));
        GXY_HIP(ctx, hipStreamSynchronize(ctx->stream));
        // walk the finished sides back into bi + bj reserved runs each
        std::vector<uint64_t> roff(ng, 0);
        uint64_t total_runs = 0;
        for (size_t y = 0; y < ng; y++) {
            roff[y] = total_runs;
            if (!pr[y].truncated) total_runs += (uint64_t)pr[y].bi + pr[y].bj;
        }
        std::vector<uint32_t> nr(ng, 0);
        std::vector<gxy_edit_run> sr;
        std::vector<uint64_t> poff(ng, 0);
        if (total_runs > 0) {
            auto *druns = (gxy_edit_run *)scratch(ctx, kSlotDpRuns, total_runs * sizeof(gxy_edit_run));
            if (!druns) return set_err(ctx, GXY_ENOMEM, "scratch allocation failed");
            GXY_HIP(ctx, hipMemcpyAsync(droff, roff.data(), ng * 8, hipMemcpyHostToDevice, ctx->stream));
            {
                KernelTimer kt_(ctx, "dp_traceback");
                k_dp_traceback_band<<<(unsigned)ng, 64, 0, ctx->stream>>>(dres, dbs, (uint32_t)ng, dtb, dwin, droff,
                                                                          druns, dnr);
            }
            GXY_LAUNCH_CHECK(ctx);
            GXY_HIP(ctx, hipMemcpyAsync(nr.data(), dnr, ng * 4, hipMemcpyDeviceToHost, ctx->stream));
            GXY_HIP(ctx, hipStreamSynchronize(ctx->stream));
            std::vector<gxy_edit_run> all(total_runs);
            GXY_HIP(ctx, hipMemcpyAsync(all.data(), druns, total_runs * sizeof(gxy_edit_run), hipMemcpyDeviceToHost,
                                        ctx->stream));
            GXY_HIP(ctx, hipStreamSynchronize(ctx->stream));
            uint64_t packed = 0;
            for (size_t y = 0; y < ng; y++) {
                poff[y] = packed;
                packed += nr[y];
            }
            sr.resize(packed);
            for (size_t y = 0; y < ng; y++)
                std::copy(all.begin() + roff[y], all.begin() + roff[y] + nr[y], sr.begin() + poff[y]);
        }
        std::vector<uint32_t> again;
        for (size_t y = 0; y < ng; y++) {
            const uint32_t x = todo[y];
            if (pr[y].truncated) {  // more rows than reserved: again, with room for every row
                cap[x] = probs[x].n + 1;
                again.push_back(x);
                ctx->dp_truncated++;
                continue;
            }
            *cells += pr[y].cells;
            (*res)[x] = pr[y];
            (*side_ptr)[x] = sr.data() + poff[y];
            (*side_n)[x] = nr[y];
        }
        run_store->push_back(std::move(sr));
        todo.erase(todo.begin(), todo.begin() + ng);
        todo.insert(todo.begin(), again.begin(), again.end());
    }
    return GXY_OK;
}
}  // namespace

namespace gxy {

// Batched gapped_extend::extend over independent (target, query, anchor)
// items held on the device.  Results in item order; runs appended to *runs.
int gapped_batch_dev(gxy_ctx *ctx, const std::vector<GappedItem> &items, const gxy_scoring &sc,
                     const gxy_gapped_params &gp, std::vector<GappedResult> *results,
                     std::vector<gxy_edit_run> *runs, uint64_t *cells_out) {
    const uint64_t n = items.size();
    results->assign(n, GappedResult{false, {}});
    if (n == 0) return GXY_OK;
    // host phase times, printed when GXY_TRACE is set: setup, persistent
    // launch (kernel), its results on the host, passes, script assembly,
    // anchor columns (part of setup)
    double tr_ms[6] = {};
    auto tr_t = std::chrono::steady_clock::now();
    auto lap = [&](int k) {
        if (!ctx->trace) return;
        const auto now = std::chrono::steady_clock::now();
        tr_ms[k] += std::chrono::duration<double, std::milli>(now - tr_t).count();
        tr_t = now;
    };
    const int32_t oe = sc.gap_open + sc.gap_extend, e = sc.gap_extend;
    // each item's target and query slice
    auto tlen_of = [](const GappedItem &it) { return it.t_len == kWholeSeq ? it.t->len - it.t_off : it.t_len; };
    auto qlen_of = [](const GappedItem &it) { return it.q_len == kWholeSeq ? it.q->len - it.q_off : it.q_len; };
    for (const auto &it : items) {
        if (it.t_off > it.t->len || it.q_off > it.q->len || tlen_of(it) > it.t->len - it.t_off ||
            qlen_of(it) > it.q->len - it.q_off)
            return set_err(ctx, GXY_EINVAL, "pair slice outside its sequences");
        if ((tlen_of(it) >> 32) || (qlen_of(it) >> 32))
            return set_err(ctx, GXY_ERANGE, "a gapped pair of 2^32 bases or more");
        if (it.a.t_pos >= tlen_of(it) || it.a.q_pos >= qlen_of(it))
            return set_err(ctx, GXY_EINVAL, "anchor outside its sequences");
    }
    // Two side-problems per item: 2k = right, 2k+1 = left.
    const uint64_t np = 2 * n;
    std::vector<DpProb> probs(np);
    for (uint64_t k = 0; k < n; k++) {
        const auto &it = items[k];
        const uint64_t at = it.a.t_pos, aq = it.a.q_pos;
        DpProb &r = probs[2 * k];
        r = DpProb{};
        r.T = it.t->d;
        r.Q = it.q->d;
        r.t0 = it.t_off + at + 1;
        r.q0 = it.q_off + aq + 1;
        r.tdir = 1;
        r.qdir = 1;
        r.n = (uint32_t)(tlen_of(it) - at - 1);
        r.m = (uint32_t)(qlen_of(it) - aq - 1);
        DpProb &l = probs[2 * k + 1];
        l = DpProb{};
        l.T = it.t->d;
        l.Q = it.q->d;
        l.t0 = it.t_off + (at >= 1 ? at - 1 : 0);
        l.q0 = it.q_off + (aq >= 1 ? aq - 1 : 0);
        l.tdir = -1;
        l.qdir = -1;
        l.n = (uint32_t)at;
        l.m = (uint32_t)aq;
    }
    // anchor column scores
    std::vector<const Blk *> tp(2 * n);
    std::vector<gxy_anchor> anc(n);
    for (uint64_t k = 0; k < n; k++) {
        // the slice's first block, and the anchor relative to it
        tp[k] = items[k].t->d + (items[k].t_off >> 5);
        tp[n + k] = items[k].q->d + (items[k].q_off >> 5);
        anc[k] = items[k].a;
        anc[k].t_pos += (uint32_t)(items[k].t_off & 31);
        anc[k].q_pos += (uint32_t)(items[k].q_off & 31);
    }
    auto *dptrs = (const Blk **)scratch(ctx, kSlotMisc0, 2 * n * sizeof(void *));
    auto *danc = (gxy_anchor *)scratch(ctx, kSlotMisc1, n * sizeof(gxy_anchor));
    auto *dcol = (int32_t *)scratch(ctx, kSlotMisc2, n * 4);
    if (!dptrs || !danc || !dcol) return set_err(ctx, GXY_ENOMEM, "scratch allocation failed");
    GXY_HIP(ctx, hipMemcpyAsync(dptrs, tp.data(), 2 * n * sizeof(void *), hipMemcpyHostToDevice, ctx->stream));
    GXY_HIP(ctx, hipMemcpyAsync(danc, anc.data(), n * sizeof(gxy_anchor), hipMemcpyHostToDevice, ctx->stream));
    if (scoring_is_i8(sc))
        k_anchor_cols<ScorerI8><<<grid_for(n, 256), 256, 0, ctx->stream>>>(dptrs, dptrs + n, danc, n,
                                                                           make_scorer_i8(sc), dcol);
    else
        k_anchor_cols<ScorerI32><<<grid_for(n, 256), 256, 0, ctx->stream>>>(dptrs, dptrs + n, danc, n,
                                                                            make_scorer_i32(sc), dcol);
    GXY_LAUNCH_CHECK(ctx);
    std::vector<int32_t> col(n);
    GXY_HIP(ctx, hipMemcpyAsync(col.data(), dcol, n * 4, hipMemcpyDeviceToHost, ctx->stream));
    GXY_HIP(ctx, hipStreamSynchronize(ctx->stream));

    lap(5);
    std::vector<DpRes> res(np);
    // each side's runs: side_ptr[x][0 .. side_n[x]), inside one of the run
    // buffers downloaded by the passes (kept until the scripts are assembled)
    std::vector<std::vector<gxy_edit_run>> run_store;
    std::vector<const gxy_edit_run *> side_ptr(np, nullptr);
    std::vector<uint32_t> side_n(np, 0);
    std::vector<uint32_t> pending, wide;  // wide: bands over kMaxW cells, swept by k_dp_band
    uint64_t need_rows = 0;
    for (uint64_t x = 0; x < np; x++) {
        if (probs[x].n == 0 || probs[x].m == 0) {
            res[x] = DpRes{0, 0, 0, kStM, 0, 0, kNoChunk, 0, 0};
        } else if (band_width_host(oe, e, gp.y_drop, probs[x].m) > (uint32_t)kMaxW) {
            wide.push_back((uint32_t)x);
        } else {
            pending.push_back((uint32_t)x);
            need_rows += std::min<uint64_t>((uint64_t)probs[x].n + 1, 1u << 20);
        }
    }
    uint64_t cells = 0;
    // Rows a side is expected to sweep: an alignment's target span stays near
    // its query span, and the band dies about y_drop / gap_extend rows past
    // the last good cell, so a contig against a long target does not count
    // the whole rest of the target.  Only pass packing uses it: a side that
    // outruns it is truncated when the pool runs out and swept again.
    std::vector<uint32_t> est_v(np);  // < 2^32: n < 2^32
    for (uint64_t x = 0; x < np; x++) {
        const uint64_t m = probs[x].m;
        est_v[x] = (uint32_t)std::min<uint64_t>((uint64_t)probs[x].n + 1, m + m / 8 + 512);
    }
    auto est_rows = [&](uint32_t x) -> uint64_t { return est_v[x]; };
    // longest first, ties in their order in v: a stable LSD radix sort on the
    // complemented estimate, two 16-bit digits
    auto longest_first = [&](std::vector<uint32_t> &v) {
        std::vector<uint32_t> tmp(v.size()), cnt(65537);
        for (int pass = 0; pass < 2; pass++) {
            auto digit = [&](uint32_t x) { return ((~est_v[x]) >> (16 * pass)) & 0xFFFFu; };
            std::fill(cnt.begin(), cnt.end(), 0u);
            for (uint32_t x : v) cnt[digit(x) + 1]++;
            for (int k = 0; k < 65536; k++) cnt[k + 1] += cnt[k];
            for (uint32_t x : v) tmp[cnt[digit(x)]++] = x;
            v.swap(tmp);
        }
    };
    if (!pending.empty()) {
        // longest sides first, so the long sweeps start early instead of
        // forming the tail of a pass
        longest_first(pending);
        // The traceback pool: up to a third of the device (at most 96 GiB),
        // no more than every side could use.  Its size bounds how many long
        // sweeps run at once (rows in flight x kRowBytes).
        size_t free_b = 0, total_b = 0;
        GXY_HIP(ctx, hipMemGetInfo(&free_b, &total_b));
        const uint64_t cap_b = ctx->dp_pool_cap ? ctx->dp_pool_cap : std::min<uint64_t>(96ull << 30, total_b / 3);
        // a row cap set by the caller (gxy_ctx_set_dp_pool_rows) bounds the pool
        // even where the slot is larger from earlier calls
        const uint64_t row_cap = ctx->dp_pool_rows ? std::max<uint64_t>(16, ctx->dp_pool_rows / kChunkRows) : ~0ull;
        uint64_t want = std::min<uint64_t>(std::min<uint64_t>(cap_b / kChunkBytes, row_cap),
                                           (need_rows + kChunkRows - 1) / kChunkRows + pending.size());
        want = std::max<uint64_t>(want, 16);
        uint8_t *dtb = nullptr;
        while (!(dtb = (uint8_t *)scratch(ctx, kSlotDpTb, want * kChunkBytes)) && want > 16) want /= 2;
        if (!dtb) return set_err(ctx, GXY_ENOMEM, "DP traceback pool allocation failed");
        const uint64_t nchunks =
            std::min<uint64_t>(std::min<uint64_t>(ctx->slot_bytes[kSlotDpTb] / kChunkBytes, row_cap), 0xFFFFFFF0ull);
        auto *dwin = (uint8_t *)scratch(ctx, kSlotDpRows, nchunks * (kChunkRows * sizeof(uint2) + 4));
        auto *dprob = (DpProb *)scratch(ctx, kSlotDpProb, np * sizeof(DpProb));
        if (!dwin || !dprob) return set_err(ctx, GXY_ENOMEM, "DP scratch allocation failed");
        DpPool pool;
        pool.tb = dtb;
        pool.win = (uint2 *)dwin;
        pool.prev = (uint32_t *)(dwin + nchunks * kChunkRows * sizeof(uint2));
        pool.nchunks = (uint32_t)nchunks;
        GXY_HIP(ctx, hipMemcpyAsync(dprob, probs.data(), np * sizeof(DpProb), hipMemcpyHostToDevice, ctx->stream));
        // a side's estimated chunks, at most 1/32 of the pool
        const uint64_t est_cap = std::max<uint64_t>(1, nchunks / 32);
        auto est = [&](uint32_t x) {
            return std::min<uint64_t>((est_rows(x) + kChunkRows - 1) / kChunkRows + 1, est_cap);
        };
        lap(0);
        bool alone = false;  // run only the first pending side (progress guarantee)
        bool persist_done = false;  // at most one persistent launch per call
        // pending[head ..] waits; a pass takes its sides from the head, and the
        // sides it truncates go back just before the new head (first in line)
        size_t head = 0;
        while (head < pending.size()) {
            // take sides in order while their estimates fit the pool
            uint64_t npp = 0, used = 0;
            while (head + npp < pending.size() && npp < (1u << 20) &&
                   (npp == 0 || used + est(pending[head + npp]) <= nchunks)) {
                used += est(pending[head + npp]);
                npp++;
                if (alone) break;
            }
            // the forward kernel: the caller's, the class's tuned one, or the
            // next candidate to measure (a big pass lends a probe-sized slice)
            int variant = ctx->dp_kernel;
            gxy_ctx::DpTune *tune = nullptr;
            if (!variant) {
                // A pass whose longest side holds more rows than an even share of
                // the pass's rows over the one-wavefront slots the GPU runs at
                // once is bound by that side's row latency, which the
                // four-wavefront kernel halves: no probe, no one-wavefront kernel.
                uint64_t rows_sum = 0;
                for (uint64_t y = 0; y < npp; y++) rows_sum += est_rows(pending[head + y]);
                if (est_rows(pending[head]) * (uint64_t)(ctx_cus(ctx) * 16) >= rows_sum) variant = GXY_DP_BLOCK4;
            }
            if (!variant) {
                tune = &ctx->dp_tune[dp_class(npp)];
                if (tune->chosen) {
                    variant = tune->chosen;
                    tune = nullptr;
                } else {
                    variant = kDpCand[tune->tried];
                    if (npp > 2 * kDpProbeSides) npp = kDpProbeSides;
                }
            }
            // Once the kernel is settled on a one-wavefront variant, the rest of a
            // large batch (or any batch, when the caller asks) runs as one
            // persistent launch; what it leaves goes on through passes.
            if (!tune && !persist_done && persist_wpe(variant) && ctx->dp_persist >= 0 &&
                (ctx->dp_persist > 0 || pending.size() - head >= kPersistMinSides)) {
                persist_done = true;
                const uint64_t slots = (uint64_t)ctx_cus(ctx) * 4 * persist_wpe(variant);
                uint64_t grid = std::min<uint64_t>(slots, pending.size() - head);
                uint64_t region = nchunks / grid;
                if (region < 2) {  // a small pool: fewer blocks, two chunks each
                    region = 2;
                    grid = nchunks / 2;
                }
                // sides expected to fit a block's chunks; the others stay for passes
                std::vector<uint32_t> plist, rest;
                uint64_t prow = 0;
                for (size_t y = head; y < pending.size(); y++) {
                    const uint32_t x = pending[y];
                    if (est_rows(x) + kChunkRows <= region * kChunkRows) {
                        plist.push_back(x);
                        prow += est_rows(x);
                    } else {
                        rest.push_back(x);
                    }
                }
                if (!plist.empty()) {
                    const uint64_t pn = plist.size();
                    grid = std::min<uint64_t>(grid, pn);
                    // results | run offsets | run counts | side list | counters
                    const size_t q_res = 0, q_roff = q_res + pn * sizeof(DpRes), q_nr = q_roff + pn * 8,
                                 q_list = q_nr + pn * 4, q_ctr = (q_list + pn * 4 + 255) & ~255ull;
                    auto *pw = (uint8_t *)scratch(ctx, kSlotDpPersist, q_ctr + 256);
                    // room for the runs: a side's runs are far fewer than its rows
                    uint64_t run_cap = std::max<uint64_t>(prow / 16, 1u << 16);
                    gxy_edit_run *pruns = nullptr;
                    while (!(pruns = (gxy_edit_run *)scratch(ctx, kSlotDpRuns, run_cap * sizeof(gxy_edit_run))) &&
                           run_cap > (1u << 16))
                        run_cap /= 2;
                    if (!pw || !pruns) return set_err(ctx, GXY_ENOMEM, "DP scratch allocation failed");
                    auto *dres = (DpRes *)(pw + q_res);
                    PersistArgs pa;
                    pa.head = (uint32_t *)(pw + q_ctr);
                    pa.region = (uint32_t)region;
                    pa.run_ctr = (unsigned long long *)(pw + q_ctr + 8);
                    pa.runs = pruns;
                    pa.run_cap = run_cap;
                    pa.run_off = (uint64_t *)(pw + q_roff);
                    pa.nruns = (uint32_t *)(pw + q_nr);
                    auto *dlist = (uint32_t *)(pw + q_list);
                    GXY_HIP(ctx, hipMemcpyAsync(dlist, plist.data(), pn * 4, hipMemcpyHostToDevice, ctx->stream));
                    GXY_HIP(ctx, hipMemsetAsync(pw + q_ctr, 0, 16, ctx->stream));
                    {
                        KernelTimer kt_(ctx, "dp_persist");
                        if (dp_scores_i8(sc))
                            launch_dp_persist(variant, (unsigned)grid, ctx->stream, dprob, dlist, (uint32_t)pn,
                                              make_scorer_i8(sc), oe, e, gp.y_drop, pool, dres, pa);
                        else
                            launch_dp_persist(variant, (unsigned)grid, ctx->stream, dprob, dlist, (uint32_t)pn,
                                              make_scorer_i32(sc), oe, e, gp.y_drop, pool, dres, pa);
                    }
                    GXY_LAUNCH_CHECK(ctx);
                    GXY_HIP(ctx, hipStreamSynchronize(ctx->stream));
                    lap(1);
                    std::vector<DpRes> pr(pn);
                    std::vector<uint64_t> proff(pn);
                    std::vector<uint32_t> pnr(pn);
                    unsigned long long nruns_all = 0;
                    GXY_HIP(ctx, hipMemcpyAsync(pr.data(), dres, pn * sizeof(DpRes), hipMemcpyDeviceToHost, ctx->stream));
                    GXY_HIP(ctx, hipMemcpyAsync(proff.data(), pa.run_off, pn * 8, hipMemcpyDeviceToHost, ctx->stream));
                    GXY_HIP(ctx, hipMemcpyAsync(pnr.data(), pa.nruns, pn * 4, hipMemcpyDeviceToHost, ctx->stream));
                    GXY_HIP(ctx, hipMemcpyAsync(&nruns_all, pa.run_ctr, 8, hipMemcpyDeviceToHost, ctx->stream));
                    GXY_HIP(ctx, hipStreamSynchronize(ctx->stream));
                    const uint64_t got = std::min<uint64_t>(nruns_all, run_cap);
                    std::vector<gxy_edit_run> prun(got);
                    if (got) {
                        GXY_HIP(ctx, hipMemcpyAsync(prun.data(), pruns, got * sizeof(gxy_edit_run),
                                                    hipMemcpyDeviceToHost, ctx->stream));
                        GXY_HIP(ctx, hipStreamSynchronize(ctx->stream));
                    }
                    ctx->dp_persist_launches++;
                    ctx->dp_persist_sides += pn;
                    for (uint64_t y = 0; y < pn; y++) {
                        const uint32_t x = plist[y];
                        if (pr[y].truncated) {  // outgrew its block's chunks, or no room for its runs
                            rest.push_back(x);
                            ctx->dp_persist_deferred++;
                            continue;
                        }
                        if (pnr[y] && proff[y] + pnr[y] > got)
                            return set_err(ctx, GXY_EDEVICE, "persistent DP: runs outside the run array");
                        cells += pr[y].cells;
                        ctx->dp_persist_cells += pr[y].cells;
                        res[x] = pr[y];
                        side_ptr[x] = prun.data() + proff[y];
                        side_n[x] = pnr[y];
                    }
                    run_store.push_back(std::move(prun));
                    // the rest (longest first) goes through passes
                    longest_first(rest);
                    pending.swap(rest);
                    head = 0;
                    lap(2);
                    continue;
                }
            }
            // per-pass arrays: results | run offsets | packed offsets | run counts | side list | counter
            const size_t o_res = 0, o_roff = o_res + npp * sizeof(DpRes), o_poff = o_roff + npp * 8,
                         o_nr = o_poff + npp * 8, o_list = o_nr + npp * 4, o_ctr = (o_list + npp * 4 + 255) & ~255ull;
            auto *work = (uint8_t *)scratch(ctx, kSlotDpOut, o_ctr + 256);
            if (!work) return set_err(ctx, GXY_ENOMEM, "DP scratch allocation failed");
            auto *dres = (DpRes *)(work + o_res);
            auto *droff = (uint64_t *)(work + o_roff);
            auto *dpoff = (uint64_t *)(work + o_poff);
            auto *dnr = (uint32_t *)(work + o_nr);
            auto *dlist = (uint32_t *)(work + o_list);
            pool.next = (uint32_t *)(work + o_ctr);
            GXY_HIP(ctx, hipMemcpyAsync(dlist, pending.data() + head, npp * 4, hipMemcpyHostToDevice, ctx->stream));
            GXY_HIP(ctx, hipMemsetAsync(pool.next, 0, 4, ctx->stream));
            if (tune && !ctx->dp_ev[0]) {
                GXY_HIP(ctx, hipEventCreate(&ctx->dp_ev[0]));
                GXY_HIP(ctx, hipEventCreate(&ctx->dp_ev[1]));
            }
            if (tune) GXY_HIP(ctx, hipEventRecord(ctx->dp_ev[0], ctx->stream));
            {
                KernelTimer kt_(ctx, "dp_forward");
                if (dp_scores_i8(sc))
                    launch_dp_forward(variant, (unsigned)npp, ctx->stream, dprob, dlist, make_scorer_i8(sc), oe, e,
                                      gp.y_drop, pool, dres);
                else
                    launch_dp_forward(variant, (unsigned)npp, ctx->stream, dprob, dlist, make_scorer_i32(sc), oe, e,
                                      gp.y_drop, pool, dres);
            }
            GXY_LAUNCH_CHECK(ctx);
            if (tune) GXY_HIP(ctx, hipEventRecord(ctx->dp_ev[1], ctx->stream));
            std::vector<DpRes> pr(npp);
            GXY_HIP(ctx, hipMemcpyAsync(pr.data(), dres, npp * sizeof(DpRes), hipMemcpyDeviceToHost, ctx->stream));
            GXY_HIP(ctx, hipStreamSynchronize(ctx->stream));
            if (tune) {  // the candidate's rate on this pass; all measured: keep the fastest
                float ms = 0;
                GXY_HIP(ctx, hipEventElapsedTime(&ms, ctx->dp_ev[0], ctx->dp_ev[1]));
                uint64_t pcells = 0;
                for (uint64_t y = 0; y < npp; y++) pcells += pr[y].cells;
                tune->gcups[tune->tried] = ms > 0 ? (double)pcells / (ms * 1e6) : 0.0;
                if (++tune->tried == gxy_ctx::kDpCands) {
                    int b = 0;
                    for (int c = 1; c < gxy_ctx::kDpCands; c++)
                        if (tune->gcups[c] > tune->gcups[b]) b = c;
                    tune->chosen = kDpCand[b];
                }
            }
            // reserve bi + bj runs per completed side and walk them back
            std::vector<uint64_t> roff(npp, 0);
            uint64_t total_runs = 0, done = 0;
            for (uint64_t y = 0; y < npp; y++) {
                roff[y] = total_runs;
                if (!pr[y].truncated) {
                    total_runs += (uint64_t)pr[y].bi + pr[y].bj;
                    done++;
                }
            }
            std::vector<uint32_t> nr(npp, 0);
            std::vector<uint64_t> poff(npp, 0);
            std::vector<gxy_edit_run> sr;
            if (total_runs > 0) {
                auto *druns = (gxy_edit_run *)scratch(ctx, kSlotDpRuns, total_runs * sizeof(gxy_edit_run));
                if (!druns) return set_err(ctx, GXY_ENOMEM, "scratch allocation failed");
                GXY_HIP(ctx, hipMemcpyAsync(droff, roff.data(), npp * 8, hipMemcpyHostToDevice, ctx->stream));
                {
                    KernelTimer kt_(ctx, "dp_traceback");
                    k_dp_traceback<<<(unsigned)npp, 64, 0, ctx->stream>>>(dres, (uint32_t)npp, pool, droff, druns,
                                                                          dnr);
                }
                GXY_LAUNCH_CHECK(ctx);
                GXY_HIP(ctx, hipMemcpyAsync(nr.data(), dnr, npp * 4, hipMemcpyDeviceToHost, ctx->stream));
                GXY_HIP(ctx, hipStreamSynchronize(ctx->stream));
                uint64_t packed = 0;
                for (uint64_t y = 0; y < npp; y++) {
                    poff[y] = packed;
                    packed += nr[y];
                }
                if (packed > 0) {
                    auto *dpk = (gxy_edit_run *)scratch(ctx, kSlotMisc3, packed * sizeof(gxy_edit_run));
                    if (!dpk) return set_err(ctx, GXY_ENOMEM, "scratch allocation failed");
                    GXY_HIP(ctx, hipMemcpyAsync(dpoff, poff.data(), npp * 8, hipMemcpyHostToDevice, ctx->stream));
                    k_pack_runs<<<(unsigned)npp, 64, 0, ctx->stream>>>(droff, dpoff, dnr, (uint32_t)npp, druns, dpk);
                    GXY_LAUNCH_CHECK(ctx);
                    sr.resize(packed);
                    GXY_HIP(ctx, hipMemcpyAsync(sr.data(), dpk, packed * sizeof(gxy_edit_run), hipMemcpyDeviceToHost,
                                                ctx->stream));
                    GXY_HIP(ctx, hipStreamSynchronize(ctx->stream));
                }
            }
            std::vector<uint32_t> again;  // truncated: swept again
            ctx->dp_passes++;
            ctx->dp_alone += alone ? 1 : 0;
            for (uint64_t y = 0; y < npp; y++) {
                const uint32_t x = pending[head + y];
                if (pr[y].truncated) {
                    again.push_back(x);
                    ctx->dp_truncated++;
                    continue;
                }
                cells += pr[y].cells;
                res[x] = pr[y];
                side_ptr[x] = sr.data() + poff[y];
                side_n[x] = nr[y];
            }
            run_store.push_back(std::move(sr));
            if (done == 0) {
                // nothing finished: the pool is shared by too many long sweeps, or
                // one sweep alone needs more than the pool
                if (alone)
                    return set_err(ctx, GXY_ENOMEM,
                                   "a gapped extension needs more than the %llu-row traceback pool",
                                   (unsigned long long)(nchunks * kChunkRows));
                alone = true;
            } else {
                alone = false;
            }
            head += npp - again.size();
            std::copy(again.begin(), again.end(), pending.begin() + head);
        }
    }
    if (!wide.empty()) {
        const int rc = band_sides_dev(ctx, probs, wide, est_v, sc, oe, e, gp.y_drop, &res, &run_store, &side_ptr,
                                      &side_n, &cells);
        if (rc != GXY_OK) return rc;
    }
    lap(3);
    if (cells_out) *cells_out += cells;

    // Assemble (src/gapped_extend.rs:107-133): score = col + L + R (saturating);
    // script = rev(L.script) + M1 + R.script.  The walk emits tip-to-root runs,
    // so the left side contributes its raw order and the right side reversed;
    // equal neighbouring ops merge (EditScript::push, src/edit_script.rs:65-74).
    // Items are independent: one pass counts each item's merged runs, a
    // prefix sum places them, a second pass writes them, both over threads.
    std::vector<uint64_t> item_off(n + 1, 0);
    auto for_items = [&](auto &&fn) {
        const uint64_t nt = std::min<uint64_t>(std::max(1u, std::thread::hardware_concurrency()),
                                               std::max<uint64_t>(1, n / 4096));
        if (nt <= 1) {
            fn(0, n);
            return;
        }
        std::vector<std::thread> th;
        for (uint64_t w = 0; w < nt; w++) th.emplace_back(fn, n * w / nt, n * (w + 1) / nt);
        for (auto &t : th) t.join();
    };
    std::vector<int32_t> total(n);
    // runs of item k in order, handed to `sink(op, n)` before merging
    auto item_runs = [&](uint64_t k, auto &&sink) {
        const gxy_edit_run *lr = side_ptr[2 * k + 1];
        for (uint32_t x = 0; x < side_n[2 * k + 1]; x++) sink(lr[x].op, lr[x].n);
        sink(GXY_OP_MATCH, 1);
        const gxy_edit_run *rr = side_ptr[2 * k];
        for (uint32_t x = side_n[2 * k]; x-- > 0;) sink(rr[x].op, rr[x].n);
    };
    for_items([&](uint64_t a, uint64_t b) {
        for (uint64_t k = a; k < b; k++) {
            const DpRes &R = res[2 * k], &L = res[2 * k + 1];
            int64_t t64 = (int64_t)col[k] + L.score;
            t64 = std::max<int64_t>(INT32_MIN, std::min<int64_t>(INT32_MAX, t64));
            t64 = std::max<int64_t>(INT32_MIN, std::min<int64_t>(INT32_MAX, t64 + R.score));
            total[k] = (int32_t)t64;
            uint64_t cnt = 0;
            if (total[k] >= gp.gapped_threshold) {
                uint32_t last = 0xFFFFFFFFu;
                item_runs(k, [&](uint32_t op, uint32_t c) {
                    if (c == 0) return;
                    if (op != last) cnt++;
                    last = op;
                });
            }
            item_off[k + 1] = cnt;
        }
    });
    for (uint64_t k = 0; k < n; k++) item_off[k + 1] += item_off[k];
    const uint64_t base = runs->size();
    runs->resize(base + item_off[n]);
    gxy_edit_run *dst = runs->data() + base;
    for_items([&](uint64_t a, uint64_t b) {
        for (uint64_t k = a; k < b; k++) {
            if (total[k] < gp.gapped_threshold) continue;
            const DpRes &R = res[2 * k], &L = res[2 * k + 1];
            GappedResult &g = (*results)[k];
            g.found = true;
            g.aln.score = total[k];
            g.aln.t_start = items[k].a.t_pos - L.bi;
            g.aln.q_start = items[k].a.q_pos - L.bj;
            g.aln.t_len = L.bi + 1 + R.bi;
            g.aln.q_len = L.bj + 1 + R.bj;
            g.aln.source = items[k].a;
            g.aln.runs_offset = base + item_off[k];
            gxy_edit_run *o = dst + item_off[k];
            uint64_t w = 0;
            item_runs(k, [&](uint32_t op, uint32_t c) {
                if (c == 0) return;
                if (w > 0 && o[w - 1].op == op)
                    o[w - 1].n += c;
                else
                    o[w++] = {op, c};
            });
            g.aln.n_runs = (uint32_t)w;
        }
    });
    lap(4);
    if (ctx->trace)
        std::fprintf(stderr, "[gxy dp] %llu items: setup %.1f (anchor columns %.1f) persistent %.1f "
                             "persistent-host %.1f passes %.1f assemble %.1f ms\n",
                     (unsigned long long)n, tr_ms[0] + tr_ms[5], tr_ms[5], tr_ms[1], tr_ms[2], tr_ms[3], tr_ms[4]);
    return GXY_OK;
}

int choose_anchor_pairs_dev(gxy_ctx *ctx, const gxy_dseq *const *ts, const gxy_dseq *const *qs,
                            const gxy_hsp *hsps_host, uint64_t n, const gxy_scoring &sc, uint32_t window,
                            gxy_anchor *out_host) {
    if (n == 0) return GXY_OK;
    std::vector<const Blk *> ptrs(2 * n);
    for (uint64_t k = 0; k < n; k++) {
        const gxy_hsp &h = hsps_host[k];
        if (h.length < 1) return set_err(ctx, GXY_EINVAL, "empty HSP cannot be anchored");  // src/anchor.rs:33
        if ((uint64_t)h.t_start + h.length > ts[k]->len || (uint64_t)h.q_start + h.length > qs[k]->len)
            return set_err(ctx, GXY_EINVAL, "HSP outside its sequences");
        ptrs[k] = ts[k]->d;
        ptrs[n + k] = qs[k]->d;
    }
    auto *dh = (gxy_hsp *)scratch(ctx, kSlotMisc2, n * sizeof(gxy_hsp));
    auto *da = (gxy_anchor *)scratch(ctx, kSlotMisc3, n * sizeof(gxy_anchor));
    auto *dp = (const Blk **)scratch(ctx, kSlotMisc1, 2 * n * sizeof(void *));
    if (!dh || !da || !dp) return set_err(ctx, GXY_ENOMEM, "scratch allocation failed");
    GXY_HIP(ctx, hipMemcpyAsync(dh, hsps_host, n * sizeof(gxy_hsp), hipMemcpyHostToDevice, ctx->stream));
    GXY_HIP(ctx, hipMemcpyAsync(dp, ptrs.data(), 2 * n * sizeof(void *), hipMemcpyHostToDevice, ctx->stream));
    if (scoring_is_i8(sc))
        k_choose_anchor<ScorerI8><<<grid_for(n, 128), 128, 0, ctx->stream>>>(dh, n, dp, dp + n, make_scorer_i8(sc),
                                                                             window, da);
    else
        k_choose_anchor<ScorerI32><<<grid_for(n, 128), 128, 0, ctx->stream>>>(dh, n, dp, dp + n,
                                                                              make_scorer_i32(sc), window, da);
    GXY_LAUNCH_CHECK(ctx);
    GXY_HIP(ctx, hipMemcpyAsync(out_host, da, n * sizeof(gxy_anchor), hipMemcpyDeviceToHost, ctx->stream));
    GXY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return GXY_OK;
}

int choose_anchor_dev(gxy_ctx *ctx, const gxy_dseq *t, const gxy_dseq *q, const gxy_hsp *hsps_host, uint64_t n,
                      const gxy_scoring &sc, uint32_t window, gxy_anchor *out_host) {
    std::vector<const gxy_dseq *> ts(n, t), qs(n, q);
    return choose_anchor_pairs_dev(ctx, ts.data(), qs.data(), hsps_host, n, sc, window, out_host);
}

}  // namespace gxy

namespace {

int make_set(gxy_ctx *ctx, const std::vector<gxy::GappedResult> &res, const std::vector<gxy_edit_run> &runs,
             uint64_t cells, gxy_alignment_set **out) {
    auto *s = (gxy_alignment_set *)std::calloc(1, sizeof(gxy_alignment_set));
    if (!s) return set_err(ctx, GXY_ENOMEM, "host allocation failed");
    s->n = res.size();
    s->alns = (gxy_alignment *)std::calloc(std::max<size_t>(1, res.size()), sizeof(gxy_alignment));
    s->found = (uint8_t *)std::calloc(std::max<size_t>(1, res.size()), 1);
    s->n_runs = runs.size();
    s->runs = (gxy_edit_run *)std::calloc(std::max<size_t>(1, runs.size()), sizeof(gxy_edit_run));
    s->cells = cells;
    if (!s->alns || !s->found || !s->runs) {
        gxy_alignment_set_free(s);
        return set_err(ctx, GXY_ENOMEM, "host allocation failed");
    }
    for (size_t k = 0; k < res.size(); k++) {
        s->found[k] = res[k].found ? 1 : 0;
        if (res[k].found) s->alns[k] = res[k].aln;
    }
    if (!runs.empty()) std::memcpy(s->runs, runs.data(), runs.size() * sizeof(gxy_edit_run));
    *out = s;
    return GXY_OK;
}

}  // namespace

extern "C" {

int gxy_choose_anchor_batch(gxy_ctx *ctx, const gxy_hsp *hsps, uint64_t n, const gxy_seq *target,
                            const gxy_seq *query, const gxy_scoring *scoring, uint32_t window, gxy_anchor *out) {
    return guarded(ctx, [&]() -> int {
        if (!ctx) return GXY_EINVAL;
        if (!target || !query || !scoring || (n && (!hsps || !out))) return set_err(ctx, GXY_EINVAL, "null argument");
        GXY_HIP(ctx, hipSetDevice(ctx->device));
        if (n == 0) return GXY_OK;
        gxy_dseq *dt = nullptr, *dq = nullptr;
        int rc = dseq_upload_impl(ctx, target, &dt);
        if (!rc) rc = dseq_upload_impl(ctx, query, &dq);
        if (!rc) rc = choose_anchor_dev(ctx, dt, dq, hsps, n, *scoring, window, out);
        gxy_dseq_free(dt);
        gxy_dseq_free(dq);
        return rc;
    });
}

int gxy_choose_anchor(gxy_ctx *ctx, const gxy_hsp *hsp, const gxy_seq *target, const gxy_seq *query,
                      const gxy_scoring *scoring, uint32_t window, gxy_anchor *out) {
    return guarded(ctx, [&]() -> int {
        return gxy_choose_anchor_batch(ctx, hsp, 1, target, query, scoring, window, out);
    });
}

int gxy_gapped_extend_batch(gxy_ctx *ctx, const gxy_anchor *anchors, uint64_t n, const gxy_seq *target,
                            const gxy_seq *query, const gxy_scoring *scoring, const gxy_gapped_params *params,
                            gxy_alignment_set **out) {
    return guarded(ctx, [&]() -> int {
        if (!ctx) return GXY_EINVAL;
        if (!target || !query || !scoring || !params || !out || (n && !anchors))
            return set_err(ctx, GXY_EINVAL, "null argument");
        GXY_HIP(ctx, hipSetDevice(ctx->device));
        gxy_dseq *dt = nullptr, *dq = nullptr;
        int rc = dseq_upload_impl(ctx, target, &dt);
        if (!rc) rc = dseq_upload_impl(ctx, query, &dq);
        std::vector<GappedResult> res;
        std::vector<gxy_edit_run> runs;
        uint64_t cells = 0;
        if (!rc) {
            std::vector<GappedItem> items(n);
            for (uint64_t k = 0; k < n; k++) items[k] = {dt, dq, anchors[k]};
            rc = gapped_batch_dev(ctx, items, *scoring, *params, &res, &runs, &cells);
        }
        gxy_dseq_free(dt);
        gxy_dseq_free(dq);
        if (rc) return rc;
        return make_set(ctx, res, runs, cells, out);
    });
}

int gxy_gapped_extend_slices_dev(gxy_ctx *ctx, const gxy_dseq *target, const gxy_dseq *query, const uint64_t *t_off,
                                 const uint64_t *t_len, const uint64_t *q_off, const uint64_t *q_len,
                                 const gxy_anchor *anchors, uint64_t n, const gxy_scoring *scoring,
                                 const gxy_gapped_params *params, gxy_alignment_set **out) {
    return guarded(ctx, [&]() -> int {
        if (!ctx) return GXY_EINVAL;
        if (!target || !query || !scoring || !params || !out ||
            (n && (!t_off || !t_len || !q_off || !q_len || !anchors)))
            return set_err(ctx, GXY_EINVAL, "null argument");
        GXY_HIP(ctx, hipSetDevice(ctx->device));
        std::vector<GappedItem> items(n);
        for (uint64_t k = 0; k < n; k++)
            items[k] = GappedItem{target, query, anchors[k], t_off[k], t_len[k], q_off[k], q_len[k]};
        std::vector<GappedResult> res;
        std::vector<gxy_edit_run> runs;
        uint64_t cells = 0;
        int rc = gapped_batch_dev(ctx, items, *scoring, *params, &res, &runs, &cells);
        if (rc) return rc;
        return make_set(ctx, res, runs, cells, out);
    });
}

int gxy_gapped_extend_pairs_dev(gxy_ctx *ctx, const gxy_dseq *const *targets, const gxy_dseq *const *queries,
                                const gxy_anchor *anchors, uint64_t n, const gxy_scoring *scoring,
                                const gxy_gapped_params *params, gxy_alignment_set **out) {
    return guarded(ctx, [&]() -> int {
        if (!ctx) return GXY_EINVAL;
        if (!scoring || !params || !out || (n && (!targets || !queries || !anchors)))
            return set_err(ctx, GXY_EINVAL, "null argument");
        GXY_HIP(ctx, hipSetDevice(ctx->device));
        std::vector<GappedItem> items(n);
        for (uint64_t k = 0; k < n; k++) {
            if (!targets[k] || !queries[k]) return set_err(ctx, GXY_EINVAL, "null sequence in pair %llu",
                                                           (unsigned long long)k);
            items[k] = {targets[k], queries[k], anchors[k]};
        }
        std::vector<GappedResult> res;
        std::vector<gxy_edit_run> runs;
        uint64_t cells = 0;
        int rc = gapped_batch_dev(ctx, items, *scoring, *params, &res, &runs, &cells);
        if (rc) return rc;
        return make_set(ctx, res, runs, cells, out);
    });
}

int gxy_gapped_extend(gxy_ctx *ctx, const gxy_anchor *anchor, const gxy_seq *target, const gxy_seq *query,
                      const gxy_scoring *scoring, const gxy_gapped_params *params, gxy_alignment *out,
                      gxy_edit_run *runs, uint64_t runs_cap, int *found) {
    return guarded(ctx, [&]() -> int {
        if (!ctx) return GXY_EINVAL;
        if (!anchor || !out || !found || (runs_cap && !runs)) return set_err(ctx, GXY_EINVAL, "null argument");
        gxy_alignment_set *s = nullptr;
        int rc = gxy_gapped_extend_batch(ctx, anchor, 1, target, query, scoring, params, &s);
        if (rc) return rc;
        *found = s->found[0];
        std::memset(out, 0, sizeof *out);
        if (s->found[0]) {
            *out = s->alns[0];
            out->runs_offset = 0;
            const uint64_t c = std::min<uint64_t>(runs_cap, s->alns[0].n_runs);
            if (c) std::memcpy(runs, s->runs + s->alns[0].runs_offset, c * sizeof(gxy_edit_run));
            if (s->alns[0].n_runs > runs_cap) rc = set_err(ctx, GXY_ERANGE, "edit script needs %u runs",
                                                            s->alns[0].n_runs);
        }
        gxy_alignment_set_free(s);
        return rc;
    });
}

void gxy_alignment_set_free(gxy_alignment_set *set) {
    if (!set) return;
    std::free(set->alns);
    std::free(set->found);
    std::free(set->runs);
    std::free(set);
}

}  // extern "C"
