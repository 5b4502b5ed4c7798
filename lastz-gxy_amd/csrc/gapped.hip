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
// Y-from: 1 bit), one 32-bit word per lane, plus the row window (jl, jh).  A
// second kernel walks back from the best cell, one lane per problem, and emits
// run-length edit runs tip-to-root.
#include <algorithm>

#include "internal.h"

using namespace gxy;

namespace {

constexpr int kMaxW = 1280;                 // widest row window held in LDS
constexpr int kDpThreads = 256;             // lanes per side-problem
constexpr int kDpWaves = kDpThreads / 64;
constexpr int kCMax = kMaxW / kDpThreads;   // cells per lane
constexpr int kRowBytes = kDpThreads * 4;   // traceback bytes per row: one word per lane

enum : uint32_t { kStM = 0, kStX = 1, kStY = 2 };

struct DpProb {
    const Blk *T;
    const Blk *Q;
    uint64_t t0, q0;   // position of T[0] / Q[0]
    int32_t tdir, qdir;
    uint32_t n, m;
    uint64_t tb_off;   // byte offset of row 0's traceback
    uint64_t row_off;  // index of row 0's window record
    uint32_t rows_cap; // traceback rows reserved; the sweep stops (truncated) there
    uint32_t pad;
};

struct DpRes {
    int32_t score;
    uint32_t bi, bj, bstate;
    uint32_t rows;      // rows swept
    uint32_t truncated; // 1 when the sweep hit rows_cap before ending
    uint64_t cells;
};

__device__ __forceinline__ int32_t shfl_up_i32(int32_t v, int d) { return __shfl_up(v, d); }

// One side-problem per block of kDpThreads lanes (four wavefronts); lane t
// owns the contiguous cells [jl + t*C, jl + (t+1)*C) of the row window,
// C = ceil(W / kDpThreads) <= 5.  Per row: read the previous row's values of
// the lane's cells (and the two cells before them) from LDS; compute X and M;
// recompute M of the cell just before the segment (the previous lane's last
// cell) so the Y scan needs no neighbour exchange; Y by the max-plus prefix
// scan (lane-local, wavefront shuffle scan, carry across wavefronts through
// LDS); row best and alive band reduced through LDS; the new row overwrites
// the old one in LDS.  Three barriers per row.
template <class S>
__global__ __launch_bounds__(kDpThreads) void k_dp_forward(const DpProb *__restrict__ probs, uint32_t nprob, S sc,
                                                           int32_t oe, int32_t e, int32_t ydrop,
                                                           uint8_t *__restrict__ tb, uint2 *__restrict__ rows,
                                                           DpRes *__restrict__ res) {
    __shared__ int32_t lM[kMaxW], lX[kMaxW], lY[kMaxW];  // the previous row, indexed by j - pjl
    __shared__ int32_t wtot[kDpWaves];                    // Y scan: wavefront totals
    __shared__ int32_t wmaxv[kDpWaves];                   // row best per wavefront
    __shared__ uint32_t wmaxj[kDpWaves], wmaxst[kDpWaves];
    __shared__ uint32_t wlo[kDpWaves], whi[kDpWaves];     // alive band per wavefront
    const uint32_t p = blockIdx.x;
    if (p >= nprob) return;  // uniform per block
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const DpProb pr = probs[p];
    const uint32_t n = pr.n, m = pr.m;
    DpRes out = {0, 0, 0, kStM, 0, 0, 0};
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

    int32_t best_score = 0;
    uint32_t bi = 0, bj = 0, bst = kStM;
    uint32_t lo = 0, hi = 0;
    uint32_t pjl = 0, pW = 0;  // previous row window (pW == 0: none)
    uint64_t cells = 0;
    uint32_t rows_done = 0;
    uint8_t *tbp = tb + pr.tb_off;
    uint2 *rowp = rows + pr.row_off;

    uint32_t truncated = 0;
    for (uint32_t i = 0; i <= n; i++) {
        if (i >= pr.rows_cap) {
            truncated = 1;
            break;
        }
        const uint32_t jl = lo >= br ? lo - br : 0;
        const uint32_t jh = (uint32_t)min(min((uint64_t)hi + br, (uint64_t)m), (uint64_t)jl + bw - 1);
        const uint32_t W = jh - jl + 1;
        const uint32_t C = (W + kDpThreads - 1) / kDpThreads;
        const uint32_t c0 = (uint32_t)t * C;
        const uint32_t cnt = c0 < W ? min(C, W - c0) : 0;
        const uint32_t js = jl + c0;  // first cell of this lane
        cells += W - ((i == 0 && jl == 0) ? 1 : 0);

        // target base of row i (T[i-1]); uniform
        uint32_t tc = 0, tv = 0;
        if (i >= 1) {
            const uint64_t tp = pr.t0 + (int64_t)pr.tdir * (int64_t)(i - 1);
            tc = base_code(pr.T, tp);
            tv = base_valid(pr.T, tp);
        }
        // query bases Q[j-1] for j in [js-1, js+cnt): window from bq
        const uint32_t bq = js >= 2 ? js - 2 : 0;
        Win qw = {0, 0, 0};
        if (cnt > 0 && bq < m) {
            if (pr.qdir > 0) {
                qw = load_win(pr.Q, pr.q0 + bq);
            } else {
                const uint64_t pp = pr.q0 - bq;
                if (pp >= 31) {
                    const Win x = load_win(pr.Q, pp - 31);
                    qw.codes = rev2(x.codes);
                    qw.valid = __builtin_bitreverse32(x.valid);
                } else {
                    const Win x = load_win(pr.Q, 0);
                    const uint32_t sh = 31 - (uint32_t)pp;
                    qw.codes = rev2(x.codes) >> (2 * sh);
                    qw.valid = __builtin_bitreverse32(x.valid) >> sh;
                }
            }
        }
        auto prev = [&](uint32_t j, int32_t &vm, int32_t &vx, int32_t &vy) {
            const uint32_t r = j - pjl;  // wraps when j < pjl
            if (i >= 1 && j >= pjl && r < pW) {
                vm = lM[r];
                vx = lX[r];
                vy = lY[r];
            } else {
                vm = vx = vy = kNegInf;
            }
        };
        // M of cell j from the previous row's diagonal (qm, qx, qy) = row i-1
        // at j-1 (src/gapped_extend.rs:339-363)
        auto mcell = [&](uint32_t j, int32_t qm, int32_t qx, int32_t qy, uint32_t *arg) -> int32_t {
            if (i >= 1 && j >= 1) {
                int32_t d;
                if (qm >= qx && qm >= qy) { d = qm; *arg = kStM; }
                else if (qx >= qy) { d = qx; *arg = kStX; }
                else { d = qy; *arg = kStY; }
                const uint32_t qi = j - 1 - bq;
                const uint32_t qc = (uint32_t)(qw.codes >> (2 * qi)) & 3u;
                const uint32_t qvb = (qw.valid >> qi) & 1u;
                return d + sc(tc, qc, tv & qvb);
            }
            *arg = kStM;
            return (i == 0 && j == 0) ? 0 : kNegInf;  // the seed cell, src/gapped_extend.rs:245
        };

        // ---- previous row, X and M --------------------------------------------
        int32_t Mv[kCMax], Xv[kCMax], Yv[kCMax];
        uint32_t nib[kCMax];
        // M at js-1: the previous lane's last cell (NEG before the window)
        int32_t firstPrevM = kNegInf;
        int32_t qm = kNegInf, qx = kNegInf, qy = kNegInf;  // row i-1 at j-1 for the next cell
        if (cnt > 0 && t > 0) {
            int32_t dm, dx, dy;
            prev(js - 2, dm, dx, dy);
            uint32_t arg;
            firstPrevM = mcell(js - 1, dm, dx, dy, &arg);
            prev(js - 1, qm, qx, qy);
        }
#pragma unroll
        for (int k = 0; k < kCMax; k++) {
            Mv[k] = kNegInf;
            Xv[k] = kNegInf;
            Yv[k] = kNegInf;
            nib[k] = 0;
            if ((uint32_t)k < cnt) {
                const uint32_t j = js + k;
                int32_t am, ax, ay;
                prev(j, am, ax, ay);
                uint32_t nb = 0;
                if (i >= 1) {  // X (src/gapped_extend.rs:293-312)
                    const int32_t fm = am - oe, fx = ax - e;
                    Xv[k] = fm >= fx ? fm : fx;
                    nb |= (fm >= fx ? 0u : 1u) << 2;
                }
                uint32_t arg;
                Mv[k] = mcell(j, qm, qx, qy, &arg);
                if (i >= 1 && j >= 1) nb |= arg;
                nib[k] = nb;
                qm = am;
                qx = ax;
                qy = ay;
            }
        }

        // ---- Y as a max-plus prefix scan ----------------------------------------
        const uint32_t j0s = jl >= 1 ? jl : 1;
        int32_t tot = INT32_MIN;
        {
            int32_t pm = firstPrevM;
#pragma unroll
            for (int k = 0; k < kCMax; k++) {
                if ((uint32_t)k < cnt) {
                    const uint32_t j = js + k;
                    if (j >= j0s) tot = max(tot, pm - oe + (int32_t)(j - jl) * e);
                    pm = Mv[k];
                }
            }
        }
        int32_t incl = tot;  // inclusive max-scan within the wavefront
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int32_t v = shfl_up_i32(incl, d);
            if (lane >= d) incl = max(incl, v);
        }
        if (lane == 63) wtot[wv] = incl;
        __syncthreads();  // (1) wavefront totals of the Y scan
        int32_t carry = shfl_up_i32(incl, 1);
        if (lane == 0) carry = INT32_MIN;
        for (int v = 0; v < wv; v++) carry = max(carry, wtot[v]);
        const int32_t zinit = kNegInf + ((int32_t)j0s - 1 - (int32_t)jl) * e;
        carry = max(carry, zinit);
        {
            int32_t z = carry;
            int32_t pm = firstPrevM;
            // Y at js-1 from the carry: Y[js-1] = Z[js-1] - (js-1-jl)*e
            int32_t py = kNegInf;
            if (t > 0 && js - 1 >= j0s) py = carry - ((int32_t)(js - 1 - jl)) * e;
#pragma unroll
            for (int k = 0; k < kCMax; k++) {
                if ((uint32_t)k < cnt) {
                    const uint32_t j = js + k;
                    if (j >= j0s) {
                        z = max(z, pm - oe + (int32_t)(j - jl) * e);
                        const int32_t y = z - (int32_t)(j - jl) * e;
                        Yv[k] = y;
                        // from-state (src/gapped_extend.rs:318-334)
                        nib[k] |= ((pm - oe >= py - e) ? 0u : 1u) << 3;
                        py = y;
                    } else {
                        py = kNegInf;
                    }
                    pm = Mv[k];
                }
            }
        }

        // ---- row best (first strict maximum in j order) ---------------------
        int32_t rmax = INT32_MIN;
        uint32_t rj = 0, rst = kStM;
        int32_t cb[kCMax];
#pragma unroll
        for (int k = 0; k < kCMax; k++) {
            cb[k] = max(max(Mv[k], Xv[k]), Yv[k]);
            if ((uint32_t)k < cnt && cb[k] > rmax) {
                rmax = cb[k];
                rj = js + k;
                rst = Mv[k] == cb[k] ? kStM : (Xv[k] == cb[k] ? kStX : kStY);
            }
        }
        int32_t wmax = rmax;
#pragma unroll
        for (int d = 32; d > 0; d >>= 1) wmax = max(wmax, __shfl_xor(wmax, d));
        {
            const unsigned long long hasmax = __ballot(cnt > 0 && rmax == wmax);
            const int src = hasmax ? __ffsll((long long)hasmax) - 1 : 0;
            const uint32_t wj = __shfl(rj, src), wst = __shfl(rst, src);
            if (lane == 0) {
                wmaxv[wv] = hasmax ? wmax : INT32_MIN;
                wmaxj[wv] = wj;
                wmaxst[wv] = wst;
            }
        }
        __syncthreads();  // (2) row best per wavefront
        int32_t gmax = INT32_MIN;
        uint32_t gj = 0, gst = kStM;
#pragma unroll
        for (int v = 0; v < kDpWaves; v++) {
            if (wmaxv[v] > gmax) {  // strict: the earliest wavefront holding the maximum
                gmax = wmaxv[v];
                gj = wmaxj[v];
                gst = wmaxst[v];
            }
        }
        if (gmax > best_score) {  // src/gapped_extend.rs:365-380
            best_score = gmax;
            bi = i;
            bj = gj;
            bst = gst;
        }
        const int32_t thr = sat_sub(best_score, ydrop);

        // ---- alive band, the new row into LDS, traceback ---------------------
        uint32_t alo = 0xFFFFFFFFu, ahi = 0;
#pragma unroll
        for (int k = 0; k < kCMax; k++) {
            if ((uint32_t)k < cnt) {
                const uint32_t j = js + k;
                if (cb[k] >= thr) {
                    alo = min(alo, j);
                    ahi = max(ahi, j);
                } else {
                    Mv[k] = Xv[k] = Yv[k] = kNegInf;
                }
                lM[c0 + k] = Mv[k];
                lX[c0 + k] = Xv[k];
                lY[c0 + k] = Yv[k];
            }
        }
        {
            const unsigned long long anyalive = __ballot(alo != 0xFFFFFFFFu);
            const int fl = anyalive ? __ffsll((long long)anyalive) - 1 : 0;
            const int ll = anyalive ? 63 - __clzll((long long)anyalive) : 0;
            const uint32_t wl = __shfl(alo, fl), wh = __shfl(ahi, ll);
            if (lane == 0) {
                wlo[wv] = anyalive ? wl : 0xFFFFFFFFu;
                whi[wv] = wh;
            }
        }
        // traceback nibbles: the lane's cells in one 32-bit word
        uint32_t word = 0;
#pragma unroll
        for (int k = 0; k < kCMax; k++) word |= nib[k] << (4 * k);
        *(uint32_t *)(tbp + (uint64_t)i * kRowBytes + (uint64_t)t * 4) = word;
        if (t == 0) rowp[i] = make_uint2(jl, jh);
        rows_done = i + 1;
        __syncthreads();  // (3) alive band; the new row is in LDS
        uint32_t nlo = 0xFFFFFFFFu, nhi = 0;
        bool alive = false;
#pragma unroll
        for (int v = 0; v < kDpWaves; v++) {
            if (wlo[v] != 0xFFFFFFFFu) {
                if (!alive) nlo = wlo[v];
                nhi = whi[v];
                alive = true;
            }
        }
        if (!alive) break;  // whole row dead (src/gapped_extend.rs:407-409)
        lo = nlo;
        hi = nhi;
        pjl = jl;
        pW = W;
    }
    if (t == 0) {
        out.score = best_score;
        out.bi = bi;
        out.bj = bj;
        out.bstate = bst;
        out.rows = rows_done;
        out.truncated = truncated;
        out.cells = cells;
        res[p] = out;
    }
}

__global__ void k_dp_traceback(const DpProb *__restrict__ probs, const DpRes *__restrict__ res, uint32_t nprob,
                               const uint8_t *__restrict__ tb, const uint2 *__restrict__ rows,
                               const uint64_t *__restrict__ run_off, gxy_edit_run *__restrict__ runs,
                               uint32_t *__restrict__ nruns) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nprob) return;
    const DpProb pr = probs[p];
    const DpRes r = res[p];
    uint32_t i = r.bi, j = r.bj, st = r.bstate;
    gxy_edit_run *out = runs + run_off[p];
    uint32_t nr = 0;
    uint32_t cur_op = 0xFFFFFFFFu, cur_n = 0;
    const uint8_t *tbp = tb + pr.tb_off;
    const uint2 *rowp = rows + pr.row_off;
    const uint32_t cap = r.bi + r.bj;  // runs reserved for this problem
    while (i > 0 || j > 0) {
        // A consistent traceback never leaves the swept cells; these guards
        // stop a walk rather than read outside the traceback store.
        if (i >= r.rows || (st == kStM && (i == 0 || j == 0)) || (st == kStX && i == 0) ||
            (st == kStY && j == 0) || st > kStY || nr >= cap)
            break;
        const uint2 win = rowp[i];
        if (j < win.x || j > win.y) break;
        const uint32_t W = win.y - win.x + 1, C = (W + kDpThreads - 1) / kDpThreads;
        const uint32_t idx = j - win.x, ln = idx / C, k = idx % C;
        const uint32_t word = *(const uint32_t *)(tbp + (uint64_t)i * kRowBytes + ln * 4);
        const uint32_t nb = (word >> (4 * k)) & 0xFu;
        uint32_t op, next;
        if (st == kStM) {
            op = GXY_OP_MATCH;
            next = nb & 3u;
            i--;
            j--;
        } else if (st == kStX) {
            op = GXY_OP_DELETE;
            next = ((nb >> 2) & 1u) ? kStX : kStM;
            i--;
        } else {
            op = GXY_OP_INSERT;
            next = ((nb >> 3) & 1u) ? kStY : kStM;
            j--;
        }
        if (op == cur_op) {
            cur_n++;
        } else {
            if (cur_n) {
                if (nr >= cap) break;
                out[nr++] = {cur_op, cur_n};
            }
            cur_op = op;
            cur_n = 1;
        }
        st = next;
    }
    if (cur_n && nr < cap) out[nr++] = {cur_op, cur_n};
    nruns[p] = nr;
}

// choose_anchor (src/anchor.rs:25-69), one lane per HSP.
template <class S>
__global__ void k_choose_anchor(const gxy_hsp *__restrict__ hsps, uint64_t n, const Blk *__restrict__ T,
                                const Blk *__restrict__ Q, S sc, uint32_t window, gxy_anchor *__restrict__ out) {
    const uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= n) return;
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
    const int32_t oe = sc.gap_open + sc.gap_extend, e = sc.gap_extend;
    for (const auto &it : items) {
        if (it.a.t_pos >= it.t->len || it.a.q_pos >= it.q->len)
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
        r.t0 = at + 1;
        r.q0 = aq + 1;
        r.tdir = 1;
        r.qdir = 1;
        r.n = (uint32_t)(it.t->len - at - 1);
        r.m = (uint32_t)(it.q->len - aq - 1);
        DpProb &l = probs[2 * k + 1];
        l = DpProb{};
        l.T = it.t->d;
        l.Q = it.q->d;
        l.t0 = at >= 1 ? at - 1 : 0;
        l.q0 = aq >= 1 ? aq - 1 : 0;
        l.tdir = -1;
        l.qdir = -1;
        l.n = (uint32_t)at;
        l.m = (uint32_t)aq;
        for (int s2 = 0; s2 < 2; s2++)
            if (band_width_host(oe, e, gp.y_drop, probs[2 * k + s2].m) > (uint32_t)kMaxW)
                return set_err(ctx, GXY_EUNSUP, "band width exceeds %d cells (y_drop/gap_extend too large)",
                               kMaxW);
    }
    // anchor column scores
    std::vector<const Blk *> tp(2 * n);
    std::vector<gxy_anchor> anc(n);
    for (uint64_t k = 0; k < n; k++) {
        tp[k] = items[k].t->d;
        tp[n + k] = items[k].q->d;
        anc[k] = items[k].a;
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

    // Sweep in passes.  Each pass reserves traceback rows for its problems
    // from a fixed budget; a side whose sweep outlives its reservation is
    // re-run in the next pass with a larger one (same result, recomputed).
    // Completed sides are walked back at once, before the next pass reuses
    // the traceback store.
    constexpr uint64_t kRowBudget = 8ull << 20;  // 8 Mi rows = 8 GiB of traceback
    std::vector<DpRes> res(np);
    std::vector<std::vector<gxy_edit_run>> side(np);
    std::vector<uint64_t> pending;
    for (uint64_t x = 0; x < np; x++) {
        if (probs[x].n == 0 || probs[x].m == 0) {
            res[x] = DpRes{0, 0, 0, kStM, 0, 0, 0};
        } else {
            pending.push_back(x);
        }
    }
    // longest sides first, so the long sweeps start early instead of forming
    // the tail of a pass
    std::stable_sort(pending.begin(), pending.end(),
                     [&](uint64_t a, uint64_t b) { return probs[a].n > probs[b].n; });
    uint64_t cells = 0;
    uint64_t cap = std::max<uint64_t>(1024, kRowBudget / std::max<uint64_t>(1, pending.size()));
    while (!pending.empty()) {
        // take as many pending problems as fit the budget at this cap
        std::vector<DpProb> pass;
        std::vector<uint64_t> ids;
        uint64_t rows_used = 0;
        std::vector<uint64_t> later;
        for (uint64_t x : pending) {
            const uint64_t want = std::min<uint64_t>((uint64_t)probs[x].n + 1, cap);
            if (!ids.empty() && rows_used + want > kRowBudget) {
                later.push_back(x);
                continue;
            }
            DpProb q = probs[x];
            q.rows_cap = (uint32_t)want;
            q.tb_off = rows_used * kRowBytes;
            q.row_off = rows_used;
            rows_used += want;
            pass.push_back(q);
            ids.push_back(x);
        }
        const uint64_t npp = pass.size();
        auto *dprob = (DpProb *)scratch(ctx, kSlotDpProb, npp * sizeof(DpProb));
        auto *dres = (DpRes *)scratch(ctx, kSlotDpOut, npp * sizeof(DpRes));
        auto *dtb = (uint8_t *)scratch(ctx, kSlotDpTb, rows_used * kRowBytes);
        auto *drows = (uint2 *)scratch(ctx, kSlotDpRows, rows_used * sizeof(uint2));
        if (!dprob || !dres || !dtb || !drows)
            return set_err(ctx, GXY_ENOMEM, "DP scratch allocation failed (%llu traceback rows)",
                           (unsigned long long)rows_used);
        GXY_HIP(ctx, hipMemcpyAsync(dprob, pass.data(), npp * sizeof(DpProb), hipMemcpyHostToDevice, ctx->stream));
        {
        KernelTimer kt_(ctx, "dp_forward");
        if (scoring_is_i8(sc))
            k_dp_forward<ScorerI8><<<(unsigned)npp, kDpThreads, 0, ctx->stream>>>(dprob, (uint32_t)npp, make_scorer_i8(sc),
                                                                          oe, e, gp.y_drop, dtb, drows, dres);
        else
            k_dp_forward<ScorerI32><<<(unsigned)npp, kDpThreads, 0, ctx->stream>>>(dprob, (uint32_t)npp,
                                                                           make_scorer_i32(sc), oe, e, gp.y_drop,
                                                                           dtb, drows, dres);
        }
        GXY_LAUNCH_CHECK(ctx);
        std::vector<DpRes> pr(npp);
        GXY_HIP(ctx, hipMemcpyAsync(pr.data(), dres, npp * sizeof(DpRes), hipMemcpyDeviceToHost, ctx->stream));
        GXY_HIP(ctx, hipStreamSynchronize(ctx->stream));
        // walk back every completed side
        std::vector<uint64_t> roff(npp, 0);
        uint64_t total_runs = 0;
        for (uint64_t y = 0; y < npp; y++) {
            roff[y] = total_runs;
            if (!pr[y].truncated) total_runs += (uint64_t)pr[y].bi + pr[y].bj;
            else pr[y].bi = pr[y].bj = 0;  // no walk for truncated sides
        }
        std::vector<uint32_t> nr(npp, 0);
        std::vector<gxy_edit_run> sr(std::max<uint64_t>(1, total_runs));
        if (total_runs > 0) {
            auto *droff = (uint64_t *)scratch(ctx, kSlotMisc3, npp * 8);
            auto *druns = (gxy_edit_run *)scratch(ctx, kSlotDpRuns, total_runs * sizeof(gxy_edit_run));
            auto *dnr = (uint32_t *)scratch(ctx, kSlotMisc2, npp * 4);
            if (!droff || !druns || !dnr) return set_err(ctx, GXY_ENOMEM, "scratch allocation failed");
            GXY_HIP(ctx, hipMemcpyAsync(dres, pr.data(), npp * sizeof(DpRes), hipMemcpyHostToDevice, ctx->stream));
            GXY_HIP(ctx, hipMemcpyAsync(droff, roff.data(), npp * 8, hipMemcpyHostToDevice, ctx->stream));
            {
                KernelTimer kt_(ctx, "dp_traceback");
                k_dp_traceback<<<grid_for(npp, 64), 64, 0, ctx->stream>>>(dprob, dres, (uint32_t)npp, dtb, drows, droff,
                                                                          druns, dnr);
            }
            GXY_LAUNCH_CHECK(ctx);
            GXY_HIP(ctx, hipMemcpyAsync(sr.data(), druns, total_runs * sizeof(gxy_edit_run), hipMemcpyDeviceToHost,
                                        ctx->stream));
            GXY_HIP(ctx, hipMemcpyAsync(nr.data(), dnr, npp * 4, hipMemcpyDeviceToHost, ctx->stream));
            GXY_HIP(ctx, hipStreamSynchronize(ctx->stream));
        }
        std::vector<uint64_t> next;
        GXY_HIP(ctx, hipMemcpyAsync(pr.data(), dres, npp * sizeof(DpRes), hipMemcpyDeviceToHost, ctx->stream));
        GXY_HIP(ctx, hipStreamSynchronize(ctx->stream));
        for (uint64_t y = 0; y < npp; y++) {
            const uint64_t x = ids[y];
            if (pr[y].truncated) {
                next.push_back(x);
                continue;
            }
            cells += pr[y].cells;
            res[x] = pr[y];
            side[x].assign(sr.begin() + roff[y], sr.begin() + roff[y] + nr[y]);
        }
        // truncated sides are retried with four times the reservation
        if (!next.empty()) cap = std::min<uint64_t>(cap * 4, kRowBudget);
        for (uint64_t x : later) next.push_back(x);
        pending.swap(next);
    }
    if (cells_out) *cells_out += cells;

    // Assemble (src/gapped_extend.rs:107-133): score = col + L + R (saturating);
    // script = rev(L.script) + M1 + R.script.  The walk emits tip-to-root runs,
    // so the left side contributes its raw order and the right side reversed.
    for (uint64_t k = 0; k < n; k++) {
        const DpRes &R = res[2 * k], &L = res[2 * k + 1];
        int64_t t64 = (int64_t)col[k] + L.score;
        t64 = std::max<int64_t>(INT32_MIN, std::min<int64_t>(INT32_MAX, t64));
        t64 = std::max<int64_t>(INT32_MIN, std::min<int64_t>(INT32_MAX, t64 + R.score));
        const int32_t tot = (int32_t)t64;
        if (tot < gp.gapped_threshold) continue;
        GappedResult &g = (*results)[k];
        g.found = true;
        g.aln.score = tot;
        g.aln.t_start = items[k].a.t_pos - L.bi;
        g.aln.q_start = items[k].a.q_pos - L.bj;
        g.aln.t_len = L.bi + 1 + R.bi;
        g.aln.q_len = L.bj + 1 + R.bj;
        g.aln.source = items[k].a;
        g.aln.runs_offset = runs->size();
        auto push = [&](uint32_t op, uint32_t cnt) {
            if (cnt == 0) return;
            if (runs->size() > g.aln.runs_offset && runs->back().op == op)
                runs->back().n += cnt;
            else
                runs->push_back({op, cnt});
        };
        for (const auto &r : side[2 * k + 1]) push(r.op, r.n);
        push(GXY_OP_MATCH, 1);
        const auto &rr = side[2 * k];
        for (size_t x = rr.size(); x-- > 0;) push(rr[x].op, rr[x].n);
        g.aln.n_runs = (uint32_t)(runs->size() - g.aln.runs_offset);
    }
    return GXY_OK;
}

int choose_anchor_dev(gxy_ctx *ctx, const gxy_dseq *t, const gxy_dseq *q, const gxy_hsp *hsps_host, uint64_t n,
                      const gxy_scoring &sc, uint32_t window, gxy_anchor *out_host) {
    if (n == 0) return GXY_OK;
    for (uint64_t k = 0; k < n; k++) {
        const gxy_hsp &h = hsps_host[k];
        if (h.length < 1) return set_err(ctx, GXY_EINVAL, "empty HSP cannot be anchored");  // src/anchor.rs:33
        if ((uint64_t)h.t_start + h.length > t->len || (uint64_t)h.q_start + h.length > q->len)
            return set_err(ctx, GXY_EINVAL, "HSP outside its sequences");
    }
    auto *dh = (gxy_hsp *)scratch(ctx, kSlotMisc2, n * sizeof(gxy_hsp));
    auto *da = (gxy_anchor *)scratch(ctx, kSlotMisc3, n * sizeof(gxy_anchor));
    if (!dh || !da) return set_err(ctx, GXY_ENOMEM, "scratch allocation failed");
    GXY_HIP(ctx, hipMemcpyAsync(dh, hsps_host, n * sizeof(gxy_hsp), hipMemcpyHostToDevice, ctx->stream));
    if (scoring_is_i8(sc))
        k_choose_anchor<ScorerI8><<<grid_for(n, 128), 128, 0, ctx->stream>>>(dh, n, t->d, q->d, make_scorer_i8(sc),
                                                                             window, da);
    else
        k_choose_anchor<ScorerI32><<<grid_for(n, 128), 128, 0, ctx->stream>>>(dh, n, t->d, q->d,
                                                                              make_scorer_i32(sc), window, da);
    GXY_LAUNCH_CHECK(ctx);
    GXY_HIP(ctx, hipMemcpyAsync(out_host, da, n * sizeof(gxy_anchor), hipMemcpyDeviceToHost, ctx->stream));
    GXY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return GXY_OK;
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
}

int gxy_choose_anchor(gxy_ctx *ctx, const gxy_hsp *hsp, const gxy_seq *target, const gxy_seq *query,
                      const gxy_scoring *scoring, uint32_t window, gxy_anchor *out) {
    return gxy_choose_anchor_batch(ctx, hsp, 1, target, query, scoring, window, out);
}

int gxy_gapped_extend_batch(gxy_ctx *ctx, const gxy_anchor *anchors, uint64_t n, const gxy_seq *target,
                            const gxy_seq *query, const gxy_scoring *scoring, const gxy_gapped_params *params,
                            gxy_alignment_set **out) {
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
}

int gxy_gapped_extend_pairs_dev(gxy_ctx *ctx, const gxy_dseq *const *targets, const gxy_dseq *const *queries,
                                const gxy_anchor *anchors, uint64_t n, const gxy_scoring *scoring,
                                const gxy_gapped_params *params, gxy_alignment_set **out) {
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
}

int gxy_gapped_extend(gxy_ctx *ctx, const gxy_anchor *anchor, const gxy_seq *target, const gxy_seq *query,
                      const gxy_scoring *scoring, const gxy_gapped_params *params, gxy_alignment *out,
                      gxy_edit_run *runs, uint64_t runs_cap, int *found) {
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
}

void gxy_alignment_set_free(gxy_alignment_set *set) {
    if (!set) return;
    std::free(set->alns);
    std::free(set->found);
    std::free(set->runs);
    std::free(set);
}

}  // extern "C"
