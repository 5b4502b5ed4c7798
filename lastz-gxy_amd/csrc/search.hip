// search.hip — seed search and ungapped x-drop (src/seed_search.rs:45-125,
// src/diag_hash.rs, src/hsp.rs:66-211).
//
// Exact parallel form of the reference's sequential scan.  DiagHash state
// changes only when an extension passes the threshold (and the entropy gate),
// it is keyed by diagonal, and hits on one diagonal reach it in ascending
// query position.  So:
//   1. every seed hit is extended speculatively (its result does not depend on
//      any state);
//   2. the passing ones are sorted by (diagonal, q) and each diagonal is
//      scanned in order: emit iff no earlier emitted HSP on that diagonal ends
//      beyond the seed's t, then move the cover to max(cover, t_end);
//   3. the emitted HSPs are stably sorted by (diagonal, t_start), which keeps
//      seed-scan order among equal keys.
// Pipeline per call: k_qcount (word + lookup counts per query window) ->
// exclusive scan -> k_qexpand (materialise (q,t) hits) -> k_extend (one lane
// per hit; passing hits appended with one atomic per wavefront) -> radix sort
// by (diag,q) -> k_admit (one lane per diagonal run) -> compaction -> radix
// sort by (diag,t_start) -> download.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>

#include "internal.h"

using namespace gxy;

namespace {

inline unsigned grid_for(uint64_t n, unsigned th) { return (unsigned)((n + th - 1) / th); }

// Persistent x-drop grid: enough 256-lane blocks to fill every CU several
// times over; lanes stride over the hits.
constexpr unsigned kXdropBlocks = 2048;

// Variants of a seed word (src/seed_search.rs:63-90) in table-index form: the
// word itself, then each single transition — a code's high bit, i.e. index bit
// W-1-i for care position i (leftmost first) — then each pair i<j when
// transitions >= 2.
struct Variants {
    uint32_t weight, transitions;
    template <class F>
    __device__ __forceinline__ void each(uint64_t idx, F &&f) const {
        f(idx);
        if (transitions >= 1)
            for (uint32_t i = 0; i < weight; i++) f(idx ^ (1ull << (weight - 1 - i)));
        if (transitions >= 2)
            for (uint32_t i = 0; i < weight; i++)
                for (uint32_t j = i + 1; j < weight; j++)
                    f(idx ^ (1ull << (weight - 1 - i)) ^ (1ull << (weight - 1 - j)));
    }
};

__device__ __forceinline__ bool seed_word(const Blk *__restrict__ q, uint64_t qlen, const DevPattern &pat,
                                          uint64_t pos, int respect_mask, uint64_t *word) {
    if (pos + pat.span > qlen) return false;
    const Win x = load_win(q, pos);
    const uint32_t care = pat.care_mask;
    const uint32_t msk = respect_mask ? x.mask : 0u;
    if ((x.valid & care) != care || (msk & care) != 0) return false;
    *word = gather_index(x.codes, pat);
    return true;
}

// The bitmap of non-empty words (4^w bits: 2 MiB for 12of19) stays in L2, so
// the ~3/4 of variant lookups that find nothing never touch `starts`.
__device__ __forceinline__ bool nonempty(const uint32_t *__restrict__ bits, uint64_t v) {
    return (bits[v >> 5] >> (v & 31)) & 1u;
}

__global__ void k_qcount(const Blk *__restrict__ q, uint64_t qlen, DevPattern pat, uint64_t step,
                         uint64_t nwin, int respect_mask, Variants var, const uint32_t *__restrict__ starts,
                         const uint32_t *__restrict__ bits, uint32_t *__restrict__ words,
                         uint32_t *__restrict__ cnt, unsigned long long *__restrict__ nvalid) {
    const uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool ok = false;
    if (w < nwin) {
        uint64_t word;
        uint32_t c = 0;
        if (seed_word(q, qlen, pat, w * step, respect_mask, &word)) {
            ok = true;
            var.each(word, [&](uint64_t v) {
                if (nonempty(bits, v)) c += starts[v + 1] - starts[v];
            });
            words[w] = (uint32_t)word;
        } else {
            words[w] = kInvalidWord;
        }
        cnt[w] = c;
    }
    const unsigned long long b = __ballot(ok);
    if ((threadIdx.x & 63) == 0 && b) spread_add(nvalid, (unsigned long long)__popcll(b));
}

// transitions <= 1 and weight <= 15: at most 16 variant words per window.
constexpr int kMaxV1 = 16;

__device__ __forceinline__ uint64_t variant1(uint64_t idx, uint32_t weight, int i) {
    // i = 0: the word; 1 <= i <= weight: transition at care position i-1
    // (src/seed_search.rs:79-83), index bit weight-i
    const uint32_t sh = (uint32_t)i <= weight ? weight - (uint32_t)i : 0u;
    return i == 0 ? idx : idx ^ (1ull << sh);
}

__global__ void k_qcount_v(const Blk *__restrict__ q, uint64_t qlen, DevPattern pat, uint64_t step,
                           uint64_t nwin, int respect_mask, uint32_t nvar, const uint32_t *__restrict__ starts,
                           const uint32_t *__restrict__ cmap, const uint32_t *__restrict__ ncnt,
                           uint32_t *__restrict__ words, uint32_t *__restrict__ cnt, uint32_t *__restrict__ vcnt,
                           uint64_t *__restrict__ bsum, unsigned long long *__restrict__ nvalid) {
    const uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool ok = false;
    uint32_t c = 0;
    if (w < nwin) {
        uint64_t word;
        uint32_t m = 0;
        if (seed_word(q, qlen, pat, w * step, respect_mask, &word)) {
            ok = true;
            // m: min(bucket size, 3) of every variant, 2 bits each (variant i
            // in bits 2i..2i+1), kept for k_qexpand_w.
            if (ncnt) {
                // one load of the word's neighbourhood counts
                m = ncnt[word] & (nvar >= 16 ? 0xFFFFFFFFu : ((1u << (2 * nvar)) - 1u));
            } else {
                // Count-map probes.  Variant i >= 1 flips index bit W-i: flips
                // of bits 0-3 stay in the word's own map word g, a flip of bit
                // 4+k reads map word g ^ (1 << k) — at most 1 + (W-4) loads,
                // issued back to back.
                const uint32_t W = pat.weight;
                const uint32_t g = (uint32_t)(word >> 4), sh = 2u * (uint32_t)(word & 15);
                const uint32_t c0w = cmap[g];
                uint32_t cf[kMaxWeight - 4];
#pragma unroll
                for (uint32_t k = 0; k < kMaxWeight - 4; k++) cf[k] = 4 + k < W ? cmap[g ^ (1u << k)] : 0u;
#pragma unroll
                for (uint32_t i = 0; i < (uint32_t)kMaxV1; i++) {
                    if (i < nvar) {
                        uint32_t cw;
                        if (i == 0) {
                            cw = (c0w >> sh) & 3u;
                        } else {
                            const uint32_t b = W - i;
                            cw = 0;
#pragma unroll
                            for (uint32_t k = 0; k < kMaxWeight - 4; k++)
                                if (b == 4 + k) cw = (cf[k] >> sh) & 3u;
                            if (b < 4) cw = (c0w >> (sh ^ (2u << b))) & 3u;
                        }
                        m |= cw << (2 * i);
                    }
                }
            }
            // sizes 1 and 2 are exact; a 3 means "3 or more": read `starts`
            const uint32_t lo = m & 0x55555555u, hi = (m >> 1) & 0x55555555u, big = lo & hi;
            c = (uint32_t)__popc(lo & ~big) + 2u * (uint32_t)__popc(hi & ~big);
            uint32_t bb = big;
            while (bb) {
                const int i = (__ffs(bb) - 1) >> 1;
                bb &= bb - 1;
                const uint64_t v = variant1(word, pat.weight, i);
                c += starts[v + 1] - starts[v];
            }
            words[w] = (uint32_t)word;
        } else {
            words[w] = kInvalidWord;
        }
        cnt[w] = c;
        vcnt[w] = m;
    }
    const unsigned long long b = __ballot(ok);
    if ((threadIdx.x & 63) == 0 && b) spread_add(nvalid, (unsigned long long)__popcll(b));
    // hits of this block of 256 windows (k_qexpand_w rebuilds the per-window
    // offsets inside the block)
    uint64_t x = c;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    __shared__ uint64_t ws[4];
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = x;
    __syncthreads();
    if (threadIdx.x == 0) bsum[blockIdx.x] = ws[0] + ws[1] + ws[2] + ws[3];
}

// Load-balanced expansion (transitions <= 1): one wavefront per 64
// consecutive windows.  Each lane lists its window's non-empty variant
// buckets in LDS; the wavefront's output range [off[w0], off[w0+64]) is then
// split evenly over the lanes, each output slot finding its window by binary
// search over the 64 window offsets and its bucket by a short scan, so hit
// stores are coalesced and a window with many hits is shared by all lanes.
constexpr int kExpWaves = 4;

// Hits are binned by target position into one bin per XCD (workgroup i runs
// on XCD i % 8), so the x-drop lanes of one XCD read one eighth of the target
// bytes — a slice its L2 holds.  bin = t * mbin >> 32, mbin = 2^35 / tlen.
constexpr uint32_t kBins = 8;

inline uint32_t bin_multiplier(uint64_t tlen) {
    return (uint32_t)std::min<uint64_t>(0xFFFFFFFFull, (8ull << 32) / std::max<uint64_t>(tlen, 1));
}
__device__ __forceinline__ uint32_t target_bin(uint32_t t, uint32_t mbin) { return __umulhi(t, mbin); }

// Load-balanced expansion (transitions <= 1): one wavefront per 64
// consecutive windows, four per block of 256 windows — the blocks of
// k_qcount_v, whose scanned hit totals `boff` give each block its first
// output slot; inside the block the window offsets are prefix sums of the
// windows' counts.  The wavefront lists the non-empty buckets of its
// windows in LDS (a wave prefix sum over the per-window bucket counts gives
// each lane its slots), then splits its output range evenly over its lanes:
// output slot j finds its bucket by binary search over the buckets' first
// slots, so hit stores are coalesced and a window with many hits is shared by
// all lanes.  One-position buckets carry the position itself.  A wavefront
// whose buckets overflow the LDS list expands lane by lane instead.
constexpr int kWaveBk = 512;

__global__ __launch_bounds__(256) void k_qexpand_w(uint64_t w0, uint64_t w1, uint64_t step,
                                                   const uint32_t *__restrict__ words,
                                                   const uint32_t *__restrict__ vcnt,
                                                   const uint32_t *__restrict__ cnt,
                                                   const uint64_t *__restrict__ boff, uint64_t base, uint32_t weight,
                                                   const uint32_t *__restrict__ starts,
                                                   const uint32_t *__restrict__ positions,
                                                   const uint32_t *__restrict__ spos,
                                                   uint32_t mbin, uint64_t *__restrict__ hits,
                                                   uint32_t *__restrict__ bincnt) {
    __shared__ uint32_t lbin[kBins];
    __shared__ uint32_t kb[kExpWaves][kWaveBk];  // bucket start in `positions`, or its one position
    __shared__ uint32_t kc[kExpWaves][kWaveBk];  // bucket's first output slot, relative to the wavefront
    __shared__ uint8_t kl[kExpWaves][kWaveBk];   // owning lane, | 0x80 for a one-position bucket
    __shared__ uint64_t wsum[kExpWaves];         // hits per wavefront
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const uint64_t wb = w0 + ((uint64_t)blockIdx.x * kExpWaves + wv) * 64;
    const uint64_t w = wb + lane;
    const bool active = wb < w1;  // uniform per wavefront
    if (threadIdx.x < kBins) lbin[threadIdx.x] = 0;
    uint32_t m = 0, nb = 0, c = 0;
    if (active && w < w1) {
        c = cnt[w];
        m = vcnt[w];  // 2 bits per variant: min(bucket size, 3)
        nb = (uint32_t)__popc((m | (m >> 1)) & 0x55555555u);
    }
    // wavefront prefix sums of the hit and bucket counts
    uint64_t xc = c;
    uint32_t x = nb;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint64_t yc = __shfl_up(xc, d);
        const uint32_t y = __shfl_up(x, d);
        if (lane >= d) {
            xc += yc;
            x += y;
        }
    }
    const uint64_t wtot = __shfl(xc, 63);
    const uint32_t kfirst = x - nb, total = __shfl(x, 63);
    if (lane == 0) wsum[wv] = wtot;
    const bool fits = total <= (uint32_t)kWaveBk;
    const uint64_t orel = xc - c;  // this window's first slot, relative to the wavefront
    if (active && fits && m) {
        const uint32_t word = words[w];
        // every non-empty variant's bucket: all loads issued before the first
        // use, so a window pays one memory round trip, not one per bucket
        // (the variant indices first, outside the branches: the word's load is
        // then waited for once, not in every branch)
        uint32_t vix[kMaxV1];
#pragma unroll
        for (int i = 0; i < kMaxV1; i++) {
            vix[i] = (uint32_t)variant1(word, weight, i);
            asm volatile("" : "+v"(vix[i]));  // kept here, not sunk into the branches
        }
        uint32_t r0[kMaxV1] = {}, r1[kMaxV1] = {};
#pragma unroll
        for (int i = 0; i < kMaxV1; i++) {
            const uint32_t cw = (m >> (2 * i)) & 3u;
            if (cw) {
                // one position: take it directly; else the bucket's bounds
                r0[i] = *(cw == 1 ? spos + vix[i] : starts + vix[i]);
                if (cw > 1) r1[i] = starts[vix[i] + 1];
            }
        }
        uint32_t cum = (uint32_t)orel, k = kfirst;
#pragma unroll
        for (int i = 0; i < kMaxV1; i++) {
            const uint32_t cw = (m >> (2 * i)) & 3u;
            if (cw) {
                kb[wv][k] = r0[i];
                kc[wv][k] = cum;
                kl[wv][k] = (uint8_t)(lane | (cw == 1 ? 0x80 : 0));
                cum += cw == 1 ? 1u : r1[i] - r0[i];
                k++;
            }
        }
        m = 0;
    }
    __syncthreads();
    if (active) {
        uint64_t o0 = boff[w0 / (64 * kExpWaves) + blockIdx.x] - base;
        for (int v = 0; v < wv; v++) o0 += wsum[v];
        const uint64_t oend = o0 + wtot, o = o0 + orel;
        uint32_t nbin[kBins] = {};  // this wavefront's hits per bin (uniform)
        if (fits) {
            for (uint64_t j = o0 + lane; j < oend; j += 64) {
                const uint32_t rel = (uint32_t)(j - o0);
                // the last bucket whose first slot is <= rel
                uint32_t k = 0;
#pragma unroll
                for (uint32_t s2 = kWaveBk / 2; s2 > 0; s2 >>= 1)
                    if (k + s2 < total && kc[wv][k + s2] <= rel) k += s2;
                const uint32_t bk = kb[wv][k], L = kl[wv][k];
                uint32_t t;
                if (L & 0x80u)
                    t = bk;
                else
                    t = positions[bk + (rel - kc[wv][k])];
                hits[j] = (((wb + (uint64_t)(L & 63u)) * step) << 32) | t;
                const uint32_t bin = target_bin(t, mbin);
#pragma unroll
                for (uint32_t b = 0; b < kBins; b++) nbin[b] += (uint32_t)__popcll(__ballot(bin == b));
            }
        } else if (m) {  // overflow: this lane writes its own window's hits
            const uint32_t word = words[w];
            uint64_t j = o;
            while (m) {
                const int i = (__ffs(m) - 1) >> 1;
                m &= ~(3u << (2 * i));
                const uint64_t v = variant1(word, weight, i);
                const uint32_t b0 = starts[v], e0 = starts[v + 1];
                for (uint32_t p = b0; p < e0; p++) {
                    const uint32_t t = positions[p];
                    hits[j++] = ((w * step) << 32) | t;
                    atomicAdd(&lbin[target_bin(t, mbin)], 1u);
                }
            }
        }
        // Lane 0 took part in every iteration of the balanced loop (it owns
        // the smallest j); lanes that left early hold partial counts.
        uint32_t c = 0;
#pragma unroll
        for (uint32_t b = 0; b < kBins; b++) {
            const uint32_t xb = __builtin_amdgcn_readfirstlane(nbin[b]);
            c = (uint32_t)lane == b ? xb : c;
        }
        if (lane < (int)kBins && c) atomicAdd(&lbin[lane], c);
    }
    __syncthreads();
    if (threadIdx.x < kBins) bincnt[threadIdx.x * gridDim.x + blockIdx.x] = lbin[threadIdx.x];
}

// Moves a batch's hits into bin order: block b re-reads the hits of its 256
// windows (range from the block offsets) and writes each to its bin's range
// for this block (binoff: the
// exclusive scan of bincnt, bin-major).  Order within a (block, bin) range is
// arbitrary; passing hits are sorted by a unique key afterwards.
__global__ __launch_bounds__(256) void k_bin_hits(uint64_t blk0, const uint64_t *__restrict__ boff,
                                                  uint64_t base, const uint64_t *__restrict__ hits_in,
                                                  const uint32_t *__restrict__ binoff, uint32_t mbin,
                                                  uint64_t *__restrict__ hits_out) {
    __shared__ uint32_t lc[kBins];
    if (threadIdx.x < kBins) lc[threadIdx.x] = binoff[threadIdx.x * gridDim.x + blockIdx.x];
    __syncthreads();
    const uint64_t ja = boff[blk0 + blockIdx.x] - base, jb = boff[blk0 + blockIdx.x + 1] - base;
    for (uint64_t j = ja + threadIdx.x; j < jb; j += blockDim.x) {
        const uint64_t h = hits_in[j];
        const uint32_t slot = atomicAdd(&lc[target_bin((uint32_t)h, mbin)], 1u);
        hits_out[slot] = h;
    }
}

struct ToU64 {
    __host__ __device__ __forceinline__ uint64_t operator()(const uint32_t &x) const { return x; }
};

__global__ void k_qexpand(uint64_t w0, uint64_t w1, uint64_t step, const uint32_t *__restrict__ words,
                          const uint64_t *__restrict__ off, uint64_t base, Variants var,
                          const uint32_t *__restrict__ starts, const uint32_t *__restrict__ bits,
                          const uint32_t *__restrict__ positions, uint64_t *__restrict__ hits) {
    const uint64_t w = w0 + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= w1) return;
    const uint32_t word = words[w];
    if (word == kInvalidWord) return;
    uint64_t o = off[w] - base;
    const uint64_t qhi = (w * step) << 32;
    var.each(word, [&](uint64_t v) {
        if (!nonempty(bits, v)) return;
        const uint32_t b = starts[v], e = starts[v + 1];
        for (uint32_t k = b; k < e; k++) hits[o++] = qhi | positions[k];
    });
}

// ---- x-drop extension (src/hsp.rs:98-195) -------------------------------------

// Window of 32 bases ending at p, reversed: index k <-> position p-k.
__device__ __forceinline__ Win load_win_rev(const Blk *__restrict__ b, uint64_t p) {
    Win w;
    if (p >= 31) {
        const Win x = load_win(b, p - 31);
        w.codes = rev2(x.codes);
        w.valid = __builtin_bitreverse32(x.valid);
    } else {
        const Win x = load_win(b, 0);
        const uint32_t sh = 31 - (uint32_t)p;
        w.codes = rev2(x.codes) >> (2 * sh);
        w.valid = __builtin_bitreverse32(x.valid) >> sh;
    }
    w.mask = 0;
    return w;
}

struct SideRes {
    int32_t best;
    uint32_t ext;
};

// One side of the x-drop walk: `maxn` columns available starting at (t0,q0)
// going forward (dir>0) or backward.  Returns best score and its extent.
template <class S, bool kFwd>
__device__ __forceinline__ SideRes xdrop_side(const Blk *__restrict__ T, const Blk *__restrict__ Q,
                                              const S &sc, int32_t xdrop, uint64_t t0, uint64_t q0,
                                              uint64_t maxn, uint32_t *cols) {
    int32_t running = 0, best = 0;
    uint32_t best_ext = 0, ext = 0;
    bool stop = false;
    for (uint64_t off = 0; off < maxn && !stop; off += 32) {
        const uint32_t n = (uint32_t)min<uint64_t>(32, maxn - off);
        Win tw, qw;
        if (kFwd) {
            tw = load_win(T, t0 + off);
            qw = load_win(Q, q0 + off);
        } else {
            tw = load_win_rev(T, t0 - off);
            qw = load_win_rev(Q, q0 - off);
        }
        const uint32_t vv = tw.valid & qw.valid;
        for (uint32_t k = 0; k < n; k++) {
            const uint32_t tc = (uint32_t)(tw.codes >> (2 * k)) & 3u;
            const uint32_t qc = (uint32_t)(qw.codes >> (2 * k)) & 3u;
            const int32_t s = sc(tc, qc, (vv >> k) & 1u);
            ext++;
            running = sat_add(running, s);
            if (running > best) {
                best = running;
                best_ext = ext;
            } else if (sat_sub(best, running) > xdrop) {
                stop = true;
                break;
            }
        }
    }
    *cols += ext;
    return {best, best_ext};
}

template <class S>
__device__ __forceinline__ bool xdrop_hit(const Blk *__restrict__ T, uint64_t tlen, const Blk *__restrict__ Q,
                                          uint64_t qlen, const S &sc, int32_t xdrop, int32_t thresh,
                                          uint32_t t_pos, uint32_t q_pos, uint32_t span, gxy_hsp *out,
                                          uint32_t *cols) {
    if ((uint64_t)t_pos + span > tlen || (uint64_t)q_pos + span > qlen) return false;  // src/hsp.rs:113-115
    int32_t seed = 0;
    for (uint32_t k0 = 0; k0 < span; k0 += 32) {
        const Win tw = load_win(T, (uint64_t)t_pos + k0);
        const Win qw = load_win(Q, (uint64_t)q_pos + k0);
        const uint32_t n = min(32u, span - k0), vv = tw.valid & qw.valid;
        for (uint32_t k = 0; k < n; k++)
            seed = sat_add(seed, sc((uint32_t)(tw.codes >> (2 * k)) & 3u, (uint32_t)(qw.codes >> (2 * k)) & 3u,
                                    (vv >> k) & 1u));
    }
    *cols += span;
    const uint64_t rt = (uint64_t)t_pos + span, rq = (uint64_t)q_pos + span;
    const uint64_t rmax = (rt < tlen && rq < qlen) ? min(tlen - rt, qlen - rq) : 0;
    const SideRes r = xdrop_side<S, true>(T, Q, sc, xdrop, rt, rq, rmax, cols);
    const uint64_t lmax = min((uint64_t)t_pos, (uint64_t)q_pos);
    const SideRes l = lmax ? xdrop_side<S, false>(T, Q, sc, xdrop, (uint64_t)t_pos - 1, (uint64_t)q_pos - 1,
                                                  lmax, cols)
                           : SideRes{0, 0};
    const int32_t score = sat_add(sat_add(seed, l.best), r.best);
    if (score < thresh) return false;
    out->t_start = t_pos - l.ext;
    out->q_start = q_pos - l.ext;
    out->length = span + l.ext + r.ext;
    out->score = score;
    return true;
}

// shannon_entropy / passes_entropy_gate (src/entropy.rs:27-66) on the device.
__device__ bool entropy_pass(const Blk *__restrict__ T, uint64_t tlen, uint32_t t_start, uint32_t t_end,
                             int32_t score, int32_t threshold) {
    uint32_t c[4] = {0, 0, 0, 0};
    uint32_t total = 0;
    const uint64_t e = min((uint64_t)t_end, tlen);
    for (uint64_t p = t_start; p < e; p += 32) {
        const Win w = load_win(T, p);
        const uint32_t n = (uint32_t)min<uint64_t>(32, e - p);
        const uint32_t vm = n == 32 ? w.valid : (w.valid & ((1u << n) - 1u));
        // count codes among valid bases
        const uint64_t sp = spread2(vm);
        const uint64_t cc = w.codes & sp;
        const uint64_t lo = cc & 0x5555555555555555ull, hi = (cc >> 1) & 0x5555555555555555ull;
        const uint64_t vlo = sp & 0x5555555555555555ull;
        const uint32_t n3 = __popcll(lo & hi), n1 = __popcll(lo & ~hi), n2 = __popcll(~lo & hi & vlo);
        const uint32_t nv = __popc(vm);
        c[1] += n1;
        c[2] += n2;
        c[3] += n3;
        c[0] += nv - n1 - n2 - n3;
        total += nv;
    }
    if (total == 0) return 0.0 * score >= (double)threshold;
    const double tt = (double)total;
    double h = 0.0;
    for (int k = 0; k < 4; k++) {
        if (c[k] == 0) continue;
        const double p = (double)c[k] / tt;
        h -= p * log2(p);
    }
    return (double)score * (h / 2.0) >= (double)threshold;
}

struct PassOut {
    uint64_t *keys;
    gxy_hsp *hsp;
    unsigned long long *count;
};

template <class S>
__global__ __launch_bounds__(256) void k_extend(const uint64_t *__restrict__ hits, uint64_t nh,
                                                const Blk *__restrict__ T, uint64_t tlen,
                                                const Blk *__restrict__ Q, uint64_t qlen, S sc,
                                                int32_t xdrop, int32_t thresh, uint32_t span, int entropy_on,
                                                int32_t entropy_thresh, PassOut po,
                                                unsigned long long *__restrict__ colsum) {
    const uint64_t h = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t cols = 0;
    bool pass = false;
    gxy_hsp r;
    uint32_t qp = 0, tp = 0;
    if (h < nh) {
        const uint64_t hit = hits[h];
        qp = (uint32_t)(hit >> 32);
        tp = (uint32_t)hit;
        pass = xdrop_hit(T, tlen, Q, qlen, sc, xdrop, thresh, tp, qp, span, &r, &cols);
        if (pass && entropy_on) pass = entropy_pass(T, tlen, r.t_start, r.t_start + r.length, r.score, entropy_thresh);
    }
    // wavefront-aggregated append of passing hits
    const unsigned long long b = __ballot(pass);
    unsigned long long base = 0;
    const int lane = threadIdx.x & 63;
    if (b) {
        if (lane == __ffsll((long long)b) - 1) base = atomicAdd(po.count, (unsigned long long)__popcll(b));
        base = __shfl(base, __ffsll((long long)b) - 1);
        if (pass) {
            const uint64_t slot = base + __popcll(b & ((1ull << lane) - 1ull));
            const uint64_t d = (uint64_t)((int64_t)tp - (int64_t)qp + (int64_t)qlen);
            po.keys[slot] = (d << 32) | qp;
            po.hsp[slot] = r;
        }
    }
    // column counter: wave reduction then one atomic
    unsigned long long c = cols;
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
    if (lane == 0 && c) atomicAdd(colsum, c);
}

// ---- persistent-lane x-drop over the byte layout ------------------------------
//
// Each lane owns hits h = lane, lane + stride, ... and steps through them as a
// small state machine — right walk, left walk — one 32-column chunk per loop
// iteration, so a lane whose walk ends early starts its next hit instead of
// idling while the rest of its wavefront finishes.  Lanes own ~stride-apart
// hits, so per-lane totals even out and the wavefront stays full until the
// end.  The seed's own score (at most 32 columns) is taken when the hit is
// fetched: score bytes masked to the seed span, summed four at a time with a
// signed dot product.  Per walk column: one signed byte extract, a saturating
// add, max/compare for the best prefix and its extent, and the x-drop test;
// after the first stop the running score is pinned to a large negative value
// so later columns of the chunk cannot move the best (src/hsp.rs:179-193
// semantics, branch-free).  A left walk loads the 32 bytes before its current
// end and reverses the score words in registers, so it touches the same cache
// lines as the seed and the right walk.
enum : uint32_t { kPhRight = 1, kPhLeft = 2, kPhDone = 3 };

struct XdropArgs {
    const uint8_t *tf, *qf;    // byte 0 of each sequence (padded buffers)
    uint64_t tlen, qlen;
    const Blk *tblk;           // block layout of the target (entropy gate)
    ByteLut lut;
    int32_t xdrop, thresh;
    uint32_t span;
    int entropy_on;
    int32_t entropy_thresh;
};

// Hit sources: next(t, q) yields the lane's next seed hit, `cur` its index
// where results are written per hit.
struct HitArraySrc {  // (q << 32 | t) records
    static constexpr bool kResume = false;
    const uint64_t *hits;
    uint64_t n, h, stride, cur;
    __device__ void init(uint32_t blk, uint32_t nblk, uint32_t tid, uint32_t bdim, uint64_t *) {
        h = (uint64_t)blk * bdim + tid;
        stride = (uint64_t)nblk * bdim;
    }
    __device__ __forceinline__ bool next(uint32_t &t, uint32_t &q) {
        if (h >= n) return false;
        const uint64_t x = hits[h];
        q = (uint32_t)(x >> 32);
        t = (uint32_t)x;
        cur = h;
        h += stride;
        return true;
    }
};

struct PairSrc {  // separate t and q position arrays
    static constexpr bool kResume = false;
    const uint32_t *tpos, *qpos;
    uint64_t n, h, stride, cur;
    __device__ void init(uint32_t blk, uint32_t nblk, uint32_t tid, uint32_t bdim, uint64_t *) {
        h = (uint64_t)blk * bdim + tid;
        stride = (uint64_t)nblk * bdim;
    }
    __device__ __forceinline__ bool next(uint32_t &t, uint32_t &q) {
        if (h >= n) return false;
        t = tpos[h];
        q = qpos[h];
        cur = h;
        h += stride;
        return true;
    }
};

// Hits of one bin per XCD: block b walks bin b % 8 with the other blocks of
// its XCD (the grid is a multiple of 8).  The bin is split into one
// contiguous segment per wavefront, and the wavefront hands out its segment
// in order: the lanes asking for a hit in one call take the next consecutive
// hits (a ballot ranks them; the head lives in LDS).  So the hits a
// wavefront holds at any time are neighbours in query order and its query
// reads share cache lines.
struct BinnedHitSrc {
    static constexpr bool kResume = false;
    const uint64_t *hits;
    const uint32_t *binoff;  // bin k starts at binoff[k * nblk_exp]
    uint32_t nblk_exp;
    uint64_t n, end, cur;
    volatile uint64_t *head;  // this wavefront's next hit (LDS)
    __device__ void init(uint32_t blk, uint32_t nblk, uint32_t tid, uint32_t bdim, uint64_t *lds) {
        const uint32_t bin = blk % kBins;
        const uint64_t lo = binoff[bin * nblk_exp], hi = bin + 1 < kBins ? binoff[(bin + 1) * nblk_exp] : n;
        const uint64_t wpb = bdim / 64, wi = (blk / kBins) * wpb + tid / 64, nw = (nblk / kBins) * wpb;
        end = lo + (hi - lo) * (wi + 1) / nw;
        head = lds;
        if ((tid & 63) == 0) *head = lo + (hi - lo) * wi / nw;
    }
    // Every lane that needs a hit must make this call together.
    __device__ __forceinline__ bool next(uint32_t &t, uint32_t &q) {
        const uint64_t asking = __ballot(1);
        const int lane = threadIdx.x & 63;
        const uint32_t rank = (uint32_t)__popcll(asking & ((1ull << lane) - 1ull));
        const uint64_t h0 = *head;
        if (rank == 0) *head = h0 + (uint64_t)__popcll(asking);
        const uint64_t h = h0 + rank;
        if (h >= end) return false;
        const uint64_t x = hits[h];
        q = (uint32_t)(x >> 32);
        t = (uint32_t)x;
        cur = h;
        return true;
    }
};

// A hit whose walks are not both over after their first 32 columns
// (k_xdrop_first): its seed score and each side's state after that chunk,
// resumed by k_xdrop through ContSrc.
struct XdropCont {
    uint32_t t, q;
    int32_t seed;
    int32_t rrun, rbest, lrun, lbest;
    uint32_t st;  // right extent (bits 0-7), left extent (8-15), kContREnded, kContLEnded
};
constexpr uint32_t kContREnded = 1u << 16, kContLEnded = 1u << 17;

// Continuation records, laid out by the first step's wavefronts: wavefront
// w of the first step writes its records from cont_base(w) on and its count
// to counts[w]; wavefront w of the resuming launch (same grid, so the same
// XCD and the same target bin) takes exactly those, in order, handing them
// to its lanes through an LDS head as BinnedHitSrc does.
struct ContLayout {
    const uint32_t *binoff;  // bin k's hits start at binoff[k * nblk_exp]; nullptr: not binned
    uint32_t nblk_exp;
    uint64_t nh;
    // first record of wavefront `wv` of block `blk` in a grid of `nblk`
    // blocks of 256: the wavefront's hit groups of 64 (one per grid-stride
    // step over its bin) bound its records; bin k's region is padded by one
    // grid-stride of hits, so regions never meet
    __device__ __forceinline__ uint64_t base(uint32_t blk, uint32_t nblk, uint32_t wv, uint64_t *lo_out,
                                             uint64_t *hi_out) const {
        uint64_t lo = 0, hi = nh, pad = 0, j, J;
        if (binoff) {
            const uint32_t bin = blk % kBins;
            lo = binoff[bin * nblk_exp];
            hi = bin + 1 < kBins ? binoff[(bin + 1) * nblk_exp] : nh;
            J = (uint64_t)(nblk / kBins) * 4;
            j = (uint64_t)(blk / kBins) * 4 + wv;
            pad = (uint64_t)bin * J * 64;
        } else {
            J = (uint64_t)nblk * 4;
            j = (uint64_t)blk * 4 + wv;
        }
        const uint64_t K = (hi - lo + J * 64 - 1) / (J * 64);  // grid-stride steps over the bin
        *lo_out = lo;
        *hi_out = hi;
        return lo + pad + j * K * 64;
    }
};
// records the layout needs for nh hits in a grid of nblk blocks
inline uint64_t cont_capacity(uint64_t nh, uint32_t nblk) { return nh + (uint64_t)nblk * 256; }

struct ContSrc {
    static constexpr bool kResume = true;
    const XdropCont *recs;
    ContLayout lay;
    const uint32_t *counts;  // per wavefront of the grid
    uint64_t end, cur;
    volatile uint64_t *head;  // this wavefront's next record (LDS)
    XdropCont rec;
    __device__ void init(uint32_t blk, uint32_t nblk, uint32_t tid, uint32_t bdim, uint64_t *lds) {
        uint64_t lo, hi;
        const uint32_t wv = tid / 64;
        const uint64_t b = lay.base(blk, nblk, wv, &lo, &hi);
        end = b + counts[(uint64_t)blk * (bdim / 64) + wv];
        head = lds;
        if ((tid & 63) == 0) *head = b;
    }
    // Every lane that needs a record must make this call together.
    __device__ __forceinline__ bool next(uint32_t &t, uint32_t &q) {
        const uint64_t asking = __ballot(1);
        const int lane = threadIdx.x & 63;
        const uint32_t rank = (uint32_t)__popcll(asking & ((1ull << lane) - 1ull));
        const uint64_t h0 = *head;
        if (rank == 0) *head = h0 + (uint64_t)__popcll(asking);
        const uint64_t h = h0 + rank;
        if (h >= end) return false;
        rec = recs[h];
        t = rec.t;
        q = rec.q;
        cur = h;
        return true;
    }
};

template <class Src, bool kAppend>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5, 8))) void k_xdrop(
    XdropArgs a, Src src, PassOut po, gxy_hsp *__restrict__ out, uint8_t *__restrict__ found,
    unsigned long long *__restrict__ colsum) {
    __shared__ uint64_t src_lds[4];  // one word per wavefront for the hit source
    src.init(blockIdx.x, gridDim.x, threadIdx.x, blockDim.x, &src_lds[threadIdx.x >> 6]);
    const int lane = threadIdx.x & 63;
    uint32_t tp = 0, qp = 0, phase = kPhDone;
    uint64_t pt = 0, pq = 0, rem = 0;
    int32_t seed = 0, run = 0, best = 0, rbest = 0;
    uint32_t bext = 0, rext = 0, base = 0;
    uint64_t cols = 0;
    constexpr int32_t kPinned = -(1 << 30);
    constexpr int32_t kSmall = 1 << 29;
    const bool xd_small = a.xdrop < kSmall;
    // seed span in score words; bytes of the last word inside the span
    const uint32_t nsw = (a.span + 3) >> 2;
    const uint32_t lastmask = (a.span & 3) ? ((1u << (8 * (a.span & 3))) - 1u) : 0xFFFFFFFFu;
    // a resumed hit's left side after its first chunk (k_xdrop_first)
    int32_t lrun = 0, lbest = 0;
    uint32_t lext = 0;
    bool lres = false, lend = false;

    auto begin_left = [&]() {  // left walk from (t-1, q-1), src/hsp.rs:129
        phase = kPhLeft;
        pt = tp;  // exclusive end: the next column is (pt-1, pq-1)
        pq = qp;
        rem = min((uint64_t)tp, (uint64_t)qp);
        run = best = 0;
        bext = base = 0;
        if (lres) {  // go on from the state after the first 32 columns
            lres = false;
            best = lbest;
            bext = lext;
            if (lend) {
                rem = 0;
            } else {
                pt -= 32;
                pq -= 32;
                rem -= 32;
                run = lrun;
                base = 32;
            }
        }
    };
    auto fetch = [&]() {
        phase = kPhDone;
        while (src.next(tp, qp)) {
            if constexpr (Src::kResume) {
                const XdropCont &c = src.rec;
                seed = c.seed;
                lres = true;
                lend = (c.st & kContLEnded) != 0;
                lrun = c.lrun;
                lbest = c.lbest;
                lext = (c.st >> 8) & 0xFFu;
                if (c.st & kContREnded) {  // only the left walk goes on
                    rbest = c.rbest;
                    rext = c.st & 0xFFu;
                    begin_left();
                } else {  // the right walk goes on from column 32 past the seed
                    phase = kPhRight;
                    pt = (uint64_t)tp + a.span + 32;
                    pq = (uint64_t)qp + a.span + 32;
                    rem = min(a.tlen - pt, a.qlen - pq);
                    run = c.rrun;
                    best = c.rbest;
                    bext = c.st & 0xFFu;
                    base = 32;
                }
                return;
            }
            if ((uint64_t)tp + a.span <= a.tlen && (uint64_t)qp + a.span <= a.qlen) {  // src/hsp.rs:113-115
                // seed score (src/hsp.rs:118-124); |sum| <= 32*128, no saturation
                uint32_t tw[8], qw[8];
                load_bytes32(a.tf, tp, tw);
                load_bytes32(a.qf, qp, qw);
                int32_t sd = 0;
#pragma unroll
                for (int w = 0; w < 8; w++) {
                    if ((uint32_t)w < nsw) {
                        uint32_t sw = score4(a.lut, tw[w], qw[w]);
                        if ((uint32_t)w == nsw - 1) sw &= lastmask;
                        sd = __builtin_amdgcn_sdot4((int)sw, 0x01010101, sd, false);
                    }
                }
                seed = sd;
                cols += a.span;
                // right walk from the seed end (src/hsp.rs:126-127)
                phase = kPhRight;
                pt = (uint64_t)tp + a.span;
                pq = (uint64_t)qp + a.span;
                rem = (pt < a.tlen && pq < a.qlen) ? min(a.tlen - pt, a.qlen - pq) : 0;
                run = best = 0;
                bext = base = 0;
                if (rem == 0) {  // no right columns at all
                    rbest = 0;
                    rext = 0;
                    begin_left();
                }
                return;
            }
            if (!kAppend) {
                out[src.cur] = gxy_hsp{0, 0, 0, 0};
                found[src.cur] = 0;
            }
        }
    };
    fetch();

    while (__any(phase != kPhDone)) {
        bool finished = false;  // this lane completed a hit this iteration
        bool pass = false;
        gxy_hsp res = {0, 0, 0, 0};
        if (phase != kPhDone) {
            const bool rv = phase == kPhLeft;
            uint32_t tw[8], qw[8], sw[8];
            load_bytes32(a.tf, rv ? (int64_t)pt - 32 : (int64_t)pt, tw);
            load_bytes32(a.qf, rv ? (int64_t)pq - 32 : (int64_t)pq, qw);
#pragma unroll
            for (int w = 0; w < 8; w++) sw[w] = score4(a.lut, tw[w], qw[w]);
            if (rv) {  // column k of a left walk is byte 31-k of the window
                uint32_t r[8];
#pragma unroll
                for (int w = 0; w < 8; w++) r[w] = __builtin_bswap32(sw[7 - w]);
#pragma unroll
                for (int w = 0; w < 8; w++) sw[w] = r[w];
            }
            const uint32_t n = (uint32_t)min<uint64_t>(32, rem);
            if (n < 32) {  // columns past the sequence end must not move the best
#pragma unroll
                for (int w = 0; w < 8; w++) {
                    const uint32_t live = n > 4u * w ? min(4u, n - 4u * w) : 0u;
                    const uint32_t keep = live >= 4 ? 0xFFFFFFFFu : ((1u << (8 * live)) - 1u);
                    sw[w] = (sw[w] & keep) | (0x80808080u & ~keep);
                }
            }
            // x_drop >= 0 here (the generic kernel takes the rest), so a
            // column that improves the best can never satisfy the stop test:
            // the reference's else-branch needs no separate guard.  After the
            // first stop every later column stops too (the pinned run stays
            // far below best - x_drop), so the stopping columns are
            // [kstop, 32).
            uint32_t kbest = 0, kstop = 32;
            const bool small = xd_small && run >= -kSmall && run <= kSmall && best <= kSmall;
            if (__all(small)) {
                // no sum can leave int32 within a chunk: plain adds, the byte
                // extract folds into the add
                uint32_t nst = 0;
#pragma unroll
                for (int k = 0; k < 32; k++) {
                    const int32_t sc = (int32_t)(int8_t)(sw[k >> 2] >> (8 * (k & 3)));
                    const int32_t r = run + sc;
                    kbest = r > best ? (uint32_t)(k + 1) : kbest;
                    best = max(best, r);
                    const bool stop = best - r > a.xdrop;
                    nst += stop ? 1u : 0u;
                    run = stop ? kPinned : r;
                }
                kstop = 32 - nst;
            } else {  // saturating form of the same walk (src/hsp.rs:186)
                // its own copy of the score words: byte extracts shared with
                // the fast form would be hoisted above the branch as 32
                // separate instructions instead of folding into the adds
                uint32_t sv[8];
#pragma unroll
                for (int w = 0; w < 8; w++) {
                    sv[w] = sw[w];
                    asm volatile("" : "+v"(sv[w]));
                }
                uint32_t stopbits = 0;
#pragma unroll
                for (int k = 0; k < 32; k++) {
                    const int32_t sc = __builtin_amdgcn_sbfe(sv[k >> 2], 8 * (k & 3), 8);
                    const int32_t r = sat_add(run, sc);
                    kbest = r > best ? (uint32_t)(k + 1) : kbest;
                    best = max(best, r);
                    const bool stop = sat_sub(best, r) > a.xdrop;
                    stopbits |= stop ? (1u << k) : 0u;
                    run = stop ? kPinned : r;
                }
                kstop = stopbits ? (uint32_t)__ffs(stopbits) - 1u : 32u;
            }
            if (kbest) bext = base + kbest;
            const bool live_stop = kstop < n;
            const bool ended = live_stop || rem <= 32;
            cols += live_stop ? kstop + 1 : n;
            rem -= n;
            if (rv) {
                pt -= n;
                pq -= n;
            } else {
                pt += n;
                pq += n;
            }
            base += 32;
            if (ended) {
                if (phase == kPhRight) {
                    rbest = best;
                    rext = bext;
                    begin_left();
                    // no left columns (left): best and bext are its result
                    if (rem == 0) finished = true;
                } else {
                    finished = true;
                }
            }
            if (finished) {  // src/hsp.rs:131-147
                const int32_t score = sat_add(sat_add(seed, best), rbest);
                res.t_start = tp - bext;
                res.q_start = qp - bext;
                res.length = a.span + bext + rext;
                res.score = score;
                pass = score >= a.thresh;
                if (pass && a.entropy_on)
                    pass = entropy_pass(a.tblk, a.tlen, res.t_start, res.t_start + res.length, res.score,
                                        a.entropy_thresh);
                if (!kAppend) {
                    out[src.cur] = pass ? res : gxy_hsp{0, 0, 0, 0};
                    found[src.cur] = pass ? 1 : 0;
                }
            }
        }
        if (kAppend) {
            const unsigned long long b = __ballot(pass);
            if (b) {
                unsigned long long basei = 0;
                const int first = __ffsll((long long)b) - 1;
                if (lane == first) basei = atomicAdd(po.count, (unsigned long long)__popcll(b));
                basei = __shfl(basei, first);
                if (pass) {
                    const uint64_t slot = basei + __popcll(b & ((1ull << lane) - 1ull));
                    const uint64_t d = (uint64_t)((int64_t)tp - (int64_t)qp + (int64_t)a.qlen);
                    po.keys[slot] = (d << 32) | qp;
                    po.hsp[slot] = res;
                }
            }
        }
        if (finished) fetch();
    }
    unsigned long long c = cols;
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
    if (lane == 0 && c) atomicAdd(colsum, c);
}

// ---- first step of every hit, straight-line --------------------------------
//
// Most seed hits are chance matches whose walks both stop within ~21 columns
// (HOXD70 averages -43.4 per random column against an x-drop of 910): in
// 82 % of them both walks end inside their first 32 columns.  This kernel
// gives every lane one hit at a time and does, with no per-lane state
// machine: the seed score, the first 32 columns right of the seed and the 32
// left of it (the same byte loads and score words as k_xdrop), both scans
// side by side.  A hit whose walks are both over is finished here (passing
// ones appended as in k_xdrop); the others are written as XdropCont records
// for k_xdrop to resume.  Needs x_drop < 2^29 (plain adds are exact then).
// query bytes a wavefront of k_xdrop_first keeps in LDS
constexpr int kQLds = 1024;
// 32 consecutive bytes at byte offset `off` of an LDS word array.
__device__ __forceinline__ void lds_bytes32(const uint32_t *w, uint32_t off, uint32_t out[8]) {
    const uint32_t *p = w + (off >> 2);
    const uint32_t sh = off & 3u;
    uint32_t d[9];
#pragma unroll
    for (int i = 0; i < 9; i++) d[i] = p[i];
#pragma unroll
    for (int i = 0; i < 8; i++) out[i] = __builtin_amdgcn_alignbyte(d[i + 1], d[i], sh);
}

// The 32-column scan of k_xdrop's fast form from (run, best): kbest is the
// 1-based column of the last improvement (0: none), kstop the first stopping
// column (32: none); run is pinned after a stop.
__device__ __forceinline__ void scan32(const uint32_t sw[8], int32_t xdrop, int32_t &run, int32_t &best,
                                       uint32_t &kbest, uint32_t &kstop) {
    constexpr int32_t kPinned = -(1 << 30);
    uint32_t nst = 0;
    kbest = 0;
#pragma unroll
    for (int k = 0; k < 32; k++) {
        const int32_t sc = (int32_t)(int8_t)(sw[k >> 2] >> (8 * (k & 3)));
        const int32_t r = run + sc;
        kbest = r > best ? (uint32_t)(k + 1) : kbest;
        best = max(best, r);
        const bool stop = best - r > xdrop;
        nst += stop ? 1u : 0u;
        run = stop ? kPinned : r;
    }
    kstop = 32 - nst;
}

// Both first chunks of a hit at once, in packed 16-bit lanes: the right walk
// in the low halves, the left walk in the high halves.  From a fresh start
// every value fits 16 bits (|run| <= 32 * 128; best - run <= 8192), so the
// walk's arithmetic is exact there; x_drop is clamped to 16383, above any
// drop a chunk can reach.  Per column and both sides: the two score bytes
// placed as the high bytes of the halves and shifted down (sign-extended),
// the run, best - run before the column (< 0: the column improves the best,
// > x_drop: it stops the walk), the new best, the last improving column
// (bit-select on the sign of best - run), a sticky stopped mask and a count
// of stopped columns.  The walk is not frozen at its first stop, so the
// result is the reference's (src/hsp.rs:179-193) unless the best improves
// after that stop — then kbest > kstop and the caller redoes the side with
// scan32.  Per side: run, best, kbest, kstop as in scan32.
typedef short s16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void scan32x2(const uint32_t rs[8], const uint32_t ls[8], int32_t xdrop, int32_t &rrun,
                                         int32_t &rbest, uint32_t &rk, uint32_t &rstop, int32_t &lrun, int32_t &lbest,
                                         uint32_t &lk, uint32_t &lstop) {
    const short X = (short)min(xdrop, 16383);
    const s16x2 xp = {X, X};
    s16x2 r = {0, 0}, b = {0, 0}, c = {0, 0};
    uint32_t kb = 0, stopped = 0;
#pragma unroll
    for (int k = 0; k < 32; k++) {
        const uint32_t by = (uint32_t)(k & 3);
        // bytes: 0, right byte `by`, 0, left byte `by` (v_perm: rs is bytes 0-3, ls 4-7, 12 gives 0)
        const uint32_t sel = 0x0Cu | (by << 8) | (0x0Cu << 16) | ((4u + by) << 24);
        const s16x2 s = __builtin_bit_cast(s16x2, __builtin_amdgcn_perm(ls[k >> 2], rs[k >> 2], sel)) >> (short)8;
        r = r + s;
        const s16x2 d = b - r;
        b = __builtin_elementwise_max(b, r);
        const uint32_t imp = __builtin_bit_cast(uint32_t, d >> (short)15);
        const uint32_t kk = (uint32_t)(k + 1) * 0x00010001u;
        kb = (imp & kk) | (~imp & kb);
        stopped |= __builtin_bit_cast(uint32_t, (xp - d) >> (short)15);
        c = c - __builtin_bit_cast(s16x2, stopped);
    }
    rrun = r.x;
    rbest = b.x;
    rk = kb & 0xFFFFu;
    rstop = 32u - (uint32_t)(uint16_t)c.x;
    lrun = r.y;
    lbest = b.y;
    lk = kb >> 16;
    lstop = 32u - (uint32_t)(uint16_t)c.y;
}

// kSpan: the seed span when it is known at compile time (the target bytes of
// a hit are then one dword-aligned window), 0 for any span.
template <int kSpan>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5, 8))) void k_xdrop_first(
    XdropArgs a, const uint64_t *__restrict__ hits, ContLayout lay, PassOut po, XdropCont *__restrict__ cont,
    uint32_t *__restrict__ cont_counts, unsigned long long *__restrict__ stats) {
    const int lane = threadIdx.x & 63;
    const uint32_t wv = threadIdx.x >> 6;
    // this block's hits: its bin's (block b walks bin b % 8 with the other
    // blocks of its XCD, 256 consecutive hits per block and grid-stride step)
    // or all of them; the wavefront's continuation records from cbase on
    uint64_t lo, hi;
    const uint64_t cbase = lay.base(blockIdx.x, gridDim.x, wv, &lo, &hi);
    const uint64_t nbin = lay.binoff ? gridDim.x / kBins : gridDim.x;
    const uint64_t h = lo + (uint64_t)(lay.binoff ? blockIdx.x / kBins : blockIdx.x) * 256 + threadIdx.x;
    const uint64_t stride = nbin * 256;
    uint32_t ncont = 0;  // this wavefront's records so far (uniform)
    uint32_t npass = 0;  // this wavefront's passing extensions (uniform)
    const uint32_t nsw = (a.span + 3) >> 2;
    const uint32_t lastmask = (a.span & 3) ? ((1u << (8 * (a.span & 3))) - 1u) : 0xFFFFFFFFu;
    uint64_t cols = 0;
    // The wavefront's 64 hits come from neighbouring query windows, so their
    // query bytes overlap: the wavefront loads the range they span once,
    // coalesced, into LDS and each lane reads its windows there; only the
    // target bytes (random per hit) go through the vector memory pipeline,
    // which bounds this kernel.  Ranges wider than kQLds bytes read the query
    // from global memory as well.
    __shared__ __align__(16) uint32_t qlds[4][kQLds / 4];
    uint32_t *ql = qlds[wv];
    const uint64_t hw0 = h - (uint64_t)lane;  // the wavefront's first hit (uniform loop below)
    for (uint64_t hw = hw0; hw < hi; hw += stride) {
        const uint64_t hh = hw + (uint64_t)lane;
        const bool act = hh < hi;
        const uint64_t x = act ? hits[hh] : 0;
        const uint32_t q = (uint32_t)(x >> 32), t = (uint32_t)x;
        // src/hsp.rs:113-115: a seed out of bounds has no HSP
        const bool ok = act && (uint64_t)t + a.span <= a.tlen && (uint64_t)q + a.span <= a.qlen;
        bool pass = false, more = false;
        gxy_hsp res = {0, 0, 0, 0};
        XdropCont c;
        // a hit reads query bytes [q - 32, q + span + 36)
        uint32_t qmin = ok ? q : 0xFFFFFFFFu, qmax = ok ? q : 0u;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            qmin = min(qmin, (uint32_t)__shfl_xor((int)qmin, o));
            qmax = max(qmax, (uint32_t)__shfl_xor((int)qmax, o));
        }
        const int64_t qb = ((int64_t)qmin - 32) & ~(int64_t)15;  // >= -47: inside the padding
        const int64_t qend = (int64_t)qmax + a.span + 36;
        const bool in_lds = qmin != 0xFFFFFFFFu && qend <= qb + kQLds;  // uniform
        if (in_lds) {
            const int64_t p = qb + 16 * lane;
            if (p < qend) *(uint4 *)&ql[4 * lane] = *(const uint4 *)(a.qf + p);
            // this wavefront's stores before its own loads below
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        }
        if (ok) {
            // bytes [t - 32, t + 32): the left chunk, then the seed; tr: the
            // right chunk, 32 bytes from the seed end
            uint32_t tw[16], qw[16], tr[8], qr[8];
            if constexpr (kSpan > 0) {
                // one window [t - 32, t + kSpan + 32) of dword-aligned loads;
                // the right chunk starts at dword r0 or r0 + 1 of it
                constexpr int ND = ((66 + kSpan) >> 2) + 2, r0 = (32 + kSpan) >> 2;
                const int64_t p0 = (int64_t)t - 32;
                const uint32_t *tp = (const uint32_t *)(a.tf + (p0 & ~(int64_t)3));
                const uint32_t sh = (uint32_t)(p0 & 3);
                uint32_t d[ND];
#pragma unroll
                for (int i = 0; i < ND; i++) d[i] = tp[i];
#pragma unroll
                for (int i = 0; i < 16; i++) tw[i] = __builtin_amdgcn_alignbyte(d[i + 1], d[i], sh);
                const bool up = ((sh + 32 + kSpan) >> 2) != (uint32_t)r0;
                const uint32_t rsh = (sh + 32 + kSpan) & 3u;
#pragma unroll
                for (int i = 0; i < 8; i++)
                    tr[i] = __builtin_amdgcn_alignbyte(up ? d[r0 + i + 2] : d[r0 + i + 1], up ? d[r0 + i + 1] : d[r0 + i],
                                                       rsh);
            } else {
                load_bytes32(a.tf, (int64_t)t - 32, tw);
                load_bytes32(a.tf, (int64_t)t, tw + 8);
                load_bytes32(a.tf, (int64_t)t + a.span, tr);
            }
            if (in_lds) {
                lds_bytes32(ql, (uint32_t)((int64_t)q - 32 - qb), qw);
                lds_bytes32(ql, (uint32_t)((int64_t)q - qb), qw + 8);
                lds_bytes32(ql, (uint32_t)((int64_t)q + a.span - qb), qr);
            } else {
                load_bytes32(a.qf, (int64_t)q - 32, qw);
                load_bytes32(a.qf, (int64_t)q, qw + 8);
                load_bytes32(a.qf, (int64_t)q + a.span, qr);
            }
            int32_t seed = 0;  // src/hsp.rs:118-124; |sum| <= 32 * 128
#pragma unroll
            for (int w = 0; w < 8; w++) {
                if ((uint32_t)w < nsw) {
                    uint32_t sw = score4(a.lut, tw[8 + w], qw[8 + w]);
                    if ((uint32_t)w == nsw - 1) sw &= lastmask;
                    seed = __builtin_amdgcn_sdot4((int)sw, 0x01010101, seed, false);
                }
            }
            uint32_t rs[8], ls[8];
#pragma unroll
            for (int w = 0; w < 8; w++) {
                rs[w] = score4(a.lut, tr[w], qr[w]);
                // column k of the left walk is byte 31 - k of [t - 32, t)
                ls[w] = __builtin_bswap32(score4(a.lut, tw[7 - w], qw[7 - w]));
            }
            // columns left on each side (src/hsp.rs:126-129)
            const uint64_t pt = (uint64_t)t + a.span, pq = (uint64_t)q + a.span;
            const uint64_t remR = (pt < a.tlen && pq < a.qlen) ? min(a.tlen - pt, a.qlen - pq) : 0;
            const uint64_t remL = min(t, q);
            const uint32_t nR = (uint32_t)min<uint64_t>(32, remR), nL = (uint32_t)min<uint64_t>(32, remL);
            if (nR < 32 || nL < 32) {  // columns past a sequence end must not move the best
#pragma unroll
                for (int w = 0; w < 8; w++) {
                    const uint32_t liveR = nR > 4u * w ? min(4u, nR - 4u * w) : 0u;
                    const uint32_t keepR = liveR >= 4 ? 0xFFFFFFFFu : ((1u << (8 * liveR)) - 1u);
                    rs[w] = (rs[w] & keepR) | (0x80808080u & ~keepR);
                    const uint32_t liveL = nL > 4u * w ? min(4u, nL - 4u * w) : 0u;
                    const uint32_t keepL = liveL >= 4 ? 0xFFFFFFFFu : ((1u << (8 * liveL)) - 1u);
                    ls[w] = (ls[w] & keepL) | (0x80808080u & ~keepL);
                }
            }
            int32_t rrun, rbest, lrun, lbest;
            uint32_t rk, rstop, lk, lstop;
            scan32x2(rs, ls, a.xdrop, rrun, rbest, rk, rstop, lrun, lbest, lk, lstop);
            // the best improved after a walk's first stop (a climb of more than
            // x_drop inside the chunk): that side again, frozen at the stop
            if (rk > rstop) {
                rrun = rbest = 0;
                scan32(rs, a.xdrop, rrun, rbest, rk, rstop);
            }
            if (lk > lstop) {
                lrun = lbest = 0;
                scan32(ls, a.xdrop, lrun, lbest, lk, lstop);
            }
            const bool rlive = rstop < nR, llive = lstop < nL;
            const bool rend = rlive || remR <= 32, lend = llive || remL <= 32;
            cols += a.span + (rlive ? rstop + 1 : nR) + (llive ? lstop + 1 : nL);
            if (rend && lend) {  // src/hsp.rs:131-147
                const int32_t score = sat_add(sat_add(seed, lbest), rbest);
                res.t_start = t - lk;
                res.q_start = q - lk;
                res.length = a.span + lk + rk;
                res.score = score;
                pass = score >= a.thresh;
                if (pass && a.entropy_on)
                    pass = entropy_pass(a.tblk, a.tlen, res.t_start, res.t_start + res.length, res.score,
                                        a.entropy_thresh);
            } else {
                more = true;
                c = XdropCont{t, q, seed, rrun, rbest, lrun, lbest,
                              rk | (lk << 8) | (rend ? kContREnded : 0u) | (lend ? kContLEnded : 0u)};
            }
        }
        // appends, one atomic per wavefront each
        const unsigned long long bp = __ballot(pass);
        if (bp) {
            unsigned long long basei = 0;
            const int first = __ffsll((long long)bp) - 1;
            if (lane == first) basei = atomicAdd(po.count, (unsigned long long)__popcll(bp));
            basei = __shfl(basei, first);
            if (pass) {
                const uint64_t slot = basei + __popcll(bp & ((1ull << lane) - 1ull));
                const uint64_t d = (uint64_t)((int64_t)t - (int64_t)q + (int64_t)a.qlen);
                po.keys[slot] = (d << 32) | q;
                po.hsp[slot] = res;
            }
            npass += (uint32_t)__popcll(bp);
        }
        const unsigned long long bm = __ballot(more);
        if (more) cont[cbase + ncont + __popcll(bm & ((1ull << lane) - 1ull))] = c;
        ncont += (uint32_t)__popcll(bm);
    }
    // every lane took part in every step: the counts are the wavefront's
    ncont = (uint32_t)__builtin_amdgcn_readfirstlane(ncont);
    npass = (uint32_t)__builtin_amdgcn_readfirstlane(npass);
    if (lane == 0) cont_counts[(uint64_t)blockIdx.x * 4 + wv] = ncont;
    // stats: [0] columns, [1] continuation records, [2] passing extensions
    unsigned long long cs = cols;
    for (int o = 32; o > 0; o >>= 1) cs += __shfl_xor(cs, o);
    if (lane == 0) {
        if (cs) atomicAdd(stats, cs);
        if (ncont) atomicAdd(stats + 1, (unsigned long long)ncont);
        if (npass) atomicAdd(stats + 2, (unsigned long long)npass);
    }
}

__global__ void k_iota(uint32_t *__restrict__ v, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = (uint32_t)i;
}

// Diagonal admission scan (src/diag_hash.rs:34-53 as used by
// src/seed_search.rs:95-114) over passing hits sorted by (diag, q).
__global__ void k_admit(const uint64_t *__restrict__ keys, const uint32_t *__restrict__ idx,
                        const gxy_hsp *__restrict__ hsp, uint64_t n, uint64_t qlen, uint32_t *__restrict__ flag) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t d = keys[i] >> 32;
    if (i > 0 && (keys[i - 1] >> 32) == d) return;  // not the head of its diagonal run
    bool has = false;
    uint32_t cover = 0;
    for (uint64_t j = i; j < n && (keys[j] >> 32) == d; j++) {
        const uint32_t q = (uint32_t)keys[j];
        const uint32_t t = (uint32_t)((int64_t)d - (int64_t)qlen + (int64_t)q);
        const bool emit = !has || t >= cover;
        flag[j] = emit ? 1u : 0u;
        if (emit) {
            const gxy_hsp r = hsp[idx[j]];
            const uint32_t te = r.t_start + r.length;
            cover = has ? max(cover, te) : te;
            has = true;
        }
    }
}

__global__ void k_emit(const uint64_t *__restrict__ keys, const uint32_t *__restrict__ idx,
                       const gxy_hsp *__restrict__ hsp, const uint32_t *__restrict__ flag,
                       const uint32_t *__restrict__ pos, uint64_t n, uint64_t *__restrict__ out_keys,
                       uint32_t *__restrict__ out_idx) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !flag[i]) return;
    const uint32_t k = idx[i];
    const uint64_t d = keys[i] >> 32;
    out_keys[pos[i]] = (d << 32) | hsp[k].t_start;
    out_idx[pos[i]] = k;
}

__global__ void k_gather_hsp(const uint32_t *__restrict__ idx, const gxy_hsp *__restrict__ hsp, uint64_t n,
                             gxy_hsp *__restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = hsp[idx[i]];
}

template <class S>
__global__ void k_extend_batch(const uint32_t *__restrict__ tpos, const uint32_t *__restrict__ qpos, uint64_t n,
                               const Blk *__restrict__ T, uint64_t tlen, const Blk *__restrict__ Q, uint64_t qlen,
                               S sc, int32_t xdrop, int32_t thresh, uint32_t span, gxy_hsp *__restrict__ out,
                               uint8_t *__restrict__ found) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t cols = 0;
    gxy_hsp r = {0, 0, 0, 0};
    const bool ok = xdrop_hit(T, tlen, Q, qlen, sc, xdrop, thresh, tpos[i], qpos[i], span, &r, &cols);
    out[i] = ok ? r : gxy_hsp{0, 0, 0, 0};
    found[i] = ok ? 1 : 0;
}

__global__ void k_find_bounds(const uint64_t *__restrict__ off, uint64_t nwin, uint64_t total, uint32_t nb,
                              uint64_t *__restrict__ bounds) {
    // bounds[k] = first window w with off[w] >= k*total/nb (k = 0..nb)
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k > nb) return;
    if (k == nb) {
        bounds[k] = nwin;
        return;
    }
    const uint64_t target = (uint64_t)((unsigned __int128)total * k / nb);
    uint64_t lo = 0, hi = nwin;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) / 2;
        if (off[mid] < target) lo = mid + 1; else hi = mid;
    }
    bounds[k] = lo;
}

}  // namespace

namespace gxy {

// Seed search against a device table and device query.  `respect_mask`
// mirrors the driver's --nomasking (masks cleared before seeding).
int search_dev_impl(gxy_ctx *ctx, const gxy_postable *tab, const gxy_dseq *query, const gxy_scoring *scm,
                    const gxy_search_params *sp, int respect_mask, std::vector<gxy_hsp> *result,
                    gxy_search_stats *stats, const uint8_t *qbytes, bool dedup) {
    if (!tab || !query || !scm || !sp || !result) return set_err(ctx, GXY_EINVAL, "null argument");
    if (sp->step < 1) return set_err(ctx, GXY_EINVAL, "step must be >= 1");
    const gxy_dseq *target = tab->target;
    const uint64_t tlen = target->len, qlen = query->len;
    if (tlen + qlen >= (1ull << 32)) return set_err(ctx, GXY_ERANGE, "target+query length must be < 4 Gbp");
    const DevPattern &pat = tab->pattern;
    Variants var{pat.weight, std::min<uint32_t>(sp->transitions, 2u)};
    const uint32_t nvar = var.transitions == 0 ? 1 : (var.transitions == 1 ? 1 + pat.weight
                                                                            : 1 + pat.weight + pat.weight * (pat.weight - 1) / 2);
    gxy_search_stats st;
    std::memset(&st, 0, sizeof st);
    result->clear();
    PhaseClock clk(ctx);

    const uint64_t max_start = qlen >= pat.span ? qlen - pat.span : 0;
    const uint64_t nwin = max_start / sp->step + 1;
    if (nwin > 0x7FFFFFFFull) return set_err(ctx, GXY_ERANGE, "too many query windows");
    auto *words = (uint32_t *)scratch(ctx, kSlotWords, nwin * 4);
    auto *cnt = (uint32_t *)scratch(ctx, kSlotCounts, nwin * 4);
    auto *off = (uint64_t *)scratch(ctx, kSlotOffsets, (nwin + 1) * 8);
    // counters: [1] passing extensions, [2] x-drop columns past the first
    // step, [3] first-step columns, [4] continuation records, [5] first-step
    // passing extensions
    auto *ctr = (unsigned long long *)scratch(ctx, kSlotCounters, 64);
    const bool fast_var = var.transitions <= 1 && nvar <= (uint32_t)kMaxV1;
    auto *vcnt = (uint32_t *)scratch(ctx, kSlotVMask, nwin * 4);  // per-variant bucket sizes, 2 bits each
    if (!words || !cnt || !off || !ctr || !vcnt) return set_err(ctx, GXY_ENOMEM, "scratch allocation failed");
    auto *spr = (unsigned long long *)scratch(ctx, kSlotSpread, kSpreadBytes);
    if (!spr) return set_err(ctx, GXY_ENOMEM, "scratch allocation failed");
    GXY_HIP(ctx, hipMemsetAsync(ctr, 0, 64, ctx->stream));
    GXY_HIP(ctx, hipMemsetAsync(spr, 0, kSpreadBytes, ctx->stream));
    // Hit offsets: per 256-window block on the fast path (k_qexpand_w
    // rebuilds the window offsets inside a block), per window otherwise.
    // `off` then has noff + 1 entries and a unit is unit_win windows.
    const uint64_t nblk_w = grid_for(nwin, 256);
    const uint64_t noff = fast_var ? nblk_w : nwin, unit_win = fast_var ? 256 : 1;
    auto *bsum = fast_var ? (uint64_t *)scratch(ctx, kSlotMisc3, nblk_w * 8) : nullptr;
    if (fast_var && !bsum) return set_err(ctx, GXY_ENOMEM, "scratch allocation failed");
    {
        KernelTimer kt_(ctx, "qcount");
        if (fast_var)
            k_qcount_v<<<(unsigned)nblk_w, 256, 0, ctx->stream>>>(query->d, qlen, pat, sp->step, nwin,
                                                                  respect_mask, nvar, tab->d_starts, tab->d_cmap,
                                                                  tab->d_ncnt, words, cnt, vcnt, bsum, spr);
        else
            k_qcount<<<grid_for(nwin, 256), 256, 0, ctx->stream>>>(query->d, qlen, pat, sp->step, nwin, respect_mask,
                                                                   var, tab->d_starts, tab->d_nonempty, words, cnt,
                                                                   spr);
    }
    GXY_LAUNCH_CHECK(ctx);
    size_t tmp_bytes = 0;
    hipcub::TransformInputIterator<uint64_t, ToU64, const uint32_t *> cnt64(cnt, ToU64());
    if (fast_var)
        (void)hipcub::DeviceScan::InclusiveSum(nullptr, tmp_bytes, bsum, off + 1, (int)noff, ctx->stream);
    else
        (void)hipcub::DeviceScan::InclusiveSum(nullptr, tmp_bytes, cnt64, off + 1, (int)noff, ctx->stream);
    void *tmp = scratch(ctx, kSlotTmp, tmp_bytes);
    if (!tmp) return set_err(ctx, GXY_ENOMEM, "scratch allocation failed");
    GXY_HIP(ctx, hipMemsetAsync(off, 0, 8, ctx->stream));
    {
        KernelTimer kt_(ctx, "qscan");
        if (fast_var)
            GXY_HIP(ctx, hipcub::DeviceScan::InclusiveSum(tmp, tmp_bytes, bsum, off + 1, (int)noff, ctx->stream));
        else
            GXY_HIP(ctx, hipcub::DeviceScan::InclusiveSum(tmp, tmp_bytes, cnt64, off + 1, (int)noff, ctx->stream));
    }
    uint64_t H = 0;
    unsigned long long nvalid = 0;
    GXY_HIP(ctx, hipMemcpyAsync(&H, off + noff, 8, hipMemcpyDeviceToHost, ctx->stream));
    {
        uint64_t nv = 0;
        int rc = spread_sum(ctx, spr, &nv);
        if (rc) return rc;
        nvalid = nv;
    }
    GXY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    clk.lap(kTrSearchCount);
    st.query_windows = nvalid;
    st.lookups = nvalid * nvar;
    st.hits = H;
    st.extended = H;

    const bool byte_ok = scoring_byte_ok(*scm) && tab->d_tbytes != nullptr && sp->hsp.x_drop >= 0;
    const ScorerI32 s32 = make_scorer_i32(*scm);
    const uint8_t *qb = qbytes;
    if (byte_ok && !qb) {
        auto *b = (uint8_t *)scratch(ctx, kSlotQBytes, byte_buffer_size(qlen));
        if (!b) return set_err(ctx, GXY_ENOMEM, "scratch allocation failed");
        int rc = seq_to_bytes(ctx, query, false, b);
        if (rc) return rc;
        qb = b + kBytePad;
    }
    uint64_t npass = 0;
    clk.lap(kTrSearchOther);
    if (H > 0) {
        // Hits are expanded and extended in batches of about `half` hits
        // (whole 256-window blocks; under 2 half): one batch for a search of
        // up to 2^27 hits, whose launches run larger and fewer.
        const uint64_t half = ctx->hit_batch ? ctx->hit_batch : (1ull << 27);
        // (no more batches than window blocks, and at most 2^20)
        const uint32_t nb = (uint32_t)std::min<uint64_t>(
            std::min<uint64_t>(std::max<uint64_t>(1, (H + half - 1) / half), std::max<uint64_t>(noff, 1)), 1u << 20);
        std::vector<uint64_t> bounds(nb + 1);  // in units of unit_win windows
        if (nb == 1) {
            bounds[0] = 0;
            bounds[1] = noff;
        } else {
            auto *db = (uint64_t *)scratch(ctx, kSlotMisc0, (nb + 1) * 8);
            if (!db) return set_err(ctx, GXY_ENOMEM, "scratch allocation failed");
            k_find_bounds<<<grid_for(nb + 1, 64), 64, 0, ctx->stream>>>(off, noff, H, nb, db);
            GXY_LAUNCH_CHECK(ctx);
            GXY_HIP(ctx, hipMemcpyAsync(bounds.data(), db, (nb + 1) * 8, hipMemcpyDeviceToHost, ctx->stream));
            GXY_HIP(ctx, hipStreamSynchronize(ctx->stream));
        }
        std::vector<uint64_t> boff(nb + 1);
        uint64_t maxb = 0;
        for (uint32_t k = 0; k <= nb; k++) {
            GXY_HIP(ctx, hipMemcpyAsync(&boff[k], off + bounds[k], 8, hipMemcpyDeviceToHost, ctx->stream));
        }
        GXY_HIP(ctx, hipStreamSynchronize(ctx->stream));
        for (uint32_t k = 0; k < nb; k++) maxb = std::max(maxb, boff[k + 1] - boff[k]);
        auto *hits = (uint64_t *)scratch(ctx, kSlotHits, maxb * 8);
        // Passing extensions: room for every hit of the batch about to run on
        // top of those already written (a hit passes at most once), grown
        // between batches — not for all H hits at once.
        uint64_t pcap = std::min<uint64_t>(H, maxb);
        auto *pkeys = (uint64_t *)scratch(ctx, kSlotPassKeys, pcap * 8);
        auto *phsp = (gxy_hsp *)scratch(ctx, kSlotPassHsp, pcap * sizeof(gxy_hsp));
        if (!hits || !pkeys || !phsp) return set_err(ctx, GXY_ENOMEM, "scratch allocation failed (%llu hits)",
                                                     (unsigned long long)H);
        // transitions <= 1 with the byte layout: hits binned by target eighth
        // for an XCD-partitioned x-drop
        const bool binned = fast_var && byte_ok;
        const uint32_t mbin = bin_multiplier(tlen);
        uint64_t maxw = 0;
        for (uint32_t k = 0; k < nb; k++) maxw = std::max(maxw, (bounds[k + 1] - bounds[k]) * unit_win);
        const uint64_t maxblk = grid_for(maxw, 64 * kExpWaves);
        auto *bincnt = (uint32_t *)scratch(ctx, kSlotBins, kBins * maxblk * 4);
        auto *binoff = (uint32_t *)scratch(ctx, kSlotBinOff, kBins * maxblk * 4);
        auto *hits2 = binned ? (uint64_t *)scratch(ctx, kSlotHits2, maxb * 8) : nullptr;
        if (!bincnt || !binoff || (binned && !hits2))
            return set_err(ctx, GXY_ENOMEM, "scratch allocation failed (%llu hits)", (unsigned long long)H);
        // the straight-line first step (k_xdrop_first) needs plain adds to be exact
        const bool first_step = byte_ok && sp->hsp.x_drop < (1 << 29) && !ctx->xdrop_single;
        auto *xcont = first_step
                          ? (XdropCont *)scratch(ctx, kSlotXCont, cont_capacity(maxb, kXdropBlocks) * sizeof(XdropCont))
                          : nullptr;
        auto *xccount = first_step ? (uint32_t *)scratch(ctx, kSlotXContCnt, kXdropBlocks * 4 * 4) : nullptr;
        if (first_step && (!xcont || !xccount))
            return set_err(ctx, GXY_ENOMEM, "scratch allocation failed (%llu hits)", (unsigned long long)H);
        uint64_t written = 0;  // passing extensions written before this batch
        for (uint32_t k = 0; k < nb; k++) {
            const uint64_t w0 = bounds[k] * unit_win, w1 = std::min(bounds[k + 1] * unit_win, nwin);
            const uint64_t nh = boff[k + 1] - boff[k];
            if (nh == 0) continue;
            if (k > 0) {
                unsigned long long c = 0;
                GXY_HIP(ctx, hipMemcpyAsync(&c, ctr + 1, 8, hipMemcpyDeviceToHost, ctx->stream));
                GXY_HIP(ctx, hipStreamSynchronize(ctx->stream));
                written = c;
            }
            if (written + nh > pcap) {
                pcap = std::min<uint64_t>(H, std::max<uint64_t>(written + nh, pcap + pcap / 2));
                pkeys = (uint64_t *)scratch_grow_keep(ctx, kSlotPassKeys, written * 8, pcap * 8);
                phsp = pkeys ? (gxy_hsp *)scratch_grow_keep(ctx, kSlotPassHsp, written * sizeof(gxy_hsp),
                                                            pcap * sizeof(gxy_hsp))
                             : nullptr;
                if (!pkeys || !phsp)
                    return set_err(ctx, GXY_ENOMEM, "scratch allocation failed (%llu passing extensions)",
                                   (unsigned long long)pcap);
            }
            PassOut po{pkeys, phsp, ctr + 1};
            const unsigned nblk = grid_for(w1 - w0, 64 * kExpWaves);
            {
                KernelTimer kt_(ctx, "qexpand");
                if (fast_var)
                    k_qexpand_w<<<nblk, 64 * kExpWaves, 0, ctx->stream>>>(w0, w1, sp->step, words, vcnt, cnt, off,
                                                                          boff[k], pat.weight, tab->d_starts,
                                                                          tab->d_positions, tab->d_spos, mbin, hits,
                                                                          bincnt);
                else
                    k_qexpand<<<grid_for(w1 - w0, 256), 256, 0, ctx->stream>>>(w0, w1, sp->step, words, off, boff[k],
                                                                              var, tab->d_starts, tab->d_nonempty,
                                                                              tab->d_positions, hits);
            }
            GXY_LAUNCH_CHECK(ctx);
            if (binned) {
                size_t tbb = 0;
                (void)hipcub::DeviceScan::ExclusiveSum(nullptr, tbb, bincnt, binoff, (int)(kBins * nblk), ctx->stream);
                void *tmpb = scratch(ctx, kSlotTmp, tbb);
                if (!tmpb) return set_err(ctx, GXY_ENOMEM, "scratch allocation failed");
                KernelTimer kt_(ctx, "bin");
                GXY_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(tmpb, tbb, bincnt, binoff, (int)(kBins * nblk),
                                                              ctx->stream));
                k_bin_hits<<<nblk, 64 * kExpWaves, 0, ctx->stream>>>(bounds[k], off, boff[k], hits, binoff, mbin, hits2);
                GXY_LAUNCH_CHECK(ctx);
            }
            if (byte_ok && first_step) {
                // the first 32 columns of both walks of every hit, then the hits
                // whose walks go on, resumed by the persistent kernel (the
                // continuation counts stay on the device)
                XdropArgs xa{tab->d_tbytes + kBytePad, qb, tlen, qlen, target->d,
                             make_byte_lut(*scm), sp->hsp.x_drop, sp->hsp.hsp_threshold, pat.span,
                             sp->entropy_enabled, sp->entropy_threshold};
                const unsigned g = std::min<unsigned>(
                    kXdropBlocks, std::max<unsigned>(kBins, (grid_for(nh, 256) + kBins - 1) / kBins * kBins));
                const ContLayout lay{binned ? binoff : nullptr, (uint32_t)nblk, nh};
                {
                    KernelTimer kt_(ctx, "extend");
                    const uint64_t *hsrc = binned ? hits2 : hits;
                    if (pat.span == 19)  // 12of19
                        k_xdrop_first<19><<<g, 256, 0, ctx->stream>>>(xa, hsrc, lay, po, xcont, xccount, ctr + 3);
                    else if (pat.span == 12)  // match12 (the tweener's inner seed)
                        k_xdrop_first<12><<<g, 256, 0, ctx->stream>>>(xa, hsrc, lay, po, xcont, xccount, ctr + 3);
                    else
                        k_xdrop_first<0><<<g, 256, 0, ctx->stream>>>(xa, hsrc, lay, po, xcont, xccount, ctr + 3);
                }
                GXY_LAUNCH_CHECK(ctx);
                {
                    KernelTimer kt_(ctx, "extend_cont");
                    const ContSrc src{xcont, lay, xccount};
                    k_xdrop<ContSrc, true><<<g, 256, 0, ctx->stream>>>(xa, src, po, nullptr, nullptr, ctr + 2);
                }
                GXY_LAUNCH_CHECK(ctx);
                continue;
            }
            KernelTimer kt_ext(ctx, "extend");
            if (byte_ok) {
                XdropArgs xa{tab->d_tbytes + kBytePad, qb, tlen, qlen, target->d,
                             make_byte_lut(*scm), sp->hsp.x_drop, sp->hsp.hsp_threshold, pat.span,
                             sp->entropy_enabled, sp->entropy_threshold};
                if (binned) {
                    // a multiple of kBins blocks, so every XCD walks its own bin
                    const unsigned g = std::min<unsigned>(
                        kXdropBlocks, std::max<unsigned>(kBins, (grid_for(nh, 256) + kBins - 1) / kBins * kBins));
                    BinnedHitSrc src{hits2, binoff, nblk, nh};
                    k_xdrop<BinnedHitSrc, true><<<g, 256, 0, ctx->stream>>>(xa, src, po, nullptr, nullptr, ctr + 2);
                } else {
                    HitArraySrc src{hits, nh};
                    const unsigned g = std::min<unsigned>(grid_for(nh, 256), kXdropBlocks);
                    k_xdrop<HitArraySrc, true><<<g, 256, 0, ctx->stream>>>(xa, src, po, nullptr, nullptr, ctr + 2);
                }
            } else {
                k_extend<ScorerI32><<<grid_for(nh, 256), 256, 0, ctx->stream>>>(
                    hits, nh, target->d, tlen, query->d, qlen, s32, sp->hsp.x_drop, sp->hsp.hsp_threshold,
                    pat.span, sp->entropy_enabled, sp->entropy_threshold, po, ctr + 2);
            }
            GXY_LAUNCH_CHECK(ctx);
        }
        unsigned long long c2[5];
        GXY_HIP(ctx, hipMemcpyAsync(c2, ctr + 1, 40, hipMemcpyDeviceToHost, ctx->stream));
        GXY_HIP(ctx, hipStreamSynchronize(ctx->stream));
        npass = c2[0];
        st.columns = c2[1] + c2[2];
        ctx->xd_first_cols += c2[2];
        ctx->xd_conts += c2[3];
        ctx->xd_first_pass += c2[4];
        clk.lap(kTrSearchExtend);
    }
    st.passing = npass;
    if (npass > 0) {
        auto *pkeys = (uint64_t *)ctx->slot[kSlotPassKeys];
        auto *phsp = (gxy_hsp *)ctx->slot[kSlotPassHsp];
        auto *k2 = (uint64_t *)scratch(ctx, kSlotPassKeys2, npass * 8);
        auto *i1 = (uint32_t *)scratch(ctx, kSlotPassIdx, npass * 4);
        auto *i2 = (uint32_t *)scratch(ctx, kSlotPassIdx2, npass * 4);
        auto *flag = (uint32_t *)scratch(ctx, kSlotFlags, npass * 4);
        auto *pos = (uint32_t *)scratch(ctx, kSlotEmit, (npass + 1) * 4);
        if (!k2 || !i1 || !i2 || !flag || !pos) return set_err(ctx, GXY_ENOMEM, "scratch allocation failed");
        {
            KernelTimer kt_(ctx, "pass_sort");
            k_iota<<<grid_for(npass, 256), 256, 0, ctx->stream>>>(i1, npass);
        }
        GXY_LAUNCH_CHECK(ctx);
        size_t tb = 0;
        (void)hipcub::DeviceRadixSort::SortPairs(nullptr, tb, pkeys, k2, i1, i2, (int)npass, 0, 64, ctx->stream);
        size_t tb2 = 0;
        (void)hipcub::DeviceScan::ExclusiveSum(nullptr, tb2, flag, pos, (int)npass + 1, ctx->stream);
        void *t2 = scratch(ctx, kSlotTmp, std::max(tb, tb2));
        if (!t2) return set_err(ctx, GXY_ENOMEM, "scratch allocation failed");
        {
            KernelTimer kt_(ctx, "pass_sort");
            GXY_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(t2, tb, pkeys, k2, i1, i2, (int)npass, 0, 64, ctx->stream));
        }
        {
            KernelTimer kt_(ctx, "admit");
            k_admit<<<grid_for(npass, 256), 256, 0, ctx->stream>>>(k2, i2, phsp, npass, qlen, flag);
        }
        GXY_LAUNCH_CHECK(ctx);
        {
            KernelTimer kt_(ctx, "admit");
            GXY_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(t2, tb2, flag, pos, (int)npass, ctx->stream));
        }
        uint32_t last_pos = 0, last_flag = 0;
        GXY_HIP(ctx, hipMemcpyAsync(&last_pos, pos + npass - 1, 4, hipMemcpyDeviceToHost, ctx->stream));
        GXY_HIP(ctx, hipMemcpyAsync(&last_flag, flag + npass - 1, 4, hipMemcpyDeviceToHost, ctx->stream));
        GXY_HIP(ctx, hipStreamSynchronize(ctx->stream));
        clk.lap(kTrSearchAdmit);
        const uint64_t nemit = (uint64_t)last_pos + last_flag;
        st.emitted = nemit;
        // emitted (diag, t_start) keys into pkeys/i1 (free now), sort into k2/i2
        (void)hipcub::DeviceRadixSort::SortPairs(nullptr, tb, pkeys, k2, i1, i2, (int)nemit, 0, 64, ctx->stream);
        t2 = scratch(ctx, kSlotTmp, tb);
        auto *outd = (gxy_hsp *)scratch(ctx, kSlotMisc1, nemit * sizeof(gxy_hsp));
        if (!t2 || !outd) return set_err(ctx, GXY_ENOMEM, "scratch allocation failed");
        {
            KernelTimer kt_(ctx, "emit");
            k_emit<<<grid_for(npass, 256), 256, 0, ctx->stream>>>(k2, i2, phsp, flag, pos, npass, pkeys, i1);
            GXY_LAUNCH_CHECK(ctx);
            GXY_HIP(ctx,
                    hipcub::DeviceRadixSort::SortPairs(t2, tb, pkeys, k2, i1, i2, (int)nemit, 0, 64, ctx->stream));
            k_gather_hsp<<<grid_for(nemit, 256), 256, 0, ctx->stream>>>(i2, phsp, nemit, outd);
            GXY_LAUNCH_CHECK(ctx);
        }
        result->resize(nemit);
        GXY_HIP(ctx,
                hipMemcpyAsync(result->data(), outd, nemit * sizeof(gxy_hsp), hipMemcpyDeviceToHost, ctx->stream));
        GXY_HIP(ctx, hipStreamSynchronize(ctx->stream));
        clk.lap(kTrSearchCopy);
        if (dedup) st.tie_pairs = dedup_hsps(result);
        clk.lap(kTrSearchDedup);
    }
    if (stats) *stats = st;
    return GXY_OK;
}

// Vec::dedup (src/seed_search.rs:123) on HSPs sorted by (diagonal, t_start),
// in place; returns the number of adjacent pairs with equal sort keys that
// survive (ties whose order the reference leaves to its unstable sort).
uint64_t dedup_hsps(std::vector<gxy_hsp> *v) {
    uint64_t ties = 0;
    size_t n = 0;
    for (size_t k = 0; k < v->size(); k++) {
        const gxy_hsp h = (*v)[k];
        if (n > 0) {
            const gxy_hsp &b = (*v)[n - 1];
            const bool same_key =
                (int64_t)h.t_start - h.q_start == (int64_t)b.t_start - b.q_start && h.t_start == b.t_start;
            if (same_key && h.length == b.length && h.score == b.score) continue;
            if (same_key) ties++;
        }
        (*v)[n++] = h;
    }
    v->resize(n);
    return ties;
}

}  // namespace gxy

namespace gxy {

bool extend_uses_bytes(const gxy_scoring &sc, const gxy_hsp_params &hp) {
    return scoring_byte_ok(sc) && hp.x_drop >= 0;
}

int extend_hits_dev(gxy_ctx *ctx, const Blk *t, uint64_t tlen, const Blk *q, uint64_t qlen, uint8_t *tbytes,
                    uint8_t *qbytes, const uint32_t *dtp, const uint32_t *dqp, uint64_t n, const gxy_scoring &sc,
                    const gxy_hsp_params &hp, uint32_t seed_len, gxy_hsp *dout, uint8_t *dfound,
                    unsigned long long *colsum) {
    if (n == 0) return GXY_OK;
    if (extend_uses_bytes(sc, hp)) {
        // persistent lanes over byte buffers (the search's x-drop kernel)
        int rc = blocks_to_bytes(ctx, t, tlen, true, tbytes);
        if (!rc) rc = blocks_to_bytes(ctx, q, qlen, false, qbytes);
        if (rc) return rc;
        XdropArgs xa{tbytes + kBytePad, qbytes + kBytePad, tlen, qlen, t, make_byte_lut(sc), hp.x_drop,
                     hp.hsp_threshold, seed_len, 0, 0};
        PairSrc src{dtp, dqp, n};
        PassOut none{nullptr, nullptr, nullptr};
        k_xdrop<PairSrc, false><<<std::min<unsigned>(grid_for(n, 256), kXdropBlocks), 256, 0, ctx->stream>>>(
            xa, src, none, dout, dfound, colsum);
    } else {
        k_extend_batch<ScorerI32><<<grid_for(n, 256), 256, 0, ctx->stream>>>(
            dtp, dqp, n, t, tlen, q, qlen, make_scorer_i32(sc), hp.x_drop, hp.hsp_threshold, seed_len, dout,
            dfound);
    }
    GXY_LAUNCH_CHECK(ctx);
    return GXY_OK;
}

}  // namespace gxy

extern "C" {

int gxy_search_dev(gxy_ctx *ctx, const gxy_postable *table, const gxy_dseq *query, const gxy_scoring *scoring,
                   const gxy_search_params *params, gxy_hsp **out, uint64_t *n_out, gxy_search_stats *stats) {
    return guarded(ctx, [&]() -> int {
        if (!ctx) return GXY_EINVAL;
        if (!out || !n_out) return set_err(ctx, GXY_EINVAL, "null output");
        GXY_HIP(ctx, hipSetDevice(ctx->device));
        std::vector<gxy_hsp> v;
        int rc = search_dev_impl(ctx, table, query, scoring, params, 1, &v, stats);
        if (rc) return rc;
        *n_out = v.size();
        *out = (gxy_hsp *)std::malloc(std::max<size_t>(1, v.size()) * sizeof(gxy_hsp));
        if (!*out) return set_err(ctx, GXY_ENOMEM, "host allocation failed");
        if (!v.empty()) std::memcpy(*out, v.data(), v.size() * sizeof(gxy_hsp));
        return GXY_OK;
    });
}

int gxy_search(gxy_ctx *ctx, const gxy_postable *table, const gxy_seq *target, const gxy_seq *query,
               const gxy_scoring *scoring, const gxy_search_params *params, gxy_hsp **out, uint64_t *n_out,
               gxy_search_stats *stats) {
    return guarded(ctx, [&]() -> int {
        if (!ctx) return GXY_EINVAL;
        if (!table || !target || !query) return set_err(ctx, GXY_EINVAL, "null argument");
        if (target->len != table->target->len)
            return set_err(ctx, GXY_EINVAL, "target is not the sequence the table was built from");
        GXY_HIP(ctx, hipSetDevice(ctx->device));
        gxy_dseq *dq;
        int rc = dseq_upload_impl(ctx, query, &dq);
        if (rc) return rc;
        rc = gxy_search_dev(ctx, table, dq, scoring, params, out, n_out, stats);
        gxy_dseq_free(dq);
        return rc;
    });
}

int gxy_extend_hit_batch(gxy_ctx *ctx, const gxy_seq *target, const gxy_seq *query, const gxy_scoring *scoring,
                         const gxy_hsp_params *params, const uint32_t *t_pos, const uint32_t *q_pos, uint64_t n,
                         uint32_t seed_len, gxy_hsp *out, uint8_t *found) {
    return guarded(ctx, [&]() -> int {
        if (!ctx) return GXY_EINVAL;
        if (!target || !query || !scoring || !params || (n && (!t_pos || !q_pos || !out || !found)))
            return set_err(ctx, GXY_EINVAL, "null argument");
        GXY_HIP(ctx, hipSetDevice(ctx->device));
        if (n == 0) return GXY_OK;
        gxy_dseq *dt = nullptr, *dq = nullptr;
        int rc = dseq_upload_impl(ctx, target, &dt);
        if (rc) return rc;
        rc = dseq_upload_impl(ctx, query, &dq);
        if (rc) {
            gxy_dseq_free(dt);
            return rc;
        }
        const bool bytes = extend_uses_bytes(*scoring, *params);
        auto *dtp = (uint32_t *)scratch(ctx, kSlotMisc0, n * 4);
        auto *dqp = (uint32_t *)scratch(ctx, kSlotMisc1, n * 4);
        auto *dout = (gxy_hsp *)scratch(ctx, kSlotMisc2, n * sizeof(gxy_hsp));
        auto *dfound = (uint8_t *)scratch(ctx, kSlotMisc3, n);
        auto *cs = (unsigned long long *)scratch(ctx, kSlotCounters, 64);
        auto *tb = bytes ? (uint8_t *)scratch(ctx, kSlotTBytes, byte_buffer_size(dt->len)) : nullptr;
        auto *qb = bytes ? (uint8_t *)scratch(ctx, kSlotQBytes, byte_buffer_size(dq->len)) : nullptr;
        if (!dtp || !dqp || !dout || !dfound || !cs || (bytes && (!tb || !qb))) {
            rc = set_err(ctx, GXY_ENOMEM, "scratch allocation failed");
        } else {
            hipError_t e = hipMemcpyAsync(dtp, t_pos, n * 4, hipMemcpyHostToDevice, ctx->stream);
            if (e == hipSuccess) e = hipMemcpyAsync(dqp, q_pos, n * 4, hipMemcpyHostToDevice, ctx->stream);
            if (e != hipSuccess) {
                rc = set_err(ctx, GXY_EDEVICE, "extend_hit_batch: %s", hipGetErrorString(e));
            } else {
                rc = extend_hits_dev(ctx, dt->d, dt->len, dq->d, dq->len, tb, qb, dtp, dqp, n, *scoring, *params,
                                     seed_len, dout, dfound, cs);
            }
            if (!rc) {
                e = hipMemcpyAsync(out, dout, n * sizeof(gxy_hsp), hipMemcpyDeviceToHost, ctx->stream);
                if (e == hipSuccess) e = hipMemcpyAsync(found, dfound, n, hipMemcpyDeviceToHost, ctx->stream);
                if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
                if (e != hipSuccess) rc = set_err(ctx, GXY_EDEVICE, "extend_hit_batch: %s", hipGetErrorString(e));
            }
        }
        gxy_dseq_free(dt);
        gxy_dseq_free(dq);
        return rc;
    });
}

int gxy_extend_hit(gxy_ctx *ctx, const gxy_seq *target, const gxy_seq *query, const gxy_scoring *scoring,
                   const gxy_hsp_params *params, uint32_t t_pos, uint32_t q_pos, uint32_t seed_len, gxy_hsp *out,
                   int *found) {
    return guarded(ctx, [&]() -> int {
        if (!ctx) return GXY_EINVAL;
        if (!out || !found) return set_err(ctx, GXY_EINVAL, "null output");
        uint8_t f = 0;
        int rc = gxy_extend_hit_batch(ctx, target, query, scoring, params, &t_pos, &q_pos, 1, seed_len, out, &f);
        *found = f;
        return rc;
    });
}

}  // extern "C"
