// postable.hip — device PosTable (src/pos_table.rs:24-137).
//
// Build = one pass that turns every window (0, step, 2*step, ... <= len-span)
// into a 64-bit item (index << 32 | pos) — `index` is the bit-permuted word of
// gather_index; invalid windows get the sentinel 4^w so they sort last — then
// a stable LSD radix sort on the index bits only.  Items enter in ascending position, so the stable sort leaves each
// word's positions ascending exactly as the reference's counting-sort scatter
// does (src/pos_table.rs:60-75).  `starts` is then filled by one thread per
// sorted item writing the run of words (key[i-1], key[i]] — every entry written
// exactly once, no atomics.
#include <hipcub/hipcub.hpp>

#include "internal.h"

using namespace gxy;

namespace {

// (word, position) of windows w0 .. w0 + nwin - 1 into items[0 .. nwin)
__global__ void k_table_items(const Blk *__restrict__ seq, uint64_t len, DevPattern pat,
                              uint64_t step, uint64_t w0, uint64_t nwin, int respect_mask,
                              uint64_t *__restrict__ items, unsigned long long *__restrict__ nvalid) {
    const uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool ok = false;
    if (w < nwin) {
        const uint64_t pos = (w0 + w) * step;
        const uint64_t sentinel = 1ull << (2 * pat.weight);
        uint64_t word = sentinel;
        // SeedExtractor::word_at (src/seeds.rs:133-147)
        if (pos + pat.span <= len) {
            const Win x = load_win(seq, pos);
            const uint32_t care = pat.care_mask;
            const uint32_t msk = respect_mask ? x.mask : 0u;
            if ((x.valid & care) == care && (msk & care) == 0) {
                word = gather_index(x.codes, pat);
                ok = true;
            }
        }
        items[w] = (word << 32) | (pos & 0xFFFFFFFFull);
    }
    // one atomic per wavefront
    const unsigned long long b = __ballot(ok);
    if ((threadIdx.x & 63) == 0 && b) spread_add(nvalid, (unsigned long long)__popcll(b));
}

__global__ void k_table_starts(const uint64_t *__restrict__ items, uint64_t nvalid, uint64_t num_words,
                               uint32_t *__restrict__ starts, uint32_t *__restrict__ positions) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i > nvalid) return;
    const uint64_t key = i < nvalid ? (items[i] >> 32) : num_words;
    const int64_t prev = i == 0 ? -1 : (int64_t)(items[i - 1] >> 32);
    for (int64_t w = prev + 1; w <= (int64_t)key; w++) starts[w] = (uint32_t)i;
    if (positions && i < nvalid) positions[i] = (uint32_t)items[i];
}

// Tables of more than 2^31 windows are sorted as two halves A (the first
// windows) and B; A's positions precede B's, so word w's bucket is A's part
// then B's: starts = startsA + startsB, A's item i of word w goes to
// startsB[w] + i and B's item j to startsA[w + 1] + j.
__global__ void k_starts_sum(const uint32_t *__restrict__ a, const uint32_t *__restrict__ b, uint64_t n,
                             uint32_t *__restrict__ out) {
    const uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w < n) out[w] = a[w] + b[w];
}
__global__ void k_place_half(const uint64_t *__restrict__ items, uint64_t nvalid, const uint32_t *__restrict__ other,
                             uint32_t shift, uint32_t *__restrict__ positions) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nvalid) return;
    const uint64_t it = items[i];
    positions[other[(it >> 32) + shift] + i] = (uint32_t)it;
}

// Bitmap of non-empty words: one thread per word (coalesced reads of
// `starts`); a wavefront's ballot is two bitmap words.
__global__ void k_nonempty(const uint32_t *__restrict__ starts, uint64_t num_words, uint32_t *__restrict__ bits) {
    const uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool ne = w < num_words && starts[w + 1] != starts[w];
    const uint64_t b = __ballot(ne);
    const int lane = threadIdx.x & 63;
    const uint64_t g = w >> 5;  // bitmap word of this lane
    if ((lane & 31) == 0 && g * 32 < num_words) bits[g] = (uint32_t)(lane ? b >> 32 : b);
}

// Count map: min(bucket size, 3) in 2 bits per word, 16 words per u32; one
// thread per word, a wavefront's two ballots are four map words.
__global__ void k_cmap(const uint32_t *__restrict__ starts, uint64_t num_words, uint32_t *__restrict__ cmap) {
    const uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t c = w < num_words ? min(starts[w + 1] - starts[w], 3u) : 0u;
    const uint64_t b0 = __ballot(c & 1u), b1 = __ballot(c >> 1);
    const int lane = threadIdx.x & 63;
    if ((lane & 15) == 0 && (w >> 4) * 16 < num_words) {
        const uint32_t lo = (uint32_t)(b0 >> lane) & 0xFFFFu, hi = (uint32_t)(b1 >> lane) & 0xFFFFu;
        // interleave: word k of the 16 takes bits 2k (lo) and 2k+1 (hi)
        uint32_t x = lo, y = hi;
        x = (x | (x << 8)) & 0x00FF00FFu;
        x = (x | (x << 4)) & 0x0F0F0F0Fu;
        x = (x | (x << 2)) & 0x33333333u;
        x = (x | (x << 1)) & 0x55555555u;
        y = (y | (y << 8)) & 0x00FF00FFu;
        y = (y | (y << 4)) & 0x0F0F0F0Fu;
        y = (y | (y << 2)) & 0x33333333u;
        y = (y | (y << 1)) & 0x55555555u;
        cmap[w >> 4] = x | (y << 1);
    }
}

// First position of every non-empty word, so a one-position bucket is read
// with one load.
__global__ void k_spos(const uint32_t *__restrict__ starts, const uint32_t *__restrict__ positions,
                       uint64_t num_words, uint32_t *__restrict__ spos) {
    const uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= num_words) return;
    const uint32_t b = starts[w];
    spos[w] = starts[w + 1] > b ? positions[b] : 0xFFFFFFFFu;
}

// Neighbourhood counts: variant 0 is the word, variant i >= 1 flips index bit
// W-i (a transition at care position i-1); each takes min(size, 3) from the
// count map.  Consecutive words read neighbouring map words, so the map stays
// in cache.
__global__ void k_ncnt(const uint32_t *__restrict__ cmap, uint64_t num_words, uint32_t weight,
                       uint32_t *__restrict__ ncnt) {
    const uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= num_words) return;
    uint32_t r = 0;
    for (uint32_t i = 0; i <= weight; i++) {
        const uint64_t v = i == 0 ? w : (w ^ (1ull << (weight - i)));
        const uint32_t c = (cmap[v >> 4] >> (2 * (uint32_t)(v & 15))) & 3u;
        r |= c << (2 * i);
    }
    ncnt[w] = r;
}

__global__ void k_prune_counts(const uint32_t *__restrict__ starts, uint64_t num_words, uint32_t limit,
                               uint32_t *__restrict__ keep_cnt) {
    const uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= num_words) return;
    const uint32_t c = starts[w + 1] - starts[w];
    keep_cnt[w] = c <= limit ? c : 0u;  // src/pos_table.rs:128
}

__global__ void k_prune_copy(const uint32_t *__restrict__ starts, const uint32_t *__restrict__ pos,
                             const uint32_t *__restrict__ new_starts, const uint32_t *__restrict__ keep_cnt,
                             uint64_t num_words, uint32_t *__restrict__ new_pos) {
    const uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= num_words) return;
    const uint32_t c = keep_cnt[w];
    const uint32_t s = starts[w], d = new_starts[w];
    for (uint32_t k = 0; k < c; k++) new_pos[d + k] = pos[s + k];
}

__global__ void k_set_last(uint32_t *__restrict__ new_starts, const uint32_t *__restrict__ keep_cnt,
                           uint64_t num_words) {
    new_starts[num_words] = new_starts[num_words - 1] + keep_cnt[num_words - 1];
}

inline unsigned grid_for(uint64_t n, unsigned th) { return (unsigned)((n + th - 1) / th); }

}  // namespace

int gxy::postable_build_impl(gxy_ctx *ctx, const gxy_dseq *t, const gxy_seed_pattern *pattern, uint64_t step,
                             int respect_mask, gxy_postable **out) {
    DevPattern pat;
    if (!make_dev_pattern(pattern, &pat)) return set_err(ctx, GXY_EINVAL, "invalid seed pattern");
    if (step < 1) return set_err(ctx, GXY_EINVAL, "step must be >= 1");  // src/pos_table.rs:36
    if (pat.weight > kMaxWeight)
        return set_err(ctx, GXY_ERANGE, "seed weight %u exceeds the dense-table limit %u", pat.weight,
                       kMaxWeight);
    if (t->len > 0xFFFFFFFFull)
        return set_err(ctx, GXY_ERANGE, "pos_table overflow (> 4 Gbp)");  // src/pos_table.rs:56
    const uint64_t max_start = t->len >= pat.span ? t->len - pat.span : 0;
    const uint64_t nwin = max_start / step + 1;  // < 2^32: len < 2^32
    // the radix sort takes at most 2^31 - 1 items: more windows sort as two halves
    const bool split = nwin > 0x7FFFFFFFull;
    const uint64_t nA = split ? nwin / 2 : nwin, nB = nwin - nA;
    auto *tab = new (std::nothrow) gxy_postable();
    if (!tab) return set_err(ctx, GXY_ENOMEM, "host allocation failed");
    tab->pattern = pat;
    tab->step = step;
    tab->num_words = 1ull << (2 * pat.weight);
    tab->target = t;

    auto *items = (uint64_t *)scratch(ctx, kSlotHits, nwin * sizeof(uint64_t));
    auto *items2 = (uint64_t *)scratch(ctx, kSlotPassKeys, nwin * sizeof(uint64_t));
    auto *cnt = (unsigned long long *)scratch(ctx, kSlotSpread, kSpreadBytes);
    auto *cntB = split ? (unsigned long long *)scratch(ctx, kSlotCounters, kSpreadBytes) : nullptr;
    if (!items || !items2 || !cnt || (split && !cntB)) {
        delete tab;
        return set_err(ctx, GXY_ENOMEM, "scratch allocation failed");
    }
    auto fail = [&](int code, const char *msg, hipError_t e) {
        gxy_postable_free(tab);
        return set_err(ctx, code, "%s: %s", msg, hipGetErrorString(e));
    };
    hipError_t e = hipMemsetAsync(cnt, 0, kSpreadBytes, ctx->stream);
    if (e == hipSuccess && split) e = hipMemsetAsync(cntB, 0, kSpreadBytes, ctx->stream);
    if (e != hipSuccess) return fail(GXY_EDEVICE, "memset", e);
    {
        KernelTimer kt_(ctx, "table_items");
        k_table_items<<<grid_for(nA, 256), 256, 0, ctx->stream>>>(t->d, t->len, pat, step, 0, nA, respect_mask,
                                                                  items, cnt);
        if (split)
            k_table_items<<<grid_for(nB, 256), 256, 0, ctx->stream>>>(t->d, t->len, pat, step, nA, nB, respect_mask,
                                                                      items + nA, cntB);
    }
    if ((e = hipGetLastError()) != hipSuccess) return fail(GXY_EDEVICE, "k_table_items", e);
    size_t tmp_bytes = 0;
    const int end_bit = 32 + 2 * (int)pat.weight + 1;
    (void)hipcub::DeviceRadixSort::SortKeys(nullptr, tmp_bytes, items, items2, (int)std::max(nA, nB), 32, end_bit,
                                            ctx->stream);
    void *tmp = scratch(ctx, kSlotTmp, tmp_bytes);
    if (!tmp) {
        gxy_postable_free(tab);
        return set_err(ctx, GXY_ENOMEM, "scratch allocation failed");
    }
    {
        KernelTimer kt_(ctx, "table_sort");
        e = hipcub::DeviceRadixSort::SortKeys(tmp, tmp_bytes, items, items2, (int)nA, 32, end_bit, ctx->stream);
        if (e == hipSuccess && split)
            e = hipcub::DeviceRadixSort::SortKeys(tmp, tmp_bytes, items + nA, items2 + nA, (int)nB, 32, end_bit,
                                                  ctx->stream);
    }
    if (e != hipSuccess) return fail(GXY_EDEVICE, "radix sort", e);
    uint64_t nvalid = 0, nvalidB = 0;
    if (spread_sum(ctx, cnt, &nvalid) != GXY_OK || (split && spread_sum(ctx, cntB, &nvalidB) != GXY_OK)) {
        gxy_postable_free(tab);
        return GXY_EDEVICE;
    }
    tab->npos = nvalid + nvalidB;
    tab->d_starts = (uint32_t *)dev_alloc(ctx, (tab->num_words + 1) * sizeof(uint32_t), &tab->pool);
    tab->d_positions = (uint32_t *)dev_alloc(ctx, std::max<uint64_t>(1, tab->npos) * sizeof(uint32_t), &tab->pool);
    if (!tab->d_starts || !tab->d_positions) {
        gxy_postable_free(tab);
        return set_err(ctx, GXY_ENOMEM, "table allocation failed");
    }
    if (!split) {
        KernelTimer kt_(ctx, "table_starts");
        k_table_starts<<<grid_for(nvalid + 1, 256), 256, 0, ctx->stream>>>(items2, nvalid, tab->num_words,
                                                                            tab->d_starts, tab->d_positions);
    } else {
        KernelTimer kt_(ctx, "table_starts");
        const uint64_t ns = tab->num_words + 1;
        auto *sA = (uint32_t *)scratch(ctx, kSlotCounts, ns * sizeof(uint32_t));
        auto *sB = (uint32_t *)scratch(ctx, kSlotOffsets, ns * sizeof(uint32_t));
        if (!sA || !sB) {
            gxy_postable_free(tab);
            return set_err(ctx, GXY_ENOMEM, "scratch allocation failed");
        }
        k_table_starts<<<grid_for(nvalid + 1, 256), 256, 0, ctx->stream>>>(items2, nvalid, tab->num_words, sA,
                                                                            nullptr);
        k_table_starts<<<grid_for(nvalidB + 1, 256), 256, 0, ctx->stream>>>(items2 + nA, nvalidB, tab->num_words,
                                                                             sB, nullptr);
        k_starts_sum<<<grid_for(ns, 256), 256, 0, ctx->stream>>>(sA, sB, ns, tab->d_starts);
        if (nvalid)
            k_place_half<<<grid_for(nvalid, 256), 256, 0, ctx->stream>>>(items2, nvalid, sB, 0, tab->d_positions);
        if (nvalidB)
            k_place_half<<<grid_for(nvalidB, 256), 256, 0, ctx->stream>>>(items2 + nA, nvalidB, sA, 1,
                                                                           tab->d_positions);
    }
    if ((e = hipGetLastError()) != hipSuccess) return fail(GXY_EDEVICE, "k_table_starts", e);
    const uint64_t nbw = (tab->num_words + 31) / 32;
    const uint64_t nbytes = byte_buffer_size(t->len);
    tab->d_nonempty = (uint32_t *)dev_alloc(ctx, nbw * 4, &tab->pool);
    tab->d_cmap = (uint32_t *)dev_alloc(ctx, ((tab->num_words + 15) / 16) * 4, &tab->pool);
    tab->d_spos = (uint32_t *)dev_alloc(ctx, tab->num_words * 4, &tab->pool);
    if (pat.weight <= kNcntMaxWeight) tab->d_ncnt = (uint32_t *)dev_alloc(ctx, tab->num_words * 4, &tab->pool);
    tab->d_tbytes = (uint8_t *)dev_alloc(ctx, nbytes, &tab->pool);
    if (!tab->d_nonempty || !tab->d_cmap || !tab->d_spos || !tab->d_tbytes ||
        (pat.weight <= kNcntMaxWeight && !tab->d_ncnt)) {
        gxy_postable_free(tab);
        return set_err(ctx, GXY_ENOMEM, "table allocation failed");
    }
    int rc;
    {
        KernelTimer kt_(ctx, "table_bytes");
        k_nonempty<<<grid_for(tab->num_words, 256), 256, 0, ctx->stream>>>(tab->d_starts, tab->num_words,
                                                                            tab->d_nonempty);
        if ((e = hipGetLastError()) != hipSuccess) return fail(GXY_EDEVICE, "k_nonempty", e);
        k_cmap<<<grid_for(tab->num_words, 256), 256, 0, ctx->stream>>>(tab->d_starts, tab->num_words, tab->d_cmap);
        if ((e = hipGetLastError()) != hipSuccess) return fail(GXY_EDEVICE, "k_cmap", e);
        k_spos<<<grid_for(tab->num_words, 256), 256, 0, ctx->stream>>>(tab->d_starts, tab->d_positions,
                                                                        tab->num_words, tab->d_spos);
        if ((e = hipGetLastError()) != hipSuccess) return fail(GXY_EDEVICE, "k_spos", e);
        if (tab->d_ncnt) {
            k_ncnt<<<grid_for(tab->num_words, 256), 256, 0, ctx->stream>>>(tab->d_cmap, tab->num_words, pat.weight,
                                                                            tab->d_ncnt);
            if ((e = hipGetLastError()) != hipSuccess) return fail(GXY_EDEVICE, "k_ncnt", e);
        }
        rc = seq_to_bytes(ctx, t, true, tab->d_tbytes);
    }
    if (rc) {
        gxy_postable_free(tab);
        return rc;
    }
    if ((e = hipStreamSynchronize(ctx->stream)) != hipSuccess) return fail(GXY_EDEVICE, "build", e);
    *out = tab;
    return GXY_OK;
}

extern "C" {

int gxy_postable_build_dev(gxy_ctx *ctx, const gxy_dseq *target, const gxy_seed_pattern *pattern,
                           uint64_t step, gxy_postable **out) {
    return guarded(ctx, [&]() -> int {
        if (!ctx) return GXY_EINVAL;
        if (!target || !out) return set_err(ctx, GXY_EINVAL, "null argument");
        GXY_HIP(ctx, hipSetDevice(ctx->device));
        return postable_build_impl(ctx, target, pattern, step, 1, out);
    });
}

int gxy_postable_build(gxy_ctx *ctx, const gxy_seq *target, const gxy_seed_pattern *pattern,
                       uint64_t step, gxy_postable **out) {
    return guarded(ctx, [&]() -> int {
        if (!ctx) return GXY_EINVAL;
        if (!target || !out) return set_err(ctx, GXY_EINVAL, "null argument");
        GXY_HIP(ctx, hipSetDevice(ctx->device));
        gxy_dseq *d;
        int rc = dseq_upload_impl(ctx, target, &d);
        if (rc) return rc;
        rc = postable_build_impl(ctx, d, pattern, step, 1, out);
        if (rc) {
            gxy_dseq_free(d);
            return rc;
        }
        (*out)->owned_target = d;
        (*out)->host_codes = target->codes;
        (*out)->host_len = target->len;
        return GXY_OK;
    });
}

int gxy_postable_prune(gxy_ctx *ctx, gxy_postable *tab, uint32_t limit) {
    return guarded(ctx, [&]() -> int {
        // PosTable::prune_hot_words (src/pos_table.rs:113-136)
        if (!ctx) return GXY_EINVAL;
        if (!tab) return set_err(ctx, GXY_EINVAL, "null table");
        const uint64_t nw = tab->num_words;
        auto *keep = (uint32_t *)scratch(ctx, kSlotCounts, nw * sizeof(uint32_t));
        uint32_t *ns = nullptr, *np = nullptr;
        if (!keep) return set_err(ctx, GXY_ENOMEM, "scratch allocation failed");
        ns = (uint32_t *)dev_alloc_from(tab->pool, (nw + 1) * sizeof(uint32_t));
        if (!ns) return set_err(ctx, GXY_ENOMEM, "table allocation failed");
        k_prune_counts<<<grid_for(nw, 256), 256, 0, ctx->stream>>>(tab->d_starts, nw, limit, keep);
        GXY_LAUNCH_CHECK(ctx);
        size_t tmp_bytes = 0;
        (void)hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, keep, ns, (int)nw, ctx->stream);
        void *tmp = scratch(ctx, kSlotTmp, tmp_bytes);
        if (!tmp) {
            dev_free(tab->pool, ns);
            return set_err(ctx, GXY_ENOMEM, "scratch allocation failed");
        }
        GXY_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, keep, ns, (int)nw, ctx->stream));
        k_set_last<<<1, 1, 0, ctx->stream>>>(ns, keep, nw);
        GXY_LAUNCH_CHECK(ctx);
        uint32_t total = 0;
        GXY_HIP(ctx, hipMemcpyAsync(&total, ns + nw, 4, hipMemcpyDeviceToHost, ctx->stream));
        GXY_HIP(ctx, hipStreamSynchronize(ctx->stream));
        np = (uint32_t *)dev_alloc_from(tab->pool, std::max<uint32_t>(1, total) * sizeof(uint32_t));
        if (!np) {
            dev_free(tab->pool, ns);
            return set_err(ctx, GXY_ENOMEM, "table allocation failed");
        }
        k_prune_copy<<<grid_for(nw, 256), 256, 0, ctx->stream>>>(tab->d_starts, tab->d_positions, ns, keep,
                                                                 nw, np);
        GXY_LAUNCH_CHECK(ctx);
        GXY_HIP(ctx, hipStreamSynchronize(ctx->stream));
        dev_free(tab->pool, tab->d_starts);
        dev_free(tab->pool, tab->d_positions);
        tab->d_starts = ns;
        tab->d_positions = np;
        tab->npos = total;
        k_nonempty<<<grid_for(nw, 256), 256, 0, ctx->stream>>>(tab->d_starts, nw, tab->d_nonempty);
        GXY_LAUNCH_CHECK(ctx);
        k_cmap<<<grid_for(nw, 256), 256, 0, ctx->stream>>>(tab->d_starts, nw, tab->d_cmap);
        GXY_LAUNCH_CHECK(ctx);
        k_spos<<<grid_for(nw, 256), 256, 0, ctx->stream>>>(tab->d_starts, tab->d_positions, nw, tab->d_spos);
        GXY_LAUNCH_CHECK(ctx);
        if (tab->d_ncnt) {
            k_ncnt<<<grid_for(nw, 256), 256, 0, ctx->stream>>>(tab->d_cmap, nw, tab->pattern.weight, tab->d_ncnt);
            GXY_LAUNCH_CHECK(ctx);
        }
        GXY_HIP(ctx, hipStreamSynchronize(ctx->stream));
        return GXY_OK;
    });
}

int gxy_postable_lookup(gxy_ctx *ctx, const gxy_postable *tab, uint64_t word, uint32_t *out, uint64_t cap,
                        uint64_t *n_out) {
    return guarded(ctx, [&]() -> int {
        // PosTable::lookup (src/pos_table.rs:102-108)
        if (!ctx) return GXY_EINVAL;
        if (!tab || !n_out || (cap && !out)) return set_err(ctx, GXY_EINVAL, "null argument");
        if (word >= tab->num_words) return set_err(ctx, GXY_EINVAL, "word out of range");
        uint32_t se[2];
        const uint64_t idx = word_to_index(word, tab->pattern.weight);
        GXY_HIP(ctx, hipMemcpyAsync(se, tab->d_starts + idx, 8, hipMemcpyDeviceToHost, ctx->stream));
        GXY_HIP(ctx, hipStreamSynchronize(ctx->stream));
        const uint64_t n = se[1] - se[0];
        *n_out = n;
        const uint64_t c = std::min(n, cap);
        if (c) {
            GXY_HIP(ctx, hipMemcpyAsync(out, tab->d_positions + se[0], c * 4, hipMemcpyDeviceToHost, ctx->stream));
            GXY_HIP(ctx, hipStreamSynchronize(ctx->stream));
        }
        return GXY_OK;
    });
}

uint64_t gxy_postable_total_positions(const gxy_postable *tab) { return tab ? tab->npos : 0; }

int gxy_postable_export(gxy_ctx *ctx, const gxy_postable *tab, uint32_t *starts, uint64_t starts_cap,
                        uint32_t *positions, uint64_t positions_cap) {
    return guarded(ctx, [&]() -> int {
        // The device buckets are in index order (gather_index); the reference
        // layout orders them by word.  Rebuild it on the host.
        if (!ctx) return GXY_EINVAL;
        if (!tab) return set_err(ctx, GXY_EINVAL, "null table");
        const uint64_t nw = tab->num_words;
        std::vector<uint32_t> ds(nw + 1), dp(positions ? std::max<uint64_t>(1, tab->npos) : 0);
        GXY_HIP(ctx, hipMemcpyAsync(ds.data(), tab->d_starts, (nw + 1) * 4, hipMemcpyDeviceToHost, ctx->stream));
        if (tab->npos && positions)
            GXY_HIP(ctx, hipMemcpyAsync(dp.data(), tab->d_positions, tab->npos * 4, hipMemcpyDeviceToHost, ctx->stream));
        GXY_HIP(ctx, hipStreamSynchronize(ctx->stream));
        uint64_t acc = 0;
        for (uint64_t w = 0; w < nw; w++) {
            const uint64_t idx = word_to_index(w, tab->pattern.weight);
            const uint32_t b = ds[idx], e = ds[idx + 1];
            if (starts && w < starts_cap) starts[w] = (uint32_t)acc;
            if (positions) {
                for (uint32_t k = b; k < e; k++, acc++)
                    if (acc < positions_cap) positions[acc] = dp[k];
            } else {
                acc += e - b;
            }
        }
        if (starts && nw < starts_cap) starts[nw] = (uint32_t)acc;
        return GXY_OK;
    });
}

void gxy_postable_free(gxy_postable *tab) {
    if (!tab) return;
    dev_free(tab->pool, tab->d_starts);
    dev_free(tab->pool, tab->d_positions);
    dev_free(tab->pool, tab->d_nonempty);
    dev_free(tab->pool, tab->d_cmap);
    dev_free(tab->pool, tab->d_spos);
    dev_free(tab->pool, tab->d_ncnt);
    dev_free(tab->pool, tab->d_tbytes);
    if (tab->owned_target) gxy_dseq_free(tab->owned_target);
    delete tab;
}

}  // extern "C"
