// internal.h — entry points shared between the library's translation units
// (not part of the C-ABI).
#pragma once

#include <vector>

#include "common.h"

namespace gxy {

// PosTable::build (src/pos_table.rs:35-83) on a device sequence; with
// respect_mask == 0 soft-masked bases seed like unmasked ones (--nomasking,
// src/main.rs:26-30).
int postable_build_impl(gxy_ctx *ctx, const gxy_dseq *t, const gxy_seed_pattern *pattern, uint64_t step,
                        int respect_mask, gxy_postable **out);

// Byte-per-base copy of `s` (x-drop layout with target or query byte codes,
// see common.h) into a buffer of byte_buffer_size(s->len) bytes; the sequence
// starts at buf + kBytePad.
// A device sequence of `len` bases (blocks uninitialized), owned by the
// caller (gxy_dseq_free); from the block cache while ctx->pool_on.
int dseq_alloc(gxy_ctx *ctx, uint64_t len, gxy_dseq **out);
int seq_to_bytes(gxy_ctx *ctx, const gxy_dseq *s, bool as_target, uint8_t *buf);
// The same from raw blocks of a sequence of `len` bases.
int blocks_to_bytes(gxy_ctx *ctx, const Blk *b, uint64_t len, bool as_target, uint8_t *buf);
// Blocks (nblocks(len) of them, the last one zero) from the reference's
// PackedSeq byte arrays already in device memory (src/sequences.rs:128-141).
int blocks_from_arrays(gxy_ctx *ctx, const uint8_t *codes, const uint8_t *valid, const uint8_t *mask,
                       uint64_t len, Blk *out);

// extend_hit (src/hsp.rs:66-211) over n hits whose positions, results and
// buffers are all in device memory the caller provides: t/q blocks of the
// pair, byte buffers of byte_buffer_size(len) for each (filled here; used
// when extend_uses_bytes), an 8-byte column counter.  Allocates nothing;
// launches on ctx->stream without waiting.
bool extend_uses_bytes(const gxy_scoring &sc, const gxy_hsp_params &hp);
int extend_hits_dev(gxy_ctx *ctx, const Blk *t, uint64_t tlen, const Blk *q, uint64_t qlen, uint8_t *tbytes,
                    uint8_t *qbytes, const uint32_t *dtp, const uint32_t *dqp, uint64_t n, const gxy_scoring &sc,
                    const gxy_hsp_params &hp, uint32_t seed_len, gxy_hsp *dout, uint8_t *dfound,
                    unsigned long long *colsum);

// Seed search on device sequences.  `qbytes` may hold the query's byte copy
// (seq_to_bytes buffer + kBytePad) when the caller searches the same query
// repeatedly; otherwise it is built per call.  With dedup false the result is
// the sorted emitted list before Vec::dedup (the caller runs dedup_hsps and
// counts the ties).
int search_dev_impl(gxy_ctx *ctx, const gxy_postable *tab, const gxy_dseq *query, const gxy_scoring *scm,
                    const gxy_search_params *sp, int respect_mask, std::vector<gxy_hsp> *result,
                    gxy_search_stats *stats, const uint8_t *qbytes = nullptr, bool dedup = true);

// Vec::dedup of a search's emitted HSPs (in place); returns the tie count.
uint64_t dedup_hsps(std::vector<gxy_hsp> *v);

constexpr uint64_t kWholeSeq = ~0ull;
struct GappedItem {
    const gxy_dseq *t;
    const gxy_dseq *q;
    gxy_anchor a;  // relative to the slices below
    // the pair is t[t_off, t_off + t_len) against q[q_off, q_off + q_len)
    // (kWholeSeq: to the end of the sequence)
    uint64_t t_off = 0, t_len = kWholeSeq, q_off = 0, q_len = kWholeSeq;
};

struct GappedResult {
    bool found;
    gxy_alignment aln;  // runs_offset indexes the call's run vector
};

int gapped_batch_dev(gxy_ctx *ctx, const std::vector<GappedItem> &items, const gxy_scoring &sc,
                     const gxy_gapped_params &gp, std::vector<GappedResult> *results,
                     std::vector<gxy_edit_run> *runs, uint64_t *cells_out);

// choose_anchor over n HSPs, HSP k on the pair (ts[k], qs[k]); one launch.
int choose_anchor_pairs_dev(gxy_ctx *ctx, const gxy_dseq *const *ts, const gxy_dseq *const *qs,
                            const gxy_hsp *hsps_host, uint64_t n, const gxy_scoring &sc, uint32_t window,
                            gxy_anchor *out_host);
int choose_anchor_dev(gxy_ctx *ctx, const gxy_dseq *t, const gxy_dseq *q, const gxy_hsp *hsps_host, uint64_t n,
                      const gxy_scoring &sc, uint32_t window, gxy_anchor *out_host);

}  // namespace gxy
