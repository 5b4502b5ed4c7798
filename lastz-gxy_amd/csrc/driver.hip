// driver.hip — host driver over the device hot path: driver::run
// (src/driver.rs:199-392) and the MAF/PAF writers (src/output/maf.rs:33-74,
// src/output/paf.rs:23-89).
//
// All sequences are uploaded once and stay resident; chunks are device slices
// (src/driver.rs:156-194), the minus strand is a device reverse complement.
// Per (chunk, query, strand) unit: device search -> primary-region filter ->
// either ungapped records or the gapped stage.
//
// Gapped stage.  The reference extends HSPs one by one in descending score
// order and skips an HSP whose seed lies in the rectangle of an alignment
// already emitted (src/driver.rs:269-338), a sequential dependency.  Every
// extension's result depends only on its own HSP, so the exact batched form
// is speculative: gather the next K HSPs (score order) that no emitted record
// covers, extend them all in one device batch, then resolve in order on the
// host — skip if covered by what has been emitted so far, else emit the
// precomputed result — until an HSP without a result is reached.  K doubles
// each round, so a unit with few surviving HSPs costs few DPs and one with
// many costs few rounds; an HSP within a few thousand diagonals of one already
// in the batch waits for that extension, since a long alignment covers most
// HSPs along it.  All units' batches go to the device together.
//
// Canonical order where the reference's is unspecified: equal scores keep
// (diagonal, t_start) order (sort_unstable_by_key(Reverse(score)),
// src/driver.rs:270) and groups of one (target, query, strand) key are merged
// in ascending chunk order (rayon push order, src/driver.rs:361-372).
#include <algorithm>
#include <atomic>
#include <charconv>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>

#include "internal.h"

using namespace gxy;

namespace {

// Sort key of a record within its group: (t_start, q_start) and
// (t_span, q_span) packed, with the record's index.
struct SortKey {
    uint64_t k1, k2;
    uint32_t i;
};

// Stable LSD radix sort by (k1, k2) — sort_by_key's order, equal keys kept in
// input order (src/driver.rs:376).  11-bit digits, least significant first;
// digits that are equal across the whole input are skipped.
void radix_sort_keys(std::vector<SortKey> &a, std::vector<SortKey> &tmp) {
    const size_t n = a.size();
    if (n < 2) return;
    tmp.resize(n);
    constexpr int kBits = 11, kBuckets = 1 << kBits;
    std::vector<uint32_t> cnt(kBuckets);
    for (int word = 1; word >= 0; word--) {
        for (int shift = 0; shift < 64; shift += kBits) {
            auto digit = [&](const SortKey &x) {
                return (uint32_t)(((word ? x.k2 : x.k1) >> shift) & (kBuckets - 1));
            };
            std::fill(cnt.begin(), cnt.end(), 0u);
            for (size_t k = 0; k < n; k++) cnt[digit(a[k])]++;
            if (cnt[digit(a[0])] == n) continue;  // constant digit
            uint32_t sum = 0;
            for (int b = 0; b < kBuckets; b++) {
                const uint32_t c = cnt[b];
                cnt[b] = sum;
                sum += c;
            }
            for (size_t k = 0; k < n; k++) tmp[cnt[digit(a[k])]++] = a[k];
            a.swap(tmp);
        }
    }
}

// A record (src/output/mod.rs:17-31) as a plain value: an ungapped record's
// script is one implicit Match run of t_span (n_runs == 0); a gapped record's
// runs live in the run's shared run store.
struct Rec {
    uint32_t ti, qi, strand;
    uint32_t t_start, q_start, t_span, q_span;
    int32_t score;
    uint32_t n_runs;
    uint64_t runs_offset;
};

// A chunk x query x strand group of records (one entry of src/driver.rs:359-366).
struct Group {
    uint32_t ti, qi, strand;
    size_t chunk;
    std::vector<Rec> recs;
};

// Sorts a group by (t_start, q_start, t_span, q_span) and drops duplicates
// (src/driver.rs:376-387).
void finalize_group(Group &g) {
    std::vector<SortKey> keys(g.recs.size()), tmp;
    for (size_t k = 0; k < g.recs.size(); k++) {
        const Rec &r = g.recs[k];
        keys[k] = {(uint64_t)r.t_start << 32 | r.q_start, (uint64_t)r.t_span << 32 | r.q_span, (uint32_t)k};
    }
    radix_sort_keys(keys, tmp);
    std::vector<Rec> out;
    out.reserve(keys.size());
    for (const SortKey &sk : keys) {
        const Rec &r = g.recs[sk.i];
        if (!out.empty()) {
            const Rec &b = out.back();
            if (r.t_start == b.t_start && r.q_start == b.q_start && r.t_span == b.t_span &&
                r.q_span == b.q_span && r.strand == b.strand)
                continue;
        }
        out.push_back(r);
    }
    g.recs.swap(out);
}

// Host jobs (record building and group sorting) run on worker threads while
// the calling thread keeps the device busy with the next search.
struct DevSeqDeleter {
    void operator()(gxy_dseq *s) const { gxy_dseq_free(s); }
};
struct TableDeleter {
    void operator()(gxy_postable *t) const { gxy_postable_free(t); }
};
using DSeq = std::unique_ptr<gxy_dseq, DevSeqDeleter>;
using Table = std::unique_ptr<gxy_postable, TableDeleter>;

struct Chunk {
    uint32_t parent;
    uint32_t global_start;
    const gxy_dseq *seq;   // the whole target or an owned slice
    DSeq owned;
    uint32_t primary_start_local, primary_end_local;
};

// chunk_target (src/driver.rs:156-194): targets longer than primary + halo
// become device slices [ps - H, pe + H) with primary region [ps, pe).
int build_chunks(gxy_ctx *ctx, const gxy_dseq *const *targets, uint64_t n_targets, const gxy_config *cfg,
                 std::vector<Chunk> *chunks) {
    for (uint64_t ti = 0; ti < n_targets; ti++) {
        const uint32_t t_len = (uint32_t)targets[ti]->len;
        const uint32_t P = cfg->chunk_primary_bp, H = cfg->chunk_halo_bp;
        if (P == 0 || (uint64_t)t_len <= (uint64_t)P + H) {
            chunks->push_back({(uint32_t)ti, 0, targets[ti], nullptr, 0, t_len});
            continue;
        }
        uint32_t ps = 0;
        while (ps < t_len) {
            const uint32_t pe = (uint32_t)std::min<uint64_t>((uint64_t)ps + P, t_len);
            const uint32_t lo = ps >= H ? ps - H : 0;
            const uint32_t hi = (uint32_t)std::min<uint64_t>((uint64_t)pe + H, t_len);
            gxy_dseq *s;
            int rc = gxy_dseq_slice(ctx, targets[ti], lo, hi, &s);
            if (rc) return rc;
            Chunk c{(uint32_t)ti, lo, s, DSeq(s), ps - lo, pe - lo};
            chunks->push_back(std::move(c));
            if (pe == t_len) break;
            ps = pe;
        }
    }
    return GXY_OK;
}

// Group merge (src/driver.rs:371-389) into a record set: groups ordered by
// (target, query, strand, chunk), records in group order.
int assemble_records(gxy_ctx *ctx, std::vector<std::unique_ptr<Group>> &groups,
                     const std::vector<gxy_edit_run> &run_store, gxy_record_set **out) {
    std::stable_sort(groups.begin(), groups.end(),
                     [](const std::unique_ptr<Group> &a, const std::unique_ptr<Group> &b) {
                         if (a->ti != b->ti) return a->ti < b->ti;
                         if (a->qi != b->qi) return a->qi < b->qi;
                         if (a->strand != b->strand) return a->strand < b->strand;
                         return a->chunk < b->chunk;
                     });
    size_t total = 0, runs_ub = 0;
    for (const auto &g : groups) {
        total += g->recs.size();
        for (const auto &r : g->recs) runs_ub += r.n_runs ? r.n_runs : 1;
    }
    auto *set = (gxy_record_set *)std::calloc(1, sizeof(gxy_record_set));
    if (!set) return set_err(ctx, GXY_ENOMEM, "host allocation failed");
    set->records = (gxy_record *)std::malloc(std::max<size_t>(1, total) * sizeof(gxy_record));
    set->runs = (gxy_edit_run *)std::malloc(std::max<size_t>(1, runs_ub) * sizeof(gxy_edit_run));
    if (!set->records || !set->runs) {
        gxy_record_set_free(set);
        return set_err(ctx, GXY_ENOMEM, "host allocation failed");
    }
    size_t nrec = 0, ro = 0;
    for (const auto &g : groups) {
        for (const Rec &r : g->recs) {
            gxy_record &o = set->records[nrec++];
            o.target_index = r.ti;
            o.query_index = r.qi;
            o.strand = r.strand;
            o.t_start = r.t_start;
            o.q_start = r.q_start;
            o.t_span = r.t_span;
            o.q_span = r.q_span;
            o.score = r.score;
            o.reserved = 0;
            o.runs_offset = ro;
            if (r.n_runs == 0) {  // build_ungapped_record: one Match run (src/driver.rs:451-452)
                o.n_runs = 1;
                set->runs[ro++] = {GXY_OP_MATCH, r.t_span};
            } else {
                o.n_runs = r.n_runs;
                for (uint32_t x = 0; x < r.n_runs; x++) set->runs[ro++] = run_store[r.runs_offset + x];
            }
        }
    }
    set->n = nrec;
    set->n_runs = ro;
    *out = set;
    return GXY_OK;
}

// Gapped work of one (chunk, query, strand) group (src/driver.rs:245-300):
// its primary-region HSPs in score order are taken in turn, an HSP whose seed
// lies in an earlier record's rectangle is skipped, the rest are anchored and
// gapped-extended.  The order decides which HSPs are skipped, so records are
// emitted strictly in that order; extensions run ahead speculatively in
// batches that double (8 .. 1024), skipping HSPs near the diagonal of one
// already in the batch, and the batches of all groups go to the device
// together.
// best_chain (src/chain.rs:19-69): the highest-scoring chain of HSPs strictly
// increasing on both axes (h_j ends before h_i starts on t and on q).  The
// reference's O(n^2) scan keeps, for each HSP in (t_start, q_start) order, the
// first earlier HSP with the strictly highest chain score; the same choice is a
// prefix-maximum query — the maximum of (chain score, -order) over HSPs whose
// t_end <= t_start(i), keyed by q_end <= q_start(i) — answered here by a
// Fenwick tree over q_end while HSPs are inserted in t_end order: O(n log n).
// Zero-length HSPs (t_end == t_start) fall back to the scan, which alone
// restricts predecessors to earlier positions in the order.
std::vector<gxy_hsp> best_chain(const std::vector<gxy_hsp> &hsps) {
    const size_t n = hsps.size();
    if (n == 0) return {};
    std::vector<gxy_hsp> h = hsps;
    std::stable_sort(h.begin(), h.end(), [](const gxy_hsp &a, const gxy_hsp &b) {
        return a.t_start != b.t_start ? a.t_start < b.t_start : a.q_start < b.q_start;
    });
    std::vector<int64_t> best(n), parent(n, -1);
    bool zero = false;
    for (const auto &x : h) zero |= x.length == 0;
    if (zero || n < 64) {
        for (size_t i = 0; i < n; i++) {
            int64_t bps = 0, bpi = -1;
            for (size_t j = 0; j < i; j++)
                if ((uint64_t)h[j].t_start + h[j].length <= h[i].t_start &&
                    (uint64_t)h[j].q_start + h[j].length <= h[i].q_start && best[j] > bps) {
                    bps = best[j];
                    bpi = (int64_t)j;
                }
            best[i] = bps + h[i].score;
            parent[i] = bpi;
        }
    } else {
        // q_end ranks for the Fenwick tree
        std::vector<uint64_t> qe(n);
        for (size_t i = 0; i < n; i++) qe[i] = (uint64_t)h[i].q_start + h[i].length;
        std::vector<uint64_t> keys = qe;
        std::sort(keys.begin(), keys.end());
        keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
        const size_t K = keys.size();
        // node value: (chain score, -order) as a pair compared lexicographically
        std::vector<std::pair<int64_t, int64_t>> fw(K + 1, {INT64_MIN, 0});
        auto upd = [&](size_t pos, std::pair<int64_t, int64_t> v) {
            for (size_t x = pos + 1; x <= K; x += x & (0 - x)) fw[x] = std::max(fw[x], v);
        };
        auto qry = [&](size_t cnt) {  // max over the first cnt keys
            std::pair<int64_t, int64_t> r{INT64_MIN, 0};
            for (size_t x = cnt; x > 0; x -= x & (0 - x)) r = std::max(r, fw[x]);
            return r;
        };
        std::vector<size_t> by_end(n);
        for (size_t i = 0; i < n; i++) by_end[i] = i;
        std::stable_sort(by_end.begin(), by_end.end(), [&](size_t a, size_t b) {
            return (uint64_t)h[a].t_start + h[a].length < (uint64_t)h[b].t_start + h[b].length;
        });
        size_t ins = 0;
        for (size_t i = 0; i < n; i++) {
            // every HSP ending at or before t_start(i) precedes i in the order
            while (ins < n && (uint64_t)h[by_end[ins]].t_start + h[by_end[ins]].length <= h[i].t_start) {
                const size_t j = by_end[ins++];
                if (best[j] > 0)  // only strictly positive chains ever beat the start value 0
                    upd((size_t)(std::lower_bound(keys.begin(), keys.end(), qe[j]) - keys.begin()),
                        {best[j], -(int64_t)j});
            }
            const size_t cnt = (size_t)(std::upper_bound(keys.begin(), keys.end(), (uint64_t)h[i].q_start) -
                                        keys.begin());
            const auto r = qry(cnt);
            const bool has = r.first != INT64_MIN;
            best[i] = (has ? r.first : 0) + h[i].score;
            parent[i] = has ? -r.second : -1;
        }
    }
    size_t end = 0;
    for (size_t i = 1; i < n; i++)
        if (best[i] > best[end]) end = i;
    std::vector<gxy_hsp> out;
    for (int64_t c = (int64_t)end; c >= 0; c = parent[(size_t)c]) out.push_back(h[(size_t)c]);
    std::reverse(out.begin(), out.end());
    return out;
}

// The tweener (src/tweener.rs:59-185) for many (target, query, strand) groups:
// every qualifying gap between neighbouring records of every group is searched
// again with the inner pattern (its own small table, step 1, relaxed
// thresholds), its HSPs chained when chaining is on, and all of them — no
// covered skip here — anchored and gapped-extended in one device batch.  Each
// group's records are then stably sorted by (t_start, q_start) and
// deduplicated.  `target` is the whole target, `query` the group's strand.
struct TweenGroup {
    Group *g;
    const gxy_dseq *target;
    const gxy_dseq *query;
};

int tweener_groups_dev(gxy_ctx *ctx, std::vector<TweenGroup> &groups, const gxy_config *cfg,
                       std::vector<gxy_edit_run> *run_store, gxy_run_stats *rs) {
    const int32_t relax = std::max(cfg->inner_relax, 1);
    gxy_search_params sp;
    sp.step = 1;
    sp.hsp.x_drop = cfg->hsp.x_drop;
    sp.hsp.hsp_threshold = std::max(cfg->hsp.hsp_threshold / relax, 100);
    sp.transitions = cfg->transitions;
    sp.entropy_enabled = cfg->entropy_enabled;
    sp.entropy_threshold = cfg->entropy_threshold;
    gxy_gapped_params gp;
    gp.y_drop = cfg->gapped.y_drop;
    gp.gapped_threshold = std::max(cfg->gapped.gapped_threshold / relax, 100);
    const int respect = cfg->respect_masking ? 1 : 0;
    const bool byte_ok = scoring_byte_ok(cfg->scoring);
    auto by_tq = [](const Rec &a, const Rec &b) {
        return a.t_start != b.t_start ? a.t_start < b.t_start : a.q_start < b.q_start;
    };
    struct Win {
        size_t group;
        uint32_t tg0, qg0;
        DSeq t, q;
        size_t first, count;  // its HSPs in the batch
    };
    std::vector<Win> wins;
    std::vector<const gxy_dseq *> ts, qs;
    std::vector<gxy_hsp> hs;
    for (size_t gi = 0; gi < groups.size(); gi++) {
        std::vector<Rec> &recs = groups[gi].g->recs;
        if (recs.size() < 2) continue;
        std::stable_sort(recs.begin(), recs.end(), by_tq);
        for (size_t k = 0; k + 1 < recs.size(); k++) {
            const Rec &a = recs[k], &b = recs[k + 1];
            const uint32_t tg0 = a.t_start + a.t_span, tg1 = b.t_start;
            const uint32_t qg0 = a.q_start + a.q_span, qg1 = b.q_start;
            if (tg1 <= tg0 || qg1 <= qg0) continue;
            const uint32_t tgap = tg1 - tg0, qgap = qg1 - qg0;
            if (tgap < cfg->inner_min_gap || qgap < cfg->inner_min_gap) continue;
            if (tgap > cfg->inner_max_gap || qgap > cfg->inner_max_gap) continue;
            gxy_dseq *t, *q;
            int rc = gxy_dseq_slice(ctx, groups[gi].target, tg0, tg1, &t);
            if (rc) return rc;
            DSeq dt(t);
            rc = gxy_dseq_slice(ctx, groups[gi].query, qg0, qg1, &q);
            if (rc) return rc;
            DSeq dq(q);
            std::vector<gxy_hsp> hsps;
            {
                gxy_postable *tp;
                rc = postable_build_impl(ctx, t, &cfg->inner_pattern, 1, respect, &tp);
                if (rc) return rc;
                Table table(tp);
                if (cfg->max_word_count > 0) {
                    rc = gxy_postable_prune(ctx, tp, cfg->max_word_count);
                    if (rc) return rc;
                }
                uint8_t *qb = nullptr;
                gxy_ctx *qb_owner = nullptr;
                if (byte_ok) {
                    qb = (uint8_t *)dev_alloc(ctx, byte_buffer_size(q->len), &qb_owner);
                    if (!qb) return set_err(ctx, GXY_ENOMEM, "query byte copy: allocation failed");
                    rc = seq_to_bytes(ctx, q, false, qb);
                }
                gxy_search_stats st;
                if (!rc)
                    rc = search_dev_impl(ctx, tp, q, &cfg->scoring, &sp, respect, &hsps, &st,
                                         qb ? qb + kBytePad : nullptr, true);
                dev_free(qb_owner, qb);
                if (rc) return rc;
                rs->hits += st.hits;
                rs->columns += st.columns;
            }
            if (hsps.empty()) continue;
            if (cfg->chain_enabled) hsps = best_chain(hsps);
            Win w{gi, tg0, qg0, std::move(dt), std::move(dq), hs.size(), hsps.size()};
            for (const auto &h : hsps) {
                ts.push_back(w.t.get());
                qs.push_back(w.q.get());
                hs.push_back(h);
            }
            wins.push_back(std::move(w));
        }
    }
    if (!hs.empty()) {
        const size_t nb = hs.size();
        std::vector<gxy_anchor> anc(nb);
        int rc = choose_anchor_pairs_dev(ctx, ts.data(), qs.data(), hs.data(), nb, cfg->scoring, cfg->anchor_window,
                                         anc.data());
        if (rc) return rc;
        std::vector<GappedItem> items(nb);
        for (size_t x = 0; x < nb; x++) items[x] = {ts[x], qs[x], anc[x]};
        std::vector<GappedResult> br;
        std::vector<gxy_edit_run> runs;
        rc = gapped_batch_dev(ctx, items, cfg->scoring, gp, &br, &runs, &rs->dp_cells);
        if (rc) return rc;
        rs->dp_problems += nb;
        // extras appended per group in window order, then HSP order
        for (const Win &w : wins) {
            Group &g = *groups[w.group].g;
            for (size_t x = w.first; x < w.first + w.count; x++) {
                if (!br[x].found) continue;
                const gxy_alignment &a = br[x].aln;
                Rec r{g.ti, g.qi, g.strand, w.tg0 + a.t_start, w.qg0 + a.q_start, a.t_len, a.q_len, a.score,
                      a.n_runs, run_store->size()};
                run_store->insert(run_store->end(), runs.begin() + a.runs_offset,
                                  runs.begin() + a.runs_offset + a.n_runs);
                g.recs.push_back(r);
            }
        }
    }
    for (auto &tg : groups) {
        std::vector<Rec> &recs = tg.g->recs;
        if (recs.size() < 2) continue;
        std::stable_sort(recs.begin(), recs.end(), by_tq);
        std::vector<Rec> out;
        for (const Rec &r : recs) {  // dedup_by on (t_start, q_start, t_span, q_span)
            if (!out.empty()) {
                const Rec &p = out.back();
                if (p.t_start == r.t_start && p.q_start == r.q_start && p.t_span == r.t_span &&
                    p.q_span == r.q_span)
                    continue;
            }
            out.push_back(r);
        }
        recs.swap(out);
    }
    return GXY_OK;
}

constexpr int64_t kSpecDiag = 4096;

struct GapGroup {
    Group *g = nullptr;
    const gxy_dseq *tseq = nullptr;  // chunk sequence (HSP target coordinates are chunk-local)
    const gxy_dseq *qseq = nullptr;  // query strand
    uint32_t global_start = 0;       // chunk offset in its target
    std::vector<gxy_hsp> kept;
    std::vector<char> have;
    std::vector<GappedResult> res;
    std::vector<std::vector<gxy_edit_run>> scripts;
    struct Cov {
        uint32_t t_lo, t_hi, q_lo, q_hi;
    };
    std::vector<Cov> covered;
    size_t pos = 0, batch = 8;
    bool done = false;
    // src/driver.rs:288-297 compares chunk-local HSP coordinates with global
    // record coordinates; kept as written.
    bool is_covered(const gxy_hsp &h) const {
        for (const auto &cv : covered)
            if (h.t_start >= cv.t_lo && h.t_start < cv.t_hi && h.q_start >= cv.q_lo && h.q_start < cv.q_hi)
                return true;
        return false;
    }
};

// Run the gapped stage of every group to completion, one device batch per
// round over all groups that still have HSPs to settle.
int gapped_groups_dev(gxy_ctx *ctx, std::vector<std::unique_ptr<GapGroup>> &groups, const gxy_config *cfg,
                      std::vector<gxy_edit_run> *run_store, gxy_run_stats *rs,
                      const std::function<void()> &after_round = {}) {
    for (auto &gp : groups) {
        const size_t K = gp->kept.size();
        gp->have.assign(K, 0);
        gp->res.assign(K, GappedResult{false, {}});
        gp->scripts.assign(K, {});
        gp->pos = 0;
        gp->batch = 8;
        gp->done = K == 0;
    }
    std::vector<const gxy_dseq *> ts, qs;
    std::vector<gxy_hsp> hs;
    std::vector<std::pair<GapGroup *, size_t>> owner;
    for (;;) {
        ts.clear();
        qs.clear();
        hs.clear();
        owner.clear();
        for (auto &gp : groups) {
            GapGroup &G = *gp;
            if (G.done) continue;
            const size_t K = G.kept.size();
            // settle HSPs in order while their extensions are known
            while (G.pos < K) {
                if (G.is_covered(G.kept[G.pos])) {
                    G.pos++;
                    continue;
                }
                if (!G.have[G.pos]) break;
                if (G.res[G.pos].found) {  // build_gapped_record (src/driver.rs:395-434)
                    const gxy_alignment &a = G.res[G.pos].aln;
                    const auto &sc = G.scripts[G.pos];
                    Rec r{G.g->ti, G.g->qi, G.g->strand, a.t_start + G.global_start, a.q_start, a.t_len, a.q_len,
                          a.score, (uint32_t)sc.size(), run_store->size()};
                    run_store->insert(run_store->end(), sc.begin(), sc.end());
                    G.covered.push_back({r.t_start, r.t_start + r.t_span, r.q_start, r.q_start + r.q_span});
                    G.g->recs.push_back(r);
                }
                G.pos++;
            }
            if (G.pos >= K) {
                G.done = true;
                continue;
            }
            // Speculate on HSPs that an alignment picked this round is
            // unlikely to cover: a candidate within kSpecDiag diagonals of one
            // already picked waits for that alignment (the next one to settle
            // is always picked).  Only which extensions run ahead changes.
            const size_t first = owner.size();
            for (size_t k = G.pos; k < K && owner.size() - first < G.batch; k++) {
                if (G.have[k] || G.is_covered(G.kept[k])) continue;
                const int64_t dk = (int64_t)G.kept[k].t_start - (int64_t)G.kept[k].q_start;
                bool near = false;
                for (size_t x = first; x < owner.size() && !near; x++) {
                    const int64_t dx = (int64_t)hs[x].t_start - (int64_t)hs[x].q_start;
                    near = (dk > dx ? dk - dx : dx - dk) <= kSpecDiag;
                }
                if (near) continue;
                ts.push_back(G.tseq);
                qs.push_back(G.qseq);
                hs.push_back(G.kept[k]);
                owner.push_back({&G, k});
            }
            G.batch = std::min<size_t>(G.batch * 2, 1024);
        }
        if (after_round) after_round();
        if (owner.empty()) break;
        const size_t nb = owner.size();
        std::vector<gxy_anchor> anc(nb);
        int rc = choose_anchor_pairs_dev(ctx, ts.data(), qs.data(), hs.data(), nb, cfg->scoring, cfg->anchor_window,
                                         anc.data());
        if (rc) return rc;
        std::vector<GappedItem> items(nb);
        for (size_t x = 0; x < nb; x++) items[x] = {ts[x], qs[x], anc[x]};
        std::vector<GappedResult> br;
        std::vector<gxy_edit_run> runs;
        rc = gapped_batch_dev(ctx, items, cfg->scoring, cfg->gapped, &br, &runs, &rs->dp_cells);
        if (rc) return rc;
        rs->dp_problems += nb;
        for (size_t x = 0; x < nb; x++) {
            GapGroup &G = *owner[x].first;
            const size_t k = owner[x].second;
            G.have[k] = 1;
            G.res[k] = br[x];
            if (br[x].found)
                G.scripts[k].assign(runs.begin() + br[x].aln.runs_offset,
                                    runs.begin() + br[x].aln.runs_offset + br[x].aln.n_runs);
        }
    }
    return GXY_OK;
}

class HostJobs {
  public:
    explicit HostJobs(unsigned n) {
        for (unsigned k = 0; k < n; k++) threads_.emplace_back([this] { loop(); });
    }
    ~HostJobs() {
        {
            std::lock_guard<std::mutex> l(m_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto &t : threads_) t.join();
    }
    void submit(std::function<void()> f) {
        {
            std::lock_guard<std::mutex> l(m_);
            q_.push_back(std::move(f));
            pending_++;
        }
        cv_.notify_one();
    }
    void wait_all() {
        std::unique_lock<std::mutex> l(m_);
        done_cv_.wait(l, [this] { return pending_ == 0; });
    }

  private:
    void loop() {
        while (true) {
            std::function<void()> f;
            {
                std::unique_lock<std::mutex> l(m_);
                cv_.wait(l, [this] { return stop_ || !q_.empty(); });
                if (q_.empty()) return;
                f = std::move(q_.front());
                q_.pop_front();
            }
            f();
            {
                std::lock_guard<std::mutex> l(m_);
                pending_--;
            }
            done_cv_.notify_all();
        }
    }
    std::vector<std::thread> threads_;
    std::deque<std::function<void()>> q_;
    std::mutex m_;
    std::condition_variable cv_, done_cv_;
    size_t pending_ = 0;
    bool stop_ = false;
};


}  // namespace

extern "C" {

void gxy_config_default(gxy_config *c) {
    if (!c) return;
    std::memset(c, 0, sizeof *c);
    gxy_pattern_parse("12of19", &c->pattern);
    static const int32_t hox[4][4] = {{91, -114, -31, -123}, {-114, 100, -125, -31},
                                      {-31, -125, 100, -114}, {-123, -31, -114, 91}};  // src/scoring.rs:17-23
    std::memcpy(c->scoring.sub, hox, sizeof hox);
    c->scoring.gap_open = 400;
    c->scoring.gap_extend = 30;
    c->scoring.ambig = -100;
    c->step = 1;
    c->hsp = {910, 3000};
    c->gapped = {9400, 3000};
    c->strand = GXY_STRAND_BOTH;
    c->respect_masking = 1;
    c->gapped_enabled = 1;
    c->anchor_window = 31;
    c->transitions = 1;
    c->chunk_halo_bp = 50000;
    gxy_pattern_parse("match12", &c->inner_pattern);  // TweenerConfig::default (src/tweener.rs:44-52)
    c->inner_min_gap = 100;
    c->inner_max_gap = 100000;
    c->inner_relax = 2;
}

int gxy_best_chain(const gxy_hsp *hsps, uint64_t n, gxy_hsp *out, uint64_t *n_out) {
    return guarded(nullptr, [&]() -> int {
        if ((!hsps && n) || (!out && n) || !n_out) return GXY_EINVAL;
        const std::vector<gxy_hsp> c = best_chain(std::vector<gxy_hsp>(hsps, hsps + n));
        std::copy(c.begin(), c.end(), out);
        *n_out = c.size();
        return GXY_OK;
    });
}

int gxy_interpolate(gxy_ctx *ctx, const gxy_record_set *records, const gxy_dseq *target, const gxy_dseq *query,
                    const gxy_config *cfg, gxy_record_set **out) {
    return guarded(ctx, [&]() -> int {
        if (!ctx) return GXY_EINVAL;
        if (!records || !target || !query || !cfg || !out) return set_err(ctx, GXY_EINVAL, "null argument");
        GXY_HIP(ctx, hipSetDevice(ctx->device));
        Group g{0, 0, 0, 0, {}};
        std::vector<gxy_edit_run> run_store;
        for (uint64_t k = 0; k < records->n; k++) {
            const gxy_record &r = records->records[k];
            if (r.runs_offset + r.n_runs > records->n_runs) return set_err(ctx, GXY_EINVAL, "record runs out of range");
            if ((uint64_t)r.t_start + r.t_span > target->len || (uint64_t)r.q_start + r.q_span > query->len)
                return set_err(ctx, GXY_EINVAL, "record outside its sequences");
            g.ti = r.target_index;
            g.qi = r.query_index;
            g.strand = r.strand;
            g.recs.push_back(Rec{r.target_index, r.query_index, r.strand, r.t_start, r.q_start, r.t_span, r.q_span,
                                 r.score, r.n_runs, run_store.size()});
            run_store.insert(run_store.end(), records->runs + r.runs_offset, records->runs + r.runs_offset + r.n_runs);
        }
        std::vector<TweenGroup> tg{{&g, target, query}};
        gxy_run_stats rs;
        std::memset(&rs, 0, sizeof rs);
        int rc = tweener_groups_dev(ctx, tg, cfg, &run_store, &rs);
        if (rc) return rc;
        std::vector<std::unique_ptr<Group>> groups;
        groups.push_back(std::make_unique<Group>(std::move(g)));
        return assemble_records(ctx, groups, run_store, out);
    });
}

int gxy_pattern_parse(const char *spec, gxy_seed_pattern *out) {
    return guarded(nullptr, [&]() -> int {
        if (!spec || !out) return GXY_EINVAL;
        std::string s(spec), bits;
        if (s == "12of19") bits = "1110100110010101111";
        else if (s == "19of20") bits = "11111111101111111111";
        else if (s.rfind("match", 0) == 0) {
            const std::string k = s.substr(5);
            if (k.empty() || k.find_first_not_of("0123456789") != std::string::npos) return GXY_EINVAL;
            const unsigned long v = std::stoul(k);
            if (v == 0 || v > 32) return GXY_EINVAL;
            bits.assign(v, '1');
        } else if (!s.empty() && s.find_first_not_of("01") == std::string::npos) {
            bits = s;
        } else {
            return GXY_EINVAL;
        }
        if (bits.empty() || bits.size() > 32) return GXY_EINVAL;
        out->length = (uint32_t)bits.size();
        out->care = 0;
        for (size_t k = 0; k < bits.size(); k++)
            if (bits[k] == '1') out->care |= 1u << k;
        return GXY_OK;
    });
}

int gxy_run_dev(gxy_ctx *ctx, const gxy_dseq *const *targets, uint64_t n_targets, const gxy_dseq *const *queries,
                uint64_t n_queries, const gxy_config *cfg, gxy_record_set **out, gxy_run_stats *stats) {
    return guarded(ctx, [&]() -> int {
        if (!ctx) return GXY_EINVAL;
        if ((!targets && n_targets) || (!queries && n_queries) || !cfg || !out)
            return set_err(ctx, GXY_EINVAL, "null argument");
        for (uint64_t i = 0; i < n_targets; i++)
            if (!targets[i]) return set_err(ctx, GXY_EINVAL, "null target %llu", (unsigned long long)i);
        for (uint64_t i = 0; i < n_queries; i++)
            if (!queries[i]) return set_err(ctx, GXY_EINVAL, "null query %llu", (unsigned long long)i);
        if (cfg->step < 1) return set_err(ctx, GXY_EINVAL, "step must be >= 1");
        GXY_HIP(ctx, hipSetDevice(ctx->device));
        PhaseClock clk(ctx);
        // Internal sequences and tables come from the context's block cache.
        struct PoolScope {
            gxy_ctx *c;
            explicit PoolScope(gxy_ctx *x) : c(x) { c->pool_on = true; }
            ~PoolScope() { c->pool_on = false; }
        } pool_scope(ctx);
        gxy_run_stats rs;
        std::memset(&rs, 0, sizeof rs);
        const int respect = cfg->respect_masking ? 1 : 0;

        std::vector<uint32_t> strands;
        if (cfg->strand == GXY_STRAND_PLUS || cfg->strand == GXY_STRAND_BOTH) strands.push_back(GXY_STRAND_PLUS);
        if (cfg->strand == GXY_STRAND_MINUS || cfg->strand == GXY_STRAND_BOTH) strands.push_back(GXY_STRAND_MINUS);
        if (strands.empty()) return set_err(ctx, GXY_EINVAL, "invalid strand selector");
        // minus strand: one device reverse complement per query (src/driver.rs:221)
        std::vector<DSeq> drc(n_queries);
        if (cfg->strand != GXY_STRAND_PLUS)
            for (uint64_t i = 0; i < n_queries; i++) {
                gxy_dseq *d;
                int rc = gxy_dseq_reverse_complement(ctx, queries[i], &d);
                if (rc) return rc;
                drc[i].reset(d);
            }

        // byte copies of every query strand, shared by all chunk searches
        struct DevBytes {
            uint8_t *p = nullptr;
            gxy_ctx *owner = nullptr;
            ~DevBytes() { dev_free(owner, p); }
        };
        const bool byte_ok = scoring_byte_ok(cfg->scoring);
        std::vector<DevBytes> qbytes(byte_ok ? 2 * n_queries : 0);
        if (byte_ok) {
            for (uint64_t i = 0; i < n_queries; i++) {
                for (int s = 0; s < 2; s++) {
                    const gxy_dseq *qs = s == 0 ? queries[i] : drc[i].get();
                    if (!qs) continue;
                    DevBytes &b = qbytes[2 * i + s];
                    b.p = (uint8_t *)dev_alloc(ctx, byte_buffer_size(qs->len), &b.owner);
                    if (!b.p) return set_err(ctx, GXY_ENOMEM, "query byte copies: allocation failed");
                    int rc = seq_to_bytes(ctx, qs, false, b.p);
                    if (rc) return rc;
                }
            }
        }

        // chunk_target (src/driver.rs:156-194)
        std::vector<Chunk> chunks;
        {
            int rc = build_chunks(ctx, targets, n_targets, cfg, &chunks);
            if (rc) return rc;
        }
        rs.chunks = chunks.size();

        gxy_search_params sp;
        sp.step = cfg->step;
        sp.hsp = cfg->hsp;
        sp.transitions = cfg->transitions;
        sp.entropy_enabled = cfg->entropy_enabled;
        sp.entropy_threshold = cfg->entropy_threshold;

        std::vector<std::unique_ptr<Group>> groups;
        std::vector<std::unique_ptr<GapGroup>> gap_groups;
        std::vector<TweenGroup> tween_groups;
        std::vector<gxy_edit_run> run_store;  // scripts of gapped records
        std::atomic<uint64_t> ties{0};        // tie pairs counted by the host jobs
        HostJobs jobs(2);                     // joined before groups and ties are destroyed
        // With both strands, a query's minus-strand search runs on a helper
        // context (its own stream and scratch) from a helper thread while the
        // plus-strand search runs here, so each search's host round trips overlap
        // the other's kernels.  The two read the same table and sequences;
        // their results are taken in strand order, as one after the other.
        // postable_build_impl and gxy_postable_prune return with ctx's stream
        // drained, so the table, the chunk sequences and the query copies are
        // complete before the helper's stream reads them.
        gxy_ctx *helper = nullptr;
        if (strands.size() == 2) {
            if (!ctx->helper && gxy_ctx_create(ctx->device, &ctx->helper) == GXY_OK)
                ctx->helper->pool_on = true;  // its own block cache: nothing it frees waits on other streams
            if (ctx->helper && !ctx->helper->stream &&
                hipStreamCreateWithFlags(&ctx->helper->stream, hipStreamNonBlocking) != hipSuccess) {
                (void)hipGetLastError();
                gxy_ctx_destroy(ctx->helper);
                ctx->helper = nullptr;
            }
            helper = ctx->helper;
            if (helper) {
                helper->timing = ctx->timing;
                helper->xdrop_single = ctx->xdrop_single;
                helper->hit_batch = ctx->hit_batch;
            }
        }
        // A second lane (its own context and helper, so four streams in all)
        // takes every other (chunk, query) item: two items' searches run at
        // once and fill each other's host round trips and table builds.  Each
        // lane builds the tables of its items' chunks and keeps its items'
        // results; they are taken here in item order (chunk-major), as one
        // lane would give them.
        const size_t nitems = chunks.size() * (size_t)n_queries;
        gxy_ctx *lane = nullptr, *lane_helper = nullptr;
        if (nitems >= 2 && helper) {
            if (!ctx->lane && gxy_ctx_create(ctx->device, &ctx->lane) == GXY_OK) ctx->lane->pool_on = true;
            if (ctx->lane && !ctx->lane->helper && gxy_ctx_create(ctx->device, &ctx->lane->helper) == GXY_OK)
                ctx->lane->helper->pool_on = true;
            bool ok = ctx->lane && ctx->lane->helper;
            for (gxy_ctx *c : {ctx->lane, ok ? ctx->lane->helper : nullptr}) {
                if (ok && !c->stream && hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
                    (void)hipGetLastError();
                    ok = false;
                }
            }
            if (ok) {
                lane = ctx->lane;
                lane_helper = lane->helper;
                for (gxy_ctx *c : {lane, lane_helper}) {
                    c->timing = ctx->timing;
                    c->xdrop_single = ctx->xdrop_single;
                    c->hit_batch = ctx->hit_batch;
                }
                // the chunk sequences, reverse complements and query byte copies
                // were written on ctx's stream
                GXY_HIP(ctx, hipStreamSynchronize(ctx->stream));
            }
        }
        struct ItemOut {
            std::vector<gxy_hsp> shs[2];  // per strand slot
            gxy_search_stats sst[2] = {};
            int rc = GXY_OK;
            bool done = false;
        };
        std::vector<ItemOut> iouts(nitems);
        std::mutex cmu;
        std::condition_variable ccv;
        // A lane's table: built for the chunk of its next item when that
        // chunk changes (the chunk's table is released then; gapped work needs
        // only the sequences).
        struct LaneTable {
            Table table;
            size_t chunk = SIZE_MAX;
        };
        // item k = (chunk k / n_queries, query k % n_queries) on a lane (lc:
        // plus strand, on this thread; lh: the other strand, on the lane's
        // helper thread)
        auto search_item = [&](size_t k, gxy_ctx *lc, gxy_ctx *lh, HostJobs &hj, LaneTable &lt) -> int {
            const size_t c = k / n_queries;
            const uint32_t qi = (uint32_t)(k % n_queries);
            const Chunk &ch = chunks[c];
            ItemOut &io = iouts[k];
            if (lt.chunk != c) {
                lt.table.reset();
                gxy_postable *tp;
                PhaseClock lclk(lc);
                int rc = postable_build_impl(lc, ch.seq, &cfg->pattern, cfg->step, respect, &tp);
                lclk.lap(kTrTable);
                if (rc) return rc;
                lt.table.reset(tp);
                lt.chunk = c;
                if (cfg->max_word_count > 0) {
                    rc = gxy_postable_prune(lc, tp, cfg->max_word_count);
                    if (rc) return rc;
                }
            }
            gxy_postable *tp = lt.table.get();
            int src[2] = {GXY_OK, GXY_OK};
            auto run_search = [&](size_t s2, gxy_ctx *sc) {
                const uint32_t strand = strands[s2];
                const gxy_dseq *qs = strand == GXY_STRAND_PLUS ? queries[qi] : drc[qi].get();
                const int si = strand == GXY_STRAND_PLUS ? 0 : 1;
                const uint8_t *qb = byte_ok ? qbytes[2 * qi + si].p + kBytePad : nullptr;
                src[s2] = search_dev_impl(sc, tp, qs, &cfg->scoring, &sp, respect, &io.shs[s2], &io.sst[s2], qb,
                                          cfg->gapped_enabled != 0);
            };
            if (lh) {
                hj.submit([&] {
                    (void)hipSetDevice(lc->device);
                    run_search(1, lh);
                });
                run_search(0, lc);
                hj.wait_all();
                if (!src[0] && src[1]) return set_err(lc, src[1], "%s", lh->err.c_str());
            } else {
                for (size_t s2 = 0; s2 < strands.size() && !src[0]; s2++) run_search(s2, lc);
            }
            return src[0];
        };
        // the searches' HSPs of item k, taken in item order
        auto take_item = [&](size_t ik) {
            const size_t c = ik / n_queries;
            const Chunk &ch = chunks[c];
            ItemOut &io = iouts[ik];
            {
                const uint32_t qi = (uint32_t)(ik % n_queries);
                for (size_t k = 0; k < strands.size(); k++) {
                    const uint32_t strand = strands[k];
                    const gxy_dseq *qs = strand == GXY_STRAND_PLUS ? queries[qi] : drc[qi].get();
                    std::vector<gxy_hsp> hsps = std::move(io.shs[k]);
                    const gxy_search_stats &st = io.sst[k];
                    rs.hits += st.hits;
                    rs.columns += st.columns;
                    rs.hsps += st.emitted;
                    rs.tie_pairs += st.tie_pairs;
                    rs.passing += st.passing;
                    rs.lookups += st.lookups;
                    if (hsps.empty()) continue;
                    groups.push_back(std::make_unique<Group>(Group{ch.parent, qi, strand, c, {}}));
                    Group &g = *groups.back();
                    if (cfg->tweener_enabled) tween_groups.push_back({&g, targets[ch.parent], qs});
                    if (!cfg->gapped_enabled) {
                        // primary-region filter (src/driver.rs:250-256), best_chain
                        // (src/driver.rs:258-260) and build_ungapped_record
                        // (src/driver.rs:437-468), off the device path
                        jobs.submit([&g, &ties, hsps = std::move(hsps), lo = ch.primary_start_local,
                                     hi = ch.primary_end_local, gs = ch.global_start, chain = cfg->chain_enabled != 0,
                                     tween = cfg->tweener_enabled != 0]() mutable {
                            ties += dedup_hsps(&hsps);
                            std::vector<gxy_hsp> kept;
                            kept.reserve(hsps.size());
                            for (const auto &h : hsps)
                                if (h.t_start >= lo && h.t_start < hi) kept.push_back(h);
                            if (chain) kept = best_chain(kept);
                            g.recs.reserve(kept.size());
                            for (const auto &h : kept)
                                g.recs.push_back(Rec{g.ti, g.qi, g.strand, h.t_start + gs, h.q_start, h.length,
                                                     h.length, h.score, 0, 0});
                            if (!tween) finalize_group(g);  // the tweener runs first otherwise
                        });
                    } else {
                        // primary-region filter (src/driver.rs:250-256); score order (src/driver.rs:270)
                        auto gg = std::make_unique<GapGroup>();
                        gg->g = &g;
                        gg->tseq = ch.seq;
                        gg->qseq = qs;
                        gg->global_start = ch.global_start;
                        gg->kept.reserve(hsps.size());
                        for (const auto &h : hsps)
                            if (h.t_start >= ch.primary_start_local && h.t_start < ch.primary_end_local)
                                gg->kept.push_back(h);
                        if (cfg->chain_enabled) gg->kept = best_chain(gg->kept);  // src/driver.rs:258-260
                        std::stable_sort(gg->kept.begin(), gg->kept.end(),
                                         [](const gxy_hsp &a, const gxy_hsp &b) { return a.score > b.score; });
                        for (size_t k2 = 1; k2 < gg->kept.size(); k2++)
                            if (gg->kept[k2].score == gg->kept[k2 - 1].score) rs.score_ties++;
                        gap_groups.push_back(std::move(gg));
                    }
                }
            }
            io.shs[0] = {};
            io.shs[1] = {};
        };

        clk.lap(kTrSetup);
        if (!lane) {
            HostJobs helper_job(helper ? 1 : 0);
            LaneTable lt;
            for (size_t k = 0; k < nitems; k++) {
                const int rc = search_item(k, ctx, helper, helper_job, lt);
                if (rc) return rc;
                take_item(k);
            }
        } else {
            // lane 0 (ctx, its helper) takes the even items, lane 1 the odd ones
            std::atomic<bool> abort_lanes{false};
            auto lane_loop = [&](int L, gxy_ctx *lc, gxy_ctx *lh) {
                (void)hipSetDevice(lc->device);
                HostJobs hj(1);
                LaneTable lt;
                for (size_t k = (size_t)L; k < nitems && !abort_lanes.load(); k += 2) {
                    const int rc = search_item(k, lc, lh, hj, lt);
                    {
                        std::lock_guard<std::mutex> lk(cmu);
                        iouts[k].rc = rc;
                        iouts[k].done = true;
                    }
                    ccv.notify_all();
                    if (rc) break;
                }
            };
            std::thread t0(lane_loop, 0, ctx, helper), t1(lane_loop, 1, lane, lane_helper);
            int rc = GXY_OK;
            for (size_t k = 0; k < nitems; k++) {
                {
                    std::unique_lock<std::mutex> lk(cmu);
                    ccv.wait(lk, [&] { return iouts[k].done; });
                    rc = iouts[k].rc;
                }
                if (rc) {
                    if (k % 2) set_err(ctx, rc, "%s", lane->err.c_str());
                    abort_lanes = true;
                    // a lane stopped by abort_lanes leaves later items undone
                    for (size_t d = k + 1; d < nitems; d++) {
                        std::lock_guard<std::mutex> lk(cmu);
                        if (!iouts[d].done) {
                            iouts[d].done = true;
                            iouts[d].rc = rc;
                        }
                    }
                    break;
                }
                take_item(k);
            }
            t0.join();
            t1.join();
            if (rc) return rc;
        }
        if (!gap_groups.empty()) {
            int rc = gapped_groups_dev(ctx, gap_groups, cfg, &run_store, &rs);
            if (rc) return rc;
            if (!cfg->tweener_enabled)
                for (auto &gg : gap_groups) jobs.submit([g = gg->g] { finalize_group(*g); });
        }
        jobs.wait_all();
        if (cfg->tweener_enabled && !tween_groups.empty()) {
            // tweener over the whole target (src/driver.rs:340-356), then the group sort
            int rc = tweener_groups_dev(ctx, tween_groups, cfg, &run_store, &rs);
            if (rc) return rc;
            for (auto &tg : tween_groups) finalize_group(*tg.g);
        }
        // the helpers' and the second lane's x-drop counts and kernel times
        // count as this context's
        for (gxy_ctx *aux : {helper, lane, lane_helper}) {
            if (!aux) continue;
            ctx->xd_first_cols += aux->xd_first_cols;
            ctx->xd_conts += aux->xd_conts;
            ctx->xd_first_pass += aux->xd_first_pass;
            aux->xd_first_cols = aux->xd_conts = aux->xd_first_pass = 0;
            if (!aux->timing) continue;
            timing_flush(aux);
            for (const auto &t : aux->totals) {
                bool found = false;
                for (auto &u : ctx->totals)
                    if (u.first == t.first) {
                        u.second.first += t.second.first;
                        u.second.second += t.second.second;
                        found = true;
                        break;
                    }
                if (!found) ctx->totals.push_back(t);
            }
            aux->totals.clear();
        }
        // The helpers' and the lane's streams end with the run (their scratch
        // and block caches stay for the next run): streams kept would take
        // shares of the device's four hardware queues from streams that other
        // contexts create later.
        for (gxy_ctx *aux : {ctx->helper, ctx->lane, ctx->lane ? ctx->lane->helper : nullptr}) {
            if (aux && aux->stream) {
                (void)hipStreamSynchronize(aux->stream);
                (void)hipStreamDestroy(aux->stream);
                aux->stream = nullptr;
            }
        }
        rs.tie_pairs += ties.load();
        for (uint64_t qi = 0; qi < n_queries; qi++) rs.query_bp += queries[qi]->len * strands.size();

        clk.lap(kTrRecords);
        {
            int rc = assemble_records(ctx, groups, run_store, out);
            if (rc) return rc;
        }
        if (stats) *stats = rs;
        clk.lap(kTrMerge);
        if (ctx->trace) {
            static const char *names[] = {"setup", "table", "s.count", "s.extend", "s.admit", "s.copy", "s.dedup",
                                          "records", "merge", "s.other"};
            std::fprintf(stderr, "gxy_run phases ms:");
            for (int k = 0; k < 10; k++) std::fprintf(stderr, " %s %.2f", names[k], ctx->trace_ms[k]);
            std::fprintf(stderr, "\n");
            std::memset(ctx->trace_ms, 0, sizeof ctx->trace_ms);
        }
        return GXY_OK;
    });
}

int gxy_run(gxy_ctx *ctx, const gxy_named_seq *targets, uint64_t n_targets, const gxy_named_seq *queries,
            uint64_t n_queries, const gxy_config *cfg, gxy_record_set **out, gxy_run_stats *stats) {
    return guarded(ctx, [&]() -> int {
        if (!ctx) return GXY_EINVAL;
        if ((!targets && n_targets) || (!queries && n_queries) || !cfg || !out)
            return set_err(ctx, GXY_EINVAL, "null argument");
        GXY_HIP(ctx, hipSetDevice(ctx->device));
        std::vector<DSeq> dt(n_targets), dq(n_queries);
        std::vector<const gxy_dseq *> tp(n_targets), qp(n_queries);
        for (uint64_t i = 0; i < n_targets; i++) {
            gxy_dseq *d;
            int rc = dseq_upload_impl(ctx, &targets[i].seq, &d);
            if (rc) return rc;
            dt[i].reset(d);
            tp[i] = d;
        }
        for (uint64_t i = 0; i < n_queries; i++) {
            gxy_dseq *d;
            int rc = dseq_upload_impl(ctx, &queries[i].seq, &d);
            if (rc) return rc;
            dq[i].reset(d);
            qp[i] = d;
        }
        return gxy_run_dev(ctx, tp.data(), n_targets, qp.data(), n_queries, cfg, out, stats);
    });
}

void gxy_record_set_free(gxy_record_set *set) {
    if (!set) return;
    std::free(set->records);
    std::free(set->runs);
    std::free(set);
}


}  // extern "C"

// ---- Serving ----------------------------------------------------------------

namespace {

struct ServeReq {
    uint64_t id = 0;
    std::vector<uint8_t> codes, valid, mask;  // the query, PackedSeq layout
    uint64_t len = 0;
    double submit_ms = 0, admit_ms = 0, done_ms = 0;
    uint64_t iteration = 0, batch_queries = 0, batch_bp = 0;
    gxy_record_set *out = nullptr;
    int rc = GXY_OK;
    std::string err;
    bool done = false;
    bool claimed = false;  // a gxy_server_wait has taken this ticket
    // between search and completion
    DSeq dq, drc;  // query strands (freed by the search thread: its block cache)
    std::vector<std::unique_ptr<Group>> groups;
    std::vector<std::unique_ptr<GapGroup>> gaps;
    std::vector<TweenGroup> tweens;  // when the tweener is on
};

// Device bytes a query holds while it is searched: both strands as 2-bit
// blocks (0.5 B/base each) and as bytes (1 B/base each), their share of the
// concatenated copies (the same again), and search scratch that scales with
// the query (window words and counts, 12 B/window; hit records are bounded by
// the batched hit buffers).
constexpr uint64_t kServeBytesPerBase = 2 * (2 * (0.5 + 1)) + 12;

}  // namespace

struct gxy_server {
    gxy_ctx *ctx = nullptr;               // search stage
    std::vector<gxy_ctx *> dp_ctx;        // gapped stage, one per worker
    std::vector<const gxy_dseq *> targets;
    gxy_config cfg;
    gxy_server_params prm;
    std::vector<Chunk> chunks;
    std::vector<Table> tables;
    std::vector<uint32_t> strands;
    uint32_t sep = 0;  // N bases between concatenated queries (0: no concatenation)
    uint32_t align = 32;
    std::mutex mu;
    std::condition_variable cv_work, cv_dp, cv_done;
    std::deque<ServeReq *> queue;   // waiting for search
    std::deque<ServeReq *> dpq;     // waiting for the gapped stage
    std::vector<DSeq> tofree;       // sequences of completed queries
    std::unordered_map<uint64_t, std::unique_ptr<ServeReq>> reqs;
    uint64_t next_id = 1, iteration = 0;
    bool stop = false, search_done = false;
    std::string err;
    gxy_run_stats totals{};  // work of every iteration so far (under mu)
    std::thread worker;
    std::vector<std::thread> dp_workers;
    std::chrono::steady_clock::time_point t0;
    double now_ms() const {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
};

namespace {

// Records of a query whose groups are all settled; call with S->mu free.
void complete(gxy_server *S, gxy_ctx *ctx, ServeReq *r, std::vector<gxy_edit_run> &run_store, int rc) {
    if (!rc && S->cfg.tweener_enabled && !r->tweens.empty()) {
        gxy_run_stats rs;
        std::memset(&rs, 0, sizeof rs);
        rc = tweener_groups_dev(ctx, r->tweens, &S->cfg, &run_store, &rs);
    }
    if (!rc) {
        for (auto &g : r->groups) finalize_group(*g);  // idempotent for groups already sorted
        rc = assemble_records(ctx, r->groups, run_store, &r->out);
    }
    std::string msg = rc ? ctx->err : std::string();
    {
        std::lock_guard<std::mutex> lk(S->mu);
        if (rc) {
            gxy_record_set_free(r->out);
            r->out = nullptr;
            r->rc = rc;
            r->err = msg;
            S->err = msg;
        }
        r->gaps.clear();
        r->tweens.clear();
        r->groups.clear();
        S->tofree.push_back(std::move(r->dq));
        S->tofree.push_back(std::move(r->drc));
        r->done_ms = S->now_ms();
        r->done = true;
    }
    S->cv_done.notify_all();
    S->cv_work.notify_one();
}

// Search stage of one iteration: the queries of `batch` through search and
// the ungapped records or the gapped groups of each query.
int serve_search(gxy_server *S, std::vector<ServeReq *> &batch) {
    gxy_ctx *ctx = S->ctx;
    const gxy_config *cfg = &S->cfg;
    const size_t n = batch.size();
    const int respect = cfg->respect_masking ? 1 : 0;
    const bool minus = cfg->strand != GXY_STRAND_PLUS;
    // queries and their reverse complements (src/driver.rs:221)
    for (size_t i = 0; i < n; i++) {
        gxy_seq v;
        v.codes = batch[i]->codes.data();
        v.valid = batch[i]->valid.data();
        v.mask = batch[i]->mask.data();
        v.len = batch[i]->len;
        gxy_dseq *d;
        int rc = gxy_dseq_upload(ctx, &v, &d);
        if (rc) return rc;
        batch[i]->dq.reset(d);
        if (minus) {
            rc = gxy_dseq_reverse_complement(ctx, d, &d);
            if (rc) return rc;
            batch[i]->drc.reset(d);
        }
    }
    // Search inputs: one concatenation per strand (queries at block-aligned
    // offsets, N between them), or the queries themselves.
    const bool concat = S->sep > 0 && n > 1;
    std::vector<uint64_t> off(n + 1, 0);
    std::vector<DSeq> cat(2);
    if (concat) {
        uint64_t o = 0;
        for (size_t i = 0; i < n; i++) {
            off[i] = o;
            o += batch[i]->len + S->sep;
            o = (o + S->align - 1) / S->align * S->align;
        }
        off[n] = o;
        const uint64_t total = off[n - 1] + batch[n - 1]->len;
        if (total > 0xFFFFFFF0ull) return set_err(ctx, GXY_EINVAL, "iteration exceeds 2^32 query bases");
        for (int sidx = 0; sidx < (minus ? 2 : 1); sidx++) {
            gxy_dseq *c;
            int rc = dseq_alloc(ctx, total, &c);
            if (rc) return rc;
            cat[sidx].reset(c);
            GXY_HIP(ctx, hipMemsetAsync(c->d, 0, c->nblk * sizeof(Blk), ctx->stream));
            for (size_t i = 0; i < n; i++) {
                const gxy_dseq *src = sidx == 0 ? batch[i]->dq.get() : batch[i]->drc.get();
                const uint64_t nb = (src->len + 31) / 32;
                if (nb)
                    GXY_HIP(ctx, hipMemcpyAsync(c->d + off[i] / 32, src->d, nb * sizeof(Blk),
                                                hipMemcpyDeviceToDevice, ctx->stream));
            }
        }
    }
    struct DevBytes {
        uint8_t *p = nullptr;
        gxy_ctx *owner = nullptr;
        ~DevBytes() { dev_free(owner, p); }
    };
    const bool byte_ok = scoring_byte_ok(cfg->scoring);
    gxy_search_params sp;
    sp.step = cfg->step;
    sp.hsp = cfg->hsp;
    sp.transitions = cfg->transitions;
    sp.entropy_enabled = cfg->entropy_enabled;
    sp.entropy_threshold = cfg->entropy_threshold;

    for (uint32_t strand : S->strands) {
        const int si = strand == GXY_STRAND_PLUS ? 0 : 1;
        std::vector<const gxy_dseq *> units;
        if (concat)
            units.push_back(cat[si].get());
        else
            for (size_t i = 0; i < n; i++) units.push_back(si == 0 ? batch[i]->dq.get() : batch[i]->drc.get());
        std::vector<DevBytes> ub(byte_ok ? units.size() : 0);
        for (size_t u = 0; u < ub.size(); u++) {
            ub[u].p = (uint8_t *)dev_alloc(ctx, byte_buffer_size(units[u]->len), &ub[u].owner);
            if (!ub[u].p) return set_err(ctx, GXY_ENOMEM, "query byte copies: allocation failed");
            int rc = seq_to_bytes(ctx, units[u], false, ub[u].p);
            if (rc) return rc;
        }
        for (size_t c = 0; c < S->chunks.size(); c++) {
            const Chunk &ch = S->chunks[c];
            std::vector<std::vector<gxy_hsp>> per(n);
            for (size_t u = 0; u < units.size(); u++) {
                std::vector<gxy_hsp> hsps;
                gxy_search_stats st;
                int rc = search_dev_impl(ctx, S->tables[c].get(), units[u], &cfg->scoring, &sp, respect, &hsps, &st,
                                         byte_ok ? ub[u].p + kBytePad : nullptr, cfg->gapped_enabled != 0);
                if (rc) return rc;
                {
                    std::lock_guard<std::mutex> lk(S->mu);
                    S->totals.hits += st.hits;
                    S->totals.columns += st.columns;
                    S->totals.passing += st.passing;
                    S->totals.hsps += hsps.size();
                }
                if (!concat) {
                    per[u] = std::move(hsps);
                    continue;
                }
                // split by query: HSPs never reach into a separator, so each
                // lies inside one query's range; order within a query is kept
                for (const auto &h : hsps) {
                    const size_t i = (size_t)(std::upper_bound(off.begin(), off.begin() + n, (uint64_t)h.q_start) -
                                              off.begin()) - 1;
                    gxy_hsp x = h;
                    x.q_start = (uint32_t)(h.q_start - off[i]);
                    if ((uint64_t)x.q_start + x.length > batch[i]->len)
                        return set_err(ctx, GXY_EDEVICE, "HSP crosses a query separator");
                    per[i].push_back(x);
                }
            }
            for (size_t i = 0; i < n; i++) {
                std::vector<gxy_hsp> &hsps = per[i];
                if (hsps.empty()) continue;
                ServeReq *r = batch[i];
                r->groups.push_back(std::make_unique<Group>(Group{ch.parent, 0, strand, c, {}}));
                Group &g = *r->groups.back();
                const gxy_dseq *qstrand = si == 0 ? r->dq.get() : r->drc.get();
                if (cfg->tweener_enabled) r->tweens.push_back({&g, S->targets[ch.parent], qstrand});
                if (!cfg->gapped_enabled) {
                    // Vec::dedup, primary-region filter (src/driver.rs:250-256),
                    // best_chain (src/driver.rs:258-260) and build_ungapped_record
                    // (src/driver.rs:437-468); the group is sorted at completion
                    dedup_hsps(&hsps);
                    std::vector<gxy_hsp> kept;
                    for (const auto &h : hsps)
                        if (h.t_start >= ch.primary_start_local && h.t_start < ch.primary_end_local)
                            kept.push_back(h);
                    if (cfg->chain_enabled) kept = best_chain(kept);
                    for (const auto &h : kept)
                        g.recs.push_back(Rec{g.ti, 0, strand, h.t_start + ch.global_start, h.q_start, h.length,
                                             h.length, h.score, 0, 0});
                } else {
                    auto gg = std::make_unique<GapGroup>();
                    gg->g = &g;
                    gg->tseq = ch.seq;
                    gg->qseq = qstrand;
                    gg->global_start = ch.global_start;
                    for (const auto &h : hsps)
                        if (h.t_start >= ch.primary_start_local && h.t_start < ch.primary_end_local)
                            gg->kept.push_back(h);
                    if (cfg->chain_enabled) gg->kept = best_chain(gg->kept);
                    std::stable_sort(gg->kept.begin(), gg->kept.end(),
                                     [](const gxy_hsp &a, const gxy_hsp &b) { return a.score > b.score; });
                    r->gaps.push_back(std::move(gg));
                }
            }
        }
    }
    return GXY_OK;
}

void search_loop(gxy_server *S) {
    (void)hipSetDevice(S->ctx->device);
    const uint64_t budget = S->prm.hbm_budget_bytes ? S->prm.hbm_budget_bytes : (16ull << 30);
    const uint64_t max_bp = S->prm.max_batch_bp ? S->prm.max_batch_bp : (256ull << 20);
    S->ctx->pool_on = true;  // this thread's sequences come from its block cache
    for (;;) {
        std::vector<ServeReq *> batch;
        std::vector<DSeq> drop;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv_work.wait(lk, [&] { return S->stop || !S->queue.empty() || !S->tofree.empty(); });
            drop.swap(S->tofree);
            if (S->queue.empty() && S->stop) {
                S->search_done = true;
                break;
            }
            uint64_t bp = 0, bytes = 0;
            while (!S->queue.empty()) {
                ServeReq *r = S->queue.front();
                const uint64_t rb = r->len * kServeBytesPerBase;
                if (!batch.empty() && (!S->prm.co_batch || bp + r->len > max_bp || bytes + rb > budget)) break;
                batch.push_back(r);
                S->queue.pop_front();
                bp += r->len;
                bytes += rb;
            }
            if (!batch.empty()) S->iteration++;
            const double t = S->now_ms();
            for (ServeReq *r : batch) {
                r->admit_ms = t;
                r->iteration = S->iteration;
                r->batch_queries = batch.size();
                r->batch_bp = bp;
            }
        }
        drop.clear();  // completed queries' sequences back to the block cache
        if (batch.empty()) continue;
        const int rc = serve_search(S, batch);
        std::vector<ServeReq *> to_dp;
        for (ServeReq *r : batch) {
            if (rc || r->gaps.empty()) {
                std::vector<gxy_edit_run> run_store;
                complete(S, S->ctx, r, run_store, rc);
            } else if (!S->prm.co_batch) {
                // the plain driver: this query's gapped stage before the next query
                std::vector<gxy_edit_run> run_store;
                gxy_run_stats rs;
                std::memset(&rs, 0, sizeof rs);
                const int rc2 = gapped_groups_dev(S->ctx, r->gaps, &S->cfg, &run_store, &rs);
                complete(S, S->ctx, r, run_store, rc2);
            } else {
                to_dp.push_back(r);
            }
        }
        if (!to_dp.empty()) {
            {
                std::lock_guard<std::mutex> lk(S->mu);
                for (ServeReq *r : to_dp) S->dpq.push_back(r);
            }
            S->cv_dp.notify_one();
        }
    }
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->tofree.clear();
    }
    S->cv_dp.notify_all();
    S->ctx->pool_on = false;
}

// Gapped stage: take the waiting queries of one length bin and extend their
// HSPs in shared rounds on this worker's stream; a query completes when its
// own groups are settled.
void dp_loop(gxy_server *S, gxy_ctx *ctx) {
    (void)hipSetDevice(ctx->device);
    // the tweener's sequences and tables come from this worker's block cache:
    // a hipFree would wait for every stream on the device
    ctx->pool_on = true;
    for (;;) {
        std::vector<ServeReq *> mine;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv_dp.wait(lk, [&] { return !S->dpq.empty() || S->search_done; });
            if (S->dpq.empty()) return;  // search finished and nothing left
            if (S->prm.co_batch) {
                // length-binned: the oldest query's bin, so a round (which lasts
                // as long as its longest extension) holds queries of similar size
                auto bin = [](uint64_t len) { return len < 10000 ? 0 : len < 100000 ? 1 : len < 1000000 ? 2 : 3; };
                const int b = bin(S->dpq.front()->len);
                for (auto it = S->dpq.begin(); it != S->dpq.end();) {
                    if (bin((*it)->len) == b) {
                        mine.push_back(*it);
                        it = S->dpq.erase(it);
                    } else {
                        ++it;
                    }
                }
            } else {
                mine.push_back(S->dpq.front());
                S->dpq.pop_front();
            }
        }
        std::vector<std::unique_ptr<GapGroup>> all;
        std::vector<std::pair<ServeReq *, size_t>> span;  // [first, count) of each query's groups in `all`
        for (ServeReq *r : mine) {
            span.push_back({r, r->gaps.size()});
            for (auto &gg : r->gaps) all.push_back(std::move(gg));
            r->gaps.clear();
        }
        std::vector<gxy_edit_run> run_store;
        std::vector<char> fin(mine.size(), 0);
        auto settle = [&](int rc) {
            size_t o = 0;
            std::vector<size_t> now;  // queries whose groups all settled since the last call
            for (size_t x = 0; x < mine.size(); x++) {
                const size_t cnt = span[x].second;
                bool done = true;
                for (size_t y = 0; y < cnt && done; y++) done = all[o + y]->done;
                if (!fin[x] && (done || rc)) {
                    // settled groups stay in `all` (done, never touched again)
                    fin[x] = 1;
                    now.push_back(x);
                }
                o += cnt;
            }
            // the tweener of all of them in one call (one batched gapped stage)
            if (!rc && S->cfg.tweener_enabled) {
                std::vector<TweenGroup> tw;
                for (size_t x : now) tw.insert(tw.end(), mine[x]->tweens.begin(), mine[x]->tweens.end());
                if (!tw.empty()) {
                    gxy_run_stats trs;
                    std::memset(&trs, 0, sizeof trs);
                    rc = tweener_groups_dev(ctx, tw, &S->cfg, &run_store, &trs);
                    std::lock_guard<std::mutex> lk(S->mu);
                    S->totals.dp_problems += trs.dp_problems;
                    S->totals.dp_cells += trs.dp_cells;
                    S->totals.hits += trs.hits;
                    S->totals.columns += trs.columns;
                }
                for (size_t x : now) mine[x]->tweens.clear();
            }
            for (size_t x : now) complete(S, ctx, mine[x], run_store, rc);
        };
        gxy_run_stats rs;
        std::memset(&rs, 0, sizeof rs);
        const int rc = gapped_groups_dev(ctx, all, &S->cfg, &run_store, &rs, [&] { settle(0); });
        {
            std::lock_guard<std::mutex> lk(S->mu);
            S->totals.dp_problems += rs.dp_problems;
            S->totals.dp_cells += rs.dp_cells;
        }
        settle(rc);
    }
}

}  // namespace

extern "C" {

int gxy_server_create(int device, const gxy_dseq *const *targets, uint64_t n_targets, const gxy_config *cfg,
                      const gxy_server_params *params, gxy_server **out) {
    return guarded(nullptr, [&]() -> int {
        if (!targets || !n_targets || !cfg || !params || !out) return GXY_EINVAL;
        for (uint64_t i = 0; i < n_targets; i++)
            if (!targets[i]) return GXY_EINVAL;
        if (cfg->step < 1) return GXY_EINVAL;
        auto S = std::make_unique<gxy_server>();
        int rc = gxy_ctx_create(device, &S->ctx);
        if (rc) return rc;
        S->targets.assign(targets, targets + n_targets);
        S->cfg = *cfg;
        S->prm = *params;
        S->t0 = std::chrono::steady_clock::now();
        auto fail = [&](int code) {
            S->chunks.clear();
            S->tables.clear();
            for (gxy_ctx *c : S->dp_ctx) gxy_ctx_destroy(c);
            gxy_ctx_destroy(S->ctx);
            S->ctx = nullptr;
            return code;
        };
        if (cfg->strand == GXY_STRAND_PLUS || cfg->strand == GXY_STRAND_BOTH) S->strands.push_back(GXY_STRAND_PLUS);
        if (cfg->strand == GXY_STRAND_MINUS || cfg->strand == GXY_STRAND_BOTH) S->strands.push_back(GXY_STRAND_MINUS);
        if (S->strands.empty()) return fail(GXY_EINVAL);
        rc = build_chunks(S->ctx, targets, n_targets, cfg, &S->chunks);
        if (rc) return fail(rc);
        // resident position tables, one per chunk
        for (const Chunk &ch : S->chunks) {
            gxy_postable *tp;
            rc = postable_build_impl(S->ctx, ch.seq, &cfg->pattern, cfg->step, cfg->respect_masking ? 1 : 0, &tp);
            if (rc) return fail(rc);
            S->tables.emplace_back(tp);
            if (cfg->max_word_count > 0) {
                rc = gxy_postable_prune(S->ctx, tp, cfg->max_word_count);
                if (rc) return fail(rc);
            }
        }
        GXY_HIP(S->ctx, hipStreamSynchronize(S->ctx->stream));
        // Separator between concatenated queries: an x-drop walk entering N bases
        // loses |ambig| per column and stops once x_drop is exceeded, and no seed
        // window contains an N; offsets keep the query windows on the step grid.
        // A window lies inside one query exactly when its first and last offsets
        // do, so concatenation needs a pattern whose first and last positions are
        // care positions: with a don't-care end a window could reach into the
        // separator with all its care positions inside a query, a window that the
        // query alone does not have (pos + span > len, src/seeds.rs:133-147).
        // Other patterns search each admitted query on its own.
        const uint32_t care_ends = cfg->pattern.length ? 1u | (1u << (cfg->pattern.length - 1)) : 1u;
        if (params->co_batch && cfg->scoring.ambig < 0 && cfg->hsp.x_drop >= 0 &&
            (cfg->pattern.care & care_ends) == care_ends) {
            const uint64_t walk = (uint64_t)cfg->hsp.x_drop / (uint64_t)(-(int64_t)cfg->scoring.ambig) + 2;
            S->sep = (uint32_t)std::max<uint64_t>(64, walk);
            S->align = (uint32_t)(32 * cfg->step / std::__gcd<uint64_t>(32, cfg->step));
        }
        // one search stream and three gapped workers by default: the device's
        // four hardware queues, one stream each
        const int nw = cfg->gapped_enabled && params->co_batch ? (params->dp_workers > 0 ? params->dp_workers : 3) : 0;
        size_t free_b = 0, total_b = 0;
        GXY_HIP(S->ctx, hipMemGetInfo(&free_b, &total_b));
        for (int w = 0; w < nw; w++) {
            gxy_ctx *c;
            rc = gxy_ctx_create(device, &c);
            if (rc) return fail(rc);
            // Together the workers' traceback pools may take half of the device
            // (each at most 48 GiB).  A pool grows only as far as its passes need,
            // and the longest queries all land in one worker's length bin.
            c->dp_pool_cap = std::min<uint64_t>(48ull << 30, total_b / (2 * (uint64_t)nw));
            S->dp_ctx.push_back(c);
        }
        S->worker = std::thread(search_loop, S.get());
        for (gxy_ctx *c : S->dp_ctx) S->dp_workers.emplace_back(dp_loop, S.get(), c);
        *out = S.release();
        return GXY_OK;
    });
}

int gxy_server_submit(gxy_server *S, const gxy_seq *query, uint64_t *ticket) {
    return guarded(nullptr, [&]() -> int {
        if (!S || !query || !ticket || (query->len && (!query->codes || !query->valid || !query->mask)))
            return GXY_EINVAL;
        auto r = std::make_unique<ServeReq>();
        r->len = query->len;
        r->codes.assign(query->codes, query->codes + (query->len + 3) / 4);
        r->valid.assign(query->valid, query->valid + (query->len + 7) / 8);
        r->mask.assign(query->mask, query->mask + (query->len + 7) / 8);
        {
            std::lock_guard<std::mutex> lk(S->mu);
            if (S->stop) return GXY_EINVAL;
            r->id = S->next_id++;
            r->submit_ms = S->now_ms();
            *ticket = r->id;
            S->queue.push_back(r.get());
            S->reqs.emplace(r->id, std::move(r));
        }
        S->cv_work.notify_one();
        return GXY_OK;
    });
}

int gxy_server_wait(gxy_server *S, uint64_t ticket, gxy_record_set **out, gxy_request_stats *stats) {
    return guarded(nullptr, [&]() -> int {
        if (!S || !out) return GXY_EINVAL;
        std::unique_lock<std::mutex> lk(S->mu);
        auto it = S->reqs.find(ticket);
        if (it == S->reqs.end() || it->second->claimed) return GXY_EINVAL;  // unknown, or waited on already
        ServeReq *r = it->second.get();
        r->claimed = true;
        S->cv_done.wait(lk, [&] { return r->done; });
        if (stats) {
            stats->submit_ms = r->submit_ms;
            stats->admit_ms = r->admit_ms;
            stats->done_ms = r->done_ms;
            stats->query_bp = r->len;
            stats->iteration = r->iteration;
            stats->batch_queries = r->batch_queries;
            stats->batch_bp = r->batch_bp;
        }
        const int rc = r->rc;
        if (rc) S->err = r->err;
        *out = r->out;
        r->out = nullptr;
        S->reqs.erase(it);
        return rc;
    });
}

const char *gxy_server_last_error(gxy_server *S) {
    if (!S) return "null server";
    // a copy per calling thread: worker threads may replace S->err at any time
    static thread_local std::string copy;
    std::lock_guard<std::mutex> lk(S->mu);
    copy = S->err;
    return copy.c_str();
}

int gxy_server_stats(gxy_server *S, gxy_run_stats *out) {
    return guarded(nullptr, [&]() -> int {
        if (!S || !out) return GXY_EINVAL;
        std::lock_guard<std::mutex> lk(S->mu);
        *out = S->totals;
        return GXY_OK;
    });
}

int gxy_server_set_timing(gxy_server *S, int enabled) {
    return guarded(nullptr, [&]() -> int {
        if (!S) return GXY_EINVAL;
        std::lock_guard<std::mutex> lk(S->mu);
        if (!S->queue.empty() || !S->dpq.empty() || !S->reqs.empty()) return GXY_EINVAL;  // requests in flight
        S->ctx->timing = enabled != 0;
        for (gxy_ctx *c : S->dp_ctx) c->timing = enabled != 0;
        return GXY_OK;
    });
}

int gxy_server_kernel_time(gxy_server *S, const char *name, double *ms, uint64_t *launches) {
    return guarded(nullptr, [&]() -> int {
        if (!S || !name || !ms || !launches) return GXY_EINVAL;
        std::lock_guard<std::mutex> lk(S->mu);
        if (!S->queue.empty() || !S->dpq.empty() || !S->reqs.empty()) return GXY_EINVAL;  // requests in flight
        *ms = 0;
        *launches = 0;
        std::vector<gxy_ctx *> all(1, S->ctx);
        all.insert(all.end(), S->dp_ctx.begin(), S->dp_ctx.end());
        for (gxy_ctx *c : all) {
            double m = 0;
            uint64_t l = 0;
            int rc = gxy_ctx_kernel_time(c, name, &m, &l);
            if (rc) return rc;
            *ms += m;
            *launches += l;
        }
        return GXY_OK;
    });
}

int gxy_server_xdrop_stats(gxy_server *S, uint64_t *first_step_columns, uint64_t *resumed,
                           uint64_t *first_step_passing) {
    return guarded(nullptr, [&]() -> int {
        if (!S || !first_step_columns || !resumed || !first_step_passing) return GXY_EINVAL;
        std::lock_guard<std::mutex> lk(S->mu);
        *first_step_columns = *resumed = *first_step_passing = 0;
        std::vector<gxy_ctx *> all(1, S->ctx);
        all.insert(all.end(), S->dp_ctx.begin(), S->dp_ctx.end());
        for (gxy_ctx *c : all) {
            *first_step_columns += c->xd_first_cols;
            *resumed += c->xd_conts;
            *first_step_passing += c->xd_first_pass;
        }
        return GXY_OK;
    });
}

int gxy_server_destroy(gxy_server *S) {
    return guarded(nullptr, [&]() -> int {
        if (!S) return GXY_EINVAL;
        {
            std::lock_guard<std::mutex> lk(S->mu);
            S->stop = true;
        }
        S->cv_work.notify_all();
        if (S->worker.joinable()) S->worker.join();
        S->cv_dp.notify_all();
        for (auto &t : S->dp_workers)
            if (t.joinable()) t.join();
        for (auto &kv : S->reqs) gxy_record_set_free(kv.second->out);
        S->reqs.clear();
        S->tofree.clear();
        S->tables.clear();
        S->chunks.clear();
        for (gxy_ctx *c : S->dp_ctx) gxy_ctx_destroy(c);
        gxy_ctx_destroy(S->ctx);
        delete S;
        return GXY_OK;
    });
}
}  // extern "C"

// ---- writers ---------------------------------------------------------------

namespace {

// Uppercase ASCII with N for invalid bases (PackedSeq::to_ascii,
// src/sequences.rs:268-272) of seq[start, start+len), optionally from the
// reverse complement of the whole sequence (src/sequences.rs:303-317).
void ascii_slice(const gxy_seq &s, bool rc, uint64_t start, uint64_t len, std::string &out) {
    static const char dec[4] = {'A', 'C', 'G', 'T'};
    out.resize(len);
    for (uint64_t k = 0; k < len; k++) {
        uint64_t i = start + k;
        if (rc) i = s.len - 1 - i;
        const bool v = (s.valid[i >> 3] >> (i & 7)) & 1;
        uint32_t c = (s.codes[i >> 2] >> (2 * (i & 3))) & 3u;
        if (rc) c = 3u - c;
        out[k] = v ? dec[c] : 'N';
    }
}

void append_u(std::string &o, uint64_t v) {
    char b[24];
    auto r = std::to_chars(b, b + sizeof b, v);
    o.append(b, r.ptr);
}
void append_i(std::string &o, int64_t v) {
    char b[24];
    auto r = std::to_chars(b, b + sizeof b, v);
    o.append(b, r.ptr);
}
// Rust `{:>w}` for integers: right-aligned, space padded, never truncated.
void append_right(std::string &o, uint64_t v, size_t w) {
    char b[24];
    auto r = std::to_chars(b, b + sizeof b, v);
    const size_t n = (size_t)(r.ptr - b);
    if (n < w) o.append(w - n, ' ');
    o.append(b, r.ptr);
}
// Rust `{:<w}` for strings: left-aligned, padded by characters.
void append_left(std::string &o, const char *s, size_t w) {
    const size_t n = std::strlen(s);
    o.append(s, n);
    // pad by Unicode scalar values like Rust's formatter
    size_t chars = 0;
    for (size_t k = 0; k < n; k++)
        if (((unsigned char)s[k] & 0xC0) != 0x80) chars++;
    if (chars < w) o.append(w - chars, ' ');
}

bool records_ok(uint64_t nt, uint64_t nq, const gxy_record_set *rs, const gxy_named_seq *ts,
                const gxy_named_seq *qs) {
    for (uint64_t k = 0; k < rs->n; k++) {
        const gxy_record &r = rs->records[k];
        if (r.target_index >= nt || r.query_index >= nq) return false;
        if (r.runs_offset + r.n_runs > rs->n_runs) return false;
        if ((uint64_t)r.t_start + r.t_span > ts[r.target_index].seq.len) return false;
        if ((uint64_t)r.q_start + r.q_span > qs[r.query_index].seq.len) return false;
        // the script must consume exactly the record's spans (EditScript
        // target_span / query_span, src/edit_script.rs:104-121)
        uint64_t tsum = 0, qsum = 0;
        for (uint32_t x = 0; x < r.n_runs; x++) {
            const gxy_edit_run &run = rs->runs[r.runs_offset + x];
            if (run.op > GXY_OP_DELETE) return false;
            tsum += run.op != GXY_OP_INSERT ? run.n : 0;
            qsum += run.op != GXY_OP_DELETE ? run.n : 0;
        }
        if (tsum != r.t_span || qsum != r.q_span) return false;
    }
    return true;
}

char *finish(const std::string &s, uint64_t *len) {
    char *p = (char *)std::malloc(s.size() + 1);
    if (!p) return nullptr;
    std::memcpy(p, s.data(), s.size());
    p[s.size()] = 0;
    *len = s.size();
    return p;
}

}  // namespace

extern "C" {

int gxy_format_maf(const gxy_named_seq *targets, uint64_t n_targets, const gxy_named_seq *queries,
                   uint64_t n_queries, const gxy_record_set *records, char **text, uint64_t *len) {
    return guarded(nullptr, [&]() -> int {
        if (!records || !text || !len || !records_ok(n_targets, n_queries, records, targets, queries))
            return GXY_EINVAL;
        std::string o, ta, qa, tt, qq;
        if (records->n > 0) o += "##maf version=1 program=lastz-gxy\n# scoring=HOXD70\n";
        for (uint64_t k = 0; k < records->n; k++) {
            const gxy_record &r = records->records[k];
            ascii_slice(targets[r.target_index].seq, false, r.t_start, r.t_span, ta);
            ascii_slice(queries[r.query_index].seq, r.strand == GXY_STRAND_MINUS, r.q_start, r.q_span, qa);
            // render_aligned_bases (src/edit_script.rs:140-172)
            tt.clear();
            qq.clear();
            size_t ti = 0, qi = 0;
            for (uint32_t x = 0; x < r.n_runs; x++) {
                const gxy_edit_run &run = records->runs[r.runs_offset + x];
                if (run.op == GXY_OP_MATCH) {
                    tt.append(ta, ti, run.n);
                    qq.append(qa, qi, run.n);
                    ti += run.n;
                    qi += run.n;
                } else if (run.op == GXY_OP_INSERT) {
                    tt.append(run.n, '-');
                    qq.append(qa, qi, run.n);
                    qi += run.n;
                } else {
                    tt.append(ta, ti, run.n);
                    qq.append(run.n, '-');
                    ti += run.n;
                }
            }
            o += "a score=";
            append_i(o, r.score);
            o += "\ns ";
            append_left(o, targets[r.target_index].name, 20);
            o += ' ';
            append_right(o, r.t_start, 10);
            o += ' ';
            append_right(o, r.t_span, 6);
            o += " + ";
            append_right(o, (uint32_t)targets[r.target_index].seq.len, 10);
            o += ' ';
            o += tt;
            o += "\ns ";
            append_left(o, queries[r.query_index].name, 20);
            o += ' ';
            append_right(o, r.q_start, 10);
            o += ' ';
            append_right(o, r.q_span, 6);
            o += r.strand == GXY_STRAND_MINUS ? " - " : " + ";
            append_right(o, (uint32_t)queries[r.query_index].seq.len, 10);
            o += ' ';
            o += qq;
            o += "\n\n";
        }
        *text = finish(o, len);
        return *text ? GXY_OK : GXY_ENOMEM;
    });
}

int gxy_format_paf(const gxy_named_seq *targets, uint64_t n_targets, const gxy_named_seq *queries,
                   uint64_t n_queries, const gxy_record_set *records, char **text, uint64_t *len) {
    return guarded(nullptr, [&]() -> int {
        if (!records || !text || !len || !records_ok(n_targets, n_queries, records, targets, queries))
            return GXY_EINVAL;
        std::string o, ta, qa, cig;
        for (uint64_t k = 0; k < records->n; k++) {
            const gxy_record &r = records->records[k];
            ascii_slice(targets[r.target_index].seq, false, r.t_start, r.t_span, ta);
            ascii_slice(queries[r.query_index].seq, r.strand == GXY_STRAND_MINUS, r.q_start, r.q_span, qa);
            // score_script (src/output/paf.rs:59-89); positions past a slice are None
            uint64_t mt = 0, mm = 0, ind = 0;
            size_t ti = 0, qi = 0;
            cig.clear();
            for (uint32_t x = 0; x < r.n_runs; x++) {
                const gxy_edit_run &run = records->runs[r.runs_offset + x];
                append_u(cig, run.n);
                cig += run.op == GXY_OP_MATCH ? 'M' : (run.op == GXY_OP_INSERT ? 'I' : 'D');
                if (run.op == GXY_OP_MATCH) {
                    for (uint32_t c = 0; c < run.n; c++, ti++, qi++) {
                        const bool th = ti < ta.size(), qh = qi < qa.size();
                        const bool eq = (th && qh) ? ta[ti] == qa[qi] : th == qh;
                        if (eq) mt++; else mm++;
                    }
                } else if (run.op == GXY_OP_INSERT) {
                    ind += run.n;
                    qi += run.n;
                } else {
                    ind += run.n;
                    ti += run.n;
                }
            }
            o += queries[r.query_index].name;
            o += '\t';
            append_u(o, (uint32_t)queries[r.query_index].seq.len);
            o += '\t';
            append_u(o, r.q_start);
            o += '\t';
            append_u(o, (uint64_t)r.q_start + r.q_span);
            o += r.strand == GXY_STRAND_MINUS ? "\t-\t" : "\t+\t";
            o += targets[r.target_index].name;
            o += '\t';
            append_u(o, (uint32_t)targets[r.target_index].seq.len);
            o += '\t';
            append_u(o, r.t_start);
            o += '\t';
            append_u(o, (uint64_t)r.t_start + r.t_span);
            o += '\t';
            append_u(o, mt);
            o += '\t';
            append_u(o, mt + mm + ind);
            o += "\t60\tNM:i:";
            append_u(o, mm + ind);
            o += "\tAS:i:";
            append_i(o, r.score);
            o += "\ttp:A:P\tcg:Z:";
            o += cig;
            o += '\n';
        }
        *text = finish(o, len);
        return *text ? GXY_OK : GXY_ENOMEM;
    });
}

}  // extern "C"
