/*
 * oracle.h — CPU restatement of lastz-gxy's seed -> HSP -> gapped path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * library declared in include/gxy.h and the CPU baseline timed beside it.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * it.  Nothing in lastz-gxy_amd/ links or calls it.
 *
 * Each function restates the reference function named in its comment
 * (paths relative to the reference crate).  The reference (Rust) cannot be
 * compiled in this environment, so the restatement is pinned by the
 * reference's own known-answer tests, closed forms, invariants and brute
 * force (tests/test_oracle_*.py).
 */
#ifndef GXY_ORACLE_H
#define GXY_ORACLE_H

#include "../include/gxy.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gxo_postable gxo_postable;

/* PackedSeq::from_ascii (src/sequences.rs:164-206). */
int gxo_pack_ascii(const char *ascii, uint64_t len, uint8_t *codes, uint8_t *valid,
                   uint8_t *mask);
/* PackedSeq::reverse_complement (src/sequences.rs:303-317) into buffers. */
int gxo_reverse_complement(const gxy_seq *in, uint8_t *codes, uint8_t *valid, uint8_t *mask);
/* PackedSeq::to_ascii (src/sequences.rs:268-272); out has len bytes. */
int gxo_to_ascii(const gxy_seq *in, char *out);

/* SeedExtractor::word_at (src/seeds.rs:133-147). */
int gxo_seed_word(const gxy_seq *seq, const gxy_seed_pattern *pattern, uint64_t pos,
                  uint64_t *word);

/* PosTable::build / prune_hot_words / lookup (src/pos_table.rs:35-136). */
int gxo_postable_build(const gxy_seq *seq, const gxy_seed_pattern *pattern, uint64_t step,
                       gxo_postable **out);
int gxo_postable_prune(gxo_postable *t, uint32_t limit);
uint64_t gxo_postable_num_words(const gxo_postable *t);
uint64_t gxo_postable_total_positions(const gxo_postable *t);
const uint32_t *gxo_postable_starts(const gxo_postable *t);
const uint32_t *gxo_postable_positions(const gxo_postable *t);
void gxo_postable_free(gxo_postable *t);

/* extend_hit_scalar (src/hsp.rs:98-148) when simd == 0, extend_hit_avx2
 * (src/hsp_simd.rs:27-95) when simd != 0 and the CPU has AVX2. */
int gxo_extend_hit(const gxy_seq *t, const gxy_seq *q, const gxy_scoring *sc,
                   const gxy_hsp_params *p, uint32_t t_pos, uint32_t q_pos, uint32_t seed_len,
                   int simd, gxy_hsp *out, int *found);

/* search (src/seed_search.rs:45-125); output sorted stably by
 * (diagonal, t_start).  *out freed with gxo_free. */
int gxo_search(const gxo_postable *table, const gxy_seq *t, const gxy_seq *q,
               const gxy_scoring *sc, const gxy_search_params *p, int simd, gxy_hsp **out,
               uint64_t *n, gxy_search_stats *stats);

/* choose_anchor (src/anchor.rs:25-69). */
int gxo_choose_anchor(const gxy_hsp *hsp, const gxy_seq *t, const gxy_seq *q,
                      const gxy_scoring *sc, uint32_t window, gxy_anchor *out);

/* gapped_extend::extend (src/gapped_extend.rs:78-463).  lazy == 0 allocates
 * (n+1) x band_width cells per side as the reference does; lazy != 0 grows
 * the rows as the sweep reaches them (identical output).  *runs freed with
 * gxo_free; *cells receives the number of DP cells evaluated. */
int gxo_gapped_extend(const gxy_anchor *a, const gxy_seq *t, const gxy_seq *q,
                      const gxy_scoring *sc, const gxy_gapped_params *p, int lazy,
                      gxy_alignment *out, gxy_edit_run **runs, uint64_t *cells, int *found);

/* shannon_entropy / passes_entropy_gate (src/entropy.rs:27-66). */
double gxo_shannon_entropy(const gxy_seq *seq, uint64_t start, uint64_t end);
int gxo_passes_entropy_gate(const gxy_seq *seq, uint32_t t_start, uint32_t t_end,
                            int32_t score, int32_t threshold);

/* best_chain (src/chain.rs:19-69); `out` has room for n HSPs. */
int gxo_best_chain(const gxy_hsp *hsps, uint64_t n, gxy_hsp *out, uint64_t *n_out);
/* tweener::interpolate (src/tweener.rs:59-185) over the records of one
 * (target, query, strand) group; `target` whole, `query` the group's strand;
 * cfg's inner_* fields are the TweenerConfig.  Free with gxo_record_set_free. */
int gxo_interpolate(const gxy_record_set *in, const gxy_seq *target, const gxy_seq *query,
                    const gxy_config *cfg, gxy_record_set **out);
/* driver::run (src/driver.rs:199-392) over `threads` worker threads (0 = all
 * cores).  Chunk groups are merged in ascending chunk order. */
int gxo_run(const gxy_named_seq *targets, uint64_t n_targets, const gxy_named_seq *queries,
            uint64_t n_queries, const gxy_config *cfg, int threads, int lazy, int simd,
            gxy_record_set **out, gxy_run_stats *stats);
/* MafWriter / PafWriter (src/output/maf.rs:33-74, src/output/paf.rs:23-89). */
int gxo_format_maf(const gxy_named_seq *targets, uint64_t n_targets,
                   const gxy_named_seq *queries, uint64_t n_queries,
                   const gxy_record_set *records, char **text, uint64_t *len);
int gxo_format_paf(const gxy_named_seq *targets, uint64_t n_targets,
                   const gxy_named_seq *queries, uint64_t n_queries,
                   const gxy_record_set *records, char **text, uint64_t *len);
void gxo_record_set_free(gxy_record_set *s);
void gxo_free(void *p);

#ifdef __cplusplus
}
#endif
#endif
