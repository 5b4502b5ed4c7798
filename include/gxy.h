/*
 * gxy.h — C-ABI of the MI355X (gfx950) seed -> HSP -> gapped hot path.
 *
 * Every entry point replaces one public Rust function of lastz-gxy that the
 * reference driver (src/driver.rs) calls on its hot path; the comment above
 * each declaration names the function it replaces (file:line, relative to the
 * reference crate).  Plain pointers and sizes only: no C++ or torch types.
 *
 * Conventions
 *  - Return value: 0 on success, a negative GXY_E* code on failure.  No
 *    exception or abort crosses the ABI; the message of the last failure on a
 *    context is returned by gxy_last_error().
 *  - Option<T> in Rust becomes an out-flag `found` (1 = Some, 0 = None).
 *  - Inputs are borrowed for the duration of a call.  Objects and arrays the
 *    library returns are freed with the matching gxy_*_free / gxy_free call,
 *    never with the caller's allocator.
 *  - One gxy_ctx owns one HIP stream and its scratch buffers.  Calls on one
 *    context are serialised by the caller (one context per host thread);
 *    distinct contexts may be used concurrently.
 *  - Sequences use the reference's PackedSeq layout (src/sequences.rs:128-141):
 *    2-bit codes A=0 C=1 G=2 T=3, four bases per byte, base i at bits
 *    2*(i&3) of byte i>>2; `valid` and `mask` are one bit per base, base i at
 *    bit (i&7) of byte i>>3.  N/ambiguous bases have code 0, valid 0, mask 0.
 */
#ifndef GXY_H
#define GXY_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define GXY_OK 0
#define GXY_EINVAL (-1)   /* invalid argument (bad size, null pointer, ...)  */
#define GXY_ENOMEM (-2)   /* host or device allocation failed                */
#define GXY_EDEVICE (-3)  /* HIP runtime / kernel launch error               */
#define GXY_EUNSUP (-4)   /* parameters outside what a kernel supports (kept for the ABI; no entry point returns it now) */
#define GXY_ERANGE (-5)   /* input exceeds a documented limit                */

/* ---- POD views ---------------------------------------------------------- */

/* Borrowed view of a PackedSeq (src/sequences.rs:128-141). */
typedef struct gxy_seq {
    const uint8_t *codes; /* (len+3)/4 bytes */
    const uint8_t *valid; /* (len+7)/8 bytes */
    const uint8_t *mask;  /* (len+7)/8 bytes */
    uint64_t len;
} gxy_seq;

/* SeedPattern (src/seeds.rs:20-23): `length` positions (1..32); bit k of
 * `care` is set when offset k is a care position. */
typedef struct gxy_seed_pattern {
    uint32_t length;
    uint32_t care;
} gxy_seed_pattern;

/* ScoringMatrix (src/scoring.rs:32-40): sub[target][query]. */
typedef struct gxy_scoring {
    int32_t sub[4][4];
    int32_t gap_open;
    int32_t gap_extend;
    int32_t ambig;
} gxy_scoring;

/* Hsp (src/hsp.rs:19-24), same field order. */
typedef struct gxy_hsp {
    uint32_t t_start;
    uint32_t q_start;
    uint32_t length;
    int32_t score;
} gxy_hsp;

/* HspParams (src/hsp.rs:46-49). */
typedef struct gxy_hsp_params {
    int32_t x_drop;
    int32_t hsp_threshold;
} gxy_hsp_params;

/* SearchParams (src/seed_search.rs:19-30).  `entropy_threshold` is used
 * only when `entropy_enabled` is non-zero (Option<i32>). */
typedef struct gxy_search_params {
    uint64_t step;
    gxy_hsp_params hsp;
    uint32_t transitions;
    int32_t entropy_enabled;
    int32_t entropy_threshold;
} gxy_search_params;

/* Anchor (src/anchor.rs:16-20). */
typedef struct gxy_anchor {
    uint32_t t_pos;
    uint32_t q_pos;
    gxy_hsp source;
} gxy_anchor;

/* GappedParams (src/gapped_extend.rs:28-31). */
typedef struct gxy_gapped_params {
    int32_t y_drop;
    int32_t gapped_threshold;
} gxy_gapped_params;

/* EditOp (src/edit_script.rs:17-22) run: op 0 = Match (M), 1 = InsertQuery
 * (I, gap in target), 2 = DeleteQuery (D, gap in query). */
#define GXY_OP_MATCH 0u
#define GXY_OP_INSERT 1u
#define GXY_OP_DELETE 2u
typedef struct gxy_edit_run {
    uint32_t op;
    uint32_t n;
} gxy_edit_run;

/* GappedAlignment (src/gapped_extend.rs:55-63).  The edit script is the run
 * range [runs_offset, runs_offset + n_runs) of the owning result's run array. */
typedef struct gxy_alignment {
    uint32_t t_start;
    uint32_t q_start;
    uint32_t t_len;
    uint32_t q_len;
    int32_t score;
    uint32_t n_runs;
    uint64_t runs_offset;
    gxy_anchor source;
} gxy_alignment;

/* Result of a batched gapped extension: found[k] says whether alignment k is
 * Some; when it is None, alns[k] is zeroed and owns no runs. */
typedef struct gxy_alignment_set {
    uint64_t n;
    gxy_alignment *alns;
    uint8_t *found;
    uint64_t n_runs;
    gxy_edit_run *runs;
    uint64_t cells; /* DP cells evaluated over the whole batch (both sides) */
} gxy_alignment_set;

/* Counters of one search call (algorithmic work, for GCUPS and roofline). */
typedef struct gxy_search_stats {
    uint64_t query_windows;  /* valid query seed windows                     */
    uint64_t lookups;        /* variant-word lookups                         */
    uint64_t hits;           /* seed hits returned by the lookups            */
    uint64_t extended;       /* hits extended (GPU: every hit, speculatively)*/
    uint64_t columns;        /* x-drop columns evaluated (of every hit)      */
    uint64_t passing;        /* extensions at or above hsp_threshold         */
    uint64_t emitted;        /* HSPs after diagonal admission, before dedup  */
    uint64_t tie_pairs;      /* adjacent equal-(diag,t_start) unequal HSPs   */
} gxy_search_stats;

/* ---- opaque handles ---------------------------------------------------- */
typedef struct gxy_ctx gxy_ctx;           /* HIP stream + scratch + error   */
typedef struct gxy_dseq gxy_dseq;         /* device-resident PackedSeq      */
typedef struct gxy_postable gxy_postable; /* device-resident PosTable       */

/* ---- context ----------------------------------------------------------- */
/* Create a context on HIP device `device` (its own stream). */
int gxy_ctx_create(int device, gxy_ctx **out);
void gxy_ctx_destroy(gxy_ctx *ctx);
/* Message of the last failed call on `ctx` ("" if none).  With ctx == NULL,
 * the message of the last failed gxy_ctx_create on this thread. */
const char *gxy_last_error(const gxy_ctx *ctx);
/* Wait for all work queued on the context's stream. */
int gxy_ctx_synchronize(gxy_ctx *ctx);
/* Library version string. */
const char *gxy_version(void);
/* Kernel timing (no counterpart in the reference; for measurement only).
 * When enabled, the library brackets its main kernels with HIP events on the
 * context's stream; gxy_ctx_kernel_time returns the summed milliseconds and
 * launch count of kernel `name` since the last reset (it synchronises the
 * stream).  Names: "table_items", "table_sort", "table_starts", "qcount",
 * "qexpand", "extend", "pass_sort", "admit", "dp_forward", "dp_traceback". */
int gxy_ctx_set_timing(gxy_ctx *ctx, int enabled);
int gxy_ctx_kernel_time(gxy_ctx *ctx, const char *name, double *ms, uint64_t *launches);
int gxy_ctx_reset_timing(gxy_ctx *ctx);
/* Traceback-pool cap of the gapped extension (no counterpart in the
 * reference, whose DP keeps every row of a side: src/gapped_extend.rs:233-242).
 * The library keeps the from-states of the rows being swept in a pool of
 * 64-row chunks, by default up to a third of the device (at most 96 GiB).
 * With rows > 0 the pool holds at most max(16, rows / 64) chunks: sides that
 * find the pool empty stop and are swept again in a later pass, a pass that
 * finishes nothing is followed by one that runs its first side alone, and a
 * side that needs more rows than the whole pool fails with GXY_ENOMEM.  Results
 * do not depend on the cap.  rows = 0 restores the default. */
int gxy_ctx_set_dp_pool_rows(gxy_ctx *ctx, uint64_t rows);
/* Returns the context's cached device memory to the device (no counterpart in
 * the reference): the scratch buffers that calls grow and keep for the next
 * call — the traceback pool among them, up to a third of the device after a
 * large gapped batch — and freed sequence blocks kept for reuse.  Sequences,
 * tables and results stay valid; the next call allocates what it needs again. */
int gxy_ctx_release_cache(gxy_ctx *ctx);
/* DP forward kernel of the gapped extension (no counterpart in the reference;
 * for measurement and tests).  Every kernel gives identical results; they
 * differ in how a side's row is spread over the GPU: GXY_DP_BLOCK4 runs each
 * side on four wavefronts with the row in LDS, GXY_DP_WAVE_K* on one
 * wavefront with the row in registers, 1, 2 or 4 columns per lane per slot.
 * GXY_DP_AUTO (the default) lets the library choose. */
#define GXY_DP_AUTO 0
#define GXY_DP_BLOCK4 1
#define GXY_DP_WAVE_K1 2
#define GXY_DP_WAVE_K2 3
#define GXY_DP_WAVE_K4 4
#define GXY_DP_WAVE_K2_W4 5 /* K2 / K4 compiled for four wavefronts per SIMD */
#define GXY_DP_WAVE_K4_W4 6
int gxy_ctx_set_dp_kernel(gxy_ctx *ctx, int kernel);
/* The autotuner behind GXY_DP_AUTO (no counterpart in the reference).  Passes
 * of the gapped extension fall in four classes by the number of sides they
 * sweep together (<= 8, <= 512, <= 8192, more); for each class the context
 * measures the candidate kernels (GXY_DP_WAVE_K2_W4, GXY_DP_WAVE_K4_W4,
 * GXY_DP_WAVE_K2, GXY_DP_WAVE_K4, GXY_DP_BLOCK4) on its first passes of that
 * class (big passes lend a slice of their sides to each probe) and keeps the
 * fastest.  chosen[4] receives each class's kernel (0 while probing) and
 * gcups[4 * 5] the forward kernel's measured rate of each candidate, in the
 * order above (0 where not yet measured).  Results never depend on it. */
int gxy_ctx_dp_autotune_state(gxy_ctx *ctx, int *chosen, double *gcups);
/* Gapped-extension passes on the context since it was created: passes run,
 * sides stopped by a full pool or by their reserved traceback rows (swept
 * again later) and passes that ran one side alone. */
int gxy_ctx_dp_pool_stats(gxy_ctx *ctx, uint64_t *passes, uint64_t *truncated, uint64_t *alone);
/* Persistent form of the gapped extension (no counterpart in the reference,
 * whose driver extends anchors one at a time: src/driver.rs:305-331).  Once
 * the forward kernel of a batch is a one-wavefront variant (GXY_DP_WAVE_K2,
 * _K4, _K2_W4, _K4_W4), its sides run as one launch of as many wavefronts as
 * the GPU holds: each takes sides from a device queue (longest first), sweeps
 * a side into its own share of the traceback pool, walks it back at once and
 * appends its edit runs, with no host work between sides.  Sides that
 * outgrow a share go on through passes.  mode 0 (the default) does this for
 * batches of 16384 sides or more, 1 for every batch, -1 never.  Results do
 * not depend on it. */
int gxy_ctx_set_dp_persist(gxy_ctx *ctx, int mode);
/* Persistent launches on the context since creation, sides given to them,
 * sides they left to passes and the DP cells of the sides they finished. */
int gxy_ctx_dp_persist_stats(gxy_ctx *ctx, uint64_t *launches, uint64_t *sides, uint64_t *deferred,
                             uint64_t *cells);
/* Form of the search's x-drop extension (no counterpart in the reference,
 * which extends one hit at a time: src/seed_search.rs:98-112).  0 (the
 * default): a straight-line first step takes the seed and the first 32
 * columns of both walks of every hit, and a persistent kernel resumes the
 * hits whose walks go on; 1: the persistent kernel alone, one 32-column chunk
 * of one walk per lane per step.  Results do not depend on it. */
int gxy_ctx_set_xdrop_form(gxy_ctx *ctx, int form);
/* Seed hits per expansion batch of a search (no counterpart in the reference,
 * which extends each hit as it is found: src/seed_search.rs:95-113).  A
 * search whose hits exceed the batch runs them in several batches of whole
 * query-window blocks, each sized to about `hits` (0: the default, 2^27), which
 * bounds its scratch memory to a few tens of bytes per batch hit.  Results do
 * not depend on it. */
int gxy_ctx_set_hit_batch(gxy_ctx *ctx, uint64_t hits);
/* First steps of the x-drop (form 0) since the context was created, the
 * minus-strand helper's included: columns they evaluated, hits they handed on
 * to be resumed and passing extensions they finished. */
int gxy_ctx_xdrop_stats(gxy_ctx *ctx, uint64_t *first_step_columns, uint64_t *resumed, uint64_t *first_step_passing);

/* ---- sequences ---------------------------------------------------------- */
/* PackedSeq::from_ascii (src/sequences.rs:164-206) into caller buffers sized
 * as in gxy_seq.  Host-only utility; needs no context. */
int gxy_pack_ascii(const char *ascii, uint64_t len, uint8_t *codes, uint8_t *valid,
                   uint8_t *mask);
/* Upload a PackedSeq to device memory (kept resident until gxy_dseq_free). */
int gxy_dseq_upload(gxy_ctx *ctx, const gxy_seq *seq, gxy_dseq **out);
/* Device-side PackedSeq::reverse_complement (src/sequences.rs:303-317). */
int gxy_dseq_reverse_complement(gxy_ctx *ctx, const gxy_dseq *seq, gxy_dseq **out);
/* Device-side PackedSeq::slice_to_new (src/sequences.rs:277-298). */
int gxy_dseq_slice(gxy_ctx *ctx, const gxy_dseq *seq, uint64_t start, uint64_t end,
                   gxy_dseq **out);
uint64_t gxy_dseq_len(const gxy_dseq *seq);
/* Copy a device PackedSeq back into caller buffers sized as in gxy_seq. */
int gxy_dseq_download(gxy_ctx *ctx, const gxy_dseq *seq, uint8_t *codes, uint8_t *valid,
                      uint8_t *mask);
void gxy_dseq_free(gxy_dseq *seq);

/* ---- position table ---------------------------------------------------- */
/* PosTable::build(&PackedSeq, &SeedPattern, step) (src/pos_table.rs:35). */
int gxy_postable_build(gxy_ctx *ctx, const gxy_seq *target, const gxy_seed_pattern *pattern,
                       uint64_t step, gxy_postable **out);
/* Same, from a device-resident target (the table keeps a reference to it:
 * free the table before the sequence). */
int gxy_postable_build_dev(gxy_ctx *ctx, const gxy_dseq *target,
                           const gxy_seed_pattern *pattern, uint64_t step,
                           gxy_postable **out);
/* PosTable::prune_hot_words(&mut self, limit) (src/pos_table.rs:113). */
int gxy_postable_prune(gxy_ctx *ctx, gxy_postable *table, uint32_t limit);
/* PosTable::lookup(&self, word) -> &[u32] (src/pos_table.rs:102).  Writes up to
 * `cap` positions into `out`; `*n_out` always receives the full count. */
int gxy_postable_lookup(gxy_ctx *ctx, const gxy_postable *table, uint64_t word, uint32_t *out,
                        uint64_t cap, uint64_t *n_out);
/* PosTable::total_positions (src/pos_table.rs:95). */
uint64_t gxy_postable_total_positions(const gxy_postable *table);
/* Copy `starts` (4^weight + 1 entries) and `positions` to host buffers of
 * the given capacities (either pointer may be NULL to skip it). */
int gxy_postable_export(gxy_ctx *ctx, const gxy_postable *table, uint32_t *starts,
                        uint64_t starts_cap, uint32_t *positions, uint64_t positions_cap);
void gxy_postable_free(gxy_postable *table);

/* ---- seed search -------------------------------------------------------- */
/* search(&PosTable, &PackedSeq target, &PackedSeq query, &ScoringMatrix,
 *        &SearchParams) -> Vec<Hsp>  (src/seed_search.rs:45-51).
 * `target` must be the sequence the table was built from.  Output sorted by
 * (diagonal, t_start); equal keys keep seed-scan order.  `*out` is freed with
 * gxy_free.  `stats` may be NULL. */
int gxy_search(gxy_ctx *ctx, const gxy_postable *table, const gxy_seq *target,
               const gxy_seq *query, const gxy_scoring *scoring,
               const gxy_search_params *params, gxy_hsp **out, uint64_t *n_out,
               gxy_search_stats *stats);
/* Same with a device-resident query; the target is the table's own. */
int gxy_search_dev(gxy_ctx *ctx, const gxy_postable *table, const gxy_dseq *query,
                   const gxy_scoring *scoring, const gxy_search_params *params,
                   gxy_hsp **out, uint64_t *n_out, gxy_search_stats *stats);

/* ---- ungapped x-drop ---------------------------------------------------- */
/* extend_hit(&PackedSeq, &PackedSeq, &ScoringMatrix, &HspParams, t_pos, q_pos,
 *            seed_len) -> Option<Hsp>  (src/hsp.rs:66-74). */
int gxy_extend_hit(gxy_ctx *ctx, const gxy_seq *target, const gxy_seq *query,
                   const gxy_scoring *scoring, const gxy_hsp_params *params, uint32_t t_pos,
                   uint32_t q_pos, uint32_t seed_len, gxy_hsp *out, int *found);
/* Batched extend_hit over n hits of one (target, query) pair. */
int gxy_extend_hit_batch(gxy_ctx *ctx, const gxy_seq *target, const gxy_seq *query,
                         const gxy_scoring *scoring, const gxy_hsp_params *params,
                         const uint32_t *t_pos, const uint32_t *q_pos, uint64_t n,
                         uint32_t seed_len, gxy_hsp *out, uint8_t *found);
/* gxy_extend_hit_batch without device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, kernels and results as gxy_extend_hit_batch; every device
 * buffer it needs (the staged sequences, their block and byte layouts, the
 * hit positions and results) lives in a 64 MiB array that the library holds
 * on each device it runs on until the process ends.  Inputs that need more
 * than that (about 2 bytes per target and query base plus 25 bytes per hit)
 * fail with GXY_ERANGE.  Calls on one device share the array: the library
 * serialises them (a per-device lock held for the whole call, which returns
 * only when the device is done with the array). */
int gxy_extend_hit_batch_noalloc(gxy_ctx *ctx, const gxy_seq *target, const gxy_seq *query,
                                 const gxy_scoring *scoring, const gxy_hsp_params *params,
                                 const uint32_t *t_pos, const uint32_t *q_pos, uint64_t n,
                                 uint32_t seed_len, gxy_hsp *out, uint8_t *found);

/* ---- anchor ------------------------------------------------------------- */
/* choose_anchor(Hsp, &PackedSeq, &PackedSeq, &ScoringMatrix, window) -> Anchor
 * (src/anchor.rs:25-31). */
int gxy_choose_anchor(gxy_ctx *ctx, const gxy_hsp *hsp, const gxy_seq *target,
                      const gxy_seq *query, const gxy_scoring *scoring, uint32_t window,
                      gxy_anchor *out);
/* Batched choose_anchor over n HSPs of one (target, query) pair. */
int gxy_choose_anchor_batch(gxy_ctx *ctx, const gxy_hsp *hsps, uint64_t n,
                            const gxy_seq *target, const gxy_seq *query,
                            const gxy_scoring *scoring, uint32_t window, gxy_anchor *out);

/* ---- gapped extension --------------------------------------------------- */
/* gapped_extend::extend(Anchor, &PackedSeq, &PackedSeq, &ScoringMatrix,
 *                       &GappedParams) -> Option<GappedAlignment>
 * (src/gapped_extend.rs:78-84).  Writes the edit script into `runs` (capacity
 * `runs_cap`); `out->n_runs` is the full run count even when it exceeds the
 * capacity (then the call returns GXY_ERANGE). */
int gxy_gapped_extend(gxy_ctx *ctx, const gxy_anchor *anchor, const gxy_seq *target,
                      const gxy_seq *query, const gxy_scoring *scoring,
                      const gxy_gapped_params *params, gxy_alignment *out, gxy_edit_run *runs,
                      uint64_t runs_cap, int *found);
/* Batched extend over n anchors of one (target, query) pair. */
int gxy_gapped_extend_batch(gxy_ctx *ctx, const gxy_anchor *anchors, uint64_t n,
                            const gxy_seq *target, const gxy_seq *query,
                            const gxy_scoring *scoring, const gxy_gapped_params *params,
                            gxy_alignment_set **out);
/* Batched extend over n independent (target_k, query_k, anchor_k) problems
 * held in device memory: targets[k]/queries[k] are device sequences. */
int gxy_gapped_extend_pairs_dev(gxy_ctx *ctx, const gxy_dseq *const *targets,
                                const gxy_dseq *const *queries, const gxy_anchor *anchors,
                                uint64_t n, const gxy_scoring *scoring,
                                const gxy_gapped_params *params, gxy_alignment_set **out);
/* The same for n pairs that are slices of two device sequences: pair k is
 * target[t_off[k], t_off[k] + t_len[k]) against query[q_off[k], q_off[k] +
 * q_len[k]), anchors[k] is relative to those slices, and so are the results.
 * (Many short pairs need no device sequence each.)  Each slice is shorter
 * than 2^32 bases; the sequences may be longer. */
int gxy_gapped_extend_slices_dev(gxy_ctx *ctx, const gxy_dseq *target, const gxy_dseq *query,
                                 const uint64_t *t_off, const uint64_t *t_len, const uint64_t *q_off,
                                 const uint64_t *q_len, const gxy_anchor *anchors, uint64_t n,
                                 const gxy_scoring *scoring, const gxy_gapped_params *params,
                                 gxy_alignment_set **out);
void gxy_alignment_set_free(gxy_alignment_set *set);

/* ---- driver (src/driver.rs:50-111, 199-392) ----------------------------- */
typedef struct gxy_named_seq {
    const char *name;
    gxy_seq seq;
} gxy_named_seq;

#define GXY_STRAND_PLUS 0u
#define GXY_STRAND_MINUS 1u
#define GXY_STRAND_BOTH 2u

/* Sequences loaded straight to the device layout, with their names. */
typedef struct gxy_dseq_list {
    uint64_t n;
    char **names;      /* NUL-terminated */
    gxy_dseq **seqs;
} gxy_dseq_list;
/* load_2bit / parse_2bit (src/bit2.rs:73-164) over the bytes of a .2bit file:
 * either byte order, N blocks invalid, mask blocks soft-masked. */
int gxy_load_2bit(gxy_ctx *ctx, const uint8_t *bytes, uint64_t len, gxy_dseq_list **out);
/* load_fasta / parse_fasta (src/sequences.rs:338-377) over FASTA text. */
int gxy_load_fasta(gxy_ctx *ctx, const char *text, uint64_t len, gxy_dseq_list **out);
/* Frees the names and the device sequences.  A caller that keeps the
 * sequences frees each name with gxy_free, sets n to 0, then frees the list. */
void gxy_dseq_list_free(gxy_dseq_list *list);

/* best_chain(&[Hsp]) -> Vec<Hsp> (src/chain.rs:19-69): the highest-scoring
 * chain strictly increasing on both axes, ascending; `out` holds n HSPs.
 * Host-side (O(n log n)); needs no context. */
int gxy_best_chain(const gxy_hsp *hsps, uint64_t n, gxy_hsp *out, uint64_t *n_out);

/* driver::Config (src/driver.rs:50-89).  tweener_enabled selects
 * Config::tweener = Some(TweenerConfig { pattern: inner_pattern, min_gap:
 * inner_min_gap, max_gap: inner_max_gap, relax_factor: inner_relax })
 * (src/tweener.rs:25-52, --inner / --inner-seed in src/cli.rs:68-80). */
typedef struct gxy_config {
    gxy_seed_pattern pattern;
    gxy_scoring scoring;
    uint64_t step;
    gxy_hsp_params hsp;
    gxy_gapped_params gapped;
    uint32_t strand;
    uint32_t max_word_count;
    int32_t respect_masking;
    int32_t gapped_enabled;
    int32_t chain_enabled;
    int32_t tweener_enabled;
    uint32_t anchor_window;
    uint32_t transitions;
    uint32_t chunk_primary_bp;
    uint32_t chunk_halo_bp;
    int32_t entropy_enabled;
    int32_t entropy_threshold;
    gxy_seed_pattern inner_pattern; /* default match12 */
    uint32_t inner_min_gap;         /* default 100     */
    uint32_t inner_max_gap;         /* default 100000  */
    int32_t inner_relax;            /* default 2       */
} gxy_config;

/* Record (src/output/mod.rs:17-31) without the owned strings: names and base
 * slices are taken from the input sequences when rendering. */
typedef struct gxy_record {
    uint32_t target_index;
    uint32_t query_index;
    uint32_t strand; /* GXY_STRAND_PLUS / GXY_STRAND_MINUS */
    uint32_t t_start;
    uint32_t q_start;
    uint32_t t_span;
    uint32_t q_span;
    int32_t score;
    uint32_t n_runs;
    uint32_t reserved;
    uint64_t runs_offset;
} gxy_record;

typedef struct gxy_record_set {
    uint64_t n;
    gxy_record *records;
    uint64_t n_runs;
    gxy_edit_run *runs;
} gxy_record_set;

/* Counters of one run() call. */
typedef struct gxy_run_stats {
    uint64_t chunks;
    uint64_t query_bp;       /* total query bases scanned per strand pass   */
    uint64_t hits;
    uint64_t columns;
    uint64_t hsps;
    uint64_t dp_problems;    /* gapped extensions computed                  */
    uint64_t dp_cells;
    uint64_t tie_pairs;      /* order-sensitive ties (see gxy_search_stats) */
    uint64_t score_ties;     /* equal-score neighbours in the gapped order  */
    uint64_t passing;        /* x-drop extensions reaching hsp_threshold    */
    uint64_t lookups;        /* table lookups: valid query windows x seed variants */
} gxy_run_stats;

/* Config::default() (src/driver.rs:91-111) with HOXD70 (src/scoring.rs:44-58). */
void gxy_config_default(gxy_config *cfg);
/* SeedPattern from "12of19", "19of20", "match<k>" or a 0/1 spec
 * (src/cli.rs:225-240). */
int gxy_pattern_parse(const char *spec, gxy_seed_pattern *out);

/* driver::run(targets, queries, &Config) -> Vec<Record> (src/driver.rs:199). */
int gxy_run(gxy_ctx *ctx, const gxy_named_seq *targets, uint64_t n_targets,
            const gxy_named_seq *queries, uint64_t n_queries, const gxy_config *cfg,
            gxy_record_set **out, gxy_run_stats *stats);
/* driver::run over device-resident sequences (names are not needed: records
 * carry indices).  The target/query arrays hold n_targets / n_queries
 * handles.  gxy_run is gxy_dseq_upload of every sequence followed by this. */
int gxy_run_dev(gxy_ctx *ctx, const gxy_dseq *const *targets, uint64_t n_targets,
                const gxy_dseq *const *queries, uint64_t n_queries, const gxy_config *cfg,
                gxy_record_set **out, gxy_run_stats *stats);
/* tweener::interpolate (src/tweener.rs:59-185) over the records of one
 * (target, query, strand) group: `target` is the whole target, `query` the
 * group's query strand, cfg's inner_* fields the TweenerConfig (chain_enabled,
 * transitions, entropy and max_word_count apply as in the reference).  The
 * output is sorted by (t_start, q_start) and deduplicated; free with
 * gxy_record_set_free. */
int gxy_interpolate(gxy_ctx *ctx, const gxy_record_set *records, const gxy_dseq *target,
                    const gxy_dseq *query, const gxy_config *cfg, gxy_record_set **out);
/* Render records as MAF (src/output/maf.rs:33-74, header as src/main.rs:42)
 * or PAF (src/output/paf.rs:23-89).  `*text` is NUL-terminated; free with
 * gxy_free. */
int gxy_format_maf(const gxy_named_seq *targets, uint64_t n_targets,
                   const gxy_named_seq *queries, uint64_t n_queries,
                   const gxy_record_set *records, char **text, uint64_t *len);
int gxy_format_paf(const gxy_named_seq *targets, uint64_t n_targets,
                   const gxy_named_seq *queries, uint64_t n_queries,
                   const gxy_record_set *records, char **text, uint64_t *len);
void gxy_record_set_free(gxy_record_set *set);

/* Free an array returned by the library (gxy_search output, texts). */
void gxy_free(void *p);

/* ---- Serving: driver::run as a request loop --------------------------------
 * Replaces the per-query loop of src/driver.rs:219-391 (run() over one query
 * at a time) for queries that arrive over time.  The server keeps the target
 * chunks and their position tables resident in HBM, admits waiting queries
 * each iteration within a query-base cap and an HBM byte budget, searches all
 * admitted queries of an iteration together (concatenated, separated by runs
 * of N long enough that neither a seed nor an x-drop walk crosses them) and
 * hands their HSPs to gapped-stage workers, each on its own device stream,
 * which take the waiting queries of one length bin (< 10 kbp, < 100 kbp,
 * < 1 Mbp, longer) and extend their HSPs in shared device batches.
 * Queries that arrive while an iteration runs are admitted by the next one;
 * a query completes as soon as its own extensions are settled.  Each query's
 * records equal those of gxy_run_dev on that query alone (query_index 0).
 *
 * The server owns a context on `device` and a worker thread; submit/wait may
 * be called from any host threads. */
typedef struct gxy_server gxy_server;

typedef struct gxy_server_params {
    uint64_t hbm_budget_bytes; /* device bytes the queries of one iteration may hold (0: 16 GiB) */
    uint64_t max_batch_bp;     /* query bases admitted per iteration (0: 256 Mbp)                */
    int32_t co_batch;          /* 0: one query per iteration (the plain per-query driver)          */
    int32_t dp_workers;        /* gapped-stage threads, one device stream each (0: 3)              */
} gxy_server_params;

typedef struct gxy_request_stats {
    double submit_ms;          /* server clock (ms since gxy_server_create) */
    double admit_ms;
    double done_ms;
    uint64_t query_bp;
    uint64_t iteration;        /* iteration that served the query          */
    uint64_t batch_queries;    /* queries served by that iteration         */
    uint64_t batch_bp;
} gxy_request_stats;

int gxy_server_create(int device, const gxy_dseq *const *targets, uint64_t n_targets,
                      const gxy_config *cfg, const gxy_server_params *params, gxy_server **out);
/* Queue a query (copied; the caller's buffers may be reused on return). */
int gxy_server_submit(gxy_server *server, const gxy_seq *query, uint64_t *ticket);
/* Block until the ticket's query is served; hands over its records (free with
 * gxy_record_set_free) and fills `stats` when non-null.  A ticket is waited on
 * once: a second wait on it (also from another thread) returns GXY_EINVAL. */
int gxy_server_wait(gxy_server *server, uint64_t ticket, gxy_record_set **out,
                    gxy_request_stats *stats);
/* Message of the last failure inside the server (or of the failed request).
 * The text is a copy owned by the calling thread, valid until its next call. */
const char *gxy_server_last_error(gxy_server *server);
/* Kernel timing over the server's contexts (no counterpart in the reference;
 * for measurement only), as gxy_ctx_set_timing / gxy_ctx_kernel_time.  Both
 * return GXY_EINVAL while requests are in flight (submitted, not yet waited). */
int gxy_server_set_timing(gxy_server *server, int enabled);
/* Work done so far over all requests (seed hits, x-drop columns, passing
 * extensions, HSPs, gapped extensions and DP cells; for measurement). */
int gxy_server_stats(gxy_server *server, gxy_run_stats *out);
int gxy_server_kernel_time(gxy_server *server, const char *name, double *ms, uint64_t *launches);
/* Stops the worker after the queued queries are served; frees everything. */
/* gxy_ctx_xdrop_stats summed over the server's contexts. */
int gxy_server_xdrop_stats(gxy_server *server, uint64_t *first_step_columns, uint64_t *resumed,
                           uint64_t *first_step_passing);
int gxy_server_destroy(gxy_server *server);

#ifdef __cplusplus
}
#endif
#endif /* GXY_H */
