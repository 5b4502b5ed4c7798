// common.h — shared host/device definitions of the gfx950 hot-path library.
//
// Device sequence layout ("block32"): a PackedSeq (src/sequences.rs:128-141)
// is re-laid in HBM as 16-byte blocks of 32 bases,
//     struct Blk { uint64 codes; uint32 valid; uint32 mask; }
// with base i at bits 2*(i&31) of codes[i>>5] and bit (i&31) of valid/mask —
// the same bit positions as the reference's byte arrays read as little-endian
// words, so the re-layout is a pure interleave.  One 16-byte load brings
// everything a lane needs for 32 bases.  Every device sequence carries one
// zero block past its end so a 32-base window read at any in-range position
// may touch block (i>>5)+1 without a bounds branch.
#pragma once

#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <new>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/gxy.h"

namespace gxy {

struct Blk {
    uint64_t codes;
    uint32_t valid;
    uint32_t mask;
};
static_assert(sizeof(Blk) == 16, "Blk must be 16 bytes");

constexpr uint32_t kMaxWeight = 14;          // dense 4^w table; 4^14+1 starts = 1 GiB
constexpr uint32_t kNcntMaxWeight = 13;      // neighbourhood counts up to 4^13 words (256 MiB)
constexpr int32_t kNegInf = INT32_MIN / 4;   // src/gapped_extend.rs:44
constexpr uint32_t kInvalidWord = 0xFFFFFFFFu;

// Care offsets of a seed pattern (src/seeds.rs:112-120) in kernel form.
struct DevPattern {
    uint32_t span;
    uint32_t weight;
    uint32_t care_mask;     // bit k set when offset k is a care position
    uint8_t off[32];        // ascending care offsets
};

inline bool make_dev_pattern(const gxy_seed_pattern *p, DevPattern *d) {
    if (!p || p->length == 0 || p->length > 32) return false;
    uint32_t care = p->length == 32 ? p->care : (p->care & ((1u << p->length) - 1u));
    d->span = p->length;
    d->care_mask = care;
    d->weight = 0;
    std::memset(d->off, 0, sizeof d->off);
    for (uint32_t k = 0; k < p->length; k++)
        if ((care >> k) & 1u) d->off[d->weight++] = (uint8_t)k;
    return d->weight > 0;
}

// ---- device helpers --------------------------------------------------------

// Table index of a seed word.  The reference's word (src/seeds.rs:139-145)
// concatenates the 2-bit codes at the care offsets, leftmost first.  The
// device tables are indexed by a bit permutation of it,
//     index = (low bits of the codes, leftmost first) << W | (high bits ...),
// because a transition (A<->G, C<->T) flips only a code's high bit: the 1+W
// single-transition variants of a word are then index ^ (1 << b), b < W — all
// inside one 4^w/2^W-entry block (512 B of the non-empty bitmap, 16 KiB of
// `starts` for 12of19) instead of 13 random places in a 64 MiB table.
// Positions per word are unchanged; only the order of the word buckets is.
// Fully unrolled so the care offsets stay in scalar registers.
__device__ __forceinline__ uint64_t gather_index(uint64_t codes, const DevPattern &pat) {
    uint32_t lo = 0, hi = 0;
#pragma unroll
    for (int k = 0; k < 32; k++) {
        if ((uint32_t)k < pat.weight) {
            const uint32_t c = (uint32_t)(codes >> (2 * pat.off[k])) & 3u;
            lo = (lo << 1) | (c & 1u);
            hi = (hi << 1) | (c >> 1);
        }
    }
    return ((uint64_t)lo << pat.weight) | hi;
}

// Reference word <-> table index (see gather_index).
__host__ __device__ __forceinline__ uint64_t word_to_index(uint64_t word, uint32_t weight) {
    uint64_t lo = 0, hi = 0;
    for (uint32_t k = 0; k < weight; k++) {
        const uint64_t c = (word >> (2 * (weight - 1 - k))) & 3u;
        lo = (lo << 1) | (c & 1u);
        hi = (hi << 1) | (c >> 1);
    }
    return (lo << weight) | hi;
}

// 32-base window starting at `pos` (any alignment): codes in bits 2k..2k+1 and
// valid/mask in bit k for base pos+k.
struct Win {
    uint64_t codes;
    uint32_t valid;
    uint32_t mask;
};

__device__ __forceinline__ Win load_win(const Blk *__restrict__ b, uint64_t pos) {
    const uint64_t k = pos >> 5;
    const uint32_t r = (uint32_t)(pos & 31);
    const Blk lo = b[k];
    if (r == 0) return {lo.codes, lo.valid, lo.mask};
    const Blk hi = b[k + 1];
    Win w;
    w.codes = (lo.codes >> (2 * r)) | (hi.codes << (64 - 2 * r));
    w.valid = (lo.valid >> r) | (hi.valid << (32 - r));
    w.mask = (lo.mask >> r) | (hi.mask << (32 - r));
    return w;
}

__device__ __forceinline__ uint32_t base_code(const Blk *__restrict__ b, uint64_t i) {
    return (uint32_t)(b[i >> 5].codes >> (2 * (i & 31))) & 3u;
}
__device__ __forceinline__ uint32_t base_valid(const Blk *__restrict__ b, uint64_t i) {
    return (b[i >> 5].valid >> (i & 31)) & 1u;
}

// Reverse the order of the 32 2-bit groups of a 64-bit word.
__host__ __device__ __forceinline__ uint64_t rev2(uint64_t x) {
    x = ((x >> 2) & 0x3333333333333333ull) | ((x & 0x3333333333333333ull) << 2);
    x = ((x >> 4) & 0x0F0F0F0F0F0F0F0Full) | ((x & 0x0F0F0F0F0F0F0F0Full) << 4);
    x = ((x >> 8) & 0x00FF00FF00FF00FFull) | ((x & 0x00FF00FF00FF00FFull) << 8);
    x = ((x >> 16) & 0x0000FFFF0000FFFFull) | ((x & 0x0000FFFF0000FFFFull) << 16);
    return (x >> 32) | (x << 32);
}

// Spread 32 validity bits to the 2-bit groups of a 64-bit code word.
__host__ __device__ __forceinline__ uint64_t spread2(uint32_t v) {
    uint64_t x = v;
    x = (x | (x << 16)) & 0x0000FFFF0000FFFFull;
    x = (x | (x << 8)) & 0x00FF00FF00FF00FFull;
    x = (x | (x << 4)) & 0x0F0F0F0F0F0F0F0Full;
    x = (x | (x << 2)) & 0x3333333333333333ull;
    x = (x | (x << 1)) & 0x5555555555555555ull;
    return x | (x << 1);
}

__device__ __forceinline__ int32_t sat_add(int32_t a, int32_t b) {
    return __builtin_elementwise_add_sat(a, b);
}
__device__ __forceinline__ int32_t sat_sub(int32_t a, int32_t b) {
    return __builtin_elementwise_sub_sat(a, b);
}

// Column score lookup (src/hsp.rs:198-211): sub[t][q] when both bases are
// valid, ambig otherwise.  Rows of sub[][] are packed as four signed bytes
// when every entry fits in int8 (HOXD70 does; src/scoring.rs:93-96), which
// makes a lookup three selects and one sign-extending shift.  Matrices with
// wider entries use the 16-entry int32 table.
struct ScorerI8 {
    uint32_t r0, r1, r2, r3;
    int32_t ambig;
    __device__ __forceinline__ int32_t operator()(uint32_t tc, uint32_t qc, uint32_t v) const {
        const uint32_t lo = (tc & 1u) ? r1 : r0;
        const uint32_t hi = (tc & 1u) ? r3 : r2;
        const uint32_t row = (tc & 2u) ? hi : lo;
        const int32_t s = (int32_t)(row << (24u - 8u * qc)) >> 24;
        return v ? s : ambig;
    }
};

struct ScorerI32 {
    int32_t s[16];
    int32_t ambig;
    __device__ __forceinline__ int32_t operator()(uint32_t tc, uint32_t qc, uint32_t v) const {
        // Selects only: a dynamically indexed register array would spill.
        int32_t row[4];
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const int32_t a = (qc & 1u) ? s[t * 4 + 1] : s[t * 4 + 0];
            const int32_t b = (qc & 1u) ? s[t * 4 + 3] : s[t * 4 + 2];
            row[t] = (qc & 2u) ? b : a;
        }
        const int32_t lo = (tc & 1u) ? row[1] : row[0];
        const int32_t hi = (tc & 1u) ? row[3] : row[2];
        const int32_t x = (tc & 2u) ? hi : lo;
        return v ? x : ambig;
    }
};

inline bool scoring_is_i8(const gxy_scoring &m) {
    auto in = [](int32_t v) { return v >= -128 && v <= 127; };
    for (int a = 0; a < 4; a++)
        for (int b = 0; b < 4; b++)
            if (!in(m.sub[a][b])) return false;
    return true;
}

inline ScorerI8 make_scorer_i8(const gxy_scoring &m) {
    ScorerI8 s;
    uint32_t r[4];
    for (int a = 0; a < 4; a++) {
        r[a] = 0;
        for (int b = 0; b < 4; b++) r[a] |= (uint32_t)(uint8_t)(int8_t)m.sub[a][b] << (8 * b);
    }
    s.r0 = r[0];
    s.r1 = r[1];
    s.r2 = r[2];
    s.r3 = r[3];
    s.ambig = m.ambig;
    return s;
}

inline ScorerI32 make_scorer_i32(const gxy_scoring &m) {
    ScorerI32 s;
    for (int a = 0; a < 4; a++)
        for (int b = 0; b < 4; b++) s.s[a * 4 + b] = m.sub[a][b];
    s.ambig = m.ambig;
    return s;
}

// ---- byte-per-base layout used by the x-drop kernels ---------------------------
//
// Byte-per-base x-drop layout.  Target and query use different byte codes
// so that a column's score is two byte permutes and one and-xor:
//   target t in {0,1}: 0xD0 | 4*t         (low nibble selects table A)
//   target t in {2,3}: 0x0D | 4*(t&1)<<4  (high nibble selects table B)
//   target invalid:    0xDD
//   query q valid:     q;   query invalid: 0x0D
// Table A holds the rows of t=0 (bytes 0-3) and t=1 (bytes 4-7), table B
// those of t=2 and t=3.  A permute selector of 13..15 yields 0xFF, so for a
// column the table not holding its target row — and both tables when either
// base is invalid — give 0xFF.  The rows are stored xor K, K = ~ambig in
// every byte, so (A & B) ^ K is the score, and ambig for invalid columns.
// A buffer holds kBytePad bytes of padding (invalid bases) on each side of
// the sequence, so a 32-base window read forward from any in-range position,
// or backward from any position >= 0, stays in bounds; callers pass
// buffer + kBytePad as the sequence's byte 0.  Backward walks read the
// forward bytes and reverse them in registers, so both sides of a hit share
// cache lines.
constexpr uint64_t kBytePad = 64;
constexpr uint8_t kTargetInvalidByte = 0xDD;
constexpr uint8_t kQueryInvalidByte = 0x0D;
__host__ __device__ constexpr uint64_t byte_buffer_size(uint64_t len) { return len + 2 * kBytePad; }

__host__ __device__ __forceinline__ uint32_t target_byte(uint32_t code, bool valid) {
    if (!valid) return kTargetInvalidByte;
    return code < 2 ? 0xD0u | (4u * code) : 0x0Du | ((4u * (code & 1u)) << 4);
}
__host__ __device__ __forceinline__ uint32_t query_byte(uint32_t code, bool valid) {
    return valid ? code : (uint32_t)kQueryInvalidByte;
}

struct ByteLut {
    uint32_t a0, a1, b0, b1;  // rows of t = 0, 1, 2, 3, each xor k
    uint32_t k;               // ~ambig in every byte
};

// Scores of four columns (four target bytes, four query bytes) as four
// signed bytes.
__device__ __forceinline__ uint32_t score4(const ByteLut &L, uint32_t t4, uint32_t q4) {
    const uint32_t sa = (t4 & 0x0F0F0F0Fu) | q4;
    const uint32_t sb = ((t4 >> 4) & 0x0F0F0F0Fu) | q4;
    const uint32_t a = __builtin_amdgcn_perm(L.a1, L.a0, sa);
    const uint32_t b = __builtin_amdgcn_perm(L.b1, L.b0, sb);
    return (a & b) ^ L.k;
}

// 32 consecutive bytes starting at any byte offset (pos may be negative down
// to -kBytePad).
__device__ __forceinline__ void load_bytes32(const uint8_t *__restrict__ b, int64_t pos, uint32_t w[8]) {
    const uint32_t *p = (const uint32_t *)(b + (pos & ~(int64_t)3));
    const uint32_t sh = (uint32_t)(pos & 3);
    uint32_t d[9];
#pragma unroll
    for (int i = 0; i < 9; i++) d[i] = p[i];
#pragma unroll
    for (int i = 0; i < 8; i++) w[i] = __builtin_amdgcn_alignbyte(d[i + 1], d[i], sh);
}

inline ByteLut make_byte_lut(const gxy_scoring &m) {
    const ScorerI8 s = make_scorer_i8(m);  // row t: byte q = sub[t][q]
    const uint32_t k = (uint32_t)(uint8_t)~(uint8_t)(int8_t)m.ambig * 0x01010101u;
    return ByteLut{s.r0 ^ k, s.r1 ^ k, s.r2 ^ k, s.r3 ^ k, k};
}

inline bool scoring_byte_ok(const gxy_scoring &m) {
    return scoring_is_i8(m) && m.ambig >= -128 && m.ambig <= 127;
}

// ---- spread counters ------------------------------------------------------------
//
// A counter incremented once per wavefront by every wavefront of a large grid
// must not be one address: same-address atomics serialise in one L2 channel.
// kSpread counters, one per 128-byte line, are added to by wavefront index and
// summed on the host.
constexpr int kSpread = 64;
constexpr int kSpreadStride = 16;  // in 8-byte words: one counter per 128-byte line
constexpr size_t kSpreadBytes = (size_t)kSpread * kSpreadStride * 8;

__device__ __forceinline__ void spread_add(unsigned long long *c, unsigned long long v) {
    const unsigned slot = (blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) & (kSpread - 1);
    atomicAdd(c + (size_t)slot * kSpreadStride, v);
}

// ---- host side ---------------------------------------------------------------

inline uint64_t nblocks(uint64_t len) { return (len + 31) / 32 + 1; }

}  // namespace gxy

// Opaque handle definitions (shared by all translation units).
struct gxy_dseq {
    uint64_t len = 0;
    uint64_t nblk = 0;
    gxy::Blk *d = nullptr;  // nblk blocks (last one zero)
    int device = 0;
    gxy_ctx *pool = nullptr;  // block cache `d` came from (driver-internal sequences)
};

struct gxy_postable {
    gxy::DevPattern pattern;
    uint64_t step = 1;
    uint64_t num_words = 0;
    uint64_t npos = 0;
    uint32_t *d_starts = nullptr;     // num_words + 1
    uint32_t *d_positions = nullptr;  // npos (>= 1 allocated)
    const gxy_dseq *target = nullptr; // sequence the table indexes
    gxy_dseq *owned_target = nullptr; // set when the table uploaded it itself
    const void *host_codes = nullptr; // identity of the host target (gxy_search)
    uint64_t host_len = 0;
    uint32_t *d_nonempty = nullptr;   // bit w set when word w has positions
    uint32_t *d_cmap = nullptr;       // 2 bits per word: min(positions, 3), 16 words per u32
    uint32_t *d_spos = nullptr;       // first position of every word (valid when it has one)
    // Neighbourhood counts (weight <= kNcntMaxWeight): for every word, the
    // 2-bit sizes of its 1+W single-transition variants (variant i in bits
    // 2i..2i+1, same order as the search), so counting a window is one load.
    uint32_t *d_ncnt = nullptr;
    uint8_t *d_tbytes = nullptr;      // target byte buffer (x-drop layout, sequence at +kBytePad)
    gxy_ctx *pool = nullptr;          // block cache the buffers came from (driver-internal tables)
};

struct gxy_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    // Growable scratch slots (device memory), reused across calls.  Where the
    // device has memory pools they are allocated and freed in stream order on
    // `stream` (hipMallocAsync / hipFreeAsync), so growing one never waits for
    // work on other contexts' streams as hipFree would.
    bool async_alloc = false;
    static constexpr int kSlots = 32;
    void *slot[kSlots] = {};
    size_t slot_bytes[kSlots] = {};
    void *pinned = nullptr;
    size_t pinned_bytes = 0;
    // Kernel timing (gxy_ctx_set_timing): event pairs pending on the stream and
    // per-name totals.
    struct Span {
        const char *name;
        hipEvent_t a, b;
    };
    bool timing = false;
    std::vector<Span> spans;
    std::vector<hipEvent_t> event_pool;
    std::vector<std::pair<std::string, std::pair<double, uint64_t>>> totals;
    // Device block cache: while pool_on is set (inside the driver), sequence
    // and table buffers are taken from and returned to these lists instead of
    // hipMalloc/hipFree, so repeated runs of the same shapes allocate nothing.
    // Reuse is ordered by the context's single stream.
    bool pool_on = false;
    // A second context for the driver's other-strand searches (created on
    // first use, destroyed with this one).
    gxy_ctx *helper = nullptr;
    // A second context (with its own helper) for the driver's second chunk
    // lane (created on first use, destroyed with this one).
    gxy_ctx *lane = nullptr;
    uint64_t dp_pool_cap = 0;  // traceback pool bytes (0: min(96 GiB, device / 3))
    uint64_t dp_pool_rows = 0; // traceback pool rows set by the caller (0: no row cap)
    int dp_kernel = 0;         // DP forward kernel (GXY_DP_*), 0: chosen by the library
    // DP autotuner (gapped.hip): per class of pass size, the forward kernel
    // chosen, or the candidates' measured rates while it is still probing.
    static constexpr int kDpClasses = 4, kDpCands = 5;
    // Passes of at most 8 sides are not probed: a lone side is bound by its
    // row latency, which the four-wavefront kernel halves (the one-wavefront
    // kernels measured 0.24-0.27 against 0.49 GCUPS on a 30 kbp pair, and a
    // probe pass there costs a whole sweep).
    struct DpTune {
        int chosen = 0;                 // GXY_DP_* once decided, else 0
        int tried = 0;                  // candidates measured so far
        double gcups[kDpCands] = {};    // forward-kernel cells / ns of each candidate's probe
    } dp_tune[kDpClasses] = {{GXY_DP_BLOCK4, 0, {}}, {}, {}, {}};
    hipEvent_t dp_ev[2] = {nullptr, nullptr};
    // gapped passes since creation: passes, sides stopped by a full pool, passes run with one side alone
    uint64_t dp_passes = 0, dp_truncated = 0, dp_alone = 0;
    // Persistent DP launches (gapped.hip, k_dp_persist): 0 for batches of at
    // least kPersistMinSides sides once the kernel is a one-wavefront variant,
    // 1 for every batch, -1 never.  Counters: launches, sides given to them,
    // sides they left to passes, DP cells of the sides they finished.
    int dp_persist = 0;
    uint64_t dp_persist_launches = 0, dp_persist_sides = 0, dp_persist_deferred = 0, dp_persist_cells = 0;
    int cus = 0;  // compute units of the device (ctx_cus)
    // x-drop in one persistent kernel per hit batch instead of a first step
    // for every hit and a kernel resuming the longer walks (for measurement)
    bool xdrop_single = false;
    // hits per expansion batch of a search (0: the default, 2^27)
    uint64_t hit_batch = 0;
    // first steps since creation: x-drop columns evaluated, hits handed on to
    // the resuming kernel, passing extensions finished
    uint64_t xd_first_cols = 0, xd_conts = 0, xd_first_pass = 0;
    std::multimap<size_t, void *> pool_free;
    std::unordered_map<void *, size_t> pool_used;
    // Host phase clock (GXY_TRACE set in the environment): milliseconds per phase.
    bool trace = std::getenv("GXY_TRACE") != nullptr;
    double trace_ms[16] = {};
};

namespace gxy {

// Adds the wall time since the previous lap to ctx->trace_ms[k] when tracing.
struct PhaseClock {
    gxy_ctx *ctx;
    std::chrono::steady_clock::time_point t = std::chrono::steady_clock::now();
    explicit PhaseClock(gxy_ctx *c) : ctx(c) {}
    void lap(int k) {
        if (!ctx->trace) return;
        auto n = std::chrono::steady_clock::now();
        ctx->trace_ms[k] += std::chrono::duration<double, std::milli>(n - t).count();
        t = n;
    }
};

// Phase slots of the host clock.
enum TracePhase : int {
    kTrSetup = 0, kTrTable, kTrSearchCount, kTrSearchExtend, kTrSearchAdmit, kTrSearchCopy, kTrSearchDedup,
    kTrRecords, kTrMerge, kTrSearchOther
};

}  // namespace gxy

namespace gxy {

int set_err(gxy_ctx *ctx, int code, const char *fmt, ...);
// Compute units of the context's device (queried once; 256 if the query fails).
inline int ctx_cus(gxy_ctx *ctx) {
    if (!ctx->cus) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, ctx->device) != hipSuccess || v <= 0)
            v = 256;
        ctx->cus = v;
    }
    return ctx->cus;
}
// Sides a batch must hold for the library to sweep it with one persistent launch.
constexpr uint64_t kPersistMinSides = 16384;
// Runs the body of a C entry point: a C++ exception (host allocation failure
// or a library invariant) becomes an error code and message instead of
// crossing the ABI.
template <class F>
int guarded(gxy_ctx *ctx, F &&body) {
    try {
        return body();
    } catch (const std::bad_alloc &) {
        return set_err(ctx, GXY_ENOMEM, "host allocation failed");
    } catch (const std::exception &ex) {
        return set_err(ctx, GXY_EINVAL, "%s", ex.what());
    } catch (...) {
        return set_err(ctx, GXY_EINVAL, "unexpected exception");
    }
}
// Device allocation for sequences and tables: from the context's block cache
// when ctx->pool_on (then *owner = ctx), else hipMalloc (*owner = nullptr).
// nullptr on failure.
void *dev_alloc(gxy_ctx *ctx, size_t bytes, gxy_ctx **owner);
// The same with the owner given: its cache when non-null, else hipMalloc.
void *dev_alloc_from(gxy_ctx *owner, size_t bytes);
// Returns p to owner's cache, or hipFree when owner is null.
void dev_free(gxy_ctx *owner, void *p);
// hipFree every cached block (after synchronising the stream).
void pool_trim(gxy_ctx *ctx);
// Sum of a spread counter block (synchronises the stream).
int spread_sum(gxy_ctx *ctx, const unsigned long long *c, uint64_t *out);

// Brackets the launches issued in its scope with timing events when the
// context's timing is enabled.
class KernelTimer {
  public:
    KernelTimer(gxy_ctx *ctx, const char *name);
    ~KernelTimer();

  private:
    gxy_ctx *ctx_;
    const char *name_;
    hipEvent_t a_ = nullptr;
};
void timing_flush(gxy_ctx *ctx);
// Device scratch slot `k` of at least `bytes` bytes (contents undefined).
void *scratch(gxy_ctx *ctx, int k, size_t bytes);
// Slot `k` grown to at least `bytes` bytes, its first `keep` bytes kept
// (copied in stream order when the slot moves).
void *scratch_grow_keep(gxy_ctx *ctx, int k, size_t keep, size_t bytes);
// Frees slot `k` (in stream order) if it holds more than `above` bytes.
void scratch_release(gxy_ctx *ctx, int k, size_t above);
int dseq_upload_impl(gxy_ctx *ctx, const gxy_seq *seq, gxy_dseq **out);

// Scratch slot assignment.
enum Slot {
    kSlotTmp = 0,      // hipcub temporary storage
    kSlotWords,
    kSlotCounts,
    kSlotOffsets,
    kSlotHits,
    kSlotPassKeys,
    kSlotPassKeys2,
    kSlotPassIdx,
    kSlotPassIdx2,
    kSlotPassHsp,
    kSlotFlags,
    kSlotEmit,
    kSlotCounters,
    kSlotMisc0,
    kSlotMisc1,
    kSlotMisc2,
    kSlotMisc3,
    kSlotDpProb,
    kSlotDpOut,
    kSlotDpTb,
    kSlotDpRows,
    kSlotDpRuns,
    kSlotQBytes,
    kSlotTBytes,
    kSlotVMask,
    kSlotSpread,
    kSlotHits2,
    kSlotBins,
    kSlotBinOff,
    kSlotDpPersist,
    kSlotXCont,
    kSlotXContCnt,
};

}  // namespace gxy

#define GXY_HIP(ctx, call)                                                              \
    do {                                                                                \
        hipError_t _e = (call);                                                         \
        if (_e != hipSuccess)                                                           \
            return gxy::set_err((ctx), GXY_EDEVICE, "%s failed: %s (%s:%d)", #call,    \
                                hipGetErrorString(_e), __FILE__, __LINE__);             \
    } while (0)

#define GXY_LAUNCH_CHECK(ctx)                                                           \
    do {                                                                                \
        hipError_t _e = hipGetLastError();                                              \
        if (_e != hipSuccess)                                                           \
            return gxy::set_err((ctx), GXY_EDEVICE, "kernel launch failed: %s (%s:%d)",   \
                                hipGetErrorString(_e), __FILE__, __LINE__);             \
    } while (0)
