// seq.hip — context, scratch, and device-resident PackedSeq (block32 layout).
//
// Reference: PackedSeq (src/sequences.rs:128-318).  reverse_complement and
// slice_to_new run as one thread per 32-base output block: each thread loads
// the (at most two) source blocks its window spans, shifts them into place and
// for RC reverses the 2-bit groups, so both kernels are single-pass,
// coalesced and HBM-bound.
#include <cstdarg>
#include <mutex>

#include "internal.h"

namespace gxy {

static thread_local std::string g_create_err;

int set_err(gxy_ctx *ctx, int code, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (ctx)
        ctx->err = buf;
    else
        g_create_err = buf;
    return code;
}

void *scratch(gxy_ctx *ctx, int k, size_t bytes) {
    if (bytes == 0) bytes = 16;
    if (ctx->slot_bytes[k] >= bytes) return ctx->slot[k];
    if (ctx->slot[k]) {
        if (ctx->async_alloc) {
            (void)hipFreeAsync(ctx->slot[k], ctx->stream);
        } else {
            (void)hipStreamSynchronize(ctx->stream);
            (void)hipFree(ctx->slot[k]);
        }
        ctx->slot[k] = nullptr;
        ctx->slot_bytes[k] = 0;
    }
    size_t sz = bytes + bytes / 8;  // headroom so slowly growing inputs do not reallocate
    sz = (sz + 255) & ~(size_t)255;
    const hipError_t e =
        ctx->async_alloc ? hipMallocAsync(&ctx->slot[k], sz, ctx->stream) : hipMalloc(&ctx->slot[k], sz);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        ctx->slot[k] = nullptr;
        return nullptr;
    }
    ctx->slot_bytes[k] = sz;
    return ctx->slot[k];
}

void *scratch_grow_keep(gxy_ctx *ctx, int k, size_t keep, size_t bytes) {
    if (ctx->slot_bytes[k] >= bytes) return ctx->slot[k];
    void *old = ctx->slot[k];
    const size_t old_bytes = ctx->slot_bytes[k];
    ctx->slot[k] = nullptr;  // keep the old block alive across scratch()
    ctx->slot_bytes[k] = 0;
    void *p = scratch(ctx, k, bytes);
    if (!p) {
        ctx->slot[k] = old;
        ctx->slot_bytes[k] = old_bytes;
        return nullptr;
    }
    if (old) {
        if (keep && hipMemcpyAsync(p, old, std::min(keep, old_bytes), hipMemcpyDeviceToDevice, ctx->stream) !=
                        hipSuccess)
            (void)hipGetLastError();
        if (ctx->async_alloc) {
            (void)hipFreeAsync(old, ctx->stream);
        } else {
            (void)hipStreamSynchronize(ctx->stream);
            (void)hipFree(old);
        }
    }
    return p;
}

void scratch_release(gxy_ctx *ctx, int k, size_t above) {
    if (!ctx->slot[k] || ctx->slot_bytes[k] <= above) return;
    if (ctx->async_alloc) {
        (void)hipFreeAsync(ctx->slot[k], ctx->stream);
    } else {
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipFree(ctx->slot[k]);
    }
    ctx->slot[k] = nullptr;
    ctx->slot_bytes[k] = 0;
}

void *dev_alloc(gxy_ctx *ctx, size_t bytes, gxy_ctx **owner) {
    *owner = ctx->pool_on ? ctx : nullptr;
    return dev_alloc_from(*owner, bytes);
}

void *dev_alloc_from(gxy_ctx *ctx, size_t bytes) {
    bytes = (std::max<size_t>(bytes, 1) + 255) & ~(size_t)255;
    void *p = nullptr;
    if (!ctx) {
        if (hipMalloc(&p, bytes) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
        return p;
    }
    // smallest cached block that fits without wasting more than a quarter
    auto it = ctx->pool_free.lower_bound(bytes);
    if (it != ctx->pool_free.end() && it->first <= bytes + bytes / 4) {
        p = it->second;
        ctx->pool_used[p] = it->first;
        ctx->pool_free.erase(it);
        return p;
    }
    if (hipMalloc(&p, bytes) != hipSuccess) {
        (void)hipGetLastError();
        pool_trim(ctx);
        if (hipMalloc(&p, bytes) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
    }
    ctx->pool_used[p] = bytes;
    return p;
}

void dev_free(gxy_ctx *owner, void *p) {
    if (!p) return;
    if (!owner) {
        (void)hipFree(p);
        return;
    }
    auto it = owner->pool_used.find(p);
    if (it == owner->pool_used.end()) return;
    owner->pool_free.emplace(it->second, p);
    owner->pool_used.erase(it);
}

// Returns the device memory pool's unused memory to the device: stream-ordered
// frees (hipFreeAsync) leave it in the pool, where allocations of other
// processes, and plain allocations of this one, cannot use it.
static void trim_device_pool(gxy_ctx *ctx) {
    if (!ctx->async_alloc) return;
    hipMemPool_t mp;
    if (hipDeviceGetDefaultMemPool(&mp, ctx->device) == hipSuccess)
        (void)hipMemPoolTrimTo(mp, 0);
    else
        (void)hipGetLastError();
}

void pool_trim(gxy_ctx *ctx) {
    if (ctx->pool_free.empty()) return;
    (void)hipStreamSynchronize(ctx->stream);
    for (auto &kv : ctx->pool_free) (void)hipFree(kv.second);
    ctx->pool_free.clear();
}

static hipEvent_t take_event(gxy_ctx *ctx) {
    if (!ctx->event_pool.empty()) {
        hipEvent_t e = ctx->event_pool.back();
        ctx->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    return e;
}

KernelTimer::KernelTimer(gxy_ctx *ctx, const char *name) : ctx_(ctx), name_(name) {
    if (!ctx_->timing) return;
    a_ = take_event(ctx_);
    if (a_) (void)hipEventRecord(a_, ctx_->stream);
}

KernelTimer::~KernelTimer() {
    if (!a_) return;
    hipEvent_t b = take_event(ctx_);
    if (!b) {
        ctx_->event_pool.push_back(a_);
        return;
    }
    (void)hipEventRecord(b, ctx_->stream);
    ctx_->spans.push_back({name_, a_, b});
    if (ctx_->spans.size() > 4096) timing_flush(ctx_);
}

void timing_flush(gxy_ctx *ctx) {
    if (ctx->spans.empty()) return;
    (void)hipStreamSynchronize(ctx->stream);
    for (auto &sp : ctx->spans) {
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, sp.a, sp.b);
        bool found = false;
        for (auto &t : ctx->totals)
            if (t.first == sp.name) {
                t.second.first += ms;
                t.second.second += 1;
                found = true;
                break;
            }
        if (!found) ctx->totals.push_back({sp.name, {(double)ms, 1}});
        ctx->event_pool.push_back(sp.a);
        ctx->event_pool.push_back(sp.b);
    }
    ctx->spans.clear();
}

int spread_sum(gxy_ctx *ctx, const unsigned long long *c, uint64_t *out) {
    unsigned long long h[kSpread * kSpreadStride];
    GXY_HIP(ctx, hipMemcpyAsync(h, c, kSpreadBytes, hipMemcpyDeviceToHost, ctx->stream));
    GXY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    uint64_t t = 0;
    for (int k = 0; k < kSpread; k++) t += h[k * kSpreadStride];
    *out = t;
    return GXY_OK;
}

// Interleave the reference byte arrays into Blk words on the host.
static void interleave(const gxy_seq *s, Blk *out, uint64_t nblk) {
    const uint64_t len = s->len;
    const uint64_t ncode = (len + 3) / 4, nbit = (len + 7) / 8;
    for (uint64_t k = 0; k < nblk; k++) {
        uint64_t c = 0;
        uint32_t v = 0, m = 0;
        const uint64_t cb = 8 * k, bb = 4 * k;
        if (cb + 8 <= ncode) {
            std::memcpy(&c, s->codes + cb, 8);
        } else {
            for (uint64_t x = cb; x < ncode && x < cb + 8; x++) c |= (uint64_t)s->codes[x] << (8 * (x - cb));
        }
        if (bb + 4 <= nbit) {
            std::memcpy(&v, s->valid + bb, 4);
            std::memcpy(&m, s->mask + bb, 4);
        } else {
            for (uint64_t x = bb; x < nbit && x < bb + 4; x++) {
                v |= (uint32_t)s->valid[x] << (8 * (x - bb));
                m |= (uint32_t)s->mask[x] << (8 * (x - bb));
            }
        }
        // Clear everything at or past `len` (block carries 32 bases).
        const uint64_t first = 32 * k;
        if (first >= len) {
            c = 0;
            v = m = 0;
        } else if (len - first < 32) {
            const uint32_t keep = (uint32_t)(len - first);
            v &= (1u << keep) - 1u;
            m &= (1u << keep) - 1u;
            c &= (keep == 32) ? ~0ull : ((1ull << (2 * keep)) - 1ull);
        }
        m &= v;                  // mask is only ever set on valid bases
        c &= spread2(v);    // invalid bases carry code 0
        out[k] = {c, v, m};
    }
}

}  // namespace gxy

using namespace gxy;

namespace {

__global__ void k_rc(const Blk *__restrict__ src, Blk *__restrict__ dst, uint64_t len, uint64_t nblk) {
    const uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nblk) return;
    const uint64_t o0 = 32 * b;  // first output position of this block
    if (o0 >= len) {
        dst[b] = {0, 0, 0};
        return;
    }
    // Output k (k < 32) comes from source position len-1-(o0+k).  The 32
    // sources are the window [s0, s0+32) with s0 = len-32-o0 (may be < 0).
    const int64_t s0 = (int64_t)len - 32 - (int64_t)o0;
    Win w;
    if (s0 >= 0) {
        w = load_win(src, (uint64_t)s0);
    } else {
        // Fewer than 32 sources remain: they sit at [0, 32+s0), shifted up.
        const Win z = load_win(src, 0);
        const uint32_t sh = (uint32_t)(-s0);
        w.codes = z.codes << (2 * sh);
        w.valid = z.valid << sh;
        w.mask = z.mask << sh;
    }
    Blk out;
    out.valid = __builtin_bitreverse32(w.valid);
    out.mask = __builtin_bitreverse32(w.mask) & out.valid;
    out.codes = (~rev2(w.codes)) & spread2(out.valid);  // complement2 (src/dna.rs:35-37)
    const uint64_t rem = len - o0;
    if (rem < 32) {
        const uint32_t keep = (uint32_t)rem;
        out.valid &= (1u << keep) - 1u;
        out.mask &= (1u << keep) - 1u;
        out.codes &= (1ull << (2 * keep)) - 1ull;
    }
    dst[b] = out;
}

__global__ void k_slice(const Blk *__restrict__ src, Blk *__restrict__ dst, uint64_t start,
                        uint64_t len, uint64_t nblk) {
    const uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nblk) return;
    const uint64_t o0 = 32 * b;
    if (o0 >= len) {
        dst[b] = {0, 0, 0};
        return;
    }
    Win w = load_win(src, start + o0);
    const uint64_t rem = len - o0;
    if (rem < 32) {
        const uint32_t keep = (uint32_t)rem;
        w.valid &= (1u << keep) - 1u;
        w.mask &= (1u << keep) - 1u;
        w.codes &= (1ull << (2 * keep)) - 1ull;
    }
    dst[b] = {w.codes, w.valid, w.mask};
}

// Blocks from the reference's byte arrays already in device memory (the
// device form of interleave above): one thread per block.
__global__ void k_from_arrays(const uint8_t *__restrict__ codes, const uint8_t *__restrict__ valid,
                              const uint8_t *__restrict__ mask, uint64_t len, uint64_t nblk, Blk *__restrict__ out) {
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nblk) return;
    const uint64_t ncode = (len + 3) / 4, nbit = (len + 7) / 8;
    uint64_t c = 0;
    uint32_t v = 0, m = 0;
    for (uint64_t x = 8 * k; x < ncode && x < 8 * k + 8; x++) c |= (uint64_t)codes[x] << (8 * (x - 8 * k));
    for (uint64_t x = 4 * k; x < nbit && x < 4 * k + 4; x++) {
        v |= (uint32_t)valid[x] << (8 * (x - 4 * k));
        m |= (uint32_t)mask[x] << (8 * (x - 4 * k));
    }
    const uint64_t first = 32 * k;
    if (first >= len) {
        c = 0;
        v = m = 0;
    } else if (len - first < 32) {
        const uint32_t keep = (uint32_t)(len - first);
        v &= (1u << keep) - 1u;
        m &= (1u << keep) - 1u;
        c &= (1ull << (2 * keep)) - 1ull;
    }
    m &= v;
    c &= spread2(v);
    out[k] = {c, v, m};
}

// Byte-per-base buffer (x-drop layout, target or query codes) including both
// pads: one thread per output dword.
__global__ void k_to_bytes(const Blk *__restrict__ src, uint64_t len, uint64_t ndw, int as_target,
                           uint32_t *__restrict__ buf) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ndw) return;
    uint32_t f = 0;
#pragma unroll
    for (int b = 0; b < 4; b++) {
        const int64_t p = (int64_t)(4 * i + b) - (int64_t)kBytePad;
        const bool in = p >= 0 && (uint64_t)p < len;
        const bool valid = in && base_valid(src, (uint64_t)p);
        const uint32_t code = in ? base_code(src, (uint64_t)p) : 0u;
        f |= (as_target ? target_byte(code, valid) : query_byte(code, valid)) << (8 * b);
    }
    buf[i] = f;
}


}  // namespace

namespace gxy {

int dseq_alloc(gxy_ctx *ctx, uint64_t len, gxy_dseq **out) {
    auto *s = new (std::nothrow) gxy_dseq();
    if (!s) return set_err(ctx, GXY_ENOMEM, "host allocation failed");
    s->len = len;
    s->nblk = nblocks(len);
    s->device = ctx->device;
    const size_t bytes = s->nblk * sizeof(Blk);
    s->d = (Blk *)dev_alloc(ctx, bytes, &s->pool);
    if (!s->d) {
        delete s;
        return set_err(ctx, GXY_ENOMEM, "device allocation of %llu bytes failed",
                       (unsigned long long)bytes);
    }
    *out = s;
    return GXY_OK;
}

}  // namespace gxy

int gxy::blocks_to_bytes(gxy_ctx *ctx, const Blk *b, uint64_t len, bool as_target, uint8_t *buf) {
    const uint64_t ndw = (byte_buffer_size(len) + 3) / 4;
    k_to_bytes<<<(unsigned)((ndw + 255) / 256), 256, 0, ctx->stream>>>(b, len, ndw, as_target ? 1 : 0,
                                                                       (uint32_t *)buf);
    GXY_LAUNCH_CHECK(ctx);
    return GXY_OK;
}

int gxy::seq_to_bytes(gxy_ctx *ctx, const gxy_dseq *s, bool as_target, uint8_t *buf) {
    return blocks_to_bytes(ctx, s->d, s->len, as_target, buf);
}

int gxy::blocks_from_arrays(gxy_ctx *ctx, const uint8_t *codes, const uint8_t *valid, const uint8_t *mask,
                            uint64_t len, Blk *out) {
    const uint64_t nblk = nblocks(len);
    k_from_arrays<<<(unsigned)((nblk + 255) / 256), 256, 0, ctx->stream>>>(codes, valid, mask, len, nblk, out);
    GXY_LAUNCH_CHECK(ctx);
    return GXY_OK;
}

int gxy::dseq_upload_impl(gxy_ctx *ctx, const gxy_seq *seq, gxy_dseq **out) {
    if (!seq || !out || (seq->len && (!seq->codes || !seq->valid || !seq->mask)))
        return set_err(ctx, GXY_EINVAL, "gxy_dseq_upload: invalid sequence view");
    gxy_dseq *s;
    int rc = dseq_alloc(ctx, seq->len, &s);
    if (rc) return rc;
    std::vector<Blk> host(s->nblk);
    interleave(seq, host.data(), s->nblk);
    hipError_t e = hipMemcpyAsync(s->d, host.data(), s->nblk * sizeof(Blk), hipMemcpyHostToDevice,
                                  ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        gxy_dseq_free(s);
        return set_err(ctx, GXY_EDEVICE, "upload failed: %s", hipGetErrorString(e));
    }
    *out = s;
    return GXY_OK;
}

extern "C" {

const char *gxy_version(void) { return "gxy-mi355x 0.1 (gfx950)"; }

int gxy_ctx_create(int device, gxy_ctx **out) {
    return guarded(nullptr, [&]() -> int {
        if (!out) return set_err(nullptr, GXY_EINVAL, "gxy_ctx_create: out is NULL");
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
            (void)hipGetLastError();
            return set_err(nullptr, GXY_EDEVICE, "no HIP device available");
        }
        if (device < 0 || device >= n) return set_err(nullptr, GXY_EINVAL, "device %d out of range", device);
        if (hipSetDevice(device) != hipSuccess) return set_err(nullptr, GXY_EDEVICE, "hipSetDevice failed");
        auto *c = new (std::nothrow) gxy_ctx();
        if (!c) return set_err(nullptr, GXY_ENOMEM, "host allocation failed");
        c->device = device;
        if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
            delete c;
            return set_err(nullptr, GXY_EDEVICE, "hipStreamCreate failed");
        }
        int pools = 0;
        if (hipDeviceGetAttribute(&pools, hipDeviceAttributeMemoryPoolsSupported, device) != hipSuccess) {
            (void)hipGetLastError();
            pools = 0;
        }
        c->async_alloc = pools != 0;
        *out = c;
        return GXY_OK;
    });
}

int gxy_ctx_set_timing(gxy_ctx *ctx, int enabled) {
    return guarded(ctx, [&]() -> int {
        if (!ctx) return GXY_EINVAL;
        ctx->timing = enabled != 0;
        return GXY_OK;
    });
}

int gxy_ctx_release_cache(gxy_ctx *ctx) {
    return guarded(ctx, [&]() -> int {
        if (!ctx) return GXY_EINVAL;
        GXY_HIP(ctx, hipSetDevice(ctx->device));
        for (gxy_ctx *c : {ctx, ctx->helper, ctx->lane, ctx->lane ? ctx->lane->helper : nullptr}) {
            if (!c) continue;
            // a helper holds no stream between runs: its work has completed
            if (c->stream) GXY_HIP(ctx, hipStreamSynchronize(c->stream));
            for (int k = 0; k < gxy_ctx::kSlots; k++) scratch_release(c, k, 0);
            pool_trim(c);
            if (c->stream) GXY_HIP(ctx, hipStreamSynchronize(c->stream));
        }
        trim_device_pool(ctx);
        return GXY_OK;
    });
}

int gxy_ctx_set_dp_pool_rows(gxy_ctx *ctx, uint64_t rows) {
    return guarded(ctx, [&]() -> int {
        if (!ctx) return GXY_EINVAL;
        ctx->dp_pool_rows = rows;
        return GXY_OK;
    });
}

int gxy_ctx_set_dp_kernel(gxy_ctx *ctx, int kernel) {
    if (!ctx) return GXY_EINVAL;
    if (kernel < 0 || kernel > GXY_DP_WAVE_K4_W4) return set_err(ctx, GXY_EINVAL, "unknown DP kernel %d", kernel);
    ctx->dp_kernel = kernel;
    return GXY_OK;
}

int gxy_ctx_dp_autotune_state(gxy_ctx *ctx, int *chosen, double *gcups) {
    return guarded(ctx, [&]() -> int {
        if (!ctx || !chosen || !gcups) return GXY_EINVAL;
        for (int c = 0; c < gxy_ctx::kDpClasses; c++) {
            chosen[c] = ctx->dp_tune[c].chosen;
            for (int k = 0; k < gxy_ctx::kDpCands; k++) gcups[c * gxy_ctx::kDpCands + k] = ctx->dp_tune[c].gcups[k];
        }
        return GXY_OK;
    });
}

int gxy_ctx_dp_pool_stats(gxy_ctx *ctx, uint64_t *passes, uint64_t *truncated, uint64_t *alone) {
    return guarded(ctx, [&]() -> int {
        if (!ctx || !passes || !truncated || !alone) return GXY_EINVAL;
        *passes = ctx->dp_passes;
        *truncated = ctx->dp_truncated;
        *alone = ctx->dp_alone;
        return GXY_OK;
    });
}

int gxy_ctx_set_dp_persist(gxy_ctx *ctx, int mode) {
    if (!ctx) return GXY_EINVAL;
    if (mode < -1 || mode > 1) return set_err(ctx, GXY_EINVAL, "unknown persistent-DP mode %d", mode);
    ctx->dp_persist = mode;
    return GXY_OK;
}

int gxy_ctx_dp_persist_stats(gxy_ctx *ctx, uint64_t *launches, uint64_t *sides, uint64_t *deferred,
                             uint64_t *cells) {
    return guarded(ctx, [&]() -> int {
        if (!ctx || !launches || !sides || !deferred || !cells) return GXY_EINVAL;
        *launches = ctx->dp_persist_launches;
        *sides = ctx->dp_persist_sides;
        *deferred = ctx->dp_persist_deferred;
        *cells = ctx->dp_persist_cells;
        return GXY_OK;
    });
}

int gxy_ctx_set_xdrop_form(gxy_ctx *ctx, int form) {
    if (!ctx) return GXY_EINVAL;
    if (form < 0 || form > 1) return set_err(ctx, GXY_EINVAL, "unknown x-drop form %d", form);
    ctx->xdrop_single = form == 1;
    return GXY_OK;
}

int gxy_ctx_set_hit_batch(gxy_ctx *ctx, uint64_t hits) {
    if (!ctx) return GXY_EINVAL;
    ctx->hit_batch = hits;
    return GXY_OK;
}

int gxy_ctx_xdrop_stats(gxy_ctx *ctx, uint64_t *first_step_columns, uint64_t *resumed, uint64_t *first_step_passing) {
    return guarded(ctx, [&]() -> int {
        if (!ctx || !first_step_columns || !resumed || !first_step_passing) return GXY_EINVAL;
        *first_step_columns = ctx->xd_first_cols;
        *resumed = ctx->xd_conts;
        *first_step_passing = ctx->xd_first_pass;
        return GXY_OK;
    });
}

int gxy_ctx_kernel_time(gxy_ctx *ctx, const char *name, double *ms, uint64_t *launches) {
    return guarded(ctx, [&]() -> int {
        if (!ctx || !name || !ms || !launches) return GXY_EINVAL;
        timing_flush(ctx);
        *ms = 0.0;
        *launches = 0;
        for (auto &t : ctx->totals)
            if (t.first == name) {
                *ms = t.second.first;
                *launches = t.second.second;
            }
        return GXY_OK;
    });
}

int gxy_ctx_reset_timing(gxy_ctx *ctx) {
    return guarded(ctx, [&]() -> int {
        if (!ctx) return GXY_EINVAL;
        timing_flush(ctx);
        ctx->totals.clear();
        return GXY_OK;
    });
}

void gxy_ctx_destroy(gxy_ctx *ctx) {
    if (!ctx) return;
    if (ctx->helper) gxy_ctx_destroy(ctx->helper);
    if (ctx->lane) gxy_ctx_destroy(ctx->lane);
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    timing_flush(ctx);
    for (hipEvent_t e : ctx->event_pool) (void)hipEventDestroy(e);
    for (hipEvent_t e : ctx->dp_ev)
        if (e) (void)hipEventDestroy(e);
    for (int k = 0; k < gxy_ctx::kSlots; k++)
        if (ctx->slot[k]) (void)(ctx->async_alloc ? hipFreeAsync(ctx->slot[k], ctx->stream) : hipFree(ctx->slot[k]));
    (void)hipStreamSynchronize(ctx->stream);
    pool_trim(ctx);
    trim_device_pool(ctx);
    if (ctx->pinned) (void)hipHostFree(ctx->pinned);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);  // a run's helper may hold none between runs
    delete ctx;
}

const char *gxy_last_error(const gxy_ctx *ctx) {
    return ctx ? ctx->err.c_str() : g_create_err.c_str();
}

int gxy_ctx_synchronize(gxy_ctx *ctx) {
    return guarded(ctx, [&]() -> int {
        if (!ctx) return GXY_EINVAL;
        GXY_HIP(ctx, hipStreamSynchronize(ctx->stream));
        return GXY_OK;
    });
}

int gxy_pack_ascii(const char *ascii, uint64_t len, uint8_t *codes, uint8_t *valid, uint8_t *mask) {
    return guarded(nullptr, [&]() -> int {
        // PackedSeq::from_ascii / push_ascii (src/sequences.rs:164-206) with
        // encode_base (src/dna.rs:12-20).
        if ((len && !ascii) || !codes || !valid || !mask) return GXY_EINVAL;
        std::memset(codes, 0, (len + 3) / 4);
        std::memset(valid, 0, (len + 7) / 8);
        std::memset(mask, 0, (len + 7) / 8);
        for (uint64_t i = 0; i < len; i++) {
            const unsigned char b = (unsigned char)ascii[i];
            uint8_t c = 0;
            bool v = true;
            switch (b) {
                case 'A': case 'a': c = 0; break;
                case 'C': case 'c': c = 1; break;
                case 'G': case 'g': c = 2; break;
                case 'T': case 't': case 'U': case 'u': c = 3; break;
                default: v = false; break;
            }
            if (!v) continue;
            codes[i >> 2] |= (uint8_t)(c << (2 * (i & 3)));
            valid[i >> 3] |= (uint8_t)(1u << (i & 7));
            if (b >= 'a' && b <= 'z') mask[i >> 3] |= (uint8_t)(1u << (i & 7));
        }
        return GXY_OK;
    });
}

int gxy_dseq_upload(gxy_ctx *ctx, const gxy_seq *seq, gxy_dseq **out) {
    return guarded(ctx, [&]() -> int {
        if (!ctx) return GXY_EINVAL;
        return dseq_upload_impl(ctx, seq, out);
    });
}

int gxy_dseq_reverse_complement(gxy_ctx *ctx, const gxy_dseq *seq, gxy_dseq **out) {
    return guarded(ctx, [&]() -> int {
        if (!ctx || !seq || !out) return ctx ? set_err(ctx, GXY_EINVAL, "null argument") : GXY_EINVAL;
        gxy_dseq *r;
        int rc = dseq_alloc(ctx, seq->len, &r);
        if (rc) return rc;
        const unsigned th = 256;
        k_rc<<<(unsigned)((r->nblk + th - 1) / th), th, 0, ctx->stream>>>(seq->d, r->d, seq->len, r->nblk);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) {
            gxy_dseq_free(r);
            return set_err(ctx, GXY_EDEVICE, "reverse complement failed: %s", hipGetErrorString(e));
        }
        *out = r;
        return GXY_OK;
    });
}

int gxy_dseq_slice(gxy_ctx *ctx, const gxy_dseq *seq, uint64_t start, uint64_t end, gxy_dseq **out) {
    return guarded(ctx, [&]() -> int {
        if (!ctx || !seq || !out) return ctx ? set_err(ctx, GXY_EINVAL, "null argument") : GXY_EINVAL;
        if (start > end || end > seq->len)  // src/sequences.rs:278
            return set_err(ctx, GXY_EINVAL, "slice [%llu, %llu) out of range (len %llu)",
                           (unsigned long long)start, (unsigned long long)end,
                           (unsigned long long)seq->len);
        gxy_dseq *r;
        int rc = dseq_alloc(ctx, end - start, &r);
        if (rc) return rc;
        const unsigned th = 256;
        k_slice<<<(unsigned)((r->nblk + th - 1) / th), th, 0, ctx->stream>>>(seq->d, r->d, start, end - start,
                                                                              r->nblk);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) {
            gxy_dseq_free(r);
            return set_err(ctx, GXY_EDEVICE, "slice failed: %s", hipGetErrorString(e));
        }
        *out = r;
        return GXY_OK;
    });
}

uint64_t gxy_dseq_len(const gxy_dseq *seq) { return seq ? seq->len : 0; }

int gxy_dseq_download(gxy_ctx *ctx, const gxy_dseq *seq, uint8_t *codes, uint8_t *valid, uint8_t *mask) {
    return guarded(ctx, [&]() -> int {
        if (!ctx || !seq || !codes || !valid || !mask) return GXY_EINVAL;
        std::vector<Blk> host(seq->nblk);
        GXY_HIP(ctx, hipMemcpyAsync(host.data(), seq->d, seq->nblk * sizeof(Blk), hipMemcpyDeviceToHost,
                                    ctx->stream));
        GXY_HIP(ctx, hipStreamSynchronize(ctx->stream));
        const uint64_t len = seq->len, ncode = (len + 3) / 4, nbit = (len + 7) / 8;
        for (uint64_t x = 0; x < ncode; x++) codes[x] = (uint8_t)(host[x / 8].codes >> (8 * (x % 8)));
        for (uint64_t x = 0; x < nbit; x++) {
            valid[x] = (uint8_t)(host[x / 4].valid >> (8 * (x % 4)));
            mask[x] = (uint8_t)(host[x / 4].mask >> (8 * (x % 4)));
        }
        return GXY_OK;
    });
}

void gxy_dseq_free(gxy_dseq *seq) {
    if (!seq) return;
    dev_free(seq->pool, seq->d);
    delete seq;
}

void gxy_free(void *p) { std::free(p); }

}  // extern "C"
