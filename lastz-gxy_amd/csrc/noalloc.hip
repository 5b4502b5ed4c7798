// noalloc.hip — batched extend_hit for callers that cannot allocate device
// memory (code under HIP graph capture, tight serving loops, harnesses that
// forbid hipMalloc).  The entry point takes the same arguments as
// gxy_extend_hit_batch and runs the same kernels, but every device buffer —
// the staged sequences, their block and byte layouts, the hit positions, the
// results and the column counter — lives in one file-scope device array of
// kNoAllocBytes, so the call makes no hipMalloc, hipFree or allocating
// library call.  Each GPU holds its own copy of the array until the process
// ends; its address is looked up once per device (hipGetSymbolAddress, with
// that device current) and only that address is handed to kernels and copies.
// Calls on one device are serialised by a per-device host lock: they share
// the array, and calls that overlapped would race on it.
#include <mutex>

#include "internal.h"

namespace {

constexpr size_t kNoAllocBytes = 64ull << 20;  // the array: 64 MiB per device
constexpr int kMaxDevices = 64;

static __device__ __align__(256) uint8_t scratch[kNoAllocBytes];

struct DeviceSlot {
    std::mutex mu;          // held for the whole call: one call at a time per device
    uint8_t *ptr = nullptr; // device address of `scratch` on this device (null until looked up)
};
DeviceSlot g_slots[kMaxDevices];

inline size_t up256(size_t x) { return (x + 255) & ~(size_t)255; }

// Regions of the array for one call, in bytes from its start.
struct Layout {
    size_t colsum, tpos, qpos, out, found, tcodes, tvalid, tmask, qcodes, qvalid, qmask, tblk, qblk, tbytes,
        qbytes, total;
};

Layout layout_for(uint64_t tlen, uint64_t qlen, uint64_t n, bool bytes) {
    Layout L;
    size_t o = 0;
    auto take = [&](size_t b) {
        const size_t at = o;
        o = up256(o + b);
        return at;
    };
    L.colsum = take(8);
    L.tpos = take(n * 4);
    L.qpos = take(n * 4);
    L.out = take(n * sizeof(gxy_hsp));
    L.found = take(n);
    L.tcodes = take((tlen + 3) / 4);
    L.tvalid = take((tlen + 7) / 8);
    L.tmask = take((tlen + 7) / 8);
    L.qcodes = take((qlen + 3) / 4);
    L.qvalid = take((qlen + 7) / 8);
    L.qmask = take((qlen + 7) / 8);
    L.tblk = take(gxy::nblocks(tlen) * sizeof(gxy::Blk));
    L.qblk = take(gxy::nblocks(qlen) * sizeof(gxy::Blk));
    L.tbytes = take(bytes ? gxy::byte_buffer_size(tlen) : 0);
    L.qbytes = take(bytes ? gxy::byte_buffer_size(qlen) : 0);
    L.total = o;
    return L;
}

}  // namespace

extern "C" {

int gxy_extend_hit_batch_noalloc(gxy_ctx *ctx, const gxy_seq *target, const gxy_seq *query,
                                 const gxy_scoring *scoring, const gxy_hsp_params *params, const uint32_t *t_pos,
                                 const uint32_t *q_pos, uint64_t n, uint32_t seed_len, gxy_hsp *out,
                                 uint8_t *found) {
    return gxy::guarded(ctx, [&]() -> int {
        using gxy::set_err;
        if (!ctx) return GXY_EINVAL;
        if (!target || !query || !scoring || !params || (n && (!t_pos || !q_pos || !out || !found)))
            return set_err(ctx, GXY_EINVAL, "null argument");
        if ((target->len && (!target->codes || !target->valid || !target->mask)) ||
            (query->len && (!query->codes || !query->valid || !query->mask)))
            return set_err(ctx, GXY_EINVAL, "invalid sequence view");
        if (ctx->device < 0 || ctx->device >= kMaxDevices) return set_err(ctx, GXY_EINVAL, "device out of range");
        if (n == 0) return GXY_OK;
        const bool bytes = gxy::extend_uses_bytes(*scoring, *params);
        const Layout L = layout_for(target->len, query->len, n, bytes);
        if (L.total > kNoAllocBytes)
            return set_err(ctx, GXY_ERANGE,
                           "gxy_extend_hit_batch_noalloc: %llu bytes of device scratch needed, the limit is %llu "
                           "(about 2 bytes per target and query base plus 25 bytes per hit)",
                           (unsigned long long)L.total, (unsigned long long)kNoAllocBytes);
        GXY_HIP(ctx, hipSetDevice(ctx->device));
        DeviceSlot &slot = g_slots[ctx->device];
        std::lock_guard<std::mutex> lock(slot.mu);
        if (!slot.ptr) {  // first call on this device: the array's address here
            void *p = nullptr;
            const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !p)
                return set_err(ctx, GXY_EDEVICE, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
            slot.ptr = (uint8_t *)p;
        }
        uint8_t *const base = slot.ptr;
        const hipStream_t st = ctx->stream;
        auto up = [&](size_t at, const void *src, size_t bytes_) {
            return bytes_ ? hipMemcpyAsync(base + at, src, bytes_, hipMemcpyHostToDevice, st) : hipSuccess;
        };
        hipError_t e = up(L.tpos, t_pos, n * 4);
        if (e == hipSuccess) e = up(L.qpos, q_pos, n * 4);
        if (e == hipSuccess) e = up(L.tcodes, target->codes, (target->len + 3) / 4);
        if (e == hipSuccess) e = up(L.tvalid, target->valid, (target->len + 7) / 8);
        if (e == hipSuccess) e = up(L.tmask, target->mask, (target->len + 7) / 8);
        if (e == hipSuccess) e = up(L.qcodes, query->codes, (query->len + 3) / 4);
        if (e == hipSuccess) e = up(L.qvalid, query->valid, (query->len + 7) / 8);
        if (e == hipSuccess) e = up(L.qmask, query->mask, (query->len + 7) / 8);
        if (e != hipSuccess) return set_err(ctx, GXY_EDEVICE, "extend_hit_batch_noalloc: %s", hipGetErrorString(e));
        auto *tblk = (gxy::Blk *)(base + L.tblk);
        auto *qblk = (gxy::Blk *)(base + L.qblk);
        int rc = gxy::blocks_from_arrays(ctx, base + L.tcodes, base + L.tvalid, base + L.tmask, target->len, tblk);
        if (!rc) rc = gxy::blocks_from_arrays(ctx, base + L.qcodes, base + L.qvalid, base + L.qmask, query->len, qblk);
        if (!rc)
            rc = gxy::extend_hits_dev(ctx, tblk, target->len, qblk, query->len, bytes ? base + L.tbytes : nullptr,
                                      bytes ? base + L.qbytes : nullptr, (const uint32_t *)(base + L.tpos),
                                      (const uint32_t *)(base + L.qpos), n, *scoring, *params, seed_len,
                                      (gxy_hsp *)(base + L.out), base + L.found,
                                      (unsigned long long *)(base + L.colsum));
        if (rc) {
            (void)hipStreamSynchronize(st);  // nothing of this call may still use the array
            return rc;
        }
        e = hipMemcpyAsync(out, base + L.out, n * sizeof(gxy_hsp), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipMemcpyAsync(found, base + L.found, n, hipMemcpyDeviceToHost, st);
        const hipError_t es = hipStreamSynchronize(st);  // the array is free again when the call returns
        if (e == hipSuccess) e = es;
        if (e != hipSuccess) return set_err(ctx, GXY_EDEVICE, "extend_hit_batch_noalloc: %s", hipGetErrorString(e));
        return GXY_OK;
    });
}

}  // extern "C"
