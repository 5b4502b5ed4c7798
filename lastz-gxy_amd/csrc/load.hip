// load.hip — input loaders straight to the device layout: UCSC .2bit
// (src/bit2.rs:73-164) and FASTA (src/sequences.rs:338-377).
//
// .2bit: the file's 2-bit DNA (4 bases per byte, first base in the high
// bits, T=00 C=01 A=10 G=11) is uploaded as it is; one thread per 32-base
// output block reverses the four fields of each byte and maps the codes to
// A=0 C=1 G=2 T=3 with two bit operations per field (lo' = ~hi,
// hi' = ~(hi ^ lo)), then clears the validity of N-block bases and sets the
// mask of mask-block bases (the blocks sorted and merged on the host; a
// block finds its intervals by binary search).  The reference decodes to
// ASCII, overlays N and lowercase, and re-packs (src/bit2.rs:134-164); the
// result is the same PackedSeq: N bases are invalid with code 0, masked
// bases are valid non-N bases inside a mask block.
//
// FASTA: the host splits records and drops line structure and whitespace
// (byte scanning); the device packs the ASCII, one thread per 32-base block
// (PackedSeq::push_ascii / encode_base, src/sequences.rs:164-206,
// src/dna.rs:12-20).
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "internal.h"

using namespace gxy;

namespace {

inline unsigned grid_for(uint64_t n, unsigned th) { return (unsigned)((n + th - 1) / th); }

struct Iv {
    uint32_t s, e;  // [s, e)
};

// Bits of block [o0, o0+32) covered by sorted disjoint intervals.
__device__ __forceinline__ uint32_t cover_bits(const Iv *__restrict__ iv, uint32_t n, uint64_t o0) {
    uint32_t lo = 0, hi = n;  // first interval with e > o0
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (iv[mid].e > o0)
            hi = mid;
        else
            lo = mid + 1;
    }
    uint32_t bits = 0;
    for (uint32_t k = lo; k < n && iv[k].s < o0 + 32; k++) {
        const uint64_t a = max((uint64_t)iv[k].s, o0), b = min((uint64_t)iv[k].e, o0 + 32);
        if (b > a) {
            const uint32_t w = (uint32_t)(b - a), sh = (uint32_t)(a - o0);
            bits |= (w >= 32 ? 0xFFFFFFFFu : ((1u << w) - 1u)) << sh;
        }
    }
    return bits;
}

__global__ void k_from_2bit(const uint8_t *__restrict__ packed, uint64_t len, const Iv *__restrict__ nb, uint32_t nn,
                            const Iv *__restrict__ mb, uint32_t nm, Blk *__restrict__ out, uint64_t nblk) {
    const uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nblk) return;
    const uint64_t o0 = 32 * b;
    if (o0 >= len) {
        out[b] = {0, 0, 0};
        return;
    }
    const uint64_t nbytes = (len + 3) / 4;
    uint64_t x = 0;
    if (8 * b + 8 <= nbytes) {
        x = *(const uint64_t *)(packed + 8 * b);
    } else {
        for (uint64_t k = 0; k < 8 && 8 * b + k < nbytes; k++) x |= (uint64_t)packed[8 * b + k] << (8 * k);
    }
    // first base of each byte in its high bits -> base i at bits 2i
    x = ((x >> 6) & 0x0303030303030303ull) | ((x >> 2) & 0x0C0C0C0C0C0C0C0Cull) |
        ((x << 2) & 0x3030303030303030ull) | ((x << 6) & 0xC0C0C0C0C0C0C0C0ull);
    const uint64_t m1 = 0x5555555555555555ull;
    const uint64_t hi = (x >> 1) & m1, lo = x & m1;
    uint64_t codes = (~hi & m1) | ((~(hi ^ lo) & m1) << 1);  // T,C,A,G -> 3,1,0,2
    const uint64_t rem = len - o0;
    uint32_t valid = rem >= 32 ? 0xFFFFFFFFu : ((1u << rem) - 1u);
    valid &= ~cover_bits(nb, nn, o0);
    const uint32_t mask = cover_bits(mb, nm, o0) & valid;
    codes &= spread2(valid);
    out[b] = {codes, valid, mask};
}

__global__ void k_pack_ascii(const uint8_t *__restrict__ ascii, uint64_t len, Blk *__restrict__ out, uint64_t nblk) {
    const uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nblk) return;
    const uint64_t o0 = 32 * b;
    uint64_t codes = 0;
    uint32_t valid = 0, mask = 0;
    for (uint32_t k = 0; k < 32 && o0 + k < len; k++) {
        const uint8_t c = ascii[o0 + k];
        const uint8_t u = c & 0xDF;  // upper case for letters
        uint32_t code = 0, v = 1;
        if (u == 'A')
            code = 0;
        else if (u == 'C')
            code = 1;
        else if (u == 'G')
            code = 2;
        else if (u == 'T' || u == 'U')
            code = 3;
        else
            v = 0;
        codes |= (uint64_t)code << (2 * k);
        valid |= v << k;
        mask |= (v && c >= 'a' && c <= 'z' ? 1u : 0u) << k;  // src/sequences.rs:126-129
    }
    out[b] = {codes, valid, mask};
}

struct Reader {
    const uint8_t *p;
    uint64_t n, pos = 0;
    bool be = false;
    bool ok = true;
    uint32_t u32() {
        if (pos + 4 > n) {
            ok = false;
            return 0;
        }
        uint32_t v;
        if (be)
            v = (uint32_t)p[pos] << 24 | (uint32_t)p[pos + 1] << 16 | (uint32_t)p[pos + 2] << 8 | p[pos + 3];
        else
            std::memcpy(&v, p + pos, 4);
        pos += 4;
        return v;
    }
};

std::vector<Iv> merged(const std::vector<uint32_t> &starts, const std::vector<uint32_t> &sizes, uint64_t len) {
    std::vector<Iv> v;
    for (size_t k = 0; k < starts.size(); k++) {
        const uint64_t s = starts[k], e = std::min<uint64_t>((uint64_t)starts[k] + sizes[k], len);
        if (e > s) v.push_back({(uint32_t)s, (uint32_t)e});
    }
    std::sort(v.begin(), v.end(), [](const Iv &a, const Iv &b) { return a.s < b.s; });
    std::vector<Iv> out;
    for (const Iv &x : v) {
        if (!out.empty() && x.s <= out.back().e)
            out.back().e = std::max(out.back().e, x.e);
        else
            out.push_back(x);
    }
    return out;
}

gxy_dseq_list *new_list(uint64_t n) {
    auto *l = (gxy_dseq_list *)std::calloc(1, sizeof(gxy_dseq_list));
    if (!l) return nullptr;
    l->names = (char **)std::calloc(std::max<uint64_t>(1, n), sizeof(char *));
    l->seqs = (gxy_dseq **)std::calloc(std::max<uint64_t>(1, n), sizeof(gxy_dseq *));
    if (!l->names || !l->seqs) {
        std::free(l->names);
        std::free(l->seqs);
        std::free(l);
        return nullptr;
    }
    return l;
}

char *dup(const uint8_t *s, size_t n) {
    char *c = (char *)std::malloc(n + 1);
    if (c) {
        std::memcpy(c, s, n);
        c[n] = 0;
    }
    return c;
}

}  // namespace

extern "C" {

void gxy_dseq_list_free(gxy_dseq_list *l) {
    if (!l) return;
    for (uint64_t k = 0; k < l->n; k++) {
        std::free(l->names[k]);
        gxy_dseq_free(l->seqs[k]);
    }
    std::free(l->names);
    std::free(l->seqs);
    std::free(l);
}

int gxy_load_2bit(gxy_ctx *ctx, const uint8_t *bytes, uint64_t len, gxy_dseq_list **out) {
    return guarded(ctx, [&]() -> int {
        if (!ctx) return GXY_EINVAL;
        if ((!bytes && len) || !out) return set_err(ctx, GXY_EINVAL, "null argument");
        GXY_HIP(ctx, hipSetDevice(ctx->device));
        // src/bit2.rs:73-95: magic decides the byte order
        if (len < 4) return set_err(ctx, GXY_EINVAL, "truncated 2bit file (needed 4 bytes at offset 0)");
        uint32_t magic;
        std::memcpy(&magic, bytes, 4);
        Reader r{bytes, len};
        if (magic == 0x1A412743u)
            r.be = false;
        else if (magic == 0x4327411Au)
            r.be = true;
        else
            return set_err(ctx, GXY_EINVAL, "not a 2bit file (magic = 0x%08x)", magic);
        r.pos = 4;
        const uint32_t version = r.u32();
        if (r.ok && version != 0) return set_err(ctx, GXY_EINVAL, "unsupported 2bit version %u", version);
        const uint32_t count = r.u32();
        (void)r.u32();  // reserved
        if (!r.ok) return set_err(ctx, GXY_EINVAL, "truncated 2bit header");
        std::vector<std::pair<std::string, uint32_t>> index;  // src/bit2.rs:98-106
        for (uint32_t k = 0; k < count; k++) {
            if (r.pos + 1 > len) return set_err(ctx, GXY_EINVAL, "truncated 2bit index");
            const uint32_t ns = bytes[r.pos++];
            if (r.pos + ns > len) return set_err(ctx, GXY_EINVAL, "truncated 2bit index");
            std::string name((const char *)bytes + r.pos, ns);
            r.pos += ns;
            const uint32_t off = r.u32();
            if (!r.ok) return set_err(ctx, GXY_EINVAL, "truncated 2bit index");
            index.push_back({std::move(name), off});
        }
        gxy_dseq_list *list = new_list(count);
        if (!list) return set_err(ctx, GXY_ENOMEM, "host allocation failed");
        auto fail = [&](int code) {
            gxy_dseq_list_free(list);
            return code;
        };
        for (const auto &ix : index) {  // parse_sequence_record (src/bit2.rs:115-164)
            Reader s{bytes, len};
            s.be = r.be;
            s.pos = ix.second;
            if (s.pos > len) return fail(set_err(ctx, GXY_EINVAL, "2bit record offset past the end"));
            const uint32_t dna = s.u32();
            const uint32_t nn = s.u32();
            // block counts are checked against the bytes left before anything is
            // sized from them (8 bytes per block: start and size)
            if (!s.ok || nn > (len - s.pos) / 8) return fail(set_err(ctx, GXY_EINVAL, "truncated 2bit record"));
            std::vector<uint32_t> ns(nn), nz(nn);
            for (auto &x : ns) x = s.u32();
            for (auto &x : nz) x = s.u32();
            const uint32_t nm = s.ok ? s.u32() : 0;
            if (!s.ok || nm > (len - s.pos) / 8) return fail(set_err(ctx, GXY_EINVAL, "truncated 2bit record"));
            std::vector<uint32_t> ms(nm), mz(nm);
            for (auto &x : ms) x = s.u32();
            for (auto &x : mz) x = s.u32();
            (void)s.u32();  // reserved
            const uint64_t pb = ((uint64_t)dna + 3) / 4;
            if (!s.ok || s.pos + pb > len) return fail(set_err(ctx, GXY_EINVAL, "truncated 2bit record"));
            const std::vector<Iv> niv = merged(ns, nz, dna), miv = merged(ms, mz, dna);
            gxy_dseq *d;
            int rc = dseq_alloc(ctx, dna, &d);
            if (rc) return fail(rc);
            list->names[list->n] = dup((const uint8_t *)ix.first.data(), ix.first.size());
            list->seqs[list->n] = d;
            list->n++;
            const size_t bytes_dev = (size_t)pb + niv.size() * sizeof(Iv) + miv.size() * sizeof(Iv) + 16;
            auto *stage = (uint8_t *)scratch(ctx, kSlotMisc0, bytes_dev);
            if (!stage) return fail(set_err(ctx, GXY_ENOMEM, "scratch allocation failed"));
            const size_t o_n = (pb + 7) & ~(size_t)7, o_m = o_n + niv.size() * sizeof(Iv);
            if (pb) GXY_HIP(ctx, hipMemcpyAsync(stage, bytes + s.pos, pb, hipMemcpyHostToDevice, ctx->stream));
            if (!niv.empty())
                GXY_HIP(ctx, hipMemcpyAsync(stage + o_n, niv.data(), niv.size() * sizeof(Iv), hipMemcpyHostToDevice,
                                            ctx->stream));
            if (!miv.empty())
                GXY_HIP(ctx, hipMemcpyAsync(stage + o_m, miv.data(), miv.size() * sizeof(Iv), hipMemcpyHostToDevice,
                                            ctx->stream));
            k_from_2bit<<<grid_for(d->nblk, 256), 256, 0, ctx->stream>>>(stage, dna, (const Iv *)(stage + o_n),
                                                                         (uint32_t)niv.size(), (const Iv *)(stage + o_m),
                                                                         (uint32_t)miv.size(), d->d, d->nblk);
            GXY_LAUNCH_CHECK(ctx);
            GXY_HIP(ctx, hipStreamSynchronize(ctx->stream));  // the staging slot is reused by the next record
        }
        *out = list;
        return GXY_OK;
    });
}

int gxy_load_fasta(gxy_ctx *ctx, const char *text, uint64_t len, gxy_dseq_list **out) {
    return guarded(ctx, [&]() -> int {
        if (!ctx) return GXY_EINVAL;
        if ((!text && len) || !out) return set_err(ctx, GXY_EINVAL, "null argument");
        GXY_HIP(ctx, hipSetDevice(ctx->device));
        // parse_fasta (src/sequences.rs:338-377): lines split at '\n' (a '\r'
        // before it dropped), empty lines skipped, '>' starts a record named by
        // its first whitespace-delimited token, other lines add their
        // non-whitespace bytes; bytes before the first header stay with the
        // first record.
        struct RecT {
            std::string name;
            std::string seq;
        };
        std::vector<RecT> recs;
        std::string cur;
        bool named = false;
        std::string cur_name;
        auto isws = [](unsigned char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == 0x0C; };
        uint64_t p = 0;
        uint64_t lineno = 0;
        while (p < len) {
            uint64_t e = p;
            while (e < len && text[e] != '\n') e++;
            uint64_t le = e;
            if (le > p && text[le - 1] == '\r') le--;
            lineno++;
            if (le > p) {
                if (text[p] == '>') {
                    if (named) {
                        recs.push_back({std::move(cur_name), std::move(cur)});
                        cur.clear();
                    }
                    uint64_t a = p + 1;
                    while (a < le && isws((unsigned char)text[a])) a++;
                    uint64_t b = a;
                    while (b < le && !isws((unsigned char)text[b])) b++;
                    if (a == b) return set_err(ctx, GXY_EINVAL, "empty FASTA header on line %llu", (unsigned long long)lineno);
                    cur_name.assign(text + a, b - a);
                    named = true;
                } else {
                    for (uint64_t k = p; k < le; k++)
                        if (!isws((unsigned char)text[k])) cur.push_back(text[k]);
                }
            }
            p = e + 1;
        }
        if (named) recs.push_back({std::move(cur_name), std::move(cur)});
        gxy_dseq_list *list = new_list(recs.size());
        if (!list) return set_err(ctx, GXY_ENOMEM, "host allocation failed");
        for (const auto &rc_ : recs) {
            const uint64_t n = rc_.seq.size();
            gxy_dseq *d;
            int rc = dseq_alloc(ctx, n, &d);
            if (rc) {
                gxy_dseq_list_free(list);
                return rc;
            }
            list->names[list->n] = dup((const uint8_t *)rc_.name.data(), rc_.name.size());
            list->seqs[list->n] = d;
            list->n++;
            auto *stage = (uint8_t *)scratch(ctx, kSlotMisc0, n + 16);
            if (!stage) {
                gxy_dseq_list_free(list);
                return set_err(ctx, GXY_ENOMEM, "scratch allocation failed");
            }
            if (n) GXY_HIP(ctx, hipMemcpyAsync(stage, rc_.seq.data(), n, hipMemcpyHostToDevice, ctx->stream));
            k_pack_ascii<<<grid_for(d->nblk, 256), 256, 0, ctx->stream>>>(stage, n, d->d, d->nblk);
            GXY_LAUNCH_CHECK(ctx);
            GXY_HIP(ctx, hipStreamSynchronize(ctx->stream));
        }
        *out = list;
        return GXY_OK;
    });
}

}  // extern "C"
