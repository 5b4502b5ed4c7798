// oracle.cpp — CPU restatement of lastz-gxy's seed -> HSP -> gapped path.
//
// TEST INFRASTRUCTURE ONLY: the parity checker and CPU baseline for the HIP
// library.  Never linked into lastz-gxy_amd/.  Every function restates the
// reference function cited beside it (paths relative to the reference crate)
// with the same integer arithmetic, including saturating adds where the
// reference saturates.
//
// Canonical choices where the reference leaves an order unspecified (these are
// the only places the reference itself is nondeterministic or depends on the
// Rust std version):
//   * `sort_unstable_by_key((diag, t_start))` (src/seed_search.rs:122): equal
//     keys keep seed-scan order (a stable sort).  Counted in tie_pairs.
//   * `sort_unstable_by_key(Reverse(score))` (src/driver.rs:270): equal scores
//     keep (diag, t_start) order.  Counted in score_ties.
//   * rayon push order of chunk groups (src/driver.rs:361-372): ascending chunk.

#include "oracle.h"

#include <immintrin.h>

#include <algorithm>
#include <atomic>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <utility>
#include <vector>

namespace {

constexpr int32_t kNegInf = INT32_MIN / 4;  // src/gapped_extend.rs:44

inline uint32_t code_at(const gxy_seq &s, uint64_t i) {  // src/sequences.rs:211-214
    return (s.codes[i >> 2] >> ((i & 3) * 2)) & 3u;
}
inline bool valid_at(const gxy_seq &s, uint64_t i) {  // src/sequences.rs:217-219
    return (s.valid[i >> 3] >> (i & 7)) & 1;
}
inline bool masked_at(const gxy_seq &s, uint64_t i) {  // src/sequences.rs:223-225
    return (s.mask[i >> 3] >> (i & 7)) & 1;
}

inline int32_t sat_add(int32_t a, int32_t b) {
    int64_t r = (int64_t)a + (int64_t)b;
    if (r > INT32_MAX) return INT32_MAX;
    if (r < INT32_MIN) return INT32_MIN;
    return (int32_t)r;
}
inline int32_t sat_sub(int32_t a, int32_t b) {
    int64_t r = (int64_t)a - (int64_t)b;
    if (r > INT32_MAX) return INT32_MAX;
    if (r < INT32_MIN) return INT32_MIN;
    return (int32_t)r;
}

// encode_base (src/dna.rs:12-20)
inline void encode_base(uint8_t b, uint8_t *code, bool *valid) {
    switch (b) {
        case 'A': case 'a': *code = 0; *valid = true; return;
        case 'C': case 'c': *code = 1; *valid = true; return;
        case 'G': case 'g': *code = 2; *valid = true; return;
        case 'T': case 't': case 'U': case 'u': *code = 3; *valid = true; return;
        default: *code = 0; *valid = false; return;
    }
}
inline uint8_t decode_base(uint32_t c) { return "ACGT"[c & 3]; }  // src/dna.rs:24-31

// Owned PackedSeq (src/sequences.rs:136-141).
struct OSeq {
    std::vector<uint8_t> codes, valid, mask;
    uint64_t len = 0;

    void reserve(uint64_t n) {
        codes.reserve((n + 3) / 4);
        valid.reserve((n + 7) / 8);
        mask.reserve((n + 7) / 8);
    }
    // PackedSeq::push_ascii (src/sequences.rs:182-206)
    void push_ascii(uint8_t b) {
        uint8_t code;
        bool v;
        encode_base(b, &code, &v);
        uint64_t bit_off = (len & 3) * 2;
        if (bit_off == 0) codes.push_back(0);
        codes.back() |= (uint8_t)((code & 3) << bit_off);
        uint64_t vbit = len & 7;
        if (vbit == 0) {
            valid.push_back(0);
            mask.push_back(0);
        }
        if (v) valid.back() |= (uint8_t)(1u << vbit);
        if (v && b >= 'a' && b <= 'z') mask.back() |= (uint8_t)(1u << vbit);
        len++;
    }
    gxy_seq view() const {
        gxy_seq s;
        s.codes = codes.data();
        s.valid = valid.data();
        s.mask = mask.data();
        s.len = len;
        return s;
    }
};

// PackedSeq::slice_to_new (src/sequences.rs:277-298)
OSeq slice_to_new(const gxy_seq &s, uint64_t start, uint64_t end) {
    OSeq out;
    out.reserve(end - start);
    for (uint64_t i = start; i < end; i++) {
        if (!valid_at(s, i)) {
            out.push_ascii('N');
        } else {
            uint8_t b = decode_base(code_at(s, i));
            if (masked_at(s, i)) b = (uint8_t)(b - 'A' + 'a');
            out.push_ascii(b);
        }
    }
    return out;
}

// PackedSeq::reverse_complement (src/sequences.rs:303-317)
OSeq reverse_complement(const gxy_seq &s) {
    OSeq out;
    out.reserve(s.len);
    for (uint64_t k = 0; k < s.len; k++) {
        uint64_t i = s.len - 1 - k;
        bool was_masked = masked_at(s, i);
        uint8_t b;
        if (valid_at(s, i)) {
            uint8_t ch = decode_base((~code_at(s, i)) & 3u);
            b = was_masked ? (uint8_t)(ch - 'A' + 'a') : ch;
        } else {
            b = 'N';
        }
        out.push_ascii(b);
    }
    return out;
}

// PackedSeq::to_ascii (src/sequences.rs:268-272)
std::string to_ascii(const gxy_seq &s) {
    std::string out(s.len, 'N');
    for (uint64_t i = 0; i < s.len; i++)
        if (valid_at(s, i)) out[i] = (char)decode_base(code_at(s, i));
    return out;
}

OSeq clear_masks_copy(const gxy_seq &s) {  // PackedSeq::clear_masks (src/sequences.rs:229-233)
    OSeq o;
    uint64_t nc = (s.len + 3) / 4, nb = (s.len + 7) / 8;
    o.codes.assign(s.codes, s.codes + nc);
    o.valid.assign(s.valid, s.valid + nb);
    o.mask.assign(nb, 0);
    o.len = s.len;
    return o;
}

// SeedPattern / SeedExtractor (src/seeds.rs:20-161)
struct Pattern {
    uint32_t len = 0;
    std::vector<uint32_t> care;  // ascending care offsets
    uint32_t weight() const { return (uint32_t)care.size(); }
};

bool make_pattern(const gxy_seed_pattern *p, Pattern *out) {
    if (!p || p->length == 0 || p->length > 32) return false;  // src/seeds.rs:27-32
    out->len = p->length;
    out->care.clear();
    for (uint32_t k = 0; k < p->length; k++)
        if ((p->care >> k) & 1u) out->care.push_back(k);
    return true;
}

// SeedExtractor::word_at (src/seeds.rs:133-147)
inline bool word_at(const gxy_seq &s, const Pattern &p, uint64_t pos, uint64_t *w) {
    uint64_t end = pos + p.len;
    if (end > s.len) return false;
    uint64_t word = 0;
    for (uint32_t off : p.care) {
        uint64_t q = pos + off;
        if (!valid_at(s, q) || masked_at(s, q)) return false;
        word = (word << 2) | code_at(s, q);
    }
    *w = word;
    return true;
}

}  // namespace

struct gxo_postable {
    Pattern pattern;
    uint64_t num_words = 0;
    std::vector<uint32_t> starts;
    std::vector<uint32_t> positions;
};

namespace {

// ---- ungapped x-drop ----------------------------------------------------------

struct SideResult {
    int32_t best_score;
    uint32_t best_extent;
};

// column_score (src/hsp.rs:198-211)
inline int32_t column_score(const gxy_seq &t, const gxy_seq &q, const gxy_scoring &m, int64_t ti,
                            int64_t qi) {
    if (!valid_at(t, (uint64_t)ti) || !valid_at(q, (uint64_t)qi)) return m.ambig;
    return m.sub[code_at(t, (uint64_t)ti)][code_at(q, (uint64_t)qi)];
}

// extend_one_side (src/hsp.rs:160-195).  *cols counts loop iterations that
// evaluated a column (the terminating one included).
SideResult extend_one_side(const gxy_seq &t, const gxy_seq &q, const gxy_scoring &m,
                           const gxy_hsp_params &p, int64_t start_t, int64_t start_q,
                           int64_t dir, uint64_t *cols) {
    const int64_t t_len = (int64_t)t.len, q_len = (int64_t)q.len;
    int32_t running = 0, best = 0;
    uint32_t best_extent = 0, extent = 0;
    int64_t ti = start_t, qi = start_q;
    for (;;) {
        if (ti < 0 || qi < 0 || ti >= t_len || qi >= q_len) break;
        extent++;
        (*cols)++;
        running = sat_add(running, column_score(t, q, m, ti, qi));
        if (running > best) {
            best = running;
            best_extent = extent;
        } else if ((int64_t)best - (int64_t)running > (int64_t)p.x_drop) {
            break;
        }
        ti += dir;
        qi += dir;
    }
    return {best, best_extent};
}

// pack_i8_lut (src/scoring.rs:102-112) and is_i8_safe (src/scoring.rs:93-96)
bool is_i8_safe(const gxy_scoring &m) {
    auto in = [](int32_t v) { return v >= -128 && v <= 127; };
    for (int a = 0; a < 4; a++)
        for (int b = 0; b < 4; b++)
            if (!in(m.sub[a][b])) return false;
    return in(m.ambig);
}

// extend_side (src/hsp_simd.rs:114-243): 16 column scores per PSHUFB, the
// prefix/x-drop scan stays scalar.
__attribute__((target("avx2,sse4.1,ssse3"))) SideResult extend_side_avx2(
    const gxy_seq &t, const gxy_seq &q, __m128i lut, int8_t ambig_i8, int32_t x_drop,
    int64_t start_t, int64_t start_q, int64_t dir, uint64_t *cols) {
    const int64_t t_len = (int64_t)t.len, q_len = (int64_t)q.len;
    uint64_t max_extent;
    if (dir > 0)
        max_extent = (start_t >= t_len || start_q >= q_len)
                         ? 0
                         : (uint64_t)std::min(t_len - start_t, q_len - start_q);
    else
        max_extent = (start_t < 0 || start_q < 0) ? 0 : (uint64_t)std::min(start_t + 1, start_q + 1);
    if (max_extent == 0) return {0, 0};
    const __m128i ambig_vec = _mm_set1_epi8(ambig_i8);
    const __m128i low4 = _mm_set1_epi8(0x0F);
    int32_t running = 0, best = 0;
    uint32_t best_extent = 0, extent = 0;
    uint64_t offset = 0;
    while (offset < max_extent) {
        uint64_t chunk = std::min<uint64_t>(max_extent - offset, 16);
        alignas(16) uint8_t tc[16] = {0}, qc[16] = {0}, inv[16];
        for (uint64_t i = 0; i < chunk; i++) {
            int64_t ti = start_t + dir * (int64_t)(offset + i);
            int64_t qi = start_q + dir * (int64_t)(offset + i);
            tc[i] = (uint8_t)code_at(t, (uint64_t)ti);
            qc[i] = (uint8_t)code_at(q, (uint64_t)qi);
            inv[i] = (valid_at(t, (uint64_t)ti) && valid_at(q, (uint64_t)qi)) ? 0 : 0xFF;
        }
        for (uint64_t i = chunk; i < 16; i++) inv[i] = 0xFF;
        __m128i tv = _mm_load_si128((const __m128i *)tc);
        __m128i qv = _mm_load_si128((const __m128i *)qc);
        __m128i idx = _mm_and_si128(_mm_or_si128(_mm_slli_epi16(tv, 2), qv), low4);
        __m128i sc = _mm_shuffle_epi8(lut, idx);
        __m128i bl = _mm_blendv_epi8(sc, ambig_vec, _mm_load_si128((const __m128i *)inv));
        alignas(16) int8_t d[16];
        _mm_store_si128((__m128i *)d, bl);
        for (uint64_t i = 0; i < chunk; i++) {
            extent++;
            (*cols)++;
            running = sat_add(running, (int32_t)d[i]);
            if (running > best) {
                best = running;
                best_extent = extent;
            } else if (sat_sub(best, running) > x_drop) {
                return {best, best_extent};
            }
        }
        offset += chunk;
    }
    return {best, best_extent};
}

bool cpu_has_avx2() {
    static const bool has = __builtin_cpu_supports("avx2");
    return has;
}

// extend_hit_scalar (src/hsp.rs:98-148) / extend_hit_avx2 (src/hsp_simd.rs:27-95)
bool extend_hit(const gxy_seq &t, const gxy_seq &q, const gxy_scoring &m, const gxy_hsp_params &p,
                uint32_t t_pos32, uint32_t q_pos32, uint32_t seed_len32, bool simd, gxy_hsp *out,
                uint64_t *cols) {
    const int64_t t_pos = t_pos32, q_pos = q_pos32, seed_len = seed_len32;
    if (t_pos + seed_len > (int64_t)t.len || q_pos + seed_len > (int64_t)q.len) return false;
    int32_t seed_score = 0;
    for (int64_t k = 0; k < seed_len; k++)
        seed_score = sat_add(seed_score, column_score(t, q, m, t_pos + k, q_pos + k));
    *cols += (uint64_t)seed_len;
    SideResult right, left;
    if (simd && is_i8_safe(m) && cpu_has_avx2()) {
        alignas(16) int8_t lut[16];
        for (int a = 0; a < 4; a++)
            for (int b = 0; b < 4; b++) lut[(a << 2) | b] = (int8_t)m.sub[a][b];
        __m128i l = _mm_load_si128((const __m128i *)lut);
        right = extend_side_avx2(t, q, l, (int8_t)m.ambig, p.x_drop, t_pos + seed_len,
                                 q_pos + seed_len, 1, cols);
        left = extend_side_avx2(t, q, l, (int8_t)m.ambig, p.x_drop, t_pos - 1, q_pos - 1, -1, cols);
    } else {
        right = extend_one_side(t, q, m, p, t_pos + seed_len, q_pos + seed_len, 1, cols);
        left = extend_one_side(t, q, m, p, t_pos - 1, q_pos - 1, -1, cols);
    }
    int32_t score = sat_add(sat_add(seed_score, left.best_score), right.best_score);
    if (score < p.hsp_threshold) return false;
    out->t_start = (uint32_t)(t_pos - (int64_t)left.best_extent);
    out->q_start = (uint32_t)(q_pos - (int64_t)left.best_extent);
    out->length = (uint32_t)(seed_len + left.best_extent + right.best_extent);
    out->score = score;
    return true;
}

// shannon_entropy (src/entropy.rs:27-49)
double shannon_entropy(const gxy_seq &s, uint64_t start, uint64_t end) {
    uint32_t counts[4] = {0, 0, 0, 0};
    uint32_t total = 0;
    uint64_t e = std::min<uint64_t>(end, s.len);
    for (uint64_t i = start; i < e; i++) {
        if (valid_at(s, i)) {
            counts[code_at(s, i)]++;
            total++;
        }
    }
    if (total == 0) return 0.0;
    double tt = (double)total, h = 0.0;
    for (uint32_t c : counts) {
        if (c == 0) continue;
        double p = (double)c / tt;
        h -= p * std::log2(p);
    }
    return h;
}

// passes_entropy_gate (src/entropy.rs:56-66)
bool passes_entropy_gate(const gxy_seq &s, uint32_t t_start, uint32_t t_end, int32_t score,
                         int32_t threshold) {
    double h = shannon_entropy(s, t_start, t_end);
    double scaled = (double)score * (h / 2.0);
    return scaled >= (double)threshold;
}

inline int64_t diagonal(const gxy_hsp &h) { return (int64_t)h.t_start - (int64_t)h.q_start; }
inline bool hsp_eq(const gxy_hsp &a, const gxy_hsp &b) {
    return a.t_start == b.t_start && a.q_start == b.q_start && a.length == b.length &&
           a.score == b.score;
}

// search (src/seed_search.rs:45-125) with DiagHash (src/diag_hash.rs:19-58).
std::vector<gxy_hsp> search(const gxo_postable &table, const gxy_seq &t, const gxy_seq &q,
                            const gxy_scoring &m, const gxy_search_params &p, bool simd,
                            gxy_search_stats *st) {
    const Pattern &pat = table.pattern;
    const uint32_t seed_len = pat.len;
    const uint32_t weight = pat.weight();
    std::vector<gxy_hsp> out;
    std::unordered_map<int64_t, uint32_t> next;  // DiagHash: d -> next t to consider
    std::vector<uint32_t> shifts(weight);
    for (uint32_t i = 0; i < weight; i++) shifts[i] = 2 * (weight - 1 - i);
    std::vector<uint64_t> variants;
    const uint64_t step = p.step;
    uint64_t max_start = q.len >= pat.len ? q.len - pat.len : 0;  // src/seeds.rs:157
    for (uint64_t qp = 0; qp <= max_start; qp += step) {
        uint64_t word;
        if (!word_at(q, pat, qp, &word)) continue;
        st->query_windows++;
        variants.clear();
        variants.push_back(word);
        if (p.transitions >= 1)
            for (uint32_t s : shifts) variants.push_back(word ^ (2ull << s));
        if (p.transitions >= 2)
            for (uint32_t i = 0; i < weight; i++)
                for (uint32_t j = i + 1; j < weight; j++)
                    variants.push_back(word ^ (2ull << shifts[i]) ^ (2ull << shifts[j]));
        const uint32_t q_pos = (uint32_t)qp;
        for (uint64_t v : variants) {
            st->lookups++;
            uint32_t b = table.starts[v], e = table.starts[v + 1];
            st->hits += e - b;
            for (uint32_t k = b; k < e; k++) {
                uint32_t t_pos = table.positions[k];
                int64_t d = (int64_t)t_pos - (int64_t)q_pos;
                auto it = next.find(d);  // DiagHash::admit (src/diag_hash.rs:34-37)
                if (it != next.end() && !(t_pos >= it->second)) continue;
                st->extended++;
                gxy_hsp h;
                if (!extend_hit(t, q, m, p.hsp, t_pos, q_pos, seed_len, simd, &h, &st->columns))
                    continue;
                st->passing++;
                if (p.entropy_enabled &&
                    !passes_entropy_gate(t, h.t_start, h.t_start + h.length, h.score,
                                         p.entropy_threshold))
                    continue;
                // DiagHash::mark_covered (src/diag_hash.rs:43-53)
                uint32_t t_end = h.t_start + h.length, q_end = h.q_start + h.length;
                int64_t dd = (int64_t)t_end - (int64_t)q_end;
                auto jt = next.find(dd);
                if (jt == next.end())
                    next.emplace(dd, t_end);
                else if (t_end > jt->second)
                    jt->second = t_end;
                out.push_back(h);
            }
        }
    }
    st->emitted += out.size();
    std::stable_sort(out.begin(), out.end(), [](const gxy_hsp &a, const gxy_hsp &b) {
        int64_t da = diagonal(a), db = diagonal(b);
        if (da != db) return da < db;
        return a.t_start < b.t_start;
    });
    for (size_t i = 1; i < out.size(); i++)
        if (diagonal(out[i]) == diagonal(out[i - 1]) && out[i].t_start == out[i - 1].t_start &&
            !hsp_eq(out[i], out[i - 1]))
            st->tie_pairs++;
    // Vec::dedup: drop consecutive equal elements.
    std::vector<gxy_hsp> dd;
    dd.reserve(out.size());
    for (const gxy_hsp &h : out)
        if (dd.empty() || !hsp_eq(dd.back(), h)) dd.push_back(h);
    return dd;
}

// choose_anchor (src/anchor.rs:25-69)
gxy_anchor choose_anchor(const gxy_hsp &hsp, const gxy_seq &t, const gxy_seq &q,
                         const gxy_scoring &m, uint32_t window) {
    const uint32_t length = hsp.length;
    uint32_t w32 = std::min(std::max(window, 1u), length);
    size_t w = w32, len = length;
    std::vector<int32_t> cols(len);
    for (size_t k = 0; k < len; k++)
        cols[k] = column_score(t, q, m, (int64_t)hsp.t_start + (int64_t)k,
                               (int64_t)hsp.q_start + (int64_t)k);
    size_t best_start = 0;
    int32_t best = 0;
    for (size_t k = 0; k < w; k++) best += cols[k];
    int32_t win = best;
    for (size_t start = 1; start + w <= len; start++) {
        win = win - cols[start - 1] + cols[start + w - 1];
        if (win > best) {
            best = win;
            best_start = start;
        }
    }
    uint32_t mid = (uint32_t)(best_start + w / 2);
    gxy_anchor a;
    a.t_pos = hsp.t_start + mid;
    a.q_pos = hsp.q_start + mid;
    a.source = hsp;
    return a;
}

// ---- gapped extension ------------------------------------------------------------

enum : uint8_t { kM = 0, kX = 1, kY = 2 };

struct Script {  // EditScript (src/edit_script.rs:46-129)
    std::vector<gxy_edit_run> runs;
    void push(uint32_t op, uint32_t n) {  // src/edit_script.rs:70-81
        if (n == 0) return;
        if (!runs.empty() && runs.back().op == op) {
            runs.back().n += n;
            return;
        }
        runs.push_back({op, n});
    }
    void extend(const Script &o) {  // src/edit_script.rs:84-88
        for (const auto &r : o.runs) push(r.op, r.n);
    }
    Script reversed() const {  // src/edit_script.rs:92-96
        Script s;
        s.runs.assign(runs.rbegin(), runs.rend());
        return s;
    }
};

struct OneSide {
    int32_t score;
    uint32_t t_span, q_span;
    Script script;
};

struct Base {
    uint8_t code;
    bool valid;
};

inline int32_t pair_score(Base a, Base b, const gxy_scoring &m) {  // src/gapped_extend.rs:161-167
    return (a.valid && b.valid) ? m.sub[a.code & 3][b.code & 3] : m.ambig;
}

// extend_one_side (src/gapped_extend.rs:179-463).  `lazy` grows the band
// storage row by row instead of allocating (n+1) x band_width up front.
OneSide gapped_one_side(const std::vector<Base> &T, const std::vector<Base> &Q,
                        const gxy_scoring &mat, const gxy_gapped_params &p, bool lazy,
                        uint64_t *cells) {
    const size_t n = T.size(), m = Q.size();
    if (n == 0 || m == 0) return {0, 0, 0, Script()};
    const int32_t oe = mat.gap_open + mat.gap_extend;
    const int32_t e = mat.gap_extend;
    const int32_t y_drop = p.y_drop;
    size_t br;
    if (e > 0) {
        int32_t r = sat_sub(y_drop, oe) / e;
        br = (size_t)(std::max(r, 0) + 2);
    } else {
        br = m + 1;
    }
    const size_t bw = std::min(4 * br + 4, m + 1);
    std::vector<size_t> row_lo(lazy ? 0 : n + 1, 0);
    std::vector<int32_t> mm, xx, yy;
    std::vector<uint8_t> mf, xf, yf;
    auto grow = [&](size_t rows) {
        size_t size = rows * bw;
        if (mm.size() >= size) return;
        mm.resize(size, kNegInf);
        xx.resize(size, kNegInf);
        yy.resize(size, kNegInf);
        mf.resize(size, kM);
        xf.resize(size, kM);
        yf.resize(size, kM);
    };
    grow(lazy ? 1 : n + 1);
    mm[0] = 0;
    int32_t best_score = 0;
    size_t bi = 0, bj = 0;
    uint8_t bstate = kM;
    size_t lo = 0, hi = 0;
    if (lazy) row_lo.push_back(0);
    for (size_t i = 0; i <= n; i++) {
        const size_t j_lo = lo >= br ? lo - br : 0;
        const size_t j_hi = std::min(std::min(hi + br, m), j_lo + bw - 1);
        if (lazy) {
            grow(i + 1);
            if (row_lo.size() < i + 1) row_lo.resize(i + 1, 0);
        }
        row_lo[i] = j_lo;
        const size_t lo_this = j_lo;
        const size_t lo_prev = i >= 1 ? row_lo[i - 1] : 0;
        const size_t hi_prev_excl = lo_prev + bw;
        const size_t base_this = i * bw;
        const size_t base_prev = i >= 1 ? (i - 1) * bw : 0;
        auto prev_in_band = [&](size_t j) { return j >= lo_prev && j < hi_prev_excl; };
        for (size_t j = j_lo; j <= j_hi; j++) {
            if (i == 0 && j == 0) continue;
            (*cells)++;
            int32_t best = kNegInf;
            const size_t here = base_this + (j - lo_this);
            if (i >= 1) {  // X: gap in query
                int32_t pm = kNegInf, px = kNegInf;
                if (prev_in_band(j)) {
                    size_t off = base_prev + (j - lo_prev);
                    pm = mm[off];
                    px = xx[off];
                }
                int32_t fm = pm - oe, fx = px - e;
                int32_t v;
                uint8_t from;
                if (fm >= fx) { v = fm; from = kM; } else { v = fx; from = kX; }
                xx[here] = v;
                xf[here] = from;
                if (v > best) best = v;
            }
            if (j >= 1) {  // Y: gap in target
                int32_t pm = kNegInf, py = kNegInf;
                if (j - 1 >= lo_this) {
                    size_t off = base_this + (j - 1 - lo_this);
                    pm = mm[off];
                    py = yy[off];
                }
                int32_t fm = pm - oe, fy = py - e;
                int32_t v;
                uint8_t from;
                if (fm >= fy) { v = fm; from = kM; } else { v = fy; from = kY; }
                yy[here] = v;
                yf[here] = from;
                if (v > best) best = v;
            }
            if (i >= 1 && j >= 1) {  // M
                int32_t s = pair_score(T[i - 1], Q[j - 1], mat);
                int32_t fm = kNegInf, fx = kNegInf, fy = kNegInf;
                if (prev_in_band(j - 1)) {
                    size_t off = base_prev + (j - 1 - lo_prev);
                    fm = mm[off];
                    fx = xx[off];
                    fy = yy[off];
                }
                int32_t prev;
                uint8_t from;
                if (fm >= fx && fm >= fy) { prev = fm; from = kM; }
                else if (fx >= fy) { prev = fx; from = kX; }
                else { prev = fy; from = kY; }
                int32_t v = prev + s;
                mm[here] = v;
                mf[here] = from;
                if (v > best) best = v;
            }
            if (best > best_score) {
                best_score = best;
                uint8_t st = mm[here] == best ? kM : (xx[here] == best ? kX : kY);
                bi = i;
                bj = j;
                bstate = st;
            }
        }
        // y-drop prune (src/gapped_extend.rs:388-412)
        const int32_t threshold = sat_sub(best_score, y_drop);
        bool any = false;
        size_t new_lo = 0, new_hi = 0;
        for (size_t j = j_lo; j <= j_hi; j++) {
            size_t here = i * bw + (j - row_lo[i]);
            int32_t alive = std::max(std::max(mm[here], xx[here]), yy[here]);
            if (alive >= threshold) {
                if (!any) {
                    new_lo = j;
                    any = true;
                }
                new_hi = j;
            } else {
                mm[here] = kNegInf;
                xx[here] = kNegInf;
                yy[here] = kNegInf;
            }
        }
        if (!any) break;
        lo = new_lo;
        hi = new_hi;
        if (lo == hi && i > 0) {
            size_t here = i * bw + (lo - row_lo[i]);
            int32_t alive = std::max(std::max(mm[here], xx[here]), yy[here]);
            if (alive < sat_sub(best_score, y_drop)) break;
        }
    }
    // traceback (src/gapped_extend.rs:426-453)
    size_t i = bi, j = bj;
    uint8_t state = bstate;
    Script script;
    auto at = [&](size_t ii, size_t jj) { return ii * bw + (jj - row_lo[ii]); };
    while (i > 0 || j > 0) {
        if (state == kM) {
            script.push(GXY_OP_MATCH, 1);
            uint8_t prev = mf[at(i, j)];
            i -= 1;
            j -= 1;
            state = prev;
        } else if (state == kX) {
            script.push(GXY_OP_DELETE, 1);
            uint8_t prev = xf[at(i, j)];
            i -= 1;
            state = prev;
        } else {
            script.push(GXY_OP_INSERT, 1);
            uint8_t prev = yf[at(i, j)];
            j -= 1;
            state = prev;
        }
    }
    return {best_score, (uint32_t)bi, (uint32_t)bj, script.reversed()};
}

// gapped_extend::extend (src/gapped_extend.rs:78-134)
bool gapped_extend(const gxy_anchor &a, const gxy_seq &t, const gxy_seq &q, const gxy_scoring &m,
                   const gxy_gapped_params &p, bool lazy, gxy_alignment *out, Script *script,
                   uint64_t *cells) {
    const uint64_t at = a.t_pos, aq = a.q_pos;
    const int32_t anchor_col = column_score(t, q, m, (int64_t)at, (int64_t)aq);
    auto slice = [](const gxy_seq &s, uint64_t b, uint64_t e) {  // src/gapped_extend.rs:139-143
        std::vector<Base> v;
        v.reserve(e - b);
        for (uint64_t i = b; i < e; i++) v.push_back({(uint8_t)code_at(s, i), valid_at(s, i)});
        return v;
    };
    OneSide right = gapped_one_side(slice(t, at + 1, t.len), slice(q, aq + 1, q.len), m, p, lazy,
                                    cells);
    std::vector<Base> tl = slice(t, 0, at), ql = slice(q, 0, aq);
    std::reverse(tl.begin(), tl.end());
    std::reverse(ql.begin(), ql.end());
    OneSide left = gapped_one_side(tl, ql, m, p, lazy, cells);
    int32_t total = sat_add(sat_add(anchor_col, left.score), right.score);
    if (total < p.gapped_threshold) return false;
    Script s;
    s.extend(left.script.reversed());
    s.push(GXY_OP_MATCH, 1);
    s.extend(right.script);
    out->t_start = a.t_pos - left.t_span;
    out->q_start = a.q_pos - left.q_span;
    out->t_len = left.t_span + 1 + right.t_span;
    out->q_len = left.q_span + 1 + right.q_span;
    out->score = total;
    out->n_runs = (uint32_t)s.runs.size();
    out->runs_offset = 0;
    out->source = a;
    *script = std::move(s);
    return true;
}

// ---- driver -------------------------------------------------------------------

struct Rec {
    uint32_t ti, qi, strand;
    uint32_t t_start, q_start, t_span, q_span;
    int32_t score;
    Script script;
};

// TargetChunk / chunk_target (src/driver.rs:137-194)
struct Chunk {
    uint32_t parent;
    uint32_t global_start;
    OSeq seq;
    uint32_t primary_start_local, primary_end_local;
};

std::vector<Chunk> chunk_target(uint32_t parent, const gxy_seq &t, uint32_t primary_bp,
                                uint32_t halo_bp) {
    std::vector<Chunk> chunks;
    const uint32_t t_len = (uint32_t)t.len;
    if (primary_bp == 0 || (uint64_t)t_len <= (uint64_t)primary_bp + halo_bp) {
        Chunk c;
        c.parent = parent;
        c.global_start = 0;
        c.seq = slice_to_new(t, 0, t.len);  // target.seq.clone()
        c.primary_start_local = 0;
        c.primary_end_local = t_len;
        chunks.push_back(std::move(c));
        return chunks;
    }
    uint32_t ps = 0;
    while (ps < t_len) {
        uint32_t pe = (uint32_t)std::min<uint64_t>((uint64_t)ps + primary_bp, t_len);
        uint32_t lo = ps >= halo_bp ? ps - halo_bp : 0;
        uint32_t hi = (uint32_t)std::min<uint64_t>((uint64_t)pe + halo_bp, t_len);
        Chunk c;
        c.parent = parent;
        c.global_start = lo;
        c.seq = slice_to_new(t, lo, hi);
        c.primary_start_local = ps - lo;
        c.primary_end_local = pe - lo;
        chunks.push_back(std::move(c));
        if (pe == t_len) break;
        ps = pe;
    }
    return chunks;
}

gxo_postable *build_table(const gxy_seq &seq, const Pattern &pat, uint64_t step) {
    // PosTable::build (src/pos_table.rs:35-83)
    auto *t = new gxo_postable();
    t->pattern = pat;
    t->num_words = 1ull << (2 * pat.weight());
    std::vector<uint32_t> counts(t->num_words, 0);
    std::vector<std::pair<uint32_t, uint32_t>> hits;
    uint64_t max_start = seq.len >= pat.len ? seq.len - pat.len : 0;
    for (uint64_t pos = 0; pos <= max_start; pos += step) {
        uint64_t w;
        if (!word_at(seq, pat, pos, &w)) continue;
        counts[w]++;
        hits.emplace_back((uint32_t)pos, (uint32_t)w);
    }
    t->starts.assign(t->num_words + 1, 0);
    uint64_t acc = 0;
    for (uint64_t i = 0; i < t->num_words; i++) {
        t->starts[i] = (uint32_t)acc;
        acc += counts[i];
    }
    t->starts[t->num_words] = (uint32_t)acc;
    for (uint64_t i = 0; i < t->num_words; i++) counts[i] = t->starts[i];
    t->positions.assign(acc, 0);
    for (auto &h : hits) t->positions[counts[h.second]++] = h.first;
    return t;
}

void prune_table(gxo_postable *t, uint32_t limit) {  // src/pos_table.rs:113-136
    std::vector<uint32_t> np, ns(t->num_words + 1, 0);
    uint32_t acc = 0;
    for (uint64_t w = 0; w < t->num_words; w++) {
        ns[w] = acc;
        uint32_t b = t->starts[w], e = t->starts[w + 1];
        if (e - b <= limit) {
            np.insert(np.end(), t->positions.begin() + b, t->positions.begin() + e);
            acc += e - b;
        }
    }
    ns[t->num_words] = acc;
    t->positions.swap(np);
    t->starts.swap(ns);
}

constexpr uint32_t kMaxWeight = 14;  // 4^14 + 1 starts = 1 GiB; the reference's
                                     // dense table is infeasible beyond this.

// best_chain (src/chain.rs:19-69): the highest-scoring chain of HSPs strictly
// increasing on both axes, by a weighted LIS over (t_start, q_start) order; a
// predecessor must score strictly more than the best so far (the first one
// found wins ties), the terminal is the first maximum.  Output ascending.
std::vector<gxy_hsp> best_chain(const std::vector<gxy_hsp> &hsps) {
    if (hsps.empty()) return {};
    std::vector<gxy_hsp> sorted = hsps;
    std::stable_sort(sorted.begin(), sorted.end(), [](const gxy_hsp &a, const gxy_hsp &b) {
        return a.t_start != b.t_start ? a.t_start < b.t_start : a.q_start < b.q_start;
    });
    const size_t n = sorted.size();
    std::vector<int64_t> best(n);
    std::vector<int64_t> parent(n, -1);
    for (size_t i = 0; i < n; i++) best[i] = sorted[i].score;
    for (size_t i = 1; i < n; i++) {
        const gxy_hsp &hi = sorted[i];
        int64_t bps = 0, bpi = -1;
        for (size_t j = 0; j < i; j++) {
            const gxy_hsp &hj = sorted[j];
            if ((uint64_t)hj.t_start + hj.length <= hi.t_start && (uint64_t)hj.q_start + hj.length <= hi.q_start &&
                best[j] > bps) {
                bps = best[j];
                bpi = (int64_t)j;
            }
        }
        best[i] = bps + hi.score;
        parent[i] = bpi;
    }
    size_t end = 0;
    for (size_t i = 1; i < n; i++)
        if (best[i] > best[end]) end = i;
    std::vector<gxy_hsp> out;
    for (int64_t c = (int64_t)end; c >= 0; c = parent[(size_t)c]) out.push_back(sorted[(size_t)c]);
    std::reverse(out.begin(), out.end());
    return out;
}

// tweener::interpolate (src/tweener.rs:59-185) over one (target, query,
// strand) group's records: each gap between neighbours in (t_start, q_start)
// order whose target and query spans both lie in [min_gap, max_gap] is
// searched again with the inner pattern (step 1) and relaxed thresholds; the
// HSPs found (chained when chaining is on) are all anchored and gapped-
// extended inside the gap, added, and the records re-sorted and deduplicated.
// `target` is the whole target, `query` the group's query strand.
void interpolate(std::vector<Rec> &recs, const gxy_seq &target, const gxy_seq &query, uint32_t ti, uint32_t qi,
                 uint32_t strand, const gxy_config *cfg, bool simd, bool lazy, uint64_t *dp, uint64_t *cells) {
    if (recs.size() < 2) return;
    auto by_tq = [](const Rec &a, const Rec &b) {
        return a.t_start != b.t_start ? a.t_start < b.t_start : a.q_start < b.q_start;
    };
    std::stable_sort(recs.begin(), recs.end(), by_tq);
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
    Pattern pat;
    if (!make_pattern(&cfg->inner_pattern, &pat)) return;
    std::vector<Rec> extras;
    for (size_t k = 0; k + 1 < recs.size(); k++) {
        const Rec &a = recs[k], &b = recs[k + 1];
        const uint32_t tg0 = a.t_start + a.t_span, tg1 = b.t_start;
        const uint32_t qg0 = a.q_start + a.q_span, qg1 = b.q_start;
        if (tg1 <= tg0 || qg1 <= qg0) continue;
        const uint32_t tgap = tg1 - tg0, qgap = qg1 - qg0;
        if (tgap < cfg->inner_min_gap || qgap < cfg->inner_min_gap) continue;
        if (tgap > cfg->inner_max_gap || qgap > cfg->inner_max_gap) continue;
        const OSeq ts = slice_to_new(target, tg0, tg1), qs = slice_to_new(query, qg0, qg1);
        const gxy_seq tv = ts.view(), qv = qs.view();
        gxo_postable *table = build_table(tv, pat, 1);
        if (cfg->max_word_count > 0) prune_table(table, cfg->max_word_count);
        gxy_search_stats st;
        std::memset(&st, 0, sizeof st);
        std::vector<gxy_hsp> hsps = search(*table, tv, qv, cfg->scoring, sp, simd, &st);
        delete table;
        if (hsps.empty()) continue;
        if (cfg->chain_enabled) hsps = best_chain(hsps);
        for (const auto &h : hsps) {
            const gxy_anchor an = choose_anchor(h, tv, qv, cfg->scoring, cfg->anchor_window);
            gxy_alignment aln;
            Script sc;
            (*dp)++;
            if (!gapped_extend(an, tv, qv, cfg->scoring, gp, lazy, &aln, &sc, cells)) continue;
            Rec r;
            r.ti = ti;
            r.qi = qi;
            r.strand = strand;
            r.t_start = tg0 + aln.t_start;
            r.q_start = qg0 + aln.q_start;
            r.t_span = aln.t_len;
            r.q_span = aln.q_len;
            r.score = aln.score;
            r.script = std::move(sc);
            extras.push_back(std::move(r));
        }
    }
    for (auto &r : extras) recs.push_back(std::move(r));
    std::stable_sort(recs.begin(), recs.end(), by_tq);
    // dedup_by on (t_start, q_start, t_span, q_span): the first of a run stays
    std::vector<Rec> out;
    for (auto &r : recs) {
        if (!out.empty()) {
            const Rec &p = out.back();
            if (p.t_start == r.t_start && p.q_start == r.q_start && p.t_span == r.t_span && p.q_span == r.q_span)
                continue;
        }
        out.push_back(std::move(r));
    }
    recs.swap(out);
}

gxy_record_set *record_set_of(const std::vector<Rec> &all);

}  // namespace

// ================================ C exports =====================================

extern "C" {

int gxo_pack_ascii(const char *ascii, uint64_t len, uint8_t *codes, uint8_t *valid,
                   uint8_t *mask) {
    if ((!ascii && len) || !codes || !valid || !mask) return GXY_EINVAL;
    OSeq s;
    s.reserve(len);
    for (uint64_t i = 0; i < len; i++) s.push_ascii((uint8_t)ascii[i]);
    std::memcpy(codes, s.codes.data(), s.codes.size());
    std::memcpy(valid, s.valid.data(), s.valid.size());
    std::memcpy(mask, s.mask.data(), s.mask.size());
    return GXY_OK;
}

int gxo_reverse_complement(const gxy_seq *in, uint8_t *codes, uint8_t *valid, uint8_t *mask) {
    if (!in) return GXY_EINVAL;
    OSeq r = reverse_complement(*in);
    std::memcpy(codes, r.codes.data(), r.codes.size());
    std::memcpy(valid, r.valid.data(), r.valid.size());
    std::memcpy(mask, r.mask.data(), r.mask.size());
    return GXY_OK;
}

int gxo_to_ascii(const gxy_seq *in, char *out) {
    if (!in || !out) return GXY_EINVAL;
    std::string s = to_ascii(*in);
    std::memcpy(out, s.data(), s.size());
    return GXY_OK;
}

int gxo_seed_word(const gxy_seq *seq, const gxy_seed_pattern *pattern, uint64_t pos,
                  uint64_t *word) {
    Pattern p;
    if (!seq || !make_pattern(pattern, &p) || !word) return GXY_EINVAL;
    return word_at(*seq, p, pos, word) ? 1 : 0;
}

int gxo_postable_build(const gxy_seq *seq, const gxy_seed_pattern *pattern, uint64_t step,
                       gxo_postable **out) {
    Pattern p;
    if (!seq || !make_pattern(pattern, &p) || !out || step < 1) return GXY_EINVAL;
    if (p.weight() > kMaxWeight) return GXY_ERANGE;
    if (seq->len > 0xFFFFFFFFull) return GXY_ERANGE;  // src/pos_table.rs:56
    *out = build_table(*seq, p, step);
    return GXY_OK;
}

int gxo_postable_prune(gxo_postable *t, uint32_t limit) {
    if (!t) return GXY_EINVAL;
    prune_table(t, limit);
    return GXY_OK;
}
uint64_t gxo_postable_num_words(const gxo_postable *t) { return t ? t->num_words : 0; }
uint64_t gxo_postable_total_positions(const gxo_postable *t) {
    return t ? t->positions.size() : 0;
}
const uint32_t *gxo_postable_starts(const gxo_postable *t) { return t ? t->starts.data() : nullptr; }
const uint32_t *gxo_postable_positions(const gxo_postable *t) {
    return t ? t->positions.data() : nullptr;
}
void gxo_postable_free(gxo_postable *t) { delete t; }

int gxo_extend_hit(const gxy_seq *t, const gxy_seq *q, const gxy_scoring *sc,
                   const gxy_hsp_params *p, uint32_t t_pos, uint32_t q_pos, uint32_t seed_len,
                   int simd, gxy_hsp *out, int *found) {
    if (!t || !q || !sc || !p || !out || !found) return GXY_EINVAL;
    uint64_t cols = 0;
    *found = extend_hit(*t, *q, *sc, *p, t_pos, q_pos, seed_len, simd != 0, out, &cols) ? 1 : 0;
    return GXY_OK;
}

int gxo_search(const gxo_postable *table, const gxy_seq *t, const gxy_seq *q,
               const gxy_scoring *sc, const gxy_search_params *p, int simd, gxy_hsp **out,
               uint64_t *n, gxy_search_stats *stats) {
    if (!table || !t || !q || !sc || !p || !out || !n || p->step < 1) return GXY_EINVAL;
    gxy_search_stats st;
    std::memset(&st, 0, sizeof st);
    std::vector<gxy_hsp> v = search(*table, *t, *q, *sc, *p, simd != 0, &st);
    *n = v.size();
    *out = (gxy_hsp *)std::malloc(std::max<size_t>(1, v.size()) * sizeof(gxy_hsp));
    if (!*out) return GXY_ENOMEM;
    if (!v.empty()) std::memcpy(*out, v.data(), v.size() * sizeof(gxy_hsp));
    if (stats) *stats = st;
    return GXY_OK;
}

int gxo_choose_anchor(const gxy_hsp *hsp, const gxy_seq *t, const gxy_seq *q,
                      const gxy_scoring *sc, uint32_t window, gxy_anchor *out) {
    if (!hsp || !t || !q || !sc || !out || hsp->length < 1) return GXY_EINVAL;
    *out = choose_anchor(*hsp, *t, *q, *sc, window);
    return GXY_OK;
}

int gxo_gapped_extend(const gxy_anchor *a, const gxy_seq *t, const gxy_seq *q,
                      const gxy_scoring *sc, const gxy_gapped_params *p, int lazy,
                      gxy_alignment *out, gxy_edit_run **runs, uint64_t *cells, int *found) {
    if (!a || !t || !q || !sc || !p || !out || !runs || !found) return GXY_EINVAL;
    if (a->t_pos >= t->len || a->q_pos >= q->len) return GXY_EINVAL;
    Script s;
    uint64_t c = 0;
    std::memset(out, 0, sizeof *out);
    *runs = nullptr;
    bool ok = gapped_extend(*a, *t, *q, *sc, *p, lazy != 0, out, &s, &c);
    if (cells) *cells = c;
    *found = ok ? 1 : 0;
    if (ok) {
        *runs = (gxy_edit_run *)std::malloc(std::max<size_t>(1, s.runs.size()) *
                                            sizeof(gxy_edit_run));
        if (!*runs) return GXY_ENOMEM;
        if (!s.runs.empty()) std::memcpy(*runs, s.runs.data(), s.runs.size() * sizeof(gxy_edit_run));
    }
    return GXY_OK;
}

double gxo_shannon_entropy(const gxy_seq *seq, uint64_t start, uint64_t end) {
    return seq ? shannon_entropy(*seq, start, end) : 0.0;
}
int gxo_passes_entropy_gate(const gxy_seq *seq, uint32_t t_start, uint32_t t_end, int32_t score,
                            int32_t threshold) {
    return seq && passes_entropy_gate(*seq, t_start, t_end, score, threshold) ? 1 : 0;
}

int gxo_best_chain(const gxy_hsp *hsps, uint64_t n, gxy_hsp *out, uint64_t *n_out) {
    if ((!hsps && n) || !n_out || (!out && n)) return GXY_EINVAL;
    const std::vector<gxy_hsp> c = best_chain(std::vector<gxy_hsp>(hsps, hsps + n));
    std::copy(c.begin(), c.end(), out);
    *n_out = c.size();
    return GXY_OK;
}

int gxo_interpolate(const gxy_record_set *in, const gxy_seq *target, const gxy_seq *query, const gxy_config *cfg,
                    gxy_record_set **out) {
    if (!in || !target || !query || !cfg || !out) return GXY_EINVAL;
    std::vector<Rec> recs;
    uint32_t ti = 0, qi = 0, strand = 0;
    for (uint64_t k = 0; k < in->n; k++) {
        const gxy_record &g = in->records[k];
        Rec r;
        r.ti = ti = g.target_index;
        r.qi = qi = g.query_index;
        r.strand = strand = g.strand;
        r.t_start = g.t_start;
        r.q_start = g.q_start;
        r.t_span = g.t_span;
        r.q_span = g.q_span;
        r.score = g.score;
        for (uint32_t x = 0; x < g.n_runs; x++) r.script.push(in->runs[g.runs_offset + x].op, in->runs[g.runs_offset + x].n);
        recs.push_back(std::move(r));
    }
    uint64_t dp = 0, cells = 0;
    interpolate(recs, *target, *query, ti, qi, strand, cfg, true, true, &dp, &cells);
    *out = record_set_of(recs);
    return *out ? GXY_OK : GXY_ENOMEM;
}

int gxo_run(const gxy_named_seq *targets, uint64_t n_targets, const gxy_named_seq *queries,
            uint64_t n_queries, const gxy_config *cfg, int threads, int lazy, int simd,
            gxy_record_set **out, gxy_run_stats *stats) {
    if ((!targets && n_targets) || (!queries && n_queries) || !cfg || !out) return GXY_EINVAL;
    Pattern pat, ipat;
    if (cfg->tweener_enabled && (!make_pattern(&cfg->inner_pattern, &ipat) || ipat.weight() > kMaxWeight))
        return GXY_EINVAL;
    if (!make_pattern(&cfg->pattern, &pat) || cfg->step < 1) return GXY_EINVAL;
    if (pat.weight() > kMaxWeight) return GXY_ERANGE;
    gxy_run_stats rs;
    std::memset(&rs, 0, sizeof rs);

    // --nomasking clears the masks at load (src/main.rs:26-30).
    std::vector<OSeq> tcopy, qcopy;
    std::vector<gxy_seq> tv(n_targets), qv(n_queries);
    for (uint64_t i = 0; i < n_targets; i++) tv[i] = targets[i].seq;
    for (uint64_t i = 0; i < n_queries; i++) qv[i] = queries[i].seq;
    if (!cfg->respect_masking) {
        for (uint64_t i = 0; i < n_targets; i++) tcopy.push_back(clear_masks_copy(tv[i]));
        for (uint64_t i = 0; i < n_queries; i++) qcopy.push_back(clear_masks_copy(qv[i]));
        for (uint64_t i = 0; i < n_targets; i++) tv[i] = tcopy[i].view();
        for (uint64_t i = 0; i < n_queries; i++) qv[i] = qcopy[i].view();
    }

    std::vector<Chunk> chunks;
    for (uint64_t ti = 0; ti < n_targets; ti++) {
        auto c = chunk_target((uint32_t)ti, tv[ti], cfg->chunk_primary_bp, cfg->chunk_halo_bp);
        for (auto &x : c) chunks.push_back(std::move(x));
    }
    rs.chunks = chunks.size();
    std::vector<OSeq> rc(n_queries);
    for (uint64_t qi = 0; qi < n_queries; qi++) rc[qi] = reverse_complement(qv[qi]);

    std::vector<uint32_t> strands;
    if (cfg->strand == GXY_STRAND_PLUS || cfg->strand == GXY_STRAND_BOTH) strands.push_back(0);
    if (cfg->strand == GXY_STRAND_MINUS || cfg->strand == GXY_STRAND_BOTH) strands.push_back(1);

    unsigned nthreads = threads > 0 ? (unsigned)threads : std::thread::hardware_concurrency();
    if (nthreads == 0) nthreads = 1;

    // Phase 1: one PosTable per chunk (src/driver.rs:212-215).
    std::vector<gxo_postable *> tables(chunks.size(), nullptr);
    {
        std::atomic<size_t> next{0};
        auto work = [&]() {
            for (size_t c; (c = next.fetch_add(1)) < chunks.size();) {
                tables[c] = build_table(chunks[c].seq.view(), pat, cfg->step);
                if (cfg->max_word_count > 0) prune_table(tables[c], cfg->max_word_count);
            }
        };
        std::vector<std::thread> pool;
        for (unsigned k = 0; k < std::min<size_t>(nthreads, chunks.size()); k++)
            pool.emplace_back(work);
        for (auto &th : pool) th.join();
    }

    // Phase 2: one unit per (chunk, query, strand) (src/driver.rs:219-367).
    struct Unit {
        size_t chunk;
        uint32_t qi, strand;
        std::vector<Rec> recs;
        gxy_search_stats st;
        uint64_t dp = 0, cells = 0, score_ties = 0;
    };
    std::vector<Unit> units;
    for (size_t c = 0; c < chunks.size(); c++)
        for (uint32_t qi = 0; qi < n_queries; qi++)
            for (uint32_t s : strands) {
                Unit u;
                u.chunk = c;
                u.qi = qi;
                u.strand = s;
                std::memset(&u.st, 0, sizeof u.st);
                units.push_back(std::move(u));
            }
    gxy_search_params sp;
    sp.step = cfg->step;
    sp.hsp = cfg->hsp;
    sp.transitions = cfg->transitions;
    sp.entropy_enabled = cfg->entropy_enabled;
    sp.entropy_threshold = cfg->entropy_threshold;
    {
        std::atomic<size_t> next{0};
        auto work = [&]() {
            for (size_t k; (k = next.fetch_add(1)) < units.size();) {
                Unit &u = units[k];
                const Chunk &ch = chunks[u.chunk];
                const gxy_seq cseq = ch.seq.view();
                const gxy_seq qseq = u.strand == 0 ? qv[u.qi] : rc[u.qi].view();
                std::vector<gxy_hsp> hsps =
                    search(*tables[u.chunk], cseq, qseq, cfg->scoring, sp, simd != 0, &u.st);
                if (hsps.empty()) continue;
                std::vector<gxy_hsp> kept;  // primary-region filter (src/driver.rs:250-256)
                for (const auto &h : hsps)
                    if (h.t_start >= ch.primary_start_local && h.t_start < ch.primary_end_local)
                        kept.push_back(h);
                if (cfg->chain_enabled) kept = best_chain(kept);  // src/driver.rs:258-260
                if (cfg->gapped_enabled) {
                    std::stable_sort(kept.begin(), kept.end(),
                                     [](const gxy_hsp &a, const gxy_hsp &b) {
                                         return a.score > b.score;
                                     });
                    for (size_t i = 1; i < kept.size(); i++)
                        if (kept[i].score == kept[i - 1].score) u.score_ties++;
                }
                struct Cov {
                    uint32_t t_lo, t_hi, q_lo, q_hi;
                };
                std::vector<Cov> covered;
                const uint32_t target_len = (uint32_t)tv[ch.parent].len;
                (void)target_len;
                for (const auto &h : kept) {
                    if (cfg->gapped_enabled) {
                        bool cov = false;  // src/driver.rs:288-297 (local vs global coords)
                        for (const auto &c : covered)
                            if (h.t_start >= c.t_lo && h.t_start < c.t_hi && h.q_start >= c.q_lo &&
                                h.q_start < c.q_hi) {
                                cov = true;
                                break;
                            }
                        if (cov) continue;
                        // build_gapped_record (src/driver.rs:395-434)
                        gxy_anchor a = choose_anchor(h, cseq, qseq, cfg->scoring, cfg->anchor_window);
                        gxy_alignment aln;
                        Script s;
                        u.dp++;
                        if (!gapped_extend(a, cseq, qseq, cfg->scoring, cfg->gapped, lazy != 0, &aln,
                                           &s, &u.cells))
                            continue;
                        Rec r;
                        r.ti = ch.parent;
                        r.qi = u.qi;
                        r.strand = u.strand;
                        r.t_start = aln.t_start + ch.global_start;
                        r.q_start = aln.q_start;
                        r.t_span = aln.t_len;
                        r.q_span = aln.q_len;
                        r.score = aln.score;
                        r.script = std::move(s);
                        covered.push_back({r.t_start, r.t_start + r.t_span, r.q_start,
                                           r.q_start + r.q_span});
                        u.recs.push_back(std::move(r));
                    } else {
                        // build_ungapped_record (src/driver.rs:437-468)
                        Rec r;
                        r.ti = ch.parent;
                        r.qi = u.qi;
                        r.strand = u.strand;
                        r.t_start = h.t_start + ch.global_start;
                        r.q_start = h.q_start;
                        r.t_span = h.length;
                        r.q_span = h.length;
                        r.score = h.score;
                        r.script.push(GXY_OP_MATCH, h.length);
                        u.recs.push_back(std::move(r));
                    }
                }
                // tweener over the whole target (src/driver.rs:340-356)
                if (cfg->tweener_enabled)
                    interpolate(u.recs, tv[ch.parent], qseq, ch.parent, u.qi, u.strand, cfg, simd != 0, lazy != 0,
                                &u.dp, &u.cells);
            }
        };
        std::vector<std::thread> pool;
        for (unsigned k = 0; k < std::min<size_t>(nthreads, units.size()); k++)
            pool.emplace_back(work);
        for (auto &th : pool) th.join();
    }
    for (auto *t : tables) delete t;

    // Group merge (src/driver.rs:371-389): groups ordered by (ti, qi, strand),
    // equal keys in ascending chunk order; each group sorted and deduped alone.
    std::vector<size_t> order;
    for (size_t k = 0; k < units.size(); k++)
        if (!units[k].recs.empty()) order.push_back(k);
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
        const Unit &ua = units[a], &ub = units[b];
        uint32_t ta = chunks[ua.chunk].parent, tb = chunks[ub.chunk].parent;
        if (ta != tb) return ta < tb;
        if (ua.qi != ub.qi) return ua.qi < ub.qi;
        if (ua.strand != ub.strand) return ua.strand < ub.strand;
        return ua.chunk < ub.chunk;
    });
    std::vector<Rec> all;
    for (size_t k : order) {
        std::vector<Rec> &recs = units[k].recs;
        std::stable_sort(recs.begin(), recs.end(), [](const Rec &a, const Rec &b) {
            if (a.t_start != b.t_start) return a.t_start < b.t_start;
            if (a.q_start != b.q_start) return a.q_start < b.q_start;
            if (a.t_span != b.t_span) return a.t_span < b.t_span;
            return a.q_span < b.q_span;
        });
        size_t first = all.size();
        for (auto &r : recs) {
            if (all.size() > first) {
                const Rec &b = all.back();
                if (r.t_start == b.t_start && r.q_start == b.q_start && r.t_span == b.t_span &&
                    r.q_span == b.q_span && r.strand == b.strand)
                    continue;
            }
            all.push_back(std::move(r));
        }
    }
    for (auto &u : units) {
        rs.hits += u.st.hits;
        rs.columns += u.st.columns;
        rs.tie_pairs += u.st.tie_pairs;
        rs.passing += u.st.passing;
        rs.hsps += u.st.emitted;
        rs.dp_problems += u.dp;
        rs.dp_cells += u.cells;
        rs.score_ties += u.score_ties;
    }
    for (uint64_t qi = 0; qi < n_queries; qi++) rs.query_bp += qv[qi].len * strands.size();

    gxy_record_set *set = record_set_of(all);
    if (!set) return GXY_ENOMEM;
    *out = set;
    if (stats) *stats = rs;
    return GXY_OK;
}

}  // extern "C"

namespace {

gxy_record_set *record_set_of(const std::vector<Rec> &all) {
    auto *set = (gxy_record_set *)std::calloc(1, sizeof(gxy_record_set));
    if (!set) return nullptr;
    size_t nruns = 0;
    for (auto &r : all) nruns += r.script.runs.size();
    set->n = all.size();
    set->records = (gxy_record *)std::calloc(std::max<size_t>(1, all.size()), sizeof(gxy_record));
    set->n_runs = nruns;
    set->runs = (gxy_edit_run *)std::calloc(std::max<size_t>(1, nruns), sizeof(gxy_edit_run));
    if (!set->records || !set->runs) {
        std::free(set->records);
        std::free(set->runs);
        std::free(set);
        return nullptr;
    }
    size_t ro = 0;
    for (size_t k = 0; k < all.size(); k++) {
        gxy_record &o = set->records[k];
        const Rec &r = all[k];
        o.target_index = r.ti;
        o.query_index = r.qi;
        o.strand = r.strand;
        o.t_start = r.t_start;
        o.q_start = r.q_start;
        o.t_span = r.t_span;
        o.q_span = r.q_span;
        o.score = r.score;
        o.n_runs = (uint32_t)r.script.runs.size();
        o.runs_offset = ro;
        for (auto &x : r.script.runs) set->runs[ro++] = x;
    }
    return set;
}

// Shared by the MAF and PAF writers: ASCII of each target and of each query
// strand, produced on demand.
struct AsciiCache {
    const gxy_named_seq *targets, *queries;
    std::vector<std::string> t, qp, qm;
    std::vector<bool> have_t, have_qp, have_qm;
    AsciiCache(const gxy_named_seq *ts, uint64_t nt, const gxy_named_seq *qs, uint64_t nq)
        : targets(ts), queries(qs), t(nt), qp(nq), qm(nq), have_t(nt), have_qp(nq), have_qm(nq) {}
    const std::string &target(uint32_t i) {
        if (!have_t[i]) { t[i] = to_ascii(targets[i].seq); have_t[i] = true; }
        return t[i];
    }
    const std::string &query(uint32_t i, uint32_t strand) {
        if (strand == 0) {
            if (!have_qp[i]) { qp[i] = to_ascii(queries[i].seq); have_qp[i] = true; }
            return qp[i];
        }
        if (!have_qm[i]) {
            OSeq r = reverse_complement(queries[i].seq);
            qm[i] = to_ascii(r.view());
            have_qm[i] = true;
        }
        return qm[i];
    }
};

bool check_records(uint64_t nt, uint64_t nq, const gxy_record_set *rs) {
    for (uint64_t k = 0; k < rs->n; k++) {
        const gxy_record &r = rs->records[k];
        if (r.target_index >= nt || r.query_index >= nq) return false;
        if (r.runs_offset + r.n_runs > rs->n_runs) return false;
    }
    return true;
}

char *dup_text(const std::string &s, uint64_t *len) {
    char *p = (char *)std::malloc(s.size() + 1);
    if (!p) return nullptr;
    std::memcpy(p, s.data(), s.size());
    p[s.size()] = 0;
    *len = s.size();
    return p;
}

}  // namespace

extern "C" {

int gxo_format_maf(const gxy_named_seq *targets, uint64_t n_targets,
                   const gxy_named_seq *queries, uint64_t n_queries,
                   const gxy_record_set *records, char **text, uint64_t *len) {
    if (!records || !text || !len || !check_records(n_targets, n_queries, records))
        return GXY_EINVAL;
    AsciiCache cache(targets, n_targets, queries, n_queries);
    std::string out;
    if (records->n > 0) {  // header written with the first record (src/output/maf.rs:33-43)
        out += "##maf version=1 program=lastz-gxy\n";
        out += "# scoring=HOXD70\n";
    }
    char buf[256];
    for (uint64_t k = 0; k < records->n; k++) {
        const gxy_record &r = records->records[k];
        const std::string &ta = cache.target(r.target_index);
        const std::string &qa = cache.query(r.query_index, r.strand);
        // render_aligned_bases (src/edit_script.rs:140-172)
        std::string tt, qq;
        size_t ti = r.t_start, qi = r.q_start;
        for (uint32_t x = 0; x < r.n_runs; x++) {
            const gxy_edit_run &run = records->runs[r.runs_offset + x];
            for (uint32_t c = 0; c < run.n; c++) {
                if (run.op == GXY_OP_MATCH) {
                    tt.push_back(ta[ti++]);
                    qq.push_back(qa[qi++]);
                } else if (run.op == GXY_OP_INSERT) {
                    tt.push_back('-');
                    qq.push_back(qa[qi++]);
                } else {
                    tt.push_back(ta[ti++]);
                    qq.push_back('-');
                }
            }
        }
        std::snprintf(buf, sizeof buf, "a score=%d\n", r.score);
        out += buf;
        std::snprintf(buf, sizeof buf, "s %-20s %10u %6u + %10u ", targets[r.target_index].name,
                      r.t_start, r.t_span, (uint32_t)targets[r.target_index].seq.len);
        out += buf;
        out += tt;
        out += '\n';
        std::snprintf(buf, sizeof buf, "s %-20s %10u %6u %c %10u ", queries[r.query_index].name,
                      r.q_start, r.q_span, r.strand == 0 ? '+' : '-',
                      (uint32_t)queries[r.query_index].seq.len);
        out += buf;
        out += qq;
        out += "\n\n";
    }
    *text = dup_text(out, len);
    return *text ? GXY_OK : GXY_ENOMEM;
}

int gxo_format_paf(const gxy_named_seq *targets, uint64_t n_targets,
                   const gxy_named_seq *queries, uint64_t n_queries,
                   const gxy_record_set *records, char **text, uint64_t *len) {
    if (!records || !text || !len || !check_records(n_targets, n_queries, records))
        return GXY_EINVAL;
    AsciiCache cache(targets, n_targets, queries, n_queries);
    std::string out;
    char buf[512];
    for (uint64_t k = 0; k < records->n; k++) {
        const gxy_record &r = records->records[k];
        const std::string &ta = cache.target(r.target_index);
        const std::string &qa = cache.query(r.query_index, r.strand);
        // score_script (src/output/paf.rs:59-89); bases outside the record's
        // slices compare as None.
        uint32_t matches = 0, mismatches = 0, indels = 0;
        size_t ti = 0, qi = 0;
        std::string cigar;
        for (uint32_t x = 0; x < r.n_runs; x++) {
            const gxy_edit_run &run = records->runs[r.runs_offset + x];
            std::snprintf(buf, sizeof buf, "%u%c", run.n,
                          run.op == GXY_OP_MATCH ? 'M' : (run.op == GXY_OP_INSERT ? 'I' : 'D'));
            cigar += buf;
            if (run.op == GXY_OP_MATCH) {
                for (uint32_t c = 0; c < run.n; c++) {
                    bool th = ti < r.t_span, qh = qi < r.q_span;
                    bool eq;
                    if (th && qh) eq = ta[r.t_start + ti] == qa[r.q_start + qi];
                    else eq = (th == qh);
                    if (eq) matches++; else mismatches++;
                    ti++;
                    qi++;
                }
            } else if (run.op == GXY_OP_INSERT) {
                indels += run.n;
                qi += run.n;
            } else {
                indels += run.n;
                ti += run.n;
            }
        }
        std::snprintf(buf, sizeof buf, "%s\t%u\t%u\t%u\t%c\t%s\t%u\t%u\t%u\t%u\t%u\t60\tNM:i:%u\tAS:i:%d\ttp:A:P\tcg:Z:",
                      queries[r.query_index].name, (uint32_t)queries[r.query_index].seq.len,
                      r.q_start, r.q_start + r.q_span, r.strand == 0 ? '+' : '-',
                      targets[r.target_index].name, (uint32_t)targets[r.target_index].seq.len,
                      r.t_start, r.t_start + r.t_span, matches, matches + mismatches + indels,
                      mismatches + indels, r.score);
        out += buf;
        out += cigar;
        out += '\n';
    }
    *text = dup_text(out, len);
    return *text ? GXY_OK : GXY_ENOMEM;
}

void gxo_record_set_free(gxy_record_set *s) {
    if (!s) return;
    std::free(s->records);
    std::free(s->runs);
    std::free(s);
}

void gxo_free(void *p) { std::free(p); }

}  // extern "C"
