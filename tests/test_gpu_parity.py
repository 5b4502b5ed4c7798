"""GPU parity: the HIP library (through the C-ABI) against the oracle.

Every comparison is bit-exact: integer scores, coordinates, edit scripts and
MAF/PAF bytes.  Inputs are seeded; sizes keep the oracle to seconds.
"""
import sys

import numpy as np
import pytest

import gxy
import gxy_types as T
import oracle_py as O
import synth

pytestmark = pytest.mark.gpu

HOX = T.scoring_hoxd70()
PS = T.PackedSeq.from_ascii


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    return gxy.default_context()


def rand_seq(rng, n, pn=0.03, plow=0.2):
    code = rng.integers(0, 4, n)
    valid = rng.random(n) >= pn
    mask = valid & (rng.random(n) < plow)
    return T.PackedSeq.from_arrays(code, valid, mask)


def homologous_pair(rng, n, sub=0.1, indel=0.01, plow=0.1):
    t = rand_seq(rng, n, 0.005, plow)
    c, v, m = t.unpack()
    lo = n // 5
    qc, qv, qm = synth.mutate(c[lo:n - lo], v[lo:n - lo], m[lo:n - lo], rng, sub, indel)
    q = T.PackedSeq.from_arrays(qc, qv, qm)
    # append a reverse-complemented copy so both strands carry homology
    rc = t.slice(lo, lo + n // 10).reverse_complement()
    c2, v2, m2 = rc.unpack()
    qc = np.concatenate([qc, c2])
    qv = np.concatenate([qv, v2])
    qm = np.concatenate([qm, m2])
    return t, T.PackedSeq.from_arrays(qc, qv, qm)


# ---- sequences -----------------------------------------------------------------

def test_dseq_roundtrip_rc_slice(ctx):
    rng = np.random.default_rng(11)
    for n in (1, 5, 31, 32, 33, 63, 64, 65, 1000, 100003):
        say("dseq n", n)
        s = rand_seq(rng, n, 0.1, 0.3)
        d = gxy.DeviceSeq.upload(s, ctx)
        back = d.download()
        assert back.unpack()[0].tolist() == s.unpack()[0].tolist()
        assert back.valid.tolist() == s.valid.tolist() and back.mask.tolist() == s.mask.tolist()
        rc = d.reverse_complement().download()
        ref = s.reverse_complement()
        assert rc.to_ascii() == ref.to_ascii()
        assert rc.mask.tolist() == ref.mask.tolist() and rc.valid.tolist() == ref.valid.tolist()
        for _ in range(5):
            a = int(rng.integers(0, n + 1))
            b = int(rng.integers(a, n + 1))
            sl = d.slice(a, b).download()
            r2 = s.slice(a, b)
            assert sl.to_ascii() == r2.to_ascii()
            assert sl.mask.tolist() == r2.mask.tolist()


# ---- pos_table -----------------------------------------------------------------

@pytest.mark.parametrize("spec,step", [("12of19", 1), ("12of19", 3), ("match8", 1),
                                       ("1101011", 2), ("match1", 1)])
def test_postable_matches_oracle(ctx, spec, step):
    say("postable", spec, step)
    rng = np.random.default_rng(sum(spec.encode()) + step)
    s = rand_seq(rng, 60000, 0.02, 0.3)
    pat = T.pattern_from_spec(spec)
    g = gxy.PosTable.build(s, pat, step, ctx)
    o = O.PosTable(s, pat, step)
    gs, gp = g.export()
    assert g.total_positions == o.total_positions
    assert np.array_equal(gs, o.starts())
    assert np.array_equal(gp, o.positions())


def test_postable_known_answers_and_prune(ctx):
    # src/pos_table.rs:143-211
    w = lambda s: O.seed_word(PS(s), T.pattern_from_spec("match3"), 0)
    t = gxy.PosTable.build(PS(b"ACGTACGT"), T.pattern_from_spec("match3"), 1, ctx)
    assert t.lookup(w(b"ACG")) == [0, 4] and t.lookup(w(b"TAC")) == [3]
    t = gxy.PosTable.build(PS(b"ACGT" * 4), T.pattern_from_spec("match3"), 3, ctx)
    assert t.lookup(w(b"ACG")) == [0, 12]
    t = gxy.PosTable.build(PS(b"AAAAAAAAAAACG"), T.pattern_from_spec("match3"), 1, ctx)
    assert len(t.lookup(w(b"AAA"))) >= 9
    t.prune_hot_words(4)
    assert t.lookup(w(b"AAA")) == [] and t.lookup(w(b"ACG")) == [10]
    rng = np.random.default_rng(5)
    s = PS(bytes(rng.choice(list(b"AAAACGT"), 20000)))
    pat = T.pattern_from_spec("match6")
    g = gxy.PosTable.build(s, pat, 1, ctx)
    o = O.PosTable(s, pat, 1)
    g.prune_hot_words(40)
    o.prune_hot_words(40)
    gs, gp = g.export()
    assert np.array_equal(gs, o.starts()) and np.array_equal(gp, o.positions())


# ---- extend_hit -------------------------------------------------------------------

def test_extend_hit_random_pairs(ctx):
    """The reference's proptest model (src/hsp.rs:283-314) on the GPU."""
    rng = np.random.default_rng(1234)
    for case in range(150):
        tb = bytes(rng.choice(list(b"ACGTN"), rng.integers(32, 200)))
        qb = bytes(rng.choice(list(b"ACGTN"), rng.integers(32, 200)))
        seed_len = int(rng.integers(4, 12))
        xdrop = int(rng.integers(50, 2000))
        t, q = PS(tb), PS(qb)
        p = T.HspParams(xdrop, 0)
        # every in-range seed position plus the reference's midpoint
        tp = rng.integers(0, max(1, len(tb) - seed_len + 3), 40)
        qp = rng.integers(0, max(1, len(qb) - seed_len + 3), 40)
        tp[0], qp[0] = max(0, len(tb) - seed_len) // 2, max(0, len(qb) - seed_len) // 2
        g = gxy.extend_hit_batch(t, q, HOX, p, tp, qp, seed_len, ctx)
        for k in range(len(tp)):
            o = O.extend_hit(t, q, HOX, p, int(tp[k]), int(qp[k]), seed_len, False)
            assert g[k] == o, (case, k)


def test_extend_hit_known_answers(ctx):
    s = PS(b"ACGTACGTACGTACGT")  # src/hsp.rs:222-261, 317-324
    assert gxy.extend_hit(s, s, HOX, T.HspParams(100, 100), 2, 2, 4, ctx) == (0, 0, 16, 1528)
    assert gxy.extend_hit(PS(b"TTTACGTTTT"), PS(b"AAAACGAAAA"), HOX, T.HspParams(50, 50), 3, 3, 3,
                          ctx) == (3, 3, 3, 291)
    assert gxy.extend_hit(PS(b"ACGT"), PS(b"TTTT"), HOX, T.HspParams(100, 10000), 0, 0, 4, ctx) is None
    assert gxy.extend_hit(PS(b"ACGT"), PS(b"ACGT"), HOX, T.HspParams(910, 3000), 2, 2, 10, ctx) is None
    s = PS(b"CCCCCCAAAAAACCCCCC")
    assert gxy.extend_hit(s, s, HOX, T.HspParams(500, 100), 8, 8, 2, ctx) == (0, 0, 18, 1746)


def test_extend_hit_long_homology_batch(ctx):
    rng = np.random.default_rng(9)
    t, q = homologous_pair(rng, 40000, 0.08, 0.004)
    n = 20000
    tp = rng.integers(0, len(t), n)
    qp = np.clip(tp - len(t) // 5 + rng.integers(-3, 4, n), 0, len(q) - 1)
    tp[:10] = [0, 1, len(t) - 19, len(t) - 1, len(t), 5, 6, 7, 8, 9]
    p = T.HspParams(910, 0)
    g = gxy.extend_hit_batch(t, q, HOX, p, tp, qp, 19, ctx)
    for k in range(0, n, 7):
        assert g[k] == O.extend_hit(t, q, HOX, p, int(tp[k]), int(qp[k]), 19, True), k


# ---- search ---------------------------------------------------------------------------

SEARCH_CASES = [
    ("12of19", 1, 1, 910, 3000, 0),
    ("12of19", 1, 0, 910, 3000, 0),
    ("12of19", 1, 2, 910, 2000, 0),
    ("12of19", 3, 1, 500, 3000, 0),
    ("match10", 1, 1, 910, 1500, 0),
    ("12of19", 1, 1, 910, 3000, 1),
]


@pytest.mark.parametrize("spec,step,trans,xdrop,thr,entropy", SEARCH_CASES)
def test_search_matches_oracle(ctx, spec, step, trans, xdrop, thr, entropy):
    say("search", spec, step, trans, xdrop, thr, entropy)
    rng = np.random.default_rng(step * 100 + trans * 10 + xdrop + entropy)
    t, q = homologous_pair(rng, 150000, 0.12, 0.01, 0.2)
    pat = T.pattern_from_spec(spec)
    sp = T.SearchParams(step, T.HspParams(xdrop, thr), trans, entropy, thr)
    g_tab = gxy.PosTable.build(t, pat, step, ctx)
    o_tab = O.PosTable(t, pat, step)
    gh, gst = gxy.search(g_tab, t, q, HOX, sp)
    oh, ost = O.search(o_tab, t, q, HOX, sp)
    assert gst.query_windows == ost.query_windows
    assert gst.hits == ost.hits
    assert gst.emitted == ost.emitted
    assert len(gh) > 0
    assert gh == oh


@pytest.mark.parametrize("trans,tlen,qlen", [(1, 150000, 0), (2, 150000, 0), (1, 700, 300), (0, 2000, 90)])
def test_search_xdrop_forms_agree(trans, tlen, qlen):
    """The two x-drop forms (a straight-line first step of every hit plus the
    resumed longer walks, and the single persistent kernel) give the same HSPs
    and the same statistics — columns included, since both extend every hit —
    and the oracle's HSPs: binned (T=1) and unbinned (T=0, 2) hit lists, and
    short sequences where most walks meet a sequence end."""
    say("x-drop forms", trans, tlen, qlen)
    rng = np.random.default_rng(4242 + trans + tlen)
    if qlen == 0:
        t, q = homologous_pair(rng, tlen, 0.12, 0.01, 0.2)
    else:
        t = rand_seq(rng, tlen, 0.01, 0.0)
        c, v, m = t.unpack()
        qc, qv, qm = synth.mutate(c[50:50 + qlen], v[50:50 + qlen], m[50:50 + qlen], rng, 0.05, 0.0)
        q = T.PackedSeq.from_arrays(qc, qv, qm)
    pat = T.pattern_from_spec("12of19")
    sp = T.SearchParams(1, T.HspParams(910, 1500), trans, 0, 1500)
    res = {}
    for form in (0, 1):
        c = gxy.Context(0)
        c.set_xdrop_form(form)
        tab = gxy.PosTable.build(t, pat, 1, c)
        res[form] = gxy.search(tab, t, q, HOX, sp)
        first, resumed, fpass = c.xdrop_stats()
        assert (first > 0) == (form == 0 and res[form][1].hits > 0)
        assert resumed <= res[form][1].hits and fpass <= res[form][1].passing
        c.close()
    (h0, s0), (h1, s1) = res[0], res[1]
    assert h0 == h1
    assert (s0.hits, s0.columns, s0.passing, s0.emitted) == (s1.hits, s1.columns, s1.passing, s1.emitted)
    oh, ost = O.search(O.PosTable(t, pat, 1), t, q, HOX, sp)
    assert h0 == oh and s0.emitted == ost.emitted


@pytest.mark.parametrize("trans", [1, 2])
def test_search_in_many_hit_batches(trans):
    """A search whose hits exceed the context's hit batch runs them in several
    batches (gxy_ctx_set_hit_batch): the passing extensions of each batch are
    kept across batches and admitted together, so HSPs and statistics equal one
    batch's and the oracle's; the driver's helper and lane contexts take the
    setting too."""
    say("hit batches, transitions", trans)
    rng = np.random.default_rng(515 + trans)
    t, q = homologous_pair(rng, 150000, 0.12, 0.01, 0.2)
    pat = T.pattern_from_spec("12of19")
    sp = T.SearchParams(1, T.HspParams(910, 2000), trans, 0, 2000)
    res = {}
    for hb in (0, 100):
        c = gxy.Context(0)
        c.set_hit_batch(hb)
        tab = gxy.PosTable.build(t, pat, 1, c)
        res[hb] = gxy.search(tab, t, q, HOX, sp)
        c.close()
    (h0, s0), (h1, s1) = res[0], res[100]
    assert s0.hits > 10 * 100
    assert h0 == h1
    assert (s0.hits, s0.columns, s0.passing, s0.emitted) == (s1.hits, s1.columns, s1.passing, s1.emitted)
    oh, ost = O.search(O.PosTable(t, pat, 1), t, q, HOX, sp)
    assert h0 == oh and s0.emitted == ost.emitted
    # through the driver: two chunks, both strands, HSP-only
    targets, queries = [("t", t)], [("q", q)]
    cfg = T.config_default()
    cfg.gapped_enabled = 0
    cfg.chunk_primary_bp = 80_000
    cfg.chunk_halo_bp = 5_000
    recs = {}
    for hb in (0, 100):
        c = gxy.Context(0)
        c.set_hit_batch(hb)
        recs[hb] = gxy.run(targets, queries, cfg, c)[0].tuples()
        c.close()
    assert recs[0] == recs[100]
    assert recs[0] == O.run(targets, queries, cfg)[0].tuples()


def test_search_dense_table(ctx):
    """A dense table (match8 on 300 kbp unmasked: ~4.6 positions per word, ~9
    non-empty variant buckets per window) overflows the expansion kernel's
    per-wavefront bucket list, so those wavefronts expand lane by lane."""
    say("search dense match8")
    rng = np.random.default_rng(808)
    t = rand_seq(rng, 300000, 0.001, 0.0)
    c, v, m = t.unpack()
    qc, qv, qm = synth.mutate(c[50000:150000], v[50000:150000], m[50000:150000], rng, 0.1, 0.01)
    q = T.PackedSeq.from_arrays(qc, qv, qm)
    pat = T.pattern_from_spec("match8")
    sp = T.SearchParams(1, T.HspParams(910, 2000), 1, 0, 2000)
    gh, gst = gxy.search(gxy.PosTable.build(t, pat, 1, ctx), t, q, HOX, sp)
    oh, ost = O.search(O.PosTable(t, pat, 1), t, q, HOX, sp)
    assert gst.hits == ost.hits and gst.hits > 20 * gst.query_windows
    assert gst.emitted == ost.emitted and len(gh) > 0
    assert gh == oh


def test_search_weight14(ctx):
    """Seed weight 14: the table carries no neighbourhood counts (they stop at
    weight 13), so counting probes the 2-bit count map per variant."""
    say("search weight 14")
    rng = np.random.default_rng(1414)
    t = rand_seq(rng, 200000, 0.001, 0.0)
    c, v, m = t.unpack()
    qc, qv, qm = synth.mutate(c[20000:120000], v[20000:120000], m[20000:120000], rng, 0.05, 0.005)
    q = T.PackedSeq.from_arrays(qc, qv, qm)
    pat = T.pattern_from_spec("11101101100110111011")  # 14 care positions
    sp = T.SearchParams(1, T.HspParams(910, 3000), 1, 0, 3000)
    gh, gst = gxy.search(gxy.PosTable.build(t, pat, 1, ctx), t, q, HOX, sp)
    oh, ost = O.search(O.PosTable(t, pat, 1), t, q, HOX, sp)
    assert gst.query_windows == ost.query_windows and gst.hits == ost.hits
    assert gst.emitted == ost.emitted and len(gh) > 0
    assert gh == oh


def test_search_device_query_and_rc(ctx):
    rng = np.random.default_rng(77)
    t, q = homologous_pair(rng, 100000)
    pat = T.pattern_from_spec("12of19")
    sp = T.SearchParams(1, T.HspParams(910, 3000), 1, 0, 0)
    dt = gxy.DeviceSeq.upload(t, ctx)
    dq = gxy.DeviceSeq.upload(q, ctx)
    tab = gxy.PosTable.build(dt, pat, 1, ctx)
    o_tab = O.PosTable(t, pat, 1)
    assert gxy.search(tab, None, dq, HOX, sp)[0] == O.search(o_tab, t, q, HOX, sp)[0]
    rc = q.reverse_complement()
    assert gxy.search(tab, None, dq.reverse_complement(), HOX, sp)[0] == \
        O.search(o_tab, t, rc, HOX, sp)[0]


def test_search_known_answers(ctx):
    def srch(tb, qb, k, x, thr, trans):
        t, q = PS(tb), PS(qb)
        tab = gxy.PosTable.build(t, T.pattern_from_spec(f"match{k}"), 1, ctx)
        return gxy.search(tab, t, q, HOX, T.SearchParams(1, T.HspParams(x, thr), trans, 0, 0))[0]
    # src/seed_search.rs:137-161
    assert srch(b"AAAAAGATTACACATGGCATGAAAAA", b"GATTACACATGGCATG", 8, 500, 1000, 0) == [(5, 0, 16, 1519)]
    assert srch(b"AAAAACGTAAAAA", b"ACATA", 5, 500, 50, 0) == []  # :164-200
    assert len(srch(b"AAAAACGTAAAAA", b"ACATA", 5, 500, 50, 1)) >= 1
    hs = srch(b"AAAAAGATTACACATGGCATGTCGAAAAAA", b"GATTACACATGGCATGTCGA", 7, 500, 500, 0)
    assert len({h[0] - h[1] for h in hs}) == len(hs) and any(h[:3] == (5, 0, 20) for h in hs)


# ---- anchor + gapped --------------------------------------------------------------------

def test_choose_anchor_matches_oracle(ctx):
    rng = np.random.default_rng(3)
    t, q = homologous_pair(rng, 20000, 0.1, 0.0)
    hs = []
    for _ in range(300):
        L = int(rng.integers(1, 400))
        ts = int(rng.integers(0, len(t) - L))
        qs = int(rng.integers(0, len(q) - L))
        hs.append((ts, qs, L, 0))
    for w in (1, 4, 31, 1000):
        g = gxy.choose_anchor_batch(hs, t, q, HOX, w, ctx)
        o = [O.choose_anchor(h, t, q, HOX, w) for h in hs]
        assert g == o
    assert gxy.choose_anchor((0, 0, 8, 0), PS(b"AACCCCAA"), PS(b"TTCCCCTT"), HOX, 4, ctx) == (4, 4)


def test_gapped_known_answers(ctx):
    s = PS(b"ACGTACGTACGTACGT")  # src/gapped_extend.rs:482-601
    r, _ = gxy.gapped_extend((8, 8), s, s, HOX, T.GappedParams(500, 100), ctx)
    assert r[:5] == (0, 0, 16, 16, 1528) and O.cigar(r[5]) == "16M"
    r, _ = gxy.gapped_extend((3, 3), PS(b"ACGTACGTACGT"), PS(b"ACGTACGGTACGT"), HOX,
                             T.GappedParams(1000, 100), ctx)
    assert (r[2], r[3], r[4]) == (12, 13, 716) and "1I" in O.cigar(r[5])
    r, _ = gxy.gapped_extend((3, 3), PS(b"ACGTACGGTACGT"), PS(b"ACGTACGTACGT"), HOX,
                             T.GappedParams(1000, 100), ctx)
    assert (r[2], r[3], r[4]) == (13, 12, 716) and "1D" in O.cigar(r[5])
    assert gxy.gapped_extend((1, 1), PS(b"ACGT"), PS(b"ACGT"), HOX, T.GappedParams(500, 100000),
                             ctx)[0] is None
    core = b"GATTACACATGGCATGTCGAATGCCTAGCATGTCA"
    t = PS(core + core)
    q = PS(core + b"AGCTAGCTAGCTAGCTAGCTAGCTAGCTAGCTAGCTAGCTAGCTAGCTAG" + core)
    r, _ = gxy.gapped_extend((5, 5), t, q, HOX, T.GappedParams(9400, 500), ctx)
    o, _ = O.gapped_extend((5, 5), t, q, HOX, T.GappedParams(9400, 500))
    assert r == o and r[2] >= 60 and r[3] >= 110


@pytest.mark.parametrize("ydrop", [300, 3000, 9400])
def test_gapped_matches_oracle_random(ctx, ydrop):
    say("gapped ydrop", ydrop)
    rng = np.random.default_rng(ydrop)
    t, q = homologous_pair(rng, 6000, 0.15, 0.02, 0.0)
    lo = len(t) // 5
    anchors = []
    for _ in range(24):
        a = int(rng.integers(0, len(q) - 1))
        anchors.append((min(len(t) - 1, a + lo + int(rng.integers(-2, 3))), a))
    anchors += [(0, 0), (len(t) - 1, len(q) - 1), (0, len(q) - 1), (len(t) - 1, 0)]
    gp = T.GappedParams(ydrop, 0)
    g, gcells = gxy.gapped_extend_batch(anchors, t, q, HOX, gp, ctx)
    ocells = 0
    for k, a in enumerate(anchors):
        o, c = O.gapped_extend(a, t, q, HOX, gp)
        ocells += c
        assert g[k] == o, (k, a)
    assert gcells == ocells


def test_gapped_pairs_dev(ctx):
    rng = np.random.default_rng(21)
    ts, qs, anc, host = [], [], [], []
    for k in range(12):
        L = int(rng.integers(50, 1500))
        t = rand_seq(rng, 2 * L + 1, 0.0, 0.0)
        c, v, m = t.unpack()
        qc, qv, qm = synth.mutate(c, v, m, rng, 0.1, 0.01)
        q = T.PackedSeq.from_arrays(qc, qv, qm)
        a = (L, min(L, len(q) - 1))
        ts.append(gxy.DeviceSeq.upload(t, ctx))
        qs.append(gxy.DeviceSeq.upload(q, ctx))
        anc.append(a)
        host.append((t, q))
    gp = T.GappedParams(9400, 3000)
    g, _ = gxy.gapped_extend_pairs_dev(ts, qs, anc, HOX, gp, ctx)
    for k in range(12):
        o, _ = O.gapped_extend(anc[k], host[k][0], host[k][1], HOX, gp)
        assert g[k] == o, k


def test_gapped_pairs_dev_large_pass(ctx):
    """320 sides in one pass (several blocks per CU, longest first) equal the
    oracle, and the same pairs a few at a time give the same alignments."""
    rng = np.random.default_rng(23)
    ts, qs, anc, host = [], [], [], []
    for k in range(160):
        L = int(rng.integers(30, 700))
        t = rand_seq(rng, 2 * L + 1, 0.01, 0.0)
        c, v, m = t.unpack()
        qc, qv, qm = synth.mutate(c, v, m, rng, float(rng.uniform(0.05, 0.3)), 0.01)
        q = T.PackedSeq.from_arrays(qc, qv, qm)
        a = (L, min(L, len(q) - 1))
        ts.append(gxy.DeviceSeq.upload(t, ctx))
        qs.append(gxy.DeviceSeq.upload(q, ctx))
        anc.append(a)
        host.append((t, q))
    gp = T.GappedParams(9400, 3000)
    say("gapped 160 pairs in one pass")
    g, gcells = gxy.gapped_extend_pairs_dev(ts, qs, anc, HOX, gp, ctx)
    ocells = 0
    for k in range(len(anc)):
        o, c = O.gapped_extend(anc[k], host[k][0], host[k][1], HOX, gp)
        ocells += c
        assert g[k] == o, k
    assert gcells == ocells
    say("gapped the same pairs 8 at a time")
    for s0 in range(0, 40, 8):
        g2, _ = gxy.gapped_extend_pairs_dev(ts[s0:s0 + 8], qs[s0:s0 + 8], anc[s0:s0 + 8], HOX, gp, ctx)
        assert g2 == g[s0:s0 + 8]


# ---- driver ----------------------------------------------------------------------------

def _cfg(**kw):
    c = T.config_default()
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def test_run_fixtures(ctx):
    targets = [("chrT", PS(b"AAAAAAAAAAAGATTACACATGGCATGTCGAATTTTTTTTTTT"))]
    queries = [("chrQ", PS(b"GATTACACATGGCATGTCGA"))]
    cfg = _cfg(pattern=T.pattern_from_spec("match12"), hsp=T.HspParams(910, 500),
               gapped=T.GappedParams(9400, 500))
    recs, _ = gxy.run(targets, queries, cfg, ctx)
    maf = gxy.format_maf(targets, queries, recs).decode()
    assert "a score=1901\ns chrT                         11     20 +         43 " \
           "GATTACACATGGCATGTCGA\n" in maf
    orec, _ = O.run(targets, queries, cfg)
    assert maf.encode() == O.format_maf(targets, queries, orec)
    assert gxy.format_paf(targets, queries, recs) == O.format_paf(targets, queries, orec)


def test_run_small_pair_maf_identical(ctx):
    say("run 30 kbp pair, gapped, MAF")
    t, q = synth.pair_30k(1)
    targets, queries = [("target", t)], [("query", q)]
    cfg = _cfg()
    g, gst = gxy.run(targets, queries, cfg, ctx)
    o, ost = O.run(targets, queries, cfg, threads=0)
    gm = gxy.format_maf(targets, queries, g)
    om = O.format_maf(targets, queries, o)
    assert len(g) > 0
    assert gm == om
    assert gxy.format_paf(targets, queries, g) == O.format_paf(targets, queries, o)
    assert gst.hits == ost.hits


def test_run_chunked_hsp_only_identical(ctx):
    say("run chunked HSP-only")
    t, q = synth.chromosome_pair(n=700_000, seed=3)
    targets, queries = [("chr", t)], [("qry", q)]
    cfg = _cfg(gapped_enabled=0, chunk_primary_bp=200_000, chunk_halo_bp=20_000)
    g, gst = gxy.run(targets, queries, cfg, ctx)
    o, ost = O.run(targets, queries, cfg)
    assert gst.chunks == ost.chunks == 4
    assert len(g) > 100
    assert gxy.format_maf(targets, queries, g) == O.format_maf(targets, queries, o)


def test_run_chunked_gapped_nomasking_identical(ctx):
    say("run chunked gapped nomasking")
    t, q = synth.chromosome_pair(n=16_000, seed=5, masked=0.4, sub=0.1, indel=0.01)
    targets, queries = [("chr", t)], [("qry", q), ("q2", t.slice(3000, 7000))]
    for respect in (1, 0):
        cfg = _cfg(chunk_primary_bp=6_000, chunk_halo_bp=1_000, respect_masking=respect)
        g, _ = gxy.run(targets, queries, cfg, ctx)
        o, _ = O.run(targets, queries, cfg)
        assert gxy.format_maf(targets, queries, g) == O.format_maf(targets, queries, o)


# ---- chain and tweener --------------------------------------------------------------

def test_interpolate_known_answer(ctx):  # src/tweener.rs:254-330
    blk = "GATTACACATGGCATGTCGA"
    t = PS(blk + "N" * 20 + "AGCTAGCTAGCTAGC" + "N" * 20 + blk)
    recs = [(0, 0, 0, 0, 0, 20, 20, 1900, [(0, 20)]), (0, 0, 0, 75, 75, 20, 20, 1900, [(0, 20)])]
    cfg = T.config_default()
    cfg.pattern = T.pattern_from_spec("match19")
    cfg.chain_enabled = 1
    cfg.hsp.hsp_threshold = 5000
    cfg.gapped.gapped_threshold = 5000
    cfg.inner_pattern = T.pattern_from_spec("match12")
    cfg.inner_min_gap, cfg.inner_max_gap, cfg.inner_relax = 10, 10_000, 100
    dt = gxy.DeviceSeq.upload(t, ctx)
    g = gxy.interpolate(recs, dt, dt, cfg, ctx)
    assert g == O.interpolate(recs, t, t, cfg)
    assert len(g) == 3 and g[1][3] == 40 and g[1][8] == "15M"


def test_interpolate_random_matches_oracle(ctx):
    rng = np.random.default_rng(31)
    t, q = homologous_pair(rng, 40000, 0.18, 0.01, 0.0)
    cfg = T.config_default()
    cfg.strand = T.STRAND_PLUS
    cfg.chain_enabled = 1
    cfg.inner_relax = 4
    dt, dq = gxy.DeviceSeq.upload(t, ctx), gxy.DeviceSeq.upload(q, ctx)
    # records: the chained gapped run's, which leaves gaps between alignments
    recs = O.run([("t", t)], [("q", q)], cfg, threads=4)[0].tuples()
    def script(c):
        import re
        return [("MID".index(op), int(n)) for n, op in re.findall(r"(\d+)([MID])", c)]
    grp = [r[:8] + (script(r[8]),) for r in recs]
    say("interpolate", len(grp), "records")
    assert gxy.interpolate(grp, dt, dq, cfg, ctx) == O.interpolate(grp, t, q, cfg)


@pytest.mark.parametrize("gapped,chunk", [(True, 0), (True, 20_000), (False, 0)])
def test_run_chain_tweener_matches_oracle(ctx, gapped, chunk):
    rng = np.random.default_rng(33 + chunk)
    t, q = homologous_pair(rng, 60000, 0.2, 0.015, 0.1)
    cfg = _cfg(gapped_enabled=1 if gapped else 0, chain_enabled=1, tweener_enabled=1, inner_relax=4,
               chunk_primary_bp=chunk, chunk_halo_bp=5000)
    say("run chain+tweener gapped", gapped, "chunk", chunk)
    grec, _ = gxy.run([("t", t)], [("q", q)], cfg, ctx)
    orec, _ = O.run([("t", t)], [("q", q)], cfg, threads=4)
    assert grec.tuples() == orec.tuples()
    assert gxy.format_maf([("t", t)], [("q", q)], grec) == O.format_maf([("t", t)], [("q", q)], orec)
    cfg.tweener_enabled = 0
    g2, _ = gxy.run([("t", t)], [("q", q)], cfg, ctx)
    o2, _ = O.run([("t", t)], [("q", q)], cfg, threads=4)
    assert g2.tuples() == o2.tuples()


def test_run_queries_across_length_bins(ctx):
    """Queries of 0.8, 3, 30 and 150 kbp in one run: the gapped stage splits
    the groups by query length and runs each bin on a context and stream of
    its own; the records must equal the oracle's."""
    rng = np.random.default_rng(44)
    t = rand_seq(rng, 400_000, 0.002, 0.05)
    c, v, m = t.unpack()
    queries = []
    for k, (lo, n) in enumerate([(1000, 3000), (50_000, 30_000), (150_000, 150_000), (10_000, 800)]):
        qc, qv, qm = synth.mutate(c[lo:lo + n], v[lo:lo + n], m[lo:lo + n], rng, 0.12, 0.01)
        queries.append((f"q{k}", T.PackedSeq.from_arrays(qc, qv, qm)))
    targets = [("t", t)]
    for chain_tw in (0, 1):
        say("run across length bins, chain+tweener", chain_tw)
        cfg = _cfg(chain_enabled=chain_tw, tweener_enabled=chain_tw)
        g, _ = gxy.run(targets, queries, cfg, ctx)
        o, _ = O.run(targets, queries, cfg, threads=4)
        assert len(g) > 0
        assert g.tuples() == o.tuples()
        assert gxy.format_maf(targets, queries, g) == O.format_maf(targets, queries, o)


# ---- input formats --------------------------------------------------------------

def _rand_ascii(rng, n):
    """Random bases with N and lowercase, plus runs of N and of soft-masking."""
    a = rng.choice(np.frombuffer(b"ACGTacgtNn", np.uint8), n,
                   p=[0.2, 0.2, 0.2, 0.2, 0.04, 0.04, 0.04, 0.04, 0.03, 0.01]).astype(np.uint8)
    for _ in range(max(1, n // 5000)):
        s0 = int(rng.integers(0, max(1, n)))
        a[s0:s0 + int(rng.integers(1, 400))] = ord("N")
        s1 = int(rng.integers(0, max(1, n)))
        seg = a[s1:s1 + int(rng.integers(1, 900))]
        seg[(seg >= 65) & (seg <= 90)] += 32  # a view: lowercases in place
    return a.tobytes()


@pytest.mark.parametrize("big_endian", [False, True])
def test_load_2bit_matches_parse(ctx, big_endian):
    import sys as _s
    _s.path.insert(0, __import__("os").path.dirname(__file__))
    from twobit import build_2bit
    rng = np.random.default_rng(17 + big_endian)
    seqs = [(f"s{k}", _rand_ascii(rng, int(n))) for k, n in enumerate([1, 3, 4, 5, 31, 32, 33, 64, 1000, 100003])]
    data = build_2bit(seqs, big_endian=big_endian)
    got = gxy.load_2bit(data, ctx)
    ref = O.parse_2bit(data)
    assert [n for n, _ in got] == [n for n, _ in ref]
    for (n, d), (_, a) in zip(got, ref):
        h = d.download()
        e = T.PackedSeq.from_ascii(a)
        assert h.len == e.len and np.array_equal(h.codes, e.codes) and np.array_equal(h.valid, e.valid) \
            and np.array_equal(h.mask, e.mask), n


def test_load_fasta_matches_parse(ctx):
    rng = np.random.default_rng(19)
    parts = [b"AC\n"]  # before the first header: joins the first record
    for k in range(6):
        s = _rand_ascii(rng, int(rng.integers(0, 70000)))
        w = int(rng.choice([60, 61, 80, 1000]))
        lines = b"\n".join(s[i:i + w] for i in range(0, len(s), w))
        nl = b"\r\n" if k % 2 else b"\n"
        parts.append(b">seq%d some description\n" % k + lines.replace(b"\n", nl) + nl + (b"\n" if k == 3 else b""))
    text = b"".join(parts)
    got = gxy.load_fasta(text, ctx)
    ref = O.parse_fasta(text)
    assert [n for n, _ in got] == [n for n, _ in ref]
    for (n, d), (_, a) in zip(got, ref):
        h = d.download()
        e = T.PackedSeq.from_ascii(a)
        assert h.len == e.len and np.array_equal(h.codes, e.codes) and np.array_equal(h.valid, e.valid) \
            and np.array_equal(h.mask, e.mask), n

