"""GPU parity on degenerate and ragged inputs, against the oracle.

Queries that are empty, one base, shorter than / exactly / one past the seed
span, all N, all soft-masked, N runs inside homology, and reverse complements,
in one run (ragged lengths in one batch), with masking respected and not,
gapped and ungapped; a target shorter than the seed span; sequences of one
base through the gapped extension; seeds at and past the sequence ends; the
entropy gate over low-complexity blocks (repeats and skewed composition).
Everything is compared bit-exactly (records as MAF/PAF bytes).
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


def _acgt(rng, n):
    return rng.choice(list(b"ACGT"), n).astype(np.uint8).tobytes()


def _degenerate_queries(base):
    return [
        ("empty", PS(b"")),
        ("one", PS(b"A")),
        ("short18", PS(base[100:118])),
        ("span19", PS(base[200:219])),
        ("span20", PS(base[300:320])),
        ("allN", PS(b"N" * 500)),
        ("lower", PS(base[500:1500].lower())),
        ("nrun", PS(base[1000:2000] + b"N" * 40 + base[2000:2600])),
        ("rc", PS(base[400:1400]).reverse_complement()),
        ("ragged", PS(base[2100:2100 + 333])),
    ]


@pytest.mark.parametrize("respect,gapped", [(1, 1), (1, 0), (0, 1), (0, 0)])
def test_run_degenerate_queries(ctx, respect, gapped):
    say("degenerate queries respect", respect, "gapped", gapped)
    rng = np.random.default_rng(5)
    base = _acgt(rng, 3000)
    targets = [("t", PS(base)), ("tiny", PS(b"ACG"))]
    queries = _degenerate_queries(base)
    cfg = T.config_default()
    cfg.respect_masking = respect
    cfg.gapped_enabled = gapped
    cfg.hsp = T.HspParams(910, 1000)
    cfg.gapped = T.GappedParams(9400, 1000)
    g, gst = gxy.run(targets, queries, cfg, ctx)
    o, ost = O.run(targets, queries, cfg)
    assert gst.hits == ost.hits
    om = O.format_maf(targets, queries, o)
    assert len(om) > 0
    assert gxy.format_maf(targets, queries, g) == om
    assert gxy.format_paf(targets, queries, g) == O.format_paf(targets, queries, o)


def test_run_chunk_larger_than_target(ctx):
    say("chunk larger than target")
    rng = np.random.default_rng(6)
    base = _acgt(rng, 5000)
    targets = [("t", PS(base))]
    queries = [("q", PS(base[1000:4000])), ("e", PS(b""))]
    cfg = T.config_default()
    cfg.chunk_primary_bp = 1_000_000
    cfg.chunk_halo_bp = 50_000
    g, _ = gxy.run(targets, queries, cfg, ctx)
    o, _ = O.run(targets, queries, cfg)
    assert gxy.format_maf(targets, queries, g) == O.format_maf(targets, queries, o)


def test_search_target_shorter_than_span(ctx):
    rng = np.random.default_rng(7)
    t = PS(b"ACGTACGTACGT")
    q = PS(_acgt(rng, 2000))
    pat = T.pattern_from_spec("12of19")
    sp = T.SearchParams(1, T.HspParams(910, 3000), 1, 0, 3000)
    gh, gst = gxy.search(gxy.PosTable.build(t, pat, 1, ctx), t, q, HOX, sp)
    oh, ost = O.search(O.PosTable(t, pat, 1), t, q, HOX, sp)
    assert gh == oh == []
    assert gst.hits == ost.hits == 0


def test_gapped_one_base_sequences(ctx):
    gp = T.GappedParams(9400, 0)
    for tb, qb in ((b"A", b"A"), (b"A", b"ACGT"), (b"ACGT", b"T"), (b"N", b"A"), (b"G", b"C")):
        t, q = PS(tb), PS(qb)
        for a in ((0, 0), (len(tb) - 1, len(qb) - 1)):
            r, _ = gxy.gapped_extend(a, t, q, HOX, gp, ctx)
            o, _ = O.gapped_extend(a, t, q, HOX, gp)
            assert r == o, (tb, qb, a)


def test_extend_hit_at_and_past_the_ends(ctx):
    rng = np.random.default_rng(8)
    tb = _acgt(rng, 97)
    t = PS(tb)
    q = PS(tb[10:90])
    p = T.HspParams(910, 0)
    span = 19
    tp = np.array([0, 1, 97 - span, 97 - span + 1, 96, 97, 10, 70, 71], dtype=np.int64)
    qp = np.array([0, 0, 80 - span, 80 - span, 79, 80, 0, 60, 61], dtype=np.int64)
    g = gxy.extend_hit_batch(t, q, HOX, p, tp, qp, span, ctx)
    for k in range(len(tp)):
        assert g[k] == O.extend_hit(t, q, HOX, p, int(tp[k]), int(qp[k]), span, True), k


def _low_complexity_pair(rng):
    """Blocks of (AT)n, (ACG)n, A-rich (0.7/0.1/0.1/0.1) and uniform bases, 4 %
    mutated, separated by uniform spacers; the query is a 6 % / 0.5 % mutant."""
    parts = []
    for k in range(24):
        n = int(rng.integers(300, 900))
        kind = k % 4
        if kind == 0:
            b = np.frombuffer(b"AT" * n, np.uint8)[:n].copy()
        elif kind == 1:
            b = np.frombuffer(b"ACG" * n, np.uint8)[:n].copy()
        elif kind == 2:
            b = rng.choice(list(b"ACGT"), n, p=[0.7, 0.1, 0.1, 0.1]).astype(np.uint8)
        else:
            b = rng.choice(list(b"ACGT"), n).astype(np.uint8)
        flip = rng.random(n) < 0.04
        b[flip] = rng.choice(list(b"ACGT"), int(flip.sum())).astype(np.uint8)
        parts.append(b)
        parts.append(rng.choice(list(b"ACGT"), 500).astype(np.uint8))
    t = PS(np.concatenate(parts).tobytes())
    c, v, m = t.unpack()
    qc, qv, qm = synth.mutate(c, v, m, rng, 0.06, 0.005)
    return t, T.PackedSeq.from_arrays(qc, qv, qm)


@pytest.mark.parametrize("thr", [3000, 6000])
def test_search_entropy_gate_on_low_complexity(ctx, thr):
    """HSPs whose target intervals span the whole entropy range (0 to 2 bits):
    the gate rejects thousands of them and the device's choices equal the
    oracle's, HSP for HSP."""
    say("entropy gate low complexity thr", thr)
    rng = np.random.default_rng(31)
    t, q = _low_complexity_pair(rng)
    pat = T.pattern_from_spec("12of19")
    g_tab = gxy.PosTable.build(t, pat, 1, ctx)
    o_tab = O.PosTable(t, pat, 1)
    off, _ = O.search(o_tab, t, q, HOX, T.SearchParams(1, T.HspParams(910, 3000), 1, 0, thr))
    sp = T.SearchParams(1, T.HspParams(910, 3000), 1, 1, thr)
    gh, gst = gxy.search(g_tab, t, q, HOX, sp)
    oh, ost = O.search(o_tab, t, q, HOX, sp)
    assert 0 < len(oh) < len(off) - 1000
    assert gst.hits == ost.hits
    assert gst.emitted == ost.emitted
    assert gh == oh


def test_load_2bit_rejects_corrupt_counts(ctx):
    """A record whose N-block or mask-block count exceeds the bytes left, or a
    file cut short, is an error (src/bit2.rs:115-164 returns Bit2Error): the
    counts are checked before anything is sized from them."""
    import os
    import struct
    sys.path.insert(0, os.path.dirname(__file__))
    from twobit import build_2bit
    good = build_2bit([("a", b"ACGTNNNNACGTacgtacgtACGT"), ("b", b"GGCCAATT")])
    assert len(gxy.load_2bit(good, ctx)) == 2
    # record "a" starts at the offset in its index entry: dna_size, then n_count
    off = struct.unpack("<I", good[16 + 2:16 + 6])[0]
    for at, val in ((off + 4, 0xFFFFFFFF), (off + 4, 0x40000000)):
        bad = bytearray(good)
        bad[at:at + 4] = struct.pack("<I", val)
        with pytest.raises(gxy.GxyError, match=r"\(-1\)"):
            gxy.load_2bit(bytes(bad), ctx)
    # mask-block count (after the one N block: start and size)
    bad = bytearray(good)
    mat = off + 8 + 8
    assert struct.unpack("<I", good[mat:mat + 4])[0] == 1  # one lowercase run
    bad[mat:mat + 4] = struct.pack("<I", 0xFFFFFFF0)
    with pytest.raises(gxy.GxyError, match=r"\(-1\)"):
        gxy.load_2bit(bytes(bad), ctx)
    for cut in (3, 17, off + 6, len(good) - 1):
        with pytest.raises(gxy.GxyError):
            gxy.load_2bit(good[:cut], ctx)
