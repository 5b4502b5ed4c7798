"""Host-side logic of the library that needs no GPU, against the oracle."""
import numpy as np

import gxy
import oracle_py as O


def test_best_chain_matches_oracle():
    """gxy_best_chain (O(n log n)) equals the oracle's restatement of the
    reference's O(n^2) scan, ties included, from 1 to 1500 HSPs."""
    rng = np.random.default_rng(7)
    for trial in range(120):
        n = int(rng.choice([1, 2, 5, 30, 63, 64, 65, 200, 1500]))
        span = int(rng.choice([50, 500, 100000]))
        hs = [(int(rng.integers(0, span)), int(rng.integers(0, span)), int(rng.integers(1, 60)),
               int(rng.choice([100, 200, 300]) if trial % 3 == 0 else rng.integers(1, 5000)))
              for _ in range(n)]
        assert gxy.best_chain(hs) == O.best_chain(hs), (trial, n)


def test_best_chain_zero_length_and_nonpositive():
    rng = np.random.default_rng(8)
    for trial in range(60):
        n = int(rng.integers(1, 120))
        hs = [(int(rng.integers(0, 80)), int(rng.integers(0, 80)), int(rng.integers(0, 5)),
               int(rng.integers(-50, 200))) for _ in range(n)]
        assert gxy.best_chain(hs) == O.best_chain(hs), trial


def _named_pairs(rng):
    import gxy_types as T
    import synth
    base = rng.choice(list(b"ACGT"), 9000).astype(np.uint8)
    lower = rng.random(9000) < 0.1
    base[lower] = base[lower] + 32
    t1 = T.PackedSeq.from_ascii(base.tobytes())
    c, v, m = t1.unpack()
    qc, qv, qm = synth.mutate(c[1000:7000], v[1000:7000], m[1000:7000], rng, 0.12, 0.01)
    q1 = T.PackedSeq.from_arrays(qc, qv, qm)
    q2 = t1.slice(2000, 5000).reverse_complement()
    noise = rng.choice(list(b"ACGTN"), 3000, p=[0.24, 0.24, 0.24, 0.24, 0.04]).astype(np.uint8)
    t2 = T.PackedSeq.from_ascii(noise[:1500].tobytes() + base[6500:8500].tobytes() + noise[1500:].tobytes())
    targets = [("chr1", t1), ("a_much_longer_target_name.v2", t2)]
    queries = [("q_fwd", q1), ("e", T.PackedSeq.from_ascii(b"")), ("query_rc_excerpt", q2)]
    return targets, queries


def test_writers_match_oracle_on_oracle_records():
    """gxy_format_maf / gxy_format_paf (host code of the library) render the
    oracle's records byte-for-byte as the oracle's writers do: the fixture
    block of src/parity.rs:405-411, and gapped (edit scripts, both strands,
    chain + tweener) and ungapped runs over several named sequences."""
    import gxy_types as T
    PS = T.PackedSeq.from_ascii
    ft = [("chrT", PS(b"AAAAAAAAAAAGATTACACATGGCATGTCGAATTTTTTTTTTT"))]
    fq = [("chrQ", PS(b"GATTACACATGGCATGTCGA"))]
    cfg = T.config_default()
    cfg.pattern = T.pattern_from_spec("match12")
    cfg.hsp = T.HspParams(910, 500)
    cfg.gapped = T.GappedParams(9400, 500)
    recs, _ = O.run(ft, fq, cfg)
    assert gxy.format_maf(ft, fq, recs) == (
        b"##maf version=1 program=lastz-gxy\n# scoring=HOXD70\n"
        b"a score=1901\n"
        b"s chrT                         11     20 +         43 GATTACACATGGCATGTCGA\n"
        b"s chrQ                          0     20 +         20 GATTACACATGGCATGTCGA\n\n")
    assert gxy.format_paf(ft, fq, recs) == O.format_paf(ft, fq, recs)

    rng = np.random.default_rng(12)
    targets, queries = _named_pairs(rng)
    for gapped, chain in ((1, 0), (1, 1), (0, 0), (0, 1)):
        cfg = T.config_default()
        cfg.gapped_enabled = gapped
        cfg.chain_enabled = chain
        cfg.tweener_enabled = chain and gapped
        cfg.hsp = T.HspParams(910, 1500)
        cfg.gapped = T.GappedParams(9400, 2000)
        recs, _ = O.run(targets, queries, cfg)
        assert len(recs.tuples()) > 0, (gapped, chain)
        assert gxy.format_maf(targets, queries, recs) == O.format_maf(targets, queries, recs)
        assert gxy.format_paf(targets, queries, recs) == O.format_paf(targets, queries, recs)


def test_writers_reject_records_whose_script_does_not_fit():
    """A record whose edit script consumes more (or fewer) bases than its spans,
    or names an unknown op, is refused with GXY_EINVAL instead of reading past
    the sequences (EditScript::target_span / query_span, src/edit_script.rs:104-121)."""
    import pytest
    from outcheck import ManualRecords, padded
    targets = [("t", padded(b"ACGTACGT", 20))]
    queries = [("q", padded(b"ACGTACGT", 20))]
    ok = ManualRecords([(0, 0, 0, 0, 0, 8, 8, 10, [(0, 8)])])
    assert gxy.format_maf(targets, queries, ok).count(b"a score=10") == 1
    cases = [([(0, 9)], 8, 8), ([(0, 4), (1, 30)], 4, 4), ([(0, 7)], 8, 8), ([(3, 8)], 8, 8),
             ([(0, 4), (2, 4)], 4, 8)]
    for script, tsp, qsp in cases:
        bad = ManualRecords([(0, 0, 0, 0, 0, tsp, qsp, 10, script)])
        for fmt in (gxy.format_maf, gxy.format_paf):
            with pytest.raises(gxy.GxyError):
                fmt(targets, queries, bad)
