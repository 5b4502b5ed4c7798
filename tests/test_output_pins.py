"""Output path pinned to the reference's own tests, fixtures and published counts.

CPU only.  Covers what byte-identical MAF/PAF rests on beyond the hot path's
kernels: the MAF writer's gap rows and strand column (src/output/maf.rs:85-188),
the PAF writer's NM/cg/strand columns (src/output/paf.rs:92-195), the minus-strand
driver path with chaining, the tweener and gapped extension
(tests/fixture_smoke.rs:53-154), the block counts the reference publishes for
its parity corpus (README.md:76-83, parity/scripts/matrix.sh: default flags)
and the 23 HSPs of the SARS pair (PLAN.md:116-121).

Both writers are checked: the oracle's (the parity checker) and the library's
host writer (gxy_format_maf / gxy_format_paf).  Every MAF/PAF text produced here
is also checked against what the formats fix on their own (check_maf /
check_paf): rows without '-' equal the source slices (the query's reverse
complement on the minus strand), sizes equal the ungapped row lengths, and the
block score equals the rows rescored with HOXD70 and affine gaps.
"""
import numpy as np
import pytest

import gxy
import gxy_types as T
import oracle_py as O
from outcheck import (CORPUS_PAIRS, HOX, IDX, PS, ManualRecords, check_maf, check_paf,
                      check_smoke_minus, corpus, fixture, padded, rescore, smoke_minus_case)

WRITERS = {"oracle": (O.format_maf, O.format_paf), "library": (gxy.format_maf, gxy.format_paf)}


def test_rescore_closed_forms():
    # what check_maf's rescoring is pinned by: the HOXD70 diagonal sum of the
    # fixture block (1901, src/parity.rs:405-411) and the DP's gap costs, O + E
    # to open and E per further base (src/gapped_extend.rs:197-198, 307, 332).
    assert rescore(b"GATTACACATGGCATGTCGA", b"GATTACACATGGCATGTCGA") == 1901
    assert rescore(b"AC-GT", b"ACAGT") == 91 + 100 + 100 + 91 - 430
    assert rescore(b"ACNGT", b"ACAGT") == 91 + 100 - 100 + 100 + 91
    assert rescore(b"A--T", b"ACGT") == 91 - 460 + 91
    assert rescore(b"A-T-", b"ACTG") == 91 - 430 + 91 - 430


# ---- src/output/maf.rs:85-188 -------------------------------------------------------

@pytest.mark.parametrize("writer", sorted(WRITERS))
def test_maf_writes_header_once(writer):  # maf.rs:92-121
    fmt_maf = WRITERS[writer][0]
    targets = [("chrT", padded(b"TTTTTTTTTTACGT", 100))]
    queries = [("chrQ", padded(b"TTTTTACGT", 50))]
    rec = (0, 0, 0, 10, 5, 4, 4, 400, [(0, 4)])
    text = fmt_maf(targets, queries, ManualRecords([rec, rec])).decode()
    assert text.count("##maf") == 1
    assert text.count("a score=400") == 2
    assert "s chrT                         10      4 +        100 ACGT\n" in text
    assert "s chrQ                          5      4 +         50 ACGT\n" in text


@pytest.mark.parametrize("writer", sorted(WRITERS))
def test_maf_renders_gap_characters_in_both_rows(writer):  # maf.rs:123-157
    fmt_maf = WRITERS[writer][0]
    targets = [("t", padded(b"AAACCCA", 50))]
    queries = [("q", padded(b"AAAGCCCC", 50))]
    script = [(0, 3), (1, 1), (0, 2), (2, 1), (0, 1)]  # 3M 1I 2M 1D 1M
    text = fmt_maf(targets, queries, ManualRecords([(0, 0, 0, 0, 0, 7, 7, 100, script)])).decode()
    assert "AAA-CC" in text  # the reference's own assertions
    assert "-C" in text
    lines = [x for x in text.split("\n") if x.startswith("s ")]
    assert lines == ["s t                             0      7 +         50 AAA-CCCA",
                     "s q                             0      7 +         50 AAAGCC-C"]


@pytest.mark.parametrize("writer", sorted(WRITERS))
def test_maf_minus_strand_marked_on_query_line(writer):  # maf.rs:159-187
    fmt_maf = WRITERS[writer][0]
    targets = [("t", padded(b"ACGT", 10))]
    # the reference's record holds query bases ACGT on the minus strand: the
    # library takes them from the query's reverse complement, so the forward
    # query ends in RC(ACGT) = ACGT
    queries = [("q", PS(b"TTTTTTACGT"))]
    text = fmt_maf(targets, queries, ManualRecords([(0, 0, 1, 0, 0, 4, 4, 100, [(0, 4)])])).decode()
    s_lines = [x for x in text.split("\n") if x.startswith("s ")]
    assert len(s_lines) == 2
    assert " + " in s_lines[0]
    assert " - " in s_lines[1]
    assert s_lines == ["s t                             0      4 +         10 ACGT",
                       "s q                             0      4 -         10 ACGT"]


# ---- src/output/paf.rs:92-195 -------------------------------------------------------

@pytest.mark.parametrize("writer", sorted(WRITERS))
def test_paf_twelve_columns_plus_tags(writer):  # paf.rs:104-136
    fmt_paf = WRITERS[writer][1]
    targets = [("chrT", padded(b"TTTTTTTTTTACGT", 100))]
    queries = [("chrQ", padded(b"TTTTTACGT", 50))]
    line = fmt_paf(targets, queries, ManualRecords([(0, 0, 0, 10, 5, 4, 4, 400, [(0, 4)])]))
    line = line.decode().split("\n")[0]
    cols = line.split("\t")
    assert len(cols) >= 12
    assert (cols[0], cols[4], cols[5], cols[7], cols[8]) == ("chrQ", "+", "chrT", "10", "14")
    assert "cg:Z:4M" in line and "NM:i:0" in line and "AS:i:400" in line
    assert line == "chrQ\t50\t5\t9\t+\tchrT\t100\t10\t14\t4\t4\t60\tNM:i:0\tAS:i:400\ttp:A:P\tcg:Z:4M"


@pytest.mark.parametrize("writer", sorted(WRITERS))
def test_paf_counts_mismatches_and_indels(writer):  # paf.rs:138-169
    fmt_paf = WRITERS[writer][1]
    targets = [("t", padded(b"ACGT", 10))]
    queries = [("q", padded(b"ACCGA", 10))]
    script = [(0, 2), (1, 1), (0, 2)]  # 2M 1I 2M: AC=AC, insert C, GT vs GA
    text = fmt_paf(targets, queries, ManualRecords([(0, 0, 0, 0, 0, 4, 5, 10, script)])).decode()
    assert "NM:i:2" in text  # one mismatch + one indel column
    assert "cg:Z:2M1I2M" in text
    assert text == "q\t10\t0\t5\t+\tt\t10\t0\t4\t3\t5\t60\tNM:i:2\tAS:i:10\ttp:A:P\tcg:Z:2M1I2M\n"


@pytest.mark.parametrize("writer", sorted(WRITERS))
def test_paf_minus_strand_renders_dash(writer):  # paf.rs:171-194
    fmt_paf = WRITERS[writer][1]
    targets = [("t", padded(b"ACGT", 10))]
    queries = [("q", PS(b"TTTTTTACGT"))]
    text = fmt_paf(targets, queries, ManualRecords([(0, 0, 1, 0, 0, 4, 4, 10, [(0, 4)])])).decode()
    line = text.split("\n")[0]
    assert line.split("\t")[4] == "-"
    assert line == "q\t10\t0\t4\t-\tt\t10\t0\t4\t4\t4\t60\tNM:i:0\tAS:i:10\ttp:A:P\tcg:Z:4M"


@pytest.mark.parametrize("writer", sorted(WRITERS))
def test_writers_on_a_hand_built_gapped_minus_record(writer):
    """A gapped record on the minus strand with N columns: the rows come from
    the target and from the query's reverse complement, and both writers'
    texts satisfy check_maf / check_paf."""
    fmt_maf, fmt_paf = WRITERS[writer]
    t = PS(b"GGGGACGTNACGTACGGGG")
    q_rc = b"CCACGTAACGTCCGTACC"  # the minus strand the record indexes
    q = PS(q_rc).reverse_complement()
    targets, queries = [("tt", t)], [("qq", q)]
    # t[4:16] = ACGTNACGTACG, rc-query[2:15] = ACGTAACGTCCGT
    script = [(0, 5), (1, 1), (0, 4), (1, 1), (0, 2), (2, 1)]
    tsp = 5 + 4 + 2 + 1
    qsp = 5 + 1 + 4 + 1 + 2
    trow = b"ACGTN-ACGT-AC" + b"G"
    qrow = b"ACGTA" + b"A" + b"CGTC" + b"C" + b"GT" + b"-"
    score = rescore(trow, qrow)
    recs = ManualRecords([(0, 0, 1, 4, 2, tsp, qsp, score, script)])
    maf = fmt_maf(targets, queries, recs)
    blocks = check_maf(maf, targets, queries)
    assert len(blocks) == 1 and blocks[0][5] == trow and blocks[0][11] == qrow
    check_paf(fmt_paf(targets, queries, recs), blocks)


# ---- tests/fixture_smoke.rs:53-154: minus strand, chain, tweener, gapped -----------

def test_fixture_smoke_minus_strand_with_tweener_and_gapped():
    targets, queries, cfg, npay = smoke_minus_case()
    recs, _ = O.run(targets, queries, cfg)
    tup = recs.tuples()
    check_smoke_minus(tup, npay, queries[0][1].len)
    # the whole query is the reverse complement of the payload: the minus
    # strand holds it exactly, so the alignment is one perfect block with the
    # HOXD70 diagonal sum as its score
    qrc = queries[0][1].reverse_complement().to_ascii()
    assert [(x[3], x[4], x[5], x[6], x[8]) for x in tup] == [(10, 0, npay, npay, f"{npay}M")]
    assert tup[0][7] == rescore(qrc, qrc)
    for _, (fmt_maf, fmt_paf) in sorted(WRITERS.items()):
        maf = fmt_maf(targets, queries, recs)
        check_paf(fmt_paf(targets, queries, recs), check_maf(maf, targets, queries))


def test_minus_strand_closed_form_record():
    """query = RC(target[a:b]): the minus-strand record is t_start = a, t_span
    = q_span = b - a, q_start = 0 in the reverse-complement frame, score = the
    HOXD70 diagonal sum of target[a:b], and both MAF rows read target[a:b]."""
    rng = np.random.default_rng(5)
    tb = bytes(rng.choice(list(b"ACGT"), 6000).astype(np.uint8))
    a, b = 2100, 2600
    t = PS(tb)
    q = t.slice(a, b).reverse_complement()
    targets, queries = [("t", t)], [("q", q)]
    for gapped in (1, 0):
        cfg = T.config_default()
        cfg.strand = T.STRAND_MINUS
        cfg.gapped_enabled = gapped
        recs, _ = O.run(targets, queries, cfg)
        tup = recs.tuples()
        seg = tb[a:b]
        assert [(x[2], x[3], x[4], x[5], x[6], x[8]) for x in tup] == [
            (T.STRAND_MINUS, a, 0, b - a, b - a, f"{b - a}M")]
        assert tup[0][7] == sum(HOX[IDX[c]][IDX[c]] for c in seg)
        for _, (fmt_maf, fmt_paf) in sorted(WRITERS.items()):
            maf = fmt_maf(targets, queries, recs)
            blocks = check_maf(maf, targets, queries)
            assert blocks[0][5] == seg and blocks[0][11] == seg and blocks[0][9] == "-"
            check_paf(fmt_paf(targets, queries, recs), blocks)


# ---- parity corpus: README.md:76-83, PLAN.md:116-121 -------------------------------

@pytest.mark.parametrize("tname,qname,blocks", CORPUS_PAIRS)
def test_corpus_block_counts(tname, qname, blocks):
    """Default flags (lastz-gxy target query --format maf): the block counts the
    reference publishes, and every block valid against the sequences."""
    targets, queries = corpus(tname), corpus(qname)
    cfg = T.config_default()
    recs, st = O.run(targets, queries, cfg)
    maf = O.format_maf(targets, queries, recs)
    assert maf.count(b"\na score=") == blocks
    got = check_maf(maf, targets, queries)
    assert len(got) == blocks
    check_paf(O.format_paf(targets, queries, recs), got)
    assert gxy.format_maf(targets, queries, recs) == maf
    if (tname, qname) == ("sars_cov2", "sars_cov1"):
        # PLAN.md:116-121: 23 gapped extensions on this pair before the
        # covered-seed skip existed, i.e. 23 HSPs; one record survives
        assert st.hsps == 23
    if tname == qname:
        # a self-alignment is one full-length perfect block on the plus strand
        (score, _, ts, tsz, _, trow, _, qs, qsz, strand, _, qrow) = got[0]
        n = targets[0][1].len
        assert (ts, tsz, qs, qsz, strand) == (0, n, 0, n, "+") and trow == qrow


def test_fixture_files():
    """parity/fixtures: the tiny pair gives the block of src/parity.rs:405-411
    (--seed match12 --hspthresh 500 --gappedthresh 500); the indel pair gives
    one block with a single 1-base insertion scored as the 20 matches minus
    400 + 30 (src/driver.rs:479-506); the multi pair's blocks are all valid."""
    fx = fixture
    cfg = T.config_default()
    cfg.pattern = T.pattern_from_spec("match12")
    cfg.hsp = T.HspParams(910, 500)
    cfg.gapped = T.GappedParams(9400, 500)
    targets, queries = fx("target_tiny"), fx("query_tiny")
    recs, _ = O.run(targets, queries, cfg)
    assert O.format_maf(targets, queries, recs) == (
        b"##maf version=1 program=lastz-gxy\n# scoring=HOXD70\n"
        b"a score=1901\n"
        b"s chrT                         11     20 +         43 GATTACACATGGCATGTCGA\n"
        b"s chrQ                          0     20 +         20 GATTACACATGGCATGTCGA\n\n")

    targets, queries = fx("target_indel"), fx("query_indel")
    c2 = T.config_default()
    c2.pattern = T.pattern_from_spec("match8")
    c2.strand = T.STRAND_PLUS
    c2.hsp = T.HspParams(910, 500)
    c2.gapped = T.GappedParams(3000, 500)
    recs, _ = O.run(targets, queries, c2)
    blocks = check_maf(O.format_maf(targets, queries, recs), targets, queries)
    best = max(blocks, key=lambda b: b[0])
    assert (best[2], best[3], best[7], best[8]) == (10, 20, 0, 21)
    assert best[0] == rescore(b"GATTACACATGGCATGTCGA", b"GATTACACATGGCATGTCGA") - 430

    targets, queries = fx("target_multi"), fx("query_multi")
    recs, _ = O.run(targets, queries, cfg)
    maf = O.format_maf(targets, queries, recs)
    blocks = check_maf(maf, targets, queries)
    assert len(blocks) >= 2
    check_paf(O.format_paf(targets, queries, recs), blocks)
