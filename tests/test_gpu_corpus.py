"""GPU parity on the reference's own inputs: the six parity-corpus pairs
(parity/scripts/matrix.sh:21-28, default flags), the three parity fixtures and
the minus-strand chain + tweener case of tests/fixture_smoke.rs:53-154, run
through gxy_run (the C-ABI driver) and compared byte for byte with the
oracle's MAF and PAF.  Each text is also checked against the format-level
invariants (outcheck.check_maf / check_paf) and the published block counts
(README.md:76-83)."""
import sys

import pytest

import gxy
import gxy_types as T
import oracle_py as O
from outcheck import (CORPUS_PAIRS, check_maf, check_paf, check_smoke_minus, corpus, fixture,
                      smoke_minus_case)

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    return gxy.default_context()


def both(targets, queries, cfg, ctx):
    recs, st = gxy.run(targets, queries, cfg, ctx)
    orec, ost = O.run(targets, queries, cfg)
    maf, omaf = gxy.format_maf(targets, queries, recs), O.format_maf(targets, queries, orec)
    paf, opaf = gxy.format_paf(targets, queries, recs), O.format_paf(targets, queries, orec)
    assert maf == omaf
    assert paf == opaf
    blocks = check_maf(maf, targets, queries)
    check_paf(paf, blocks)
    assert (st.hsps, st.dp_problems) == (ost.hsps, ost.dp_problems)
    return recs, st, blocks


@pytest.mark.parametrize("tname,qname,nblocks", CORPUS_PAIRS)
def test_corpus_pair_maf_paf_identical(ctx, tname, qname, nblocks):
    say("corpus", tname, qname)
    targets, queries = corpus(tname), corpus(qname)
    _, st, blocks = both(targets, queries, T.config_default(), ctx)
    assert len(blocks) == nblocks
    if (tname, qname) == ("sars_cov2", "sars_cov1"):
        assert st.hsps == 23  # PLAN.md:116-121


@pytest.mark.parametrize("chunk", [0, 7000])
def test_corpus_cross_pair_chunked_and_hsp_only(ctx, chunk):
    """The multi-contig cross pair chunked (both strands, gapped) and without
    gapped extension: MAF/PAF identical to the oracle's."""
    targets, queries = corpus("pseudocat"), corpus("pseudopig")
    for gapped in (1, 0):
        cfg = T.config_default()
        cfg.gapped_enabled = gapped
        if chunk:
            cfg.chunk_primary_bp, cfg.chunk_halo_bp = chunk, 1500
        say("corpus cat x pig chunk", chunk, "gapped", gapped)
        both(targets, queries, cfg, ctx)


def test_fixture_files(ctx):
    cfg = T.config_default()
    cfg.pattern = T.pattern_from_spec("match12")
    cfg.hsp = T.HspParams(910, 500)
    cfg.gapped = T.GappedParams(9400, 500)
    for name in ("tiny", "indel", "multi"):
        say("fixture", name)
        targets, queries = fixture("target_" + name), fixture("query_" + name)
        both(targets, queries, cfg, ctx)


def test_fixture_smoke_minus_strand(ctx):
    targets, queries, cfg, npay = smoke_minus_case()
    recs, _, _ = both(targets, queries, cfg, ctx)
    check_smoke_minus(recs.tuples(), npay, queries[0][1].len)
