"""Gapped extension with bands wider than the register and LDS kernels hold
(1280 cells): a large y-drop, where the band radius (y_drop - O - E) / E + 2
makes rows of up to ~4000 cells, and gap_extend = 0, where the band is the
whole query (src/gapped_extend.rs:214-233).  Those sides run through the
library's band kernel; alignments and DP cell counts must equal the oracle's,
also when narrow and wide sides share a batch, when a side outruns its
reserved traceback rows and is swept again, and through the driver."""
import sys

import numpy as np
import pytest

import gxy
import gxy_types as T
import oracle_py as O
import synth
from test_gpu_dp_kernels import oracle_all, pairs, run_all

pytestmark = pytest.mark.gpu

HOX = T.scoring_hoxd70()


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def band_width(sc, gp, m):
    oe, e = sc.gap_open + sc.gap_extend, sc.gap_extend
    br = max((gp.y_drop - oe) // e, 0) + 2 if e > 0 else m + 1
    return min(4 * br + 4, m + 1)


@pytest.fixture(scope="module")
def ctx():
    return gxy.Context(0)


def test_large_ydrop_random_pairs(ctx):
    rng = np.random.default_rng(301)
    cases = pairs(rng, [int(x) for x in rng.integers(1400, 4000, 10)] + [int(x) for x in rng.integers(20, 900, 10)],
                  indel_blocks=True)
    for ydrop in (30000, 15000):
        gp = T.GappedParams(ydrop, 1000)
        wide = sum(band_width(HOX, gp, a[1]) > 1280 or band_width(HOX, gp, len(q) - a[1] - 1) > 1280
                   for _, q, a in cases)
        say("y_drop", ydrop, "band", band_width(HOX, gp, 10 ** 6), "pairs with a wide side", wide)
        assert wide >= 5
        got, cells = run_all(ctx, cases, HOX, gp)
        exp, ecells = oracle_all(cases, HOX, gp)
        assert cells == ecells
        for k in range(len(cases)):
            assert got[k] == exp[k], (ydrop, k)


def test_gap_extend_zero(ctx):
    """gap_extend = 0: the band radius is the query length, every row spans the
    whole query; sides of up to 2600 columns next to short ones."""
    sc = T.scoring_hoxd70()
    sc.gap_extend = 0
    rng = np.random.default_rng(302)
    cases = pairs(rng, [int(x) for x in rng.integers(1300, 2600, 6)] + [int(x) for x in rng.integers(30, 500, 6)])
    gp = T.GappedParams(9400, 1000)
    say("gap_extend 0:", len(cases), "pairs")
    got, cells = run_all(ctx, cases, sc, gp)
    exp, ecells = oracle_all(cases, sc, gp)
    assert cells == ecells
    for k in range(len(cases)):
        assert got[k] == exp[k], k


def test_wide_side_outrunning_its_rows_is_swept_again(ctx):
    """Sides of ~2.1 kbp of query (band 2100 cells at y_drop 30000) with ten
    250-base deletions: the sweep runs ~4.6 k rows where the row estimate
    (query span + 1/8 + 512) reserves ~2.9 k, so the side stops and is swept
    again with room for every row."""
    rng = np.random.default_rng(303)
    cases = []
    for _ in range(3):
        seg = []
        for _ in range(10):
            seg.append((200, True))
            seg.append((250, False))
        tl = sum(n for n, _ in seg) + 100
        tcode = rng.integers(0, 4, 2 * tl + 1)
        keep = np.ones(2 * tl + 1, bool)
        p = tl + 1
        for n, kept in seg:
            if not kept:
                keep[p:p + n] = False
            p += n
        p = tl
        for n, kept in seg:
            if not kept:
                keep[p - n:p] = False
            p -= n
        t = T.PackedSeq.from_arrays(tcode, np.ones(len(tcode), bool), np.zeros(len(tcode), bool))
        qcode = tcode[keep]
        q = T.PackedSeq.from_arrays(qcode, np.ones(len(qcode), bool), np.zeros(len(qcode), bool))
        cases.append((t, q, (tl, int(keep[:tl].sum()))))
    gp = T.GappedParams(30000, 1000)
    truncated0 = ctx.dp_pool_stats()[1]
    got, cells = run_all(ctx, cases, HOX, gp)
    exp, ecells = oracle_all(cases, HOX, gp)
    assert cells == ecells
    assert got == exp
    assert all(g is not None for g in got)
    assert all(band_width(HOX, gp, a[1]) > 1280 for _, _, a in cases)
    assert ctx.dp_pool_stats()[1] - truncated0 >= 6  # both sides of every pair stopped once


def test_driver_large_ydrop_matches_oracle(ctx):
    """The whole driver (both strands, gapped, MAF) on the 30 kbp pair with
    y_drop 30000: every gapped side is a wide one."""
    target, query = synth.pair_30k(seed=3)
    cfg = T.config_default()
    cfg.gapped.y_drop = 30000
    say("driver, y_drop 30000")
    recs, _ = gxy.run([("t", target)], [("q", query)], cfg, ctx)
    orecs, _ = O.run([("t", target)], [("q", query)], cfg)
    assert len(recs) > 0
    assert recs.tuples() == orecs.tuples()
    assert gxy.format_maf([("t", target)], [("q", query)], recs) == O.format_maf(
        [("t", target)], [("q", query)], orecs)
