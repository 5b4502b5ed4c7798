"""The gapped extension's traceback pool when it is too small for the work:
sides that find it empty stop and are swept again, a pass that finishes
nothing is followed by one that runs a side alone, and a side longer than the
whole pool is an error.  With a 2048-row pool (32 chunks of 64 rows) every one
of these paths runs, and the alignments still equal the oracle's and those of
the default pool."""
import sys

import numpy as np
import pytest

import gxy
import gxy_types as T
import oracle_py as O
import synth

pytestmark = pytest.mark.gpu

HOX = T.scoring_hoxd70()
POOL_ROWS = 2048


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def pairs(rng, lens, ctx):
    ts, qs, anc, host = [], [], [], []
    for L in lens:
        code = rng.integers(0, 4, 2 * L + 1)
        valid = rng.random(2 * L + 1) >= 0.005
        t = T.PackedSeq.from_arrays(code, valid, np.zeros(2 * L + 1, bool))
        c, v, m = t.unpack()
        qc, qv, qm = synth.mutate(c, v, m, rng, float(rng.uniform(0.03, 0.2)), 0.01)
        q = T.PackedSeq.from_arrays(qc, qv, qm)
        ts.append(gxy.DeviceSeq.upload(t, ctx))
        qs.append(gxy.DeviceSeq.upload(q, ctx))
        anc.append((L, min(L, len(q) - 1)))
        host.append((t, q))
    return ts, qs, anc, host


def test_small_pool_truncates_resweeps_and_runs_alone():
    ctx = gxy.Context(0)
    ctx.set_dp_pool_rows(POOL_ROWS)
    rng = np.random.default_rng(41)
    # 40 sides of 600-1990 rows (10-32 chunks each) and 20 of 30-300 rows
    lens = [int(x) for x in rng.integers(600, 1990, 20)] + [int(x) for x in rng.integers(30, 300, 10)]
    ts, qs, anc, host = pairs(rng, lens, ctx)
    gp = T.GappedParams(9400, 3000)
    say("gapped", len(anc), "pairs through a", POOL_ROWS, "row pool")
    g, cells = gxy.gapped_extend_pairs_dev(ts, qs, anc, HOX, gp, ctx)
    passes, truncated, alone = ctx.dp_pool_stats()
    say("passes", passes, "truncated", truncated, "alone", alone)
    assert truncated > 0 and alone > 0 and passes > alone
    ocells = 0
    for k in range(len(anc)):
        o, c = O.gapped_extend(anc[k], host[k][0], host[k][1], HOX, gp)
        ocells += c
        assert g[k] == o, k
    assert cells == ocells
    say("same pairs through the default pool")
    big = gxy.Context(0)
    g2, cells2 = gxy.gapped_extend_pairs_dev(ts, qs, anc, HOX, gp, big)
    assert g2 == g and cells2 == cells
    assert big.dp_pool_stats()[1] == 0


def test_side_longer_than_the_pool_is_an_error():
    ctx = gxy.Context(0)
    ctx.set_dp_pool_rows(POOL_ROWS)
    rng = np.random.default_rng(42)
    ts, qs, anc, host = pairs(rng, [2600], ctx)
    gp = T.GappedParams(9400, 3000)
    say("gapped one 2600-row side through a", POOL_ROWS, "row pool")
    with pytest.raises(gxy.GxyError, match=r"\(-2\).*traceback pool"):
        gxy.gapped_extend_pairs_dev(ts, qs, anc, HOX, gp, ctx)
    # the context stays usable, and a pool of the default size takes the side
    ts2, qs2, anc2, host2 = pairs(rng, [400, 900], ctx)
    g, _ = gxy.gapped_extend_pairs_dev(ts2, qs2, anc2, HOX, gp, ctx)
    for k in range(2):
        assert g[k] == O.gapped_extend(anc2[k], host2[k][0], host2[k][1], HOX, gp)[0]
    ctx.set_dp_pool_rows(0)
    g, _ = gxy.gapped_extend_pairs_dev(ts, qs, anc, HOX, gp, ctx)
    assert g[0] == O.gapped_extend(anc[0], host[0][0], host[0][1], HOX, gp)[0]


def test_release_cache_keeps_sequences_and_results():
    """gxy_ctx_release_cache returns the scratch buffers and the traceback pool
    between calls; resident sequences stay valid and the next call, which
    allocates again, gives the same alignments (including a two-strand driver
    run, whose helper context is released too)."""
    ctx = gxy.Context(0)
    rng = np.random.default_rng(43)
    ts, qs, anc, host = pairs(rng, [int(x) for x in rng.integers(200, 1500, 12)], ctx)
    gp = T.GappedParams(9400, 3000)
    g1, c1 = gxy.gapped_extend_pairs_dev(ts, qs, anc, HOX, gp, ctx)
    say("released after a gapped batch")
    ctx.release_cache()
    g2, c2 = gxy.gapped_extend_pairs_dev(ts, qs, anc, HOX, gp, ctx)
    assert g2 == g1 and c2 == c1
    target, query = synth.pair_30k(seed=5)
    dt, dq = gxy.DeviceSeq.upload(target, ctx), gxy.DeviceSeq.upload(query, ctx)
    cfg = T.config_default()
    r1 = gxy.run_dev([dt], [dq], cfg, ctx)[0].tuples()
    say("released after a two-strand driver run")
    ctx.release_cache()
    r2 = gxy.run_dev([dt], [dq], cfg, ctx)[0].tuples()
    assert r1 == r2 and len(r1) > 0
    ctx.release_cache()
    g3, c3 = gxy.gapped_extend_pairs_dev(ts, qs, anc, HOX, gp, ctx)
    assert g3 == g1 and c3 == c1
