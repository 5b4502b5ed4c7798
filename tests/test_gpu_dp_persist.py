"""The persistent form of the gapped batch (gxy_ctx_set_dp_persist): one launch
in which every wavefront takes sides from a device queue, sweeps each into its
own share of the traceback pool, walks it back and appends its runs.  Against
the oracle for every one-wavefront kernel it runs, with sides whose runs
overflow the LDS stage (the walk runs twice), sides that outgrow their share
(left to passes) and the automatic trigger on large batches."""
import sys

import numpy as np
import pytest

import gxy
import gxy_types as T
import oracle_py as O
import synth
from test_gpu_dp_kernels import oracle_all, pairs, run_all, slices

pytestmark = pytest.mark.gpu

ONE_WAVE = {"wave_k2": 3, "wave_k4": 4, "wave_k2_w4": 5, "wave_k4_w4": 6}
HOX = T.scoring_hoxd70()


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("kernel", sorted(ONE_WAVE))
def test_persistent_random_pairs(kernel):
    ctx = gxy.Context(0)
    ctx.set_dp_kernel(ONE_WAVE[kernel])
    ctx.set_dp_persist(1)
    rng = np.random.default_rng(201)
    cases = pairs(rng, [int(x) for x in rng.integers(10, 1500, 60)])
    for ydrop in (9400, 3000, 800):
        gp = T.GappedParams(ydrop, 1000)
        say(kernel, "persistent, random pairs, ydrop", ydrop)
        got, cells = run_all(ctx, cases, HOX, gp)
        exp, ecells = oracle_all(cases, HOX, gp)
        assert cells == ecells
        for k in range(len(cases)):
            assert got[k] == exp[k], (kernel, ydrop, k)
    launches, sides, deferred, _ = ctx.dp_persist_stats()
    say("persistent launches", launches, "sides", sides, "deferred", deferred)
    assert launches == 3 and sides > 0
    ctx.close()


def test_persistent_many_runs_walk_twice():
    """Sides of ~5 kbp with indel events at 6 %: hundreds of runs per side, more
    than the block stages in LDS, so the walk runs again into the reserved range."""
    ctx = gxy.Context(0)
    ctx.set_dp_kernel(ONE_WAVE["wave_k4_w4"])
    ctx.set_dp_persist(1)
    rng = np.random.default_rng(202)
    cases = []
    for L in (5000, 4500, 5200, 3800, 4900, 5100):
        code = rng.integers(0, 4, 2 * L + 1)
        t = T.PackedSeq.from_arrays(code, np.ones(2 * L + 1, bool), np.zeros(2 * L + 1, bool))
        c, v, m = t.unpack()
        qc, qv, qm = synth.mutate(c, v, m, rng, 0.04, 0.06)
        q = T.PackedSeq.from_arrays(qc, qv, qm)
        cases.append((t, q, (L, min(L, len(q) - 1))))
    gp = T.GappedParams(9400, 1000)
    got, cells = run_all(ctx, cases, HOX, gp)
    exp, ecells = oracle_all(cases, HOX, gp)
    assert cells == ecells
    assert got == exp
    # the assembled scripts hold two sides' runs each; one side alone exceeds 256
    nruns = [len(g[-1]) if g else 0 for g in got]
    say("runs per alignment", nruns)
    assert max(nruns) > 2 * 256
    assert ctx.dp_persist_stats()[0] == 1
    ctx.close()


def test_persistent_sides_outgrowing_their_share_go_to_passes():
    """Query sides with ten 250-base deletions: the sweep runs ~3.5 k rows where
    the row estimate (query span + 1/8 + 512) says ~1.75 k, so with a pool of
    30 chunks (1920 rows) per block the sides stop in the persistent launch and
    are swept again through passes.  Alignments equal the oracle's."""
    ctx = gxy.Context(0)
    ctx.set_dp_kernel(ONE_WAVE["wave_k2_w4"])
    ctx.set_dp_persist(1)
    rng = np.random.default_rng(203)
    cases = []
    for _ in range(4):
        seg = []
        for _ in range(10):  # 100 kept bases, then 250 deleted from the query
            seg.append((100, True))
            seg.append((250, False))
        tl = sum(n for n, _ in seg) + 100
        tcode = rng.integers(0, 4, 2 * tl + 1)
        keep = np.ones(2 * tl + 1, bool)
        p = tl + 1
        for n, kept in seg:  # right of the anchor
            if not kept:
                keep[p:p + n] = False
            p += n
        p = tl
        for n, kept in seg:  # left of the anchor, mirrored
            if not kept:
                keep[p - n:p] = False
            p -= n
        qcode = tcode[keep]
        t = T.PackedSeq.from_arrays(tcode, np.ones(len(tcode), bool), np.zeros(len(tcode), bool))
        q = T.PackedSeq.from_arrays(qcode, np.ones(len(qcode), bool), np.zeros(len(qcode), bool))
        cases.append((t, q, (tl, int(keep[:tl].sum()))))
    ctx.set_dp_pool_rows(8 * 30 * 64)  # 240 chunks over 8 sides: 30 each
    gp = T.GappedParams(9400, 1000)
    got, cells = run_all(ctx, cases, HOX, gp)
    exp, ecells = oracle_all(cases, HOX, gp)
    launches, sides, deferred, _ = ctx.dp_persist_stats()
    say("persistent launches", launches, "sides", sides, "deferred", deferred, "passes", ctx.dp_pool_stats())
    assert cells == ecells
    assert got == exp
    assert launches == 1 and sides == 8 and deferred == 8
    ctx.close()


def test_persistent_auto_on_large_batches():
    """With the default mode a batch of at least 16384 sides runs persistently
    once its kernel is a one-wavefront variant; a smaller one does not."""
    ctx = gxy.Context(0)
    ctx.set_dp_kernel(ONE_WAVE["wave_k4_w4"])
    rng = np.random.default_rng(204)
    cases = pairs(rng, [int(x) for x in rng.integers(20, 120, 9000)])
    t, q, to, tl, qo, ql, anc = slices(cases)
    dt, dq = gxy.DeviceSeq.upload(t, ctx), gxy.DeviceSeq.upload(q, ctx)
    gp = T.GappedParams(9400, 300)
    exp, ecells = oracle_all(cases, HOX, gp)
    say("auto: 9000 pairs")
    got, cells = gxy.gapped_extend_slices_dev(dt, dq, to, tl, qo, ql, anc, HOX, gp, ctx)
    assert cells == ecells and got == exp
    launches, sides, _, pcells = ctx.dp_persist_stats()
    assert launches == 1 and sides >= 16384 and 0 < pcells <= cells
    say("auto: 3000 pairs")
    n = 3000
    got, cells = gxy.gapped_extend_slices_dev(dt, dq, to[:n], tl[:n], qo[:n], ql[:n], anc[:n], HOX, gp, ctx)
    assert got == exp[:n]
    assert ctx.dp_persist_stats()[0] == 1
    ctx.set_dp_persist(-1)
    say("off: 9000 pairs")
    got, cells = gxy.gapped_extend_slices_dev(dt, dq, to, tl, qo, ql, anc, HOX, gp, ctx)
    assert cells == ecells and got == exp
    assert ctx.dp_persist_stats()[0] == 1
    ctx.close()
