"""gxy_extend_hit_batch_noalloc against gxy_extend_hit_batch: bit-identical
HSPs and found flags on two calls in a row, the second with smaller inputs
(the shared per-device array keeps the first call's contents), for the byte
x-drop kernel and the int32-matrix kernel; and inputs beyond the array's size
are refused before any device work."""
import sys

import numpy as np
import pytest

import gxy
import gxy_types as T
import synth

pytestmark = pytest.mark.gpu

HOX = T.scoring_hoxd70()


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rand_seq(rng, n, pn=0.02, plow=0.2):
    code = rng.integers(0, 4, n)
    valid = rng.random(n) >= pn
    mask = valid & (rng.random(n) < plow)
    return T.PackedSeq.from_arrays(code, valid, mask)


def case(rng, nt, nq, nhits):
    t = rand_seq(rng, nt)
    c, v, m = t.unpack()
    lo = nt // 4
    qc, qv, qm = synth.mutate(c[lo:lo + nq], v[lo:lo + nq], m[lo:lo + nq], rng, 0.1, 0.005)
    q = T.PackedSeq.from_arrays(qc, qv, qm)
    # half the hits on the homologous diagonal band, half anywhere (some past the ends)
    k = nhits // 2
    qp = rng.integers(0, len(q), nhits).astype(np.uint32)
    tp = np.empty(nhits, np.uint32)
    tp[:k] = np.minimum(qp[:k].astype(np.int64) + lo + rng.integers(-3, 4, k), nt + 5).clip(0).astype(np.uint32)
    tp[k:] = rng.integers(0, nt + 20, nhits - k).astype(np.uint32)
    return t, q, tp, qp


def int32_matrix():
    s = T.scoring_hoxd70()
    for a in range(4):
        for b in range(4):
            s.sub[a][b] = 2 * s.sub[a][b] + (3 if a < b else 0)  # outside int8, asymmetric
    return s


@pytest.mark.parametrize("scoring", ["hoxd70", "int32"])
def test_noalloc_equals_batch_twice(scoring):
    ctx = gxy.default_context()
    sc = HOX if scoring == "hoxd70" else int32_matrix()
    rng = np.random.default_rng(71)
    for nt, nq, nh, seed_len, xd in ((2_000_000, 1_500_000, 400_000, 19, 910),
                                     (30_000, 9_000, 3_000, 12, 400)):
        say("noalloc", scoring, nt, nq, nh)
        t, q, tp, qp = case(rng, nt, nq, nh)
        hp = T.HspParams(xd, 1500)
        ref = gxy.extend_hit_batch(t, q, sc, hp, tp, qp, seed_len, ctx)
        got = gxy.extend_hit_batch_noalloc(t, q, sc, hp, tp, qp, seed_len, ctx)
        assert got == ref
        assert sum(x is not None for x in ref) > 0


def test_noalloc_rejects_inputs_beyond_its_array():
    ctx = gxy.default_context()
    rng = np.random.default_rng(72)
    t = rand_seq(rng, 40_000_000)
    q = rand_seq(rng, 1000)
    with pytest.raises(gxy.GxyError, match=r"\(-5\)"):
        gxy.extend_hit_batch_noalloc(t, q, HOX, T.HspParams(910, 3000), [5], [5], 19, ctx)
    # the array is still usable afterwards
    t, q, tp, qp = case(rng, 50_000, 20_000, 5_000)
    hp = T.HspParams(910, 1500)
    assert (gxy.extend_hit_batch_noalloc(t, q, HOX, hp, tp, qp, 19, ctx)
            == gxy.extend_hit_batch(t, q, HOX, hp, tp, qp, 19, ctx))
