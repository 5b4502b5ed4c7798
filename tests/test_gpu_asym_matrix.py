"""GPU parity with an asymmetric substitution matrix (the target base indexes
the row, src/scoring.rs:17-64): a transposed lookup anywhere on the device —
x-drop, anchor choice, gapped DP, driver — changes scores and fails here."""
import sys

import numpy as np
import pytest

import gxy
import gxy_types as T
import oracle_py as O
from test_gpu_parity import homologous_pair

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def asym_scoring():
    """HOXD70 with its off-diagonal entries perturbed differently above and
    below the diagonal (matches unchanged, so homology still scores)."""
    sc = T.scoring_hoxd70()
    rng = np.random.default_rng(99)
    for i in range(4):
        for j in range(4):
            if i != j:
                sc.sub[i][j] = int(sc.sub[i][j]) + int(rng.integers(-60, 61))
    assert any(sc.sub[i][j] != sc.sub[j][i] for i in range(4) for j in range(4))
    return sc


@pytest.fixture(scope="module")
def ctx():
    return gxy.default_context()


def test_asym_extend_anchor_gapped(ctx):
    say("asymmetric matrix: extend_hit, choose_anchor, gapped_extend")
    sc = asym_scoring()
    rng = np.random.default_rng(41)
    t, q = homologous_pair(rng, 20000, 0.12, 0.0)
    n = 400
    tp = rng.integers(0, len(t) - 19, n).astype(np.int64)
    qp = rng.integers(0, len(q) - 19, n).astype(np.int64)
    p = T.HspParams(910, 0)
    g = gxy.extend_hit_batch(t, q, sc, p, tp, qp, 19, ctx)
    for k in range(n):
        assert g[k] == O.extend_hit(t, q, sc, p, int(tp[k]), int(qp[k]), 19, False), k
    hs = []
    for _ in range(200):
        L = int(rng.integers(1, 400))
        hs.append((int(rng.integers(0, len(t) - L)), int(rng.integers(0, len(q) - L)), L, 0))
    for w in (1, 31):
        assert gxy.choose_anchor_batch(hs, t, q, sc, w, ctx) == [O.choose_anchor(h, t, q, sc, w) for h in hs]
    t2, q2 = homologous_pair(rng, 6000, 0.15, 0.02, 0.0)
    lo = len(t2) // 5
    anchors = [(min(len(t2) - 1, a + lo), a) for a in rng.integers(0, len(q2) - 1, 16).tolist()]
    gp = T.GappedParams(3000, 0)
    gg, _ = gxy.gapped_extend_batch(anchors, t2, q2, sc, gp, ctx)
    for k, a in enumerate(anchors):
        assert gg[k] == O.gapped_extend(a, t2, q2, sc, gp)[0], (k, a)


@pytest.mark.parametrize("gapped", [0, 1])
def test_asym_search_and_run(ctx, gapped):
    say("asymmetric matrix: search and run, gapped", gapped)
    sc = asym_scoring()
    rng = np.random.default_rng(42 + gapped)
    t, q = homologous_pair(rng, 60000, 0.12, 0.01, 0.2)
    pat = T.pattern_from_spec("12of19")
    sp = T.SearchParams(1, T.HspParams(910, 3000), 1, 0, 3000)
    gh, _ = gxy.search(gxy.PosTable.build(t, pat, 1, ctx), t, q, sc, sp)
    oh, _ = O.search(O.PosTable(t, pat, 1), t, q, sc, sp)
    assert len(gh) > 0 and gh == oh
    targets, queries = [("t", t)], [("q", q)]
    cfg = T.config_default()
    cfg.scoring = sc
    cfg.gapped_enabled = gapped
    g, _ = gxy.run(targets, queries, cfg, ctx)
    o, _ = O.run(targets, queries, cfg)
    om = O.format_maf(targets, queries, o)
    assert len(o.tuples()) > 0
    assert gxy.format_maf(targets, queries, g) == om
