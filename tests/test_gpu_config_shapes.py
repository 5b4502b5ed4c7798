"""Parity at the benchmark configurations' own shapes.

* The chromosome-scale HSP-only run cuts the target into 10 Mbp chunks with
  50 kbp halos (src/driver.rs:156-194).  Here a 20.5 Mbp target of the same
  kind (50 % soft-masked, 2 % N, segmental duplications) gives three chunks
  at exactly that shape, searched with a 6 Mbp query excerpt that spans the
  boundary at 10 Mbp, both strands, 12of19, one transition, x-drop 910,
  hspthresh 3000, --nogapped: records, statistics and MAF bytes must equal
  the oracle's, and records must fall in the halos on both sides of the
  boundary (where the primary-range filter of src/driver.rs:288-297 decides
  which chunk reports them).
* The batched gapped DP draws extension lengths from LogNormal(median 2 kbp,
  sigma 1) clipped to 100 bp - 50 kbp, identity U[0.70, 0.95]; here 48 pairs
  from that law plus 3 pairs of 20-50 kbp per side go through the slices
  entry point in one call (the persistent launch forced on, the
  four-wavefront kernel in passes, and the automatic choice): alignments and
  cell counts must equal the oracle's."""
import sys
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import gxy
import gxy_types as T
import oracle_py as O
import synth

pytestmark = pytest.mark.gpu

HOX = T.scoring_hoxd70()


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    return gxy.Context(0)


def test_chrom_chunk_shape(ctx):
    target, query = synth.chromosome_pair(n=20_500_000, seed=9)
    query = query.slice(7_000_000, 13_000_000)
    targets, queries = [("chr", target)], [("qry", query)]
    cfg = T.config_default()
    cfg.gapped_enabled = 0
    cfg.chunk_primary_bp = 10_000_000
    cfg.chunk_halo_bp = 50_000
    say("10 Mbp chunks / 50 kbp halo:", len(target), "bp target,", len(query), "bp query")
    g, gst = gxy.run(targets, queries, cfg, ctx)
    say("gpu:", len(g), "records")
    o, ost = O.run(targets, queries, cfg)
    say("oracle:", len(o.tuples()), "records")
    assert gst.chunks == ost.chunks == 3
    assert gst.hits == ost.hits
    assert len(g) > 1000
    assert g.tuples() == o.tuples()
    assert gxy.format_maf(targets, queries, g) == O.format_maf(targets, queries, o)
    # records reported from the halos on both sides of the 10 Mbp boundary
    ts = np.array([r[3] for r in g.tuples()], np.int64)
    assert ((ts >= 10_000_000 - 50_000) & (ts < 10_000_000)).sum() > 0
    assert ((ts >= 10_000_000) & (ts < 10_000_000 + 50_000)).sum() > 0


def _cases(n, seed, **kw):
    return [(t, q, (tp, qp)) for t, q, tp, qp in synth.dp_anchors(n, seed, **kw)]


def test_dp_config_length_law(ctx):
    cases = _cases(48, 21) + _cases(3, 22, median=45_000.0, sigma=0.1, lo=20_000, hi=50_000)
    Ls = [a[0] for _, _, a in cases]
    say("pairs:", len(cases), "extension lengths median", int(np.median(Ls)), "max", max(Ls))
    assert max(Ls) >= 20_000
    gp = T.GappedParams(9400, 3000)
    # the oracle, pairs in parallel (the long sides dominate)
    with ThreadPoolExecutor(8) as ex:
        ores = list(ex.map(lambda c: O.gapped_extend(c[2], c[0], c[1], HOX, gp), cases))
    exp = [r for r, _ in ores]
    ecells = sum(c for _, c in ores)
    say("oracle done:", ecells, "cells")
    # all pairs as slices of two sequences, one call
    tcode = np.concatenate([t.unpack()[0] for t, _, _ in cases])
    qcode = np.concatenate([q.unpack()[0] for _, q, _ in cases])
    tlen = np.array([t.len for t, _, _ in cases], np.int64)
    qlen = np.array([q.len for _, q, _ in cases], np.int64)
    toff = np.concatenate([[0], np.cumsum(tlen)[:-1]])
    qoff = np.concatenate([[0], np.cumsum(qlen)[:-1]])
    anchors = np.array([a for _, _, a in cases], np.int64)
    T_ = T.PackedSeq.from_arrays(tcode, np.ones(len(tcode), bool), np.zeros(len(tcode), bool))
    Q_ = T.PackedSeq.from_arrays(qcode, np.ones(len(qcode), bool), np.zeros(len(qcode), bool))
    dT = gxy.DeviceSeq.upload(T_, ctx)
    dQ = gxy.DeviceSeq.upload(Q_, ctx)
    for kernel, persist in ((6, 1), (1, -1), (0, 0)):
        ctx.set_dp_kernel(kernel)
        ctx.set_dp_persist(persist)
        got, cells = gxy.gapped_extend_slices_dev(dT, dQ, toff, tlen, qoff, qlen, anchors, HOX, gp, ctx)
        say("kernel", kernel, "persist", persist, ":", cells, "cells")
        assert cells == ecells
        for k in range(len(cases)):
            assert got[k] == exp[k], (kernel, k)
    ctx.set_dp_kernel(0)
    ctx.set_dp_persist(0)
