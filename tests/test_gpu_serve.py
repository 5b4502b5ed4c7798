"""Serving loop (gxy_server) on the GPU: every query's records, served in a mix
of concurrently submitted queries, equal gxy_run_dev on that query alone, in
the co-batched and the plain (one query per iteration) modes, gapped and
ungapped, with chunked targets and both strands; and equal the oracle's
driver on that query (the per-query loop of src/driver.rs:219-391), also for
seed patterns whose first or last position is a don't-care."""
import sys
import threading

import numpy as np
import pytest

import gxy
import gxy_types as T
import oracle_py as O
import synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    return gxy.default_context()


@pytest.fixture(scope="module")
def workload():
    # 3 Mbp target in 1 Mbp chunks (halo 50 kbp), 40 % soft-masked; 14
    # queries of 1-60 kbp, excerpts on both strands and decoys, plus an
    # empty-hit query and a query shorter than the seed span
    target, reqs = synth.serve_workload(n_target=3_000_000, n_req=14, seed=5, lo=1000, hi=60_000)
    queries = [q for q, _, _ in reqs]
    queries.append(T.PackedSeq.from_ascii("ACGTACGTAC"))
    queries.append(T.PackedSeq.from_ascii("N" * 500 + "ACGT" * 100))
    return target, queries


def cfg_for(gapped, chain=False):
    cfg = T.config_default()
    cfg.gapped_enabled = 1 if gapped else 0
    cfg.chunk_primary_bp = 1_000_000
    cfg.chunk_halo_bp = 50_000
    if chain:  # --chain --inner (the tweener relaxed further so that gaps get searched)
        cfg.chain_enabled = 1
        cfg.tweener_enabled = 1
        cfg.inner_relax = 6
    return cfg


@pytest.mark.parametrize("gapped,co_batch,chain", [(True, True, False), (True, False, False), (False, True, False),
                                                   (True, True, True)])
def test_served_equals_alone(ctx, workload, gapped, co_batch, chain):
    target, queries = workload
    cfg = cfg_for(gapped, chain)
    dt = gxy.DeviceSeq.upload(target, ctx)
    alone = []
    for k, q in enumerate(queries):
        say("alone", k, len(q))
        recs, _ = gxy.run_dev([dt], [gxy.DeviceSeq.upload(q, ctx)], cfg, ctx)
        alone.append(recs.tuples())
    say("serve gapped", gapped, "co_batch", co_batch)
    srv = gxy.Server([dt], cfg, co_batch=co_batch)
    tickets = [None] * len(queries)

    def submit(idx):
        for k in idx:
            tickets[k] = srv.submit(queries[k])

    # submit from two threads at once
    th = [threading.Thread(target=submit, args=(range(s, len(queries), 2),)) for s in (0, 1)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    iters = set()
    for k in range(len(queries)):
        recs, st = srv.wait(tickets[k])
        assert st.query_bp == len(queries[k])
        assert st.done_ms >= st.admit_ms >= st.submit_ms
        iters.add(st.iteration)
        assert recs.tuples() == alone[k], f"query {k} differs served vs alone"
    srv.close()
    if not co_batch:
        assert len(iters) == len(queries)  # one query per iteration
    assert sum(len(a) for a in alone) > 0


def test_serve_maf_identical(ctx, workload):
    target, queries = workload
    cfg = cfg_for(True)
    dt = gxy.DeviceSeq.upload(target, ctx)
    srv = gxy.Server([dt], cfg, co_batch=True, max_batch_bp=50_000)  # several iterations
    tickets = [srv.submit(q) for q in queries[:6]]
    for k, tk in enumerate(tickets):
        recs, st = srv.wait(tk)
        alone, _ = gxy.run_dev([dt], [gxy.DeviceSeq.upload(queries[k], ctx)], cfg, ctx)
        assert (gxy.format_maf([("t", target)], [("q%d" % k, queries[k])], recs)
                == gxy.format_maf([("t", target)], [("q%d" % k, queries[k])], alone))
    srv.close()


@pytest.fixture(scope="module")
def small_workload():
    # 1.2 Mbp target in 400 kbp chunks (halo 20 kbp), 40 % soft-masked; 14
    # queries of 1-12 kbp (excerpts on both strands and decoys) and one
    # shorter than the seed span: small enough for the oracle in seconds
    target, reqs = synth.serve_workload(n_target=1_200_000, n_req=14, seed=9, lo=1000, hi=12_000)
    queries = [q for q, _, _ in reqs]
    queries.append(T.PackedSeq.from_ascii("ACGTACGTAC"))
    return target, queries


@pytest.mark.parametrize("spec,gapped,chain", [
    ("12of19", True, False), ("12of19", False, False), ("12of19", True, True),
    ("01110100110010101111", True, False), ("11101001100101011110", False, False),
    ("11101001100101011110", True, True)])
def test_served_equals_oracle(ctx, small_workload, spec, gapped, chain):
    """Every served query's records (co-batched, all submitted at once) equal
    the oracle's records for that query; the oracle runs all queries in one
    call and its records are split by query index (each query's records in
    the reference depend on that query alone: src/driver.rs:219-391)."""
    target, queries = small_workload
    cfg = cfg_for(gapped, chain)
    cfg.chunk_primary_bp, cfg.chunk_halo_bp = 400_000, 20_000
    cfg.pattern = T.pattern_from_spec(spec)
    say("oracle", spec, "gapped", gapped, "chain", chain)
    orec, _ = O.run([("t", target)], [("q%d" % k, q) for k, q in enumerate(queries)], cfg)
    per = [[] for _ in queries]
    for x in orec.tuples():
        per[x[1]].append((x[0], 0) + tuple(x[2:]))
    assert sum(len(p) for p in per) > 0
    dt = gxy.DeviceSeq.upload(target, ctx)
    srv = gxy.Server([dt], cfg, co_batch=True)
    tickets = [srv.submit(q) for q in queries]
    for k, tk in enumerate(tickets):
        recs, _ = srv.wait(tk)
        assert recs.tuples() == per[k], f"query {k} differs from the oracle"
    srv.close()
    if spec != "12of19":
        # and equal to the query run alone through the driver
        for k in (0, 6, 8):
            alone, _ = gxy.run_dev([dt], [gxy.DeviceSeq.upload(queries[k], ctx)], cfg, ctx)
            assert alone.tuples() == per[k]


def test_server_ticket_waited_once(ctx, small_workload):
    target, queries = small_workload
    dt = gxy.DeviceSeq.upload(target, ctx)
    cfg = cfg_for(False)
    srv = gxy.Server([dt], cfg, co_batch=True)
    tk = srv.submit(queries[0])
    srv.wait(tk)
    with pytest.raises(gxy.GxyError):
        srv.wait(tk)
    srv.close()
