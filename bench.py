"""Benchmark: the seed -> HSP -> gapped hot path on one MI355X, all five workloads.

The default run (`--workload all`) prints one JSON line whose headline
(`value`, `roofline`, `cpu_baseline`) is the chromosome-scale HSP-only pair,
with one sub-object per other workload, each carrying its own `roofline` (for
its dominant kernel, measured with HIP events over its timed region) and
`cpu_baseline` (the oracle, the C++ restatement of the reference's path, on a
bounded sample, whose output is also compared with the GPU's on that sample):

* headline (`--workload chrom`): 48 Mbp uniform target with 50 % soft-masked
  blocks and 2 % N-runs, query = target plus 5 % segmental duplications at
  15 % substitutions and 2 % indel events (~50.4 Mbp); 12of19, transitions 1,
  x-drop 910, hspthresh 3000, --nogapped, --chunk-size 10000000 --halo 50000.
  One step is a full driver::run over the resident pair.
* `pair30k` (`--workload pair30k`): the 30 kbp pair, gapped, MAF.
* `batched_dp` (`--workload dp`): anchored pairs with extension lengths
  LogNormal(median 2 kbp, sigma 1) clipped to 100 bp-50 kbp at 70-95 %
  identity, HOXD70, gap 400/30, ydrop 9400, gappedthresh 3000, all pairs in one
  gxy_gapped_extend_slices_dev call (1,000,000 pairs in the default run).
* `serve` (`--workload serve`): the serving loop (gxy_server) on a resident
  100 Mbp target with 40 % soft-masked blocks; queries of log-uniform length,
  80 % mutated excerpts (identity 75-95 %, either strand) and 20 % decoys,
  full seed -> HSP -> gapped per query, Poisson arrivals; p50 / p99 latency
  (arrival -> records) at fractions of the served and of the plain capacity.
* `contigs` (`--workload contigs`): a 200 Mbp unmasked target against contigs
  of log-uniform length 1 kbp-1 Mbp at 70-90 % identity with ~3 % of bases in
  indels; 12of19, --chain, --inner (match12 tweener), gapped, PAF; batched (one
  gxy_run_dev) and served.

Size cuts that keep the default run inside the harness's time limit are named
in each sub-object's `config` (`full_size` gives the configuration's own size).
A leg that would not finish inside `--budget-s` is skipped and says so.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "lastz-gxy_amd"))

METRIC = "end-to-end query Mbp/s + HSP/DP GCUPS; p50/p99 serve latency vs plain GPU"
HBM_PEAK_GBS = 8000.0
# int32 VALU: 256 CUs x 128 lane-ops per clock x 2.4 GHz = 78.6 T ops/s, the
# lane-op rate behind the published 157.3 TFLOP/s vector fp32 figure (an FMA
# counts 2).  A 16-lane-per-SIMD model (39.3 T) is too low: an int32 add/xor
# loop over 8192 wavefronts ran 51.4 T lane-ops/s on the MI355X.
VALU_I32_PEAK_TOPS = 256 * 128 * 2.4e9 / 1e12
DP_OPS_PER_CELL = 11  # X: 2 sub + max; Y: 2 sub + max; M: add + 2 max; cell best: 2 max


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def measured_copy_gbs(nbytes=1 << 30, reps=10):
    """HBM rate of a plain device-to-device copy (the runtime's streaming
    copy kernel), read + write bytes over HIP-event time; None on any error."""
    import ctypes
    try:
        hip = ctypes.CDLL("libamdhip64.so")
    except OSError:
        return None
    vp = ctypes.c_void_p
    a, b, e0, e1 = vp(), vp(), vp(), vp()
    ok = (hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0 and
          hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0 and
          hip.hipEventCreate(ctypes.byref(e0)) == 0 and hip.hipEventCreate(ctypes.byref(e1)) == 0)
    ms = ctypes.c_float(0)
    if ok:
        ok = hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) == 0
        ok = ok and hip.hipMemcpyAsync(b, a, ctypes.c_size_t(nbytes), 3, None) == 0  # warm
        ok = ok and hip.hipEventRecord(e0, None) == 0
        for _ in range(reps if ok else 0):
            ok = ok and hip.hipMemcpyAsync(b, a, ctypes.c_size_t(nbytes), 3, None) == 0
        ok = ok and hip.hipEventRecord(e1, None) == 0 and hip.hipEventSynchronize(e1) == 0
        ok = ok and hip.hipEventElapsedTime(ctypes.byref(ms), e0, e1) == 0
    for p in (a, b):
        if p.value:
            hip.hipFree(p)
    for e in (e0, e1):
        if e.value:
            hip.hipEventDestroy(e)
    if not ok or ms.value <= 0:
        return None
    return round(2.0 * nbytes * reps / (ms.value / 1e3) / 1e9, 1)


def valu_roofline(cells, kernel_ms, launches, kernel="k_dp_forward"):
    """VALU roofline of the DP forward kernel: 11 int32 ops per cell against
    the int32 VALU peak."""
    in_kernel = cells / (kernel_ms / 1e3) / 1e9 if kernel_ms else None
    peak_gcups = VALU_I32_PEAK_TOPS * 1e3 / DP_OPS_PER_CELL
    return {"bound": "valu", "kernel": kernel,
            "achieved": round(in_kernel * DP_OPS_PER_CELL / 1e3, 3) if in_kernel else None,
            "peak": round(VALU_I32_PEAK_TOPS, 2), "unit": "T int32 ops/s",
            "frac": round(in_kernel / peak_gcups, 5) if in_kernel else None,
            "gcups_in_kernel": round(in_kernel, 2) if in_kernel else None,
            "peak_gcups": round(peak_gcups, 1), "ops_per_cell": DP_OPS_PER_CELL,
            "launches": launches, "avg_launch_ms": round(kernel_ms / max(1, launches), 4) if kernel_ms else None,
            "traffic": None}


def dp_leg(n_anchors, steps, warmup, cpu_anchors, seed=3, kernel=0, workers=16):
    """Batched gapped DP over n_anchors synthetic anchored pairs held as slices
    of two resident device sequences.  Returns a dict (GCUPS, kernel times,
    VALU roofline, oracle baseline on a stratified sample whose alignments are
    compared with those the timed call returned)."""
    import numpy as np
    import gxy
    import gxy_types as T
    import synth

    t0 = time.time()
    Tq, Qq, toff, tlen, qoff, qlen, tpos, qpos = synth.dp_anchor_slices(n_anchors, seed,
                                                                        workers=min(workers, os.cpu_count() or 1))
    log(f"dp: generated {n_anchors} pairs ({int(tlen.sum())} + {int(qlen.sum())} bp) in {time.time() - t0:.1f}s")
    ctx = gxy.Context(0)  # its own context: its traceback pool goes with it
    ctx.set_dp_kernel(kernel)
    t0 = time.time()
    dT = gxy.DeviceSeq.upload(Tq, ctx)
    dQ = gxy.DeviceSeq.upload(Qq, ctx)
    ctx.synchronize()
    log(f"dp: resident in {time.time() - t0:.1f}s")
    hox = T.scoring_hoxd70()
    gp = T.GappedParams(9400, 3000)
    anchors = np.stack([tpos, qpos], axis=1)
    # stratified sample (equal counts from each decile of extension length)
    # for the CPU baseline, checked against the timed call's own results
    order = np.argsort(tpos, kind="stable")
    per = max(1, cpu_anchors // 10)
    pick = []
    if cpu_anchors:
        for b in range(10):
            seg = order[b * n_anchors // 10:(b + 1) * n_anchors // 10]
            if len(seg):
                pick.extend(seg[np.linspace(0, len(seg) - 1, min(per, len(seg))).astype(int)].tolist())

    def once(pk=None):
        return gxy.gapped_extend_slices_dev(dT, dQ, toff, tlen, qoff, qlen, anchors, hox, gp, ctx, raw=True,
                                            pick=pk)

    for _ in range(warmup):
        once()
    ctx.reset_timing()
    ctx.set_timing(True)
    ctx.synchronize()
    p0 = ctx.dp_persist_stats()
    t1 = time.perf_counter()
    for st in range(steps):
        found, cells, picked = once(pick if st == steps - 1 else None)
    ctx.synchronize()
    el = time.perf_counter() - t1
    ctx.set_timing(False)
    kt = {nm: ctx.kernel_time(nm) for nm in ("dp_forward", "dp_traceback", "dp_persist")}
    gcups = cells * steps / el / 1e9
    fwd_ms, fwd_n = kt["dp_forward"]
    p_launch, p_sides, p_deferred, p_cells = (b - a for a, b in zip(p0, ctx.dp_persist_stats()))
    if kt["dp_persist"][1]:
        # the persistent launch (sweep and walk of most sides) dominates; its
        # cells are those of the sides it finished in the timed steps
        p_ms, p_n = kt["dp_persist"]
        roof = valu_roofline(p_cells, p_ms, p_n, kernel="k_dp_persist")
        roof["note"] = "k_dp_persist's time includes each side's traceback walk"
    else:
        roof = valu_roofline(cells * steps, fwd_ms, fwd_n)
    out = {
        "workload": f"{n_anchors} anchored pairs, extension LogNormal(median 2000, sigma 1) clipped 100-50000 bp, "
                    f"identity U[0.70, 0.95], HOXD70, gap 400/30, ydrop 9400, gappedthresh 3000, one call",
        "anchors": n_anchors, "cells_per_step": cells, "found": found, "steps": steps, "warmup": warmup,
        "ms_per_step": round(1000 * el / steps, 2), "dp_gcups": round(gcups, 3),
        "kernel_ms_per_step": {k: round(v[0] / steps, 3) for k, v in kt.items() if v[1]},
        "roofline": roof,
        "dp_autotune": ctx.dp_autotune_state(),
        "dp_pool_passes": ctx.dp_pool_stats()[0],
        "dp_persistent_timed_steps": {"launches": p_launch, "sides": p_sides, "sides_left_to_passes": p_deferred,
                                      "cells": p_cells},
    }
    log(f"dp: {cells} cells/step, {found}/{n_anchors} found, {gcups:.2f} GCUPS end to end, kernels {kt}")

    if cpu_anchors:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_py as O
        from concurrent.futures import ThreadPoolExecutor
        hp = [(synth.slice_of(Tq, toff[k], tlen[k]), synth.slice_of(Qq, qoff[k], qlen[k]), int(tpos[k]), int(qpos[k]))
              for k in pick]
        nthreads = os.cpu_count() or 1

        def one(p):
            return O.gapped_extend((p[2], p[3]), p[0], p[1], hox, gp)

        t1 = time.perf_counter()
        with ThreadPoolExecutor(nthreads) as ex:
            ores = list(ex.map(one, hp))
        cpu_s = time.perf_counter() - t1
        ocells = sum(c for _, c in ores)
        same = all(picked[k] == r for k, (r, _) in zip(pick, ores))
        out["cpu_baseline"] = {
            "value": round(ocells / cpu_s / 1e9, 4), "unit": "GCUPS", "cores": nthreads, "kind": "port",
            "sample": f"oracle (lazy-row DP) on {len(pick)} of the pairs, {per} from each decile of "
                      f"extension length: {ocells} cells in {cpu_s:.2f} s",
            "gpu_output_identical_on_sample": bool(same),
            "gpu_output_checked": "the sampled pairs' alignments as returned by the last timed call"}
        log(f"dp: cpu baseline {ocells / cpu_s / 1e9:.3f} GCUPS on {len(pick)} pairs, identical: {same}")
    del dT, dQ
    ctx.close()
    return out


def hbm_xdrop_roofline(hits, xds, kernel_ms, launches, traffic=None):
    """HBM roofline of k_xdrop_first, the x-drop's first step (the kernel
    timed as "extend"): per hit 8 B of record, per column it evaluates 0.75 B
    (2 code bits + 1 valid bit on each side), 36 B per continuation record it
    writes and 24 B per passing extension it appends.  xds: the context's
    (first-step columns, resumed hits, first-step passing) over the same calls."""
    cols, resumed, fpass = xds
    alg = 8 * hits + 0.75 * cols + 36 * resumed + 24 * fpass
    achieved = alg / (kernel_ms / 1e3) / 1e9 if kernel_ms else None
    return {"bound": "hbm", "kernel": "k_xdrop_first", "achieved": round(achieved, 2) if achieved else None,
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5) if achieved else None,
            "launches": launches, "avg_launch_ms": round(kernel_ms / max(1, launches), 4) if kernel_ms else None,
            "algorithmic_bytes_per_launch": round(alg / max(1, launches)), "traffic": traffic,
            "bytes_model": "8 B/hit + 0.75 B/first-step column + 36 B/resumed hit + 24 B/first-step passing HSP"}


def hbm_qscan_roofline(lookups, hits, kernel_ms, launches, traffic=None):
    """HBM roofline of k_qexpand_w, the query scan's expansion (timed as
    "qexpand"): the reference's bucket reads for every lookup (8 B: a word's two
    `starts` entries, src/pos_table.rs:102-108), 4 B per hit position and 8 B
    per hit record written — the lookups' bytes counted on this kernel, though
    the count pass (k_qcount_v) does part of that work."""
    alg = 8 * lookups + 12 * hits
    achieved = alg / (kernel_ms / 1e3) / 1e9 if kernel_ms else None
    return {"bound": "hbm", "kernel": "k_qexpand_w", "achieved": round(achieved, 2) if achieved else None,
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5) if achieved else None,
            "launches": launches, "avg_launch_ms": round(kernel_ms / max(1, launches), 4) if kernel_ms else None,
            "algorithmic_bytes_per_launch": round(alg / max(1, launches)), "traffic": traffic,
            "bytes_model": "8 B/lookup (query windows x seed variants) + 4 B/hit position + 8 B/hit record"}


def dominant_roofline(ktimes, st, xds):
    """Roofline of whichever of the DP forward and x-drop first-step kernels took longer."""
    dp_ms, dp_n = ktimes.get("dp_forward", (0.0, 0))
    xd_ms, xd_n = ktimes.get("extend", (0.0, 0))
    if dp_ms >= xd_ms and dp_ms > 0:
        return valu_roofline(st.dp_cells, dp_ms, dp_n)
    if xd_ms > 0:
        return hbm_xdrop_roofline(st.hits, xds, xd_ms, xd_n)
    return None


def serve_leg(args, budget_left=None):
    """The serving workload; returns its measurements."""
    import threading
    import numpy as np
    import gxy
    import gxy_types as T
    import synth

    t0 = time.time()
    target, reqs = synth.serve_workload(n_target=args.serve_target_bp, n_req=args.serve_requests, seed=11,
                                        lo=1000, hi=args.serve_max_len)
    queries = [q for q, _, _ in reqs]
    lens = np.array([len(q) for q in queries], np.float64)
    log(f"serve: target {len(target)} bp, {len(queries)} queries, {int(lens.sum())} bp "
        f"(generated in {time.time() - t0:.1f}s)")
    ctx = gxy.default_context()
    dt = gxy.DeviceSeq.upload(target, ctx)
    ctx.synchronize()
    cfg = T.config_default()
    rng = np.random.default_rng(12)
    gaps = rng.exponential(1.0, len(queries))
    keep = {}  # records of sampled queries from the first served run: {index: tuples}

    def run(co_batch, idx, rate, timing=False, keep_idx=()):
        """Serve queries idx (arriving as a Poisson stream of `rate` queries/s,
        or all at once when rate is None); latency = arrival -> records."""
        srv = gxy.Server([dt], cfg, co_batch=co_batch)
        if timing:
            srv.set_timing(True)
        arrivals = None if rate is None else np.cumsum(gaps[:len(idx)]) / rate
        tickets = [None] * len(idx)
        start = time.perf_counter()

        def submitter():
            for x, k in enumerate(idx):
                if arrivals is not None:
                    dt_s = arrivals[x] - (time.perf_counter() - start)
                    if dt_s > 0:
                        time.sleep(dt_s)
                tickets[x] = srv.submit(queries[k])

        th = threading.Thread(target=submitter)
        th.start()
        th.join()
        lat, sub, done, nrec = [], [], [], 0
        for x, k in enumerate(idx):
            recs, rst = srv.wait(tickets[x])
            nrec += len(recs)
            lat.append(rst.done_ms - rst.submit_ms)
            sub.append(rst.submit_ms)
            done.append(rst.done_ms)
            if k in keep_idx:
                keep[k] = recs.tuples()
        extra = {}
        if timing:
            st = srv.stats()
            kt = {nm: srv.kernel_time(nm) for nm in ("extend", "extend_cont", "qexpand", "qcount", "dp_forward",
                                                      "dp_traceback")}
            extra = {"kernel_ms": {k: round(v[0], 1) for k, v in kt.items() if v[1]},
                     "dp_cells": st.dp_cells, "seed_hits": st.hits,
                     "roofline": dominant_roofline(kt, st, srv.xdrop_stats())}
        srv.close()
        span_s = (max(done) - min(sub)) / 1e3
        lat = np.array(lat)
        qbp = float(lens[list(idx)].sum())
        r = {"queries": len(idx), "query_bp": int(qbp), "throughput_mbps": round(qbp / 1e6 / span_s, 4),
             "span_s": round(span_s, 2), "p50_ms": round(float(np.percentile(lat, 50)), 1),
             "p99_ms": round(float(np.percentile(lat, 99)), 1), "max_ms": round(float(lat.max()), 1),
             "mean_ms": round(float(lat.mean()), 1), "records": nrec}
        r.update(extra)
        return r

    allq = list(range(len(queries)))
    sub_n = min(args.serve_plain_requests, len(queries))
    subq = allq[:sub_n]
    # oracle sample: queries from the 20th length percentile up, ~args.serve_cpu_bp bases
    order = np.argsort(lens)
    pick, tot = [], 0
    for k in order[len(order) // 5:]:
        if tot + lens[k] > args.serve_cpu_bp:
            break
        pick.append(int(k))
        tot += lens[k]
    runs = {}
    scap = run(True, allq, None, timing=True, keep_idx=set(pick))
    log(f"serve: served capacity {scap}")
    runs["served_closed_loop"] = scap
    rate_s = scap["throughput_mbps"] * 1e6 / lens.mean()  # queries per second
    for f in args.serve_loads:
        r = run(True, allq, f * rate_s)
        r["load_of_served_capacity"] = f
        r["qps"] = round(f * rate_s, 4)
        runs[f"served_load{f:g}_of_served"] = r
        log(f"serve: served at {f} of its capacity {r}")
    pcap = run(False, subq, None)
    log(f"serve: plain capacity (first {sub_n} queries) {pcap}")
    runs["plain_closed_loop"] = pcap
    rate_p = pcap["throughput_mbps"] * 1e6 / lens[subq].mean()
    for co in (False, True):
        r = run(co, subq, args.serve_plain_load * rate_p)
        r["load_of_plain_capacity"] = args.serve_plain_load
        r["qps"] = round(args.serve_plain_load * rate_p, 4)
        runs[f"{'served' if co else 'plain'}_load{args.serve_plain_load:g}_of_plain"] = r
        log(f"serve: {'served' if co else 'plain'} at {args.serve_plain_load} of plain capacity {r}")
    cpu = None
    if not args.no_cpu and pick:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_py as O
        nthreads = os.cpu_count() or 1
        names = [(f"q{k}", queries[k]) for k in pick]
        t1 = time.perf_counter()
        orec, _ = O.run([("t", target)], names, cfg, threads=nthreads)
        cpu_s = time.perf_counter() - t1
        per = {k: [] for k in pick}
        for x in orec.tuples():
            per[pick[x[1]]].append((x[0], 0) + tuple(x[2:]))
        same = all(keep.get(k) == per[k] for k in pick)
        cpu = {"value": round(tot / 1e6 / cpu_s, 4), "unit": "query-Mbp/s", "cores": nthreads, "kind": "port",
               "sample": f"oracle (lazy-row DP) on {len(pick)} queries ({int(tot)} bp, the 20th length percentile "
                         f"up) against the whole target, one run: {cpu_s:.2f} s (throughput only, no latency claim)",
               "gpu_output_identical_on_sample": bool(same),
               "gpu_output_checked": "each sampled query's records from the served capacity run"}
        log(f"serve: cpu {cpu}")
    best = runs.get(f"served_load{max(args.serve_loads):g}_of_served", scap)
    wl = {"workload": f"serving: {len(target)} bp uniform target, 40% soft-masked; {len(queries)} queries, "
                      f"length log-uniform 1 kbp-{args.serve_max_len} bp, 80% excerpts (identity 75-95%, both "
                      f"strands) / 20% decoys, full seed->HSP->gapped, Poisson arrivals; loads as fractions of the "
                      f"served capacity (all queries) and of the plain per-query driver's capacity (first "
                      f"{sub_n} queries)",
          "target_bp": len(target), "queries": len(queries), "query_bp": int(lens.sum()),
          "full_size": "the configuration's >= 1000 requests of log-uniform 1 kbp-5 Mbp (--serve-max-len 5000000)"}
    return {"config": wl, "value": best["throughput_mbps"], "unit": "query-Mbp/s",
            "p50_ms": best["p50_ms"], "p99_ms": best["p99_ms"],
            "capacity_mbps": {"served": scap["throughput_mbps"], "plain": pcap["throughput_mbps"]},
            "runs": runs, "roofline": scap.get("roofline"), "cpu_baseline": cpu}


def contigs_leg(args):
    """The many-contig workload; returns its measurements."""
    import numpy as np
    import gxy
    import gxy_types as T
    import synth

    t0 = time.time()
    target, contigs = synth.contig_workload(n_target=args.contig_target_bp, n_contigs=args.contigs,
                                            hi=args.contig_max_len)
    qbp = sum(len(q) for q in contigs)
    log(f"contigs: target {len(target)} bp, {len(contigs)} contigs, {qbp} bp (generated in {time.time() - t0:.1f}s)")
    ctx = gxy.default_context()
    dt = gxy.DeviceSeq.upload(target, ctx)
    dqs = [gxy.DeviceSeq.upload(q, ctx) for q in contigs]
    ctx.synchronize()
    cfg = T.config_default()
    cfg.chain_enabled = 1
    cfg.tweener_enabled = 1
    names_t = [("target", target)]
    names_q = [(f"contig{k}", q) for k, q in enumerate(contigs)]

    bctx = gxy.Context(0)
    bctx.reset_timing()
    bctx.set_timing(True)
    t1 = time.perf_counter()
    recs, st = gxy.run_dev([dt], dqs, cfg, bctx)
    bctx.synchronize()
    batched_s = time.perf_counter() - t1
    bctx.set_timing(False)
    t1 = time.perf_counter()
    paf = gxy.format_paf(names_t, names_q, recs)
    paf_s = time.perf_counter() - t1
    log(f"contigs: batched {batched_s:.2f}s, {len(recs)} records, PAF {len(paf)} bytes in {paf_s:.2f}s")
    kt = {nm: bctx.kernel_time(nm) for nm in ("extend", "extend_cont", "qexpand", "qcount", "dp_forward",
                                               "dp_traceback")}
    roof = dominant_roofline(kt, st, bctx.xdrop_stats())
    tune = bctx.dp_autotune_state()
    bctx.close()
    rec_by_q = {}
    for x in recs.tuples():
        rec_by_q.setdefault(x[1], []).append(x)

    srv = gxy.Server([dt], cfg)
    srv.set_timing(True)
    t1 = time.perf_counter()
    tickets = [srv.submit(q) for q in contigs]
    lat, nrec = [], 0
    step = max(1, len(tickets) // 10)
    for k, tk in enumerate(tickets):
        r, sst = srv.wait(tk)
        nrec += len(r)
        lat.append(sst.done_ms - sst.submit_ms)
        if (k + 1) % step == 0:
            log(f"contigs: served {k + 1}/{len(tickets)} answered, {time.perf_counter() - t1:.1f}s")
    served_s = time.perf_counter() - t1
    skt = {nm: srv.kernel_time(nm) for nm in ("extend", "dp_forward")}
    srv.close()
    lat = np.array(lat)
    log(f"contigs: served {served_s:.2f}s, {nrec} records")

    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_py as O
        # sample: contigs from the 20th length percentile up, up to ~contig_cpu_bp bases in all
        order = np.argsort([len(q) for q in contigs])
        pick, tot = [], 0
        for k in order[len(order) // 5:]:
            if tot + len(contigs[k]) > args.contig_cpu_bp:
                break
            pick.append(int(k))
            tot += len(contigs[k])
        sample = [names_q[k] for k in pick]
        nthreads = os.cpu_count() or 1
        # fixed cost: the target's table (a 1 kbp query adds almost no search work)
        t1 = time.perf_counter()
        O.run(names_t, [("tiny", contigs[int(order[0])].slice(0, min(1000, len(contigs[int(order[0])]))))], cfg,
              threads=nthreads)
        build_s = time.perf_counter() - t1
        t1 = time.perf_counter()
        orec, _ = O.run(names_t, sample, cfg, threads=nthreads)
        cpu_s = time.perf_counter() - t1
        oper = {}
        for x in orec.tuples():
            oper.setdefault(pick[x[1]], []).append((x[0], pick[x[1]]) + tuple(x[2:]))
        same = all(rec_by_q.get(k, []) == oper.get(k, []) for k in pick)
        per_bp = max(cpu_s - build_s, 1e-9) / tot
        full_s = build_s + per_bp * qbp
        cpu = {"value": round(qbp / 1e6 / full_s, 4), "unit": "query-Mbp/s", "cores": nthreads, "kind": "port",
               "sample": f"oracle (lazy-row DP, chain + tweener) on {len(pick)} contigs ({tot} bp, the 20th "
                         f"percentile of length upward) against the same target: {cpu_s:.2f} s, of which "
                         f"{build_s:.2f} s the table; value = all contigs extrapolated ({full_s:.1f} s)",
               "sample_query_mbps": round(tot / 1e6 / cpu_s, 4),
               "gpu_output_identical_on_sample": bool(same),
               "gpu_output_checked": "the sampled contigs' records from the batched run (coordinates, scores, "
                                     "edit scripts; the PAF writer is pinned separately)"}
        log(f"contigs: cpu {cpu}")
    wl = {"workload": f"many contigs: {len(target)} bp uniform target vs {len(contigs)} contigs, length log-uniform "
                      f"1 kbp-{args.contig_max_len} bp, identity 70-90%, ~3% of bases in indels; 12of19, --chain, "
                      f"--inner (match12), gapped, PAF",
          "target_bp": len(target), "contigs": len(contigs), "query_bp": qbp,
          "full_size": "the configuration's 10,000 contigs (--contigs 10000)"}
    return {"config": wl, "value": round(qbp / 1e6 / batched_s, 4), "unit": "query-Mbp/s",
            "ms_per_step": round(batched_s * 1e3, 1),
            "records": len(recs), "paf_bytes": len(paf), "paf_format_s": round(paf_s, 3),
            "dp_cells": st.dp_cells, "dp_gcups": round(st.dp_cells / 1e9 / batched_s, 3),
            "kernel_ms": {k: round(v[0], 1) for k, v in kt.items() if v[1]}, "dp_autotune": tune,
            "served": {"query_mbps": round(qbp / 1e6 / served_s, 4), "wall_s": round(served_s, 2),
                       "records": nrec, "p50_ms": round(float(np.percentile(lat, 50)), 1),
                       "p99_ms": round(float(np.percentile(lat, 99)), 1),
                       "kernel_ms": {k: round(v[0], 1) for k, v in skt.items() if v[1]}},
            "roofline": roof, "cpu_baseline": cpu}


def plain_dropin_leg(target, query, cfg, steps, recs):
    """The plain drop-in on the chromosome pair: the reference driver's loop
    (src/driver.rs:199-300) with the GPU behind its two calls.  Per target chunk
    (src/driver.rs:156-194), one host thread each as rayon runs them, each with
    its own context: PosTable::build on the chunk, then per strand search, each
    a synchronous library call that uploads its host inputs, launches its
    stages and downloads its HSPs; the primary-range filter and the shift to
    global coordinates on the host, then one sort.  The chunk slices and the
    query's reverse complement are made before the timed region (the
    reference recomputes the reverse complement per chunk), in the drop-in's
    favour.  Its HSPs must equal the batched run's records."""
    from concurrent.futures import ThreadPoolExecutor

    import numpy as np
    import gxy
    import gxy_types as T
    tlen = len(target)
    prim, halo = cfg.chunk_primary_bp, cfg.chunk_halo_bp
    bounds = []  # (lo, hi, primary start, primary end), chunk-local primary range
    if prim == 0 or tlen <= prim + halo:
        bounds.append((0, tlen, 0, tlen))
    else:
        ps = 0
        while ps < tlen:
            pe = min(ps + prim, tlen)
            lo, hi = max(ps - halo, 0), min(pe + halo, tlen)
            bounds.append((lo, hi, ps - lo, pe - lo))
            if pe == tlen:
                break
            ps = pe
    chunks = [target.slice(lo, hi) for lo, hi, _, _ in bounds]
    strands = [(0, query), (1, query.reverse_complement())]
    sp = T.SearchParams(cfg.step, cfg.hsp, cfg.transitions, cfg.entropy_enabled, cfg.entropy_threshold)
    ctxs = [gxy.Context(0) for _ in bounds]

    def one(ci):
        lo, _, pl, pe = bounds[ci]
        tab = gxy.PosTable.build(chunks[ci], cfg.pattern, cfg.step, ctxs[ci])
        out = []
        for strand, qs in strands:
            h, _ = gxy.search_array(tab, chunks[ci], qs, cfg.scoring, sp)
            h = h[(h["t_start"] >= pl) & (h["t_start"] < pe)]
            out.append((strand, lo, h))
        return out

    def step():
        with ThreadPoolExecutor(len(bounds)) as ex:
            parts = [x for r in ex.map(one, range(len(bounds))) for x in r]
        n = sum(len(h) for _, _, h in parts)
        rec = np.empty(n, [("strand", np.uint32), ("t_start", np.uint32), ("q_start", np.uint32),
                           ("length", np.uint32), ("score", np.int32)])
        o = 0
        for strand, lo, h in parts:
            k = len(h)
            rec["strand"][o:o + k] = strand
            rec["t_start"][o:o + k] = h["t_start"] + lo
            rec["q_start"][o:o + k] = h["q_start"]
            rec["length"][o:o + k] = h["length"]
            rec["score"][o:o + k] = h["score"]
            o += k
        rec.sort(order=["strand", "t_start", "q_start", "length", "score"])
        return rec

    step()  # warmup: contexts, scratch
    t0 = time.perf_counter()
    for _ in range(steps):
        rec = step()
    dt = (time.perf_counter() - t0) / steps
    g = np.array([(r[2], r[3], r[4], r[5], r[7]) for r in recs.tuples()],
                 dtype=[("strand", np.uint32), ("t_start", np.uint32), ("q_start", np.uint32),
                        ("length", np.uint32), ("score", np.int32)])
    g.sort(order=["strand", "t_start", "q_start", "length", "score"])
    same = len(g) == len(rec) and bool(np.array_equal(g, rec))
    del ctxs
    return {"query_mbps": round(len(query) / 1e6 / dt, 3), "ms_per_step": round(1000 * dt, 2),
            "steps": steps, "host_threads": len(bounds), "records": int(len(rec)),
            "records_identical_to_batched": same,
            "form": "reference driver loop (chunks on host threads, one context each) over host "
                    "sequences; per chunk one PosTable::build call and one search call per strand, "
                    "each uploading its inputs and downloading its HSPs; primary-range filter, "
                    "coordinate shift and sort on the host"}


def pair_leg(args, workload, steps, warmup):
    """The chromosome-scale HSP-only pair (workload "chrom") or the 30 kbp pair
    ("pair30k"): `steps` driver runs over the resident pair."""
    import numpy as np
    import gxy
    import gxy_types as T
    import synth

    t0 = time.time()
    if workload == "chrom":
        target, query = synth.chromosome_pair(n=args.target_bp, seed=7)
        cfg = T.config_default()
        cfg.gapped_enabled = 0
        cfg.chunk_primary_bp = 10_000_000
        cfg.chunk_halo_bp = 50_000
        wl = {"workload": "chromosome-scale HSP-only pair, 12of19, --nogapped, chunk 10 Mbp / "
                          "halo 50 kbp, both strands",
              "target_bp": len(target), "query_bp": len(query), "masked_frac": 0.5,
              "n_frac": 0.02, "sub": 0.15, "indel_events": 0.02, "segdup_frac": 0.05}
    else:
        target, query = synth.pair_30k(seed=1)
        cfg = T.config_default()
        wl = {"workload": "30 kbp pair, 20% substitutions, 1% indel events, gapped, MAF",
              "target_bp": len(target), "query_bp": len(query)}
    log(f"{workload}: generated {len(target)} + {len(query)} bp in {time.time() - t0:.1f}s")

    ctx = gxy.default_context()
    dt = gxy.DeviceSeq.upload(target, ctx)
    dq = gxy.DeviceSeq.upload(query, ctx)
    ctx.synchronize()

    for w in range(warmup):
        t1 = time.time()
        recs, st = gxy.run_dev([dt], [dq], cfg, ctx)
        log(f"warmup {w}: {time.time() - t1:.3f}s, {len(recs)} records")

    names = ["table_items", "table_sort", "table_starts", "table_bytes", "qcount", "qscan", "qexpand", "bin",
             "extend", "extend_cont", "pass_sort", "admit", "emit", "dp_forward", "dp_traceback"]
    ctx.reset_timing()
    ctx.set_timing(True)
    ctx.synchronize()
    xd0 = ctx.xdrop_stats()
    tstart = time.perf_counter()
    for s in range(steps):
        recs, st = gxy.run_dev([dt], [dq], cfg, ctx)
        log(f"step {s}: {len(recs)} records")
    ctx.synchronize()
    elapsed = time.perf_counter() - tstart
    ctx.set_timing(False)
    ktime = {n: ctx.kernel_time(n) for n in names}
    ms_per_step = 1000.0 * elapsed / steps
    qbp = len(query)
    value = qbp / 1e6 / (elapsed / steps)
    hsp_gcups = st.columns / 1e9 / (elapsed / steps)
    dp_gcups = st.dp_cells / 1e9 / (elapsed / steps)

    xds = tuple(b - a for a, b in zip(xd0, ctx.xdrop_stats()))
    dom_name, (dom_ms, dom_launches) = max(ktime.items(), key=lambda kv: kv[1][0])
    roof = None

    def prof_of(name):
        # the committed rocprofv3 summary of this workload: HBM bytes per launch
        # (FETCH_SIZE/WRITE_SIZE, scaled by a 1 GiB copy) and the kernel's
        # duration in the counter runs, where kernels run one at a time
        prof = os.path.join(ROOT, "profiles", name)
        if workload == "chrom" and os.path.exists(prof):
            return json.load(open(prof))
        return {}

    def with_serialized(r, prof):
        # the summary's per-launch figures hold only for launches of the same
        # size (the same hit batch): otherwise they are left out
        pb, lb = prof.get("algorithmic_bytes_per_launch"), r.get("algorithmic_bytes_per_launch")
        if not pb or not lb or abs(pb - lb) > 0.02 * pb:
            r["traffic"] = None
            if prof:
                r["profile_mismatch"] = f"{prof.get('source')} was taken on launches of {pb} algorithmic bytes"
            return r
        r["traffic"] = prof.get("traffic_bytes_per_launch")
        ser = prof.get("avg_ns_serialized")
        if ser:
            # the same bytes over the kernel's own duration (no other stream on the device)
            r["serialized"] = {"avg_launch_ms": round(ser / 1e6, 4),
                               "achieved": round(r["algorithmic_bytes_per_launch"] / (ser / 1e9) / 1e9, 2),
                               "source": prof.get("source")}
            r["serialized"]["frac"] = round(r["serialized"]["achieved"] / HBM_PEAK_GBS, 5)
        return r

    def xdrop_roof():
        xms, xn = ktime.get("extend", (0.0, 0))
        r = with_serialized(hbm_xdrop_roofline(st.hits * steps, xds, xms, xn),
                            prof_of("k_xdrop_first_chrom48M.json"))
        ems, en = ktime.get("extend_cont", (0.0, 0))
        if en:
            r["resume_kernel"] = {"kernel": "k_xdrop (resumed walks)", "launches": en,
                                  "avg_launch_ms": round(ems / en, 4)}
        return r

    if dom_name == "extend":
        roof = xdrop_roof()
    elif dom_name == "qexpand":
        roof = with_serialized(hbm_qscan_roofline(st.lookups * steps, st.hits * steps, dom_ms, dom_launches),
                               prof_of("k_qexpand_w_chrom48M.json"))
        roof["second_kernel"] = xdrop_roof()
    elif dom_name == "dp_forward":
        roof = valu_roofline(st.dp_cells * steps, dom_ms, dom_launches)
    else:
        roof = {"bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": None, "traffic": None, "kernel": dom_name,
                "avg_launch_ms": round(dom_ms / max(1, dom_launches), 4), "launches": dom_launches}

    # kernels of the driver's four streams overlap: their launch intervals
    # (the `avg_launch_ms` above) are shared with the other streams' kernels
    roof["stream_concurrency"] = round(sum(v[0] for v in ktime.values()) / max(1e-9, 1000.0 * elapsed), 2)
    if roof["bound"] == "hbm":
        # `peak` stays the published 8 TB/s; a measured streaming-copy rate rides beside it
        copy_gbs = measured_copy_gbs()
        roof["copy_peak_measured_gbs"] = copy_gbs
        if copy_gbs and roof.get("achieved"):
            roof["frac_of_measured_copy"] = round(roof["achieved"] / copy_gbs, 5)

    # PCIe-inclusive rate (host buffers handed over): reported, never `value`
    t1 = time.perf_counter()
    gxy.run([("t", target)], [("q", query)], cfg, ctx)
    pcie_s = time.perf_counter() - t1

    plain = None
    if workload == "chrom" and not args.no_plain:
        plain = plain_dropin_leg(target, query, cfg, 2, recs)
        plain["batched_over_plain"] = round(value / plain["query_mbps"], 2)
        plain["pcie_inclusive_over_plain"] = round(qbp / 1e6 / pcie_s / plain["query_mbps"], 2)
        log(f"plain drop-in: {plain['query_mbps']} query-Mbp/s ({plain['ms_per_step']} ms/step), "
            f"batched/plain {plain['batched_over_plain']}x, identical {plain['records_identical_to_batched']}")

    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_py as O
        nthreads = os.cpu_count() or 1
        if workload == "chrom":
            sample = query.slice(0, min(args.cpu_sample_bp, len(query)))
        else:
            sample = query
        # fixed cost: the chunk tables (a 1 kbp query barely adds search work);
        # the 30 kbp pair runs the reference's as-written DP storage
        lazy = workload != "pair30k"
        t1 = time.perf_counter()
        O.run([("t", target)], [("q", query.slice(0, 1000))], cfg, threads=nthreads, lazy=lazy)
        build_s = time.perf_counter() - t1
        t1 = time.perf_counter()
        orec, ost = O.run([("t", target)], [("q", sample)], cfg, threads=nthreads, lazy=lazy)
        cpu_s = time.perf_counter() - t1
        # the whole query: the last timed run's records; a prefix: a run of its own
        grec = recs if len(sample) == len(query) else gxy.run_dev([dt], [gxy.DeviceSeq.upload(sample, ctx)],
                                                                   cfg, ctx)[0]
        same = (gxy.format_maf([("t", target)], [("q", sample)], grec)
                == O.format_maf([("t", target)], [("q", sample)], orec))
        per_bp = max(cpu_s - build_s, 1e-9) / len(sample)
        full_s = build_s + per_bp * len(query)
        cpu = {"value": round(len(query) / 1e6 / full_s, 4), "unit": "query-Mbp/s",
               "cores": nthreads, "kind": "port",
               "sample": f"oracle ({'lazy-row' if lazy else 'as-written'} DP storage) on the same target "
                         f"({len(target)} bp, all chunks, both strands) "
                         f"with the first {len(sample)} bp of the query: {cpu_s:.2f} s wall, of which "
                         f"{build_s:.2f} s table builds; value = whole query extrapolated "
                         f"({full_s:.1f} s)",
               "sample_query_mbps": round(len(sample) / 1e6 / cpu_s, 4),
               "gpu_output_identical_on_sample": bool(same)}
        log(f"cpu baseline {cpu_s:.2f}s on {len(sample)} bp, GPU output identical: {same}")

    out = {
        "value": round(value, 3), "unit": "query-Mbp/s", "steps": steps, "warmup": warmup,
        "ms_per_step": round(ms_per_step, 2), "config": wl,
        "hsp_gcups": round(hsp_gcups, 3), "dp_gcups": round(dp_gcups, 3),
        "records": len(recs), "seed_hits_per_step": st.hits, "xdrop_columns_per_step": st.columns,
        "pcie_inclusive_query_mbps": round(qbp / 1e6 / pcie_s, 3),
        "kernel_ms_per_step": {n: round(v[0] / steps, 3) for n, v in ktime.items() if v[1]},
        "timed_kernels_ms_per_step": round(sum(v[0] for v in ktime.values()) / steps, 3),
        "roofline": roof, "cpu_baseline": cpu,
    }
    if plain is not None:
        out["plain_dropin"] = plain
    if workload == "pair30k":
        out["dp_autotune"] = ctx.dp_autotune_state()
        out["dp_cells_per_step"] = st.dp_cells
        out["dp_pool_passes_truncated_alone"] = ctx.dp_pool_stats()
    return out


LEG_ESTIMATE_S = {"pair30k": 20, "dp": 150, "serve": 240, "contigs": 120}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="all", choices=["all", "chrom", "pair30k", "dp", "serve", "contigs"])
    ap.add_argument("--budget-s", type=float, default=540.0,
                    help="wall-clock budget of an `all` run: a leg whose estimate would overrun it is skipped")
    ap.add_argument("--contig-target-bp", type=int, default=200_000_000)
    ap.add_argument("--contigs", type=int, default=2000)
    ap.add_argument("--contig-max-len", type=int, default=1_000_000)
    ap.add_argument("--contig-cpu-bp", type=int, default=1_500_000)
    ap.add_argument("--serve-target-bp", type=int, default=100_000_000)
    ap.add_argument("--serve-requests", type=int, default=1000)
    ap.add_argument("--serve-max-len", type=int, default=1_000_000)
    ap.add_argument("--serve-loads", type=float, nargs="+", default=[0.5, 0.8])
    ap.add_argument("--serve-plain-requests", type=int, default=100)
    ap.add_argument("--serve-plain-load", type=float, default=0.8)
    ap.add_argument("--serve-cpu-bp", type=int, default=1_000_000)
    ap.add_argument("--dp-anchors", type=int, default=1_000_000)
    ap.add_argument("--dp-cpu-anchors", type=int, default=400)
    ap.add_argument("--dp-kernel", type=int, default=0,
                    help="DP forward kernel (gxy_ctx_set_dp_kernel: 0 autotuned, 1 four wavefronts per side, "
                         "2/3/4 one wavefront per side with 1/2/4 columns per lane per slot, 5/6 = 3/4 at four "
                         "wavefronts per SIMD)")
    ap.add_argument("--xdrop-form", type=int, default=0,
                    help="x-drop form of the search (gxy_ctx_set_xdrop_form: 0 first step + resumed walks, "
                         "1 one persistent kernel)")
    ap.add_argument("--pair-steps", type=int, default=5)
    ap.add_argument("--target-bp", type=int, default=48_000_000)
    ap.add_argument("--cpu-sample-bp", type=int, default=16_000_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="omit the plain drop-in leg of the chrom line")
    ap.add_argument("--legs", nargs="+", default=["pair30k", "dp", "serve", "contigs"],
                    help="sub-workloads of an `all` run")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(this library drives one device per process)"}))
        sys.exit(2)
    t_start = time.time()

    import gxy
    gxy.default_context().set_dp_kernel(args.dp_kernel)
    gxy.default_context().set_xdrop_form(args.xdrop_form)
    head = {"metric": METRIC, "n_gpus": 1, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "int32", "data": "synthetic"}

    def run_leg(name, steps, warmup):
        if name == "pair30k":
            return pair_leg(args, "pair30k", steps, warmup)
        if name == "dp":
            return dp_leg(args.dp_anchors, steps, warmup, 0 if args.no_cpu else args.dp_cpu_anchors,
                          kernel=args.dp_kernel)
        if name == "serve":
            return serve_leg(args)
        return contigs_leg(args)

    if args.workload != "all":
        if args.workload == "chrom":
            d = pair_leg(args, "chrom", args.steps, args.warmup)
        else:
            d = run_leg(args.workload, args.steps, args.warmup)
        out = dict(head)
        out.update({"steps": args.steps, "warmup": args.warmup})
        out.update(d)
        if args.workload == "dp":
            out["value"], out["unit"] = d["dp_gcups"], "GCUPS"
            out["config"] = {"workload": d["workload"], "anchors": d["anchors"]}
        print(json.dumps(out), flush=True)
        return

    d = pair_leg(args, "chrom", args.steps, args.warmup)
    out = dict(head)
    out.update(d)
    for name in args.legs:
        left = args.budget_s - (time.time() - t_start)
        if left < LEG_ESTIMATE_S.get(name, 60):
            out[name] = {"skipped": f"time budget: {left:.0f} s left of {args.budget_s:.0f} s, the leg needs "
                                    f"about {LEG_ESTIMATE_S.get(name, 60)} s"}
            log(f"{name}: skipped ({left:.0f} s left)")
            continue
        # each configuration starts on a device without the scratch memory the
        # previous one left cached in the default context (its lanes' search
        # buffers, its traceback pool): gxy_ctx_release_cache
        import gxy
        gxy.default_context().release_cache()
        t1 = time.time()
        key = {"dp": "batched_dp"}.get(name, name)
        if name == "pair30k":
            out[key] = run_leg(name, args.pair_steps, 2)
        elif name == "dp":
            out[key] = run_leg(name, 1, 1)
        else:
            out[key] = run_leg(name, 1, 0)
        out[key]["wall_s"] = round(time.time() - t1, 1)
        log(f"{name}: done in {time.time() - t1:.1f}s")
    out["bench_wall_s"] = round(time.time() - t_start, 1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
