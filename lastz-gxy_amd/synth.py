"""Deterministic synthetic workloads (numpy, seeded).

* `lcg_dna` — the reference bench generator (benches/hotpath.rs:22-30):
  x <- x*6364136223846793005 + 1442695040888963407 (mod 2^64), base
  ACGT[(x >> 33) & 3].
* `mutate` — per source base: with probability `indel` an indel event
  (insertion or deletion with equal odds, length Geometric(1/3), mean 3,
  inserted bases uniform uppercase); otherwise a substitution with probability
  `sub` to one of the other three bases.  Case and N carry over from source
  bases.
* `pair_30k` — 30 kbp LCG target, query = target at 20 % substitutions and 1 %
  indel events (the small gapped pair).
* `chromosome_pair` — uniform target with 50 % soft-masked blocks
  (LogNormal median 300, sigma 1) and 2 % N-runs (LogNormal median 1 kbp,
  sigma 1); query = target plus segmental duplications (copies of target
  segments, LogNormal median 5 kbp, half reverse-complemented, inserted at
  random points until the query is 1.05x), then 15 % substitutions and 2 %
  indel events.  The 48 Mbp default is the chromosome-scale HSP-only workload.
* `dp_anchors` — the batched gapped-DP workload (anchored pairs of log-normal
  extension length at 70-95 % identity).
"""
from __future__ import annotations

import numpy as np

from gxy_types import PackedSeq

_M64 = (1 << 64) - 1


def lcg_dna(n: int, seed: int) -> bytes:
    """synthetic_dna (benches/hotpath.rs:22-30); vectorised by jumping the LCG."""
    a, c = 6364136223846793005, 1442695040888963407
    out = np.empty(n, np.uint8)
    # advance in blocks of 4096 lanes: x_{k+B} = A_B x_k + C_B
    B = min(4096, max(1, n))
    xs = np.empty(B, np.uint64)
    x = seed & _M64
    for k in range(B):
        x = (x * a + c) & _M64
        xs[k] = x
    AB, CB = 1, 0
    for _ in range(B):
        AB, CB = (AB * a) & _M64, (CB * a + c) & _M64
    AB, CB = np.uint64(AB), np.uint64(CB)
    acgt = np.frombuffer(b"ACGT", np.uint8)
    with np.errstate(over="ignore"):
        for s in range(0, n, B):
            m = min(B, n - s)
            out[s:s + m] = acgt[((xs[:m] >> np.uint64(33)) & np.uint64(3)).astype(np.int64)]
            xs = xs * AB + CB
    return out.tobytes()


def mutate(code, valid, mask, rng, sub: float, indel: float, with_map: bool = False):
    """Return mutated (code, valid, mask) arrays (see module docstring); with
    `with_map`, also (kept, qpos): whether each source base survives and its
    output position."""
    n = len(code)
    u = rng.random(n)
    is_indel = u < indel
    is_sub = (~is_indel) & (rng.random(n) < sub)
    new = code.copy()
    off = rng.integers(1, 4, n).astype(np.uint8)
    new[is_sub] = (code[is_sub] + off[is_sub]) % 4
    is_ins = is_indel & (rng.random(n) < 0.5)
    is_del = is_indel & ~is_ins
    L = rng.geometric(1.0 / 3.0, n)
    cov = np.zeros(n + 1, np.int64)
    di = np.nonzero(is_del)[0]
    np.add.at(cov, di, 1)
    np.add.at(cov, np.minimum(di + L[di], n), -1)
    keep = np.cumsum(cov[:n]) == 0
    ins = np.where(is_ins, L, 0)
    counts = keep.astype(np.int64) + ins
    total = int(counts.sum())
    idx = np.repeat(np.arange(n), counts)
    starts = np.cumsum(counts) - counts
    rank = np.arange(total) - starts[idx]
    is_src = keep[idx] & (rank == 0)
    rnd = rng.integers(0, 4, total).astype(np.uint8)
    oc = np.where(is_src, new[idx], rnd).astype(np.uint8)
    ov = np.where(is_src, valid[idx], True)
    om = np.where(is_src, mask[idx], False)
    if with_map:
        return oc, ov, om, keep, starts
    return oc, ov, om


def _mutate_segments(code, lens, sub, indel, rng):
    """Mutate concatenated independent segments (segment r: lens[r] bases,
    substitution rate sub[r], indel event rate indel[r]) with the model of
    `mutate` (all bases valid and uppercase), drawing the random numbers only
    where events happen.  Returns (out_code, out_seg_start, opos) with opos(p)
    giving the output position of source base p and keep(p) whether it
    survives, as a pair of arrays over all source bases."""
    n = len(code)
    off = np.zeros(len(lens) + 1, np.int64)
    np.cumsum(lens, out=off[1:])
    u = rng.random(n, dtype=np.float32)
    thr_indel = np.repeat(indel.astype(np.float32), lens)
    thr_sub = np.repeat((indel + (1 - indel) * sub).astype(np.float32), lens)
    is_indel = u < thr_indel
    del thr_indel
    is_sub = (~is_indel) & (u < thr_sub)
    del thr_sub, u
    new = code.copy()
    sp = np.flatnonzero(is_sub)
    new[sp] = (code[sp] + rng.integers(1, 4, len(sp)).astype(np.uint8)) % 4
    ip = np.flatnonzero(is_indel)
    L = rng.geometric(1.0 / 3.0, len(ip))
    is_ins = rng.random(len(ip)) < 0.5
    seg_of = np.searchsorted(off, ip, side="right") - 1
    dpos, dend = ip[~is_ins], np.minimum(ip[~is_ins] + L[~is_ins], off[seg_of[~is_ins] + 1])
    cov = np.bincount(dpos, minlength=n + 1)[:n + 1] - np.bincount(dend, minlength=n + 1)[:n + 1]
    keep = np.cumsum(cov[:n]) == 0
    del cov
    ck = np.cumsum(keep)  # kept bases up to and including p
    ipos, ilen = ip[is_ins], L[is_ins]
    kept = new[keep]
    out = np.insert(kept, np.repeat(ck[ipos], ilen), rng.integers(0, 4, int(ilen.sum())).astype(np.uint8))
    icum = np.concatenate([[0], np.cumsum(ilen)])

    def opos(p):
        p = np.asarray(p)
        return (ck[p] - keep[p]) + icum[np.searchsorted(ipos, p, side="left")]

    seg_start = np.append(opos(off[:-1]), len(out))
    return out, seg_start, opos, keep


def dp_anchor_arrays(n: int, seed: int = 3, median: float = 2000.0, sigma: float = 1.0, lo: int = 100,
                     hi: int = 50_000, packed: bool = True):
    """Batched gapped-DP workload: n anchored pairs.  Per anchor, extension
    length L ~ LogNormal(median, sigma) clipped to [lo, hi] and identity
    I ~ U[0.70, 0.95]; target = 2L+1 uniform ACGT; query = target mutated at
    0.9(1-I) substitutions and 0.1(1-I)/3 indel events (Geometric mean 3);
    anchor = the target's centre base and its query position (the mutation is
    re-drawn while the centre is deleted).

    Returns (target, query, t_off, q_off, t_pos, q_pos): the targets and the
    queries concatenated into two PackedSeqs, pair k being
    target[t_off[k]:t_off[k+1]] and query[q_off[k]:q_off[k+1]], anchored at
    (t_pos[k], q_pos[k]) relative to those slices.  With packed=False the
    two sequences are returned as 2-bit code arrays (all bases valid and
    uppercase)."""
    rng = np.random.default_rng(seed)
    Ls = np.clip(np.round(median * np.exp(sigma * rng.standard_normal(n))), lo, hi).astype(np.int64)
    ids = rng.uniform(0.70, 0.95, n)
    tlen = 2 * Ls + 1
    t_off = np.zeros(n + 1, np.int64)
    np.cumsum(tlen, out=t_off[1:])
    code = rng.integers(0, 4, int(t_off[-1])).astype(np.uint8)
    qparts = [None] * n
    qpos = np.zeros(n, np.int64)
    for c0 in range(0, n, 4096):  # chunks of pairs keep the temporaries small
        todo = np.arange(c0, min(n, c0 + 4096))
        while len(todo):
            lens = tlen[todo]
            src = np.concatenate([code[t_off[k]:t_off[k + 1]] for k in todo])
            out, qbeg, opos, keep = _mutate_segments(src, lens, 0.9 * (1 - ids[todo]), 0.1 * (1 - ids[todo]) / 3,
                                                     rng)
            soff = np.concatenate([[0], np.cumsum(lens)])
            centre = soff[:-1] + Ls[todo]
            ok = keep[centre]
            cq = opos(centre)
            for r in np.flatnonzero(ok):
                k = todo[r]
                qparts[k] = out[qbeg[r]:qbeg[r + 1]]
                qpos[k] = cq[r] - qbeg[r]
            todo = todo[~ok]
    q_off = np.zeros(n + 1, np.int64)
    np.cumsum([len(x) for x in qparts], out=q_off[1:])
    qcode = np.concatenate(qparts)
    if not packed:
        return code, qcode, t_off, q_off, Ls.copy(), qpos
    return _packed(code), _packed(qcode), t_off, q_off, Ls.copy(), qpos


def _packed(code):
    return PackedSeq.from_arrays(code, np.ones(len(code), bool), np.zeros(len(code), bool))


def dp_anchors(n: int, seed: int = 3, **kw):
    """dp_anchor_arrays as a list of per-pair (target, query, t_pos, q_pos)."""
    t, q, toff, qoff, tp, qp = dp_anchor_arrays(n, seed, packed=False, **kw)
    return [(_packed(t[toff[k]:toff[k + 1]]), _packed(q[qoff[k]:qoff[k + 1]]), int(tp[k]), int(qp[k]))
            for k in range(n)]


def pair_30k(seed: int = 1, n: int = 30000):
    """Small gapped pair: (target, query) PackedSeqs."""
    t = PackedSeq.from_ascii(lcg_dna(n, seed))
    rng = np.random.default_rng(seed)
    c, v, m = t.unpack()
    qc, qv, qm = mutate(c, v, m, rng, 0.20, 0.01)
    return t, PackedSeq.from_arrays(qc, qv, qm)


def _lognormal_runs(rng, median: float, sigma: float, total: int):
    lens = []
    s = 0
    while s < total:
        x = np.maximum(1, rng.lognormal(np.log(median), sigma, 4096).astype(np.int64))
        lens.append(x)
        s += int(x.sum())
    return np.concatenate(lens)


def chromosome_pair(n: int = 48_000_000, seed: int = 7, masked: float = 0.5, nfrac: float = 0.02,
                    dup_frac: float = 0.05, sub: float = 0.15, indel: float = 0.02):
    """Chromosome-scale pair (target, query) PackedSeqs (see module docstring)."""
    rng = np.random.default_rng(seed)
    code = rng.integers(0, 4, n).astype(np.uint8)
    valid = np.ones(n, bool)
    mask = np.zeros(n, bool)
    if masked > 0:
        # alternate unmasked / masked blocks; masked share set by scaling
        runs = _lognormal_runs(rng, 300.0, 1.0, 2 * n)
        scale_m = masked / (1 - masked)
        ends = np.cumsum(np.where(np.arange(len(runs)) % 2 == 1,
                                  np.maximum(1, (runs * scale_m).astype(np.int64)), runs))
        ends = ends[ends < n]
        flags = np.zeros(n + 1, np.int8)
        flags[ends] = 1
        state = np.cumsum(flags[:n]) % 2 == 1
        mask = state
    if nfrac > 0:
        mean_len = 1000.0 * np.exp(0.5)
        k = max(1, int(nfrac * n / mean_len))
        starts = rng.integers(0, n, k)
        lens = np.maximum(1, rng.lognormal(np.log(1000.0), 1.0, k).astype(np.int64))
        cov = np.zeros(n + 1, np.int64)
        np.add.at(cov, starts, 1)
        np.add.at(cov, np.minimum(starts + lens, n), -1)
        isn = np.cumsum(cov[:n]) > 0
        valid = ~isn
        code = np.where(valid, code, 0).astype(np.uint8)
        mask = mask & valid
    target = PackedSeq.from_arrays(code, valid, mask)
    # segmental duplications
    pieces_c, pieces_v, pieces_m = [], [], []
    need = int(dup_frac * n)
    dups = []
    got = 0
    while got < need:
        L = int(max(100, min(n // 2, rng.lognormal(np.log(5000.0), 1.0))))
        s = int(rng.integers(0, n - L))
        rc = bool(rng.random() < 0.5)
        at = int(rng.integers(0, n))
        dups.append((at, s, L, rc))
        got += L
    dups.sort()
    prev = 0
    for at, s, L, rc in dups:
        pieces_c.append(code[prev:at])
        pieces_v.append(valid[prev:at])
        pieces_m.append(mask[prev:at])
        c, v, m = code[s:s + L], valid[s:s + L], mask[s:s + L]
        if rc:
            c, v, m = np.where(v, 3 - c, 0)[::-1].astype(np.uint8), v[::-1], m[::-1]
        pieces_c.append(c)
        pieces_v.append(v)
        pieces_m.append(m)
        prev = at
    pieces_c.append(code[prev:])
    pieces_v.append(valid[prev:])
    pieces_m.append(mask[prev:])
    qc0 = np.concatenate(pieces_c)
    qv0 = np.concatenate(pieces_v)
    qm0 = np.concatenate(pieces_m)
    qc, qv, qm = mutate(qc0, qv0, qm0, rng, sub, indel)
    return target, PackedSeq.from_arrays(qc, qv, qm)


def _masked_blocks(rng, n: int, masked: float):
    """Soft-mask flags: alternating unmasked / masked blocks with lengths
    LogNormal(median 300, sigma 1), masked blocks scaled to the share `masked`."""
    runs = _lognormal_runs(rng, 300.0, 1.0, 2 * n)
    scale_m = masked / (1 - masked)
    ends = np.cumsum(np.where(np.arange(len(runs)) % 2 == 1,
                              np.maximum(1, (runs * scale_m).astype(np.int64)), runs))
    ends = ends[ends < n]
    flags = np.zeros(n + 1, np.int8)
    flags[ends] = 1
    return np.cumsum(flags[:n]) % 2 == 1


def serve_workload(n_target: int = 100_000_000, n_req: int = 1000, seed: int = 11, masked: float = 0.4,
                   lo: int = 1000, hi: int = 5_000_000, decoy: float = 0.2, rate: float = 1.0):
    """Serving workload: a uniform target with `masked` of its bases in
    soft-masked blocks, and n_req queries with lengths log-uniform in [lo, hi]:
    a share `decoy` of them uniform random, the rest excerpts of the target at
    a random start on a random strand, mutated at identity I ~ U[0.75, 0.95]
    (substitutions 0.9(1-I), indel events 0.1(1-I)/3, Geometric mean 3).
    Arrival times are a Poisson process of `rate` queries per time unit.

    Returns (target PackedSeq, [(query PackedSeq, kind, arrival)]), kind being
    "excerpt+", "excerpt-" or "decoy"."""
    rng = np.random.default_rng(seed)
    code = rng.integers(0, 4, n_target).astype(np.uint8)
    mask = _masked_blocks(rng, n_target, masked) if masked > 0 else np.zeros(n_target, bool)
    target = PackedSeq.from_arrays(code, np.ones(n_target, bool), mask)
    lens = np.exp(rng.uniform(np.log(lo), np.log(hi), n_req)).astype(np.int64)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n_req))
    reqs = []
    for k in range(n_req):
        L = int(min(lens[k], n_target))
        if rng.random() < decoy:
            q = rng.integers(0, 4, L).astype(np.uint8)
            kind = "decoy"
        else:
            st = int(rng.integers(0, n_target - L + 1))
            src = code[st:st + L]
            minus = bool(rng.random() < 0.5)
            if minus:
                src = (3 - src[::-1]).astype(np.uint8)
            ident = rng.uniform(0.75, 0.95)
            q, _, _, _ = _mutate_segments(src, np.array([L]), np.array([0.9 * (1 - ident)]),
                                          np.array([0.1 * (1 - ident) / 3]), rng)
            kind = "excerpt-" if minus else "excerpt+"
        reqs.append((_packed(q), kind, float(arrivals[k])))
    return target, reqs


def contig_workload(n_target: int = 200_000_000, n_contigs: int = 10_000, seed: int = 13, lo: int = 1000,
                    hi: int = 1_000_000):
    """Many-contig workload: a uniform unmasked target and n_contigs queries of
    length log-uniform in [lo, hi], each an excerpt of the target at a random
    start on a random strand, mutated at identity I ~ U[0.70, 0.90] with
    indel events at 0.01 per base (Geometric mean 3: ~3 % of bases) and
    substitutions at 1 - I.  Returns (target, [query PackedSeq])."""
    rng = np.random.default_rng(seed)
    code = rng.integers(0, 4, n_target).astype(np.uint8)
    target = _packed(code)
    lens = np.exp(rng.uniform(np.log(lo), np.log(hi), n_contigs)).astype(np.int64)
    out = []
    for k in range(n_contigs):
        L = int(min(lens[k], n_target))
        st = int(rng.integers(0, n_target - L + 1))
        src = code[st:st + L]
        if rng.random() < 0.5:
            src = (3 - src[::-1]).astype(np.uint8)
        ident = rng.uniform(0.70, 0.90)
        q, _, _, _ = _mutate_segments(src, np.array([L]), np.array([1 - ident]), np.array([0.01]), rng)
        out.append(_packed(q))
    return target, out



def _pack2(code: np.ndarray) -> np.ndarray:
    """2-bit codes (length a multiple of 4) into PackedSeq code bytes."""
    c4 = code.reshape(-1, 4)
    return c4[:, 0] | (c4[:, 1] << 2) | (c4[:, 2] << 4) | (c4[:, 3] << 6)


def _dp_slice_worker(args):
    """Pairs [a, b) of dp_anchor_slices: generate them (dp_anchor_arrays'
    model, its own seed) and write their packed codes at the worker's offsets
    of the shared target and query code buffers.  Returns the query lengths
    and anchor query positions."""
    from multiprocessing import shared_memory
    a, b, seed, tname, qname, tbase, qbase, qcap = args
    t, q, toff, qoff, tpos, qpos = dp_anchor_arrays(b - a, seed=(seed, a), packed=False)
    tsh = shared_memory.SharedMemory(name=tname)
    qsh = shared_memory.SharedMemory(name=qname)
    try:
        tb = np.ndarray((tsh.size,), np.uint8, tsh.buf)
        qb = np.ndarray((qsh.size,), np.uint8, qsh.buf)
        if len(q) > qcap:
            raise RuntimeError("query segment bound exceeded")
        pad = lambda x: np.concatenate([x, np.zeros((-len(x)) % 4, np.uint8)])
        pt, pq = _pack2(pad(t)), _pack2(pad(q))
        tb[tbase // 4: tbase // 4 + len(pt)] = pt
        qb[qbase // 4: qbase // 4 + len(pq)] = pq
        del tb, qb
    finally:
        tsh.close()
        qsh.close()
    return np.diff(toff), np.diff(qoff), tpos, qpos


def dp_anchor_slices(n: int, seed: int = 3, workers: int = 16):
    """dp_anchor_arrays' workload for many pairs, generated by `workers`
    processes: pairs are split in contiguous ranges, range r drawn with seed
    (seed, first pair of r).  Returns (target, query, t_off, t_len, q_off,
    q_len, t_pos, q_pos): two PackedSeqs holding the pairs (each range's
    pairs end to end, ranges 32-base aligned, queries in segments sized for
    their range with room to spare) and each pair's slices and anchor."""
    from concurrent.futures import ProcessPoolExecutor
    from multiprocessing import get_context, shared_memory
    workers = max(1, min(workers, n))
    bounds = [n * w // workers for w in range(workers + 1)]
    # each range's target length is fixed by its draw of extension lengths:
    # draw them first (dp_anchor_arrays draws them first from its seed)
    tlens = []
    for w in range(workers):
        a, b = bounds[w], bounds[w + 1]
        rng = np.random.default_rng((seed, a))
        Ls = np.clip(np.round(2000.0 * np.exp(rng.standard_normal(b - a))), 100, 50_000).astype(np.int64)
        tlens.append(int((2 * Ls + 1).sum()))
    al = lambda x: (x + 31) // 32 * 32
    tbase = np.concatenate([[0], np.cumsum([al(x) for x in tlens])])
    qcap = [al(int(x * 1.1) + 4096) for x in tlens]
    qbase = np.concatenate([[0], np.cumsum(qcap)])
    tsh = shared_memory.SharedMemory(create=True, size=int(tbase[-1]) // 4)
    qsh = shared_memory.SharedMemory(create=True, size=int(qbase[-1]) // 4)
    try:
        jobs = [(bounds[w], bounds[w + 1], seed, tsh.name, qsh.name, int(tbase[w]), int(qbase[w]), qcap[w])
                for w in range(workers)]
        with ProcessPoolExecutor(workers, mp_context=get_context("fork")) as ex:
            parts = list(ex.map(_dp_slice_worker, jobs))
        tcodes = np.ndarray((tsh.size,), np.uint8, tsh.buf).copy()
        qcodes = np.ndarray((qsh.size,), np.uint8, qsh.buf).copy()
    finally:
        tsh.close()
        tsh.unlink()
        qsh.close()
        qsh.unlink()
    t_len = np.concatenate([p[0] for p in parts])
    q_len = np.concatenate([p[1] for p in parts])
    t_off = np.concatenate([tbase[w] + np.concatenate([[0], np.cumsum(parts[w][0])[:-1]])
                            for w in range(workers)])
    q_off = np.concatenate([qbase[w] + np.concatenate([[0], np.cumsum(parts[w][1])[:-1]])
                            for w in range(workers)])
    t_pos = np.concatenate([p[2] for p in parts])
    q_pos = np.concatenate([p[3] for p in parts])
    ones = lambda m: np.full((m + 7) // 8, 0xFF, np.uint8)
    T = PackedSeq(tcodes, ones(int(tbase[-1])), np.zeros((int(tbase[-1]) + 7) // 8, np.uint8), int(tbase[-1]))
    Q = PackedSeq(qcodes, ones(int(qbase[-1])), np.zeros((int(qbase[-1]) + 7) // 8, np.uint8), int(qbase[-1]))
    return T, Q, t_off, t_len, q_off, q_len, t_pos, q_pos


def slice_of(seq: PackedSeq, off: int, length: int) -> PackedSeq:
    """seq[off, off + length) as a PackedSeq, unpacking only that range."""
    off, length = int(off), int(length)
    a, b = off // 32 * 32, min(seq.len, (off + length + 31) // 32 * 32)
    part = PackedSeq(seq.codes[a // 4:(b + 3) // 4], seq.valid[a // 8:(b + 7) // 8], seq.mask[a // 8:(b + 7) // 8],
                     b - a)
    return part.slice(off - a, off - a + length)
