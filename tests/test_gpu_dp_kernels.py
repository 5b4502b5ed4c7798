"""Every DP forward kernel (gxy_ctx_set_dp_kernel) against the oracle: the
four-wavefront kernel with the row in LDS and the one-wavefront kernels with
the row in registers (1, 2 and 4 columns per lane per slot).  Alignments
(coordinates, score, edit script) and DP cell counts must be identical.  The
cases cover short and long sides, windows that move right across many slots
and jump left (long insertions and deletions), N bases, three y-drops, a
substitution matrix outside int8 (the int32 score path) and an asymmetric one."""
import sys

import numpy as np
import pytest

import gxy
import gxy_types as T
import oracle_py as O
import synth

pytestmark = pytest.mark.gpu

KERNELS = {"block4": 1, "wave_k1": 2, "wave_k2": 3, "wave_k4": 4, "wave_k2_w4": 5, "wave_k4_w4": 6}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def pairs(rng, lens, indel_blocks=False, pn=0.01):
    out = []
    for k, L in enumerate(lens):
        code = rng.integers(0, 4, 2 * L + 1)
        valid = rng.random(2 * L + 1) >= pn
        t = T.PackedSeq.from_arrays(code, valid, np.zeros(2 * L + 1, bool))
        c, v, m = t.unpack()
        qc, qv, qm = synth.mutate(c, v, m, rng, float(rng.uniform(0.02, 0.3)), float(rng.uniform(0.0, 0.03)))
        if indel_blocks and len(qc) > 600:
            p = int(rng.integers(len(qc) // 2 + 1, len(qc) - 250))
            if k % 2 == 0:  # a long insertion in the query: the window jumps right
                ins = rng.integers(0, 4, int(rng.integers(60, 250)))
                qc = np.concatenate([qc[:p], ins, qc[p:]])
                qv = np.concatenate([qv[:p], np.ones(len(ins), bool), qv[p:]])
                qm = np.concatenate([qm[:p], np.zeros(len(ins), bool), qm[p:]])
            else:  # a long deletion: the band grows left
                d = int(rng.integers(60, 240))
                qc, qv, qm = (np.concatenate([a[:p], a[p + d:]]) for a in (qc, qv, qm))
        q = T.PackedSeq.from_arrays(qc, qv, qm)
        out.append((t, q, (L, min(L, len(q) - 1))))
    return out


def run_all(ctx, cases, sc, gp):
    ts = [gxy.DeviceSeq.upload(t, ctx) for t, _, _ in cases]
    qs = [gxy.DeviceSeq.upload(q, ctx) for _, q, _ in cases]
    anc = [a for _, _, a in cases]
    return gxy.gapped_extend_pairs_dev(ts, qs, anc, sc, gp, ctx)


def oracle_all(cases, sc, gp):
    res, cells = [], 0
    for t, q, a in cases:
        r, c = O.gapped_extend(a, t, q, sc, gp)
        res.append(r)
        cells += c
    return res, cells


@pytest.fixture(scope="module")
def ctx():
    return gxy.Context(0)


@pytest.mark.parametrize("kernel", sorted(KERNELS))
def test_dp_kernel_random_pairs(ctx, kernel):
    ctx.set_dp_kernel(KERNELS[kernel])
    rng = np.random.default_rng(101)
    hox = T.scoring_hoxd70()
    cases = pairs(rng, [int(x) for x in rng.integers(10, 1500, 60)])
    for ydrop in (9400, 3000, 800):
        gp = T.GappedParams(ydrop, 1000)
        say(kernel, "random pairs ydrop", ydrop)
        got, cells = run_all(ctx, cases, hox, gp)
        exp, ecells = oracle_all(cases, hox, gp)
        assert cells == ecells
        for k in range(len(cases)):
            assert got[k] == exp[k], (kernel, ydrop, k)
    ctx.set_dp_kernel(0)


@pytest.mark.parametrize("kernel", sorted(KERNELS))
def test_dp_kernel_long_sides_and_indel_blocks(ctx, kernel):
    ctx.set_dp_kernel(KERNELS[kernel])
    rng = np.random.default_rng(102)
    hox = T.scoring_hoxd70()
    cases = pairs(rng, [3000, 2500, 4000, 1800, 3500, 600, 5000, 2200], indel_blocks=True)
    gp = T.GappedParams(9400, 1000)
    say(kernel, "long sides with indel blocks")
    got, cells = run_all(ctx, cases, hox, gp)
    exp, ecells = oracle_all(cases, hox, gp)
    assert cells == ecells
    for k in range(len(cases)):
        assert got[k] == exp[k], (kernel, k)
    ctx.set_dp_kernel(0)


@pytest.mark.parametrize("kernel", sorted(KERNELS))
def test_dp_kernel_int32_and_asymmetric_matrices(ctx, kernel):
    ctx.set_dp_kernel(KERNELS[kernel])
    rng = np.random.default_rng(103)
    wide = T.scoring_hoxd70()  # scores outside int8: the int32 row table
    asym = T.scoring_hoxd70()
    for a in range(4):
        for b in range(4):
            wide.sub[a][b] = 3 * wide.sub[a][b]
            if a != b:
                asym.sub[a][b] += 7 if a < b else -11
    wide.gap_open, wide.gap_extend = 1200, 90
    cases = pairs(rng, [int(x) for x in rng.integers(20, 900, 30)], pn=0.03)
    for sc, gp in ((wide, T.GappedParams(9400 * 3, 3000)), (asym, T.GappedParams(9400, 1000))):
        say(kernel, "matrix", sc.sub[0][0], sc.sub[0][1], sc.sub[1][0])
        got, cells = run_all(ctx, cases, sc, gp)
        exp, ecells = oracle_all(cases, sc, gp)
        assert cells == ecells
        for k in range(len(cases)):
            assert got[k] == exp[k], (kernel, k)
    ctx.set_dp_kernel(0)


def test_slices_equal_pairs_and_oracle(ctx):
    """gxy_gapped_extend_slices_dev over pairs laid end to end in two device
    sequences equals the oracle on each pair (and the per-pair entry point)."""
    ctx.set_dp_kernel(0)
    rng = np.random.default_rng(104)
    hox = T.scoring_hoxd70()
    cases = pairs(rng, [int(x) for x in rng.integers(1, 1200, 80)])
    tcat = np.concatenate([t.unpack()[0] for t, _, _ in cases])
    qcat = np.concatenate([q.unpack()[0] for _, q, _ in cases])
    tval = np.concatenate([t.unpack()[1] for t, _, _ in cases])
    qval = np.concatenate([q.unpack()[1] for _, q, _ in cases])
    tl = [len(t) for t, _, _ in cases]
    ql = [len(q) for _, q, _ in cases]
    to = np.concatenate([[0], np.cumsum(tl)[:-1]])
    qo = np.concatenate([[0], np.cumsum(ql)[:-1]])
    z = np.zeros
    dt = gxy.DeviceSeq.upload(T.PackedSeq.from_arrays(tcat, tval, z(len(tcat), bool)), ctx)
    dq = gxy.DeviceSeq.upload(T.PackedSeq.from_arrays(qcat, qval, z(len(qcat), bool)), ctx)
    gp = T.GappedParams(9400, 500)
    got, cells = gxy.gapped_extend_slices_dev(dt, dq, to, tl, qo, ql, [a for _, _, a in cases], hox, gp, ctx)
    exp, ecells = oracle_all(cases, hox, gp)
    assert cells == ecells
    assert got == exp
    per, pcells = run_all(ctx, cases, hox, gp)
    assert per == got and pcells == cells


def test_autotuner_probes_then_keeps_the_fastest():
    """With no kernel forced, a context measures every candidate on its first
    passes of a size class (one pass each), keeps the fastest, and every pass —
    probe or not — gives the oracle's alignments.  A pass bound by its longest
    side's row latency (fewer rows in all than that side's rows times the
    one-wavefront slots of the GPU) uses the four-wavefront kernel without
    probing: 80 sides of up to 800 rows here."""
    ctx = gxy.Context(0)
    rng = np.random.default_rng(105)
    hox = T.scoring_hoxd70()
    gp = T.GappedParams(9400, 300)
    few = pairs(rng, [int(x) for x in rng.integers(50, 800, 40)])  # 80 long-ish sides
    got, cells = run_all(ctx, few, hox, gp)
    exp, ecells = oracle_all(few, hox, gp)
    assert cells == ecells and got == exp
    assert all(c == (None, {}) for c in ctx.dp_autotune_state()[1:])
    # 8400 sides of 40-60 rows: rows enough to keep every slot busy (class 3)
    cases = pairs(rng, [int(x) for x in rng.integers(40, 60, 4200)])
    t, q, to, tl, qo, ql, anc = slices(cases)
    dt, dq = gxy.DeviceSeq.upload(t, ctx), gxy.DeviceSeq.upload(q, ctx)
    exp, ecells = oracle_all(cases, hox, gp)
    n_cand = len(gxy.Context.DP_CANDIDATES)
    for k in range(n_cand + 1):
        got, cells = gxy.gapped_extend_slices_dev(dt, dq, to, tl, qo, ql, anc, hox, gp, ctx)
        assert cells == ecells and got == exp, k
    state = ctx.dp_autotune_state()
    assert state[0][0] == "block4" and state[0][1] == {}
    chosen, rates = state[3]
    assert set(rates) == set(gxy.Context.DP_CANDIDATES) and all(v > 0 for v in rates.values())
    assert chosen == max(rates, key=rates.get)
    ctx.close()


def slices(cases):
    """Pairs laid end to end in two host sequences: (t, q, t_off, t_len, q_off, q_len, anchors)."""
    tu = [t.unpack() for t, _, _ in cases]
    qu = [q.unpack() for _, q, _ in cases]
    tl = [len(u[0]) for u in tu]
    ql = [len(u[0]) for u in qu]
    cat = lambda us, k: np.concatenate([u[k] for u in us])
    t = T.PackedSeq.from_arrays(cat(tu, 0), cat(tu, 1), np.zeros(sum(tl), bool))
    q = T.PackedSeq.from_arrays(cat(qu, 0), cat(qu, 1), np.zeros(sum(ql), bool))
    to = np.concatenate([[0], np.cumsum(tl)[:-1]])
    qo = np.concatenate([[0], np.cumsum(ql)[:-1]])
    return t, q, to, tl, qo, ql, [a for _, _, a in cases]
