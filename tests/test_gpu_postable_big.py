"""A position table over more than 2^31 windows (the reference's PosTable
takes targets up to 4 Gbp: src/pos_table.rs:56): a 2.26 Gbp target made of
540 copies of a random 4 Mbp block (0.1% N), 12of19, step 1: 2.26 G windows,
2.2 G of them in the table.  The table is
built as two sorted halves merged by per-word offsets; its bucket sizes and
sampled buckets must follow from the block's own table: word w of the block
at positions P (windows of the block extended periodically) sits in the big
table at p + k * L for every copy k, except windows that would run past the
end of the last copy."""
import sys

import numpy as np
import pytest

import gxy
import gxy_types as T

pytestmark = pytest.mark.gpu

L = 1 << 22   # block length, a multiple of 32
K = 540       # copies: 2 264 924 160 bases


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_table_over_2g_windows():
    ctx = gxy.Context(0)
    rng = np.random.default_rng(401)
    code = rng.integers(0, 4, L).astype(np.uint8)
    valid = rng.random(L) >= 0.001
    block = T.PackedSeq.from_arrays(code, valid, np.zeros(L, bool))
    pat = T.pattern_from_spec("12of19")
    span = 19
    # the block's windows, periodic: the block followed by its first span - 1 bases
    ext = T.PackedSeq.from_arrays(np.concatenate([code, code[:span - 1]]),
                                  np.concatenate([valid, valid[:span - 1]]), np.zeros(L + span - 1, bool))
    small = gxy.PosTable.build(ext, pat, 1, ctx)
    s_starts, s_pos = small.export()
    assert s_pos.size == 0 or int(s_pos.max()) < L
    say("block table:", s_pos.size, "positions")

    big_seq = T.PackedSeq(np.tile(block.codes, K), np.tile(block.valid, K), np.tile(block.mask, K), L * K)
    say("uploading", big_seq.len, "bases")
    dt = gxy.DeviceSeq.upload(big_seq, ctx)
    del big_seq
    say("building the table over", L * K - span + 1, "windows")
    big = gxy.PosTable.build(dt, pat, 1, ctx)
    b_starts = np.zeros(big.num_words + 1, np.uint32)
    # bucket sizes: K copies of every block window, minus the last copy's
    # windows past the end (block positions > L - span)
    cnt_s = np.diff(s_starts.astype(np.int64))
    tail = np.zeros(big.num_words, np.int64)
    words_of = np.repeat(np.arange(big.num_words), cnt_s)
    np.add.at(tail, words_of[s_pos > L - span], 1)
    expect = K * cnt_s - tail
    assert big.total_positions == int(expect.sum())
    assert big.total_positions > (1 << 31)
    say("total positions", big.total_positions)
    nonempty = np.flatnonzero(cnt_s)
    pick = rng.choice(nonempty, 300, replace=False)
    heavy = nonempty[np.argsort(cnt_s[nonempty])[-5:]]
    for w in np.concatenate([pick, heavy]):
        got = big.lookup(int(w))
        p = s_pos[s_starts[w]:s_starts[w + 1]].astype(np.int64)
        exp = (p[None, :] + L * np.arange(K, dtype=np.int64)[:, None]).reshape(-1)
        exp = exp[exp <= L * K - span]
        assert len(got) == expect[w], int(w)
        assert np.array_equal(np.array(got, np.int64), exp), int(w)
    say("sampled", len(pick) + len(heavy), "buckets equal")
    # every bucket size, through the exported starts of the big table
    ctx.check(gxy.lib().gxy_postable_export(ctx._h, big._h, b_starts.ctypes.data, b_starts.size, None, 0),
              "gxy_postable_export")
    assert np.array_equal(np.diff(b_starts.astype(np.int64)), expect)
    say("all bucket sizes equal")
    ctx.close()
