"""ctypes binding of oracle/liboracle.so — TEST INFRASTRUCTURE ONLY.

The CPU restatement of the reference path, used as the parity checker by
tests/, by __graft_entry__.smoke() and as the CPU baseline of bench.py.  The
product (lastz-gxy_amd/) never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(_HERE), "lastz-gxy_amd"))
import gxy_types as T  # noqa: E402

LIB_PATH = os.path.join(_HERE, "liboracle.so")

_P = C.POINTER


def _load():
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"oracle library missing: {LIB_PATH} (run __graft_entry__.build())")
    lib = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    sig = {
        "gxo_pack_ascii": (C.c_int, [C.c_char_p, C.c_uint64, vp, vp, vp]),
        "gxo_seed_word": (C.c_int, [_P(T.Seq), _P(T.SeedPatternC), C.c_uint64, _P(C.c_uint64)]),
        "gxo_postable_build": (C.c_int, [_P(T.Seq), _P(T.SeedPatternC), C.c_uint64, _P(vp)]),
        "gxo_postable_prune": (C.c_int, [vp, C.c_uint32]),
        "gxo_postable_num_words": (C.c_uint64, [vp]),
        "gxo_postable_total_positions": (C.c_uint64, [vp]),
        "gxo_postable_starts": (_P(C.c_uint32), [vp]),
        "gxo_postable_positions": (_P(C.c_uint32), [vp]),
        "gxo_postable_free": (None, [vp]),
        "gxo_extend_hit": (C.c_int, [_P(T.Seq), _P(T.Seq), _P(T.Scoring), _P(T.HspParams),
                                     C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, _P(T.Hsp),
                                     _P(C.c_int)]),
        "gxo_search": (C.c_int, [vp, _P(T.Seq), _P(T.Seq), _P(T.Scoring), _P(T.SearchParams),
                                 C.c_int, _P(_P(T.Hsp)), _P(C.c_uint64), _P(T.SearchStats)]),
        "gxo_choose_anchor": (C.c_int, [_P(T.Hsp), _P(T.Seq), _P(T.Seq), _P(T.Scoring),
                                        C.c_uint32, _P(T.Anchor)]),
        "gxo_gapped_extend": (C.c_int, [_P(T.Anchor), _P(T.Seq), _P(T.Seq), _P(T.Scoring),
                                        _P(T.GappedParams), C.c_int, _P(T.Alignment),
                                        _P(_P(T.EditRun)), _P(C.c_uint64), _P(C.c_int)]),
        "gxo_shannon_entropy": (C.c_double, [_P(T.Seq), C.c_uint64, C.c_uint64]),
        "gxo_passes_entropy_gate": (C.c_int, [_P(T.Seq), C.c_uint32, C.c_uint32, C.c_int32,
                                              C.c_int32]),
        "gxo_run": (C.c_int, [_P(T.NamedSeq), C.c_uint64, _P(T.NamedSeq), C.c_uint64,
                              _P(T.Config), C.c_int, C.c_int, C.c_int, _P(_P(T.RecordSet)),
                              _P(T.RunStats)]),
        "gxo_format_maf": (C.c_int, [_P(T.NamedSeq), C.c_uint64, _P(T.NamedSeq), C.c_uint64,
                                     _P(T.RecordSet), _P(C.c_void_p), _P(C.c_uint64)]),
        "gxo_format_paf": (C.c_int, [_P(T.NamedSeq), C.c_uint64, _P(T.NamedSeq), C.c_uint64,
                                     _P(T.RecordSet), _P(C.c_void_p), _P(C.c_uint64)]),
        "gxo_record_set_free": (None, [_P(T.RecordSet)]),
        "gxo_free": (None, [vp]),
        "gxo_best_chain": (C.c_int, [vp, C.c_uint64, vp, _P(C.c_uint64)]),
        "gxo_interpolate": (C.c_int, [_P(T.RecordSet), _P(T.Seq), _P(T.Seq), _P(T.Config),
                                      _P(_P(T.RecordSet))]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed with status {rc}")


def cigar(runs):
    return "".join(f"{n}{'MID'[op]}" for op, n in runs)


class PosTable:
    """PosTable (src/pos_table.rs:24-137) built by the oracle."""

    def __init__(self, seq: T.PackedSeq, pattern: T.SeedPatternC, step: int = 1):
        self._h = C.c_void_p()
        _check(lib().gxo_postable_build(C.byref(seq.view()), C.byref(pattern), step,
                                        C.byref(self._h)), "gxo_postable_build")
        self.pattern = pattern

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.gxo_postable_free(self._h)
            self._h = None

    @property
    def num_words(self):
        return int(lib().gxo_postable_num_words(self._h))

    @property
    def total_positions(self):
        return int(lib().gxo_postable_total_positions(self._h))

    def starts(self) -> np.ndarray:
        p = lib().gxo_postable_starts(self._h)
        return np.ctypeslib.as_array(p, shape=(self.num_words + 1,)).copy()

    def positions(self) -> np.ndarray:
        n = self.total_positions
        if n == 0:
            return np.zeros(0, np.uint32)
        p = lib().gxo_postable_positions(self._h)
        return np.ctypeslib.as_array(p, shape=(n,)).copy()

    def lookup(self, word: int):
        s = self.starts()
        return self.positions()[s[word]:s[word + 1]].tolist()

    def prune_hot_words(self, limit: int):
        _check(lib().gxo_postable_prune(self._h, limit), "gxo_postable_prune")


def seed_word(seq: T.PackedSeq, pattern: T.SeedPatternC, pos: int):
    w = C.c_uint64()
    r = lib().gxo_seed_word(C.byref(seq.view()), C.byref(pattern), pos, C.byref(w))
    if r < 0:
        raise RuntimeError("gxo_seed_word failed")
    return int(w.value) if r == 1 else None


def extend_hit(t, q, sc, params, t_pos, q_pos, seed_len, simd=False):
    out = T.Hsp()
    found = C.c_int()
    _check(lib().gxo_extend_hit(C.byref(t.view()), C.byref(q.view()), C.byref(sc),
                                C.byref(params), t_pos, q_pos, seed_len, 1 if simd else 0,
                                C.byref(out), C.byref(found)), "gxo_extend_hit")
    return out.astuple() if found.value else None


def shannon_entropy(seq: T.PackedSeq, start: int, end: int) -> float:
    """src/entropy.rs:27-49: bits of A/C/G/T frequency over the window's valid bases."""
    return lib().gxo_shannon_entropy(C.byref(seq.view()), start, end)


def passes_entropy_gate(seq: T.PackedSeq, t_start: int, t_end: int, score: int,
                        threshold: int) -> bool:
    """src/entropy.rs:56-66: score * H / 2 >= threshold."""
    return bool(lib().gxo_passes_entropy_gate(C.byref(seq.view()), t_start, t_end, score,
                                              threshold))


def search(table: PosTable, t, q, sc, sp, simd=True):
    out = _P(T.Hsp)()
    n = C.c_uint64()
    st = T.SearchStats()
    _check(lib().gxo_search(table._h, C.byref(t.view()), C.byref(q.view()), C.byref(sc),
                            C.byref(sp), 1 if simd else 0, C.byref(out), C.byref(n),
                            C.byref(st)), "gxo_search")
    res = [out[i].astuple() for i in range(n.value)]
    lib().gxo_free(out)
    return res, st


def choose_anchor(hsp, t, q, sc, window):
    h = T.Hsp(*hsp)
    a = T.Anchor()
    _check(lib().gxo_choose_anchor(C.byref(h), C.byref(t.view()), C.byref(q.view()),
                                   C.byref(sc), window, C.byref(a)), "gxo_choose_anchor")
    return (a.t_pos, a.q_pos)


def gapped_extend(anchor, t, q, sc, gp, lazy=True):
    """Returns (t_start, q_start, t_len, q_len, score, runs) or None, and cells."""
    a = T.Anchor()
    a.t_pos, a.q_pos = anchor[0], anchor[1]
    if len(anchor) > 2:
        a.source = T.Hsp(*anchor[2])
    out = T.Alignment()
    runs = _P(T.EditRun)()
    cells = C.c_uint64()
    found = C.c_int()
    _check(lib().gxo_gapped_extend(C.byref(a), C.byref(t.view()), C.byref(q.view()),
                                   C.byref(sc), C.byref(gp), 1 if lazy else 0, C.byref(out),
                                   C.byref(runs), C.byref(cells), C.byref(found)),
           "gxo_gapped_extend")
    if not found.value:
        return None, cells.value
    rr = [(runs[i].op, runs[i].n) for i in range(out.n_runs)]
    lib().gxo_free(runs)
    return (out.t_start, out.q_start, out.t_len, out.q_len, out.score, rr), cells.value


class Records:
    """Owning wrapper of a gxy_record_set produced by the oracle."""

    def __init__(self, ptr):
        self.ptr = ptr

    def __del__(self):
        if getattr(self, "ptr", None) and _lib is not None:
            _lib.gxo_record_set_free(self.ptr)
            self.ptr = None

    def tuples(self):
        return T.record_tuples(self.ptr.contents)


def run(targets, queries, cfg: T.Config, threads=0, lazy=True, simd=True):
    """driver::run (src/driver.rs:199); targets/queries are [(name, PackedSeq)]."""
    ta, tk = T.named_seq_array(targets)
    qa, qk = T.named_seq_array(queries)
    out = _P(T.RecordSet)()
    st = T.RunStats()
    _check(lib().gxo_run(ta, len(targets), qa, len(queries), C.byref(cfg), threads,
                         1 if lazy else 0, 1 if simd else 0, C.byref(out), C.byref(st)),
           "gxo_run")
    del tk, qk
    return Records(out), st


def best_chain(hsps):
    """best_chain (src/chain.rs:19); hsps are (t_start, q_start, length, score)."""
    n = len(hsps)
    arr = (T.Hsp * max(1, n))(*[T.Hsp(*h) for h in hsps])
    out = (T.Hsp * max(1, n))()
    m = C.c_uint64()
    _check(lib().gxo_best_chain(C.addressof(arr), n, C.addressof(out), C.byref(m)), "gxo_best_chain")
    return [(out[k].t_start, out[k].q_start, out[k].length, out[k].score) for k in range(m.value)]


def interpolate(records, target, query, cfg: T.Config):
    """tweener::interpolate (src/tweener.rs:59) over one group's records, given as
    (ti, qi, strand, t_start, q_start, t_span, q_span, score, [(op, n), ...])."""
    n = len(records)
    recs = (T.Record * max(1, n))()
    runs = []
    for k, r in enumerate(records):
        ti, qi, strand, ts, qs, tsp, qsp, score, script = r
        recs[k] = T.Record(ti, qi, strand, ts, qs, tsp, qsp, score, len(script), 0, len(runs))
        runs.extend(script)
    rarr = (T.EditRun * max(1, len(runs)))(*[T.EditRun(op, m) for op, m in runs])
    rs = T.RecordSet(n, C.cast(recs, _P(T.Record)), len(runs), C.cast(rarr, _P(T.EditRun)))
    out = _P(T.RecordSet)()
    _check(lib().gxo_interpolate(C.byref(rs), C.byref(target.view()), C.byref(query.view()), C.byref(cfg),
                                 C.byref(out)), "gxo_interpolate")
    return Records(out).tuples()


def _format(fn, targets, queries, recs: Records) -> bytes:
    ta, tk = T.named_seq_array(targets)
    qa, qk = T.named_seq_array(queries)
    txt = C.c_void_p()
    n = C.c_uint64()
    _check(fn(ta, len(targets), qa, len(queries), recs.ptr, C.byref(txt), C.byref(n)), "format")
    data = C.string_at(txt, n.value)
    lib().gxo_free(txt)
    del tk, qk
    return data


def format_maf(targets, queries, recs: Records) -> bytes:
    return _format(lib().gxo_format_maf, targets, queries, recs)


def format_paf(targets, queries, recs: Records) -> bytes:
    return _format(lib().gxo_format_paf, targets, queries, recs)


# ---- input formats (pure Python restatements; small inputs) ---------------------

def parse_2bit(data: bytes):
    """parse_2bit (src/bit2.rs:78-164): [(name, ASCII bytes)] with N overlays and
    lowercase mask overlays, either byte order."""
    import struct
    if len(data) < 4:
        raise ValueError("truncated 2bit file")
    magic = struct.unpack("<I", data[:4])[0]
    if magic == 0x1A412743:
        fmt = "<I"
    elif magic == 0x4327411A:
        fmt = ">I"
    else:
        raise ValueError(f"not a 2bit file (magic = 0x{magic:08x})")
    pos = 4

    def u32():
        nonlocal pos
        v = struct.unpack(fmt, data[pos:pos + 4])[0]
        pos += 4
        return v

    if u32() != 0:
        raise ValueError("unsupported 2bit version")
    count = u32()
    u32()
    index = []
    for _ in range(count):
        ns = data[pos]
        pos += 1
        name = data[pos:pos + ns].decode()
        pos += ns
        index.append((name, u32()))
    out = []
    for name, off in index:
        pos = off
        dna = u32()
        nn = u32()
        nst = [u32() for _ in range(nn)]
        nsz = [u32() for _ in range(nn)]
        nm = u32()
        mst = [u32() for _ in range(nm)]
        msz = [u32() for _ in range(nm)]
        u32()
        packed = data[pos:pos + (dna + 3) // 4]
        asc = bytearray(b"TCAG"[(packed[i >> 2] >> (6 - (i & 3) * 2)) & 3] for i in range(dna))
        for s_, z in zip(nst, nsz):
            for i in range(s_, min(s_ + z, dna)):
                asc[i] = ord("N")
        for s_, z in zip(mst, msz):
            for i in range(s_, min(s_ + z, dna)):
                if asc[i] != ord("N"):
                    asc[i] = ord(chr(asc[i]).lower())
        out.append((name, bytes(asc)))
    return out


def parse_fasta(text: bytes):
    """parse_fasta (src/sequences.rs:338-377): [(name, ASCII bytes)]; sequence
    lines before the first header stay with the first record."""
    recs, cur, name = [], bytearray(), None
    for line in text.split(b"\n"):
        if line.endswith(b"\r"):
            line = line[:-1]
        if not line:
            continue
        if line.startswith(b">"):
            if name is not None:
                recs.append((name, bytes(cur)))
                cur = bytearray()
            toks = line[1:].split()
            if not toks:
                raise ValueError("empty FASTA header")
            name = toks[0].decode()
        else:
            cur.extend(b for b in line if b not in b" \t\r\n\x0c")
    if name is not None:
        recs.append((name, bytes(cur)))
    return recs

