"""ctypes mirrors of include/gxy.h and a numpy PackedSeq.

PackedSeq follows the reference layout (src/sequences.rs:128-141): 2-bit codes
A=0 C=1 G=2 T=3, four bases per byte (base i at bits 2*(i&3) of byte i>>2), and
one `valid` and one `mask` bit per base (bit i&7 of byte i>>3).  Lowercase ACGT
is valid and soft-masked; anything but ACGTU/acgtu is invalid with code 0
(src/dna.rs:12-20).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

GXY_OK = 0
GXY_EINVAL = -1
GXY_ENOMEM = -2
GXY_EDEVICE = -3
GXY_EUNSUP = -4
GXY_ERANGE = -5

OP_MATCH, OP_INSERT, OP_DELETE = 0, 1, 2
STRAND_PLUS, STRAND_MINUS, STRAND_BOTH = 0, 1, 2


class Seq(C.Structure):
    _fields_ = [("codes", C.c_void_p), ("valid", C.c_void_p), ("mask", C.c_void_p),
                ("len", C.c_uint64)]


class SeedPatternC(C.Structure):
    _fields_ = [("length", C.c_uint32), ("care", C.c_uint32)]


class Scoring(C.Structure):
    _fields_ = [("sub", (C.c_int32 * 4) * 4), ("gap_open", C.c_int32),
                ("gap_extend", C.c_int32), ("ambig", C.c_int32)]


class Hsp(C.Structure):
    _fields_ = [("t_start", C.c_uint32), ("q_start", C.c_uint32), ("length", C.c_uint32),
                ("score", C.c_int32)]

    def astuple(self):
        return (self.t_start, self.q_start, self.length, self.score)


class HspParams(C.Structure):
    _fields_ = [("x_drop", C.c_int32), ("hsp_threshold", C.c_int32)]


class SearchParams(C.Structure):
    _fields_ = [("step", C.c_uint64), ("hsp", HspParams), ("transitions", C.c_uint32),
                ("entropy_enabled", C.c_int32), ("entropy_threshold", C.c_int32)]


class Anchor(C.Structure):
    _fields_ = [("t_pos", C.c_uint32), ("q_pos", C.c_uint32), ("source", Hsp)]


class GappedParams(C.Structure):
    _fields_ = [("y_drop", C.c_int32), ("gapped_threshold", C.c_int32)]


class EditRun(C.Structure):
    _fields_ = [("op", C.c_uint32), ("n", C.c_uint32)]


class Alignment(C.Structure):
    _fields_ = [("t_start", C.c_uint32), ("q_start", C.c_uint32), ("t_len", C.c_uint32),
                ("q_len", C.c_uint32), ("score", C.c_int32), ("n_runs", C.c_uint32),
                ("runs_offset", C.c_uint64), ("source", Anchor)]


class AlignmentSet(C.Structure):
    _fields_ = [("n", C.c_uint64), ("alns", C.POINTER(Alignment)),
                ("found", C.POINTER(C.c_uint8)), ("n_runs", C.c_uint64),
                ("runs", C.POINTER(EditRun)), ("cells", C.c_uint64)]


class SearchStats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("query_windows", "lookups", "hits", "extended",
                                           "columns", "passing", "emitted", "tie_pairs")]


class NamedSeq(C.Structure):
    _fields_ = [("name", C.c_char_p), ("seq", Seq)]


class Config(C.Structure):
    _fields_ = [("pattern", SeedPatternC), ("scoring", Scoring), ("step", C.c_uint64),
                ("hsp", HspParams), ("gapped", GappedParams), ("strand", C.c_uint32),
                ("max_word_count", C.c_uint32), ("respect_masking", C.c_int32),
                ("gapped_enabled", C.c_int32), ("chain_enabled", C.c_int32),
                ("tweener_enabled", C.c_int32), ("anchor_window", C.c_uint32),
                ("transitions", C.c_uint32), ("chunk_primary_bp", C.c_uint32),
                ("chunk_halo_bp", C.c_uint32), ("entropy_enabled", C.c_int32),
                ("entropy_threshold", C.c_int32), ("inner_pattern", SeedPatternC),
                ("inner_min_gap", C.c_uint32), ("inner_max_gap", C.c_uint32), ("inner_relax", C.c_int32)]


class Record(C.Structure):
    _fields_ = [("target_index", C.c_uint32), ("query_index", C.c_uint32),
                ("strand", C.c_uint32), ("t_start", C.c_uint32), ("q_start", C.c_uint32),
                ("t_span", C.c_uint32), ("q_span", C.c_uint32), ("score", C.c_int32),
                ("n_runs", C.c_uint32), ("reserved", C.c_uint32), ("runs_offset", C.c_uint64)]


class RecordSet(C.Structure):
    _fields_ = [("n", C.c_uint64), ("records", C.POINTER(Record)), ("n_runs", C.c_uint64),
                ("runs", C.POINTER(EditRun))]


class RunStats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("chunks", "query_bp", "hits", "columns", "hsps",
                                           "dp_problems", "dp_cells", "tie_pairs",
                                           "score_ties", "passing", "lookups")]


# ---------------------------------------------------------------------------
# Scoring and seed patterns
# ---------------------------------------------------------------------------

HOXD70 = ((91, -114, -31, -123), (-114, 100, -125, -31), (-31, -125, 100, -114),
          (-123, -31, -114, 91))  # src/scoring.rs:17-23


def scoring_hoxd70() -> Scoring:
    """ScoringMatrix::hoxd70 (src/scoring.rs:44-58)."""
    s = Scoring()
    for a in range(4):
        for b in range(4):
            s.sub[a][b] = HOXD70[a][b]
    s.gap_open, s.gap_extend, s.ambig = 400, 30, -100
    return s


def pattern_from_spec(spec: str) -> SeedPatternC:
    """parse_seed_spec (src/cli.rs:225-240) / SeedPattern::parse (src/seeds.rs:38-51)."""
    if spec == "12of19":
        bits = "1110100110010101111"
    elif spec == "19of20":
        bits = "11111111101111111111"
    elif spec.startswith("match"):
        bits = "1" * int(spec[5:])
    elif spec and all(c in "01" for c in spec):
        bits = spec
    else:
        raise ValueError(f"unrecognised seed spec {spec!r}")
    if not 1 <= len(bits) <= 32:
        raise ValueError("seed pattern length must be 1..32")
    p = SeedPatternC()
    p.length = len(bits)
    p.care = sum(1 << k for k, c in enumerate(bits) if c == "1")
    return p


def pattern_span_weight(p: SeedPatternC):
    return p.length, bin(p.care).count("1")


def config_default() -> Config:
    """Config::default() (src/driver.rs:91-111) with the CLI defaults (src/cli.rs)."""
    c = Config()
    c.pattern = pattern_from_spec("12of19")
    c.scoring = scoring_hoxd70()
    c.step = 1
    c.hsp = HspParams(910, 3000)
    c.gapped = GappedParams(9400, 3000)
    c.strand = STRAND_BOTH
    c.max_word_count = 0
    c.respect_masking = 1
    c.gapped_enabled = 1
    c.chain_enabled = 0
    c.tweener_enabled = 0
    c.anchor_window = 31
    c.transitions = 1
    c.chunk_primary_bp = 0
    c.chunk_halo_bp = 50000
    c.entropy_enabled = 0
    c.entropy_threshold = 0
    c.inner_pattern = pattern_from_spec("match12")  # TweenerConfig::default (src/tweener.rs:44-52)
    c.inner_min_gap = 100
    c.inner_max_gap = 100_000
    c.inner_relax = 2
    return c


# ---------------------------------------------------------------------------
# PackedSeq
# ---------------------------------------------------------------------------

_ENC = np.full(256, 255, np.uint8)
for _ch, _c in zip(b"ACGTUacgtu", (0, 1, 2, 3, 3, 0, 1, 2, 3, 3)):
    _ENC[_ch] = _c
_LOWER = np.zeros(256, bool)
_LOWER[ord("a"):ord("z") + 1] = True
_DEC = np.frombuffer(b"ACGT", np.uint8)


class PackedSeq:
    """2-bit PackedSeq with `valid` and `mask` bitsets (src/sequences.rs:128-318)."""

    __slots__ = ("codes", "valid", "mask", "len")

    def __init__(self, codes: np.ndarray, valid: np.ndarray, mask: np.ndarray, length: int):
        self.codes = np.ascontiguousarray(codes, np.uint8)
        self.valid = np.ascontiguousarray(valid, np.uint8)
        self.mask = np.ascontiguousarray(mask, np.uint8)
        self.len = int(length)

    @classmethod
    def from_arrays(cls, code: np.ndarray, valid: np.ndarray, mask: np.ndarray) -> "PackedSeq":
        """Pack per-base arrays (code 0..3, valid bool, mask bool)."""
        n = len(code)
        code = np.where(valid, code, 0).astype(np.uint8)
        mask = np.logical_and(mask, valid)
        c4 = np.zeros(((n + 3) // 4) * 4, np.uint8)
        c4[:n] = code
        c4 = c4.reshape(-1, 4)
        codes = c4[:, 0] | (c4[:, 1] << 2) | (c4[:, 2] << 4) | (c4[:, 3] << 6)
        return cls(codes, np.packbits(valid.astype(bool), bitorder="little"),
                   np.packbits(mask, bitorder="little"), n)

    @classmethod
    def from_ascii(cls, text) -> "PackedSeq":
        """PackedSeq::from_ascii (src/sequences.rs:164-170)."""
        if isinstance(text, str):
            text = text.encode()
        a = np.frombuffer(bytes(text), np.uint8)
        enc = _ENC[a]
        valid = enc != 255
        return cls.from_arrays(np.where(valid, enc, 0), valid, valid & _LOWER[a])

    def unpack(self):
        """Per-base (code, valid, mask) arrays."""
        n = self.len
        c = self.codes
        code = np.stack([c & 3, (c >> 2) & 3, (c >> 4) & 3, (c >> 6) & 3], axis=1).reshape(-1)[:n]
        valid = np.unpackbits(self.valid, bitorder="little")[:n].astype(bool)
        mask = np.unpackbits(self.mask, bitorder="little")[:n].astype(bool)
        return code, valid, mask

    def to_ascii(self) -> bytes:
        """PackedSeq::to_ascii (src/sequences.rs:268-272)."""
        code, valid, _ = self.unpack()
        out = np.where(valid, _DEC[code], ord("N")).astype(np.uint8)
        return out.tobytes()

    def reverse_complement(self) -> "PackedSeq":
        """PackedSeq::reverse_complement (src/sequences.rs:303-317)."""
        code, valid, mask = self.unpack()
        return PackedSeq.from_arrays((3 - code)[::-1], valid[::-1], mask[::-1])

    def slice(self, start: int, end: int) -> "PackedSeq":
        """PackedSeq::slice_to_new (src/sequences.rs:277-298)."""
        code, valid, mask = self.unpack()
        return PackedSeq.from_arrays(code[start:end], valid[start:end], mask[start:end])

    def cleared_masks(self) -> "PackedSeq":
        """PackedSeq::clear_masks (src/sequences.rs:229-233), as a copy."""
        return PackedSeq(self.codes, self.valid, np.zeros_like(self.mask), self.len)

    def view(self) -> Seq:
        s = Seq()
        s.codes = self.codes.ctypes.data if self.codes.size else None
        s.valid = self.valid.ctypes.data if self.valid.size else None
        s.mask = self.mask.ctypes.data if self.mask.size else None
        s.len = self.len
        return s

    def __len__(self):
        return self.len


def named_seq_array(seqs):
    """[(name, PackedSeq)] -> (NamedSeq array, keepalive)."""
    arr = (NamedSeq * max(1, len(seqs)))()
    keep = []
    for i, (name, ps) in enumerate(seqs):
        nb = name.encode() if isinstance(name, str) else bytes(name)
        keep.append(nb)
        arr[i].name = nb
        arr[i].seq = ps.view()
    return arr, keep


def record_tuples(rs: RecordSet):
    """Records as plain tuples with their CIGAR text."""
    out = []
    for k in range(rs.n):
        r = rs.records[k]
        cig = "".join(f"{rs.runs[r.runs_offset + x].n}{'MID'[rs.runs[r.runs_offset + x].op]}"
                      for x in range(r.n_runs))
        out.append((r.target_index, r.query_index, r.strand, r.t_start, r.q_start, r.t_span,
                    r.q_span, r.score, cig))
    return out


class ServerParams(C.Structure):
    """gxy_server_params."""
    _fields_ = [("hbm_budget_bytes", C.c_uint64), ("max_batch_bp", C.c_uint64),
                ("co_batch", C.c_int32), ("dp_workers", C.c_int32)]


class RequestStats(C.Structure):
    """gxy_request_stats (server clock, ms)."""
    _fields_ = [("submit_ms", C.c_double), ("admit_ms", C.c_double), ("done_ms", C.c_double),
                ("query_bp", C.c_uint64), ("iteration", C.c_uint64), ("batch_queries", C.c_uint64),
                ("batch_bp", C.c_uint64)]


class DSeqList(C.Structure):
    """gxy_dseq_list."""
    _fields_ = [("n", C.c_uint64), ("names", C.POINTER(C.c_void_p)), ("seqs", C.POINTER(C.c_void_p))]

