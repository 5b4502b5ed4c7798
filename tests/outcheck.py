"""MAF/PAF checks fixed by the formats themselves, shared by the CPU and GPU
tests of the output path, and the reference's corpus and fixture inputs.

check_maf: every block's rows without '-' equal the source slices (the
query's reverse complement on the minus strand), the size fields equal the
ungapped row lengths and the block score equals the rows rescored with the
substitution matrix and affine gaps (O + E to open, E per further base:
src/gapped_extend.rs:197-198).  check_paf: every PAF line equals what its MAF
block fixes (src/output/paf.rs:23-89).
"""
import ctypes as C
import os
import re

import gxy_types as T
import oracle_py as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
PS = T.PackedSeq.from_ascii
HOX = T.HOXD70
IDX = {ord("A"): 0, ord("C"): 1, ord("G"): 2, ord("T"): 3}


class ManualRecords:
    """A gxy_record_set built from explicit records, for the writers.  Each
    record is (ti, qi, strand, t_start, q_start, t_span, q_span, score,
    [(op, n), ...])."""

    def __init__(self, records):
        runs = []
        self._recs = (T.Record * max(1, len(records)))()
        for k, (ti, qi, strand, ts, qs, tsp, qsp, score, script) in enumerate(records):
            self._recs[k] = T.Record(ti, qi, strand, ts, qs, tsp, qsp, score, len(script), 0,
                                     len(runs))
            runs.extend(script)
        self._runs = (T.EditRun * max(1, len(runs)))(*[T.EditRun(op, n) for op, n in runs])
        self._set = T.RecordSet(len(records), C.cast(self._recs, C.POINTER(T.Record)),
                                len(runs), C.cast(self._runs, C.POINTER(T.EditRun)))
        self.ptr = C.pointer(self._set)


def padded(prefix: bytes, n: int) -> T.PackedSeq:
    return PS(prefix + b"T" * (n - len(prefix)))


def read_fasta(path):
    with open(path, "rb") as f:
        return [(name, PS(seq)) for name, seq in O.parse_fasta(f.read())]


# ---- what the formats fix on their own -------------------------------------------

def col_score(a, b):
    if a not in IDX or b not in IDX:
        return -100
    return HOX[IDX[a]][IDX[b]]


def rescore(trow: bytes, qrow: bytes, gap_open=400, gap_extend=30):
    """Score of an alignment given as two MAF rows: substitution scores over the
    aligned columns, O + k*E per maximal gap run of length k in either row."""
    score = 0
    k = 0
    while k < len(trow):
        if trow[k] == ord("-") or qrow[k] == ord("-"):
            row = trow if trow[k] == ord("-") else qrow
            e = k
            while e < len(row) and row[e] == ord("-"):
                e += 1
            score -= gap_open + gap_extend * (e - k)
            k = e
        else:
            score += col_score(trow[k], qrow[k])
            k += 1
    return score


def check_maf(text: bytes, targets, queries):
    """Validate every block of a MAF text against the sequences.  Returns the
    blocks as (score, tname, tstart, tsize, tsrc, trow, qname, qstart, qsize,
    strand, qsrc, qrow)."""
    tmap = {n: s for n, s in targets}
    qmap = {n: s for n, s in queries}
    lines = text.split(b"\n")
    assert lines[0] == b"##maf version=1 program=lastz-gxy"
    blocks = []
    k = 1
    while k < len(lines):
        if not lines[k].startswith(b"a score="):
            k += 1
            continue
        score = int(lines[k][8:])
        st = lines[k + 1].split()
        sq = lines[k + 2].split()
        assert lines[k + 3] == b"", "a block ends with an empty line"
        assert st[0] == b"s" and sq[0] == b"s" and len(st) == 7 and len(sq) == 7
        tname, ts, tsz, tstr, tsrc, trow = st[1].decode(), int(st[2]), int(st[3]), st[4], int(st[5]), st[6]
        qname, qs, qsz, qstr, qsrc, qrow = sq[1].decode(), int(sq[2]), int(sq[3]), sq[4], int(sq[5]), sq[6]
        assert tstr == b"+" and qstr in (b"+", b"-")
        assert len(trow) == len(qrow)
        assert not any(a == b == ord("-") for a, b in zip(trow, qrow))
        t, q = tmap[tname], qmap[qname]
        assert tsrc == t.len and qsrc == q.len
        qstrand = q if qstr == b"+" else q.reverse_complement()
        assert trow.replace(b"-", b"") == t.to_ascii()[ts:ts + tsz]
        assert qrow.replace(b"-", b"") == qstrand.to_ascii()[qs:qs + qsz]
        assert score == rescore(trow, qrow), (tname, qname, ts, qs)
        blocks.append((score, tname, ts, tsz, tsrc, trow, qname, qs, qsz, qstr.decode(), qsrc, qrow))
        k += 4
    return blocks


def check_paf(text: bytes, maf_blocks):
    """Each PAF line equals what its MAF block (same record, same order) fixes:
    coordinates, strand, matching columns, aligned columns, NM = mismatches +
    gap columns, AS = score and the CIGAR read off the rows."""
    lines = [x for x in text.split(b"\n") if x]
    assert len(lines) == len(maf_blocks)
    for line, blk in zip(lines, maf_blocks):
        score, tname, ts, tsz, tsrc, trow, qname, qs, qsz, strand, qsrc, qrow = blk
        c = line.decode().split("\t")
        assert len(c) == 16
        ops = ["I" if a == ord("-") else "D" if b == ord("-") else "M" for a, b in zip(trow, qrow)]
        cig, k = "", 0
        while k < len(ops):
            e = k
            while e < len(ops) and ops[e] == ops[k]:
                e += 1
            cig += f"{e - k}{ops[k]}"
            k = e
        nmatch = sum(1 for a, b, o in zip(trow, qrow, ops) if o == "M" and a == b)
        assert c == [qname, str(qsrc), str(qs), str(qs + qsz), strand, tname, str(tsrc), str(ts),
                     str(ts + tsz), str(nmatch), str(len(ops)), "60", f"NM:i:{len(ops) - nmatch}",
                     f"AS:i:{score}", "tp:A:P", f"cg:Z:{cig}"], line


def smoke_minus_case():
    query = b"GATTACACATGGCATGTCGA" + b"AGCTAGCTAGCTAGCTAGCTAGCTA" + b"CCCGGGAATTCCGGTTAAGGCC"
    rc_payload = PS(query).reverse_complement().to_ascii()
    target = b"TTTTTTTTTT" + rc_payload + b"TTTTTTTTTT"
    cfg = T.config_default()
    cfg.pattern = T.pattern_from_spec("match12")
    cfg.strand = T.STRAND_MINUS
    cfg.hsp = T.HspParams(910, 500)
    cfg.gapped = T.GappedParams(9400, 500)
    cfg.gapped_enabled = 1
    cfg.chain_enabled = 1
    cfg.tweener_enabled = 1
    cfg.inner_pattern = T.pattern_from_spec("match8")
    cfg.inner_min_gap = 5
    cfg.inner_max_gap = 1000
    cfg.inner_relax = 2
    return [("chrT", PS(target))], [("chrQ", PS(query))], cfg, len(rc_payload)


def check_smoke_minus(tuples, n_payload, qlen):
    """The assertions of tests/fixture_smoke.rs:115-154 on (ti, qi, strand,
    t_start, q_start, t_span, q_span, score, cigar) tuples."""
    assert tuples, "expected at least one minus-strand record"
    for ti, qi, strand, ts, qs, tsp, qsp, score, cig in tuples:
        assert strand == T.STRAND_MINUS
        assert ts >= 10 and ts + tsp <= 10 + n_payload
        assert qs + qsp <= qlen
        runs = [(int(n), op) for n, op in re.findall(r"(\d+)([MID])", cig)]
        assert qsp == sum(n for n, op in runs if op in "MI")
        assert tsp == sum(n for n, op in runs if op in "MD")
    for a, b in zip(tuples, tuples[1:]):
        assert a[3] <= b[3], "records sorted by t_start"


CORPUS_PAIRS = [  # parity/scripts/matrix.sh:21-28 and the "gxy" column of README.md:78-83
    ("pseudocat", "pseudocat", 1),
    ("pseudopig1", "pseudopig1", 1),
    ("pseudocat", "pseudopig1", 6),
    ("pseudocat", "pseudopig", 21),
    ("sars_cov2", "sars_cov2", 1),
    ("sars_cov2", "sars_cov1", 1),
]


def corpus(name):
    return read_fasta(os.path.join(GOLDEN, "corpus", name + ".fa"))


def fixture(name):
    return read_fasta(os.path.join(GOLDEN, "fixtures", name + ".fa"))
