"""Python host mirror of lastz-gxy's hot-path interface over the gfx950 library.

Names and argument meanings follow the reference crate: `PosTable.build`
(src/pos_table.rs:35), `PosTable.lookup` (:102), `PosTable.prune_hot_words`
(:113), `search` (src/seed_search.rs:45), `extend_hit` (src/hsp.rs:66),
`choose_anchor` (src/anchor.rs:25), `gapped_extend` (src/gapped_extend.rs:78)
and `run` (src/driver.rs:199).  Every call goes through the C-ABI of
include/gxy.h into lastz-gxy_amd/libgxy.so; there is no CPU fallback — if the
library or the GPU is missing, the calls raise.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

import gxy_types as T
from gxy_types import PackedSeq  # noqa: F401  (re-export)

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libgxy.so")
HEADER_PATH = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                           "include", "gxy.h")

_P = C.POINTER
_vp = C.c_void_p

_SIGS = {
    "gxy_ctx_create": (C.c_int, [C.c_int, _P(_vp)]),
    "gxy_ctx_destroy": (None, [_vp]),
    "gxy_last_error": (C.c_char_p, [_vp]),
    "gxy_ctx_synchronize": (C.c_int, [_vp]),
    "gxy_version": (C.c_char_p, []),
    "gxy_ctx_set_timing": (C.c_int, [_vp, C.c_int]),
    "gxy_ctx_set_dp_pool_rows": (C.c_int, [_vp, C.c_uint64]),
    "gxy_ctx_release_cache": (C.c_int, [_vp]),
    "gxy_ctx_set_dp_kernel": (C.c_int, [_vp, C.c_int]),
    "gxy_ctx_dp_autotune_state": (C.c_int, [_vp, _P(C.c_int), _P(C.c_double)]),
    "gxy_ctx_dp_pool_stats": (C.c_int, [_vp, _P(C.c_uint64), _P(C.c_uint64), _P(C.c_uint64)]),
    "gxy_ctx_set_dp_persist": (C.c_int, [_vp, C.c_int]),
    "gxy_ctx_set_xdrop_form": (C.c_int, [_vp, C.c_int]),
    "gxy_ctx_set_hit_batch": (C.c_int, [_vp, C.c_uint64]),
    "gxy_ctx_xdrop_stats": (C.c_int, [_vp, _P(C.c_uint64), _P(C.c_uint64), _P(C.c_uint64)]),
    "gxy_ctx_dp_persist_stats": (C.c_int, [_vp, _P(C.c_uint64), _P(C.c_uint64), _P(C.c_uint64), _P(C.c_uint64)]),
    "gxy_ctx_kernel_time": (C.c_int, [_vp, C.c_char_p, _P(C.c_double), _P(C.c_uint64)]),
    "gxy_ctx_reset_timing": (C.c_int, [_vp]),
    "gxy_pack_ascii": (C.c_int, [C.c_char_p, C.c_uint64, _vp, _vp, _vp]),
    "gxy_dseq_upload": (C.c_int, [_vp, _P(T.Seq), _P(_vp)]),
    "gxy_dseq_reverse_complement": (C.c_int, [_vp, _vp, _P(_vp)]),
    "gxy_dseq_slice": (C.c_int, [_vp, _vp, C.c_uint64, C.c_uint64, _P(_vp)]),
    "gxy_dseq_len": (C.c_uint64, [_vp]),
    "gxy_dseq_download": (C.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "gxy_dseq_free": (None, [_vp]),
    "gxy_postable_build": (C.c_int, [_vp, _P(T.Seq), _P(T.SeedPatternC), C.c_uint64, _P(_vp)]),
    "gxy_postable_build_dev": (C.c_int, [_vp, _vp, _P(T.SeedPatternC), C.c_uint64, _P(_vp)]),
    "gxy_postable_prune": (C.c_int, [_vp, _vp, C.c_uint32]),
    "gxy_postable_lookup": (C.c_int, [_vp, _vp, C.c_uint64, _vp, C.c_uint64, _P(C.c_uint64)]),
    "gxy_postable_total_positions": (C.c_uint64, [_vp]),
    "gxy_postable_export": (C.c_int, [_vp, _vp, _vp, C.c_uint64, _vp, C.c_uint64]),
    "gxy_postable_free": (None, [_vp]),
    "gxy_search": (C.c_int, [_vp, _vp, _P(T.Seq), _P(T.Seq), _P(T.Scoring), _P(T.SearchParams),
                             _P(_P(T.Hsp)), _P(C.c_uint64), _P(T.SearchStats)]),
    "gxy_search_dev": (C.c_int, [_vp, _vp, _vp, _P(T.Scoring), _P(T.SearchParams),
                                 _P(_P(T.Hsp)), _P(C.c_uint64), _P(T.SearchStats)]),
    "gxy_extend_hit": (C.c_int, [_vp, _P(T.Seq), _P(T.Seq), _P(T.Scoring), _P(T.HspParams),
                                 C.c_uint32, C.c_uint32, C.c_uint32, _P(T.Hsp), _P(C.c_int)]),
    "gxy_extend_hit_batch": (C.c_int, [_vp, _P(T.Seq), _P(T.Seq), _P(T.Scoring), _P(T.HspParams),
                                       _vp, _vp, C.c_uint64, C.c_uint32, _vp, _vp]),
    "gxy_extend_hit_batch_noalloc": (C.c_int, [_vp, _P(T.Seq), _P(T.Seq), _P(T.Scoring), _P(T.HspParams),
                                       _vp, _vp, C.c_uint64, C.c_uint32, _vp, _vp]),
    "gxy_choose_anchor": (C.c_int, [_vp, _P(T.Hsp), _P(T.Seq), _P(T.Seq), _P(T.Scoring),
                                    C.c_uint32, _P(T.Anchor)]),
    "gxy_choose_anchor_batch": (C.c_int, [_vp, _vp, C.c_uint64, _P(T.Seq), _P(T.Seq),
                                          _P(T.Scoring), C.c_uint32, _vp]),
    "gxy_gapped_extend": (C.c_int, [_vp, _P(T.Anchor), _P(T.Seq), _P(T.Seq), _P(T.Scoring),
                                    _P(T.GappedParams), _P(T.Alignment), _vp, C.c_uint64,
                                    _P(C.c_int)]),
    "gxy_gapped_extend_batch": (C.c_int, [_vp, _vp, C.c_uint64, _P(T.Seq), _P(T.Seq),
                                          _P(T.Scoring), _P(T.GappedParams),
                                          _P(_P(T.AlignmentSet))]),
    "gxy_gapped_extend_pairs_dev": (C.c_int, [_vp, _vp, _vp, _vp, C.c_uint64, _P(T.Scoring),
                                              _P(T.GappedParams), _P(_P(T.AlignmentSet))]),
    "gxy_gapped_extend_slices_dev": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, C.c_uint64,
                                               _P(T.Scoring), _P(T.GappedParams), _P(_P(T.AlignmentSet))]),
    "gxy_alignment_set_free": (None, [_P(T.AlignmentSet)]),
    "gxy_config_default": (None, [_P(T.Config)]),
    "gxy_pattern_parse": (C.c_int, [C.c_char_p, _P(T.SeedPatternC)]),
    "gxy_run": (C.c_int, [_vp, _P(T.NamedSeq), C.c_uint64, _P(T.NamedSeq), C.c_uint64,
                          _P(T.Config), _P(_P(T.RecordSet)), _P(T.RunStats)]),
    "gxy_run_dev": (C.c_int, [_vp, _vp, C.c_uint64, _vp, C.c_uint64, _P(T.Config),
                              _P(_P(T.RecordSet)), _P(T.RunStats)]),
    "gxy_format_maf": (C.c_int, [_P(T.NamedSeq), C.c_uint64, _P(T.NamedSeq), C.c_uint64,
                                 _P(T.RecordSet), _P(_vp), _P(C.c_uint64)]),
    "gxy_format_paf": (C.c_int, [_P(T.NamedSeq), C.c_uint64, _P(T.NamedSeq), C.c_uint64,
                                 _P(T.RecordSet), _P(_vp), _P(C.c_uint64)]),
    "gxy_record_set_free": (None, [_P(T.RecordSet)]),
    "gxy_free": (None, [_vp]),
    "gxy_load_2bit": (C.c_int, [_vp, C.c_char_p, C.c_uint64, _P(_P(T.DSeqList))]),
    "gxy_load_fasta": (C.c_int, [_vp, C.c_char_p, C.c_uint64, _P(_P(T.DSeqList))]),
    "gxy_dseq_list_free": (None, [_P(T.DSeqList)]),
    "gxy_best_chain": (C.c_int, [_vp, C.c_uint64, _vp, _P(C.c_uint64)]),
    "gxy_interpolate": (C.c_int, [_vp, _P(T.RecordSet), _vp, _vp, _P(T.Config), _P(_P(T.RecordSet))]),
    "gxy_server_create": (C.c_int, [C.c_int, _vp, C.c_uint64, _P(T.Config), _P(T.ServerParams), _P(_vp)]),
    "gxy_server_set_timing": (C.c_int, [_vp, C.c_int]),
    "gxy_server_stats": (C.c_int, [_vp, _P(T.RunStats)]),
    "gxy_server_kernel_time": (C.c_int, [_vp, C.c_char_p, _P(C.c_double), _P(C.c_uint64)]),
    "gxy_server_xdrop_stats": (C.c_int, [_vp, _P(C.c_uint64), _P(C.c_uint64), _P(C.c_uint64)]),
    "gxy_server_submit": (C.c_int, [_vp, _P(T.Seq), _P(C.c_uint64)]),
    "gxy_server_wait": (C.c_int, [_vp, C.c_uint64, _P(_P(T.RecordSet)), _P(T.RequestStats)]),
    "gxy_server_last_error": (C.c_char_p, [_vp]),
    "gxy_server_destroy": (C.c_int, [_vp]),
}

_lib = None


def lib():
    """Load libgxy.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"HIP library missing: {LIB_PATH}; build it with "
                               "__graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


class GxyError(RuntimeError):
    pass


class Context:
    """A HIP stream plus scratch on one device (gxy_ctx)."""

    def __init__(self, device: int = 0):
        self._h = _vp()
        rc = lib().gxy_ctx_create(device, C.byref(self._h))
        if rc != 0:
            raise GxyError(f"gxy_ctx_create failed ({rc}): {lib().gxy_last_error(None).decode()}")

    def check(self, rc, what):
        if rc != 0:
            raise GxyError(f"{what} failed ({rc}): {lib().gxy_last_error(self._h).decode()}")

    def synchronize(self):
        self.check(lib().gxy_ctx_synchronize(self._h), "gxy_ctx_synchronize")

    def set_timing(self, on: bool = True):
        self.check(lib().gxy_ctx_set_timing(self._h, 1 if on else 0), "gxy_ctx_set_timing")

    def set_dp_pool_rows(self, rows: int):
        """Cap the gapped extension's traceback pool at `rows` rows (0: default)."""
        self.check(lib().gxy_ctx_set_dp_pool_rows(self._h, rows), "gxy_ctx_set_dp_pool_rows")

    def set_dp_kernel(self, kernel: int):
        """Choose the DP forward kernel (GXY_DP_*: 0 auto, 1 four wavefronts per side,
        2/3/4 one wavefront per side with 1/2/4 columns per lane per slot)."""
        self.check(lib().gxy_ctx_set_dp_kernel(self._h, kernel), "gxy_ctx_set_dp_kernel")

    DP_CANDIDATES = ("wave_k2_w4", "wave_k4_w4", "wave_k2", "wave_k4", "block4")
    DP_KERNEL_NAMES = {1: "block4", 2: "wave_k1", 3: "wave_k2", 4: "wave_k4", 5: "wave_k2_w4", 6: "wave_k4_w4"}

    def dp_autotune_state(self):
        """The DP autotuner per pass-size class (<= 8, <= 512, <= 8192, more sides):
        [(chosen kernel name or None, {candidate: measured GCUPS})]."""
        chosen = (C.c_int * 4)()
        g = (C.c_double * 20)()
        self.check(lib().gxy_ctx_dp_autotune_state(self._h, chosen, g), "gxy_ctx_dp_autotune_state")
        return [(self.DP_KERNEL_NAMES.get(chosen[c]),
                 {n: round(g[5 * c + k], 2) for k, n in enumerate(self.DP_CANDIDATES) if g[5 * c + k] > 0})
                for c in range(4)]

    def release_cache(self):
        """Return cached device memory (scratch buffers, traceback pool, freed
        sequence blocks) to the device."""
        self.check(lib().gxy_ctx_release_cache(self._h), "gxy_ctx_release_cache")

    def dp_pool_stats(self):
        """(passes, truncated sides, passes run with one side alone) since creation."""
        v = [C.c_uint64() for _ in range(3)]
        self.check(lib().gxy_ctx_dp_pool_stats(self._h, *[C.byref(x) for x in v]), "gxy_ctx_dp_pool_stats")
        return tuple(x.value for x in v)

    def set_xdrop_form(self, form: int):
        """x-drop form of the search: 0 first step + resumed walks (default), 1 one persistent kernel."""
        self.check(lib().gxy_ctx_set_xdrop_form(self._h, form), "gxy_ctx_set_xdrop_form")

    def set_hit_batch(self, hits: int):
        """Seed hits per expansion batch of a search (0: the default, 2^27)."""
        self.check(lib().gxy_ctx_set_hit_batch(self._h, hits), "gxy_ctx_set_hit_batch")

    def xdrop_stats(self):
        """First steps of the x-drop since creation: (columns evaluated, hits
        handed on to be resumed, passing extensions finished)."""
        v = [C.c_uint64() for _ in range(3)]
        self.check(lib().gxy_ctx_xdrop_stats(self._h, *[C.byref(x) for x in v]), "gxy_ctx_xdrop_stats")
        return tuple(x.value for x in v)

    def set_dp_persist(self, mode: int):
        """Persistent DP launches: 0 for large batches (default), 1 always, -1 never."""
        self.check(lib().gxy_ctx_set_dp_persist(self._h, mode), "gxy_ctx_set_dp_persist")

    def dp_persist_stats(self):
        """(persistent launches, sides given to them, sides left to passes, DP
        cells of the sides they finished) since creation."""
        v = [C.c_uint64() for _ in range(4)]
        self.check(lib().gxy_ctx_dp_persist_stats(self._h, *[C.byref(x) for x in v]), "gxy_ctx_dp_persist_stats")
        return tuple(x.value for x in v)

    def reset_timing(self):
        self.check(lib().gxy_ctx_reset_timing(self._h), "gxy_ctx_reset_timing")

    def kernel_time(self, name: str):
        """(milliseconds, launches) summed for kernel `name` since the last reset."""
        ms = C.c_double()
        n = C.c_uint64()
        self.check(lib().gxy_ctx_kernel_time(self._h, name.encode(), C.byref(ms), C.byref(n)),
                   "gxy_ctx_kernel_time")
        return ms.value, n.value

    def close(self):
        if getattr(self, "_h", None):
            lib().gxy_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


class DeviceSeq:
    """A PackedSeq resident in HBM (gxy_dseq)."""

    def __init__(self, handle, ctx: Context):
        self._h = handle
        self.ctx = ctx

    @classmethod
    def upload(cls, seq: T.PackedSeq, ctx: Context | None = None) -> "DeviceSeq":
        ctx = ctx or default_context()
        h = _vp()
        ctx.check(lib().gxy_dseq_upload(ctx._h, C.byref(seq.view()), C.byref(h)), "gxy_dseq_upload")
        return cls(h, ctx)

    def __len__(self):
        return int(lib().gxy_dseq_len(self._h))

    def reverse_complement(self) -> "DeviceSeq":
        h = _vp()
        self.ctx.check(lib().gxy_dseq_reverse_complement(self.ctx._h, self._h, C.byref(h)),
                       "gxy_dseq_reverse_complement")
        return DeviceSeq(h, self.ctx)

    def slice(self, start: int, end: int) -> "DeviceSeq":
        h = _vp()
        self.ctx.check(lib().gxy_dseq_slice(self.ctx._h, self._h, start, end, C.byref(h)),
                       "gxy_dseq_slice")
        return DeviceSeq(h, self.ctx)

    def download(self) -> T.PackedSeq:
        n = len(self)
        codes = np.zeros((n + 3) // 4, np.uint8)
        valid = np.zeros((n + 7) // 8, np.uint8)
        mask = np.zeros((n + 7) // 8, np.uint8)
        self.ctx.check(lib().gxy_dseq_download(self.ctx._h, self._h, codes.ctypes.data,
                                               valid.ctypes.data, mask.ctypes.data),
                       "gxy_dseq_download")
        return T.PackedSeq(codes, valid, mask, n)

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.gxy_dseq_free(self._h)
            self._h = None


class PosTable:
    """PosTable (src/pos_table.rs:24-137) resident in HBM."""

    def __init__(self, handle, ctx, pattern, keep=None):
        self._h = handle
        self.ctx = ctx
        self.pattern = pattern
        self._keep = keep

    @classmethod
    def build(cls, seq, pattern: T.SeedPatternC, step: int = 1, ctx: Context | None = None):
        """PosTable::build(&PackedSeq, &SeedPattern, step); `seq` may be a
        PackedSeq (uploaded) or a DeviceSeq (kept resident)."""
        ctx = ctx or default_context()
        h = _vp()
        if isinstance(seq, DeviceSeq):
            ctx.check(lib().gxy_postable_build_dev(ctx._h, seq._h, C.byref(pattern), step,
                                                   C.byref(h)), "gxy_postable_build_dev")
            return cls(h, ctx, pattern, keep=seq)
        ctx.check(lib().gxy_postable_build(ctx._h, C.byref(seq.view()), C.byref(pattern), step,
                                           C.byref(h)), "gxy_postable_build")
        return cls(h, ctx, pattern, keep=seq)

    @property
    def num_words(self):
        return 1 << (2 * bin(self.pattern.care).count("1"))

    @property
    def total_positions(self):
        return int(lib().gxy_postable_total_positions(self._h))

    def export(self):
        starts = np.zeros(self.num_words + 1, np.uint32)
        pos = np.zeros(max(1, self.total_positions), np.uint32)
        self.ctx.check(lib().gxy_postable_export(self.ctx._h, self._h, starts.ctypes.data,
                                                 starts.size, pos.ctypes.data, pos.size),
                       "gxy_postable_export")
        return starts, pos[:self.total_positions]

    def lookup(self, word: int):
        n = C.c_uint64()
        self.ctx.check(lib().gxy_postable_lookup(self.ctx._h, self._h, word, None, 0, C.byref(n)),
                       "gxy_postable_lookup")
        out = np.zeros(max(1, n.value), np.uint32)
        self.ctx.check(lib().gxy_postable_lookup(self.ctx._h, self._h, word, out.ctypes.data,
                                                 out.size, C.byref(n)), "gxy_postable_lookup")
        return out[:n.value].tolist()

    def prune_hot_words(self, limit: int):
        self.ctx.check(lib().gxy_postable_prune(self.ctx._h, self._h, limit), "gxy_postable_prune")

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.gxy_postable_free(self._h)
            self._h = None


def search(table: PosTable, target, query, matrix: T.Scoring, params: T.SearchParams):
    """search(&PosTable, &PackedSeq, &PackedSeq, &ScoringMatrix, &SearchParams)
    -> (list of (t_start, q_start, length, score), SearchStats)."""
    ctx = table.ctx
    out = _P(T.Hsp)()
    n = C.c_uint64()
    st = T.SearchStats()
    if isinstance(query, DeviceSeq):
        ctx.check(lib().gxy_search_dev(ctx._h, table._h, query._h, C.byref(matrix),
                                       C.byref(params), C.byref(out), C.byref(n), C.byref(st)),
                  "gxy_search_dev")
    else:
        ctx.check(lib().gxy_search(ctx._h, table._h, C.byref(target.view()),
                                   C.byref(query.view()), C.byref(matrix), C.byref(params),
                                   C.byref(out), C.byref(n), C.byref(st)), "gxy_search")
    res = [out[i].astuple() for i in range(n.value)]
    lib().gxy_free(out)
    return res, st


HSP_DTYPE = np.dtype([("t_start", np.uint32), ("q_start", np.uint32), ("length", np.uint32),
                      ("score", np.int32)])


def search_array(table: PosTable, target: T.PackedSeq, query: T.PackedSeq, matrix: T.Scoring,
                 params: T.SearchParams):
    """search() on host sequences, the HSPs as a numpy record array (HSP_DTYPE)
    copied straight from the library's buffer."""
    ctx = table.ctx
    out = _P(T.Hsp)()
    n = C.c_uint64()
    st = T.SearchStats()
    ctx.check(lib().gxy_search(ctx._h, table._h, C.byref(target.view()), C.byref(query.view()),
                               C.byref(matrix), C.byref(params), C.byref(out), C.byref(n), C.byref(st)),
              "gxy_search")
    res = np.empty(n.value, HSP_DTYPE)
    if n.value:
        C.memmove(res.ctypes.data, out, n.value * res.itemsize)
    lib().gxy_free(out)
    return res, st


def extend_hit_batch(target, query, matrix, params, t_pos, q_pos, seed_len, ctx=None):
    """Batched extend_hit; returns a list of HSP tuples or None."""
    ctx = ctx or default_context()
    t_pos = np.ascontiguousarray(t_pos, np.uint32)
    q_pos = np.ascontiguousarray(q_pos, np.uint32)
    n = len(t_pos)
    out = (T.Hsp * max(1, n))()
    found = np.zeros(max(1, n), np.uint8)
    ctx.check(lib().gxy_extend_hit_batch(ctx._h, C.byref(target.view()), C.byref(query.view()),
                                         C.byref(matrix), C.byref(params), t_pos.ctypes.data,
                                         q_pos.ctypes.data, n, seed_len, C.addressof(out),
                                         found.ctypes.data), "gxy_extend_hit_batch")
    return [out[i].astuple() if found[i] else None for i in range(n)]


def extend_hit_batch_noalloc(target, query, matrix, params, t_pos, q_pos, seed_len, ctx=None):
    """gxy_extend_hit_batch_noalloc: extend_hit_batch with every device buffer
    in the library's fixed per-device array (no device allocation)."""
    ctx = ctx or default_context()
    t_pos = np.ascontiguousarray(t_pos, np.uint32)
    q_pos = np.ascontiguousarray(q_pos, np.uint32)
    n = len(t_pos)
    out = (T.Hsp * max(1, n))()
    found = np.zeros(max(1, n), np.uint8)
    ctx.check(lib().gxy_extend_hit_batch_noalloc(ctx._h, C.byref(target.view()), C.byref(query.view()),
                                                 C.byref(matrix), C.byref(params), t_pos.ctypes.data,
                                                 q_pos.ctypes.data, n, seed_len, C.addressof(out),
                                                 found.ctypes.data), "gxy_extend_hit_batch_noalloc")
    return [out[i].astuple() if found[i] else None for i in range(n)]


def extend_hit(target, query, matrix, params, t_pos, q_pos, seed_len, ctx=None):
    """extend_hit (src/hsp.rs:66-74) -> HSP tuple or None."""
    ctx = ctx or default_context()
    out = T.Hsp()
    found = C.c_int()
    ctx.check(lib().gxy_extend_hit(ctx._h, C.byref(target.view()), C.byref(query.view()),
                                   C.byref(matrix), C.byref(params), t_pos, q_pos, seed_len,
                                   C.byref(out), C.byref(found)), "gxy_extend_hit")
    return out.astuple() if found.value else None


def choose_anchor_batch(hsps, target, query, matrix, window, ctx=None):
    ctx = ctx or default_context()
    n = len(hsps)
    arr = (T.Hsp * max(1, n))(*[T.Hsp(*h) for h in hsps])
    out = (T.Anchor * max(1, n))()
    ctx.check(lib().gxy_choose_anchor_batch(ctx._h, C.addressof(arr), n, C.byref(target.view()),
                                            C.byref(query.view()), C.byref(matrix), window,
                                            C.addressof(out)), "gxy_choose_anchor_batch")
    return [(out[i].t_pos, out[i].q_pos) for i in range(n)]


def choose_anchor(hsp, target, query, matrix, window, ctx=None):
    """choose_anchor (src/anchor.rs:25-31) -> (t_pos, q_pos)."""
    return choose_anchor_batch([hsp], target, query, matrix, window, ctx)[0]


def _alignment_list(s: T.AlignmentSet):
    out = []
    for k in range(s.n):
        if not s.found[k]:
            out.append(None)
            continue
        a = s.alns[k]
        runs = [(s.runs[a.runs_offset + x].op, s.runs[a.runs_offset + x].n) for x in range(a.n_runs)]
        out.append((a.t_start, a.q_start, a.t_len, a.q_len, a.score, runs))
    return out


def gapped_extend_batch(anchors, target, query, matrix, params, ctx=None):
    """Batched gapped_extend::extend over anchors [(t_pos, q_pos), ...]; returns
    ([(t_start, q_start, t_len, q_len, score, runs) or None], cells)."""
    ctx = ctx or default_context()
    n = len(anchors)
    arr = (T.Anchor * max(1, n))()
    for i, a in enumerate(anchors):
        arr[i].t_pos, arr[i].q_pos = a[0], a[1]
        if len(a) > 2:
            arr[i].source = T.Hsp(*a[2])
    sp = _P(T.AlignmentSet)()
    ctx.check(lib().gxy_gapped_extend_batch(ctx._h, C.addressof(arr), n, C.byref(target.view()),
                                            C.byref(query.view()), C.byref(matrix), C.byref(params),
                                            C.byref(sp)), "gxy_gapped_extend_batch")
    res = _alignment_list(sp.contents)
    cells = sp.contents.cells
    lib().gxy_alignment_set_free(sp)
    return res, cells


def gapped_extend_pairs_dev(targets, queries, anchors, matrix, params, ctx=None):
    """Batched extend over independent device-resident (target, query, anchor)
    triples; returns (results, cells)."""
    ctx = ctx or default_context()
    n = len(anchors)
    tv = (_vp * max(1, n))(*[t._h for t in targets])
    qv = (_vp * max(1, n))(*[q._h for q in queries])
    arr = (T.Anchor * max(1, n))()
    for i, a in enumerate(anchors):
        arr[i].t_pos, arr[i].q_pos = a[0], a[1]
    sp = _P(T.AlignmentSet)()
    ctx.check(lib().gxy_gapped_extend_pairs_dev(ctx._h, C.addressof(tv), C.addressof(qv),
                                                C.addressof(arr), n, C.byref(matrix),
                                                C.byref(params), C.byref(sp)),
              "gxy_gapped_extend_pairs_dev")
    res = _alignment_list(sp.contents)
    cells = sp.contents.cells
    lib().gxy_alignment_set_free(sp)
    return res, cells


def gapped_extend_slices_dev(target, query, t_off, t_len, q_off, q_len, anchors, matrix, params, ctx=None,
                             raw=False, pick=None):
    """Batched extend over pairs that are slices of two device sequences
    (gxy_gapped_extend_slices_dev); anchors relative to the slices.  Returns
    (results, cells), or with raw=True (found count, cells, {k: result for k
    in pick}) decoding only the alignments listed in `pick`."""
    ctx = ctx or default_context()
    n = len(anchors)
    arrs = [np.ascontiguousarray(a, np.uint64) for a in (t_off, t_len, q_off, q_len)]
    an = np.asarray(anchors, np.uint32).reshape(n, 2) if n else np.zeros((0, 2), np.uint32)
    arr = (T.Anchor * max(1, n))()
    view = np.ctypeslib.as_array(C.cast(arr, _P(C.c_uint32)), (max(1, n) * C.sizeof(T.Anchor) // 4,))
    view = view.reshape(max(1, n), C.sizeof(T.Anchor) // 4)
    view[:n, 0] = an[:, 0]
    view[:n, 1] = an[:, 1]
    sp = _P(T.AlignmentSet)()
    ctx.check(lib().gxy_gapped_extend_slices_dev(ctx._h, target._h, query._h, *[a.ctypes.data for a in arrs],
                                                 C.addressof(arr), n, C.byref(matrix), C.byref(params),
                                                 C.byref(sp)), "gxy_gapped_extend_slices_dev")
    cells = sp.contents.cells
    if raw:
        st = sp.contents
        found = int(np.ctypeslib.as_array(st.found, (n,)).astype(bool).sum()) if n else 0
        picked = {}
        for k in (pick or ()):
            if not st.found[k]:
                picked[k] = None
                continue
            a = st.alns[k]
            runs = [(st.runs[a.runs_offset + x].op, st.runs[a.runs_offset + x].n) for x in range(a.n_runs)]
            picked[k] = (a.t_start, a.q_start, a.t_len, a.q_len, a.score, runs)
        lib().gxy_alignment_set_free(sp)
        return found, cells, picked
    res = _alignment_list(sp.contents)
    lib().gxy_alignment_set_free(sp)
    return res, cells


def gapped_extend(anchor, target, query, matrix, params, ctx=None):
    """gapped_extend::extend (src/gapped_extend.rs:78-84)."""
    res, cells = gapped_extend_batch([anchor], target, query, matrix, params, ctx)
    return res[0], cells


class Records:
    """Owning wrapper of a gxy_record_set."""

    def __init__(self, ptr):
        self.ptr = ptr

    def __len__(self):
        return self.ptr.contents.n

    def tuples(self):
        return T.record_tuples(self.ptr.contents)

    def __del__(self):
        if getattr(self, "ptr", None) and _lib is not None:
            _lib.gxy_record_set_free(self.ptr)
            self.ptr = None


def run(targets, queries, cfg: T.Config, ctx: Context | None = None):
    """driver::run (src/driver.rs:199); targets/queries are [(name, PackedSeq)]."""
    ctx = ctx or default_context()
    ta, tk = T.named_seq_array(targets)
    qa, qk = T.named_seq_array(queries)
    out = _P(T.RecordSet)()
    st = T.RunStats()
    ctx.check(lib().gxy_run(ctx._h, ta, len(targets), qa, len(queries), C.byref(cfg),
                            C.byref(out), C.byref(st)), "gxy_run")
    del tk, qk
    return Records(out), st


def run_dev(targets, queries, cfg: T.Config, ctx: Context | None = None):
    """driver::run over device-resident sequences ([DeviceSeq])."""
    ctx = ctx or default_context()
    tv = (_vp * max(1, len(targets)))(*[t._h for t in targets])
    qv = (_vp * max(1, len(queries)))(*[q._h for q in queries])
    out = _P(T.RecordSet)()
    st = T.RunStats()
    ctx.check(lib().gxy_run_dev(ctx._h, C.addressof(tv), len(targets), C.addressof(qv),
                                len(queries), C.byref(cfg), C.byref(out), C.byref(st)),
              "gxy_run_dev")
    return Records(out), st


def _take_list(lp):
    """[(name, DeviceSeq)] from a gxy_dseq_list; the DeviceSeqs own the sequences."""
    L = lib()
    lst = lp.contents
    out = []
    for k in range(lst.n):
        name = C.string_at(lst.names[k]).decode()
        L.gxy_free(lst.names[k])
        out.append((name, DeviceSeq(C.c_void_p(lst.seqs[k]), default_context())))
    lst.n = 0
    L.gxy_dseq_list_free(lp)
    return out


def load_2bit(data: bytes, ctx=None):
    """load_2bit (src/bit2.rs:73) over a .2bit file's bytes: [(name, DeviceSeq)]."""
    ctx = ctx or default_context()
    lp = _P(T.DSeqList)()
    ctx.check(lib().gxy_load_2bit(ctx._h, data, len(data), C.byref(lp)), "gxy_load_2bit")
    return _take_list(lp)


def load_fasta(text: bytes, ctx=None):
    """load_fasta (src/sequences.rs:338) over FASTA text: [(name, DeviceSeq)]."""
    ctx = ctx or default_context()
    lp = _P(T.DSeqList)()
    ctx.check(lib().gxy_load_fasta(ctx._h, text, len(text), C.byref(lp)), "gxy_load_fasta")
    return _take_list(lp)


def best_chain(hsps):
    """best_chain (src/chain.rs:19); hsps are (t_start, q_start, length, score)."""
    n = len(hsps)
    arr = (T.Hsp * max(1, n))(*[T.Hsp(*h) for h in hsps])
    out = (T.Hsp * max(1, n))()
    m = C.c_uint64()
    rc = lib().gxy_best_chain(C.addressof(arr), n, C.addressof(out), C.byref(m))
    if rc != 0:
        raise GxyError(f"gxy_best_chain failed ({rc})")
    return [(out[k].t_start, out[k].q_start, out[k].length, out[k].score) for k in range(m.value)]


def interpolate(records, target, query, cfg: T.Config, ctx=None):
    """tweener::interpolate (src/tweener.rs:59) over one group's records given as
    (ti, qi, strand, t_start, q_start, t_span, q_span, score, [(op, n), ...]);
    target/query are DeviceSeqs (whole target, the group's query strand)."""
    ctx = ctx or default_context()
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
    ctx.check(lib().gxy_interpolate(ctx._h, C.byref(rs), target._h, query._h, C.byref(cfg), C.byref(out)),
              "gxy_interpolate")
    return Records(out).tuples()


class Server:
    """A serving loop over resident targets (gxy_server): submit PackedSeq
    queries from any thread, wait for each one's records."""

    def __init__(self, targets, cfg: T.Config, co_batch: bool = True, hbm_budget_bytes: int = 0,
                 max_batch_bp: int = 0, dp_workers: int = 0, device: int = 0):
        self._targets = list(targets)  # DeviceSeqs must outlive the server
        tv = (_vp * len(self._targets))(*[t._h for t in self._targets])
        prm = T.ServerParams(hbm_budget_bytes, max_batch_bp, 1 if co_batch else 0, dp_workers)
        h = _vp()
        rc = lib().gxy_server_create(device, C.addressof(tv), len(self._targets), C.byref(cfg), C.byref(prm),
                                     C.byref(h))
        if rc != 0:
            raise GxyError(f"gxy_server_create failed ({rc})")
        self._h = h

    def submit(self, query: T.PackedSeq) -> int:
        t = C.c_uint64()
        v = query.view()
        rc = lib().gxy_server_submit(self._h, C.byref(v), C.byref(t))
        if rc != 0:
            raise GxyError(f"gxy_server_submit failed ({rc})")
        return t.value

    def wait(self, ticket: int):
        """(Records, RequestStats) of a submitted query."""
        out = _P(T.RecordSet)()
        st = T.RequestStats()
        rc = lib().gxy_server_wait(self._h, ticket, C.byref(out), C.byref(st))
        if rc != 0:
            msg = lib().gxy_server_last_error(self._h)
            raise GxyError(f"gxy_server_wait failed ({rc}): {msg.decode() if msg else ''}")
        return Records(out), st

    def set_timing(self, on: bool = True):
        """Kernel timing on the server's contexts (only with no request in flight)."""
        rc = lib().gxy_server_set_timing(self._h, 1 if on else 0)
        if rc != 0:
            raise GxyError(f"gxy_server_set_timing failed ({rc})")

    def stats(self) -> T.RunStats:
        """Work done so far over all requests."""
        st = T.RunStats()
        rc = lib().gxy_server_stats(self._h, C.byref(st))
        if rc != 0:
            raise GxyError(f"gxy_server_stats failed ({rc})")
        return st

    def kernel_time(self, name: str):
        """(milliseconds, launches) of kernel `name` over the server's contexts."""
        ms = C.c_double()
        n = C.c_uint64()
        rc = lib().gxy_server_kernel_time(self._h, name.encode(), C.byref(ms), C.byref(n))
        if rc != 0:
            raise GxyError(f"gxy_server_kernel_time failed ({rc})")
        return ms.value, n.value

    def xdrop_stats(self):
        """Context.xdrop_stats summed over the server's contexts."""
        v = [C.c_uint64() for _ in range(3)]
        rc = lib().gxy_server_xdrop_stats(self._h, *[C.byref(x) for x in v])
        if rc != 0:
            raise GxyError(f"gxy_server_xdrop_stats failed ({rc})")
        return tuple(x.value for x in v)

    def close(self):
        if getattr(self, "_h", None):
            lib().gxy_server_destroy(self._h)
            self._h = None

    def __del__(self):
        if _lib is not None:
            self.close()


def _format(fn, targets, queries, recs: Records) -> bytes:
    ta, tk = T.named_seq_array(targets)
    qa, qk = T.named_seq_array(queries)
    txt = _vp()
    n = C.c_uint64()
    rc = fn(ta, len(targets), qa, len(queries), recs.ptr, C.byref(txt), C.byref(n))
    if rc != 0:
        raise GxyError(f"format failed ({rc})")
    data = C.string_at(txt, n.value)
    lib().gxy_free(txt)
    del tk, qk
    return data


def format_maf(targets, queries, recs: Records) -> bytes:
    """MafWriter (src/output/maf.rs:33-74) with the "HOXD70" header of src/main.rs:42."""
    return _format(lib().gxy_format_maf, targets, queries, recs)


def format_paf(targets, queries, recs: Records) -> bytes:
    """PafWriter (src/output/paf.rs:23-89)."""
    return _format(lib().gxy_format_paf, targets, queries, recs)


def declared_symbols():
    """Function names declared in include/gxy.h."""
    import re
    text = open(HEADER_PATH).read()
    return sorted(set(re.findall(r"\b(gxy_[a-z0-9_]+)\s*\(", text)))
