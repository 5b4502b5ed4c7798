"""The C-ABI library loads and exports every function include/gxy.h declares
(no compute calls: runs without a GPU)."""
import ctypes
import os

import gxy


def test_header_symbols_exported():
    names = gxy.declared_symbols()
    assert len(names) >= 30
    lib = ctypes.CDLL(gxy.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert missing == []


def test_python_signatures_cover_header():
    assert sorted(gxy._SIGS) == gxy.declared_symbols()


def test_pack_ascii_matches_numpy_packer():
    import numpy as np
    import gxy_types as T
    text = b"ACGTacgtNNnRYACGUu-" * 11
    ps = T.PackedSeq.from_ascii(text)
    codes = np.zeros(len(ps.codes), np.uint8)
    valid = np.zeros(len(ps.valid), np.uint8)
    mask = np.zeros(len(ps.mask), np.uint8)
    assert gxy.lib().gxy_pack_ascii(text, len(text), codes.ctypes.data, valid.ctypes.data,
                                    mask.ctypes.data) == 0
    assert codes.tolist() == ps.codes.tolist()
    assert valid.tolist() == ps.valid.tolist()
    assert mask.tolist() == ps.mask.tolist()


def test_config_default_and_pattern_parse():
    import gxy_types as T
    c = T.Config()
    gxy.lib().gxy_config_default(ctypes.byref(c))
    ref = T.config_default()
    assert bytes(c) == bytes(ref)
    for spec in ("12of19", "19of20", "match12", "1101"):
        p = T.SeedPatternC()
        assert gxy.lib().gxy_pattern_parse(spec.encode(), ctypes.byref(p)) == 0
        r = T.pattern_from_spec(spec)
        assert (p.length, p.care) == (r.length, r.care)
    p = T.SeedPatternC()
    assert gxy.lib().gxy_pattern_parse(b"deadbeef", ctypes.byref(p)) != 0


def test_context_without_gpu_fails_cleanly():
    if os.path.exists("/dev/kfd"):
        return
    h = ctypes.c_void_p()
    assert gxy.lib().gxy_ctx_create(0, ctypes.byref(h)) != 0
    assert b"device" in gxy.lib().gxy_last_error(None)


def test_hsp_record_array_matches_c_layout():
    """search_array copies the library's gxy_hsp buffer byte for byte into a
    numpy record array: the two layouts must agree field by field."""
    import gxy_types as T
    dt = gxy.HSP_DTYPE
    assert dt.itemsize == ctypes.sizeof(T.Hsp)
    for name, ctype in T.Hsp._fields_:
        assert dt.fields[name][1] == getattr(T.Hsp, name).offset
        assert dt.fields[name][0].itemsize == ctypes.sizeof(ctype)
    h = T.Hsp(7, 11, 19, -3000)
    a = gxy.np.frombuffer(bytes(h), dt)
    assert (int(a["t_start"][0]), int(a["q_start"][0]), int(a["length"][0]), int(a["score"][0])) == \
        (7, 11, 19, -3000)
