"""Writes .2bit byte streams (UCSC layout: header, index, records with N and
mask blocks, 4 bases per byte first base in the high bits, T=00 C=01 A=10
G=11) for the loader tests."""
import struct


def build_2bit(seqs, big_endian=False, n_code=2):
    """seqs: [(name, ASCII bytes)]; lowercase -> mask blocks, N -> N blocks."""
    fmt = ">I" if big_endian else "<I"
    u32 = lambda v: struct.pack(fmt, v)
    out = bytearray(u32(0x1A412743) + u32(0) + u32(len(seqs)) + u32(0))
    patches = []
    for name, _ in seqs:
        nb = name.encode()
        out += bytes([len(nb)]) + nb
        patches.append(len(out))
        out += u32(0)
    for (name, bases), at in zip(seqs, patches):
        out[at:at + 4] = u32(len(out))

        def runs(pred):
            r, start = [], None
            for i, b in enumerate(bases):
                if pred(b):
                    if start is None:
                        start = i
                elif start is not None:
                    r.append((start, i - start))
                    start = None
            if start is not None:
                r.append((start, len(bases) - start))
            return r

        nr = runs(lambda b: b in b"Nn")
        mr = runs(lambda b: chr(b).islower() and b not in b"Nn")
        out += u32(len(bases)) + u32(len(nr))
        out += b"".join(u32(s) for s, _ in nr) + b"".join(u32(z) for _, z in nr)
        out += u32(len(mr))
        out += b"".join(u32(s) for s, _ in mr) + b"".join(u32(z) for _, z in mr)
        out += u32(0)
        packed = bytearray((len(bases) + 3) // 4)
        code = {ord("T"): 0, ord("C"): 1, ord("A"): 2, ord("G"): 3}
        for i, b in enumerate(bases):
            c = code.get(ord(chr(b).upper()), n_code)
            packed[i >> 2] |= c << (6 - (i & 3) * 2)
        out += packed
    return bytes(out)
