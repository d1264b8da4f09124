"""A minimal OpenEXR reader for the tests: single-part scanline files with
NO_COMPRESSION (0), ZIPS (2) or ZIP (3) blocks and HALF / FLOAT channels,
written from the published file format (magic, header attributes, offset
table, scanline blocks; ZIP blocks are deflated after splitting even/odd
bytes and delta-coding them).  Independent of the writer under test."""
import struct
import zlib

import numpy as np

_TYPES = {1: np.float16, 2: np.float32, 0: np.uint32}


def _cstr(buf, pos):
    end = buf.index(b"\0", pos)
    return buf[pos:end].decode(), end + 1


def read_exr(path):
    """-> (header dict, {channel name: (H, W) array})"""
    buf = open(path, "rb").read()
    assert buf[:4] == bytes([0x76, 0x2F, 0x31, 0x01]), "not an EXR file"
    version = struct.unpack_from("<i", buf, 4)[0]
    assert version & 0xFF == 2 and not (version & 0x200), "only single-part scanline files"
    pos = 8
    header = {}
    while buf[pos] != 0:
        name, pos = _cstr(buf, pos)
        typ, pos = _cstr(buf, pos)
        size = struct.unpack_from("<i", buf, pos)[0]
        pos += 4
        header[name] = (typ, buf[pos:pos + size])
        pos += size
    pos += 1
    chans = []
    raw = header["channels"][1]
    p = 0
    while raw[p] != 0:
        name, p = _cstr(raw, p)
        ptype, _plin, xs, ys = struct.unpack_from("<iB3xii", raw, p)
        p += 16
        chans.append((name, ptype))
    x0, y0, x1, y1 = struct.unpack("<4i", header["dataWindow"][1])
    w, h = x1 - x0 + 1, y1 - y0 + 1
    comp = header["compression"][1][0]
    lines = {0: 1, 2: 1, 3: 16}[comp]
    nblocks = (h + lines - 1) // lines
    offsets = struct.unpack_from(f"<{nblocks}Q", buf, pos)
    out = {n: np.zeros((h, w), dtype=_TYPES[t]) for n, t in chans}
    for off in offsets:
        yb, size = struct.unpack_from("<ii", buf, off)
        data = buf[off + 8:off + 8 + size]
        nl = min(lines, y0 + h - yb)
        expect = sum(w * (2 if t == 1 else 4) for _, t in chans) * nl
        if comp != 0 and size < expect:
            d = np.frombuffer(zlib.decompress(data), dtype=np.uint8).astype(np.int64)
            # undo the delta coding: b[0] = d[0], b[i] = b[i-1] + d[i] - 128 (mod 256)
            d[1:] -= 128
            b = (np.cumsum(d) % 256).astype(np.uint8)
            half = (b.size + 1) // 2
            inter = np.empty_like(b)
            inter[0::2] = b[:half]
            inter[1::2] = b[half:]
            data = inter.tobytes()
        assert len(data) == expect, (len(data), expect)
        q = 0
        for yy in range(nl):
            for n, t in chans:
                nb = w * (2 if t == 1 else 4)
                out[n][yb - y0 + yy] = np.frombuffer(data[q:q + nb], dtype=_TYPES[t])
                q += nb
    return header, out
