"""A minimal TIFF 6.0 / BigTIFF + GeoTIFF 1.0 reader for the tests (written
from the file format, independently of the library's writer)."""
import struct

import numpy as np

F = np.float32(np.finfo(np.float32).max)


def read_tiff(path):
    """Minimal TIFF 6.0 / BigTIFF reader: returns (tags dict, raster)."""
    b = open(path, "rb").read()
    assert b[:2] == b"II"
    big = struct.unpack_from("<H", b, 2)[0] == 43
    types = {2: ("s", 1), 3: ("H", 2), 4: ("I", 4), 12: ("d", 8), 16: ("Q", 8)}
    if big:
        off = struct.unpack_from("<Q", b, 8)[0]
        n = struct.unpack_from("<Q", b, off)[0]
        pos, esz, inl = off + 8, 20, 8
    else:
        off = struct.unpack_from("<I", b, 4)[0]
        n = struct.unpack_from("<H", b, off)[0]
        pos, esz, inl = off + 2, 12, 4
    tags = {}
    last = -1
    for i in range(n):
        e = pos + i * esz
        tag, typ = struct.unpack_from("<HH", b, e)
        assert tag > last
        last = tag
        cnt = struct.unpack_from("<Q" if big else "<I", b, e + 4)[0]
        fmt, sz = types[typ]
        at = e + (12 if big else 8)
        if cnt * sz > inl:
            at = struct.unpack_from("<Q" if big else "<I", b, at)[0]
        if typ == 2:
            tags[tag] = b[at:at + cnt].decode()
        else:
            tags[tag] = list(struct.unpack_from("<" + fmt * cnt, b, at))
    W, H = tags[256][0], tags[257][0]
    tw, th = tags[322][0], tags[323][0]
    out = np.empty((H, W), np.float32)
    tx = (W + tw - 1) // tw
    for k, (o, c) in enumerate(zip(tags[324], tags[325])):
        assert c == tw * th * 4
        tile = np.frombuffer(b, np.float32, tw * th, o).reshape(th, tw)
        y0, x0 = (k // tx) * th, (k % tx) * tw
        h, w = min(th, H - y0), min(tw, W - x0)
        out[y0:y0 + h, x0:x0 + w] = tile[:h, :w]
        assert np.all(tile[h:, :] == F) and np.all(tile[:, w:] == F)   # padding holds nodata
    return tags, out


def geokeys(tags):
    d = tags[34735]
    assert d[:3] == [1, 1, 0]
    keys = {}
    for i in range(d[3]):
        k, loc, cnt, val = d[4 + 4 * i: 8 + 4 * i]
        keys[k] = tags[loc][val:val + cnt] if loc == 34737 else val
    return keys
