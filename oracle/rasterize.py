"""GDAL's default rasterisation, restated -- TEST INFRASTRUCTURE ONLY.

Only tests/ may import this module, as the checker of the library's
rasteriser (pg_rasterize_times / pg_rasterize_fuels).

What it restates.  The reference burns its ignition features
(crates/propag/src/propag.rs:837-890, ``rasterize_times``) and fuel features
(propag.rs:929-978, ``rasterize_fuels``) with ``GDALRasterizeGeometries``
into a one-band MEM dataset carrying the raster's geotransform, with no
options (ALL_TOUCHED off, MERGE_ALG=REPLACE), no transformer and one burn
value per feature, after filling the band with the nodata value (FLT_MAX for
times, 255 for fuel codes).  GDAL is a third-party dependency absent from the
reference checkout: the ``gdal`` 0.17.1 and ``gdal-sys`` 0.10.0 crates
(crates/Cargo.lock:548-567) bind the system libgdal, whose version the lock
does not pin.  The rules below are those of GDAL 3.x ``alg/gdalrasterize.cpp``
(``GDALRasterizeGeometries``, ``gv_rasterize_one_shape``,
``GDALCollectRingsFromGeometry``) and ``alg/llrasterize.cpp``
(``GDALdllImagePoint``, ``GDALdllImageLine``, ``GDALdllImageFilledPolygon``):

* pixel/line coordinates: with no transformer GDAL builds a GenImgProj
  transformer from the dataset's geotransform, i.e. applies the inverse
  geotransform of ``GDALInvGeoTransform`` in double precision (no-rotation
  special case, else the adjoint over the determinant);
* a geometry is flattened into parts (a point, a line string, each ring of a
  polygon; members of multi-geometries and collections in order) and
  dispatched on its own type: points and multipoints burn the pixel
  (floor(x), floor(y)); line strings and multi line strings burn an integer
  Bresenham line between the floored vertices of each segment; everything
  else (polygons, multipolygons, collections) is scan-filled;
* scan fill: for every raster row y (from int(min y) to int(max y), clipped)
  the scanline at y + 0.5 intersects each non-horizontal edge whose lower end
  is at or below it and whose upper end is above it; the crossings x become
  floor(x + 0.5), are sorted, and each pair [a, b) of them burns pixels a to
  b - 1; a horizontal edge lying exactly on the scanline burns
  [floor(x_lo + 0.5), floor(x_hi + 0.5)) when it runs right to left and
  nothing otherwise;
* burns are clipped to the raster; features burn in order, later ones
  replacing earlier ones; an empty geometry burns nothing.

Beyond GDAL: a feature with a vertex that is not finite or lies 2^30 pixels
or more from the origin burns nothing (GDAL's integer casts are undefined
there).

Invalid WKB is skipped without error, as the reference's ``filter_map`` of
``TryFrom`` does (propag.rs:208-215, 902-907).

Parity to GDAL itself is unpinned: no GDAL is installed here and the
reference holds no rasterisation fixture.  The restatement is pinned by the
tests against what the mathematics fixes: the pixel that contains a point,
the pixels whose centres lie inside a polygon (even-odd rule) when no centre
lies on an edge, and the defining properties of a Bresenham line.
"""
from __future__ import annotations

import math
import struct

import numpy as np

FLT_MAX = float(np.finfo(np.float32).max)

POINT, LINESTRING, POLYGON, MULTIPOINT, MULTILINESTRING, MULTIPOLYGON, COLLECTION = range(1, 8)


# ------------------------------------------------------------------ WKB

class WKBError(ValueError):
    pass


def _read_geom(buf: bytes, off: int, depth: int = 0):
    """Returns (flat type, parts, new offset); parts is a list of
    (xs, ys) coordinate lists."""
    if depth > 32:
        raise WKBError("nesting too deep")
    if off + 5 > len(buf):
        raise WKBError("truncated header")
    order = buf[off]
    if order not in (0, 1):
        raise WKBError("bad byte order")
    e = "<" if order == 1 else ">"
    (code,) = struct.unpack_from(e + "I", buf, off + 1)
    off += 5
    # OGC 2.5D (0x80000000) and M (0x40000000) flags, or ISO codes +1000 (Z),
    # +2000 (M), +3000 (ZM); the EWKB SRID flag and mixed forms are rejected.
    flag_z, flag_m = bool(code & 0x80000000), bool(code & 0x40000000)
    if code & 0x30000000:
        raise WKBError("not OGC/ISO WKB")
    code &= 0x0FFFFFFF
    if (flag_z or flag_m) and code >= 1000:
        raise WKBError("mixed dimension flags")
    if code >= 4000:
        raise WKBError("unknown dimension")
    base = code % 1000
    dims = 2 + flag_z + flag_m + (0, 1, 1, 2)[code // 1000]

    def count(o):
        if o + 4 > len(buf):
            raise WKBError("truncated count")
        return struct.unpack_from(e + "I", buf, o)[0], o + 4

    def points(n, o):
        need = n * dims * 8
        if o + need > len(buf) or n > len(buf):
            raise WKBError("truncated points")
        vals = struct.unpack_from(e + "d" * (n * dims), buf, o)
        xs = [vals[i * dims] for i in range(n)]
        ys = [vals[i * dims + 1] for i in range(n)]
        return xs, ys, o + need

    if base == POINT:
        xs, ys, off = points(1, off)
        if math.isnan(xs[0]) and math.isnan(ys[0]):
            return POINT, [], off           # POINT EMPTY
        return POINT, [(xs, ys)], off
    if base == LINESTRING:
        n, off = count(off)
        xs, ys, off = points(n, off)
        return LINESTRING, ([(xs, ys)] if n else []), off
    if base == POLYGON:
        nr, off = count(off)
        parts = []
        for _ in range(nr):
            n, off = count(off)
            xs, ys, off = points(n, off)
            if n:
                parts.append((xs, ys))
        return POLYGON, parts, off
    if base in (MULTIPOINT, MULTILINESTRING, MULTIPOLYGON, COLLECTION):
        ng, off = count(off)
        want = {MULTIPOINT: POINT, MULTILINESTRING: LINESTRING, MULTIPOLYGON: POLYGON}.get(base)
        parts = []
        for _ in range(ng):
            t, p, off = _read_geom(buf, off, depth + 1)
            if want is not None and t != want:
                raise WKBError("member of the wrong type")
            parts.extend(p)
        return base, parts, off
    raise WKBError(f"unsupported geometry type {code}")


def parse_wkb(buf: bytes):
    t, parts, off = _read_geom(bytes(buf), 0)
    return t, parts


def wkb(kind: int, coords, big_endian: bool = False) -> bytes:
    """Encodes 2-D WKB: POINT (x, y); LINESTRING [(x, y), ...]; POLYGON
    [ring, ...]; MULTI*: [member coords, ...]; COLLECTION: [(kind, coords)]."""
    e = ">" if big_endian else "<"
    head = bytes([0 if big_endian else 1]) + struct.pack(e + "I", kind)
    if kind == POINT:
        return head + struct.pack(e + "dd", *coords)
    if kind == LINESTRING:
        return head + struct.pack(e + "I", len(coords)) + b"".join(struct.pack(e + "dd", *p) for p in coords)
    if kind == POLYGON:
        out = head + struct.pack(e + "I", len(coords))
        for ring in coords:
            out += struct.pack(e + "I", len(ring)) + b"".join(struct.pack(e + "dd", *p) for p in ring)
        return out
    if kind in (MULTIPOINT, MULTILINESTRING, MULTIPOLYGON):
        sub = {MULTIPOINT: POINT, MULTILINESTRING: LINESTRING, MULTIPOLYGON: POLYGON}[kind]
        return head + struct.pack(e + "I", len(coords)) + b"".join(wkb(sub, c, big_endian) for c in coords)
    if kind == COLLECTION:
        return head + struct.pack(e + "I", len(coords)) + b"".join(wkb(k, c, big_endian) for k, c in coords)
    raise ValueError(kind)


# ------------------------------------------------------- geotransform

def inv_geotransform(gt):
    """GDALInvGeoTransform in double (gt as GDAL orders it)."""
    g = [float(v) for v in gt]
    if g[2] == 0.0 and g[4] == 0.0 and g[1] != 0.0 and g[5] != 0.0:
        return [-g[0] / g[1], 1.0 / g[1], 0.0, -g[3] / g[5], 0.0, 1.0 / g[5]]
    det = g[1] * g[5] - g[2] * g[4]
    mag = max(abs(g[1]), abs(g[2]), abs(g[4]), abs(g[5]))
    if abs(det) <= 1e-10 * mag * mag:
        return None
    inv = 1.0 / det
    return [(g[2] * g[3] - g[0] * g[5]) * inv, g[5] * inv, -g[2] * inv,
            (-g[1] * g[3] + g[0] * g[4]) * inv, -g[4] * inv, g[1] * inv]


def to_pixel(inv, x, y):
    return inv[0] + x * inv[1] + y * inv[2], inv[3] + x * inv[4] + y * inv[5]


# ------------------------------------------------------------ burning

def _burn_points(out, parts, v):
    H, W = out.shape
    for xs, ys in parts:
        x, y = math.floor(xs[0]), math.floor(ys[0])
        if 0 <= x < W and 0 <= y < H:
            out[y, x] = v


def _burn_lines(out, parts, v):
    H, W = out.shape
    for xs, ys in parts:
        for j in range(1, len(xs)):
            ix, iy = math.floor(xs[j - 1]), math.floor(ys[j - 1])
            ix1, iy1 = math.floor(xs[j]), math.floor(ys[j])
            ddx, ddy = abs(ix1 - ix), abs(iy1 - iy)
            sx = -1 if ix > ix1 else 1
            sy = -1 if iy > iy1 else 1
            major, minor = (ddx, ddy) if ddx >= ddy else (ddy, ddx)
            e_minor = minor << 1
            e_major = e_minor - (major << 1)
            err = e_minor - major
            for _ in range(major + 1):
                if 0 <= ix < W and 0 <= iy < H:
                    out[iy, ix] = v
                if ddx >= ddy:
                    ix += sx
                    if err > 0:
                        iy += sy
                        err += e_major
                    else:
                        err += e_minor
                else:
                    iy += sy
                    if err > 0:
                        ix += sx
                        err += e_major
                    else:
                        err += e_minor


def _burn_span(out, y, a, b, v):
    W = out.shape[1]
    a, b = max(a, 0), min(b, W - 1)
    if a <= b:
        out[y, a:b + 1] = v


def _burn_polygon(out, parts, v):
    H, W = out.shape
    xs_all = [x for xs, _ in parts for x in xs]
    ys_all = [y for _, ys in parts for y in ys]
    if not ys_all:
        return
    miny, maxy = int(min(ys_all)), int(max(ys_all))
    miny = max(miny, 0)
    maxy = min(maxy, H - 1)
    maxx = W - 1
    for y in range(miny, maxy + 1):
        dy = y + 0.5
        ints = []
        for xs, ys in parts:
            n = len(xs)
            for i in range(n):
                i1, i2 = (n - 1, 0) if i == 0 else (i - 1, i)
                y1, y2 = ys[i1], ys[i2]
                if (y1 < dy and y2 < dy) or (y1 > dy and y2 > dy):
                    continue
                if y1 < y2:
                    x1, x2 = xs[i1], xs[i2]
                elif y1 > y2:
                    y1, y2 = y2, y1
                    x1, x2 = xs[i2], xs[i1]
                else:
                    if xs[i1] > xs[i2]:
                        a = math.floor(xs[i2] + 0.5)
                        b = math.floor(xs[i1] + 0.5)
                        if not (a > maxx or b <= 0):
                            _burn_span(out, y, a, b - 1, v)
                    continue
                if y1 <= dy < y2:
                    ints.append(math.floor((dy - y1) * (x2 - x1) / (y2 - y1) + x1 + 0.5))
        ints.sort()
        for k in range(0, len(ints) - 1, 2):
            if ints[k] <= maxx and ints[k + 1] > 0:
                _burn_span(out, y, ints[k], ints[k + 1] - 1, v)


def rasterize(features, gt, width: int, height: int, fill, dtype):
    """features: iterable of (burn value, WKB bytes); gt: GDAL-ordered
    geotransform (the f32 GeoReference values widened to double)."""
    out = np.full((height, width), fill, dtype=dtype)
    inv = inv_geotransform(gt)
    if inv is None:
        raise ValueError("geotransform is not invertible")
    for value, blob in features:
        try:
            t, parts = parse_wkb(blob)
        except (WKBError, struct.error):
            continue
        px = []
        for xs, ys in parts:
            pts = [to_pixel(inv, x, y) for x, y in zip(xs, ys)]
            px.append(([p[0] for p in pts], [p[1] for p in pts]))
        if not all(abs(c) < 2.0 ** 30 for xs, ys in px for c in xs + ys):
            continue
        v = np.dtype(dtype).type(value)
        if t in (POINT, MULTIPOINT):
            _burn_points(out, px, v)
        elif t in (LINESTRING, MULTILINESTRING):
            _burn_lines(out, px, v)
        else:
            _burn_polygon(out, px, v)
    return out


def init_refs(time):
    """refs = own cell where time holds a value, 0xFFFF elsewhere
    (propag.rs:433-449)."""
    H, W = time.shape
    rx = np.full((H, W), 0xFFFF, np.uint16)
    ry = np.full((H, W), 0xFFFF, np.uint16)
    ys, xs = np.nonzero(time != np.float32(FLT_MAX))
    rx[ys, xs] = xs
    ry[ys, xs] = ys
    return rx, ry
