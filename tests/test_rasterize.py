"""Ignition / fuel rasterisation and the initial state (CPU; no GPU calls).

The reference burns WKB features with GDALRasterizeGeometries
(crates/propag/src/propag.rs:837-890, 929-978).  GDAL is not installed here
and the reference holds no rasterisation fixture, so parity with GDAL itself
is unpinned; oracle/rasterize.py restates GDAL's default rules.  This file:

* pins that restatement against what the mathematics fixes, independently of
  how the rules are written: the pixel containing a point; the pixels whose
  centres lie inside a polygon by the even-odd rule (no centre on an edge);
  Bresenham's closed form y_k = ceil((2 k dy - dx) / (2 dx)) for a line
  between floored vertices; GDAL's inverse geotransform against a direct
  linear solve;
* checks the library (pg_rasterize_times / pg_rasterize_fuels /
  pg_init_refs through py_propag) against the oracle bit for bit on seeded
  random features, including vertices on pixel edges and centres (the tie
  cases), every WKB encoding and invalid WKB;
* checks that WKB ignitions give the same initial state as raster cells.
"""
import math
import struct

import numpy as np
import pytest

from oracle import rasterize as R

F = R.FLT_MAX


def _geo(width, height, gt):
    import py_propag as pp

    g = pp.GeoReference(width, height, "+proj=utm +zone=30 +datum=WGS84 +units=m +no_defs", gt)
    return g, [float(v) for v in g.transform]


# ------------------------------------------------- the oracle, pinned

def _even_odd(px, py, rings):
    inside = False
    for ring in rings:
        n = len(ring)
        for i in range(n):
            (x1, y1), (x2, y2) = ring[i - 1], ring[i]
            if (y1 > py) != (y2 > py):
                xc = x1 + (py - y1) * (x2 - x1) / (y2 - y1)
                if xc > px:
                    inside = not inside
    return inside


def _edge_dist(px, py, rings):
    best = math.inf
    for ring in rings:
        for i in range(len(ring)):
            (x1, y1), (x2, y2) = ring[i - 1], ring[i]
            vx, vy = x2 - x1, y2 - y1
            L = vx * vx + vy * vy
            t = 0.0 if L == 0 else max(0.0, min(1.0, ((px - x1) * vx + (py - y1) * vy) / L))
            best = min(best, math.hypot(px - x1 - t * vx, py - y1 - t * vy))
    return best


@pytest.mark.parametrize("north_up", [False, True])
def test_oracle_points_burn_the_pixel_that_contains_them(north_up):
    r = np.random.Generator(np.random.PCG64(1))
    W, H = 40, 30
    gt = [500.0, 5.0, 0.0, 8000.0 if north_up else 2000.0, 0.0, -5.0 if north_up else 5.0]
    for _ in range(300):
        i, j = int(r.integers(-3, W + 3)), int(r.integers(-3, H + 3))
        fx, fy = r.uniform(0.01, 0.99, 2)
        X = gt[0] + (i + fx) * gt[1]
        Y = gt[3] + (j + fy) * gt[5]
        out = R.rasterize([(1.0, R.wkb(R.POINT, (X, Y)))], gt, W, H, F, np.float32)
        want = np.full((H, W), F, np.float32)
        if 0 <= i < W and 0 <= j < H:
            want[j, i] = 1.0
        assert np.array_equal(out, want), (i, j, fx, fy)


def _random_ring(r, cx, cy, rad, n):
    ang = np.sort(r.uniform(0, 2 * np.pi, n))
    rr = r.uniform(0.3, 1.0, n) * rad
    pts = [(float(cx + a * math.cos(t)), float(cy + a * math.sin(t))) for a, t in zip(rr, ang)]
    return pts + [pts[0]]


@pytest.mark.parametrize("seed", range(6))
def test_oracle_polygons_burn_the_pixel_centres_inside(seed):
    """Even-odd rule at pixel centres, with holes and self-intersections."""
    r = np.random.Generator(np.random.PCG64(100 + seed))
    W, H = 48, 40
    gt = [0.0, 1.0, 0.0, 0.0, 0.0, 1.0]  # world == pixel coordinates
    rings = [_random_ring(r, r.uniform(10, 38), r.uniform(10, 30), r.uniform(6, 16), int(r.integers(3, 12)))]
    if seed % 2:
        cx, cy = np.mean([p[0] for p in rings[0][:-1]]), np.mean([p[1] for p in rings[0][:-1]])
        rings.append(_random_ring(r, cx, cy, 3.0, 5))       # a hole (or an island, by even-odd)
    if seed == 4:
        rings = [[(5.3, 5.2), (30.7, 25.9), (30.2, 5.6), (5.1, 25.4), (5.3, 5.2)]]  # bow tie
    if seed == 5:
        rings = [[(-6.2, -3.7), (60.3, 10.1), (20.4, 52.8), (-6.2, -3.7)]]         # beyond the raster
    out = R.rasterize([(2.0, R.wkb(R.POLYGON, rings))], gt, W, H, F, np.float32)
    checked = 0
    for y in range(H):
        for x in range(W):
            c = (x + 0.5, y + 0.5)
            if _edge_dist(*c, rings) < 1e-9:
                continue
            assert (out[y, x] == 2.0) == _even_odd(*c, rings), (seed, x, y)
            checked += 1
    assert checked > W * H - 5


def _bresenham_closed_form(x0, y0, x1, y1):
    dx, dy = abs(x1 - x0), abs(y1 - y0)
    sx, sy = (1 if x1 >= x0 else -1), (1 if y1 >= y0 else -1)
    if dx >= dy:
        return {(x0 + sx * k, y0 + sy * (-((-(2 * k * dy - dx)) // (2 * dx)) if dx else 0)) for k in range(dx + 1)}
    return {(x0 + sx * (-((-(2 * k * dx - dy)) // (2 * dy))), y0 + sy * k) for k in range(dy + 1)}


def test_oracle_lines_follow_bresenhams_closed_form():
    r = np.random.Generator(np.random.PCG64(7))
    W, H = 64, 64
    gt = [0.0, 1.0, 0.0, 0.0, 0.0, 1.0]
    for _ in range(200):
        n = int(r.integers(2, 5))
        pts = [(float(v[0]), float(v[1])) for v in r.uniform(-8, 72, (n, 2))]
        out = R.rasterize([(1.0, R.wkb(R.LINESTRING, pts))], gt, W, H, F, np.float32)
        want = set()
        for a, b in zip(pts, pts[1:]):
            want |= _bresenham_closed_form(math.floor(a[0]), math.floor(a[1]), math.floor(b[0]), math.floor(b[1]))
        want = {(x, y) for x, y in want if 0 <= x < W and 0 <= y < H}
        got = {(int(x), int(y)) for y, x in zip(*np.nonzero(out == 1.0))}
        assert got == want, pts
    # axis-aligned and diagonal lines burn exactly their cells
    out = R.rasterize([(1.0, R.wkb(R.LINESTRING, [(2.5, 3.5), (9.5, 3.5)])),
                       (2.0, R.wkb(R.LINESTRING, [(20.2, 1.1), (20.7, 9.9)])),
                       (3.0, R.wkb(R.LINESTRING, [(30.5, 30.5), (40.5, 40.5)]))], gt, W, H, F, np.float32)
    assert {(x, y) for y, x in zip(*np.nonzero(out == 1))} == {(x, 3) for x in range(2, 10)}
    assert {(x, y) for y, x in zip(*np.nonzero(out == 2))} == {(20, y) for y in range(1, 10)}
    assert {(x, y) for y, x in zip(*np.nonzero(out == 3))} == {(k, k) for k in range(30, 41)}


def test_oracle_inverse_geotransform_matches_a_linear_solve():
    r = np.random.Generator(np.random.PCG64(3))
    for _ in range(50):
        gt = [float(r.uniform(-1e5, 1e5)), float(r.uniform(0.5, 30)), float(r.uniform(-3, 3)),
              float(r.uniform(-1e5, 1e5)), float(r.uniform(-3, 3)), float(-r.uniform(0.5, 30))]
        inv = R.inv_geotransform(gt)
        A = np.array([[gt[1], gt[2]], [gt[4], gt[5]]])
        for _ in range(5):
            p = r.uniform(-100, 100, 2)
            X, Y = gt[0] + p[0] * gt[1] + p[1] * gt[2], gt[3] + p[0] * gt[4] + p[1] * gt[5]
            q = np.linalg.solve(A, [X - gt[0], Y - gt[3]])
            assert np.allclose(R.to_pixel(inv, X, Y), q, atol=1e-6)
    assert R.inv_geotransform([0, 1, 1, 0, 1, 1]) is None


def test_oracle_later_features_replace_earlier_ones():
    gt = [0.0, 1.0, 0.0, 0.0, 0.0, 1.0]
    sq = lambda a, b: [[(a, a), (b, a), (b, b), (a, b), (a, a)]]
    out = R.rasterize([(1.0, R.wkb(R.POLYGON, sq(0.2, 6.2))), (2.0, R.wkb(R.POLYGON, sq(3.2, 9.2)))],
                      gt, 10, 10, F, np.float32)
    assert out[0, 0] == 1 and out[4, 4] == 2 and out[8, 8] == 2 and out[9, 0] == F and out[0, 9] == F
    assert (out == 1).sum() == 36 - 9 and (out == 2).sum() == 36


# ------------------------------------------- the library vs the oracle

def _rand_coord(r, lo, hi):
    """Mostly random, sometimes exactly on a pixel edge or centre."""
    u = r.random()
    if u < 0.25:
        return float(r.integers(lo, hi))
    if u < 0.5:
        return float(r.integers(lo, hi)) + 0.5
    return float(r.uniform(lo, hi))


def _rand_feature(r, W, H):
    k = int(r.integers(0, 8))
    pt = lambda: (_rand_coord(r, -3, W + 3), _rand_coord(r, -3, H + 3))
    if k == 0:
        return R.wkb(R.POINT, pt())
    if k == 1:
        return R.wkb(R.MULTIPOINT, [pt() for _ in range(int(r.integers(1, 5)))])
    if k == 2:
        return R.wkb(R.LINESTRING, [pt() for _ in range(int(r.integers(2, 6)))])
    if k == 3:
        return R.wkb(R.MULTILINESTRING, [[pt() for _ in range(3)] for _ in range(2)])
    ring = lambda: (lambda p: p + [p[0]])([pt() for _ in range(int(r.integers(3, 8)))])
    if k == 4:
        return R.wkb(R.POLYGON, [ring()])
    if k == 5:
        a, b = sorted((_rand_coord(r, -2, W + 2), _rand_coord(r, -2, W + 2)))
        c, d = sorted((_rand_coord(r, -2, H + 2), _rand_coord(r, -2, H + 2)))
        box = [(a, c), (b, c), (b, d), (a, d), (a, c)]
        return R.wkb(R.POLYGON, [box if r.random() < 0.5 else box[::-1], ring()])
    if k == 6:
        return R.wkb(R.MULTIPOLYGON, [[ring()], [ring()]])
    return R.wkb(R.COLLECTION, [(R.POLYGON, [ring()]), (R.POINT, pt())])


GTS = {
    "south_up": [1000.0, 5.0, 0.0, 3000.0, 0.0, 5.0],
    "north_up": [1000.0, 5.0, 0.0, 3300.0, 0.0, -5.0],
    "rotated": [1000.0, 4.0, 1.5, 3000.0, -1.0, 4.5],
}


def _world(gt, pts):
    return [(gt[0] + x * gt[1] + y * gt[2], gt[3] + x * gt[4] + y * gt[5]) for x, y in pts]


@pytest.mark.parametrize("name", list(GTS))
def test_library_matches_oracle_on_random_features(name):
    import py_propag as pp

    r = np.random.Generator(np.random.PCG64(11))
    W, H = 37, 29
    g, gt = _geo(W, H, GTS[name])
    for trial in range(25):
        feats = []
        for i in range(int(r.integers(1, 7))):
            blob = _rand_feature(r, W, H)
            t, parts = R.parse_wkb(blob)
            # re-encode the pixel-space coordinates in world coordinates
            if name != "rotated" or trial % 2:
                blob = _reencode_world(blob, gt)
            feats.append((float(r.uniform(0, 3600)), blob))
        got, n = pp.rasterize_times(g, feats)
        want = R.rasterize(feats, gt, W, H, F, np.float32)
        assert n == len(feats)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (name, trial)
        fuels = [(int(r.integers(0, 14)), b) for _, b in feats]
        got8, _ = pp.rasterize_fuels(g, fuels)
        assert np.array_equal(got8, R.rasterize(fuels, gt, W, H, 255, np.uint8)), (name, trial)


def _reencode_world(blob, gt):
    """Maps every vertex of a 2-D little-endian WKB blob from pixel to world
    coordinates (same structure, new doubles)."""
    out = bytearray(blob)

    def walk(off):
        code = struct.unpack_from("<I", out, off + 1)[0]
        off += 5
        if code == R.POINT:
            x, y = struct.unpack_from("<dd", out, off)
            struct.pack_into("<dd", out, off, *_world(gt, [(x, y)])[0])
            return off + 16
        if code == R.LINESTRING:
            n = struct.unpack_from("<I", out, off)[0]
            off += 4
            for i in range(n):
                x, y = struct.unpack_from("<dd", out, off)
                struct.pack_into("<dd", out, off, *_world(gt, [(x, y)])[0])
                off += 16
            return off
        if code == R.POLYGON:
            nr = struct.unpack_from("<I", out, off)[0]
            off += 4
            for _ in range(nr):
                n = struct.unpack_from("<I", out, off)[0]
                off += 4
                for i in range(n):
                    x, y = struct.unpack_from("<dd", out, off)
                    struct.pack_into("<dd", out, off, *_world(gt, [(x, y)])[0])
                    off += 16
            return off
        ng = struct.unpack_from("<I", out, off)[0]
        off += 4
        for _ in range(ng):
            off = walk(off)
        return off

    walk(0)
    return bytes(out)


def _recode(blob, big_endian=False, dims=""):
    """The same geometry in another WKB encoding: byte order and Z/M/ZM
    (ISO codes, or the OGC 2.5D flag for 'Z25')."""
    t, _ = R.parse_wkb(blob)
    e = ">" if big_endian else "<"
    src = memoryview(blob)

    def code_of(base):
        if dims == "Z25":
            return base | 0x80000000
        return base + {"": 0, "Z": 1000, "M": 2000, "ZM": 3000}[dims]

    extra = {"": 0, "Z": 1, "M": 1, "ZM": 2, "Z25": 1}[dims]

    def walk(off):
        code = struct.unpack_from("<I", src, off + 1)[0]
        out = bytes([0 if big_endian else 1]) + struct.pack(e + "I", code_of(code))
        off += 5

        def pts(n, off):
            o = b""
            for _ in range(n):
                x, y = struct.unpack_from("<dd", src, off)
                o += struct.pack(e + "d" * (2 + extra), x, y, *([7.25] * extra))
                off += 16
            return o, off

        if code == R.POINT:
            o, off = pts(1, off)
            return out + o, off
        if code == R.LINESTRING:
            n = struct.unpack_from("<I", src, off)[0]
            o, off = pts(n, off + 4)
            return out + struct.pack(e + "I", n) + o, off
        if code == R.POLYGON:
            nr = struct.unpack_from("<I", src, off)[0]
            out += struct.pack(e + "I", nr)
            off += 4
            for _ in range(nr):
                n = struct.unpack_from("<I", src, off)[0]
                o, off = pts(n, off + 4)
                out += struct.pack(e + "I", n) + o
            return out, off
        ng = struct.unpack_from("<I", src, off)[0]
        out += struct.pack(e + "I", ng)
        off += 4
        for _ in range(ng):
            o, off = walk(off)
            out += o
        return out, off

    return walk(0)[0]


def test_every_wkb_encoding_burns_the_same_raster():
    import py_propag as pp

    r = np.random.Generator(np.random.PCG64(5))
    W, H = 24, 20
    g, gt = _geo(W, H, [0.0, 1.0, 0.0, 0.0, 0.0, 1.0])
    feats = [(float(i + 1), _rand_feature(r, W, H)) for i in range(12)]
    base, _ = pp.rasterize_times(g, feats)
    for be in (False, True):
        for dims in ("", "Z", "M", "ZM", "Z25"):
            alt = [(t, _recode(b, be, dims)) for t, b in feats]
            got, n = pp.rasterize_times(g, alt)
            assert n == len(feats) and np.array_equal(got, base), (be, dims)
            assert np.array_equal(R.rasterize(alt, gt, W, H, F, np.float32), base), (be, dims)


def test_invalid_wkb_is_skipped_like_the_reference():
    import py_propag as pp

    W, H = 8, 8
    g, gt = _geo(W, H, [0.0, 1.0, 0.0, 0.0, 0.0, 1.0])
    good = R.wkb(R.POINT, (2.5, 3.5))
    srid = bytes([1]) + struct.pack("<II", 0x20000001, 4326) + struct.pack("<dd", 1.5, 1.5)
    bad = [b"", b"\x01", bytes([2]) + good[1:], good[:-1], srid,
           bytes([1]) + struct.pack("<I", 99) + good[5:],                      # unknown type
           R.wkb(R.MULTIPOINT, [(1.0, 1.0)])[:9] + R.wkb(R.LINESTRING, [(1, 1), (2, 2)]),  # wrong member
           R.wkb(R.LINESTRING, [(0.5, 0.5), (5.5, 5.5)])[:-8],                # truncated vertex
           bytes([1]) + struct.pack("<II", 2, 0xFFFFFFF0)]                    # absurd count
    feats = [(9.0, b) for b in bad] + [(1.0, good)]
    got, n = pp.rasterize_times(g, feats)
    assert n == 1
    want = np.full((H, W), F, np.float32)
    want[3, 2] = 1.0
    assert np.array_equal(got, want)
    assert np.array_equal(R.rasterize(feats, gt, W, H, F, np.float32), want)
    empty = bytes([1]) + struct.pack("<I", 1) + struct.pack("<dd", math.nan, math.nan)
    got, n = pp.rasterize_times(g, [(4.0, empty), (5.0, R.wkb(R.POLYGON, []))])
    assert n == 2 and (got == F).all()                       # empty geometries parse, burn nothing
    far = R.wkb(R.LINESTRING, [(1.5, 1.5), (1e12, 1.5)])
    got, n = pp.rasterize_times(g, [(4.0, far)])
    assert n == 1 and (got == F).all()
    assert (R.rasterize([(4.0, far)], gt, W, H, F, np.float32) == F).all()


def test_init_refs_point_ignited_cells_at_themselves():
    import py_propag as pp
    from py_propag import _native
    import ctypes as C

    r = np.random.Generator(np.random.PCG64(2))
    W, H = 33, 17
    g, _ = _geo(W, H, [0.0, 1.0, 0.0, 0.0, 0.0, 1.0])
    time = np.full((H, W), F, np.float32)
    time[r.integers(0, H, 40), r.integers(0, W, 40)] = r.uniform(0, 100, 40).astype(np.float32)
    rx = np.empty((H, W), np.uint16)
    ry = np.empty((H, W), np.uint16)
    lib = _native.load(require_torch=False)
    gc = g.to_c()
    assert lib.pg_init_refs(C.byref(gc), time.ctypes.data, rx.ctypes.data, ry.ctypes.data) == 0
    wx, wy = R.init_refs(time)
    assert np.array_equal(rx, wx) and np.array_equal(ry, wy)
    assert lib.pg_init_refs(C.byref(gc), None, rx.ctypes.data, ry.ctypes.data) != 0


def test_wkb_ignitions_give_the_initial_state_of_their_cells():
    import py_propag as pp

    W, H = 50, 40
    g, gt = _geo(W, H, [700000.0, 5.0, 0.0, 4400000.0, 0.0, -5.0])
    cells = [(3, 4, 0.0), (20, 30, 60.0), (49, 0, 5.0)]
    t0, rx0, ry0 = pp.initial_state(W, H, cells)
    feats = [pp.TimeFeature(t, R.wkb(R.POINT, _world(gt, [(x + 0.5, y + 0.5)])[0])) for x, y, t in cells]
    t1, rx1, ry1 = pp.initial_state(W, H, feats, g)
    assert np.array_equal(t0, t1) and np.array_equal(rx0, rx1) and np.array_equal(ry0, ry1)
    # a line ignition: every cell of it ignites at its time and points at itself
    line = pp.TimeFeature(30.0, R.wkb(R.LINESTRING, _world(gt, [(10.5, 10.5), (30.5, 15.5)])))
    t2, rx2, ry2 = pp.initial_state(W, H, [line], g)
    lit = np.flatnonzero(t2 != F)
    assert len(lit) == 21 and np.all(t2[lit] == 30.0)
    assert np.array_equal(rx2[lit], lit % W) and np.array_equal(ry2[lit], lit // W)
    with pytest.raises(pp.PropagValidationError):
        pp.initial_state(W, H, feats)                         # WKB needs the GeoReference
