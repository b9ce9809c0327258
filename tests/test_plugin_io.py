"""The plugin library libpropag.so and the output writers (CPU; no GPU calls).

* libpropag.so loads and exports every function of include/propag_host.h
  and include/geometry.h; the FFI structs have the reference's #[repr(C)]
  layout (FFIPropagation 1160 B with the terrain loader at 1144,
  crates/propag/src/propag.rs:177-188);
* GeoReference_* follow crates/geometry/src/lib.rs:195-332;
* the GeoTIFF of pg_write_times_geotiff reads back, through this file's own
  TIFF 6.0 / GeoTIFF 1.0 parser and through the system's libtiff, as the
  raster, geotransform, nodata and CRS keys it was given (the file layout of
  write_times, propag.rs:576-600: f32, 256 x 256 tiles, nodata FLT_MAX);
* the GeoPackages of pg_write_refs_gpkg / pg_write_tiles_gpkg read back
  through Python's sqlite3 as the reference's layers (propag.rs:601-750):
  one feature per burned cell in column-major order, lines from the
  reference's centre to the cell's, with the coordinates the reference
  forms (f32 geotransform arithmetic printed as the shortest round-trip
  decimal, read as a double);
* propag_rasterize_fuels and the error paths of FFIPropagation_run that end
  before the GPU is used: CRS checks, the required output path, a terrain
  callback that refuses, and the NUL-terminated error message.
The GPU run of FFIPropagation_run is in tests/test_gpu_plugin.py.
"""
import ctypes as C
import math
import os
import re
import sqlite3
import struct

import numpy as np
import pytest

from oracle import rasterize as R
from tiffread import geokeys as _geokeys, read_tiff as _read_tiff

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
UTM = "+proj=utm +zone=30 +ellps=GRS80 +towgs84=0,0,0,0,0,0,0 +units=m +no_defs +type=crs"
F = np.float32(np.finfo(np.float32).max)


@pytest.fixture(scope="module")
def plug():
    from py_propag import plugin

    return plugin


def _geo(W, H, gt, proj=UTM):
    import py_propag as pp

    return pp.GeoReference(W, H, proj, gt)


def test_plugin_exports_the_header_functions(plug):
    lib = plug.load()
    src = open(os.path.join(ROOT, "include", "propag_host.h")).read() + \
        open(os.path.join(ROOT, "include", "geometry.h")).read()
    declared = set(re.findall(r"^\s*(?:bool|float|uint32_t|Settings|FireSimpleCuda|PluginT)\s+(\w+)\s*\(", src, re.M))
    assert declared == set(plug.EXPORTS)
    for name in declared:
        assert hasattr(lib, name), name
    assert C.sizeof(plug.FFIPropagation) == 1160 and plug.FFIPropagation.terrain_loader.offset == 1144
    assert C.sizeof(plug.FFITerrain) == 80 and C.sizeof(plug.FireSimpleCuda) == 12
    p = lib.create_plugin()
    assert C.cast(p.run, C.c_void_p).value == C.cast(lib.FFIPropagation_run, C.c_void_p).value
    f = lib.create_fire(1.5, 2.0, 0.25)
    assert (f.speed_max, f.azimuth_max, f.eccentricity) == (1.5, 2.0, 0.25)
    s = lib.create_settings(_geo(5, 4, [0, 1, 0, 0, 0, 1]).to_c(), 600.0)
    assert s.geo_ref.width == 5 and s.max_time == 600.0


def test_georeference_functions(plug):
    from py_propag import _native

    lib = plug.load()
    g = _native.GeoReferenceC()
    assert lib.GeoReference_south_up(1000.0, 2000.0, 1512.5, 2301.0, 5.0, 5.0, UTM.encode(), C.byref(g))
    assert (g.width, g.height) == (103, 61)           # ceil(512.5 / 5), ceil(301 / 5)
    t = g.transform.gt
    assert (t.x0, t.dx, t.rx, t.y0, t.ry, t.dy) == (1000.0, 5.0, 0.0, 2000.0, 0.0, 5.0)
    inv = g.transform.inv
    assert inv.dx == np.float32(1) / np.float32(5) and inv.x0 == -np.float32(1000) / np.float32(5)
    assert bytes(g.proj).rstrip(b"\0").decode() == UTM
    assert lib.GeoReference_len(C.byref(g)) == 103 * 61
    assert (lib.GeoReference_x0(C.byref(g)), lib.GeoReference_x1(C.byref(g))) == (1000.0, 1515.0)
    assert (lib.GeoReference_y0(C.byref(g)), lib.GeoReference_y1(C.byref(g))) == (2000.0, 2305.0)
    n = _native.GeoReferenceC()
    assert lib.GeoReference_north_up(1000.0, 2000.0, 1512.5, 2301.0, 5.0, 5.0, b"", C.byref(n))
    assert (n.transform.gt.y0, n.transform.gt.dy, n.height) == (2301.0, -5.0, 61)
    assert (lib.GeoReference_y0(C.byref(n)), lib.GeoReference_y1(C.byref(n))) == (1996.0, 2301.0)
    bad = _native.GeoReferenceC()
    assert not lib.GeoReference_south_up(10.0, 0.0, 0.0, 5.0, 1.0, 1.0, b"", C.byref(bad))   # negative width
    assert not lib.GeoReference_south_up(0.0, 0.0, 5.0, 5.0, 0.0, 1.0, b"", C.byref(bad))    # singular
    assert not lib.GeoReference_south_up(0.0, 0.0, 5.0, 5.0, 1.0, 1.0, b"\xff\xfe", C.byref(bad))  # not UTF-8


# ------------------------------------------------------------ GeoTIFF

def _libtiff_read(path, W, H):
    """The same file through the system libtiff (None when it is absent)."""
    try:
        lt = C.CDLL("libtiff.so.5")
    except OSError:
        return None
    lt.TIFFSetWarningHandler.restype = C.c_void_p
    lt.TIFFSetWarningHandler.argtypes = [C.c_void_p]
    lt.TIFFSetWarningHandler(None)
    lt.TIFFOpen.restype = C.c_void_p
    lt.TIFFOpen.argtypes = [C.c_char_p, C.c_char_p]
    lt.TIFFReadTile.restype = C.c_int64
    lt.TIFFReadTile.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint16]
    lt.TIFFClose.argtypes = [C.c_void_p]
    lt.TIFFIsTiled.argtypes = [C.c_void_p]
    tif = lt.TIFFOpen(os.fsencode(path), b"r")
    assert tif
    try:
        assert lt.TIFFIsTiled(tif)
        vals = {}
        for tag, ct in ((256, C.c_uint32), (257, C.c_uint32), (322, C.c_uint32), (323, C.c_uint32),
                        (258, C.c_uint16), (339, C.c_uint16)):
            v = ct()
            assert lt.TIFFGetField(C.c_void_p(tif), C.c_uint32(tag), C.byref(v)) == 1, tag
            vals[tag] = v.value
        assert (vals[256], vals[257], vals[322], vals[323], vals[258], vals[339]) == (W, H, 256, 256, 32, 3)
        out = np.empty((H, W), np.float32)
        tile = np.empty((256, 256), np.float32)
        for y in range(0, H, 256):
            for x in range(0, W, 256):
                assert lt.TIFFReadTile(tif, tile.ctypes.data, x, y, 0, 0) == tile.nbytes
                h, w = min(256, H - y), min(256, W - x)
                out[y:y + h, x:x + w] = tile[:h, :w]
        return out
    finally:
        lt.TIFFClose(tif)


@pytest.mark.parametrize("north_up", [False, True])
def test_geotiff_reads_back(tmp_path, north_up):
    import py_propag as pp

    r = np.random.Generator(np.random.PCG64(8))
    W, H = 300, 517                                  # partial tiles on both edges
    gt = [712345.5, 5.0, 0.0, 4401234.0, 0.0, -5.0 if north_up else 5.0]
    g = _geo(W, H, gt)
    time = r.uniform(0, 36000, (H, W)).astype(np.float32)
    time[r.random((H, W)) < 0.3] = F
    path = tmp_path / "t.tif"
    pp.write_times_geotiff(path, g, time)
    tags, got = _read_tiff(path)
    assert np.array_equal(got.view(np.uint32), time.view(np.uint32))
    assert tags[258] == [32] and tags[339] == [3] and tags[259] == [1] and tags[277] == [1]
    assert np.float32(float(tags[42113].rstrip("\0"))) == F
    if north_up:
        assert tags[33550] == [5.0, 5.0, 0.0] and tags[33922] == [0, 0, 0, 712345.5, 4401234.0, 0]
        assert 34264 not in tags
    else:
        m = tags[34264]
        assert (m[0], m[1], m[3], m[4], m[5], m[7], m[15]) == (5.0, 0.0, 712345.5, 0.0, 5.0, 4401234.0, 1.0)
        assert 33550 not in tags
    k = _geokeys(tags)
    # the reference's CRS (ETRS89-like GRS80 UTM 30): user-defined projected CS on UTM zone 30N
    assert k[1024] == 1 and k[1025] == 1 and k[3072] == 32767 and k[3074] == 16030 and k[2056] == 7019
    assert k[3076] == 9001 and k[1026] == UTM + "|"
    lt = _libtiff_read(path, W, H)
    if lt is not None:
        assert np.array_equal(lt.view(np.uint32), time.view(np.uint32))


@pytest.mark.parametrize("proj,keys", [
    ("+proj=utm +zone=30 +datum=WGS84 +units=m +no_defs", {1024: 1, 3072: 32630}),
    ("+proj=utm +zone=19 +south +ellps=WGS84", {1024: 1, 3072: 32719}),
    ("EPSG:25830", {1024: 1, 3072: 25830}),
    ("+init=epsg:4258", {1024: 2, 2048: 4258}),
    ("+proj=longlat +datum=WGS84 +no_defs", {1024: 2, 2048: 4326}),
    ("+proj=tmerc +lat_0=0 +lon_0=3", {1024: 32767}),
])
def test_geotiff_crs_keys(tmp_path, proj, keys):
    import py_propag as pp

    g = _geo(4, 3, [0.0, 1.0, 0.0, 10.0, 0.0, -1.0], proj)
    pp.write_times_geotiff(tmp_path / "c.tif", g, np.zeros((3, 4), np.float32))
    tags, _ = _read_tiff(tmp_path / "c.tif")
    k = _geokeys(tags)
    for key, v in keys.items():
        assert k[key] == v, (proj, key)
    assert k[1026] == proj + "|"


def test_geotiff_rejects_bad_arguments(tmp_path):
    import py_propag as pp

    g = _geo(4, 3, [0.0, 1.0, 0.0, 0.0, 0.0, 1.0])
    with pytest.raises(pp.PropagValidationError):
        pp.write_times_geotiff(tmp_path / "x.tif", g, np.zeros(5, np.float32))
    with pytest.raises(pp.PropagGdalError):
        pp.write_times_geotiff(tmp_path / "no" / "dir" / "x.tif", g, np.zeros((3, 4), np.float32))


# ------------------------------------------------------------ GeoPackage

def _gp_geometry(blob):
    assert blob[:2] == b"GP" and blob[2] == 0 and blob[3] == 0x03
    srs = struct.unpack_from("<i", blob, 4)[0]
    env = struct.unpack_from("<4d", blob, 8)
    w = blob[40:]
    assert w[0] == 1
    kind = struct.unpack_from("<I", w, 1)[0]
    if kind == 2:
        n = struct.unpack_from("<I", w, 5)[0]
        pts = struct.unpack_from("<" + "d" * 2 * n, w, 9)
    else:
        assert kind == 3 and struct.unpack_from("<I", w, 5)[0] == 1
        n = struct.unpack_from("<I", w, 9)[0]
        pts = struct.unpack_from("<" + "d" * 2 * n, w, 13)
    xy = np.array(pts).reshape(-1, 2)
    assert env == (xy[:, 0].min(), xy[:, 0].max(), xy[:, 1].min(), xy[:, 1].max())
    return srs, kind, xy


def _f32_decimal(v):
    """f32 printed as its shortest round-trip decimal, read as a double."""
    return float(str(np.float32(v)))


def test_refs_gpkg_has_the_reference_layer(tmp_path):
    import py_propag as pp

    r = np.random.Generator(np.random.PCG64(9))
    W, H = 23, 17
    gt = [712345.5, 5.0, 0.0, 4401234.0, 0.0, -5.0]
    g = _geo(W, H, gt)
    time = np.full((H, W), F, np.float32)
    lit = r.random((H, W)) < 0.6
    time[lit] = r.uniform(0, 100, lit.sum()).astype(np.float32)
    rx = r.integers(0, W, (H, W)).astype(np.uint16)
    ry = r.integers(0, H, (H, W)).astype(np.uint16)
    path = tmp_path / "refs.gpkg"
    pp.write_refs_gpkg(path, g, time, rx, ry)
    db = sqlite3.connect(path)
    assert db.execute("PRAGMA application_id").fetchone()[0] == 0x47504B47
    assert db.execute("SELECT table_name, data_type FROM gpkg_contents").fetchall() == [("fire_references", "features")]
    assert db.execute("SELECT * FROM gpkg_geometry_columns").fetchall() == [("fire_references", "geom", "LINESTRING", 100000, 0, 0)]
    srs = db.execute("SELECT organization, description FROM gpkg_spatial_ref_sys WHERE srs_id = 100000").fetchone()
    assert srs == ("NONE", UTM)
    rows = db.execute("SELECT fid, geom FROM fire_references ORDER BY fid").fetchall()
    cells = [(i, j) for i in range(W) for j in range(H) if time[j, i] < F]      # column-major, as the reference
    assert len(rows) == len(cells)
    f = np.float32
    hx, hy = f(gt[1]) / f(2), f(gt[5]) / f(2)
    cx = lambda x, y: f(f(f(gt[0]) + f(x) * f(gt[1])) + f(y) * f(gt[2])) + hx
    cy = lambda x, y: f(f(f(gt[3]) + f(x) * f(gt[4])) + f(y) * f(gt[5])) + hy
    for (fid, blob), (i, j) in zip(rows, cells):
        srs_id, kind, xy = _gp_geometry(blob)
        assert srs_id == 100000 and kind == 2
        sx, sy = int(rx[j, i]), int(ry[j, i])
        want = [[_f32_decimal(cx(sx, sy)), _f32_decimal(cy(sx, sy))], [_f32_decimal(cx(i, j)), _f32_decimal(cy(i, j))]]
        assert xy.tolist() == want, (i, j)
    env = db.execute("SELECT min_x, min_y, max_x, max_y FROM gpkg_contents").fetchone()
    allxy = np.concatenate([_gp_geometry(b)[2] for _, b in rows])
    assert env == (allxy[:, 0].min(), allxy[:, 1].min(), allxy[:, 0].max(), allxy[:, 1].max())


def test_tiles_gpkg(tmp_path):
    import py_propag as pp

    W, H = 21, 10
    gt = [100.0, 2.0, 0.0, 50.0, 0.0, 3.0]
    g = _geo(W, H, gt, "+proj=utm +zone=30 +datum=WGS84")
    pp.write_tiles_gpkg(tmp_path / "b.gpkg", "blocks", g, 8, 8)
    db = sqlite3.connect(tmp_path / "b.gpkg")
    assert db.execute("SELECT geometry_type_name, srs_id FROM gpkg_geometry_columns").fetchone() == ("POLYGON", 32630)
    polys = [_gp_geometry(b) for (b,) in db.execute("SELECT geom FROM blocks ORDER BY fid")]
    assert len(polys) == 3 * 2
    for k, (srs, kind, xy) in enumerate(polys):
        x0, y0 = (k % 3) * 8, (k // 3) * 8
        corners = [(x0, y0), (x0 + 8, y0), (x0 + 8, y0 + 8), (x0, y0 + 8), (x0, y0)]
        assert kind == 3 and xy.tolist() == [[100.0 + 2.0 * x, 50.0 + 3.0 * y] for x, y in corners]
    pp.write_tiles_gpkg(tmp_path / "g.gpkg", "grids", g, W, H)
    assert sqlite3.connect(tmp_path / "g.gpkg").execute("SELECT count(*) FROM grids").fetchone()[0] == 1
    with pytest.raises(pp.PropagValidationError):
        pp.write_tiles_gpkg(tmp_path / "x.gpkg", "bad name;", g, 8, 8)
    with pytest.raises(pp.PropagGdalError):
        pp.write_tiles_gpkg(tmp_path / "no" / "dir" / "x.gpkg", "blocks", g, 8, 8)


# ----------------------------------------------- plugin entry points (CPU)

def test_propag_rasterize_fuels(plug):
    import py_propag as pp
    from py_propag import _native

    lib = plug.load()
    W, H = 30, 20
    g = _geo(W, H, [1000.0, 5.0, 0.0, 2000.0, 0.0, 5.0])
    gt = [float(v) for v in g.transform]
    r = np.random.Generator(np.random.PCG64(4))
    feats = []
    for code in range(1, 9):
        x, y = r.uniform(1000, 1150), r.uniform(2000, 2100)
        feats.append((code, R.wkb(R.POLYGON, [[(x, y), (x + 40, y), (x + 40, y + 30), (x, y + 30), (x, y)]])))
    arr, keep = pp._feature_array(feats, _native.FuelFeatureC, int)
    out = np.zeros((H, W), np.uint8)
    err = C.create_string_buffer(256)
    gc = g.to_c()
    assert lib.propag_rasterize_fuels(arr, len(feats), UTM.encode(), C.byref(gc), out.ctypes.data, err, 256)
    assert np.array_equal(out, R.rasterize(feats, gt, W, H, 255, np.uint8))
    # the same CRS written with its parameters in another order is the same CRS
    shuffled = " ".join(reversed(UTM.split()))
    assert lib.propag_rasterize_fuels(arr, len(feats), shuffled.encode(), C.byref(gc), out.ctypes.data, err, 256)
    # CRSs outside the transformable set, and pairs that need a datum shift, are refused
    assert not lib.propag_rasterize_fuels(arr, len(feats), b"+proj=merc +datum=WGS84", C.byref(gc),
                                          out.ctypes.data, err, 256)
    assert b"not supported" in err.value
    assert not lib.propag_rasterize_fuels(arr, len(feats), b"+proj=longlat +ellps=airy +towgs84=446.4,-125.2,542.1",
                                          C.byref(gc), out.ctypes.data, err, 256)
    assert b"datum shift" in err.value
    small = C.create_string_buffer(b"\xff" * 8)
    assert not lib.propag_rasterize_fuels(arr, len(feats), b"", C.byref(gc), out.ctypes.data, small, 8)
    assert small.raw[7] == 0 and len(small.value) == 7                  # truncated and terminated
    assert lib.propag_rasterize_fuels(None, 0, UTM.encode(), C.byref(gc), out.ctypes.data, err, 256)
    assert (out == 255).all()


def test_run_reports_errors_before_the_gpu(plug, tmp_path):
    import py_propag as pp

    W, H = 16, 12
    s = pp.Settings(_geo(W, H, [0.0, 5.0, 0.0, 0.0, 0.0, 5.0]), 3600.0)
    ign = [(0.0, R.wkb(R.POINT, (40.0, 30.0)))]
    calls = []

    def fill(geo, cols):
        calls.append((geo.width, geo.height, {k: (v.dtype, v.size, int(np.count_nonzero(v))) for k, v in cols.items()}))
        return False

    ok, msg = plug.run(s, str(tmp_path / "t.tif"), ign, UTM, fill)
    assert not ok and msg == "load_extent_error"
    w, h, cols = calls[0]
    assert (w, h) == (W, H) and cols["fuel"] == (np.uint8, W * H, 0)   # TerrainCuda::NULL columns
    assert all(cols[k] == (np.float32, W * H, 0) for k in plug.TERRAIN_COLUMNS)
    ok, msg = plug.run(s, "", ign, UTM, fill)
    assert not ok and msg == "invalid output_path"
    ok, msg = plug.run(s, str(tmp_path / "t.tif"), ign, "+proj=lcc +lat_1=40 +datum=WGS84", fill)
    assert not ok and "not supported" in msg
    ok, msg = plug.run(s, str(tmp_path / "t.tif"), ign, None, fill)
    assert not ok and "CRS" in msg
    assert len(calls) == 1 and not os.path.exists(tmp_path / "t.tif")
