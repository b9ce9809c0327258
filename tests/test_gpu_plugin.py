"""FFIPropagation_run (libpropag.so, the reference's plugin entry point) on
the GPU, end to end, the way the QGIS plugin calls it
(qgis-plugin/propag-algo/src/qgsalgorithmpropag.cpp:162-258): WKB ignitions
in the raster's CRS, a terrain callback filling the library's columns, and
output paths.

The GeoTIFF it writes must hold, bit for bit, (a) the arrival times that
py_propag.propagate (pg_propagate_host: the same GPU burn + propagation)
returns from the initial state of the oracle's rasterisation of the same WKB
(oracle/rasterize.py) and its initial references, and (b) the serial
oracle's propagation from that initial state over the fire rasters of the
GPU burn pass.  The burn pass itself is pinned separately to 1e-5 relative
(tests/test_gpu_burn.py); its last-ulp differences from the f32 oracle are
not carried through here, because the update rule's discrete choices (ties
between neighbours, line-of-sight outcomes of different references) can
turn a one-ulp change of a fire raster into seconds of arrival time.  The
GeoPackages must hold one reference line per burned cell and the tile
layers.
"""
import sqlite3

import numpy as np
import pytest

import oracle
from oracle import rasterize as R
from py_propag import synth
from tiffread import read_tiff

pytestmark = pytest.mark.gpu

UTM = "+proj=utm +zone=30 +ellps=GRS80 +towgs84=0,0,0,0,0,0,0 +units=m +no_defs +type=crs"


@pytest.mark.parametrize("north_up", [False, True])
def test_ffipropagation_run_matches_the_oracle(oracle_lib, tmp_path, north_up):
    import py_propag as pp
    from py_propag import plugin

    side = 96
    sc = synth.single_scenario(seed=5, side=side)
    dx = float(sc["dx"])
    dy = -dx if north_up else dx
    x0, y0 = 700000.0, 4400000.0 + (side * dx if north_up else 0.0)
    gt = [x0, dx, 0.0, y0, 0.0, dy]
    geo = pp.GeoReference(side, side, UTM, gt)
    settings = pp.Settings(geo, float(sc["max_time"]))
    world = lambda px, py: (x0 + px * dx, y0 + py * dy)
    ign = [(0.0, R.wkb(R.POINT, world(40.5, 41.5))),
           (600.0, R.wkb(R.LINESTRING, [world(10.2, 70.7), world(30.9, 60.1)])),
           (1200.0, R.wkb(R.POLYGON, [[world(70.3, 20.2), world(80.6, 20.2), world(80.6, 27.9),
                                       world(70.3, 27.9), world(70.3, 20.2)]])),
           (5.0, b"\x01\x02")]                                   # invalid WKB: skipped
    ter = sc["terrain"]

    def fill(g, cols):
        assert (g.width, g.height) == (side, side)
        cols["fuel"][:] = ter["fuel"]
        for k in plugin.TERRAIN_COLUMNS:
            cols[k][:] = ter[k]
        return True

    tif, refs, blocks, grids = (str(tmp_path / n) for n in ("t.tif", "r.gpkg", "b.gpkg", "g.gpkg"))
    ok, msg = plugin.run(settings, tif, ign, UTM, fill, refs, blocks, grids)
    assert ok, msg

    # the oracle on the same inputs
    t0 = R.rasterize(ign, [float(v) for v in geo.transform], side, side, R.FLT_MAX, np.float32)
    assert (t0 < R.FLT_MAX).sum() > 20
    rx0, ry0 = R.init_refs(t0)
    same = pp.propagate(settings, ter, time=t0, refs_x=rx0, refs_y=ry0, return_fire=True)
    g = oracle.grid(side, side, dx, dy, sc["max_time"])
    ref = oracle.propagate(g, *same.fire, t0.ravel().copy(), rx0.ravel().copy(), ry0.ravel().copy(),
                           nthreads=16)

    tags, got = read_tiff(tif)
    assert np.array_equal(got.view(np.uint32), same.arrival_times.view(np.uint32))
    assert np.array_equal(got.ravel().view(np.uint32), ref["time"].view(np.uint32))
    assert np.array_equal(same.refs_x, ref["refs_x"]) and np.array_equal(same.refs_y, ref["refs_y"])
    burned = int((got < R.FLT_MAX).sum())
    assert burned > 1000
    db = sqlite3.connect(refs)
    assert db.execute("SELECT count(*) FROM fire_references").fetchone()[0] == burned
    assert sqlite3.connect(blocks).execute("SELECT count(*) FROM blocks").fetchone()[0] == (side // 8) ** 2
    assert sqlite3.connect(grids).execute("SELECT count(*) FROM grids").fetchone()[0] == 1

    # a second run in the same process works (the reference's MPI init forbids it)
    ok, msg = plugin.run(settings, str(tmp_path / "t2.tif"), ign, UTM, fill)
    assert ok, msg
    assert np.array_equal(read_tiff(str(tmp_path / "t2.tif"))[1].view(np.uint32), got.view(np.uint32))


def test_ignitions_in_another_crs_are_reprojected(oracle_lib, tmp_path):
    """FFIPropagation_run with the ignitions in longitude/latitude (the QGIS
    plugin passes the ignition layer's own CRS, qgsalgorithmpropag.cpp:
    188-216,247-250) writes the same GeoTIFF as with the same ignitions in
    the raster's UTM coordinates: the features are reprojected before they
    are rasterised (propag.rs:854-857).  Vertices sit at least 0.1 cell from
    pixel centres and edges, so the round trip's sub-millimetre differences
    cannot move a burn."""
    import ctypes as C

    import py_propag as pp
    from py_propag import _native, plugin

    lib = _native.load()
    side = 96
    sc = synth.single_scenario(seed=5, side=side)
    dx = float(sc["dx"])
    x0, y0 = 440000.0, 4470000.0
    geo = pp.GeoReference(side, side, UTM, [x0, dx, 0.0, y0, 0.0, dx])
    settings = pp.Settings(geo, float(sc["max_time"]))
    world = lambda px, py: (x0 + px * dx, y0 + py * dx)  # noqa: E731
    parts = [(0.0, R.POINT, world(40.5, 41.5)),
             (600.0, R.LINESTRING, [world(10.2, 70.7), world(30.9, 60.1)]),
             (1200.0, R.POLYGON, [[world(70.3, 20.2), world(80.6, 20.2), world(80.6, 27.9), world(70.3, 27.9),
                                   world(70.3, 20.2)]])]
    wgs84 = "+proj=longlat +datum=WGS84 +no_defs"

    def to_lonlat(pts):
        pts = np.asarray(pts, float).reshape(-1, 2)
        x, y = pts[:, 0].copy(), pts[:, 1].copy()
        ok = np.zeros(len(pts), np.uint8)
        assert lib.pg_transform_points(UTM.encode(), wgs84.encode(), len(pts), x.ctypes.data, y.ctypes.data,
                                       ok.ctypes.data) == 0
        assert ok.all()
        return [tuple(p) for p in np.column_stack([x, y])]

    ign_utm = [(t, R.wkb(kind, coords)) for t, kind, coords in parts]
    ign_ll = [(0.0, R.wkb(R.POINT, to_lonlat(parts[0][2])[0])),
              (600.0, R.wkb(R.LINESTRING, to_lonlat(parts[1][2]))),
              (1200.0, R.wkb(R.POLYGON, [to_lonlat(parts[2][2][0])]))]
    ter = sc["terrain"]

    def fill(g, cols):
        cols["fuel"][:] = ter["fuel"]
        for k in plugin.TERRAIN_COLUMNS:
            cols[k][:] = ter[k]
        return True

    out = []
    for ign, crs, name in ((ign_utm, UTM, "utm.tif"), (ign_ll, wgs84, "ll.tif")):
        ok, msg = plugin.run(settings, str(tmp_path / name), ign, crs, fill)
        assert ok, msg
        out.append(read_tiff(str(tmp_path / name))[1])
    assert (out[0] < R.FLT_MAX).sum() > 1000
    assert np.array_equal(out[0].view(np.uint32), out[1].view(np.uint32))
