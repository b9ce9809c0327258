"""CRS transformation (pg_transform_points, csrc/reproject.h) and the
plugin's reprojection of features (FFIPropagation_run /
propag_rasterize_fuels), on the CPU.

The reference hands every ignition and fuel geometry to OGR/PROJ
(Geometry::transform_to_inplace, crates/propag/src/propag.rs:854-857,
964-969).  PROJ is absent here; the transverse Mercator implementation is
pinned by what does not depend on it:
  * published worked examples: Snyder 1987 (USGS Professional Paper 1395,
    p. 269-270), Clarke 1866, lat 40.5, lon -73.5, lon_0 -75, k 0.9996:
    x = 127,106.5 m, y = 4,484,124.4 m; Ordnance Survey, "A guide to
    coordinate systems in Great Britain", Annex C: Airy 1830, lat
    52 39 27.2531 N, lon 1 43 4.5177 E on the National Grid projection
    (lat_0 49, lon_0 -2, k 0.9996012717, false origin 400000, -100000):
    E = 651,409.903 m, N = 313,177.270 m; UTM 31N of (0, 0) on WGS84:
    E = 166,021.443 m;
  * on the central meridian the northing is k0 times the meridian arc, here
    integrated numerically from the ellipsoid's meridian radius (this pins
    the series' A and alpha terms);
  * on a sphere (f = 0) the projection has the closed form
    x = R k0 atanh(cos(lat) sin(dlon)), y = R k0 atan2(tan(lat), cos(dlon));
  * conformality: off the meridian the scale factors east and north agree
    (a wrong term of the complex series breaks this);
  * round trips and the projection's symmetry.
"""
import ctypes as C
import math

import numpy as np
import pytest

WGS84 = "+proj=longlat +datum=WGS84 +no_defs"


@pytest.fixture(scope="module")
def lib():
    from py_propag import _native

    return _native.load(require_torch=False)


def transform(lib, src, dst, pts):
    pts = np.asarray(pts, float).reshape(-1, 2)
    x, y = np.ascontiguousarray(pts[:, 0]), np.ascontiguousarray(pts[:, 1])
    ok = np.zeros(len(pts), np.uint8)
    rc = lib.pg_transform_points(src.encode(), dst.encode(), len(pts), x.ctypes.data, y.ctypes.data,
                                 ok.ctypes.data)
    if rc != 0:
        raise ValueError(lib.pg_last_error().decode())
    return x, y, ok.astype(bool)


def test_snyder_worked_example(lib):
    x, y, ok = transform(lib, "+proj=longlat +ellps=clrk66",
                         "+proj=tmerc +lat_0=0 +lon_0=-75 +k=0.9996 +x_0=0 +y_0=0 +ellps=clrk66", [(-73.5, 40.5)])
    assert ok[0]
    assert abs(x[0] - 127106.5) < 0.06 and abs(y[0] - 4484124.4) < 0.06


def test_ordnance_survey_worked_example(lib):
    lat = 52 + 39 / 60 + 27.2531 / 3600
    lon = 1 + 43 / 60 + 4.5177 / 3600
    osgb = "+proj=tmerc +lat_0=49 +lon_0=-2 +k=0.9996012717 +x_0=400000 +y_0=-100000 +ellps=airy +units=m"
    x, y, ok = transform(lib, "+proj=longlat +ellps=airy", osgb, [(lon, lat)])
    assert ok[0]
    assert abs(x[0] - 651409.903) < 0.002 and abs(y[0] - 313177.270) < 0.002


def test_utm_known_points(lib):
    utm31 = "+proj=utm +zone=31 +datum=WGS84"
    x, y, ok = transform(lib, WGS84, utm31, [(0.0, 0.0), (3.0, 0.0)])
    assert ok.all()
    assert abs(x[0] - 166021.4431) < 0.001 and y[0] == 0.0
    assert x[1] == 500000.0 and y[1] == 0.0
    x, y, _ = transform(lib, WGS84, utm31 + " +south", [(3.0, 0.0), (3.0, -10.0)])
    assert y[0] == 10000000.0 and y[1] < 10000000.0


def _meridian_arc(a, f, lat):
    from scipy.integrate import quad

    e2 = f * (2 - f)
    val, _ = quad(lambda p: a * (1 - e2) / (1 - e2 * math.sin(p) ** 2) ** 1.5, 0.0, lat, epsabs=1e-7,
                  epsrel=1e-13, limit=200)
    return val


@pytest.mark.parametrize("ellps,a,rf", [("GRS80", 6378137.0, 298.257222101), ("intl", 6378388.0, 297.0)])
def test_central_meridian_is_the_meridian_arc(lib, ellps, a, rf):
    lats = np.arange(0.0, 85.0, 7.0)
    x, y, ok = transform(lib, f"+proj=longlat +ellps={ellps}", f"+proj=utm +zone=30 +ellps={ellps}",
                         [(-3.0, v) for v in lats])
    assert ok.all() and np.all(x == 500000.0)
    for lat, n in zip(lats, y):
        assert abs(n - 0.9996 * _meridian_arc(a, 1 / rf, math.radians(lat))) < 1e-4


def test_sphere_closed_form(lib):
    R, k0, lon0 = 6370997.0, 0.9996, 9.0
    r = np.random.Generator(np.random.PCG64(3))
    pts = np.column_stack([lon0 + r.uniform(-20, 20, 200), r.uniform(-80, 80, 200)])
    x, y, ok = transform(lib, f"+proj=longlat +R={R}", f"+proj=tmerc +lon_0={lon0} +k={k0} +R={R}", pts)
    assert ok.all()
    lam, phi = np.radians(pts[:, 0] - lon0), np.radians(pts[:, 1])
    assert np.allclose(x, R * k0 * np.arctanh(np.cos(phi) * np.sin(lam)), atol=1e-6, rtol=0)
    assert np.allclose(y, R * k0 * np.arctan2(np.tan(phi), np.cos(lam)), atol=1e-6, rtol=0)


def test_conformal_off_the_meridian(lib):
    """Scale east = scale north, against the ellipsoid's radii of curvature."""
    a, f = 6378137.0, 1 / 298.257222101
    e2 = f * (2 - f)
    h = 1e-5  # degrees
    for lon, lat in [(-1.0, 40.0), (2.5, 10.0), (-4.0, 65.0), (6.0, -30.0)]:
        pts = [(lon, lat), (lon + h, lat), (lon, lat + h)]
        x, y, _ = transform(lib, "+proj=longlat +ellps=GRS80", "+proj=utm +zone=30 +ellps=GRS80", pts)
        p = math.radians(lat)
        nu = a / math.sqrt(1 - e2 * math.sin(p) ** 2)
        rho = a * (1 - e2) / (1 - e2 * math.sin(p) ** 2) ** 1.5
        k_east = math.hypot(x[1] - x[0], y[1] - y[0]) / (nu * math.cos(p) * math.radians(h))
        k_north = math.hypot(x[2] - x[0], y[2] - y[0]) / (rho * math.radians(h))
        assert abs(k_east - k_north) < 1e-7, (lon, lat, k_east, k_north)


def test_round_trips_and_symmetry(lib):
    r = np.random.Generator(np.random.PCG64(11))
    utm = "+proj=utm +zone=30 +ellps=GRS80 +towgs84=0,0,0,0,0,0,0 +units=m +no_defs"
    pts = np.column_stack([-3.0 + r.uniform(-6, 6, 500), r.uniform(-84, 84, 500)])
    x, y, ok = transform(lib, WGS84, utm, pts)
    assert ok.all()
    lon, lat, ok2 = transform(lib, utm, WGS84, np.column_stack([x, y]))
    assert ok2.all()
    assert np.abs(lon - pts[:, 0]).max() < 1e-9 and np.abs(lat - pts[:, 1]).max() < 1e-9
    # mirror images about the central meridian and the equator
    xm, ym, _ = transform(lib, WGS84, utm, np.column_stack([-6.0 - pts[:, 0], -pts[:, 1]]))
    assert np.allclose(xm - 500000.0, -(x - 500000.0), atol=1e-6) and np.allclose(ym, -y, atol=1e-6)


def test_units_identity_domain_and_refusals(lib):
    x, y, _ = transform(lib, WGS84, "+proj=utm +zone=30 +datum=WGS84", [(-2.0, 41.0)])
    xk, yk, _ = transform(lib, WGS84, "+proj=utm +zone=30 +datum=WGS84 +units=km", [(-2.0, 41.0)])
    assert abs(xk[0] * 1000 - x[0]) < 1e-6 and abs(yk[0] * 1000 - y[0]) < 1e-6
    # the same CRS in another parameter order: coordinates pass unchanged
    same = transform(lib, "+proj=utm +zone=30 +datum=WGS84", "+datum=WGS84 +zone=30 +proj=utm", [(1.5e5, 4.4e6)])
    assert same[0][0] == 1.5e5 and same[1][0] == 4.4e6
    # a point 100 degrees from the central meridian is outside the projection
    _, _, ok = transform(lib, WGS84, "+proj=utm +zone=30 +datum=WGS84", [(97.0, 10.0), (-3.0, 10.0)])
    assert list(ok) == [False, True]
    for src, dst, msg in [("+proj=merc +datum=WGS84", WGS84, "not supported"),
                          (WGS84, "+proj=longlat +ellps=airy +towgs84=446.4,-125.2,542.1", "datum shift"),
                          (WGS84, "+proj=utm +zone=61 +datum=WGS84", "zone"),
                          ("+proj=longlat +datum=NAD27", WGS84, "datum shift"),
                          ("+proj=utm +zone=30 +ellps=foo", WGS84, "ellipsoid")]:
        with pytest.raises(ValueError, match=msg):
            transform(lib, src, dst, [(0.0, 0.0)])
    # the same shift on both sides cancels
    shifted = "+proj=longlat +ellps=airy +towgs84=446.4,-125.2,542.1"
    transform(lib, shifted, shifted.replace("longlat", "tmerc +lon_0=-2 +k=0.9996"), [(-2.0, 50.0)])


def test_plugin_reprojects_fuel_features(lib):
    """propag_rasterize_fuels with the features in longitude/latitude burns
    the same raster as with the same features in the raster's UTM
    coordinates (vertices a quarter cell from pixel centres and edges, so
    the round trip's sub-millimetre differences cannot move a burn)."""
    import py_propag as pp
    from py_propag import _native, plugin as plug

    import oracle.rasterize as R

    utm = "+proj=utm +zone=30 +ellps=GRS80 +towgs84=0,0,0,0,0,0,0 +units=m +no_defs"
    W, H, cell = 60, 40, 5.0
    x0, y0 = 440000.0, 4470000.0
    g = pp.GeoReference(W, H, utm, [x0, cell, 0.0, y0, 0.0, cell])
    r = np.random.Generator(np.random.PCG64(8))
    utm_feats, ll_feats = [], []
    for code in range(1, 10):
        cx, cy = r.integers(2, W - 12), r.integers(2, H - 10)
        w, h = r.integers(3, 10), r.integers(3, 8)
        ring = [(x0 + (cx + dx + 0.25) * cell, y0 + (cy + dy + 0.25) * cell)
                for dx, dy in [(0, 0), (w, 0), (w, h), (0, h), (0, 0)]]
        lon, lat, ok = transform(lib, utm, WGS84, ring)
        assert ok.all()
        utm_feats.append((code, R.wkb(R.POLYGON, [ring])))
        ll_feats.append((code, R.wkb(R.POLYGON, [list(zip(lon, lat))])))
    pl = plug.load()
    gc = g.to_c()
    err = C.create_string_buffer(256)
    out = []
    for feats, crs in ((utm_feats, utm), (ll_feats, WGS84)):
        arr, keep = pp._feature_array(feats, _native.FuelFeatureC, int)
        o = np.zeros((H, W), np.uint8)
        assert pl.propag_rasterize_fuels(arr, len(feats), crs.encode(), C.byref(gc), o.ctypes.data, err, 256), \
            err.value
        out.append(o)
    assert (out[0] != 255).sum() > 100
    assert np.array_equal(out[0], out[1])
