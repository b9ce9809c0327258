"""CPU checks of the C-ABI library and the Python host layer (no GPU calls).

* libpropag_hip.so is built, loads, and exports every entry point declared
  in include/propag_hip.h;
* the ctypes structures have the reference's #[repr(C)] sizes
  (GeoReference 1080 B, Settings 1084 B: crates/geometry/src/lib.rs:180-187,
  crates/propag/src/propag.rs:38-43);
* host-side validation mirrors py-propag (crates/py-propag/src/lib.rs);
* the product package never imports the oracle.
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "propag_hip.h")
PKG = os.path.join(ROOT, "propag-hidalgo_amd")


def declared_entry_points():
    src = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(pg_\w+)\s*\(", src, re.M)))


def test_library_exports_every_declared_symbol():
    from py_propag import _native

    names = declared_entry_points()
    assert {"pg_burn", "pg_propagate", "pg_propagate_host", "pg_last_error"} <= set(names)
    lib = _native.load(require_torch=False)
    for name in names:
        assert hasattr(lib, name), name
    assert set(_native.EXPORTS) == set(names)
    assert lib.pg_version().decode().startswith("propag-hip")


def test_struct_layouts_match_reference():
    from py_propag import _native

    assert C.sizeof(_native.GT) == 24
    assert C.sizeof(_native.GeoTransform) == 48
    assert C.sizeof(_native.GeoReferenceC) == 1080
    assert C.sizeof(_native.SettingsC) == 1084
    assert _native.SettingsC.max_time.offset == 1080
    assert _native.GeoReferenceC.transform.offset == 1032


def test_georeference_and_settings_validation():
    import py_propag as pp

    g = pp.GeoReference.south_up(((0.0, 0.0), (5120.0, 5120.0)), (5.0, 5.0), "+proj=utm +zone=30")
    assert (g.width, g.height) == (1024, 1024) and not g.is_north_up
    gn = pp.GeoReference.north_up(((0.0, 0.0), (50.0, 30.0)), (5.0, 5.0), "+proj=utm +zone=30")
    assert (gn.width, gn.height) == (10, 6) and gn.is_north_up
    c = g.to_c()
    assert c.transform.inv.dx == pytest.approx(0.2)
    with pytest.raises(pp.PropagValidationError):
        pp.GeoReference(0, 10, "x", [0, 1, 0, 0, 0, 1])
    with pytest.raises(pp.PropagValidationError):
        pp.GeoReference(10, 10, "x", [0, 0, 0, 0, 0, 1])
    with pytest.raises(pp.PropagValidationError):
        pp.GeoReference(10, 10, "x" * 1024, [0, 1, 0, 0, 0, 1])
    with pytest.raises(pp.PropagValidationError):
        pp.Settings(g, -1.0)
    s = pp.Settings(g, 36000.0)
    assert s.to_c().max_time == 36000.0


def test_initial_state_and_fuel_validation():
    import py_propag as pp

    t, rx, ry = pp.initial_state(8, 4, [(1, 2, 0.0), pp.TimeFeature(5.0, cells=[(7, 3)])])
    assert t[1 + 2 * 8] == 0.0 and rx[1 + 2 * 8] == 1 and ry[1 + 2 * 8] == 2
    assert t[7 + 3 * 8] == 5.0 and (t < pp.FLT_MAX).sum() == 2
    assert rx[0] == 0xFFFF
    with pytest.raises(pp.PropagValidationError):
        pp.initial_state(8, 4, [(8, 0, 0.0)])
    with pytest.raises(pp.PropagValidationError):
        pp.load_terrain_data(4, 4, "p", [0, 1, 0, 0, 0, 1], fuel_model=np.full((4, 4), 14, np.uint8))
    with pytest.raises(pp.PropagValidationError):
        pp.load_terrain_data(4, 4, "p", [0, 1, 0, 0, 0, 1], slope=np.zeros((3, 4), np.float32))


def test_product_never_imports_the_oracle():
    for dirpath, _, files in os.walk(PKG):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".h")):
                src = open(os.path.join(dirpath, f)).read()
                assert not re.search(r"^\s*(import|from)\s+oracle\b", src, re.M), f
                assert "liboracle" not in src, f
                assert "oracle.h" not in src, f


def test_catalog_bits_equal_oracle_catalog():
    """The device catalog (csrc/catalog.cpp) and the oracle's f32 catalog
    are derived independently from the table of lib.rs:125-268 under the
    reference's const-eval rounding rules; every constant must agree bit for
    bit.  Both are host code, so this runs without a GPU."""
    import oracle
    from py_propag import _native

    oracle.build(ref=False)
    lib = _native.load(require_torch=False)
    for code in range(14):
        out = np.zeros(16, np.float32)
        n = lib.pg_catalog_f32(code, out.ctypes.data)
        n_o, c_o = oracle.catalog_f32(code)
        assert n == n_o
        if n:
            assert np.array_equal(out.view(np.uint32), c_o.view(np.uint32)), code
    assert lib.pg_catalog_f32(14, np.zeros(16, np.float32).ctypes.data) == -1
