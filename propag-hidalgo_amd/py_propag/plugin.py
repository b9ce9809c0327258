"""ctypes binding of libpropag.so, the reference's plugin ABI
(include/propag_host.h; crates/propag/src/propag.rs:120-311, 752-927).

``run`` drives FFIPropagation_run the way the QGIS plugin does
(qgis-plugin/propag-algo/src/qgsalgorithmpropag.cpp:162-258): a Settings, WKB
ignitions with their CRS, a terrain callback, and output paths.  The library
must be built; there is no fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from ._native import PKG_DIR, FuelFeatureC, GeoReferenceC, SettingsC, TimeFeatureC
from ._native import load as _load_hip

PLUGIN_PATH = os.path.join(PKG_DIR, "libpropag.so")

F32P = C.POINTER(C.c_float)


class FFITerrain(C.Structure):
    """propag.rs:763-775."""
    _fields_ = [("fuel_code", C.POINTER(C.c_uint8))] + [
        (k, F32P) for k in ("d1hr", "d10hr", "d100hr", "herb", "wood", "wind_speed", "wind_azimuth",
                            "slope", "aspect")]


LoadFn = C.CFUNCTYPE(C.c_bool, C.c_void_p, C.POINTER(GeoReferenceC), C.POINTER(FFITerrain))


class FFITerrainLoader(C.Structure):
    _fields_ = [("load", LoadFn), ("context", C.c_void_p)]


class FFIPropagation(C.Structure):
    """propag.rs:177-188."""
    _fields_ = [("settings", SettingsC), ("output_path", C.c_char_p), ("refs_output_path", C.c_char_p),
                ("block_boundaries_out_path", C.c_char_p), ("grid_boundaries_out_path", C.c_char_p),
                ("initial_ignited_elements", C.POINTER(TimeFeatureC)),
                ("initial_ignited_elements_len", C.c_size_t),
                ("initial_ignited_elements_crs", C.c_char_p), ("terrain_loader", FFITerrainLoader)]


class FireSimpleCuda(C.Structure):
    _fields_ = [("speed_max", C.c_float), ("azimuth_max", C.c_float), ("eccentricity", C.c_float)]


RunFn = C.CFUNCTYPE(C.c_bool, FFIPropagation, C.c_size_t, C.c_char_p)
RasterizeFuelsFn = C.CFUNCTYPE(C.c_bool, C.POINTER(FuelFeatureC), C.c_size_t, C.c_char_p,
                               C.POINTER(GeoReferenceC), C.c_void_p, C.c_char_p, C.c_size_t)


class PluginT(C.Structure):
    _fields_ = [("run", RunFn), ("rasterize_fuels", RasterizeFuelsFn)]


EXPORTS = ("FFIPropagation_run", "propag_rasterize_fuels", "create_settings", "create_fire",
           "create_plugin", "GeoReference_len", "GeoReference_x0", "GeoReference_x1",
           "GeoReference_y0", "GeoReference_y1", "GeoReference_south_up", "GeoReference_north_up")

_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    _load_hip(require_torch=True)      # the data-plane library first (shares torch's HIP runtime)
    if not os.path.exists(PLUGIN_PATH):
        raise OSError(f"{PLUGIN_PATH} is not built; run __graft_entry__.build()")
    lib = C.CDLL(PLUGIN_PATH)
    gp = C.POINTER(GeoReferenceC)
    lib.FFIPropagation_run.restype = C.c_bool
    lib.FFIPropagation_run.argtypes = [FFIPropagation, C.c_size_t, C.c_char_p]
    lib.propag_rasterize_fuels.restype = C.c_bool
    lib.propag_rasterize_fuels.argtypes = [C.POINTER(FuelFeatureC), C.c_size_t, C.c_char_p, gp,
                                           C.c_void_p, C.c_char_p, C.c_size_t]
    lib.create_settings.restype = SettingsC
    lib.create_settings.argtypes = [GeoReferenceC, C.c_float]
    lib.create_fire.restype = FireSimpleCuda
    lib.create_fire.argtypes = [C.c_float] * 3
    lib.create_plugin.restype = PluginT
    lib.create_plugin.argtypes = []
    lib.GeoReference_len.restype = C.c_uint32
    lib.GeoReference_len.argtypes = [gp]
    for n in ("x0", "x1", "y0", "y1"):
        f = getattr(lib, "GeoReference_" + n)
        f.restype = C.c_float
        f.argtypes = [gp]
    for n in ("south_up", "north_up"):
        f = getattr(lib, "GeoReference_" + n)
        f.restype = C.c_bool
        f.argtypes = [C.c_float] * 6 + [C.c_char_p, gp]
    _lib = lib
    return lib


TERRAIN_COLUMNS = ("d1hr", "d10hr", "d100hr", "herb", "wood", "wind_speed", "wind_azimuth", "slope", "aspect")


def make_loader(fill):
    """Wraps ``fill(geo: GeoReferenceC, columns: dict of numpy views) -> bool``
    as an FFITerrainLoader; the views alias the library's buffers (fuel u8
    and the nine f32 columns, W*H each).  Keep the returned object alive
    for the duration of the run."""

    def _load(_ctx, geo_p, ter_p):
        geo = geo_p.contents
        n = geo.width * geo.height
        t = ter_p.contents
        cols = {"fuel": np.ctypeslib.as_array(t.fuel_code, (n,))}
        for k in TERRAIN_COLUMNS:
            cols[k] = np.ctypeslib.as_array(getattr(t, k), (n,))
        return bool(fill(geo, cols))

    cb = LoadFn(_load)
    return FFITerrainLoader(cb, None), cb


def run(settings, output_path, ignitions, crs: str, fill, refs_output_path=None,
        block_boundaries_out_path=None, grid_boundaries_out_path=None, err_len: int = 1024):
    """FFIPropagation_run.  ``ignitions``: (time s, WKB bytes) pairs;
    ``fill``: the terrain callback of make_loader.  Returns (ok, message)."""
    lib = load()
    keep = [bytes(b) for _, b in ignitions]
    feats = (TimeFeatureC * max(1, len(keep)))()
    for i, ((t, _), b) in enumerate(zip(ignitions, keep)):
        feats[i].time = float(t)
        feats[i].geom_wkb = C.cast(C.c_char_p(b), C.c_void_p) if b else None
        feats[i].geom_wkb_len = len(b)
    loader, cb = make_loader(fill)
    enc = lambda p: None if p is None else os.fsencode(p)
    p = FFIPropagation(settings.to_c() if hasattr(settings, "to_c") else settings, enc(output_path),
                       enc(refs_output_path), enc(block_boundaries_out_path), enc(grid_boundaries_out_path),
                       feats, len(keep), crs.encode() if crs is not None else None, loader)
    err = C.create_string_buffer(max(1, err_len))
    ok = lib.FFIPropagation_run(p, err_len, err)
    del cb
    return bool(ok), err.value.decode(errors="replace")
