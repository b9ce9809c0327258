"""ctypes binding of libpropag_hip.so (include/propag_hip.h).

The library must be built (``__graft_entry__.build()``); there is no CPU
fallback: loading fails loudly when the .so is missing.  torch is imported
before the library so that both share one HIP runtime (torch's bundled
libamdhip64.so.7 satisfies the library's dependency by SONAME).
"""
from __future__ import annotations

import ctypes as C
import os

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "libpropag_hip.so")

EXPORTS = ("pg_last_error", "pg_version", "pg_catalog_f32", "pg_burn", "pg_propagate",
           "pg_propagate_host", "pg_ctx_create", "pg_ctx_destroy", "pg_terrain_add",
           "pg_scenario_bytes", "pg_scenario_admit", "pg_batch_run", "pg_scenario_status",
           "pg_scenario_result", "pg_scenario_release", "pg_ctx_take_finished", "pg_moisture",
           "pg_rasterize_times", "pg_rasterize_fuels", "pg_init_refs", "pg_write_times_geotiff",
           "pg_write_refs_gpkg", "pg_write_tiles_gpkg", "pg_propagate_noalloc", "pg_shutdown", "pg_transform_points",
           "pg_ctx_set_tile")


class GT(C.Structure):
    _fields_ = [("x0", C.c_float), ("dx", C.c_float), ("rx", C.c_float),
                ("y0", C.c_float), ("ry", C.c_float), ("dy", C.c_float)]


class TimeFeatureC(C.Structure):
    """FFITimeFeature (crates/propag/src/propag.rs:120-125)."""
    _fields_ = [("time", C.c_float), ("geom_wkb", C.c_void_p), ("geom_wkb_len", C.c_size_t)]


class FuelFeatureC(C.Structure):
    """FFIFuelFeature (crates/propag/src/propag.rs:150-155)."""
    _fields_ = [("code", C.c_uint8), ("geom_wkb", C.c_void_p), ("geom_wkb_len", C.c_size_t)]


class GeoTransform(C.Structure):
    _fields_ = [("gt", GT), ("inv", GT)]


class GeoReferenceC(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32),
                ("proj", C.c_uint8 * 1024), ("transform", GeoTransform)]


class SettingsC(C.Structure):
    _fields_ = [("geo_ref", GeoReferenceC), ("max_time", C.c_float)]


class Weather(C.Structure):
    _fields_ = [("d1hr", C.c_float), ("d10hr", C.c_float), ("d100hr", C.c_float),
                ("herb", C.c_float), ("wood", C.c_float), ("wind_speed", C.c_float),
                ("wind_azimuth", C.c_float), ("wind_speed_field", C.c_void_p),
                ("wind_azimuth_field", C.c_void_p)]


class Ignition(C.Structure):
    _fields_ = [("x", C.c_uint32), ("y", C.c_uint32), ("time", C.c_float)]


class Tuning(C.Structure):
    _fields_ = [("tile_w", C.c_uint32), ("tile_h", C.c_uint32), ("workgroups", C.c_uint32),
                ("max_iterations", C.c_uint32), ("mode", C.c_uint32),
                ("spin_limit_log2", C.c_uint32), ("waves_per_group", C.c_uint32),
                ("profile", C.c_uint32)]


class DayWeather(C.Structure):
    _fields_ = [("temperature", C.c_float * 24), ("humidity", C.c_float * 24),
                ("cloud_cover", C.c_float * 24), ("rain_6_days", C.c_float * 6), ("month", C.c_int32)]


class Stats(C.Structure):
    _fields_ = [("iterations", C.c_uint64), ("relaxations", C.c_uint64),
                ("improvements", C.c_uint64), ("dda_steps", C.c_uint64),
                ("tiles_visited", C.c_uint64), ("elapsed_ms", C.c_double),
                ("relax_ms", C.c_double), ("work_ms", C.c_double), ("wait_ms", C.c_double),
                ("crit_ms", C.c_double),
                ("workgroups", C.c_uint32), ("lds_bitset", C.c_uint32),
                ("phase_ms", C.c_double * 4), ("barrier_ms", C.c_double * 4),
                ("tile_marks", C.c_uint64), ("phase_total_ms", C.c_double * 4), ("idle_tiles", C.c_uint64),
                ("idle_far_tiles", C.c_uint64)]

    def as_dict(self):
        arrays = ("phase_ms", "barrier_ms", "phase_total_ms")
        d = {f: getattr(self, f) for f, _ in self._fields_ if f not in arrays}
        for f in arrays:
            d[f] = list(getattr(self, f))
        return d


STATUS = {0: "PG_OK", 1: "PG_ERR_INVALID", 2: "PG_ERR_HIP", 3: "PG_ERR_NO_DEVICE",
          4: "PG_ERR_NOT_CONVERGED", 5: "PG_ERR_OUT_OF_MEMORY", 6: "PG_ERR_TIMEOUT",
          7: "PG_ERR_BUSY", 8: "PG_ERR_IO"}

_lib = None


def load(require_torch: bool = True):
    """Loads the library once.  Raises OSError when it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if require_torch:
        try:
            import torch  # noqa: F401  (share torch's HIP runtime)
        except ImportError:
            pass
    if not os.path.exists(LIB_PATH):
        raise OSError(f"{LIB_PATH} is not built; run __graft_entry__.build() "
                      "(there is no CPU fallback for the propagation path)")
    lib = C.CDLL(LIB_PATH)
    vp, sz = C.c_void_p, C.c_size_t
    lib.pg_last_error.restype = C.c_char_p
    lib.pg_version.restype = C.c_char_p
    lib.pg_catalog_f32.restype = C.c_int
    lib.pg_catalog_f32.argtypes = [C.c_int, C.c_void_p]
    lib.pg_burn.restype = C.c_int
    lib.pg_burn.argtypes = [sz] + [vp] * 14
    lib.pg_moisture.restype = C.c_int
    lib.pg_moisture.argtypes = [sz, vp, vp, vp, C.POINTER(DayWeather), C.c_int] + [vp] * 6
    lib.pg_propagate.restype = C.c_int
    lib.pg_propagate.argtypes = [C.POINTER(SettingsC)] + [vp] * 7 + [C.POINTER(Tuning),
                                                                     C.POINTER(Stats), vp]
    lib.pg_transform_points.restype = C.c_int
    lib.pg_transform_points.argtypes = [C.c_char_p, C.c_char_p, sz, vp, vp, vp]
    lib.pg_shutdown.restype = C.c_int
    lib.pg_shutdown.argtypes = []
    lib.pg_propagate_noalloc.restype = C.c_int
    lib.pg_propagate_noalloc.argtypes = lib.pg_propagate.argtypes
    lib.pg_propagate_host.restype = C.c_int
    lib.pg_propagate_host.argtypes = [C.POINTER(SettingsC)] + [vp] * 15 + [C.POINTER(Tuning),
                                                                           C.POINTER(Stats)]
    i32, u32 = C.c_int, C.c_uint32
    for name, args in (("pg_ctx_create", [sz, C.POINTER(vp)]),
                       ("pg_ctx_destroy", [vp]),
                       ("pg_ctx_set_tile", [vp, u32, u32]),
                       ("pg_terrain_add", [vp, C.POINTER(GeoReferenceC), vp, vp, vp, C.POINTER(i32)]),
                       ("pg_scenario_admit", [vp, i32, C.POINTER(Weather), vp, sz, C.c_float, u32,
                                              C.POINTER(i32)]),
                       ("pg_batch_run", [vp, u32, C.POINTER(Tuning), C.POINTER(Stats), C.POINTER(u32)]),
                       ("pg_scenario_status", [vp, i32, C.POINTER(i32), C.POINTER(C.c_uint64)]),
                       ("pg_scenario_result", [vp, i32, vp, vp, vp, vp]),
                       ("pg_scenario_release", [vp, i32]),
                       ("pg_ctx_take_finished", [vp, C.POINTER(i32), u32, C.POINTER(u32)])):
        f = getattr(lib, name)
        f.restype = C.c_int
        f.argtypes = args
    lib.pg_scenario_bytes.restype = sz
    lib.pg_scenario_bytes.argtypes = [u32, u32]
    lib.pg_rasterize_times.restype = C.c_int
    lib.pg_rasterize_times.argtypes = [C.POINTER(GeoReferenceC), C.POINTER(TimeFeatureC), sz, vp,
                                       C.POINTER(u32)]
    lib.pg_rasterize_fuels.restype = C.c_int
    lib.pg_rasterize_fuels.argtypes = [C.POINTER(GeoReferenceC), C.POINTER(FuelFeatureC), sz, vp,
                                       C.POINTER(u32)]
    lib.pg_init_refs.restype = C.c_int
    lib.pg_init_refs.argtypes = [C.POINTER(GeoReferenceC), vp, vp, vp]
    lib.pg_write_times_geotiff.restype = C.c_int
    lib.pg_write_times_geotiff.argtypes = [C.c_char_p, C.POINTER(GeoReferenceC), vp]
    lib.pg_write_refs_gpkg.restype = C.c_int
    lib.pg_write_refs_gpkg.argtypes = [C.c_char_p, C.POINTER(GeoReferenceC), vp, vp, vp]
    lib.pg_write_tiles_gpkg.restype = C.c_int
    lib.pg_write_tiles_gpkg.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(GeoReferenceC), u32, u32]
    _lib = lib
    return lib


def exported_symbols():
    """Names from include/propag_hip.h that the loaded library exports."""
    lib = load(require_torch=False)
    found = []
    for name in EXPORTS:
        try:
            getattr(lib, name)
            found.append(name)
        except AttributeError:
            pass
    return found


def last_error() -> str:
    msg = load().pg_last_error()
    return msg.decode() if msg else ""
