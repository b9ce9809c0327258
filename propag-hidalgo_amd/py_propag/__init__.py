"""py_propag -- Python front of the MI355X implementation of propag25's GPU stages.

Names follow the reference's Python package (crates/py-propag/py_propag/
__init__.py:6-29 and crates/py-propag/src/lib.rs): ``GeoReference``,
``Settings``, ``TimeFeature``, ``PropagResults``, ``propagate``,
``load_terrain_data``, ``check_cuda_available`` and the exception types.
Unlike the reference's mock (py-propag/src/lib.rs:676-691), ``propagate``
runs the real burn + propagation on the GPU through libpropag_hip.so.
Ignitions are WKB features (rasterised by the library with GDAL's default
burn rules, ``rasterize_times``), raster cells, or both.

Two levels:
  * host arrays (numpy): ``burn``, ``propagate`` -- the whole device data
    plane of the reference's propagate() (crates/propag/src/propag.rs:375-573);
  * device tensors (torch, cuda:N): ``burn_device``, ``propagate_device`` --
    the C-ABI entry points pg_burn / pg_propagate on resident HBM buffers.
"""
from __future__ import annotations

import atexit
import ctypes as C
import math
import os
import struct
import weakref
from dataclasses import dataclass, field

import numpy as np

from . import _native
from ._native import Stats, Tuning

__version__ = "0.1.0"

FLT_MAX = np.float32(np.finfo(np.float32).max)
NO_REF = 0xFFFF
TERRAIN_FIELDS = ("d1hr", "d10hr", "d100hr", "herb", "wood", "wind_speed",
                  "wind_azimuth", "slope", "aspect")


class PropagCudaError(RuntimeError):
    """A GPU runtime failure (name kept from the reference; this is HIP)."""


PropagGpuError = PropagCudaError


class PropagGdalError(RuntimeError):
    """Kept for API compatibility; raster/vector I/O is not part of this package."""


class PropagValidationError(ValueError):
    """Invalid arguments (shapes, ranges, layout)."""


class PropagMPIError(RuntimeError):
    """Kept for API compatibility; no MPI is used."""


class PropagNotConverged(PropagCudaError):
    """The Jacobi sweep cap was reached before a quiet sweep."""


def _raise_for(status: int, what: str):
    if status == 0:
        return
    msg = f"{what}: {_native.STATUS.get(status, status)}: {_native.last_error()}"
    if status == 1:
        raise PropagValidationError(msg)
    if status == 4:
        raise PropagNotConverged(msg)
    if status == 8:
        raise PropagGdalError(msg)     # the reference's error type for its output files
    raise PropagCudaError(msg)


# ------------------------------------------------------------ geometry

def _invert(gt):
    """GT::invert (crates/geometry/src/lib.rs:86-125) in f32."""
    x0, dx, rx, y0, ry, dy = (np.float32(v) for v in gt)
    if rx == 0 and ry == 0 and dx != 0 and dy != 0:
        return [-x0 / dx, np.float32(1) / dx, np.float32(0), -y0 / dy, np.float32(0),
                np.float32(1) / dy]
    det = dx * dy - rx * ry
    mag = max(abs(dx), abs(rx), abs(ry), abs(dy))
    if abs(det) <= 1e-10 * mag * mag:
        raise PropagValidationError("geo transform is not invertible")
    inv = np.float32(1) / det
    return [(rx * y0 - x0 * dy) * inv, dy * inv, -rx * inv, (-dx * y0 + x0 * ry) * inv,
            -ry * inv, dx * inv]


class GeoReference:
    """Raster geometry (crates/geometry/src/lib.rs:180-187).

    transform = [x_origin, pixel_width, x_rotation, y_origin, y_rotation,
    pixel_height]; a negative pixel_height is north-up."""

    def __init__(self, width: int, height: int, proj: str, transform):
        # The checks and messages of PyGeoReference::new
        # (crates/py-propag/src/lib.rs:272-330).
        if width <= 0 or height <= 0:
            raise PropagValidationError("Grid dimensions must be positive (width > 0, height > 0)")
        if width > 32768 or height > 32768:
            raise PropagValidationError("Grid dimensions too large (max 32768x32768)")
        if not proj:
            raise PropagValidationError("Projection string cannot be empty")
        pb = proj.encode()
        if len(pb) >= 1024:
            raise PropagValidationError("Projection string too large (max 1023 characters)")
        t = [float(v) for v in np.asarray(transform, dtype=np.float64).ravel()]
        if len(t) != 6:
            raise PropagValidationError("Geo transform must have exactly 6 elements [x_origin, pixel_width, "
                                        "x_rotation, y_origin, y_rotation, pixel_height]")
        if t[1] == 0.0 or t[5] == 0.0:
            raise PropagValidationError("Invalid geo transform: pixel size cannot be zero")
        if not all(math.isfinite(float(np.float32(v))) for v in t):
            raise PropagValidationError("Invalid geo transform parameters - ensure pixel sizes are "
                                        "non-zero and finite")
        self._width, self._height, self._proj = int(width), int(height), proj
        self._gt = [np.float32(v) for v in t]
        self._inv = _invert(self._gt)

    @classmethod
    def south_up(cls, bbox, pixel_size, proj):
        """GeoReference::south_up (crates/geometry/src/lib.rs:239-257)."""
        (x0, y0), (x1, y1) = bbox
        dx, dy = pixel_size
        w = math.ceil(np.float32(np.float32(x1 - x0) / np.float32(dx)))
        h = math.ceil(np.float32(np.float32(y1 - y0) / np.float32(dy)))
        return cls(w, h, proj, [x0, dx, 0.0, y0, 0.0, dy])

    @classmethod
    def north_up(cls, bbox, pixel_size, proj):
        """GeoReference::north_up (crates/geometry/src/lib.rs:258-276)."""
        (x0, y0), (x1, y1) = bbox
        dx, dy = pixel_size
        w = math.ceil(np.float32(np.float32(x1 - x0) / np.float32(dx)))
        h = math.ceil(np.float32(np.float32(y1 - y0) / np.float32(dy)))
        return cls(w, h, proj, [x0, dx, 0.0, y1, 0.0, -dy])

    width = property(lambda self: self._width)
    height = property(lambda self: self._height)
    proj = property(lambda self: self._proj)
    transform = property(lambda self: np.array(self._gt, dtype=np.float64))
    is_north_up = property(lambda self: self._gt[5] < 0)

    def __len__(self):
        return self._width * self._height

    def to_c(self) -> _native.GeoReferenceC:
        g = _native.GeoReferenceC()
        g.width, g.height = self._width, self._height
        pb = self._proj.encode()
        C.memmove(g.proj, pb, len(pb))
        g.transform.gt = _native.GT(*[float(v) for v in self._gt])
        g.transform.inv = _native.GT(*[float(v) for v in self._inv])
        return g

    def __repr__(self):
        return f"GeoReference({self._width}x{self._height}, transform={list(map(float, self._gt))})"


class Settings:
    """Settings{geo_ref, max_time} (crates/propag/src/propag.rs:38-43)."""

    def __init__(self, geo_ref: GeoReference, max_time: float):
        if not isinstance(geo_ref, GeoReference):
            raise PropagValidationError("geo_ref must be a GeoReference")
        self.geo_ref = geo_ref
        self.max_time = max_time

    @property
    def max_time(self):
        return self._max_time

    @max_time.setter
    def max_time(self, v):
        # PySettings::new / set_max_time (crates/py-propag/src/lib.rs:183-233)
        v = np.float32(v)
        if v <= 0:
            raise PropagValidationError("Maximum simulation time must be positive")
        if not np.isfinite(v):
            raise PropagValidationError("Maximum simulation time must be a finite number")
        if v > np.float32(10.0 * 365.0 * 24.0 * 3600.0):
            raise PropagValidationError("Maximum simulation time too large (limit: 10 years)")
        self._max_time = v

    def to_c(self) -> _native.SettingsC:
        s = _native.SettingsC()
        s.geo_ref = self.geo_ref.to_c()
        s.max_time = float(self._max_time)
        return s


@dataclass
class TimeFeature:
    """An ignition: a time (s) and WKB geometry (TimeFeature, propag.rs:115-118)
    and/or raster cells (an iterable of (x, y)).  WKB is in the raster's
    coordinates and burns as the reference's rasterize_times does
    (propag.rs:837-890)."""
    time: float
    geom_wkb: bytes = b""
    cells: list = field(default_factory=list)

    def __post_init__(self):
        # PyTimeFeature::new (crates/py-propag/src/lib.rs:392-418); a feature
        # given only as raster cells carries no WKB.
        t = np.float32(self.time)
        if t < 0:
            raise PropagValidationError("Time must be non-negative")
        if not np.isfinite(t):
            raise PropagValidationError("Time must be a finite number")
        self.geom_wkb = bytes(self.geom_wkb)
        if not self.cells:
            if not self.geom_wkb:
                raise PropagValidationError("Geometry WKB cannot be empty")
            if len(self.geom_wkb) < 9:
                raise PropagValidationError("Invalid geometry WKB - too short for valid geometry")


@dataclass
class PropagResults:
    """Outputs of one propagation (crates/propag/src/propag.rs:311-321)."""
    arrival_times: np.ndarray     # (height, width) f32, FLT_MAX = not burned
    boundary_change: np.ndarray   # ceil(W*H/32) u32 words
    refs_x: np.ndarray            # W*H u16
    refs_y: np.ndarray            # W*H u16
    geo_ref: GeoReference
    stats: dict
    fire: tuple | None = None     # (speed_max, azimuth_max, eccentricity) rasters

    @property
    def shape(self):
        return self.geo_ref.height, self.geo_ref.width


def check_cuda_available() -> bool:
    """True when a gfx950 GPU is visible (name kept from the reference,
    crates/py-propag/src/lib.rs:26-56, which asks nvidia-smi and falls back
    to the device node): no KFD device node means no GPU."""
    if not os.path.exists("/dev/kfd"):
        return False
    try:
        import torch
        return bool(torch.cuda.is_available()) and "gfx950" in torch.cuda.get_device_properties(0).gcnArchName
    except Exception:
        return False


check_gpu_available = check_cuda_available


def library_version() -> str:
    return _native.load().pg_version().decode()


def load_terrain_data(width, height, proj, transform, elevation=None, slope=None, aspect=None,
                      fuel_model=None) -> GeoReference:
    """Validation front of the reference (py-propag/src/lib.rs:525-577)."""
    if width <= 0 or height <= 0:
        raise PropagValidationError("Grid dimensions must be positive")
    if width * height > 100_000_000:
        raise PropagValidationError("Grid too large (limit: 100 million cells).")
    for name, arr, lo, hi in (("Elevation", elevation, -1000.0, 9000.0),
                              ("Slope", slope, 0.0, 90.0), ("Aspect", aspect, 0.0, 360.0),
                              ("Fuel model", fuel_model, None, None)):
        if arr is None:
            continue
        arr = np.asarray(arr)
        if arr.shape != (height, width):
            raise PropagValidationError(f"{name} array shape mismatch: expected ({height}, {width}), got {arr.shape}")
        if lo is not None and (arr < lo).any():
            raise PropagValidationError(f"{name} value below minimum {lo}")
        if hi is not None and (arr > hi).any():
            raise PropagValidationError(f"{name} value exceeds maximum {hi}")
    if fuel_model is not None:
        _validate_fuel(np.asarray(fuel_model))
    return GeoReference(width, height, proj, transform)


def _validate_fuel(fuel):
    bad = np.nonzero(fuel.ravel() > 13)[0]
    if bad.size:
        i = int(bad[0])
        raise PropagValidationError(f"Invalid fuel model ID {int(fuel.ravel()[i])} at index {i}. "
                                    "Valid NFFL fuel models are 1-13.")


def _terrain_columns(terrain: dict, n: int):
    fuel = np.ascontiguousarray(np.asarray(terrain["fuel"], dtype=np.uint8).ravel())
    if fuel.size != n:
        raise PropagValidationError(f"fuel raster has {fuel.size} cells, expected {n}")
    cols = []
    for k in TERRAIN_FIELDS:
        if k not in terrain:
            raise PropagValidationError(f"terrain is missing '{k}'")
        v = np.asarray(terrain[k], dtype=np.float32)
        if v.size == 1:
            v = np.full(n, v.item(), np.float32)
        v = np.ascontiguousarray(v.ravel())
        if v.size != n:
            raise PropagValidationError(f"terrain '{k}' has {v.size} cells, expected {n}")
        cols.append(v)
    return fuel, cols


def _feature_array(features, cls, value_of):
    """ctypes array of FFI feature records; keeps the WKB bytes alive."""
    keep = [bytes(wkb) for _, wkb in features]
    arr = (cls * max(1, len(features)))()
    for i, ((v, _), b) in enumerate(zip(features, keep)):
        setattr(arr[i], cls._fields_[0][0], value_of(v))
        arr[i].geom_wkb = C.cast(C.c_char_p(b), C.c_void_p) if b else None
        arr[i].geom_wkb_len = len(b)
    return arr, keep


def rasterize_times(geo_ref: GeoReference, features):
    """Burns (time, WKB) features into a (height, width) f32 raster (FLT_MAX
    where nothing burned) with GDAL's default rules, later features
    replacing earlier ones; returns (raster, number of features that
    parsed).  Replaces rasterize_times (crates/propag/src/propag.rs:837-890)."""
    feats = [(f.time, f.geom_wkb) if isinstance(f, TimeFeature) else tuple(f) for f in features]
    arr, keep = _feature_array(feats, _native.TimeFeatureC, float)
    out = np.empty((geo_ref.height, geo_ref.width), np.float32)
    n = C.c_uint32(0)
    lib = _native.load(require_torch=False)
    g = geo_ref.to_c()
    _raise_for(lib.pg_rasterize_times(C.byref(g), arr, len(feats), _ptr(out), C.byref(n)),
               "pg_rasterize_times")
    return out, int(n.value)


def rasterize_fuels(geo_ref: GeoReference, features):
    """Burns (fuel code, WKB) features into a (height, width) u8 raster (255
    where nothing burned): propag_rasterize_fuels (propag.rs:893-978)."""
    feats = [tuple(f) for f in features]
    for code, _ in feats:
        if not 0 <= int(code) <= 255:
            raise PropagValidationError(f"fuel code {code} does not fit a byte")
    arr, keep = _feature_array(feats, _native.FuelFeatureC, int)
    out = np.empty((geo_ref.height, geo_ref.width), np.uint8)
    n = C.c_uint32(0)
    lib = _native.load(require_torch=False)
    g = geo_ref.to_c()
    _raise_for(lib.pg_rasterize_fuels(C.byref(g), arr, len(feats), _ptr(out), C.byref(n)),
               "pg_rasterize_fuels")
    return out, int(n.value)


def write_times_geotiff(path, geo_ref: GeoReference, time) -> None:
    """Arrival times as the reference's GeoTIFF (pg_write_times_geotiff)."""
    t = np.ascontiguousarray(np.asarray(time, np.float32).reshape(-1))
    if t.size != len(geo_ref):
        raise PropagValidationError("time raster does not match the geo reference")
    g = geo_ref.to_c()
    _raise_for(_native.load(require_torch=False).pg_write_times_geotiff(os.fsencode(path), C.byref(g), _ptr(t)),
               "pg_write_times_geotiff")


def write_refs_gpkg(path, geo_ref: GeoReference, time, refs_x, refs_y) -> None:
    """The reference's "fire_references" GeoPackage layer (pg_write_refs_gpkg)."""
    arrs = [np.ascontiguousarray(np.asarray(a, dt).reshape(-1))
            for a, dt in ((time, np.float32), (refs_x, np.uint16), (refs_y, np.uint16))]
    if any(a.size != len(geo_ref) for a in arrs):
        raise PropagValidationError("rasters do not match the geo reference")
    g = geo_ref.to_c()
    _raise_for(_native.load(require_torch=False).pg_write_refs_gpkg(os.fsencode(path), C.byref(g),
                                                                    *[_ptr(a) for a in arrs]),
               "pg_write_refs_gpkg")


def write_tiles_gpkg(path, layer: str, geo_ref: GeoReference, tile_w: int, tile_h: int) -> None:
    """A "blocks" / "grids" polygon layer (pg_write_tiles_gpkg)."""
    g = geo_ref.to_c()
    _raise_for(_native.load(require_torch=False).pg_write_tiles_gpkg(os.fsencode(path), layer.encode(), C.byref(g),
                                                                     int(tile_w), int(tile_h)),
               "pg_write_tiles_gpkg")


def _point_wkb(x: float, y: float) -> bytes:
    return struct.pack("<BIdd", 1, 1, x, y)


def initial_state(width: int, height: int, ignitions, geo_ref: GeoReference | None = None):
    """time / refs as propag.rs:433-449 builds them from the ignitions.
    ``ignitions``: iterable of (x, y, t) cells or TimeFeature.  When any
    TimeFeature carries WKB (needs ``geo_ref``), every ignition burns through
    the library's rasteriser in the given order (cells as points at their
    centres) and the references come from pg_init_refs."""
    ignitions = list(ignitions)
    if any(isinstance(i, TimeFeature) and i.geom_wkb for i in ignitions):
        if geo_ref is None or (geo_ref.width, geo_ref.height) != (width, height):
            raise PropagValidationError("WKB ignitions need the raster's GeoReference")
        gt = [float(v) for v in geo_ref.transform]
        feats = []
        for ign in ignitions:
            if isinstance(ign, TimeFeature):
                if ign.geom_wkb:
                    feats.append((ign.time, ign.geom_wkb))
                pts = [(x, y, ign.time) for x, y in ign.cells]
            else:
                pts = [ign]
            for x, y, t in pts:
                if not (0 <= x < width and 0 <= y < height):
                    raise PropagValidationError(f"ignition cell ({x}, {y}) is outside the raster")
                cx, cy = x + 0.5, y + 0.5
                feats.append((t, _point_wkb(gt[0] + cx * gt[1] + cy * gt[2], gt[3] + cx * gt[4] + cy * gt[5])))
        time, _ = rasterize_times(geo_ref, feats)
        time = time.ravel()
        rx = np.empty(width * height, np.uint16)
        ry = np.empty(width * height, np.uint16)
        lib = _native.load(require_torch=False)
        g = geo_ref.to_c()
        _raise_for(lib.pg_init_refs(C.byref(g), _ptr(time), _ptr(rx), _ptr(ry)), "pg_init_refs")
        return time, rx, ry
    n = width * height
    time = np.full(n, FLT_MAX, np.float32)
    rx = np.full(n, NO_REF, np.uint16)
    ry = np.full(n, NO_REF, np.uint16)
    for ign in ignitions:
        if isinstance(ign, TimeFeature):
            pts = [(x, y, ign.time) for x, y in ign.cells]
        else:
            pts = [ign]
        for x, y, t in pts:
            if not (0 <= x < width and 0 <= y < height):
                raise PropagValidationError(f"ignition cell ({x}, {y}) is outside the raster")
            i = int(x) + int(y) * width
            time[i] = np.float32(t)
            rx[i], ry[i] = x, y
    return time, rx, ry


def _tuning(tuning):
    if tuning is None:
        return None
    if isinstance(tuning, Tuning):
        return tuning
    t = Tuning()
    names = {f[0] for f in Tuning._fields_}
    for k, v in dict(tuning).items():
        if k not in names:
            raise PropagValidationError(f"unknown tuning field '{k}'")
        setattr(t, k, int(v))
    return t


def _ptr(a):
    return C.c_void_p(a.ctypes.data)


def propagate(settings: Settings, terrain: dict, ignitions=None, time=None, refs_x=None,
              refs_y=None, tuning=None, return_fire: bool = False) -> PropagResults:
    """Burn pass + propagation on the GPU from host rasters.

    ``terrain``: dict with 'fuel' (u8) and the TerrainCuda columns
    (TERRAIN_FIELDS; scalars broadcast).  Either ``ignitions`` or an initial
    (time, refs_x, refs_y) state must be given."""
    g = settings.geo_ref
    n = len(g)
    fuel, cols = _terrain_columns(terrain, n)
    _validate_fuel(fuel)
    if time is None:
        if not ignitions:
            raise PropagValidationError("At least one ignition element is required")
        time, refs_x, refs_y = initial_state(g.width, g.height, ignitions, g)
    time = np.ascontiguousarray(np.asarray(time, np.float32).ravel()).copy()
    refs_x = np.ascontiguousarray(np.asarray(refs_x, np.uint16).ravel()).copy()
    refs_y = np.ascontiguousarray(np.asarray(refs_y, np.uint16).ravel()).copy()
    if time.size != n or refs_x.size != n or refs_y.size != n:
        raise PropagValidationError("initial state rasters do not match the geo reference")
    bits = np.zeros((n + 31) // 32, np.uint32)
    fire = np.zeros(3 * n, np.float32) if return_fire else None
    st = Stats()
    tn = _tuning(tuning)
    lib = _native.load()
    sc = settings.to_c()
    rc = lib.pg_propagate_host(C.byref(sc), _ptr(fuel), *[_ptr(c) for c in cols], _ptr(time),
                               _ptr(refs_x), _ptr(refs_y), _ptr(bits),
                               _ptr(fire) if fire is not None else None,
                               C.byref(tn) if tn is not None else None, C.byref(st))
    _raise_for(rc, "pg_propagate_host")
    fr = tuple(fire.reshape(3, n)) if fire is not None else None
    return PropagResults(time.reshape(g.height, g.width), bits, refs_x, refs_y, g, st.as_dict(), fr)


# --------------------------------------------------------- device level

def _dptr(t):
    import torch
    if not isinstance(t, torch.Tensor) or not t.is_cuda or not t.is_contiguous():
        raise PropagValidationError("expected a contiguous GPU tensor")
    return C.c_void_p(t.data_ptr())


def _stream_ptr(stream):
    if stream is None:
        import torch
        return C.c_void_p(torch.cuda.current_stream().cuda_stream)
    return C.c_void_p(getattr(stream, "cuda_stream", stream))


def burn_device(fuel, d1hr, d10hr, d100hr, herb, wood, wind_speed, wind_azimuth, slope, aspect,
                speed_max, azimuth_max, eccentricity, stream=None):
    """pg_burn on resident tensors (u8 fuel, f32 columns, f32 outputs the
    caller pre-zeroed).  Asynchronous on ``stream`` (torch's current stream
    by default)."""
    n = fuel.numel()
    for t in (d1hr, d10hr, d100hr, herb, wood, wind_speed, wind_azimuth, slope, aspect,
              speed_max, azimuth_max, eccentricity):
        if t.numel() != n:
            raise PropagValidationError("all burn rasters must have the same number of cells")
    rc = _native.load().pg_burn(n, _dptr(fuel), _dptr(d1hr), _dptr(d10hr), _dptr(d100hr),
                                _dptr(herb), _dptr(wood), _dptr(wind_speed), _dptr(wind_azimuth),
                                _dptr(slope), _dptr(aspect), _dptr(speed_max), _dptr(azimuth_max),
                                _dptr(eccentricity), _stream_ptr(stream))
    _raise_for(rc, "pg_burn")


def day_weather(temperature, humidity, cloud_cover, rain_6_days, month) -> _native.DayWeather:
    """pg_day_weather of one day: hourly temperature (C), relative humidity
    (%), cloud cover (%) (24 values or a scalar each), the rain (mm) of the
    six previous days (index 0 = the day before) and the month (1..12)."""
    w = _native.DayWeather()
    for name, v, k in (("temperature", temperature, 24), ("humidity", humidity, 24),
                       ("cloud_cover", cloud_cover, 24), ("rain_6_days", rain_6_days, 6)):
        a = np.broadcast_to(np.asarray(v, np.float32), (k,))
        getattr(w, name)[:] = [float(x) for x in a]
    if not 1 <= int(month) <= 12:
        raise PropagValidationError("month must be 1..12")
    w.month = int(month)
    return w


def moisture_device(fuel, slope, aspect, weather: _native.DayWeather, hour: int, d1hr, d10hr, d100hr,
                    herb, wood, stream=None):
    """pg_moisture on resident tensors: the moisture ratios of every cell at
    ``hour`` (0..23) of the day ``weather`` (see day_weather), written into
    the five f32 outputs (the burn pass's moisture columns).  Asynchronous on
    ``stream``."""
    n = fuel.numel()
    for t in (slope, aspect, d1hr, d10hr, d100hr, herb, wood):
        if t.numel() != n:
            raise PropagValidationError("all moisture rasters must have the same number of cells")
    rc = _native.load().pg_moisture(n, _dptr(fuel), _dptr(slope), _dptr(aspect), C.byref(weather),
                                    int(hour), _dptr(d1hr), _dptr(d10hr), _dptr(d100hr), _dptr(herb),
                                    _dptr(wood), _stream_ptr(stream))
    _raise_for(rc, "pg_moisture")


def propagate_device(settings: Settings, speed_max, azimuth_max, eccentricity, time, refs_x,
                     refs_y, boundaries, tuning=None, stream=None, noalloc: bool = False) -> dict:
    """pg_propagate on resident tensors: time (f32) and refs (16-bit, e.g.
    torch.int16 holding the u16 bits) are in-out, boundaries (32-bit words)
    out.  Synchronises ``stream``; returns the stats dict.  ``noalloc``
    calls pg_propagate_noalloc instead (no device allocation; rasters of at
    most 1,048,576 cells)."""
    n = len(settings.geo_ref)
    for t, want in ((speed_max, n), (azimuth_max, n), (eccentricity, n), (time, n), (refs_x, n),
                    (refs_y, n), (boundaries, (n + 31) // 32)):
        if t.numel() != want:
            raise PropagValidationError("raster sizes do not match the geo reference")
    if refs_x.element_size() != 2 or refs_y.element_size() != 2 or boundaries.element_size() != 4:
        raise PropagValidationError("refs must be 16-bit and boundaries 32-bit tensors")
    st = Stats()
    tn = _tuning(tuning)
    sc = settings.to_c()
    lib = _native.load()
    name = "pg_propagate_noalloc" if noalloc else "pg_propagate"
    rc = getattr(lib, name)(C.byref(sc), _dptr(speed_max), _dptr(azimuth_max),
                            _dptr(eccentricity), _dptr(time), _dptr(refs_x),
                            _dptr(refs_y), _dptr(boundaries),
                            C.byref(tn) if tn is not None else None, C.byref(st),
                            _stream_ptr(stream))
    _raise_for(rc, name)
    return st.as_dict()


def burn(terrain: dict):
    """Stage F on the GPU from host rasters; returns numpy (speed_max,
    azimuth_max, eccentricity), zero where the fuel is not a catalog fuel."""
    import torch
    fuel, cols = _terrain_columns(terrain, np.asarray(terrain["fuel"]).size)
    dev = torch.device("cuda")
    tf = torch.from_numpy(fuel).to(dev)
    tc = [torch.from_numpy(c).to(dev) for c in cols]
    out = [torch.zeros(fuel.size, dtype=torch.float32, device=dev) for _ in range(3)]
    burn_device(tf, *tc, *out)
    torch.cuda.synchronize()
    return tuple(o.cpu().numpy() for o in out)


# The reference module's class names (crates/py-propag/py_propag/__init__.py).
PySettings, PyGeoReference, PyTimeFeature, PyPropagResults = Settings, GeoReference, TimeFeature, PropagResults

__all__ = [
    "Settings", "PySettings", "GeoReference", "PyGeoReference", "TimeFeature", "PyTimeFeature",
    "PropagResults", "PyPropagResults", "propagate", "load_terrain_data",
    "check_cuda_available", "check_gpu_available", "PropagCudaError", "PropagGpuError",
    "PropagGdalError", "PropagValidationError", "PropagMPIError", "PropagNotConverged", "burn",
    "burn_device", "propagate_device", "initial_state", "rasterize_times", "rasterize_fuels",
    "write_times_geotiff", "write_refs_gpkg", "write_tiles_gpkg", "Tuning", "Stats", "FLT_MAX", "NO_REF",
]


# ------------------------------------------------------------ batching / serving

RUNNING, DONE, NOT_CONVERGED, FAILED = 0, 1, 2, 3

_open_contexts = weakref.WeakSet()


@atexit.register
def _release_at_exit():
    """Closes the contexts still open and releases the library's device
    resources while the interpreter and the HIP runtime are whole (instead
    of leaving them to interpreter finalisation and the runtime's exit)."""
    for ctx in list(_open_contexts):
        try:
            ctx.close()
        except Exception:
            pass
    if _native._lib is not None:
        try:
            _native._lib.pg_shutdown()
        except Exception:
            pass


class Context:
    """Device-resident batching context (the pg_ctx_* handle API).

    Terrains (fuel, slope, aspect) are uploaded once; scenarios on them carry
    their own weather, ignitions and max_time, are admitted while the memory
    budget allows, and advance together in shared persistent launches
    (``run``).  A scenario's result equals what ``propagate`` returns for the
    same inputs, whatever else shares its launches."""

    def __init__(self, hbm_budget: int = 0):
        lib = _native.load()
        h = C.c_void_p()
        _raise_for(lib.pg_ctx_create(C.c_size_t(hbm_budget), C.byref(h)), "pg_ctx_create")
        self._lib, self._h = lib, h
        self._shapes = {}     # terrain id -> (height, width)
        self._scen = {}       # scenario id -> terrain id
        _open_contexts.add(self)

    def close(self):
        if self._h:
            self._lib.pg_ctx_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_tile(self, tile_w: int, tile_h: int) -> None:
        """Tile shape (8x8 default, 16x4 or 32x2) of the scenarios admitted
        from now on; refused (PG_ERR_BUSY) while any scenario is admitted.
        Results do not depend on it; ``run``'s tuning may name it or leave
        ``tile_w``/``tile_h`` out."""
        _raise_for(self._lib.pg_ctx_set_tile(self._h, int(tile_w), int(tile_h)), "pg_ctx_set_tile")

    def add_terrain(self, geo_ref: GeoReference, fuel, slope, aspect) -> int:
        n = len(geo_ref)
        fuel = np.ascontiguousarray(np.asarray(fuel, np.uint8).ravel())
        slope = np.ascontiguousarray(np.asarray(slope, np.float32).ravel())
        aspect = np.ascontiguousarray(np.asarray(aspect, np.float32).ravel())
        for name, a in (("fuel", fuel), ("slope", slope), ("aspect", aspect)):
            if a.size != n:
                raise PropagValidationError(f"terrain '{name}' has {a.size} cells, expected {n}")
        _validate_fuel(fuel)
        g = geo_ref.to_c()
        tid = C.c_int()
        _raise_for(self._lib.pg_terrain_add(self._h, C.byref(g), _ptr(fuel), _ptr(slope), _ptr(aspect),
                                            C.byref(tid)), "pg_terrain_add")
        self._shapes[tid.value] = (geo_ref.height, geo_ref.width)
        return tid.value

    def admit(self, terrain_id: int, weather: dict, ignitions, max_time: float,
              max_iterations: int = 0) -> int | None:
        """Admits a scenario; returns its id, or None when the memory budget
        has no room (PG_ERR_BUSY).  ``weather``: d1hr, d10hr, d100hr, herb,
        wood, wind_speed (m/s), wind_azimuth (rad); wind_speed /
        wind_azimuth may be rasters.  ``ignitions``: (x, y, t) cells."""
        w = _native.Weather()
        keep = []
        for k in ("d1hr", "d10hr", "d100hr", "herb", "wood"):
            setattr(w, k, float(weather[k]))
        ws, waz = np.asarray(weather["wind_speed"], np.float32), np.asarray(weather["wind_azimuth"], np.float32)
        if ws.size == 1 and waz.size == 1:
            w.wind_speed, w.wind_azimuth = float(ws.item()), float(waz.item())
        else:
            h, wd = self._shapes[terrain_id]
            n = h * wd

            def field(v):
                v = np.asarray(v, np.float32)
                v = v.reshape(-1) if v.size == n else np.broadcast_to(v, (h, wd))
                return np.ascontiguousarray(v, np.float32).ravel()

            ws, waz = field(ws), field(waz)
            keep += [ws, waz]
            w.wind_speed_field, w.wind_azimuth_field = ws.ctypes.data, waz.ctypes.data
        ign = (_native.Ignition * max(1, len(ignitions)))()
        for i, (x, y, t) in enumerate(ignitions):
            ign[i].x, ign[i].y, ign[i].time = int(x), int(y), float(t)
        sid = C.c_int()
        st = self._lib.pg_scenario_admit(self._h, int(terrain_id), C.byref(w), C.cast(ign, C.c_void_p),
                                         C.c_size_t(len(ignitions)), C.c_float(max_time),
                                         C.c_uint32(max_iterations), C.byref(sid))
        if st == 7:
            return None
        _raise_for(st, "pg_scenario_admit")
        self._scen[sid.value] = terrain_id
        return sid.value

    def run(self, max_sweeps: int = 0, tuning=None):
        """One shared launch of all unfinished scenarios (max_sweeps 0: to
        completion).  Returns (number finished, stats dict).  If the launch
        fails, the scenarios it ran become FAILED (listed by
        ``take_finished``, without results) and the error is raised."""
        stats = Stats()
        fin = C.c_uint32()
        t = _tuning(tuning)
        _raise_for(self._lib.pg_batch_run(self._h, C.c_uint32(max_sweeps), C.byref(t) if t else None,
                                          C.byref(stats), C.byref(fin)), "pg_batch_run")
        return fin.value, stats.as_dict()

    def status(self, sid: int):
        st, sw = C.c_int(), C.c_uint64()
        _raise_for(self._lib.pg_scenario_status(self._h, int(sid), C.byref(st), C.byref(sw)), "pg_scenario_status")
        return st.value, sw.value

    def result(self, sid: int):
        """(time, refs_x, refs_y, boundaries) of a finished scenario."""
        h, w = self._shapes[self._scen[sid]]
        n = h * w
        time = np.empty(n, np.float32)
        rx = np.empty(n, np.uint16)
        ry = np.empty(n, np.uint16)
        bits = np.empty((n + 31) // 32, np.uint32)
        _raise_for(self._lib.pg_scenario_result(self._h, int(sid), _ptr(time), _ptr(rx), _ptr(ry), _ptr(bits)),
                   "pg_scenario_result")
        return time.reshape(h, w), rx, ry, bits

    def time_into(self, sid: int, address: int):
        """Copies a finished scenario's arrival times (W*H f32) to host memory
        at ``address`` (for instance a pinned buffer) and waits for it."""
        _raise_for(self._lib.pg_scenario_result(self._h, int(sid), C.c_void_p(address), None, None, None),
                   "pg_scenario_result")

    def take_finished(self, capacity: int = 4096):
        """Ids of scenarios that finished since the previous call, oldest
        first."""
        ids = (C.c_int * capacity)()
        n = C.c_uint32()
        _raise_for(self._lib.pg_ctx_take_finished(self._h, ids, C.c_uint32(capacity), C.byref(n)),
                   "pg_ctx_take_finished")
        return list(ids[:n.value])

    def release(self, sid: int):
        _raise_for(self._lib.pg_scenario_release(self._h, int(sid)), "pg_scenario_release")
        self._scen.pop(sid, None)

    @staticmethod
    def scenario_bytes(width: int, height: int) -> int:
        return int(_native.load().pg_scenario_bytes(C.c_uint32(width), C.c_uint32(height)))
