"""Seeded synthetic inputs for the benchmark workloads.

Every workload is a pure function of its seed (numpy PCG64), so the GPU path
and the CPU checker see the same rasters.  Angles follow TerrainCuda
(crates/firelib/src/firelib.rs:95-124): radians, compass convention, wind
azimuth = direction the wind blows toward, aspect = downslope direction.
The rasters are south-up (row index grows with geographic y) with square
cells, like the reference CLI scenario (crates/propag/src/main.rs:14-25).

Workloads:
  rothermel_field   1024x1024 i.i.d. per-cell firelib inputs
  landscape         DEM of Gaussian hills, Voronoi fuel patches, smooth wind
  single_scenario   1024x1024, 20 hills, 200 patches, ignition (500,500), 10 h
  large_scenario    4096x4096, 80 hills, 3000 patches, 16 points + 200-cell line, 12 h
  ensemble          scenarios of uneven wind / moisture / ignitions / max_time
                    over one shared landscape
  serving_stream    Poisson arrivals of such scenarios over several resident
                    landscapes of mixed raster sizes
"""
from __future__ import annotations

import math

import numpy as np

FLT_MAX = np.float32(np.finfo(np.float32).max)
NO_REF = 0xFFFF
TWO_PI = 2.0 * math.pi


def _rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def rothermel_field(seed: int = 0, side: int = 1024) -> dict:
    """i.i.d. per-cell inputs: fuel U{1..13}; d1 U(0.03,0.12), d10 = d1+0.01,
    d100 = d1+0.02; herb U(0.3,1.5); wood U(0.6,1.5); wind U(0,10) m/s;
    wind azimuth U(0,360) deg; slope U(0,0.6); aspect U(0,360) deg."""
    r = _rng(seed)
    n = side * side
    d1 = r.uniform(0.03, 0.12, n).astype(np.float32)
    return {
        "fuel": r.integers(1, 14, n, dtype=np.uint8),
        "d1hr": d1,
        "d10hr": (d1 + np.float32(0.01)).astype(np.float32),
        "d100hr": (d1 + np.float32(0.02)).astype(np.float32),
        "herb": r.uniform(0.3, 1.5, n).astype(np.float32),
        "wood": r.uniform(0.6, 1.5, n).astype(np.float32),
        "wind_speed": r.uniform(0.0, 10.0, n).astype(np.float32),
        "wind_azimuth": np.radians(r.uniform(0.0, 360.0, n)).astype(np.float32),
        "slope": r.uniform(0.0, 0.6, n).astype(np.float32),
        "aspect": np.radians(r.uniform(0.0, 360.0, n)).astype(np.float32),
        "shape": (side, side),
    }


def _dem(r, side, cell, hills):
    """Sum of Gaussian hills, height U(50,400) m, sigma U(200,1500) m, built
    separably (each hill is an outer product of two 1-D Gaussians)."""
    coords = (np.arange(side, dtype=np.float64) + 0.5) * cell
    z = np.zeros((side, side), np.float64)
    ext = side * cell
    for _ in range(hills):
        h = r.uniform(50.0, 400.0)
        s = r.uniform(200.0, 1500.0)
        cx, cy = r.uniform(0.0, ext, 2)
        gx = np.exp(-((coords - cx) ** 2) / (2 * s * s))
        gy = np.exp(-((coords - cy) ** 2) / (2 * s * s))
        z += h * np.outer(gy, gx)
    return z


def _slope_aspect(z, cell):
    """slope = |grad z| (central differences), aspect = compass azimuth of
    -grad z (downslope) in [0, 2pi).  Rows grow with y (south-up)."""
    dzdy, dzdx = np.gradient(z, cell)
    slope = np.hypot(dzdx, dzdy)
    aspect = np.mod(np.arctan2(-dzdx, -dzdy), TWO_PI)
    return slope.astype(np.float32), aspect.astype(np.float32)


def _voronoi_fuel(r, side, patches):
    from scipy.spatial import cKDTree

    seeds = r.uniform(0.0, side, (patches, 2))
    codes = r.integers(1, 14, patches, dtype=np.uint8)
    tree = cKDTree(seeds)
    out = np.empty(side * side, np.uint8)
    xs = np.arange(side, dtype=np.float64) + 0.5
    rows = max(1, (1 << 22) // side)
    for y0 in range(0, side, rows):
        y1 = min(side, y0 + rows)
        yy, xx = np.meshgrid(np.arange(y0, y1, dtype=np.float64) + 0.5, xs, indexing="ij")
        _, idx = tree.query(np.column_stack([xx.ravel(), yy.ravel()]), workers=-1)
        out[y0 * side:y1 * side] = codes[idx]
    return out.reshape(side, side)


def _smooth_field(r, side, cell, k=4):
    """Normalised sum of k plane sines in [-1, 1] (wavelengths 1-4 km)."""
    coords = (np.arange(side, dtype=np.float64) + 0.5) * cell
    yy, xx = np.meshgrid(coords, coords, indexing="ij")
    acc = np.zeros((side, side), np.float64)
    for _ in range(k):
        th = r.uniform(0.0, TWO_PI)
        lam = r.uniform(1000.0, 4000.0)
        ph = r.uniform(0.0, TWO_PI)
        acc += np.sin(TWO_PI * (xx * math.cos(th) + yy * math.sin(th)) / lam + ph)
    return acc / k


def _moisture(r):
    d1 = np.float32(r.uniform(0.03, 0.12))
    return {"d1hr": d1, "d10hr": np.float32(d1 + np.float32(0.01)),
            "d100hr": np.float32(d1 + np.float32(0.02)),
            "herb": np.float32(r.uniform(0.3, 1.5)), "wood": np.float32(r.uniform(0.6, 1.5))}


def landscape(seed: int, side: int, hills: int, patches: int, cell: float = 5.0) -> dict:
    """Terrain shared by scenarios: fuel, slope, aspect (per cell)."""
    r = _rng(seed)
    z = _dem(r, side, cell, hills)
    slope, aspect = _slope_aspect(z, cell)
    fuel = _voronoi_fuel(r, side, patches)
    return {"fuel": fuel.ravel(), "slope": slope.ravel(), "aspect": aspect.ravel(),
            "shape": (side, side), "cell": cell, "seed": seed}


def initial_state(side_w: int, side_h: int, ignitions):
    """time / refs rasters as propag.rs:433-449 initialises them: ignited
    cells hold their time and reference themselves, the rest FLT_MAX and
    0xFFFF.  ``ignitions`` is an iterable of (x, y, t)."""
    n = side_w * side_h
    time = np.full(n, FLT_MAX, np.float32)
    rx = np.full(n, NO_REF, np.uint16)
    ry = np.full(n, NO_REF, np.uint16)
    for x, y, t in ignitions:
        i = int(x) + int(y) * side_w
        time[i] = np.float32(t)
        rx[i] = x
        ry[i] = y
    return time, rx, ry


def line_cells(x0, y0, x1, y1):
    """Cells of a straight segment (Bresenham), endpoints included."""
    cells = []
    dx, dy = abs(x1 - x0), -abs(y1 - y0)
    sx, sy = (1 if x0 < x1 else -1), (1 if y0 < y1 else -1)
    err = dx + dy
    x, y = x0, y0
    while True:
        cells.append((x, y))
        if x == x1 and y == y1:
            break
        e2 = 2 * err
        if e2 >= dy:
            err += dy
            x += sx
        if e2 <= dx:
            err += dx
            y += sy
    return cells


def scenario(seed: int, side: int, hills: int, patches: int, ignitions, max_time: float,
             cell: float = 5.0, land: dict | None = None) -> dict:
    """A full scenario: landscape + constant moisture + smooth wind field
    (6 m/s +- 2, toward 225 deg +- 20) + ignitions."""
    if land is None:
        land = landscape(seed, side, hills, patches, cell)
    r = _rng(seed + 7919)
    m = _moisture(r)
    s1 = _smooth_field(r, side, cell)
    s2 = _smooth_field(r, side, cell)
    n = side * side
    terrain = {
        "fuel": land["fuel"], "slope": land["slope"], "aspect": land["aspect"],
        "wind_speed": (6.0 + 2.0 * s1).astype(np.float32).ravel(),
        "wind_azimuth": np.radians(225.0 + 20.0 * s2).astype(np.float32).ravel(),
    }
    for k, v in m.items():
        terrain[k] = np.full(n, v, np.float32)
    time, rx, ry = initial_state(side, side, ignitions)
    return {"terrain": terrain, "width": side, "height": side, "dx": cell, "dy": cell,
            "max_time": np.float32(max_time), "time": time, "refs_x": rx, "refs_y": ry,
            "ignitions": list(ignitions)}


def single_scenario(seed: int = 1, side: int = 1024) -> dict:
    """1024x1024 at 5 m, 20 hills, 200 Voronoi patches, point ignition at
    (500, 500), max_time 36000 s.  Smaller ``side`` scales the landscape
    down for parity tests (ignition at the centre)."""
    hills = max(1, round(20 * (side / 1024) ** 2)) if side < 1024 else 20
    patches = max(4, round(200 * (side / 1024) ** 2)) if side < 1024 else 200
    ign = (500, 500) if side == 1024 else (side // 2 - 12 * side // 1024, side // 2 - 12 * side // 1024)
    return scenario(seed, side, hills, patches, [(ign[0], ign[1], 0.0)], 36000.0)


def large_scenario(seed: int = 2, side: int = 4096) -> dict:
    """4096x4096 at 5 m, 80 hills, 3000 patches, 16 interior point ignitions
    plus one 200-cell straight line ignition, max_time 43200 s."""
    r = _rng(seed + 104729)
    lo, hi = side // 8, side - side // 8
    pts = [(int(x), int(y), 0.0) for x, y in r.integers(lo, hi, (16, 2))]
    x0, y0 = (int(v) for v in r.integers(lo, hi, 2))
    ang = r.uniform(0.0, TWO_PI)
    length = 199
    x1 = int(np.clip(round(x0 + length * math.cos(ang)), 0, side - 1))
    y1 = int(np.clip(round(y0 + length * math.sin(ang)), 0, side - 1))
    line = [(x, y, 0.0) for x, y in line_cells(x0, y0, x1, y1)][:200]
    seen = {}
    for x, y, t in pts + line:
        seen[(x, y)] = t
    ign = [(x, y, t) for (x, y), t in seen.items()]
    hills = 80 if side == 4096 else max(1, round(80 * (side / 4096) ** 2))
    patches = 3000 if side == 4096 else max(4, round(3000 * (side / 4096) ** 2))
    return scenario(seed, side, hills, patches, ign, 43200.0)


def uniform_scenario(side=1024, fuel=1, moisture=0.1, wind=5.0, wind_az=0.0, slope=0.0,
                     aspect=0.0, ignition=None, max_time=36000.0, cell=5.0) -> dict:
    """Spatially constant terrain, as the reference CLI (main.rs:36-56):
    fuel 1, all moistures 0.1, wind 5 m/s toward azimuth 0, flat, ignition at
    (side/2 - 12, side/2 - 12), 10 h."""
    n = side * side
    if ignition is None:
        ignition = (side // 2 - 12, side // 2 - 12)
    terrain = {
        "fuel": np.full(n, fuel, np.uint8),
        "d1hr": np.full(n, moisture, np.float32), "d10hr": np.full(n, moisture, np.float32),
        "d100hr": np.full(n, moisture, np.float32), "herb": np.full(n, moisture, np.float32),
        "wood": np.full(n, moisture, np.float32),
        "wind_speed": np.full(n, wind, np.float32), "wind_azimuth": np.full(n, wind_az, np.float32),
        "slope": np.full(n, slope, np.float32), "aspect": np.full(n, aspect, np.float32),
    }
    time, rx, ry = initial_state(side, side, [(ignition[0], ignition[1], 0.0)])
    return {"terrain": terrain, "width": side, "height": side, "dx": cell, "dy": cell,
            "max_time": np.float32(max_time), "time": time, "refs_x": rx, "refs_y": ry,
            "ignitions": [(ignition[0], ignition[1], 0.0)]}


def scenario_weather(r, ign_count, w, h, max_time, margin=8):
    """Weather and ignitions of one ensemble / serving scenario: 1-h moisture
    U(0.03,0.12), 10-h/100-h +0.01/+0.02, herb U(0.3,1.5), wood U(0.6,1.5),
    uniform wind U(2,12) m/s toward U(0,360) deg, ``ign_count`` distinct
    point ignitions at t = 0 inside the central (1 - 2/margin) of the raster."""
    d1 = np.float32(r.uniform(0.03, 0.12))
    wx = {"d1hr": d1, "d10hr": np.float32(d1 + np.float32(0.01)),
          "d100hr": np.float32(d1 + np.float32(0.02)), "herb": np.float32(r.uniform(0.3, 1.5)),
          "wood": np.float32(r.uniform(0.6, 1.5)), "wind_speed": np.float32(r.uniform(2.0, 12.0)),
          "wind_azimuth": np.float32(np.radians(r.uniform(0.0, 360.0)))}
    cells = set()
    while len(cells) < ign_count:
        cells.add((int(r.integers(w // margin, w - w // margin)), int(r.integers(h // margin, h - h // margin))))
    return {"wx": wx, "ign": [(x, y, 0.0) for x, y in sorted(cells)], "max_time": float(max_time)}


def serving_stream(seed: int = 4, rate: float = 50.0, seconds: float = 60.0, terrains: int = 4,
                   side_lo: int = 512, side_hi: int = 4096) -> dict:
    """Serving workload: ``terrains`` landscapes with sides log-uniform in
    [side_lo, side_hi] (config-1 generator, hills scaled with the side and
    fuel patches with the area as between the 1024 and 4096 landscapes:
    20 hills and 200 patches at 1024), and Poisson arrivals at ``rate`` per second
    for ``seconds``; each request picks a landscape uniformly, draws its
    weather and 1-4 ignitions as ``scenario_weather`` and max_time
    U(3600, 36000) s.  Landscapes are generated lazily by ``land(i)``."""
    r = _rng(seed)
    sides = [int(round(math.exp(r.uniform(math.log(side_lo), math.log(side_hi))))) for _ in range(terrains)]
    t, reqs = 0.0, []
    while True:
        t += r.exponential(1.0 / rate)
        if t >= seconds:
            break
        k = int(r.integers(0, terrains))
        case = scenario_weather(r, int(r.integers(1, 5)), sides[k], sides[k], r.uniform(3600.0, 36000.0))
        case.update({"arrival": t, "terrain": k})
        reqs.append(case)

    def land(i):
        a = sides[i] / 1024.0
        return landscape(seed * 1000 + i, sides[i], max(1, round(20 * a)), max(4, round(200 * a * a)))

    return {"sides": sides, "requests": reqs, "land": land, "rate": rate, "seconds": seconds}


def ensemble_2048(seed: int = 3, count: int = 256, side: int = 2048) -> dict:
    """Batched-ensemble workload: ``count`` scenarios sharing one
    ``side``-square landscape (config-1 generator scaled as in
    serving_stream), each with its own weather and 1-4 ignitions as
    ``scenario_weather`` and max_time log-uniform in [3600, 43200] s."""
    r = _rng(seed)
    cases = []
    for _ in range(count):
        mt = math.exp(r.uniform(math.log(3600.0), math.log(43200.0)))
        cases.append(scenario_weather(r, int(r.integers(1, 5)), side, side, mt))
    a = side / 1024.0
    land = landscape(seed * 1000, side, max(1, round(20 * a)), max(4, round(200 * a * a)))
    return {"side": side, "land": land, "cases": cases}
