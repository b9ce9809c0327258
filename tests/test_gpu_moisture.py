"""pg_moisture (the moisture pre-pass kernel) against the oracle.

Bar: bit-exact moisture ratios for every cell, except where the slope in
degrees (f32 atan of rise/run x 180/pi) lies within 1e-4 degree of the 30
degree class boundary, where the device's and the host's atanf may round
differently by one ulp (a handful of the 262,144 cells per draw; they are
excluded from the comparison and their number bounded).  Covers every month and the hours that switch the shading rules
(night, 06-08, day), flat cells, aspect at 0 and near 360 degrees, fuel 0
and fuels 7-12 (own shading rule), dry and rainy weeks.
"""
import math

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _cells(seed, n):
    r = np.random.Generator(np.random.PCG64(seed))
    fuel = r.integers(0, 14, n).astype(np.uint8)
    slope = r.uniform(0.0, 1.5, n).astype(np.float32)
    slope[: n // 50] = 0.0
    aspect = np.radians(r.uniform(0.0, 360.0, n)).astype(np.float32)
    aspect[n // 50: n // 25] = 0.0
    aspect[n // 25: n // 20] = np.float32(np.nextafter(np.float32(2 * math.pi), np.float32(0)))
    return fuel, slope, aspect


def _near_class_edge(slope):
    deg = np.degrees(np.arctan(slope.astype(np.float64)))
    return np.abs(deg - 30.0) < 1e-4


@pytest.mark.parametrize("month", list(range(1, 13)))
def test_moisture_matches_oracle(month, oracle_lib):
    import torch
    import py_propag as pp

    n = 1 << 18
    fuel, slope, aspect = _cells(100 + month, n)
    r = np.random.Generator(np.random.PCG64(month))
    rain = [0.0, 0.4, 3.0, 8.0, 20.0, 50.0] if month % 2 else [0.0] * 6
    wx = dict(temperature=r.uniform(-8, 45, 24), humidity=r.uniform(0, 100, 24), cloud_cover=r.uniform(0, 100, 24),
              rain_6_days=rain, month=month)
    dev = torch.device("cuda:0")
    t_fuel = torch.from_numpy(fuel).to(dev)
    t_slope = torch.from_numpy(slope).to(dev)
    t_aspect = torch.from_numpy(aspect).to(dev)
    outs = [torch.empty(n, dtype=torch.float32, device=dev) for _ in range(5)]
    w = pp.day_weather(**wx)
    ow = oracle.day_weather(wx["temperature"], wx["humidity"], wx["cloud_cover"], rain, month)
    edge = _near_class_edge(slope)
    for hour in (0, 5, 6, 7, 8, 12, 15, 19, 20, 23):
        pp.moisture_device(t_fuel, t_slope, t_aspect, w, hour, *outs)
        torch.cuda.synchronize()
        ref = oracle.moisture(fuel, slope, aspect, ow, hour)
        for name, got, want in zip(("d1hr", "d10hr", "d100hr", "herb", "wood"), outs, ref):
            g = got.cpu().numpy()
            same = (g.view(np.uint32) == want.view(np.uint32)) | (np.isnan(g) & np.isnan(want))
            bad = ~same & ~edge
            assert not bad.any(), (month, hour, name, int(bad.sum()), np.flatnonzero(bad)[:5])
    assert edge.sum() <= n // 10000  # the excluded band stays tiny


def test_moisture_rejects_bad_arguments():
    import torch
    import py_propag as pp

    with pytest.raises(pp.PropagValidationError):
        pp.day_weather(20.0, 50.0, 10.0, [0.0] * 6, 13)
    w = pp.day_weather(20.0, 50.0, 10.0, [0.0] * 6, 7)
    t = torch.zeros(16, device="cuda:0")
    f = torch.zeros(16, dtype=torch.uint8, device="cuda:0")
    with pytest.raises(pp.PropagValidationError):
        pp.moisture_device(f, t, t, w, 24, t, t, t, t, t)


def test_moisture_feeds_the_burn_pass(oracle_lib):
    """The moisture rasters of pg_moisture feed pg_burn directly on the device
    (the per-cell form of create_terrain_with_fuel_moisture,
    crates/firelib/src/fuel_moisture.rs:26-63): the GPU chain matches the
    oracle's chain (oracle moisture -> f32 oracle burn) to the Stage F bar of
    tests/test_gpu_burn.py.  Slopes in the atanf edge band are left out so
    that both chains burn identical moisture."""
    import torch
    import py_propag as pp
    from test_gpu_burn import _check

    n = 1 << 16
    fuel, slope, aspect = _cells(77, n)
    keep = ~_near_class_edge(slope)
    fuel, slope, aspect = fuel[keep], slope[keep], aspect[keep]
    n = fuel.size
    r = np.random.Generator(np.random.PCG64(78))
    ws = r.uniform(0.0, 10.0, n).astype(np.float32)
    waz = np.radians(r.uniform(0.0, 360.0, n)).astype(np.float32)
    wx = dict(temperature=r.uniform(5, 38, 24), humidity=r.uniform(10, 90, 24), cloud_cover=r.uniform(0, 80, 24),
              rain_6_days=[0.0, 0.0, 1.5, 0.0, 6.0, 0.0], month=7)
    hour = 14
    dev = torch.device("cuda:0")
    put = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    t_fuel, t_slope, t_aspect = put(fuel), put(slope), put(aspect)
    moist = [torch.empty(n, dtype=torch.float32, device=dev) for _ in range(5)]
    pp.moisture_device(t_fuel, t_slope, t_aspect, pp.day_weather(**wx), hour, *moist)
    out = [torch.zeros(n, dtype=torch.float32, device=dev) for _ in range(3)]
    pp.burn_device(t_fuel, *moist, put(ws), put(waz), t_slope, t_aspect, *out)
    torch.cuda.synchronize()
    gpu = tuple(o.cpu().numpy() for o in out)

    ow = oracle.day_weather(wx["temperature"], wx["humidity"], wx["cloud_cover"], wx["rain_6_days"], 7)
    d1, d10, d100, herb, wood = oracle.moisture(fuel, slope, aspect, ow, hour)
    for got, want in zip(moist, (d1, d10, d100, herb, wood)):
        assert np.array_equal(got.cpu().numpy().view(np.uint32), want.view(np.uint32))
    terrain = {"fuel": fuel, "d1hr": d1, "d10hr": d10, "d100hr": d100, "herb": herb, "wood": wood,
               "wind_speed": ws, "wind_azimuth": waz, "slope": slope, "aspect": aspect}
    cpu = oracle.burn_f32(terrain, nthreads=16)
    assert (cpu[0] > 0).sum() > n // 2          # most cells spread
    _check(gpu, cpu, terrain)
