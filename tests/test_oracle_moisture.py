"""The moisture pre-pass oracle against the reference's own fixtures.

tests/golden/dmoist/*.txt are the data files of the reference's dmoist tests
(crates/dmoist/fixtures; lines "(args) -> expected", "[month] + range(6) ->
expected", "(month,) -> expected").  The comparisons are the reference's
(crates/dmoist/src/lib.rs:814-1046): exact, except d1hr and the thicker-class
interpolation which compare the result rounded to an integer.  Two fixture
files (dir_viento, mod_viento) belong to wind helpers that are not part of
the moisture model and are not used.  Beyond the fixtures, the hourly driver
is checked for its invariants on whole days of weather.
"""
import math
import os
import re

import numpy as np
import pytest

import oracle

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "dmoist")


def _rows(name):
    """(args, expected) of the fixture lines, tuple or bare-list form."""
    out = []
    with open(os.path.join(GOLDEN, name)) as f:
        for line in f:
            line = line.strip()
            if not line or line.startswith("#") or " -> " not in line:
                continue
            left, right = (x.strip() for x in line.split(" -> ", 1))
            left = left.strip("()")
            args = [float(v) for v in left.split(",") if v.strip()]
            out.append((args, float(right)))
    return out


def _rust_round(x):
    """f32::round: half away from zero."""
    return math.copysign(math.floor(abs(x) + 0.5), x)


@pytest.fixture(scope="module")
def fns(oracle_lib):
    return {n: oracle.moisture_fn(n) for n in (
        "base_moisture", "shading", "ignition_probability", "corrected_moisture",
        "rain_corrected_probability", "fine_moisture", "thicker_moisture", "live_moisture")}


def test_base_moisture_fixtures(fns):
    rows = _rows("hcb.txt")
    assert len(rows) > 7000
    for (t, rh, hour), want in rows:
        assert fns["base_moisture"](t, rh) == want, (t, rh, hour)
    # the reference's own edge cases (lib.rs:1032-1037)
    assert fns["base_moisture"](50.0, 0.0) == 1 and fns["base_moisture"](-5.0, 100.0) == 14


def test_shading_fixtures(fns):
    rows = _rows("sombreado.txt")
    assert len(rows) > 200
    for (cloud, fuel), want in rows:
        assert fns["shading"](cloud, int(fuel)) == want, (cloud, fuel)


def test_corrected_moisture_fixtures(fns):
    rows = _rows("corrige_hcb_por_sombreado.txt")
    assert len(rows) > 50
    for (hcb, cloud, month, aspect, slope, fuel, hour), want in rows:
        got = fns["corrected_moisture"](hcb, cloud, int(month), aspect, slope, int(fuel), int(hour))
        assert got == want, (hcb, cloud, month, aspect, slope, fuel, hour, got)


def test_ignition_probability_fixtures(fns):
    rows = _rows("probabilidad_ignicion.txt")
    assert len(rows) >= 100
    for (t, cloud, hcs, fuel), want in rows:
        assert fns["ignition_probability"](t, cloud, hcs, int(fuel)) == want, (t, cloud, hcs, fuel)


def test_rain_corrected_probability_fixtures(fns):
    rows = _rows("corrige_prob_por_pp.txt")
    assert len(rows) > 40
    for (p, e), want in rows:
        assert fns["rain_corrected_probability"](p, e) == want, (p, e)


def test_fine_moisture_fixtures(fns):
    rows = _rows("d1hr.txt")
    assert len(rows) > 500
    for (hcs, p_dry, p_wet), want in rows:
        assert _rust_round(fns["fine_moisture"](hcs, p_dry, p_wet)) == want, (hcs, p_dry, p_wet)


def test_thicker_moisture_fixtures(fns):
    rows = _rows("hco_x10hr.txt")
    assert len(rows) > 400
    for (hour, m, m6, m15), want in rows:
        assert _rust_round(fns["thicker_moisture"](int(hour), m, m6, m15)) == want, (hour, m, m6, m15)
    assert fns["thicker_moisture"](12, 0.0, 10.0, 20.0) == 0.0


def test_rain_effect_and_live_moisture_fixtures(fns):
    path = os.path.join(GOLDEN, "efecto_precipitacion_maximo.txt")
    pat = re.compile(r"\[(\d+)\] \+ range\(6\) -> (\S+)")
    n = 0
    with open(path) as f:
        for line in f:
            m = pat.match(line.strip())
            if m:
                n += 1
                got = oracle.rain_effect_max(int(m.group(1)), [0.0, 1.0, 2.0, 3.0, 4.0, 5.0])
                assert got == float(m.group(2)), line
    assert n == 12
    for (month,), want in _rows("humedad_vivo.txt"):
        assert fns["live_moisture"](int(month)) == want
    assert math.isnan(fns["live_moisture"](13))


def test_hourly_driver_invariants(oracle_lib):
    """Whole days on random terrain: outputs are ratios of the per-value
    functions, non-combustible cells are dry, live moisture is the month's,
    and 10-h / 100-h equal the interpolation of the finer class."""
    r = np.random.Generator(np.random.PCG64(5))
    n = 2000
    fuel = r.integers(0, 14, n).astype(np.uint8)
    slope = r.uniform(0, 1.2, n).astype(np.float32)
    slope[:50] = 0.0
    aspect = np.radians(r.uniform(0, 359.9, n)).astype(np.float32)
    for month in (1, 4, 7, 10):
        w = oracle.day_weather(r.uniform(-5, 40, 24), r.uniform(5, 100, 24), r.uniform(0, 100, 24),
                               r.choice([0.0, 0.5, 3.0, 12.0, 40.0], 6), month)
        for hour in (0, 6, 11, 15, 22):
            d1, d10, d100, herb, wood = oracle.moisture(fuel, slope, aspect, w, hour)
            assert np.all(d1[fuel == 0] == 0) and np.all(d10[fuel == 0] == 0)
            assert np.all(herb == np.float32(oracle.moisture_fn("live_moisture")(month)) / np.float32(100))
            assert np.array_equal(herb, wood)
            assert np.all(d1 >= 0)
            d1_6 = oracle.moisture(fuel, slope, aspect, w, 6)[0]
            d1_15 = oracle.moisture(fuel, slope, aspect, w, 15)[0]
            thick = oracle.moisture_fn("thicker_moisture")
            for i in range(0, n, 97):
                want = np.float32(thick(hour, float(np.float32(d1[i] * 100)), float(np.float32(d1_6[i] * 100)),
                                        float(np.float32(d1_15[i] * 100)))) / np.float32(100)
                assert abs(float(d10[i]) - float(want)) <= 1e-6, (month, hour, i)
