"""Pins the Stage F oracle (oracle/firelib_oracle.c).

* FP64 twin vs the reference's own C fireLib (golden vectors written by
  tests/golden/make_firelib_golden.py from crates/firelib-sys/fireLib.c with
  the Rust particle table): the equations, catalog derivation and branch
  logic must agree to double rounding.
* f32 path vs FP64: the error the f32 transcription carries, stated here and
  bounded well inside the reference's own 2e-2 quickcheck tolerance
  (crates/firelib/src/f32.rs:5-6).
* Closed forms of firelib.rs:1100-1107 (no wind no slope, wind only, slope
  only, moisture above extinction).
"""
import numpy as np
import pytest

import oracle

FIELDS = oracle.TERRAIN_FIELDS


def _terrain(z):
    return {k: z[k] for k in FIELDS + ("fuel",)}


def _angdiff(a, b):
    return np.abs(np.mod(a - b + np.pi, 2 * np.pi) - np.pi)


def test_f64_matches_firelib_c(oracle_lib, golden):
    t = _terrain(golden)
    s, a, e = oracle.burn_f64(t)
    valid = (t["fuel"] >= 1) & (t["fuel"] <= 13)
    rs, ra, re = golden["ref_speed_max"], golden["ref_azimuth_max"], golden["ref_eccentricity"]
    rel = np.abs(s[valid] - rs[valid]) / np.maximum(rs[valid], 1e-300)
    assert rel.max() < 1e-12
    assert np.abs(e[valid] - re[valid]).max() < 1e-12
    # fireLib returns azimuth 0 when there is no spread; the reference only
    # compares azimuths above 1 ft/min (firelib.rs:473-490).
    spreading = valid & (rs > 0.00508)
    assert _angdiff(a[spreading], ra[spreading]).max() < 1e-6
    # non-catalog code 0 writes nothing
    assert np.all(s[~valid] == 0)


def test_f32_error_vs_fp64(oracle_lib, golden):
    t = _terrain(golden)
    s32, a32, e32 = oracle.burn_f32(t)
    s64, a64, e64 = oracle.burn_f64(t)
    valid = (t["fuel"] >= 1) & (t["fuel"] <= 13) & (s64 > 0)
    rel = np.abs(s32[valid] - s64[valid]) / s64[valid]
    # Stated f32 error of the transcription: <= 5e-5 relative speed,
    # <= 1e-5 absolute eccentricity, <= 1e-3 rad azimuth where spreading.
    assert rel.max() < 5e-5
    assert np.abs(e32[valid] - e64[valid]).max() < 1e-5
    spreading = valid & (s64 > 0.00508)
    assert _angdiff(a32[spreading], a64[spreading]).max() < 1e-3


def test_f32_matches_firelib_c_within_reference_tolerance(oracle_lib, golden):
    t = _terrain(golden)
    s, _, e = oracle.burn_f32(t)
    valid = (t["fuel"] >= 1) & (t["fuel"] <= 13)
    rs, re = golden["ref_speed_max"], golden["ref_eccentricity"]
    # reference quickcheck: diff < 1e-2 (ft/min units) or diff/min < 2e-2
    ft = s[valid] / 0.00508
    rft = rs[valid] / 0.00508
    diff = np.abs(ft - rft)
    assert np.all((diff < 1e-2) | (diff / np.maximum(np.minimum(ft, rft), 1e-30) < 2e-2))
    assert np.all(np.abs(e[valid] - re[valid]) < 2e-2)


def _uniform(code, n=1, **kw):
    base = dict(d1hr=0.06, d10hr=0.07, d100hr=0.08, herb=0.9, wood=1.0, wind_speed=0.0,
                wind_azimuth=0.0, slope=0.0, aspect=0.0)
    base.update(kw)
    t = {k: np.full(n, v, np.float32) for k, v in base.items()}
    t["fuel"] = np.full(n, code, np.uint8)
    return t


@pytest.mark.parametrize("code", range(1, 14))
def test_closed_forms(oracle_lib, code):
    # no wind, no slope: speed = speed0, az = 0, ecc = 0 (NoSlopeNoWind)
    s0, a0, e0 = oracle.burn_f32(_uniform(code))
    assert s0[0] > 0 and a0[0] == 0 and e0[0] == 0
    # wind only: azimuth = wind azimuth, faster than speed0 (WindNoSlope)
    s1, a1, e1 = oracle.burn_f32(_uniform(code, wind_speed=3.0, wind_azimuth=1.25))
    assert a1[0] == np.float32(1.25) and s1[0] > s0[0] and 0 < e1[0] < 1
    # slope only: azimuth = upslope = aspect + pi (SlopeNoWind)
    s2, a2, _ = oracle.burn_f32(_uniform(code, slope=0.4, aspect=0.5))
    assert a2[0] == np.float32(np.float32(0.5) + np.float32(np.pi)) and s2[0] > s0[0]
    # dead moisture at extinction: no dead reaction; fuels without live
    # particles stop spreading (NoSpread)
    wet = _uniform(code, d1hr=0.9, d10hr=0.9, d100hr=0.9, herb=5.0, wood=5.0, wind_speed=4.0,
                   wind_azimuth=1.0)
    s3, a3, e3 = oracle.burn_f32(wet)
    assert s3[0] == 0 and a3[0] == 0 and e3[0] == 0


def test_non_catalog_codes_write_nothing(oracle_lib):
    t = _uniform(0, n=4)
    t["fuel"] = np.array([0, 14, 19, 255], np.uint8)
    s, a, e = oracle.burn_f32(t)
    assert not s.any() and not a.any() and not e.any()


def test_catalog_f32_close_to_f64(oracle_lib):
    for code in range(1, 14):
        n32, c32 = oracle.catalog_f32(code)
        n64, c64 = oracle.catalog_f64(code)
        assert n32 == n64 > 0
        rel = np.abs(c32.astype(np.float64) - c64) / np.maximum(np.abs(c64), 1e-30)
        assert rel.max() < 2e-5, (code, rel)


def test_rothermel_field_parallel_equals_serial(oracle_lib):
    from py_propag import synth

    t = synth.rothermel_field(seed=3, side=64)
    a = oracle.burn_f32(t, nthreads=1)
    b = oracle.burn_f32(t, nthreads=4)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
