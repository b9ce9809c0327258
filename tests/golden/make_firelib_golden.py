"""Writes tests/golden/firelib_fp64.npz: FP64 golden vectors from the
reference's own C fireLib (crates/firelib-sys/fireLib.c), compiled from
/root/reference into oracle/_ref/ by oracle/Makefile.

The catalog is built through fireLib's public API with the particle table of
crates/firelib/src/lib.rs:125-268 (the one the reference's GPU path uses;
NFFL13 10-h load 0.0580).  Inputs: the i.i.d. per-cell Rothermel field
distribution (fuel U{0..13} here, so code 0 is covered), plus hand-picked
special cases (no wind, no slope, wind exactly upslope, moisture at and above
extinction, zero wind and slope).

Run from the repo root after `make -C oracle`:
    python tests/golden/make_firelib_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402


def inputs(seed=20261203, n=4096):
    r = np.random.Generator(np.random.PCG64(seed))
    d1 = r.uniform(0.03, 0.12, n).astype(np.float32)
    t = {
        "fuel": r.integers(0, 14, n, dtype=np.uint8),
        "d1hr": d1,
        "d10hr": (d1 + np.float32(0.01)).astype(np.float32),
        "d100hr": (d1 + np.float32(0.02)).astype(np.float32),
        "herb": r.uniform(0.3, 1.5, n).astype(np.float32),
        "wood": r.uniform(0.6, 1.5, n).astype(np.float32),
        "wind_speed": r.uniform(0.0, 10.0, n).astype(np.float32),
        "wind_azimuth": np.radians(r.uniform(0.0, 360.0, n)).astype(np.float32),
        "slope": r.uniform(0.0, 0.6, n).astype(np.float32),
        "aspect": np.radians(r.uniform(0.0, 360.0, n)).astype(np.float32),
    }
    # special cases: every fuel 1..13 under each situation
    rows = []
    for code in range(1, 14):
        base = dict(fuel=code, d1hr=0.06, d10hr=0.07, d100hr=0.08, herb=0.9, wood=1.0)
        rows.append({**base, "wind_speed": 0.0, "wind_azimuth": 0.0, "slope": 0.0, "aspect": 0.0})
        rows.append({**base, "wind_speed": 4.0, "wind_azimuth": 1.0, "slope": 0.0, "aspect": 0.0})
        rows.append({**base, "wind_speed": 0.0, "wind_azimuth": 0.0, "slope": 0.4, "aspect": 2.0})
        # wind blowing exactly upslope: aspect pi/2 -> upslope 3pi/2
        up = np.float32(np.float32(np.pi / 2) + np.float32(np.pi))
        rows.append({**base, "wind_speed": 3.0, "wind_azimuth": float(up), "slope": 0.3,
                     "aspect": float(np.float32(np.pi / 2))})
        rows.append({**base, "wind_speed": 6.0, "wind_azimuth": 5.5, "slope": 0.5, "aspect": 0.7})
        wet = dict(base, d1hr=0.6, d10hr=0.6, d100hr=0.6, herb=3.0, wood=3.0)
        rows.append({**wet, "wind_speed": 5.0, "wind_azimuth": 1.0, "slope": 0.2, "aspect": 1.0})
    for k in t:
        extra = np.array([row[k] for row in rows], dtype=t[k].dtype)
        t[k] = np.concatenate([t[k], extra])
    return t


def main():
    oracle.build(ref=True)
    t = inputs()
    out = oracle.ref_firelib(t)
    if out is None:
        raise SystemExit("oracle/_ref/libfirelib_ref.so is not built (reference checkout missing)")
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "firelib_fp64.npz")
    np.savez_compressed(path, **t, ref_speed_max=out["speed_max"], ref_azimuth_max=out["azimuth_max"],
                        ref_eccentricity=out["eccentricity"], ref_spread0_ftmin=out["spread0_ftmin"],
                        ref_rx_int=out["rx_int"])
    print("wrote", path, "rows", t["fuel"].size)


if __name__ == "__main__":
    main()
