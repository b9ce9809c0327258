"""Stage F on the GPU (pg_burn through the C ABI) against the f32 oracle.

Bar (north star): GPU within 1e-5 relative of the host f32 oracle for
speed_max and eccentricity (absolute floors 1e-9 m/s and 1e-7).  Azimuth, as
an angular difference where the fire spreads faster than 1 ft/min (the
reference's own azimuth waiver, crates/firelib/src/firelib.rs:473-490):
1e-5 relative to a full turn, widened where the reference's asin(|y|/rv) is
ill-conditioned (see _az_tolerance; the f32 oracle is itself up to 4e-4 rad
from FP64 there).  Differences come only from the device arithmetic of the
transcendental functions vs the host's: the device library's sincosf and
asinf, and for the positive-base powers the kernel's own exp2(y * log2 x)
(hardware log2 / exp2 with the product in double, about 3e-7 relative).
"""
import numpy as np
import pytest

import oracle
from py_propag import synth

pytestmark = pytest.mark.gpu

REL = 1e-5


def _gpu_burn(t, prefill=0.0, offset=0):
    import torch
    import py_propag as pp

    n = np.asarray(t["fuel"]).size
    dev = torch.device("cuda")

    def put(a, dtype):
        buf = torch.empty(n + offset, dtype=dtype, device=dev)
        buf[offset:] = torch.from_numpy(np.ascontiguousarray(a)).to(dev)
        return buf[offset:]

    fuel = put(np.asarray(t["fuel"], np.uint8), torch.uint8)
    cols = [put(np.asarray(t[k], np.float32), torch.float32) for k in oracle.TERRAIN_FIELDS]
    out = [torch.full((n + offset,), prefill, dtype=torch.float32, device=dev)[offset:]
           for _ in range(3)]
    pp.burn_device(fuel, *cols, *out)
    torch.cuda.synchronize()
    return tuple(o.cpu().numpy() for o in out)


def _angdiff(a, b):
    return np.abs(np.mod(np.asarray(a, np.float64) - b + np.pi, 2 * np.pi) - np.pi)


def _az_tolerance(terrain, a64):
    """Azimuth bar: 1e-5 relative to a full turn (2*pi*1e-5 rad) plus the
    conditioning of the reference's asin(|y|/rv) (firelib.rs:1122): with
    u = |y|/rv = |sin(az - upslope)|, an f32 error of a few ulps in u moves
    asin(u) by ~ 8 eps / sqrt(1 - u^2), ~ sqrt(16 eps) as u -> 1.  The f32
    oracle meets this bar against the FP64 twin on the whole field."""
    eps = 2.0 ** -23
    asp = np.asarray(terrain["aspect"], np.float64)
    up = np.where(asp >= np.float32(np.pi), asp - np.pi, asp + np.pi)
    u = np.abs(np.sin(a64 - up))
    return 2 * np.pi * 1e-5 + 8 * eps / np.sqrt(np.maximum(1 - u * u, 8 * eps))


def _check(gpu, cpu, terrain):
    """speed/ecc: 1e-5 relative; azimuth: see _az_tolerance."""
    s_g, a_g, e_g = gpu
    s_c, a_c, e_c = cpu
    assert np.all(np.abs(s_g - s_c) <= REL * np.abs(s_c) + 1e-9)
    assert np.all(np.abs(e_g - e_c) <= REL * np.abs(e_c) + 1e-7)
    _, a64, _ = oracle.burn_f64(terrain)
    spreading = s_c > 0.00508
    tol = _az_tolerance(terrain, a64)
    assert np.all(_angdiff(a_c, a64)[spreading] <= tol[spreading])   # the bar holds for the oracle
    d = _angdiff(a_g, a_c)
    assert np.all(d[spreading] <= tol[spreading])
    print(f"azimuth: max |GPU - oracle| {d[spreading].max():.2e} rad; cells beyond 1e-5 rad: "
          f"{int((d[spreading] > 1e-5).sum())} of {int(spreading.sum())}")
    return (np.abs(s_g - s_c) / np.maximum(s_c, 1e-30)).max()


def test_burn_rothermel_field_1024(oracle_lib):
    t = synth.rothermel_field(seed=0, side=1024)
    cpu = oracle.burn_f32(t, nthreads=16)
    gpu = _gpu_burn(t)
    worst = _check(gpu, cpu, t)
    print(f"max relative speed difference GPU vs f32 oracle: {worst:.3e}")
    s64, _, _ = oracle.burn_f64(t)
    ok = s64 > 0
    print("max relative speed error vs FP64: GPU %.3e, oracle %.3e" % (
        (np.abs(gpu[0][ok] - s64[ok]) / s64[ok]).max(), (np.abs(cpu[0][ok] - s64[ok]) / s64[ok]).max()))


def test_burn_golden_vectors_and_untouched_cells(oracle_lib, golden):
    t = {k: golden[k] for k in oracle.TERRAIN_FIELDS + ("fuel",)}
    gpu = _gpu_burn(t, prefill=-7.0)
    cpu = oracle.burn_f32(t)
    invalid = (t["fuel"] == 0) | (t["fuel"] > 13)
    for g in gpu:
        assert np.all(g[invalid] == -7.0)  # cuda.rs:89-93 writes only valid fuel
    tv = {k: np.asarray(v)[~invalid] for k, v in t.items()}
    _check(tuple(g[~invalid] for g in gpu), tuple(c[~invalid] for c in cpu), tv)
    # against fireLib.c FP64: the reference quickcheck tolerance
    rs = golden["ref_speed_max"][~invalid]
    s = gpu[0][~invalid]
    assert np.all((np.abs(s - rs) / 0.00508 < 1e-2) | (np.abs(s - rs) / np.maximum(rs, 1e-30) < 2e-2))


@pytest.mark.parametrize("n,offset", [(1023, 0), (4096, 1), (5, 3)])
def test_burn_scalar_paths(oracle_lib, n, offset):
    t = synth.rothermel_field(seed=9, side=64)
    t = {k: (v[:n] if isinstance(v, np.ndarray) else v) for k, v in t.items()}
    t["fuel"] = t["fuel"].copy()
    t["fuel"][::7] = 0
    t["fuel"][::11] = 200
    gpu = _gpu_burn(t, offset=offset)
    cpu = oracle.burn_f32(t)
    _check(gpu, cpu, t)
