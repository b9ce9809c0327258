"""Stage F on the GPU (pg_burn through the C ABI) against the f32 oracle.

Bar (north star): GPU within 1e-5 relative of the host f32 oracle for
speed_max and eccentricity (absolute floors 1e-9 m/s and 1e-7), azimuth as an
angular difference within 1e-5 rad where the fire spreads faster than 1 ft/min
(the reference's own azimuth waiver, crates/firelib/src/firelib.rs:473-490).
The error of both against the FP64 twin is stated as well.  Differences come
only from the device libm (powf, sinf, cosf, asinf) vs the host's.
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


def _check(gpu, cpu):
    s_g, a_g, e_g = gpu
    s_c, a_c, e_c = cpu
    assert np.all(np.abs(s_g - s_c) <= REL * np.abs(s_c) + 1e-9)
    assert np.all(np.abs(e_g - e_c) <= REL * np.abs(e_c) + 1e-7)
    spreading = s_c > 0.00508
    d = np.abs(np.mod(a_g.astype(np.float64) - a_c + np.pi, 2 * np.pi) - np.pi)
    assert np.all(d[spreading] <= 1e-5)
    return (np.abs(s_g - s_c) / np.maximum(s_c, 1e-30)).max()


def test_burn_rothermel_field_1024(oracle_lib):
    t = synth.rothermel_field(seed=0, side=1024)
    cpu = oracle.burn_f32(t, nthreads=16)
    gpu = _gpu_burn(t)
    worst = _check(gpu, cpu)
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
    _check(tuple(g[~invalid] for g in gpu), tuple(c[~invalid] for c in cpu))
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
    _check(gpu, cpu)
