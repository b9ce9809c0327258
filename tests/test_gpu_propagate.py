"""Stage P on the GPU (pg_propagate through the C ABI) against the oracle.

Both sides get the same fire rasters (the f32 oracle's burn output), so the
comparison isolates the propagation.  The bar is bit-exact arrival times,
references and boundary bits: the kernel runs the oracle's exact Jacobi
schedule with the same IEEE arithmetic.  That is stricter than the north
star's tolerance (burned mask exact except cells within 0.5 s of max_time,
max |dt| <= 0.5 s), which is also asserted so a relaxation is visible.

At sizes the serial oracle cannot finish in seconds (4096^2) the checks are
size-independent: determinism across schedules, and that the result is a
fixed point (propagating again from it changes nothing).
"""
import numpy as np
import pytest

import oracle
from py_propag import synth

pytestmark = pytest.mark.gpu

FLT_MAX = oracle.FLT_MAX


def _settings(sc):
    import py_propag as pp

    g = pp.GeoReference(sc["width"], sc["height"], "+proj=utm +zone=30",
                        [0.0, sc["dx"], 0.0, 0.0, 0.0, sc["dy"]])
    return pp.Settings(g, float(sc["max_time"]))


def gpu_propagate(sc, fire, tuning=None, time=None, refs=None):
    import torch
    import py_propag as pp

    n = sc["width"] * sc["height"]
    dev = torch.device("cuda")
    f = [torch.from_numpy(np.ascontiguousarray(x, np.float32)).to(dev) for x in fire]
    t0 = sc["time"] if time is None else time
    rx0, ry0 = (sc["refs_x"], sc["refs_y"]) if refs is None else refs
    t = torch.from_numpy(t0.copy()).to(dev)
    rx = torch.from_numpy(rx0.view(np.int16).copy()).to(dev)
    ry = torch.from_numpy(ry0.view(np.int16).copy()).to(dev)
    b = torch.zeros((n + 31) // 32, dtype=torch.int32, device=dev)
    stats = pp.propagate_device(_settings(sc), *f, t, rx, ry, b, tuning=tuning)
    return {"time": t.cpu().numpy(), "refs_x": rx.cpu().numpy().view(np.uint16),
            "refs_y": ry.cpu().numpy().view(np.uint16), "boundaries": b.cpu().numpy().view(np.uint32),
            "stats": stats}


def oracle_propagate(sc, fire, nthreads=16):
    g = oracle.grid(sc["width"], sc["height"], sc["dx"], sc["dy"], sc["max_time"])
    return oracle.propagate(g, *fire, sc["time"], sc["refs_x"], sc["refs_y"], nthreads=nthreads)


def assert_parity(gpu, ref, max_time):
    # north-star tolerance first (so a failure says how far off it is) ...
    mg, mr = gpu["time"] < FLT_MAX, ref["time"] < FLT_MAX
    near = np.abs(ref["time"].astype(np.float64) - max_time) <= 0.5
    assert np.array_equal(mg[~near], mr[~near]), "burned mask differs"
    both = mg & mr
    dt = np.abs(gpu["time"][both].astype(np.float64) - ref["time"][both])
    assert dt.size == 0 or dt.max() <= 0.5, f"max |dt| = {dt.max()}"
    # ... then the bit-exact bar
    assert np.array_equal(gpu["time"].view(np.uint32), ref["time"].view(np.uint32))
    assert np.array_equal(gpu["refs_x"], ref["refs_x"])
    assert np.array_equal(gpu["refs_y"], ref["refs_y"])
    assert np.array_equal(gpu["boundaries"], ref["boundaries"])


def _fire(sc):
    return oracle.burn_f32(sc["terrain"], nthreads=16)


@pytest.mark.parametrize("side,seed", [(64, 1), (200, 2), (256, 3)])
def test_landscape_small(oracle_lib, side, seed):
    sc = synth.single_scenario(seed=seed, side=side)
    fire = _fire(sc)
    ref = oracle_propagate(sc, fire)
    gpu = gpu_propagate(sc, fire)
    assert_parity(gpu, ref, float(sc["max_time"]))
    assert gpu["stats"]["iterations"] == ref["sweeps"]
    assert gpu["stats"]["improvements"] == ref["improvements"]


def test_readme_known_answer_1024(oracle_lib):
    sc = synth.uniform_scenario(side=1024)
    fire = _fire(sc)
    ref = oracle_propagate(sc, fire)
    gpu = gpu_propagate(sc, fire)
    assert_parity(gpu, ref, 36000.0)
    assert int((gpu["time"] < FLT_MAX).sum()) == 450599


def test_single_scenario_1024(oracle_lib):
    sc = synth.single_scenario(seed=1, side=1024)
    fire = _fire(sc)
    ref = oracle_propagate(sc, fire)
    gpu = gpu_propagate(sc, fire)
    assert_parity(gpu, ref, 36000.0)
    print("GPU stats:", gpu["stats"])


@pytest.mark.parametrize("dx,dy", [(5.0, -5.0), (5.0, 10.0), (4.0, -3.0)])
def test_orientation_and_cell_shapes(oracle_lib, dx, dy):
    sc = synth.single_scenario(seed=5, side=160)
    sc["dx"], sc["dy"] = dx, dy
    fire = _fire(sc)
    assert_parity(gpu_propagate(sc, fire), oracle_propagate(sc, fire), 36000.0)


def test_multiple_ignitions_and_ragged_raster(oracle_lib):
    land = synth.landscape(seed=8, side=300, hills=6, patches=40)
    ign = [(5, 7, 0.0), (250, 40, 600.0), (150, 290, 1200.0), (299, 299, 0.0)]
    sc = synth.scenario(8, 300, 6, 40, ign, 7200.0, land=land)
    # a ragged (non-square, non-multiple-of-16) window of it
    W, H = 237, 181
    for k in ("fuel", "slope", "aspect", "wind_speed", "wind_azimuth", "d1hr", "d10hr", "d100hr",
              "herb", "wood"):
        sc["terrain"][k] = np.ascontiguousarray(sc["terrain"][k].reshape(300, 300)[:H, :W].ravel())
    sc["width"], sc["height"] = W, H
    ign = [(x, y, t) for x, y, t in ign if x < W and y < H]
    sc["time"], sc["refs_x"], sc["refs_y"] = synth.initial_state(W, H, ign)
    fire = _fire(sc)
    assert_parity(gpu_propagate(sc, fire), oracle_propagate(sc, fire), 7200.0)


def test_nofuel_ignition_and_tiny_rasters(oracle_lib):
    for side, fuel in ((1, 1), (3, 1), (40, 0), (17, 5)):
        sc = synth.uniform_scenario(side=side, fuel=fuel, ignition=(side // 2, side // 2),
                                    max_time=600.0)
        fire = _fire(sc)
        assert_parity(gpu_propagate(sc, fire), oracle_propagate(sc, fire), 600.0)


def test_schedule_independence(oracle_lib):
    sc = synth.single_scenario(seed=11, side=256)
    fire = _fire(sc)
    base = gpu_propagate(sc, fire)
    for tuning in ({"workgroups": 7}, {"workgroups": 1}, {"mode": 1}):
        other = gpu_propagate(sc, fire, tuning=tuning)
        for k in ("time", "refs_x", "refs_y", "boundaries"):
            assert np.array_equal(base[k], other[k]), (tuning, k)


def test_invalid_reference_is_rejected(oracle_lib):
    import py_propag as pp

    sc = synth.uniform_scenario(side=32, ignition=(4, 4))
    sc["refs_x"][4 + 4 * 32] = 40  # outside the raster
    with pytest.raises(pp.PropagValidationError):
        gpu_propagate(sc, _fire(sc))


def test_iteration_cap_reports_not_converged(oracle_lib):
    import py_propag as pp

    sc = synth.uniform_scenario(side=64, ignition=(2, 2))
    with pytest.raises(pp.PropagNotConverged):
        gpu_propagate(sc, _fire(sc), tuning={"max_iterations": 5})


def test_host_entry_point_matches_device_path(oracle_lib):
    import py_propag as pp

    sc = synth.single_scenario(seed=13, side=192)
    res = pp.propagate(_settings(sc), sc["terrain"], time=sc["time"], refs_x=sc["refs_x"],
                       refs_y=sc["refs_y"], return_fire=True)
    ref = oracle_propagate(sc, res.fire)
    got = {"time": res.arrival_times.ravel(), "refs_x": res.refs_x, "refs_y": res.refs_y,
           "boundaries": res.boundary_change}
    assert_parity(got, ref, 36000.0)


def test_large_scenario_4096_fixed_point(oracle_lib):
    """4096^2, 16 points + a 200-cell line, 12 h: deterministic across
    workgroup counts, and a fixed point of the update rule (propagating again
    from the result reproduces it)."""
    sc = synth.large_scenario(seed=2, side=4096)
    import torch
    import py_propag as pp

    fire = [np.zeros(4096 * 4096, np.float32) for _ in range(3)]
    dev = torch.device("cuda")
    cols = [torch.from_numpy(np.ascontiguousarray(sc["terrain"][k], np.float32)).to(dev)
            for k in oracle.TERRAIN_FIELDS]
    fuel = torch.from_numpy(sc["terrain"]["fuel"]).to(dev)
    out = [torch.zeros(4096 * 4096, dtype=torch.float32, device=dev) for _ in range(3)]
    pp.burn_device(fuel, *cols, *out)
    torch.cuda.synchronize()
    fire = [o.cpu().numpy() for o in out]
    first = gpu_propagate(sc, fire)
    second = gpu_propagate(sc, fire, tuning={"workgroups": 97})
    for k in ("time", "refs_x", "refs_y", "boundaries"):
        assert np.array_equal(first[k], second[k]), k
    burned = first["time"] < FLT_MAX
    assert burned.sum() > 100000
    # resume from the converged state: nothing earlier can be found
    again = gpu_propagate(sc, fire, time=first["time"], refs=(first["refs_x"], first["refs_y"]))
    assert np.array_equal(again["time"].view(np.uint32), first["time"].view(np.uint32))
    print("4096^2 stats:", first["stats"])



def _wide_scenario(W, H, seed):
    """A W x H band (sides past 32767 select the kernels with full 16-bit
    reference coordinates and line-of-sight walks at evaluation time):
    fuel codes in 64 x 8 blocks, one moisture draw, wind 6 m/s toward
    80 deg, slope 0.05 with blockwise aspect, 2 h, two ignitions near the east
    end (past x = 32767 when the band is wider than that)."""
    r = np.random.Generator(np.random.PCG64(seed))
    bx, by = (W + 63) // 64, (H + 7) // 8
    blocks = lambda v: np.kron(v, np.ones((8, 64), v.dtype))[:H, :W].ravel()  # noqa: E731
    n = W * H
    d1 = np.float32(r.uniform(0.03, 0.12))
    terrain = {"fuel": np.ascontiguousarray(blocks(r.integers(1, 14, (by, bx)).astype(np.uint8))),
               "d1hr": np.full(n, d1, np.float32), "d10hr": np.full(n, d1 + np.float32(0.01), np.float32),
               "d100hr": np.full(n, d1 + np.float32(0.02), np.float32),
               "herb": np.full(n, np.float32(r.uniform(0.3, 1.5)), np.float32),
               "wood": np.full(n, np.float32(r.uniform(0.6, 1.5)), np.float32),
               "wind_speed": np.full(n, 6.0, np.float32),
               "wind_azimuth": np.full(n, np.float32(np.radians(80.0)), np.float32),
               "slope": np.full(n, 0.05, np.float32),
               "aspect": np.ascontiguousarray(blocks(r.uniform(0, 2 * np.pi, (by, bx)).astype(np.float32)))}
    ign = [(W - 6500, H // 2, 0.0), (W - 3800, 3, 300.0)]
    time, rx, ry = synth.initial_state(W, H, ign)
    return {"terrain": terrain, "width": W, "height": H, "dx": 5.0, "dy": 5.0, "max_time": np.float32(7200.0),
            "time": time, "refs_x": rx, "refs_y": ry, "ignitions": ign}


def _raw_settings(sc):
    """Settings through the C structs directly: the Python GeoReference
    keeps the reference's 32768 limit, the C ABI takes sides up to 65535."""
    from py_propag import _native

    s = _native.SettingsC()
    s.geo_ref.width, s.geo_ref.height = sc["width"], sc["height"]
    s.geo_ref.proj[:len(b"+proj=utm +zone=30")] = b"+proj=utm +zone=30"
    gt = s.geo_ref.transform.gt
    gt.x0, gt.dx, gt.rx, gt.y0, gt.ry, gt.dy = 0.0, sc["dx"], 0.0, 0.0, 0.0, sc["dy"]
    inv = s.geo_ref.transform.inv
    inv.x0, inv.dx, inv.rx, inv.y0, inv.ry, inv.dy = 0.0, 1.0 / sc["dx"], 0.0, 0.0, 0.0, 1.0 / sc["dy"]
    s.max_time = float(sc["max_time"])
    return s


def _gpu_propagate_raw(sc, fire, tuning):
    import ctypes as C
    import torch
    import py_propag as pp
    from py_propag import _native

    n = sc["width"] * sc["height"]
    dev = torch.device("cuda")
    f = [torch.from_numpy(np.ascontiguousarray(x, np.float32)).to(dev) for x in fire]
    t = torch.from_numpy(sc["time"].copy()).to(dev)
    rx = torch.from_numpy(sc["refs_x"].view(np.int16).copy()).to(dev)
    ry = torch.from_numpy(sc["refs_y"].view(np.int16).copy()).to(dev)
    b = torch.zeros((n + 31) // 32, dtype=torch.int32, device=dev)
    tn = _native.Tuning()
    for k, v in tuning.items():
        setattr(tn, k, v)
    st = _native.Stats()
    s = _raw_settings(sc)
    ptr = lambda x: C.c_void_p(x.data_ptr())  # noqa: E731
    rc = _native.load().pg_propagate(C.byref(s), *[ptr(x) for x in f], ptr(t), ptr(rx), ptr(ry), ptr(b),
                                     C.byref(tn), C.byref(st), C.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 0, _native.last_error()
    return {"time": t.cpu().numpy(), "refs_x": rx.cpu().numpy().view(np.uint16),
            "refs_y": ry.cpu().numpy().view(np.uint16), "boundaries": b.cpu().numpy().view(np.uint32),
            "stats": st.as_dict()}


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("W,H,lds", [(32768, 16, 1), (32768, 48, 0), (40000, 16, 1)])
def test_wide_rasters_uncached_path(oracle_lib, W, H, lds, mode):
    """Bands 32768 and 40000 cells wide, where references need all 16 bits
    of a coordinate: the boundary bitset in LDS (H = 16) and in L2
    (H = 48), persistent and per-sweep launches, bit-exact vs the oracle.
    The 40000-wide band goes through the C ABI directly (the Python
    GeoReference keeps the reference's 32768 limit)."""
    sc = _wide_scenario(W, H, seed=H)
    fire = _fire(sc)
    ref = oracle_propagate(sc, fire)
    if W > 32768:
        gpu = _gpu_propagate_raw(sc, fire, {"mode": mode})
        assert ref["refs_x"][ref["time"] < FLT_MAX].max() > 32767
    else:
        gpu = gpu_propagate(sc, fire, tuning={"mode": mode})
    assert (ref["time"] < FLT_MAX).sum() > 5000
    assert_parity(gpu, ref, 7200.0)
    assert gpu["stats"]["iterations"] == ref["sweeps"]
    if mode == 0:
        assert gpu["stats"]["lds_bitset"] == lds


def test_shutdown_releases_and_the_next_call_recreates(oracle_lib):
    """pg_shutdown frees the workspaces, cached launch graphs (built by a
    mode-2 batched launch first) and the device catalog; the calls after it
    rebuild them and give the same results."""
    import py_propag as pp
    from py_propag import _native

    sc = synth.single_scenario(seed=6, side=128)
    fire = _fire(sc)
    before = gpu_propagate(sc, fire)
    burn_before = pp.burn(sc["terrain"])
    w = 96
    land = synth.landscape(3, w, 2, 8)
    g = pp.GeoReference(w, w, "+proj=utm +zone=30", [0.0, 5.0, 0.0, 0.0, 0.0, 5.0])
    wx = {"d1hr": 0.06, "d10hr": 0.07, "d100hr": 0.08, "herb": 0.9, "wood": 1.0, "wind_speed": 5.0,
          "wind_azimuth": 1.0}
    with pp.Context() as ctx:
        tid = ctx.add_terrain(g, land["fuel"], land["slope"], land["aspect"])
        ctx.admit(tid, wx, [(40, 40, 0.0)], 3600.0)
        ctx.run(tuning={"mode": 2})
    assert _native.load().pg_shutdown() == 0
    after = gpu_propagate(sc, fire)
    for k in ("time", "refs_x", "refs_y", "boundaries"):
        assert np.array_equal(before[k], after[k]), k
    for a, b in zip(burn_before, pp.burn(sc["terrain"])):
        assert np.array_equal(a, b)
