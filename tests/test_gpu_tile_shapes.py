"""Tile shapes 16 x 4 and 32 x 2 beside the default 8 x 8 (pg_tuning.tile_w /
tile_h for one scenario per call, pg_ctx_set_tile for the batching context).

Every shape is one cell per lane of a 64-lane wave and the schedule is the
oracle's exact Jacobi whatever the tiling, so each shape must give the
oracle's arrival times, references and boundary bits bit for bit, with the
oracle's sweep and improvement counts: in both launch forms, with the
boundary bitset in LDS (config-1's 1024^2 scenario, where the larger halo
slices of 16 x 4 and 32 x 2 leave 7 and 6 waves per workgroup beside it)
and in L2, on ragged rasters, with the full 16-bit references of a band
wider than 32767 cells, batched, and through the allocation-free entry point.
"""
import numpy as np
import pytest

import oracle
from py_propag import synth

from test_gpu_propagate import (_fire, _gpu_propagate_raw, _wide_scenario, assert_parity, gpu_propagate,
                                oracle_propagate)
from test_gpu_batch import _alone, _case, _geo, _same, _terrain

pytestmark = pytest.mark.gpu

FLT_MAX = oracle.FLT_MAX
SHAPES = [(16, 4), (32, 2)]


def _ragged():
    land = synth.landscape(seed=8, side=300, hills=6, patches=40)
    ign = [(5, 7, 0.0), (250, 40, 600.0), (150, 290, 1200.0), (299, 299, 0.0)]
    sc = synth.scenario(8, 300, 6, 40, ign, 7200.0, land=land)
    W, H = 237, 181  # not a multiple of any tile edge
    for k in ("fuel", "slope", "aspect", "wind_speed", "wind_azimuth", "d1hr", "d10hr", "d100hr",
              "herb", "wood"):
        sc["terrain"][k] = np.ascontiguousarray(sc["terrain"][k].reshape(300, 300)[:H, :W].ravel())
    sc["width"], sc["height"] = W, H
    ign = [(x, y, t) for x, y, t in ign if x < W and y < H]
    sc["time"], sc["refs_x"], sc["refs_y"] = synth.initial_state(W, H, ign)
    return sc


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("tw,th", SHAPES)
def test_shapes_match_the_oracle(oracle_lib, tw, th, mode):
    for sc in (synth.single_scenario(seed=2, side=200), _ragged()):
        fire = _fire(sc)
        ref = oracle_propagate(sc, fire)
        gpu = gpu_propagate(sc, fire, tuning={"tile_w": tw, "tile_h": th, "mode": mode})
        assert_parity(gpu, ref, float(sc["max_time"]))
        assert gpu["stats"]["iterations"] == ref["sweeps"]
        assert gpu["stats"]["improvements"] == ref["improvements"]


@pytest.mark.parametrize("tw,th", SHAPES)
def test_shapes_config1_bitset_in_lds(oracle_lib, tw, th):
    sc = synth.single_scenario(seed=1, side=1024)
    fire = _fire(sc)
    ref = oracle_propagate(sc, fire)
    gpu = gpu_propagate(sc, fire, tuning={"tile_w": tw, "tile_h": th})
    assert gpu["stats"]["lds_bitset"] == 1
    assert_parity(gpu, ref, 36000.0)
    print(f"{tw}x{th} config-1:", {k: gpu["stats"][k] for k in ("relax_ms", "tiles_visited", "iterations")})


@pytest.mark.parametrize("tw,th", SHAPES)
def test_shapes_wide_band_full_references(oracle_lib, tw, th):
    sc = _wide_scenario(40000, 16, seed=16)
    fire = _fire(sc)
    ref = oracle_propagate(sc, fire)
    gpu = _gpu_propagate_raw(sc, fire, {"tile_w": tw, "tile_h": th})
    assert ref["refs_x"][ref["time"] < FLT_MAX].max() > 32767
    assert_parity(gpu, ref, 7200.0)


@pytest.mark.parametrize("tw,th", SHAPES)
def test_shapes_batched_match_alone(tw, th):
    import py_propag as pp

    w, h = 200, 150
    land = _terrain(11, w, h)
    cases = [_case(s, w, h, field=(s == 3)) for s in range(6)]
    with pp.Context() as ctx:
        ctx.set_tile(tw, th)
        tid = ctx.add_terrain(_geo(pp, w, h), land["fuel"], land["slope"], land["aspect"])
        ids = [ctx.admit(tid, c["wx"], c["ign"], c["max_time"]) for c in cases]
        # another shape while scenarios are admitted: refused, nothing changes
        with pytest.raises(pp.PropagCudaError, match="BUSY"):
            ctx.set_tile(8, 8)
        # a tuning naming another shape: refused before any launch
        with pytest.raises(pp.PropagValidationError):
            ctx.run(tuning={"tile_w": 8, "tile_h": 8})
        fin, _ = ctx.run(tuning={"tile_w": tw, "tile_h": th, "workgroups": 40})
        assert fin == len(cases)
        for sid, c in zip(ids, cases):
            assert ctx.status(sid)[0] == pp.DONE
            _same(ctx.result(sid), _alone(pp, land, w, h, c["wx"], c["ign"], c["max_time"]))
            ctx.release(sid)
        ctx.set_tile(8, 8)  # allowed again once every scenario is released


def test_invalid_shapes_are_refused():
    import py_propag as pp

    sc = synth.uniform_scenario(side=32, ignition=(4, 4))
    fire = _fire(sc)
    for tw, th in ((8, 4), (64, 1), (4, 16), (16, 0)):
        with pytest.raises(pp.PropagValidationError, match="tile shape"):
            gpu_propagate(sc, fire, tuning={"tile_w": tw, "tile_h": th})
    with pp.Context() as ctx:
        with pytest.raises(pp.PropagValidationError, match="tile shape"):
            ctx.set_tile(2, 32)


@pytest.mark.parametrize("tw,th", SHAPES)
def test_shapes_through_the_allocation_free_entry(oracle_lib, tw, th):
    import py_propag as pp
    import torch

    sc = synth.single_scenario(seed=3, side=256)
    fire = _fire(sc)
    ref = oracle_propagate(sc, fire)
    n = sc["width"] * sc["height"]
    dev = torch.device("cuda")
    f = [torch.from_numpy(np.ascontiguousarray(x, np.float32)).to(dev) for x in fire]
    t = torch.from_numpy(sc["time"].copy()).to(dev)
    rx = torch.from_numpy(sc["refs_x"].view(np.int16).copy()).to(dev)
    ry = torch.from_numpy(sc["refs_y"].view(np.int16).copy()).to(dev)
    b = torch.zeros((n + 31) // 32, dtype=torch.int32, device=dev)
    from test_gpu_propagate import _settings

    pp.propagate_device(_settings(sc), *f, t, rx, ry, b, tuning={"tile_w": tw, "tile_h": th}, noalloc=True)
    gpu = {"time": t.cpu().numpy(), "refs_x": rx.cpu().numpy().view(np.uint16),
           "refs_y": ry.cpu().numpy().view(np.uint16), "boundaries": b.cpu().numpy().view(np.uint32)}
    assert_parity(gpu, ref, float(sc["max_time"]))
