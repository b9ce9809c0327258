"""pg_propagate_noalloc (the allocation-free entry point) against pg_propagate.

Both entry points run the same kernels; the allocation-free one carves every
workspace buffer from the library's fixed device array instead of
hipMalloc'ed memory.  The bar is bit-exact arrival times, references,
boundary bits and sweep counts, on two calls in a row where the second raster
is smaller (the array keeps the first call's contents, so anything the call
fails to clear shows up), in both launch forms it accepts, plus the size
limit and the refused launch forms.
"""
import numpy as np
import pytest

import oracle
from py_propag import synth

pytestmark = pytest.mark.gpu


def _settings(sc):
    import py_propag as pp

    g = pp.GeoReference(sc["width"], sc["height"], "+proj=utm +zone=30",
                        [0.0, sc["dx"], 0.0, 0.0, 0.0, sc["dy"]])
    return pp.Settings(g, float(sc["max_time"]))


def _run(sc, fire, noalloc, tuning=None):
    import torch
    import py_propag as pp

    n = sc["width"] * sc["height"]
    dev = torch.device("cuda")
    f = [torch.from_numpy(np.ascontiguousarray(x, np.float32)).to(dev) for x in fire]
    t = torch.from_numpy(sc["time"].copy()).to(dev)
    rx = torch.from_numpy(sc["refs_x"].view(np.int16).copy()).to(dev)
    ry = torch.from_numpy(sc["refs_y"].view(np.int16).copy()).to(dev)
    b = torch.zeros((n + 31) // 32, dtype=torch.int32, device=dev)
    st = pp.propagate_device(_settings(sc), *f, t, rx, ry, b, tuning=tuning, noalloc=noalloc)
    return {"time": t.cpu().numpy().view(np.uint32), "refs_x": rx.cpu().numpy().view(np.uint16),
            "refs_y": ry.cpu().numpy().view(np.uint16), "boundaries": b.cpu().numpy().view(np.uint32),
            "stats": st}


def _same(a, b):
    for k in ("time", "refs_x", "refs_y", "boundaries"):
        assert np.array_equal(a[k], b[k]), k
    # sweeps and commits are fixed by the exact schedule; how many cells are
    # evaluated is not (a stamp written in the same sweep may or may not be
    # seen, which only adds evaluations that commit nothing)
    for k in ("iterations", "improvements"):
        assert a["stats"][k] == b["stats"][k], k


@pytest.mark.parametrize("mode", [0, 1])
def test_noalloc_matches_pg_propagate_twice(oracle_lib, mode):
    """Two calls in a row, the second on a smaller raster."""
    tuning = {"mode": mode}
    for side, seed in ((320, 11), (120, 12)):
        sc = synth.single_scenario(seed=seed, side=side)
        fire = oracle.burn_f32(sc["terrain"], nthreads=16)
        want = _run(sc, fire, False, tuning)
        got = _run(sc, fire, True, tuning)
        assert (got["time"] != np.float32(oracle.FLT_MAX).view(np.uint32)).sum() > 100
        _same(got, want)
        assert got["stats"]["elapsed_ms"] == 0.0


def test_noalloc_at_its_cell_limit(oracle_lib):
    """1024 x 1024 (the config-1 landscape) is at the limit; one more row is
    refused before any device work."""
    import py_propag as pp

    sc = synth.single_scenario(seed=1, side=1024)
    fire = oracle.burn_f32(sc["terrain"], nthreads=16)
    sc["max_time"] = 3000.0
    _same(_run(sc, fire, True), _run(sc, fire, False))
    big = synth.uniform_scenario(side=8, ignition=(2, 2))
    big["width"], big["height"] = 1024, 1025
    n = 1024 * 1025
    for k in ("time", "refs_x", "refs_y"):
        big[k] = np.resize(big[k], n)
    with pytest.raises(pp.PropagValidationError, match="exceeds its maximum"):
        _run(big, [np.zeros(n, np.float32)] * 3, True)


def test_noalloc_refuses_allocating_forms(oracle_lib):
    import py_propag as pp

    sc = synth.single_scenario(seed=2, side=64)
    fire = oracle.burn_f32(sc["terrain"])
    for tuning in ({"mode": 2}, {"profile": 1}):
        with pytest.raises(pp.PropagValidationError, match="allocate"):
            _run(sc, fire, True, tuning)
