"""The batched (2048^2 ensemble) and served (512-768 landscapes, request
loop) workloads checked against the oracle directly, at their real raster
sizes.

Every batched or served result is first compared with the one-scenario GPU
path (bit-exact); then the scenarios the serial oracle can finish in a few
seconds (estimated from the result: burned bounding box x sweeps) are run
through the oracle (16 threads) on the same fire rasters - the GPU burn's,
which the one-scenario path returns - and compared bit for bit, after the
north star's tolerance (burned mask exact except within 0.5 s of max_time,
max |dt| <= 0.5 s).  The oracle's f32 burn agrees with the GPU burn only to
1e-5 relative, which the propagation's ties can amplify, hence the GPU's
fire rasters on both sides.
"""
import numpy as np
import pytest

import oracle
from py_propag import synth

pytestmark = pytest.mark.gpu

FLT_MAX = oracle.FLT_MAX
ORACLE_BUDGET = 5e8     # cell relaxations the oracle may spend on one scenario (~3 s on 16 threads)


def _geo(pp, side, cell=5.0):
    return pp.GeoReference(side, side, "+proj=utm +zone=30", [0.0, cell, 0.0, 0.0, 0.0, cell])


def _full_terrain(land, wx, n):
    t = {"fuel": land["fuel"], "slope": land["slope"], "aspect": land["aspect"]}
    for k in ("d1hr", "d10hr", "d100hr", "herb", "wood", "wind_speed", "wind_azimuth"):
        t[k] = np.full(n, np.float32(wx[k]), np.float32)
    return t


def _cost(t, side, sweeps):
    """Relaxations the oracle spends: it sweeps the burned bounding box grown
    by one cell, once per sweep (bounded above by the final box)."""
    burned = np.nonzero(t.ravel() < FLT_MAX)[0]
    if burned.size == 0:
        return 0
    xs, ys = burned % side, burned // side
    return (int(xs.max() - xs.min()) + 3) * (int(ys.max() - ys.min()) + 3) * int(sweeps)


def _check(pp, got, land, side, case):
    """got = (time, refs_x, refs_y, bits) of a batched / served scenario.
    Returns True if the oracle was run on it."""
    t, rx, ry, bits = got
    settings = pp.Settings(_geo(pp, side), case["max_time"])
    alone = pp.propagate(settings, _full_terrain(land, case["wx"], side * side), ignitions=case["ign"],
                         return_fire=True)
    assert np.array_equal(t.view(np.uint32).ravel(), alone.arrival_times.view(np.uint32).ravel())
    assert np.array_equal(rx, alone.refs_x) and np.array_equal(ry, alone.refs_y)
    assert np.array_equal(bits, alone.boundary_change)
    if _cost(t, side, alone.stats["iterations"]) > ORACLE_BUDGET:
        return False
    fire = tuple(np.ascontiguousarray(f, np.float32).ravel() for f in alone.fire)
    t0, rx0, ry0 = synth.initial_state(side, side, case["ign"])
    ref = oracle.propagate(oracle.grid(side, side, 5.0, 5.0, case["max_time"]), *fire, t0, rx0, ry0,
                           nthreads=16)
    tg = t.ravel()
    mg, mr = tg < FLT_MAX, ref["time"] < FLT_MAX
    near = np.abs(ref["time"].astype(np.float64) - case["max_time"]) <= 0.5
    assert np.array_equal(mg[~near], mr[~near]), "burned mask differs"
    both = mg & mr
    if both.any():
        assert np.abs(tg[both].astype(np.float64) - ref["time"][both]).max() <= 0.5
    assert np.array_equal(tg.view(np.uint32), ref["time"].view(np.uint32))
    assert np.array_equal(rx, ref["refs_x"]) and np.array_equal(ry, ref["refs_y"])
    assert np.array_equal(bits, ref["boundaries"])
    return True


def test_ensemble_2048_vs_oracle():
    """Twelve scenarios of the 2048^2 ensemble workload (shared landscape,
    per-scenario wind, moisture, 1-4 ignitions, max_time log-uniform in
    1-12 h) in one shared launch; each equals the one-scenario path, and
    those the oracle can afford equal the oracle."""
    import py_propag as pp

    e = synth.ensemble_2048(seed=3, count=12)
    side, land, cases = e["side"], e["land"], e["cases"]
    with pp.Context() as ctx:
        tid = ctx.add_terrain(_geo(pp, side), land["fuel"], land["slope"], land["aspect"])
        ids = [ctx.admit(tid, c["wx"], c["ign"], c["max_time"]) for c in cases]
        fin, _ = ctx.run()
        assert fin == len(cases)
        checked = sum(_check(pp, ctx.result(i), land, side, c) for i, c in zip(ids, cases))
    print(f"2048^2 ensemble: {checked} of {len(cases)} scenarios checked against the oracle")
    assert checked >= 2


def test_served_requests_vs_oracle():
    """The request loop over a 512^2 and a 768^2 landscape (the serving
    workload's generator and weather; max_time U(3600, 36000) s), arrivals
    spread over 0.1 s so requests join launches already running, launches
    of a few sweeps; every delivered result equals the one-scenario path,
    and every one the oracle can afford (at least two of the 512^2 ones)
    equals the oracle."""
    import py_propag as pp
    from py_propag.serve import Request, Server

    sides = [512, 768]
    lands = [synth.landscape(4000 + i, s, max(1, round(20 * s / 1024)), max(4, round(200 * (s / 1024) ** 2)))
             for i, s in enumerate(sides)]
    r = np.random.Generator(np.random.PCG64(77))
    cases = []
    for k in range(8):
        j = k % 2
        c = synth.scenario_weather(r, int(r.integers(1, 5)), sides[j], sides[j], r.uniform(3600.0, 36000.0))
        c.update({"terrain": j, "arrival": 0.0125 * k})
        cases.append(c)
    got = {}
    with pp.Context() as ctx:
        tids = [ctx.add_terrain(_geo(pp, s), L["fuel"], L["slope"], L["aspect"]) for s, L in zip(sides, lands)]

        def on_done(req, sid):
            got[req.tag] = ctx.result(sid)

        server = Server(ctx, on_done=on_done, target_ms=1.0, min_sweeps=8)
        reqs = [Request(tids[c["terrain"]], c["wx"], c["ign"], c["max_time"], c["arrival"], tag=i)
                for i, c in enumerate(cases)]
        done = server.play(reqs)
        assert len(done) == len(cases) and all(q.state == pp.DONE for q in done)
        assert server.launches > 2
    checked = {0: 0, 1: 0}
    for i, c in enumerate(cases):
        j = c["terrain"]
        checked[j] += _check(pp, got[i], lands[j], sides[j], c)
    print(f"served: oracle checked {checked[0]} of 4 on 512^2, {checked[1]} of 4 on 768^2")
    assert checked[0] >= 2
