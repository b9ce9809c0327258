"""Batched and served forms (the pg_ctx_* handle API) on the GPU.

The bar: every scenario's arrival times, references and boundary bits are
bit-identical to the one-scenario path (pg_propagate_host, itself pinned to
the oracle by test_gpu_propagate) on the same inputs, whatever else shares
its launches - other scenarios, other raster sizes, scenarios admitted
while it runs (continuous batching), launches cut after a few sweeps, and a
running neighbour being released.  One batched scenario is also compared
with the oracle directly.
"""
import math

import numpy as np
import pytest

import oracle
from py_propag import synth

pytestmark = pytest.mark.gpu

FLT_MAX = oracle.FLT_MAX


def _geo(pp, w, h, cell=5.0):
    return pp.GeoReference(w, h, "+proj=utm +zone=30", [0.0, cell, 0.0, 0.0, 0.0, cell])


def _terrain(seed, w, h):
    side = max(w, h)
    land = synth.landscape(seed, side, hills=max(1, side * side // 50000), patches=max(4, side * side // 5000))
    sl = lambda a: np.ascontiguousarray(np.asarray(a).reshape(side, side)[:h, :w].ravel())  # noqa: E731
    return {"fuel": sl(land["fuel"]), "slope": sl(land["slope"]), "aspect": sl(land["aspect"])}


def _weather(seed, w=None, h=None, field=False):
    r = np.random.Generator(np.random.PCG64(seed))
    d1 = np.float32(r.uniform(0.03, 0.12))
    wx = {"d1hr": d1, "d10hr": np.float32(d1 + np.float32(0.01)), "d100hr": np.float32(d1 + np.float32(0.02)),
          "herb": np.float32(r.uniform(0.3, 1.5)), "wood": np.float32(r.uniform(0.6, 1.5))}
    if field:
        yy, xx = np.meshgrid(np.arange(h), np.arange(w), indexing="ij")
        wx["wind_speed"] = (6.0 + 2.0 * np.sin(xx / 37.0) * np.cos(yy / 23.0)).astype(np.float32)
        wx["wind_azimuth"] = np.radians(200.0 + 30.0 * np.sin(yy / 41.0)).astype(np.float32)
    else:
        wx["wind_speed"] = np.float32(r.uniform(2.0, 12.0))
        wx["wind_azimuth"] = np.float32(math.radians(r.uniform(0.0, 360.0)))
    return wx


def _ignitions(seed, w, h, k):
    r = np.random.Generator(np.random.PCG64(seed + 17))
    cells = set()
    while len(cells) < k:
        cells.add((int(r.integers(w // 6, w - w // 6)), int(r.integers(h // 6, h - h // 6))))
    return [(x, y, float(i) * 60.0) for i, (x, y) in enumerate(sorted(cells))]


def _full_terrain(land, wx, n):
    t = {"fuel": land["fuel"], "slope": land["slope"], "aspect": land["aspect"]}
    for k in ("d1hr", "d10hr", "d100hr", "herb", "wood", "wind_speed", "wind_azimuth"):
        v = np.asarray(wx[k], np.float32)
        t[k] = np.full(n, v.item(), np.float32) if v.size == 1 else np.ascontiguousarray(v.ravel())
    return t


def _alone(pp, land, w, h, wx, ign, max_time):
    res = pp.propagate(pp.Settings(_geo(pp, w, h), max_time), _full_terrain(land, wx, w * h), ignitions=ign)
    return res


def _same(batched, alone):
    t, rx, ry, bits = batched
    assert np.array_equal(t.view(np.uint32), alone.arrival_times.view(np.uint32))
    assert np.array_equal(rx, alone.refs_x) and np.array_equal(ry, alone.refs_y)
    assert np.array_equal(bits, alone.boundary_change)


def _case(seed, w, h, field=False):
    return {"wx": _weather(seed, w, h, field), "ign": _ignitions(seed, w, h, 1 + seed % 3),
            "max_time": float(1800.0 * (1 + seed % 7))}


def test_batch_matches_alone_and_oracle(oracle_lib):
    import py_propag as pp

    w, h = 200, 150  # not a multiple of the tile edge
    land = _terrain(11, w, h)
    cases = [_case(s, w, h, field=(s == 3)) for s in range(6)]
    with pp.Context() as ctx:
        tid = ctx.add_terrain(_geo(pp, w, h), land["fuel"], land["slope"], land["aspect"])
        ids = [ctx.admit(tid, c["wx"], c["ign"], c["max_time"]) for c in cases]
        fin, stats = ctx.run()
        assert fin == len(cases)
        for sid, c in zip(ids, cases):
            state, sweeps = ctx.status(sid)
            assert state == pp.DONE
            alone = _alone(pp, land, w, h, c["wx"], c["ign"], c["max_time"])
            _same(ctx.result(sid), alone)
            assert sweeps == alone.stats["iterations"]
        # scenario 0's propagation against the oracle directly, on the same
        # fire rasters (the GPU burn of the one-scenario path: the oracle's
        # own f32 burn agrees with it only to 1e-5, which the propagation can
        # amplify)
        c = cases[0]
        res = pp.propagate(pp.Settings(_geo(pp, w, h), c["max_time"]), _full_terrain(land, c["wx"], w * h),
                           ignitions=c["ign"], return_fire=True)
        fire = tuple(np.ascontiguousarray(f, np.float32).ravel() for f in res.fire)
        time0, rx0, ry0 = synth.initial_state(w, h, c["ign"])
        ref = oracle.propagate(oracle.grid(w, h, 5.0, 5.0, c["max_time"]), *fire, time0, rx0, ry0, nthreads=16)
        t, rx, ry, bits = ctx.result(ids[0])
        assert np.array_equal(t.ravel().view(np.uint32), ref["time"].view(np.uint32))
        assert np.array_equal(rx, ref["refs_x"]) and np.array_equal(ry, ref["refs_y"])
        assert np.array_equal(bits, ref["boundaries"])
    print(f"batched 6 scenarios: {stats['iterations']} sweeps, {stats['tiles_visited']} tile visits")


@pytest.mark.parametrize("w,h", [(160, 96), (150, 97)])
def test_admission_paths_match_alone(w, h):
    """Constant-weather scenarios with at most 16 ignitions are admitted in
    two launches (burn, coefficients and state in one pass; the bitset pass
    seeds the ignitions), others through the general set-up (burn pass,
    device ignition list, prep): both give the one-per-call results, on a
    width that is a multiple of 32 (fused bitset pass) and one that is not."""
    import py_propag as pp

    land = _terrain(23, w, h)
    many = _ignitions(5, w, h, 20)  # above the 16 passed by value
    few = _ignitions(6, w, h, 16)
    cases = [{"wx": _weather(40), "ign": few, "max_time": 5400.0},
             {"wx": _weather(41), "ign": many, "max_time": 5400.0},
             {"wx": _weather(42, w, h, field=True), "ign": few[:3], "max_time": 3600.0}]
    with pp.Context() as ctx:
        tid = ctx.add_terrain(_geo(pp, w, h), land["fuel"], land["slope"], land["aspect"])
        ids = [ctx.admit(tid, c["wx"], c["ign"], c["max_time"]) for c in cases]
        fin, _ = ctx.run()
        assert fin == len(cases)
        for sid, c in zip(ids, cases):
            _same(ctx.result(sid), _alone(pp, land, w, h, c["wx"], c["ign"], c["max_time"]))


def test_continuous_batching_admits_between_short_launches():
    import py_propag as pp

    w, h = 160, 160
    land = _terrain(5, w, h)
    cases = [_case(s, w, h) for s in range(10, 17)]
    with pp.Context() as ctx:
        tid = ctx.add_terrain(_geo(pp, w, h), land["fuel"], land["slope"], land["aspect"])
        pending = list(range(len(cases)))
        ids = {}
        launches = 0
        while pending or any(ctx.status(s)[0] == pp.RUNNING for s in ids.values()):
            for _ in range(2):  # two arrivals per launch
                if pending:
                    i = pending.pop(0)
                    ids[i] = ctx.admit(tid, cases[i]["wx"], cases[i]["ign"], cases[i]["max_time"])
            ctx.run(max_sweeps=9)
            launches += 1
            assert launches < 1000
        for i, sid in ids.items():
            assert ctx.status(sid)[0] == pp.DONE
            c = cases[i]
            _same(ctx.result(sid), _alone(pp, land, w, h, c["wx"], c["ign"], c["max_time"]))
    print(f"{len(cases)} scenarios over {launches} launches of at most 9 sweeps")


def test_mixed_raster_sizes_share_a_launch():
    import py_propag as pp

    shapes = [(96, 72), (256, 200), (64, 300)]
    with pp.Context() as ctx:
        lands, tids, ids, cases = [], [], [], []
        for j, (w, h) in enumerate(shapes):
            land = _terrain(20 + j, w, h)
            lands.append(land)
            tids.append(ctx.add_terrain(_geo(pp, w, h), land["fuel"], land["slope"], land["aspect"]))
        for s in range(7):
            j = s % len(shapes)
            w, h = shapes[j]
            c = _case(30 + s, w, h)
            cases.append((j, c))
            ids.append(ctx.admit(tids[j], c["wx"], c["ign"], c["max_time"]))
        ctx.run()
        for sid, (j, c) in zip(ids, cases):
            w, h = shapes[j]
            _same(ctx.result(sid), _alone(pp, lands[j], w, h, c["wx"], c["ign"], c["max_time"]))


def test_release_of_a_running_scenario_and_slot_reuse():
    import py_propag as pp

    w, h = 128, 128
    land = _terrain(8, w, h)
    a, b, c = (_case(s, w, h) for s in (40, 41, 42))
    a["max_time"] = b["max_time"] = 30000.0
    with pp.Context() as ctx:
        tid = ctx.add_terrain(_geo(pp, w, h), land["fuel"], land["slope"], land["aspect"])
        ia = ctx.admit(tid, a["wx"], a["ign"], a["max_time"])
        ib = ctx.admit(tid, b["wx"], b["ign"], b["max_time"])
        ctx.run(max_sweeps=5)
        assert ctx.status(ia)[0] == pp.RUNNING
        ctx.release(ia)                      # its tiles are still queued
        ic = ctx.admit(tid, c["wx"], c["ign"], c["max_time"])
        assert ic != ia                      # the slot waits for the next launch
        ctx.run()
        _same(ctx.result(ib), _alone(pp, land, w, h, b["wx"], b["ign"], b["max_time"]))
        _same(ctx.result(ic), _alone(pp, land, w, h, c["wx"], c["ign"], c["max_time"]))
        idd = ctx.admit(tid, a["wx"], a["ign"], a["max_time"])  # slot reusable now
        ctx.run()
        _same(ctx.result(idd), _alone(pp, land, w, h, a["wx"], a["ign"], a["max_time"]))


def test_memory_budget_refuses_then_admits():
    import py_propag as pp

    w, h = 96, 96
    land = _terrain(9, w, h)
    c = _case(50, w, h)
    budget = int(pp.Context.scenario_bytes(w, h) * 1.5)
    with pp.Context(hbm_budget=budget) as ctx:
        tid = ctx.add_terrain(_geo(pp, w, h), land["fuel"], land["slope"], land["aspect"])
        s1 = ctx.admit(tid, c["wx"], c["ign"], c["max_time"])
        assert s1 is not None
        assert ctx.admit(tid, c["wx"], c["ign"], c["max_time"]) is None
        ctx.run()
        ctx.release(s1)
        s2 = ctx.admit(tid, c["wx"], c["ign"], c["max_time"])
        assert s2 is not None
        ctx.run()
        _same(ctx.result(s2), _alone(pp, land, w, h, c["wx"], c["ign"], c["max_time"]))


def test_sweep_cap_reports_not_converged():
    import py_propag as pp

    w, h = 96, 96
    land = _terrain(10, w, h)
    c = _case(60, w, h)
    c["max_time"] = 20000.0
    with pp.Context() as ctx:
        tid = ctx.add_terrain(_geo(pp, w, h), land["fuel"], land["slope"], land["aspect"])
        s = ctx.admit(tid, c["wx"], c["ign"], c["max_time"], max_iterations=3)
        ok = ctx.admit(tid, c["wx"], c["ign"], c["max_time"])
        ctx.run()
        assert ctx.status(s) == (pp.NOT_CONVERGED, 3)
        assert ctx.status(ok)[0] == pp.DONE


def test_single_scenario_1024_through_the_context():
    import py_propag as pp

    sc = synth.single_scenario(seed=1, side=1024)
    t = sc["terrain"]
    w = h = 1024
    wx = {k: np.float32(t[k][0]) for k in ("d1hr", "d10hr", "d100hr", "herb", "wood")}
    wx["wind_speed"], wx["wind_azimuth"] = t["wind_speed"], t["wind_azimuth"]
    land = {"fuel": t["fuel"], "slope": t["slope"], "aspect": t["aspect"]}
    with pp.Context() as ctx:
        tid = ctx.add_terrain(_geo(pp, w, h), land["fuel"], land["slope"], land["aspect"])
        sid = ctx.admit(tid, wx, sc["ignitions"], float(sc["max_time"]))
        ctx.run()
        _same(ctx.result(sid), _alone(pp, land, w, h, wx, sc["ignitions"], float(sc["max_time"])))


def test_server_stream_each_result_same_as_alone():
    """The request loop: arrivals spread in time over two landscapes, launches
    of a few sweeps, a tight memory budget (admission waits for releases);
    every delivered result equals the one-scenario path's."""
    import py_propag as pp
    from py_propag.serve import Request, Server

    shapes = [(120, 120), (72, 200)]
    lands = [_terrain(70 + j, w, h) for j, (w, h) in enumerate(shapes)]
    cases = []
    for s in range(12):
        j = s % 2
        w, h = shapes[j]
        c = _case(80 + s, w, h)
        c.update({"terrain": j, "arrival": 0.004 * s})
        cases.append(c)
    budget = 4 * max(pp.Context.scenario_bytes(w, h) for w, h in shapes)
    got = {}
    with pp.Context(hbm_budget=budget) as ctx:
        tids = [ctx.add_terrain(_geo(pp, w, h), L["fuel"], L["slope"], L["aspect"]) for (w, h), L in zip(shapes, lands)]

        def on_done(req, sid):
            got[req.tag] = ctx.result(sid)

        server = Server(ctx, on_done=on_done, target_ms=0.2, min_sweeps=4)
        reqs = [Request(tids[c["terrain"]], c["wx"], c["ign"], c["max_time"], c["arrival"], tag=i)
                for i, c in enumerate(cases)]
        done = server.play(reqs)
        assert len(done) == len(cases) and server.idle()
        assert server.launches > 2
        assert ctx.take_finished() == []
    for i, c in enumerate(cases):
        w, h = shapes[c["terrain"]]
        _same(got[i], _alone(pp, lands[c["terrain"]], w, h, c["wx"], c["ign"], c["max_time"]))
    print(f"served {len(cases)} requests in {server.launches} launches, {server.busy_refusals} refusals")


def test_graph_launches_match_direct_launches():
    """mode 2 (the persistent launch through a cached, re-parameterised
    hipGraph), replayed over many short launches with host work between
    them, gives the same results as direct launches."""
    import py_propag as pp

    w, h = 140, 110
    land = _terrain(90, w, h)
    cases = [_case(s, w, h) for s in range(90, 95)]
    with pp.Context() as ctx:
        tid = ctx.add_terrain(_geo(pp, w, h), land["fuel"], land["slope"], land["aspect"])
        out = {}
        for mode in (0, 2):
            ids = [ctx.admit(tid, c["wx"], c["ign"], c["max_time"]) for c in cases]
            launches = 0
            while any(ctx.status(i)[0] == pp.RUNNING for i in ids):
                ctx.run(max_sweeps=7, tuning={"mode": mode})
                launches += 1
                assert launches < 2000
            out[mode] = [ctx.result(i) for i in ids]
            for i in ids:
                ctx.release(i)
        assert launches > 3
        for a, b in zip(out[0], out[2]):
            _same(a, type("R", (), {"arrival_times": b[0], "refs_x": b[1], "refs_y": b[2],
                                    "boundary_change": b[3]})())
        c = cases[0]
        _same(out[2][0], _alone(pp, land, w, h, c["wx"], c["ign"], c["max_time"]))


def test_autotuner_picks_a_tuning_without_changing_results():
    import py_propag as pp
    from py_propag.autotune import Autotuner

    w, h = 128, 96
    land = _terrain(91, w, h)
    cases = [_case(s, w, h) for s in range(100, 106)]
    with pp.Context() as ctx:
        tid = ctx.add_terrain(_geo(pp, w, h), land["fuel"], land["slope"], land["aspect"])
        tuner = Autotuner(ctx)
        best = tuner.tune(tid, cases, w, h)
        assert set(best) == {"tile_w", "tile_h", "waves_per_group", "workgroups", "mode"}
        assert {(t["tile_w"], t["tile_h"]) for t, _ in tuner.tables[tuner.key(w, h, len(cases))]} == \
            {(8, 8), (16, 4), (32, 2)}
        assert len(tuner.tables[tuner.key(w, h, len(cases))]) == len(tuner.space)
        assert tuner.tune(tid, cases, w, h) is best   # cached per workload class
        ids = [ctx.admit(tid, c["wx"], c["ign"], c["max_time"]) for c in cases]
        ctx.run(tuning=best)
        for sid, c in zip(ids, cases):
            _same(ctx.result(sid), _alone(pp, land, w, h, c["wx"], c["ign"], c["max_time"]))


def test_failed_launch_fails_its_running_scenarios_only():
    """A launch that returns an error (here the grid barrier told to give up
    at its first wait: spin_limit_log2 = 32, so every kernel still exits
    normally) fails the scenarios it was running: they report FAILED through
    take_finished, have no result and are released; a scenario finished
    before keeps its result, and later admissions run normally.  The plain
    entry point recovers the same way on its next call."""
    import py_propag as pp
    from py_propag.serve import Request, Server

    w, h = 160, 120
    land = _terrain(12, w, h)
    a, b, c, d = (_case(s, w, h) for s in (110, 111, 112, 113))
    stop = {"spin_limit_log2": 32}
    with pp.Context() as ctx:
        tid = ctx.add_terrain(_geo(pp, w, h), land["fuel"], land["slope"], land["aspect"])
        ia = ctx.admit(tid, a["wx"], a["ign"], a["max_time"])
        ctx.run()
        assert ctx.status(ia)[0] == pp.DONE
        assert ctx.take_finished() == [ia]
        ib = ctx.admit(tid, b["wx"], b["ign"], b["max_time"])
        ic = ctx.admit(tid, c["wx"], c["ign"], c["max_time"])
        with pytest.raises(pp.PropagCudaError, match="PG_ERR_TIMEOUT"):
            ctx.run(tuning=stop)
        assert ctx.status(ib)[0] == pp.FAILED and ctx.status(ic)[0] == pp.FAILED
        assert sorted(ctx.take_finished()) == sorted([ib, ic])
        with pytest.raises(pp.PropagValidationError, match="failed"):
            ctx.result(ib)
        ctx.release(ib)
        ctx.release(ic)
        _same(ctx.result(ia), _alone(pp, land, w, h, a["wx"], a["ign"], a["max_time"]))
        idd = ctx.admit(tid, d["wx"], d["ign"], d["max_time"])
        ctx.run()
        assert ctx.status(idd)[0] == pp.DONE
        _same(ctx.result(idd), _alone(pp, land, w, h, d["wx"], d["ign"], d["max_time"]))
        ctx.release(idd)
        ctx.release(ia)
        # the request loop delivers failed requests as FAILED, without on_done
        got = []
        server = Server(ctx, on_done=lambda req, sid: got.append(req.tag), tuning=stop)
        done = server.play([Request(tid, b["wx"], b["ign"], b["max_time"], 0.0, tag="b")])
        assert [r.state for r in done] == [pp.FAILED] and got == [] and server.failed_launches == 1
        assert ctx.take_finished() == []
    # the one-scenario entry point: the failing call raises, the next one is exact
    settings = pp.Settings(_geo(pp, w, h), b["max_time"])
    with pytest.raises(pp.PropagCudaError, match="PG_ERR_TIMEOUT"):
        pp.propagate(settings, _full_terrain(land, b["wx"], w * h), ignitions=b["ign"], tuning=stop)
    again = pp.propagate(settings, _full_terrain(land, b["wx"], w * h), ignitions=b["ign"])
    ref = _alone(pp, land, w, h, b["wx"], b["ign"], b["max_time"])
    assert np.array_equal(again.arrival_times.view(np.uint32), ref.arrival_times.view(np.uint32))
