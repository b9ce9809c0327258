"""Host logic of the serving loop and the autotuner, on a simulated context.

FakeContext stands in for py_propag.Context with the same calls: a scenario
needs a fixed number of sweeps (its weather's "work"), each run advances
every admitted, unfinished scenario by at most max_sweeps, admission fails
(None) when the simulated memory budget is full, finished ids are drained by
take_finished.  The GPU-side behaviour of the real context is covered by
tests/test_gpu_batch.py.
"""
import itertools

import numpy as np
import pytest

from py_propag import synth
from py_propag.autotune import Autotuner, default_space
from py_propag.serve import Request, Server


class FakeContext:
    def __init__(self, slots=3, sweep_ms=0.0, clock=None):
        self.slots = slots
        self.left = {}        # id -> sweeps left
        self.done = []
        self.state = {}
        self.sweeps = {}
        self.ids = itertools.count()
        self.launches = []    # (max_sweeps, ids running)
        self.released = []
        self.clock = clock

    def admit(self, terrain, weather, ignitions, max_time, max_iterations=0):
        if len([i for i, s in self.state.items() if s is not None]) >= self.slots:
            return None
        sid = next(self.ids)
        self.left[sid] = int(weather["work"])
        self.state[sid] = 0
        self.sweeps[sid] = 0
        return sid

    def run(self, max_sweeps=0, tuning=None):
        running = [i for i, s in self.state.items() if s == 0]
        self.launches.append((max_sweeps, tuple(running)))
        n = max(self.left[i] for i in running) if running else 0
        if max_sweeps:
            n = min(n, max_sweeps)
        fin = 0
        for i in running:
            step = min(n, self.left[i])
            self.left[i] -= step
            self.sweeps[i] += step
            if self.left[i] == 0:
                self.state[i] = 1
                self.done.append(i)
                fin += 1
        if self.clock is not None:
            self.clock.t += 0.001 * n
        return fin, {"relax_ms": float(n), "iterations": n, "workgroups": 256}

    def take_finished(self):
        out, self.done = self.done, []
        return out

    def status(self, sid):
        return self.state[sid], self.sweeps[sid]

    def release(self, sid):
        self.state[sid] = None
        self.released.append(sid)


class Clock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t

    def sleep(self, dt):
        self.t += dt


def _req(work, arrival=0.0, tag=None):
    return Request(0, {"work": work}, [(1, 1, 0.0)], 3600.0, arrival, tag=tag)


def test_server_delivers_every_request_once_and_releases_after_delivery():
    ctx = FakeContext(slots=2)
    seen = []

    def on_done(req, sid):
        assert ctx.state[sid] == 1          # still resident when delivered
        seen.append(req.tag)

    srv = Server(ctx, on_done=on_done, min_sweeps=4, max_sweeps=4, clock=Clock())
    for i, w in enumerate([10, 3, 7, 12, 1]):
        srv.submit(_req(w, tag=i))
    while not srv.idle():
        srv.step()
    assert sorted(seen) == list(range(5))
    assert sorted(ctx.released) == sorted(ctx.state)
    # never more than the budget's two scenarios in one launch
    assert all(len(r) <= 2 for _, r in ctx.launches)
    assert srv.busy_refusals > 0


def test_server_admits_arrivals_into_the_next_launch():
    ctx = FakeContext(slots=8)
    srv = Server(ctx, min_sweeps=5, max_sweeps=5, clock=Clock())
    srv.submit(_req(20, tag="a"))
    srv.step()                              # a runs alone
    srv.submit(_req(6, tag="b"))            # arrives while a is unfinished
    srv.step()
    assert ctx.launches[0][1] == (0,)
    assert ctx.launches[1][1] == (0, 1)     # b joined a's next launch
    done = []
    while not srv.idle():
        done += [r.tag for r in srv.step()]
    assert set(done) == {"a", "b"}


def test_server_play_serves_in_arrival_order_with_latencies():
    clock = Clock()
    ctx = FakeContext(slots=4, clock=clock)
    srv = Server(ctx, min_sweeps=8, max_sweeps=8, clock=clock, sleep=clock.sleep)
    reqs = [_req(w, arrival=0.004 * i, tag=i) for i, w in enumerate([30, 5, 5, 40, 2, 9])]
    out = srv.play(reqs, start=0.0)
    assert sorted(r.tag for r in out) == list(range(6))
    for r in out:
        assert r.finished >= r.admitted >= r.arrival
        assert r.state == 1 and r.sweeps == r.weather["work"]


def test_server_adapts_launch_length_within_bounds():
    clock = Clock()
    ctx = FakeContext(slots=4, clock=clock)
    srv = Server(ctx, target_ms=20.0, min_sweeps=4, max_sweeps=64, clock=clock)
    srv.submit(_req(10_000))
    lengths = []
    for _ in range(12):
        srv.step()
        lengths.append(ctx.launches[-1][0])
    assert all(4 <= n <= 64 for n in lengths)
    assert lengths[-1] > lengths[0]         # 1 ms per sweep simulated: launches grow toward 20 ms


def test_server_raises_when_a_request_can_never_fit():
    ctx = FakeContext(slots=0)
    srv = Server(ctx, clock=Clock())
    srv.submit(_req(3))
    with pytest.raises(RuntimeError):
        srv.step()


class TunerContext(FakeContext):
    """Trials cost what the tuning says: 16x4 tiles, mode 2 with 8 waves is
    fastest.  Like the real context it takes a tile shape only while it
    holds no scenario, and a tuning must name the shape it holds."""

    def __init__(self):
        super().__init__(slots=1000)
        self.clock = Clock()
        self.tile = (8, 8)

    def set_tile(self, tw, th):
        if any(s is not None for s in self.state.values()):
            raise RuntimeError("set_tile with scenarios admitted")
        self.tile = (tw, th)

    def run(self, max_sweeps=0, tuning=None):
        t = dict(tuning or {})
        if t.get("tile_w") and (t["tile_w"], t["tile_h"]) != self.tile:
            raise RuntimeError("tuning names another tile shape than the context's")
        fin, st = super().run(max_sweeps, tuning)
        w = t.get("waves_per_group", 16) or 16
        cost = {16: 3.0, 8: 1.0, 4: 2.0}[w] + (0.0 if t.get("mode", 0) == 2 else 0.5)
        cost += {(8, 8): 0.2, (16, 4): 0.0, (32, 2): 0.4}[self.tile]
        cost *= 1.0 if not t.get("workgroups") or t["workgroups"] >= 4096 // w else 1.7
        self.clock.t += cost
        st["workgroups"] = t.get("workgroups") or 4096 // w
        return fin, st


def test_autotuner_picks_the_fastest_and_caches_per_class():
    ctx = TunerContext()
    cases = [{"wx": {"work": 5 + i}, "ign": [(1, 1, 0.0)], "max_time": 3600.0} for i in range(6)]
    tuner = Autotuner(ctx, clock=ctx.clock)
    best = tuner.tune(0, cases, 1000, 1000, count=128)
    assert best == {"tile_w": 16, "tile_h": 4, "waves_per_group": 8, "mode": 2, "workgroups": 512}
    assert ctx.tile == (16, 4)      # the context is left in the best shape
    table = tuner.tables[tuner.key(1000, 1000, 128)]
    assert len(table) == len(default_space())
    assert [dt for _, dt in table] == sorted(dt for _, dt in table)
    assert tuner.key(1000, 1000, 128) == (1024, 1024, 128)
    n = len(ctx.launches)
    assert tuner.tune(0, cases, 1000, 1000, count=100) is best    # same class: cached
    assert len(ctx.launches) == n


def test_autotuner_rejects_a_tuning_that_changes_sweep_counts():
    class Drifting(TunerContext):
        def run(self, max_sweeps=0, tuning=None):
            if tuning and tuning.get("waves_per_group") == 4:
                for i in [i for i, s in self.state.items() if s == 0]:
                    self.left[i] += 1
            return super().run(max_sweeps, tuning)

    ctx = Drifting()
    cases = [{"wx": {"work": 5}, "ign": [(1, 1, 0.0)], "max_time": 3600.0}] * 3
    with pytest.raises(RuntimeError, match="sweep counts"):
        Autotuner(ctx, clock=ctx.clock).tune(0, cases, 64, 64)


def test_workloads_are_seeded_and_shaped_as_documented():
    a = synth.serving_stream(seed=4, rate=50.0, seconds=10.0)
    b = synth.serving_stream(seed=4, rate=50.0, seconds=10.0)
    assert a["sides"] == b["sides"] and len(a["requests"]) == len(b["requests"])
    assert all(512 <= s <= 4096 for s in a["sides"])
    arr = np.array([r["arrival"] for r in a["requests"]])
    assert np.all(np.diff(arr) > 0) and arr[-1] < 10.0
    assert 350 < len(arr) < 650                        # Poisson(500)
    for r in a["requests"]:
        side = a["sides"][r["terrain"]]
        assert 3600.0 <= r["max_time"] <= 36000.0 and 1 <= len(r["ign"]) <= 4
        assert all(side // 8 <= x < side - side // 8 and side // 8 <= y < side - side // 8 for x, y, _ in r["ign"])
        assert 2.0 <= r["wx"]["wind_speed"] <= 12.0 and 0.03 <= r["wx"]["d1hr"] <= 0.12
    r = np.random.Generator(np.random.PCG64(3))
    c = synth.scenario_weather(r, 3, 64, 64, 1000.0)
    assert len(c["ign"]) == 3 and c["max_time"] == 1000.0
