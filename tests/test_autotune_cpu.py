"""The autotuner's choice rule, on the CPU with a stand-in context whose
launch time depends only on the tuning: a candidate replaces the default
tuning (8x8 tiles, 16 waves, every workgroup, direct launch) only when it is
faster by more than the margin on the sample."""
import pytest

from py_propag.autotune import Autotuner


class _FakeContext:
    """admit / run / status / release / set_tile of py_propag.Context; run()
    advances a fake clock by the cost its tuning is given."""

    def __init__(self, cost):
        self.cost = cost
        self.now = 0.0
        self.tile = (8, 8)
        self.held = set()
        self.next_id = 0

    def clock(self):
        return self.now

    def set_tile(self, tw, th):
        assert not self.held
        self.tile = (tw, th)

    def admit(self, tid, wx, ign, max_time):
        self.next_id += 1
        self.held.add(self.next_id)
        return self.next_id

    def run(self, tuning=None):
        t = tuning or {}
        waves = t.get("waves_per_group", 16)
        full = 256 if waves == 16 else 512
        groups = t.get("workgroups") or full
        key = (self.tile, waves, groups == full, t.get("mode", 0))
        self.now += self.cost(key)
        return len(self.held), {"workgroups": groups}

    def status(self, sid):
        return ("done", 100)

    def release(self, sid):
        self.held.discard(sid)


def _tune(cost):
    ctx = _FakeContext(cost)
    tuner = Autotuner(ctx, clock=ctx.clock)
    cases = [{"wx": None, "ign": None, "max_time": 1.0}] * 4
    return tuner.tune(0, cases, 1024, 1024)


DEFAULT = ((8, 8), 16, True, 0)
OTHER = ((16, 4), 16, True, 2)


@pytest.mark.parametrize("lead, picked", [(0.01, DEFAULT), (0.05, OTHER)])
def test_default_kept_unless_a_candidate_leads_by_the_margin(lead, picked):
    def cost(key):
        if key == DEFAULT:
            return 1.0
        if key == OTHER:
            return 1.0 - lead
        return 2.0

    best = _tune(cost)
    got = ((best["tile_w"], best["tile_h"]), best["waves_per_group"], best["workgroups"] == 256, best["mode"])
    assert got == picked


def test_fastest_candidate_when_the_default_is_slow():
    def cost(key):
        return {((32, 2), 8, True, 0): 0.5, DEFAULT: 1.0}.get(key, 0.9)

    best = _tune(cost)
    assert (best["tile_w"], best["tile_h"], best["waves_per_group"], best["mode"]) == (32, 2, 8, 0)
