"""Autotuner of the batched propagation launch.

It picks the tile shape (8x8, 16x4 or 32x2 cells, each one cell per lane of
a 64-lane wave), the persistent kernel's super-grid — workgroups per launch
and waves per workgroup, i.e. how many tiles are in flight per CU and how
the per-sweep tile queue is split — and the launch form (direct cooperative
launch or the cached hipGraph), by timing candidates on a sample of the
ensemble to be run, each ``repeats`` times (the median counts).  A
candidate replaces the default tuning (8x8 tiles, 16 waves per workgroup,
every workgroup that fits, direct launch) only when it is faster by more
than ``margin`` on the sample: candidates within a percent of each other on
a 32-scenario sample have been seen to order differently on the full
ensemble (a 16x4 pick 0.7% ahead on the sample ran 4.5% behind on 256
2048x2048 scenarios).  The sweep
schedule is exact Jacobi, so results never depend on the tuning and any
sample is a fair trial; the tuner checks that every trial finishes every
scenario in the same number of sweeps.

One shape is fixed rather than tuned: one sweep per grid barrier (more
local sweeps between barriers would change the arrival times the exact
schedule defines).

The reference sizes its super-grid from the occupancy of its cooperative
kernel (crates/propag/src/propag.rs:403-418) and has no tuner.
"""
from __future__ import annotations

import itertools
import time

import numpy as np

from . import PropagCudaError


TILE_SHAPES = ((8, 8), (16, 4), (32, 2))


def default_space():
    """Candidate tunings: tile shape x waves per workgroup x fraction of the
    workgroups that fit x launch form."""
    out = []
    for (tw, th), waves, frac, mode in itertools.product(TILE_SHAPES, (16, 8, 4), (1.0, 0.5), (0, 2)):
        out.append({"tile_w": tw, "tile_h": th, "waves_per_group": waves, "group_fraction": frac, "mode": mode})
    return out


class Autotuner:
    """Times candidate tunings of ``Context.run`` on sample ensembles and
    remembers the best per workload key."""

    DEFAULT = {"tile_w": 8, "tile_h": 8, "waves_per_group": 16, "mode": 0}

    def __init__(self, ctx, space=None, repeats: int = 2, clock=time.perf_counter, margin: float = 0.02):
        self.ctx = ctx
        self.clock = clock
        self.space = space if space is not None else default_space()
        self.repeats = repeats
        self.margin = margin
        self.best = {}     # key -> tuning dict
        self.tables = {}   # key -> [(tuning, seconds)]
        self.rejected = []  # (tuning, error) of candidates whose launch failed

    @staticmethod
    def key(width: int, height: int, count: int):
        """Workload class: raster size and ensemble size, each rounded to a
        power of two."""
        def p2(v):
            return 1 << max(0, (int(v) - 1).bit_length())
        return (p2(width), p2(height), p2(count))

    def _trial(self, terrain_id, cases, tuning):
        """Runs the sample once with ``tuning`` (in its tile shape, set on
        the context while it holds no scenario); the admitted scenarios are
        released whatever happens.  A launch error propagates (the
        scenarios are then FAILED and released here)."""
        ids = []
        if tuning is not None and tuning.get("tile_w"):
            self.ctx.set_tile(tuning["tile_w"], tuning["tile_h"])
        try:
            for c in cases:
                i = self.ctx.admit(terrain_id, c["wx"], c["ign"], c["max_time"])
                if i is None:
                    raise RuntimeError("autotune sample does not fit the memory budget")
                ids.append(i)
            t0 = self.clock()
            fin, stats = self.ctx.run(tuning=tuning)
            dt = self.clock() - t0
            sweeps = [self.ctx.status(i)[1] for i in ids]
        finally:
            for i in ids:
                self.ctx.release(i)
        if fin != len(ids):
            raise RuntimeError("autotune trial left scenarios unfinished")
        return dt, sweeps, stats

    def tune(self, terrain_id: int, cases, width: int, height: int, count: int | None = None):
        """Runs every candidate on ``cases`` (a sample of the ensemble) and
        returns the fastest tuning (a dict for Context.run).  The context
        must hold no scenario; it is left in the best tuning's tile shape."""
        k = self.key(width, height, count if count is not None else len(cases))
        if k in self.best:
            self._set_tile(self.best[k])
            return self.best[k]
        # reference sweep counts from an untuned run; the device's full
        # workgroup count for each waves-per-group choice is what a launch
        # with workgroups = 0 reports
        _, ref_sweeps, _ = self._trial(terrain_id, cases, None)
        full = {}
        table = []
        for cand in sorted(self.space, key=lambda c: -c["group_fraction"]):
            w = cand["waves_per_group"]
            tile = (cand.get("tile_w", 8), cand.get("tile_h", 8))
            if cand["group_fraction"] < 1.0 and (w, tile) not in full:
                continue  # needs the full count of this launch shape first
            groups = 0 if cand["group_fraction"] >= 1.0 else max(1, int(full[(w, tile)] * cand["group_fraction"]))
            tuning = {"tile_w": tile[0], "tile_h": tile[1], "waves_per_group": w, "mode": cand["mode"],
                      "workgroups": groups}
            times = []
            try:
                for _ in range(self.repeats):
                    dt, sweeps, st = self._trial(terrain_id, cases, tuning)
                    if sweeps != ref_sweeps:
                        raise RuntimeError(f"tuning {tuning} changed the sweep counts")
                    times.append(dt)
            except PropagCudaError as e:
                # a launch error (e.g. a barrier past its spin limit) rejects
                # the candidate; its sample was released
                self.rejected.append((tuning, str(e)))
                continue
            full.setdefault((w, tile), st["workgroups"])
            table.append(({**tuning, "workgroups": st["workgroups"]}, float(np.median(times))))
        table.sort(key=lambda e: e[1])
        self.tables[k] = table
        pick = table[0]
        full_default = [e for e in table
                        if all(e[0].get(f) == v for f, v in self.DEFAULT.items()) and e[0]["workgroups"] == full.get(
                            (self.DEFAULT["waves_per_group"], (self.DEFAULT["tile_w"], self.DEFAULT["tile_h"])))]
        if full_default and pick[1] > full_default[0][1] * (1.0 - self.margin):
            pick = full_default[0]
        self.best[k] = pick[0]
        self._set_tile(pick[0])
        return pick[0]

    def _set_tile(self, tuning):
        self.ctx.set_tile(tuning.get("tile_w", 8), tuning.get("tile_h", 8))
