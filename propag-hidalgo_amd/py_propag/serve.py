"""Request loop over a batching context: continuous batching of fire scenarios.

Requests (a scenario on a resident terrain: weather, ignitions, max_time)
wait in a FIFO.  Each turn of the loop admits requests from the head while
the context's memory budget has room for them, runs every unfinished
scenario together in one shared persistent launch cut after a bounded number
of sweeps, and hands finished scenarios to the caller before releasing
their device memory.  A request that arrives while others are on the GPU
therefore joins the next launch, at most one launch after its arrival; its
result does not depend on what else shares the launches.

The launch length adapts so that one launch lasts about ``target_ms``: short
enough that arrivals are admitted promptly, long enough that the per-launch
host overhead stays small.

The reference has no serving form; its device layer is one scenario per
call (crates/propag/src/propag.rs:375-573).
"""
from __future__ import annotations

import collections
import time
from dataclasses import dataclass, field

from . import FAILED, Context, PropagCudaError


@dataclass
class Request:
    terrain: int                # terrain id in the context
    weather: dict               # d1hr, d10hr, d100hr, herb, wood, wind_speed, wind_azimuth
    ignitions: list             # (x, y, t) cells
    max_time: float
    arrival: float = 0.0        # seconds on the loop's clock
    tag: object = None
    admitted: float = field(default=None)
    finished: float = field(default=None)
    sweeps: int = 0
    state: int = 0


class Server:
    """Continuous-batching request loop.

    ``on_done(request, scenario_id)`` is called for each finished request
    while its results are still resident (fetch them there, e.g. with
    ``Context.time_into``); the scenario is released afterwards."""

    def __init__(self, ctx: Context, on_done=None, target_ms: float = 2.0, min_sweeps: int = 16,
                 max_sweeps: int = 4096, tuning=None, clock=time.perf_counter, sleep=time.sleep):
        self.ctx = ctx
        self.on_done = on_done
        self.target_ms = target_ms
        self.min_sweeps, self.max_sweeps = min_sweeps, max_sweeps
        self.sweeps = min_sweeps
        self.tuning = tuning
        self.clock = clock
        self.sleep = sleep
        self.queue = collections.deque()
        self.running = {}            # scenario id -> Request
        self.launches = 0
        self.launch_ms = 0.0         # host wall time inside launches
        self.kernel_ms = 0.0         # relaxation kernel time (HIP events)
        self.sweeps_run = 0
        self.busy_refusals = 0
        self.failed_launches = 0
        self.last_error = None

    def submit(self, req: Request):
        self.queue.append(req)

    def idle(self) -> bool:
        return not self.queue and not self.running

    def _admit(self, now):
        while self.queue:
            req = self.queue[0]
            sid = self.ctx.admit(req.terrain, req.weather, req.ignitions, req.max_time)
            if sid is None:          # memory budget full: wait for releases
                self.busy_refusals += 1
                if not self.running:
                    raise RuntimeError("a request does not fit the memory budget of an empty context")
                return
            self.queue.popleft()
            req.admitted = now
            self.running[sid] = req

    def step(self):
        """Admits what fits, runs one launch, delivers what finished.
        Returns the requests finished in this step.  A launch that fails
        fails the requests it was running (state FAILED, ``on_done`` not
        called); the loop goes on with the others."""
        self._admit(self.clock())
        if not self.running:
            return []
        t0 = self.clock()
        try:
            _, stats = self.ctx.run(max_sweeps=self.sweeps, tuning=self.tuning)
        except PropagCudaError as e:
            self.failed_launches += 1
            self.last_error = str(e)
            stats = None
        ms = (self.clock() - t0) * 1e3
        self.launches += 1
        self.launch_ms += ms
        if stats is not None:
            self.kernel_ms += stats["relax_ms"]
            self.sweeps_run += stats["iterations"]
            if ms > 0:
                want = int(self.sweeps * self.target_ms / ms)
                self.sweeps = max(self.min_sweeps, min(self.max_sweeps, (self.sweeps + want) // 2))
        done = []
        for sid in self.ctx.take_finished():
            req = self.running.pop(sid, None)
            if req is None:
                continue
            req.state, req.sweeps = self.ctx.status(sid)
            if self.on_done is not None and req.state != FAILED:
                self.on_done(req, sid)
            req.finished = self.clock()
            self.ctx.release(sid)
            done.append(req)
        return done

    def play(self, requests, start=None):
        """Serves ``requests`` (sorted by arrival, seconds from ``start``) in
        real time: each is submitted once the clock passes its arrival.
        Returns the requests in order of completion."""
        start = self.clock() if start is None else start
        pending = collections.deque(requests)
        out = []
        while pending or not self.idle():
            now = self.clock() - start
            while pending and pending[0].arrival <= now:
                req = pending.popleft()
                req.arrival += start
                self.submit(req)
            if self.idle():
                wait = pending[0].arrival - (self.clock() - start)
                if wait > 0:
                    self.sleep(min(wait, 0.001))
                continue
            out.extend(self.step())
        return out
