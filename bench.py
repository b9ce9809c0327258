"""Benchmark of the propagation hot path on one MI355X.

Headline (value): scenarios/s of 1024x1024, 10-hour scenarios run batched.
A step is one ensemble of --batch scenarios on the single-scenario
landscape (1024x1024 cells of 5 m, 20 Gaussian hills, 200 Voronoi fuel
patches, resident in HBM before the timed region), each with its own
weather (1-h moisture U(0.03,0.12), 10-h/100-h +0.01/+0.02, herb U(0.3,1.5),
wood U(0.6,1.5), wind U(2,12) m/s toward U(0,360) deg) and 1-4 point
ignitions, max_time 36000 s: admission (burn pass, boundaries, initial
state on the device) and one shared persistent launch to convergence
through the batching context.

Beside it, on the same line:
  * plain: the same scenarios one per call (pg_burn + pg_propagate, the
    drop-in form; the caller's reset of its buffers between calls is not
    timed), scenarios/s and p50/p99 latency per scenario, plus the
    config-1 scenario itself (smooth wind field, ignition at (500,500));
    the first scenarios of the ensemble are checked bit-identical between
    the two forms;
  * roofline of the dominant kernel (the relaxation kernel), with
    algorithmic bytes = 24 B per cell relaxation + 8 B per committed
    improvement, timed with HIP events around that kernel on its stream;
  * the Rothermel burn pass alone on the 1024x1024 i.i.d. field (Mcells/s,
    49 B/cell);
  * the CPU baseline: the oracle (a serial C transcription) on the config-1
    scenario, one thread, and its burn pass on the host cores.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--batch B]
                       [--no-cpu-baseline]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "propag-hidalgo_amd"))

METRIC = ("scenarios/sec + p50/p99 latency (1024² 10 h) vs plain launch; "
          "Rothermel Mcells/s")
HBM_PEAK_GBS = 8000.0
FIELDS = ("d1hr", "d10hr", "d100hr", "herb", "wood", "wind_speed", "wind_azimuth", "slope",
          "aspect")


TRAFFIC_PROFILE = "profiles/propagate_hbm_traffic.md"


def roofline(relaxations, improvements, kernel_ms, workload=None):
    """HBM roofline of a relaxation kernel: algorithmic bytes (24 B per cell
    relaxation + 8 B per committed improvement) over its HIP-event time, with
    the HBM bytes per launch that rocprofv3 measured for `workload`."""
    alg = 24.0 * relaxations + 8.0 * improvements
    gbs = alg / (kernel_ms * 1e-3) / 1e9 if kernel_ms > 0 else 0.0
    out = {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
           "algorithmic_bytes": alg, "kernel_ms": kernel_ms, "kernel": "propagate_persistent"}
    if workload is not None:
        t = measured_traffic(workload)
        out.update({"traffic": t, "traffic_source": f"{TRAFFIC_PROFILE} ({workload})",
                    "traffic_over_algorithmic": t / alg if t and alg else None})
    return out


def measured_traffic(workload):
    """HBM bytes per launch of the relaxation kernel on `workload`, from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (counters cannot be
    read from inside this process); None if it is absent."""
    key = f"traffic_bytes_per_launch {workload}:"
    try:
        with open(os.path.join(ROOT, TRAFFIC_PROFILE)) as f:
            for line in f:
                if line.startswith(key):
                    return int(line.split(":")[1])
    except OSError:
        pass
    return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def ensemble_cases(seed, count, W, H, max_time):
    """Per-scenario weather and ignitions of the ensemble (seeded)."""
    r = np.random.Generator(np.random.PCG64(seed))
    cases = []
    for _ in range(count):
        d1 = np.float32(r.uniform(0.03, 0.12))
        wx = {"d1hr": d1, "d10hr": np.float32(d1 + np.float32(0.01)),
              "d100hr": np.float32(d1 + np.float32(0.02)), "herb": np.float32(r.uniform(0.3, 1.5)),
              "wood": np.float32(r.uniform(0.6, 1.5)), "wind_speed": np.float32(r.uniform(2.0, 12.0)),
              "wind_azimuth": np.float32(np.radians(r.uniform(0.0, 360.0)))}
        k = int(r.integers(1, 5))
        cells = set()
        while len(cells) < k:
            cells.add((int(r.integers(W // 8, W - W // 8)), int(r.integers(H // 8, H - H // 8))))
        cases.append({"wx": wx, "ign": [(x, y, 0.0) for x, y in sorted(cells)], "max_time": max_time})
    return cases


def autotuned(args, ctx, tid, cases, W, H, tuning):
    """The ensemble's launch tuning: the autotuner's pick on a sample of
    --tune-sample scenarios (or the fixed bench tuning with --no-autotune /
    --profile-phases)."""
    if args.tuning or args.no_autotune or args.profile_phases or args.mode == 1:
        if tuning.get("tile_w"):
            ctx.set_tile(tuning["tile_w"], tuning["tile_h"])
        if args.tuning:
            w, g, m = (int(v) for v in args.tuning.split(","))
            return {**tuning, "waves_per_group": w, "workgroups": g, "mode": m}, None
        return dict(tuning), None
    from py_propag.autotune import Autotuner

    t0 = time.perf_counter()
    tuner = Autotuner(ctx)
    best = tuner.tune(tid, cases[:args.tune_sample], W, H, count=len(cases))
    table = tuner.tables[tuner.key(W, H, len(cases))]
    log(f"autotune ({len(table)} candidates on {min(len(cases), args.tune_sample)} scenarios, "
        f"{time.perf_counter() - t0:.1f}s): best {best}; "
        + ", ".join(f"{t['tile_w']}x{t['tile_h']}/{t['waves_per_group']}w/{t['workgroups']}g/m{t['mode']} "
                    f"{dt * 1e3:.1f}ms" for t, dt in table))
    return dict(best), {"best": best, "sample": min(len(cases), args.tune_sample),
                        "table": [{"tuning": t, "ms": dt * 1e3} for t, dt in table]}


def device_columns(torch, dev, terrain, n, fields):
    return [torch.from_numpy(np.ascontiguousarray(np.broadcast_to(np.asarray(terrain[k], np.float32), (n,))))
            .to(dev) for k in fields]


def plain_run(pp, torch, dev, stream, settings, fuel, cols, init, tuning):
    """One scenario through the drop-in form: pg_burn + pg_propagate on
    resident columns; returns (wall ms, stats, time tensor)."""
    n = fuel.numel()
    fire = [torch.zeros(n, dtype=torch.float32, device=dev) for _ in range(3)]
    t, rx, ry = (x.clone() for x in init)
    bits = torch.zeros((n + 31) // 32, dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    s0 = time.perf_counter()
    pp.burn_device(fuel, *cols, *fire, stream=stream)
    st = pp.propagate_device(settings, *fire, t, rx, ry, bits, tuning=tuning, stream=stream)
    return (time.perf_counter() - s0) * 1e3, st, t


def large_single(args, pp, synth, torch, tuning, stream):
    """The 4096x4096 12-h scenario (16 point ignitions + a 200-cell line),
    plain form, resident inputs."""
    dev = torch.device("cuda:0")
    t0 = time.time()
    sc = synth.large_scenario(seed=2, side=4096)
    W, H = sc["width"], sc["height"]
    n = W * H
    log(f"large: generated 4096^2 scenario in {time.time() - t0:.1f}s")
    g = pp.GeoReference(W, H, "+proj=utm +zone=30", [0.0, sc["dx"], 0.0, 0.0, 0.0, sc["dy"]])
    settings = pp.Settings(g, float(sc["max_time"]))
    fuel = torch.from_numpy(sc["terrain"]["fuel"]).to(dev)
    cols = device_columns(torch, dev, sc["terrain"], n, FIELDS)
    init = (torch.from_numpy(sc["time"]).to(dev), torch.from_numpy(sc["refs_x"].view(np.int16)).to(dev),
            torch.from_numpy(sc["refs_y"].view(np.int16)).to(dev))
    plain_run(pp, torch, dev, stream, settings, fuel, cols, init, tuning)
    runs = [plain_run(pp, torch, dev, stream, settings, fuel, cols, init, tuning) for _ in range(3)]
    ms = [r[0] for r in runs]
    st, t = runs[-1][1], runs[-1][2]
    burned = int((t < pp.FLT_MAX).sum().item())
    same = all(torch.equal(r[2], t) for r in runs)
    out = {"workload": "large_scenario_4096 (16 points + 200-cell line, 12 h)", "ms": float(np.mean(ms)),
           "ms_min": float(np.min(ms)), "relax_ms": float(np.mean([r[1]["relax_ms"] for r in runs])),
           "sweeps": st["iterations"], "relaxations": st["relaxations"], "improvements": st["improvements"],
           "tiles_visited": st["tiles_visited"], "tile_marks": st["tile_marks"], "idle_tiles": st["idle_tiles"],
           "dda_steps": st["dda_steps"], "burned_cells": burned,
           "repeatable": bool(same), "ignitions": len(sc["ignitions"])}
    if tuning.get("profile"):
        out.update({k: st[k] for k in ("crit_ms", "phase_ms", "phase_total_ms", "barrier_ms", "idle_far_tiles")})
    out["roofline"] = roofline(out["relaxations"], out["improvements"], out["relax_ms"], "large_4096")
    log(f"large 4096^2: {out['ms']:.2f} ms per scenario (relax {out['relax_ms']:.2f} ms), {out['sweeps']} sweeps, "
        f"{burned} burned, {out['tiles_visited']} tile visits; repeatable {same}")
    return out, (sc, t.cpu().numpy())


def ensemble_2048_section(args, pp, synth, torch, tuning, stream):
    """256 uneven scenarios on one 2048x2048 landscape in one shared launch,
    beside a sample of them one per call."""
    dev = torch.device("cuda:0")
    t0 = time.time()
    wl = synth.ensemble_2048(seed=3, count=args.batch2048)
    side = wl["side"]
    n = side * side
    land = wl["land"]
    log(f"ensemble 2048: generated in {time.time() - t0:.1f}s")
    g = pp.GeoReference(side, side, "+proj=utm +zone=30", [0.0, 5.0, 0.0, 0.0, 0.0, 5.0])
    ctx = pp.Context()
    tid = ctx.add_terrain(g, land["fuel"], land["slope"], land["aspect"])
    btuning, tune = autotuned(args, ctx, tid, wl["cases"], side, side, tuning)

    def ensemble():
        ids = [ctx.admit(tid, c["wx"], c["ign"], c["max_time"]) for c in wl["cases"]]
        if any(i is None for i in ids):
            raise RuntimeError("ensemble does not fit the memory budget")
        fin, stats = ctx.run(tuning=btuning)
        if fin != len(ids):
            raise RuntimeError(f"{len(ids) - fin} scenarios did not finish")
        return ids, stats

    ids, _ = ensemble()
    for i in ids:
        ctx.release(i)
    torch.cuda.synchronize()
    ms, stats = [], None
    for k in range(2):
        s0 = time.perf_counter()
        ids, stats = ensemble()
        ms.append((time.perf_counter() - s0) * 1e3)
        if k == 1:
            kept = [ctx.result(ids[i])[0].ravel().copy() for i in range(2)]
        for i in ids:
            ctx.release(i)
    ctx.close()
    value = len(wl["cases"]) / (np.mean(ms) / 1e3)
    # a sample one per call, drop-in form
    fuel = torch.from_numpy(land["fuel"]).to(dev)
    slope = torch.from_numpy(land["slope"]).to(dev)
    aspect = torch.from_numpy(land["aspect"]).to(dev)
    plain_ms, parity = [], True
    for i, c in enumerate(wl["cases"][:args.plain2048]):
        cols = [slope if k == "slope" else aspect if k == "aspect" else
                torch.full((n,), float(c["wx"][k]), dtype=torch.float32, device=dev) for k in FIELDS]
        tm, rx, ry = synth.initial_state(side, side, c["ign"])
        init = (torch.from_numpy(tm).to(dev), torch.from_numpy(rx.view(np.int16)).to(dev),
                torch.from_numpy(ry.view(np.int16)).to(dev))
        wall, _, t = plain_run(pp, torch, dev, stream, pp.Settings(g, c["max_time"]), fuel, cols, init, tuning)
        plain_ms.append(wall)
        if i < len(kept):
            parity &= bool(np.array_equal(t.cpu().numpy().view(np.uint32), kept[i].view(np.uint32)))
    plain_sps = 1e3 / float(np.mean(plain_ms[1:] if len(plain_ms) > 1 else plain_ms))
    out = {"workload": f"ensemble of {len(wl['cases'])} scenarios on one 2048x2048 landscape, max_time "
                       "log-uniform 3600-43200 s, 1-4 ignitions, uniform wind U(2,12) m/s",
           "value": value, "unit": "scenarios/s", "ms_per_ensemble": float(np.mean(ms)),
           "relax_ms": stats["relax_ms"], "sweeps": stats["iterations"], "tiles_visited": stats["tiles_visited"],
           "tile_marks": stats["tile_marks"], "idle_tiles": stats["idle_tiles"],
           "relaxations": stats["relaxations"], "improvements": stats["improvements"], "dda_steps": stats["dda_steps"],
           "roofline": roofline(stats["relaxations"], stats["improvements"], stats["relax_ms"], "ensemble_2048"),
           "plain_value": plain_sps, "plain_sample": len(plain_ms), "speedup_batched": value / plain_sps,
           "parity_first_scenarios": parity, "autotune": tune}
    if tuning.get("profile"):
        out.update({k: stats[k] for k in ("crit_ms", "phase_ms", "phase_total_ms", "barrier_ms", "idle_far_tiles")})
    log(f"ensemble 2048: {out['ms_per_ensemble']:.1f} ms per ensemble of {len(wl['cases'])} ({value:.1f} "
        f"scenarios/s; one per call {plain_sps:.1f}/s over {len(plain_ms) - 1} timed calls, "
        f"{value / plain_sps:.1f}x), relax kernel {stats['relax_ms']:.1f} ms, {stats['iterations']} sweeps, "
        f"parity {parity}")
    return out, wl


def serve_stream(args, pp, synth, torch, tuning):
    """Serving: Poisson arrivals over four resident landscapes of mixed sizes,
    continuous batching through py_propag.serve.Server; latency = arrival to
    arrival-time raster in (pinned) host memory."""
    from py_propag.serve import Request, Server

    wl = synth.serving_stream(seed=4, rate=args.serve_rate, seconds=args.serve_seconds)
    sides = wl["sides"]
    t0 = time.time()
    lands = [wl["land"](i) for i in range(len(sides))]
    log(f"serving: generated landscapes {sides} in {time.time() - t0:.1f}s; "
        f"{len(wl['requests'])} requests over {args.serve_seconds:.0f} s")
    geos = [pp.GeoReference(s, s, "+proj=utm +zone=30", [0.0, 5.0, 0.0, 0.0, 0.0, 5.0]) for s in sides]
    ctx = pp.Context()
    tids = [ctx.add_terrain(g, L["fuel"], L["slope"], L["aspect"]) for g, L in zip(geos, lands)]
    pinned = torch.empty(max(sides) ** 2, dtype=torch.float32, pin_memory=True)
    host = pinned.numpy()
    check = {}
    for i, c in enumerate(wl["requests"]):     # parity sample: the first request on each landscape
        check.setdefault(c["terrain"], i)
    check = set(check.values())
    kept = {}

    def on_done(req, sid):
        n = sides[req.tag[1]] ** 2
        ctx.time_into(sid, pinned.data_ptr())
        if req.tag[0] in check:
            kept[req.tag[0]] = host[:n].copy()

    def requests(cases, warm=False):
        return [Request(tids[c["terrain"]], c["wx"], c["ign"], c["max_time"], 0.0 if warm else c["arrival"],
                        tag=(-1 if warm else i, c["terrain"])) for i, c in enumerate(cases)]

    if tuning.get("tile_w"):
        ctx.set_tile(tuning["tile_w"], tuning["tile_h"])
    server = Server(ctx, on_done=on_done, tuning=tuning)
    server.play(requests(wl["requests"][:8], warm=True))        # untimed: pools, code paths
    server.launches, server.launch_ms = 0, 0.0
    done = server.play(requests(wl["requests"]))
    lat = np.array([(r.finished - r.arrival) * 1e3 for r in done])
    span = max(r.finished for r in done) - min(r.arrival for r in done)
    capped = sum(1 for r in done if r.state != pp.DONE)
    ok = True
    for i in sorted(check):
        c = wl["requests"][i]
        k = c["terrain"]
        n = sides[k] ** 2
        full = {"fuel": lands[k]["fuel"], "slope": lands[k]["slope"], "aspect": lands[k]["aspect"]}
        for f in ("d1hr", "d10hr", "d100hr", "herb", "wood", "wind_speed", "wind_azimuth"):
            full[f] = np.full(n, c["wx"][f], np.float32)
        alone = pp.propagate(pp.Settings(geos[k], c["max_time"]), full, ignitions=c["ign"])
        ok &= bool(np.array_equal(kept[i].view(np.uint32), alone.arrival_times.ravel().view(np.uint32)))
    ctx.close()
    out = {"value": len(done) / span, "unit": "scenarios/s", "requests": len(done),
           "rate_per_s": args.serve_rate, "seconds": args.serve_seconds, "sides": sides,
           "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
           "mean_ms": float(lat.mean()), "max_ms": float(lat.max()),
           "mean_in_flight": float(lat.sum() / 1e3 / span), "launches": server.launches,
           "gpu_launch_busy": server.launch_ms / 1e3 / span, "not_converged": capped,
           "busy_refusals": server.busy_refusals, "parity_sample": len(check), "parity": ok,
           "workload": f"Poisson {args.serve_rate:g}/s for {args.serve_seconds:g} s over 4 landscapes "
                       f"{sides} at 5 m, max_time U(3600,36000) s, 1-4 ignitions, uniform wind"}
    log(f"served: {out['value']:.1f} scenarios/s, latency p50 {out['p50_ms']:.2f} p99 {out['p99_ms']:.2f} "
        f"max {out['max_ms']:.2f} ms, {out['mean_in_flight']:.2f} in flight, {server.launches} launches, "
        f"GPU launches {100 * out['gpu_launch_busy']:.0f}% of the time; parity on {len(check)} requests: {ok}")
    return out


def readme_section(pp, synth, torch, tuning, stream, calls=10):
    """The reference README's CLI scenario (crates/propag/src/main.rs: 1024 x
    1024 at 5 m, fuel model 1, all moistures 0.1, wind 5 m/s toward azimuth
    0, flat, ignition at (500, 500), 10 h), timed per call in the drop-in
    form (pg_burn + pg_propagate on resident inputs).  Context for the one
    figure the reference publishes, README.md:98-108's "10 loops: 79.46 ms"
    on an unnamed 128-SM NVIDIA GPU from an older build (whether it covers
    one run or ten is unclear) - not a target."""
    dev = torch.device("cuda:0")
    sc = synth.uniform_scenario(side=1024)
    n = sc["width"] * sc["height"]
    g = pp.GeoReference(sc["width"], sc["height"], "+proj=utm +zone=30", [0.0, sc["dx"], 0.0, 0.0, 0.0, sc["dy"]])
    settings = pp.Settings(g, float(sc["max_time"]))
    fuel = torch.from_numpy(sc["terrain"]["fuel"]).to(dev)
    cols = device_columns(torch, dev, sc["terrain"], n, FIELDS)
    init = (torch.from_numpy(sc["time"]).to(dev), torch.from_numpy(sc["refs_x"].view(np.int16)).to(dev),
            torch.from_numpy(sc["refs_y"].view(np.int16)).to(dev))
    plain_run(pp, torch, dev, stream, settings, fuel, cols, init, tuning)
    runs = [plain_run(pp, torch, dev, stream, settings, fuel, cols, init, tuning) for _ in range(calls)]
    ms = [r[0] for r in runs]
    t = runs[-1][2]
    burned = int((t < pp.FLT_MAX).sum().item())
    out = {"workload": "reference README CLI scenario (main.rs): 1024x1024 at 5 m, fuel 1, moisture 0.1, "
                       "wind 5 m/s toward 0, flat, ignition (500,500), 10 h; pg_burn + pg_propagate per call",
           "ms": float(np.mean(ms)), "p50_ms": float(np.percentile(ms, 50)), "calls": calls,
           "relax_ms": float(np.mean([r[1]["relax_ms"] for r in runs])), "sweeps": runs[-1][1]["iterations"],
           "burned_cells": burned, "max_time_s": float(t[t < pp.FLT_MAX].max().item()),
           "reference_published_ms": 79.46, "reference_burned_cells": 447822,
           "reference_hardware": "unnamed NVIDIA GPU, 128 SMs, older build (README.md:98-112)",
           "note": "context only: the reference's figure is from another GPU and build"}
    log(f"README scenario: {out['ms']:.2f} ms per call (relax {out['relax_ms']:.2f} ms), {out['sweeps']} sweeps, "
        f"{burned} cells burned; the reference README printed 79.46 ms and 447822 cells")
    return out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_section(args, sc, field, large=None, ens2048=None):
    """The oracle on the host cores, beside the GPU figures (a reported
    baseline, not a target).  The oracle relaxes every cell of the burned
    bounding box (grown by one) in every sweep and runs the same sweeps as
    the GPU (the schedule is the same exact Jacobi)."""
    import oracle

    cores = len(os.sched_getaffinity(0))
    out = {"host": {"cpu_model": cpu_model(), "cores_available": cores, "nproc": os.cpu_count()}}
    W, H = sc["width"], sc["height"]
    # config-1: burn + serial propagation to convergence, one thread
    t0 = time.perf_counter()
    fire = oracle.burn_f32(sc["terrain"], nthreads=1)
    og = oracle.grid(W, H, sc["dx"], sc["dy"], sc["max_time"])
    ref = oracle.propagate(og, *fire, sc["time"], sc["refs_x"], sc["refs_y"], nthreads=1)
    c1_s = time.perf_counter() - t0
    out["config1"] = {"value": 1.0 / c1_s, "unit": "scenarios/s", "cores": 1, "kind": "port",
                      "sample": "one full config-1 scenario (burn + serial exact-Jacobi propagation to "
                                f"convergence, {ref['sweeps']} sweeps) on 1 thread, {c1_s:.2f} s",
                      "seconds": c1_s, "sweeps": ref["sweeps"], "relaxations": ref["relaxations"]}
    log(f"cpu config-1: {c1_s:.2f} s per scenario (1 thread, {ref['sweeps']} sweeps)")
    # Stage F on every host core (after one untimed call that starts the threads)
    oracle.burn_f32(field, nthreads=cores)
    t1 = time.perf_counter()
    oracle.burn_f32(field, nthreads=cores)
    burn_s = time.perf_counter() - t1
    n0 = field["fuel"].size
    out["rothermel"] = {"value": n0 / burn_s / 1e6, "unit": "Mcells/s", "cores": cores, "kind": "port",
                        "sample": f"the 1024x1024 i.i.d. field, f32 oracle burn on {cores} threads, {burn_s * 1e3:.1f} ms"}
    log(f"cpu rothermel: {n0 / burn_s / 1e6:.1f} Mcells/s on {cores} threads ({cpu_model()})")
    # 4096^2 (C2): first sweeps on one thread, extrapolated
    if large is not None:
        lsc, lres = large
        K = args.cpu_c2_sweeps
        LW, LH = lsc["width"], lsc["height"]
        lfire = oracle.burn_f32(lsc["terrain"], nthreads=cores)
        lg = oracle.grid(LW, LH, lsc["dx"], lsc["dy"], lsc["max_time"])
        t2 = time.perf_counter()
        part = oracle.propagate(lg, *lfire, lsc["time"], lsc["refs_x"], lsc["refs_y"], nthreads=1, max_sweeps=K)
        part_s = time.perf_counter() - t2
        burned = np.nonzero(lres["time_host"] < oracle.FLT_MAX)[0]
        bx, by = burned % LW, burned // LW
        box = (min(int(bx.max()) + 1, LW - 1) - max(int(bx.min()) - 1, 0) + 1) * \
              (min(int(by.max()) + 1, LH - 1) - max(int(by.min()) - 1, 0) + 1)
        per_relax = part_s / max(1, part["relaxations"])
        est = per_relax * box * lres["sweeps"]
        out["large_4096"] = {"value": 1.0 / est, "unit": "scenarios/s", "cores": 1, "kind": "port",
                             "estimated": True, "seconds_est": est,
                             "sample": f"the first {K} of {lres['sweeps']} sweeps on 1 thread ({part_s:.2f} s, "
                                       f"{part['relaxations']} cell relaxations); full run estimated as the "
                                       "time per relaxation x the final burned box grown by one cell x the "
                                       "GPU's sweep count (an upper bound: the box only grows)"}
        log(f"cpu 4096^2: first {K} sweeps {part_s:.2f} s on 1 thread; full run est {est:.0f} s")
    # 2048^2 ensemble (C3): a seeded sample of its scenarios, all cores
    if ens2048 is not None:
        side, land, cases = ens2048["side"], ens2048["land"], ens2048["cases"]
        done, spent = 0, 0.0
        for c in cases:
            if done >= args.cpu_c3_sample or (done > 0 and spent > args.cpu_c3_seconds):
                break
            terr = {"fuel": land["fuel"], "slope": land["slope"], "aspect": land["aspect"]}
            for k in FIELDS[:7]:
                terr[k] = np.full(side * side, np.float32(c["wx"][k]), np.float32)
            from py_propag import synth
            t3 = time.perf_counter()
            f3 = oracle.burn_f32(terr, nthreads=cores)
            tm, rx, ry = synth.initial_state(side, side, c["ign"])
            oracle.propagate(oracle.grid(side, side, 5.0, 5.0, c["max_time"]), *f3, tm, rx, ry, nthreads=cores)
            spent += time.perf_counter() - t3
            done += 1
        out["ensemble_2048"] = {"value": done / spent, "unit": "scenarios/s", "cores": cores, "kind": "port",
                                "sample": f"the first {done} scenarios of the seeded ensemble (burn + propagation "
                                          f"to convergence, the oracle on {cores} threads), {spent:.2f} s"}
        log(f"cpu 2048^2 ensemble: {done} scenarios in {spent:.2f} s on {cores} threads")
    return out


SECTIONS = ("config1", "plain", "ensemble", "serving", "ensemble2048", "large4096", "readme", "rothermel",
            "moisture", "cpu")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3, help="timed ensembles")
    ap.add_argument("--warmup", type=int, default=1, help="untimed ensembles")
    ap.add_argument("--batch", type=int, default=128, help="scenarios per ensemble")
    ap.add_argument("--plain", type=int, default=16, help="ensemble scenarios timed one per call")
    ap.add_argument("--mode", type=int, default=0, help="0 persistent kernel, 1 one launch per sweep")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="run the ensembles with the default tuning")
    ap.add_argument("--tune-sample", type=int, default=32, help="scenarios per autotuner trial")
    ap.add_argument("--tuning", default="", help="fixed ensemble tuning WAVES,GROUPS,MODE instead of the autotuner")
    ap.add_argument("--tile", default="", help="tile shape WxH (8x8, 16x4, 32x2) of the fixed tunings and the "
                                               "one-per-call runs (the autotuner searches all three)")
    ap.add_argument("--profile-phases", action="store_true",
                    help="time the sweep phases inside the kernel (slower; diagnostics only)")
    ap.add_argument("--serve-seconds", type=float, default=60.0,
                    help="length of the served Poisson request stream (0: skip serving)")
    ap.add_argument("--serve-rate", type=float, default=50.0, help="request arrivals per second")
    ap.add_argument("--batch2048", type=int, default=256, help="scenarios of the 2048x2048 ensemble (0: skip)")
    ap.add_argument("--plain2048", type=int, default=17, help="of which timed one per call (the first is untimed)")
    ap.add_argument("--no-large", action="store_true", help="skip the 4096x4096 scenario")
    ap.add_argument("--cpu-c2-sweeps", type=int, default=12, help="sweeps of the 4096^2 CPU sample")
    ap.add_argument("--cpu-c3-sample", type=int, default=4, help="scenarios of the 2048^2 CPU sample")
    ap.add_argument("--cpu-c3-seconds", type=float, default=20.0, help="time budget of the 2048^2 CPU sample")
    ap.add_argument("--only", choices=SECTIONS, default=None,
                    help="run one section alone and print its result (profiling; not the benchmark line)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="also run a device copy of this many MiB (rocprof counter scaling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for {args.gpus}); "
              "scenarios are independent and shard one-per-GPU, see DESIGN.md", file=sys.stderr)
        sys.exit(2)
    want = (lambda name: True) if args.only is None else (lambda name: name == args.only)

    import torch
    import py_propag as pp
    from py_propag import synth

    dev = torch.device("cuda:0")
    torch.cuda.set_device(0)
    stream = torch.cuda.current_stream()
    tuning = {"mode": args.mode, "profile": 1 if args.profile_phases else 0}
    if args.tile:
        tw, th = (int(v) for v in args.tile.lower().split("x"))
        tuning.update({"tile_w": tw, "tile_h": th})

    # ---------------------------------------------------------- workload
    t0 = time.time()
    sc = synth.single_scenario(seed=1, side=1024)
    W, H = sc["width"], sc["height"]
    n = W * H
    log(f"generated single_scenario_1024 in {time.time() - t0:.1f}s")
    g = pp.GeoReference(W, H, "+proj=utm +zone=30", [0.0, sc["dx"], 0.0, 0.0, 0.0, sc["dy"]])
    settings = pp.Settings(g, float(sc["max_time"]))
    land = {k: sc["terrain"][k] for k in ("fuel", "slope", "aspect")}
    cases = ensemble_cases(2024, args.batch, W, H, 36000.0)

    fuel = torch.from_numpy(land["fuel"]).to(dev)
    slope = torch.from_numpy(land["slope"]).to(dev)
    aspect = torch.from_numpy(land["aspect"]).to(dev)
    fire_all = torch.zeros(3 * n, dtype=torch.float32, device=dev)   # one fill pre-zeroes all three
    fire = [fire_all[i * n:(i + 1) * n] for i in range(3)]
    t_buf = torch.empty(n, dtype=torch.float32, device=dev)
    rx_buf = torch.empty(n, dtype=torch.int16, device=dev)
    ry_buf = torch.empty(n, dtype=torch.int16, device=dev)
    bits = torch.zeros((n + 31) // 32, dtype=torch.int32, device=dev)

    def resident_case(c):
        """Device columns and initial state of one ensemble scenario."""
        cols = []
        for k in FIELDS:
            if k in ("slope", "aspect"):
                cols.append(slope if k == "slope" else aspect)
            else:
                cols.append(torch.full((n,), float(c["wx"][k]), dtype=torch.float32, device=dev))
        tm, rx, ry = synth.initial_state(W, H, c["ign"])
        return cols, torch.from_numpy(tm).to(dev), torch.from_numpy(rx.view(np.int16)).to(dev), \
            torch.from_numpy(ry.view(np.int16)).to(dev)

    def plain_reset(t0_, rx0, ry0):
        """The caller's buffers before a call: initial state, pre-zeroed fire
        (the drop-in contract); not part of the timed call."""
        t_buf.copy_(t0_)
        rx_buf.copy_(rx0)
        ry_buf.copy_(ry0)
        fire_all.zero_()
        torch.cuda.synchronize()

    def plain_call(cols, st=settings):
        """One call of the drop-in form: pg_burn + pg_propagate."""
        pp.burn_device(fuel, *cols, *fire, stream=stream)
        return pp.propagate_device(st, *fire, t_buf, rx_buf, ry_buf, bits, tuning=tuning, stream=stream)

    def plain_step(cols, t0_, rx0, ry0, st=settings):
        """One call timed (wall ms) after its reset; returns (ms, stats)."""
        plain_reset(t0_, rx0, ry0)
        s0 = time.perf_counter()
        stats = plain_call(cols, st)
        return (time.perf_counter() - s0) * 1e3, stats

    results = {}

    # ------------------------------------------- plain: config-1 scenario
    if want("config1"):
        c1_cols = [torch.from_numpy(np.ascontiguousarray(sc["terrain"][k], np.float32)).to(dev) for k in FIELDS]
        c1_init = (torch.from_numpy(sc["time"]).to(dev), torch.from_numpy(sc["refs_x"].view(np.int16)).to(dev),
                   torch.from_numpy(sc["refs_y"].view(np.int16)).to(dev))
        for _ in range(2):
            plain_step(c1_cols, *c1_init)
        torch.cuda.synchronize()
        c1_lat, c1_stats = [], []
        for _ in range(10):
            ms, stc = plain_step(c1_cols, *c1_init)
            c1_stats.append(stc)
            c1_lat.append(ms)
        burned = int((t_buf < pp.FLT_MAX).sum().item())
        st1 = c1_stats[-1]
        relax_ms = float(np.mean([s["relax_ms"] for s in c1_stats]))
        log(f"config-1 plain: {np.mean(c1_lat):.3f} ms/scenario, relax kernel {relax_ms:.3f} ms, "
            f"{st1['iterations']} sweeps, {st1['relaxations']} relaxations, {st1['tiles_visited']} tile visits, "
            f"{burned} burned; critical path {st1['crit_ms']:.3f} ms, phases (stage/rule/walk/tail) "
            f"{st1['phase_ms']}, barrier (start/span/gather/book) {st1['barrier_ms']}")
        results["config1"] = {
            "ms": float(np.mean(c1_lat)), "p50_ms": float(np.percentile(c1_lat, 50)),
            "p99_ms": float(np.percentile(c1_lat, 99)),
            "stats": {"sweeps": st1["iterations"], "relaxations": st1["relaxations"],
                      "improvements": st1["improvements"], "dda_steps": st1["dda_steps"],
                      "tiles_visited": st1["tiles_visited"], "burned_cells": burned,
                      "relax_ms": relax_ms, "crit_ms": st1["crit_ms"],
                      "phase_ms": st1["phase_ms"], "barrier_ms": st1["barrier_ms"],
                      "lds_bitset": st1["lds_bitset"], "workgroups": st1["workgroups"]},
            "roofline": roofline(st1["relaxations"], st1["improvements"], relax_ms, "config1")}

    # ------------------------------------------ plain: ensemble scenarios
    plain_out = []
    if want("plain") or want("ensemble"):
        n_plain = min(args.plain, args.batch)
        res_cases = [resident_case(c) for c in cases[:n_plain]]
        plain_step(*res_cases[0])
        torch.cuda.synchronize()
        plain_lat = []
        for i, rc in enumerate(res_cases):
            ms, _ = plain_step(*rc)
            plain_lat.append(ms)
            if i < 2:
                plain_out.append(t_buf.cpu().numpy().copy())
        plain_sps = 1e3 / float(np.mean(plain_lat))
        log(f"ensemble scenarios one per call: {np.mean(plain_lat):.3f} ms/scenario ({plain_sps:.1f} scenarios/s), "
            f"p50 {np.percentile(plain_lat, 50):.3f} p99 {np.percentile(plain_lat, 99):.3f} ms")
        results["plain"] = {"value": plain_sps, "lat": plain_lat, "scenarios": n_plain}
        del res_cases

    # ------------------------------------------------ batched: ensembles
    if want("ensemble"):
        ctx = pp.Context()
        tid = ctx.add_terrain(g, land["fuel"], land["slope"], land["aspect"])
        btuning, tune_1024 = autotuned(args, ctx, tid, cases, W, H, tuning)

        def ensemble():
            ids = [ctx.admit(tid, c["wx"], c["ign"], c["max_time"]) for c in cases]
            if any(i is None for i in ids):
                raise RuntimeError("ensemble does not fit the memory budget")
            fin, stats = ctx.run(tuning=btuning)
            return ids, fin, stats

        for _ in range(args.warmup):
            ids, _, _ = ensemble()
            for i in ids:
                ctx.release(i)
        torch.cuda.synchronize()
        ens_ms, ens_stats = [], []
        wall0 = time.perf_counter()
        for k in range(args.steps):
            s0 = time.perf_counter()
            ids, fin, stats = ensemble()
            ens_ms.append((time.perf_counter() - s0) * 1e3)
            ens_stats.append(stats)
            if fin != len(cases):
                raise RuntimeError(f"{len(cases) - fin} scenarios did not finish")
            if k == args.steps - 1:
                parity = all(np.array_equal(ctx.result(ids[i])[0].ravel().view(np.uint32),
                                            plain_out[i].view(np.uint32)) for i in range(len(plain_out)))
            for i in ids:
                ctx.release(i)
        wall = time.perf_counter() - wall0
        ctx.close()
        value = args.batch * args.steps / wall
        es = ens_stats[-1]
        plain_sps = results["plain"]["value"]
        log(f"batched: {np.mean(ens_ms):.1f} ms per ensemble of {args.batch} ({value:.1f} scenarios/s, "
            f"{value / plain_sps:.1f}x one per call); relax kernel {es['relax_ms']:.1f} ms, {es['iterations']} sweeps, "
            f"{es['relaxations']} relaxations, {es['tiles_visited']} tile visits ({es['idle_tiles']} idle, "
            f"{es['idle_far_tiles']} of them with no change within two cells), {es['tile_marks']} marks, "
            f"{es['dda_steps']} walk steps; parity with one-per-call: {parity}; phases (stage/rule/walk/tail) "
            f"{es['phase_ms']}, phase totals over waves {es['phase_total_ms']}, "
            f"barrier (start/span/gather/book) {es['barrier_ms']}; per workgroup: work {es['work_ms']:.1f} ms, "
            f"wait {es['wait_ms']:.1f} ms, slowest-workgroup work summed over sweeps {es['crit_ms']:.1f} ms")
        ens_relax_ms = float(np.mean([s["relax_ms"] for s in ens_stats]))
        rl = roofline(es["relaxations"], es["improvements"], ens_relax_ms, "ensemble_1024")
        if args.mode == 1:
            rl["kernel"] = "propagate_sweep"
        results["ensemble"] = {"value": value, "wall": wall, "ens_ms": ens_ms, "stats": es, "parity": parity,
                               "tuning": btuning, "autotune": tune_1024, "roofline": rl}

    serving = serve_stream(args, pp, synth, torch, tuning) if want("serving") and args.serve_seconds > 0 else None
    ens2048, wl2048 = (ensemble_2048_section(args, pp, synth, torch, tuning, stream)
                       if want("ensemble2048") and args.batch2048 > 0 else (None, None))
    large, large_sc = (large_single(args, pp, synth, torch, tuning, stream)
                       if want("large4096") and not args.no_large else (None, None))
    readme = readme_section(pp, synth, torch, tuning, stream) if want("readme") else None

    # ---------------------------------------------------- Rothermel field
    field = synth.rothermel_field(seed=0, side=1024)
    rothermel = None
    if want("rothermel"):
        bf = torch.from_numpy(field["fuel"]).to(dev)
        bc = [torch.from_numpy(field[k]).to(dev) for k in FIELDS]
        bo = [torch.zeros(n, dtype=torch.float32, device=dev) for _ in range(3)]
        for _ in range(3):
            pp.burn_device(bf, *bc, *bo, stream=stream)
        torch.cuda.synchronize()
        reps = 50
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            pp.burn_device(bf, *bc, *bo, stream=stream)
        e1.record(stream)
        torch.cuda.synchronize()
        burn_ms = e0.elapsed_time(e1) / reps
        burn_mcells = n / (burn_ms * 1e-3) / 1e6
        burn_gbs = 49.0 * n / (burn_ms * 1e-3) / 1e9
        log(f"rothermel: {burn_ms * 1e3:.1f} us per 1024^2 pass, {burn_mcells:.0f} Mcells/s, {burn_gbs:.0f} GB/s")
        rothermel = {"value": burn_mcells, "unit": "Mcells/s", "us_per_pass": burn_ms * 1e3,
                     "achieved_gbs": burn_gbs, "frac": burn_gbs / HBM_PEAK_GBS, "workload": "rothermel_field_1024"}

    # ------------------------------------------------ moisture pre-pass
    moisture = None
    if want("moisture"):
        mf = torch.from_numpy(field["fuel"]).to(dev)
        mo_slope = torch.from_numpy(field["slope"]).to(dev)
        mo_aspect = torch.from_numpy(field["aspect"]).to(dev)
        mo_outs = [torch.empty(n, dtype=torch.float32, device=dev) for _ in range(5)]
        hours = np.arange(24)
        mo_w = pp.day_weather(12 + 10 * np.sin(hours / 24 * 2 * np.pi), 60 - 25 * np.sin(hours / 24 * 2 * np.pi),
                              30.0, [0.0, 0.0, 3.0, 0.0, 12.0, 0.0], 7)
        for _ in range(3):
            pp.moisture_device(mf, mo_slope, mo_aspect, mo_w, 14, *mo_outs, stream=stream)
        reps = 50
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            pp.moisture_device(mf, mo_slope, mo_aspect, mo_w, 14, *mo_outs, stream=stream)
        e1.record(stream)
        torch.cuda.synchronize()
        mo_ms = e0.elapsed_time(e1) / reps
        mo_gbs = 29.0 * n / (mo_ms * 1e-3) / 1e9
        log(f"moisture pre-pass: {mo_ms * 1e3:.1f} us per 1024^2 hour, {n / (mo_ms * 1e-3) / 1e6:.0f} Mcells/s, "
            f"{mo_gbs:.0f} GB/s")
        moisture = {"value": n / (mo_ms * 1e-3) / 1e6, "unit": "Mcells/s", "us_per_pass": mo_ms * 1e3,
                    "achieved_gbs": mo_gbs, "frac": mo_gbs / HBM_PEAK_GBS,
                    "workload": "moisture pre-pass, 1024x1024 i.i.d. fuel/slope/aspect, July day, 14:00"}

    if args.copy_probe_mib:
        src = torch.empty(args.copy_probe_mib << 18, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy probe: {args.copy_probe_mib} MiB read + written")

    # -------------------------------------------------------- CPU baseline
    cpu = None
    if not args.no_cpu_baseline and want("cpu"):
        c2 = None
        if large is not None:
            c2 = (large_sc[0], {"time_host": large_sc[1], "sweeps": large["sweeps"]})
        cpu = cpu_section(args, sc, field, c2, wl2048)

    if args.only is not None:
        sect = {"config1": results.get("config1"), "plain": results.get("plain"),
                "ensemble": {k: v for k, v in (results.get("ensemble") or {}).items() if k != "stats"},
                "serving": serving, "ensemble2048": ens2048, "large4096": large, "readme": readme,
                "rothermel": rothermel, "moisture": moisture, "cpu": cpu}[args.only]
        print(json.dumps({"section": args.only, "result": sect}, default=float), flush=True)
        return

    ens = results["ensemble"]
    es = ens["stats"]
    c1 = results["config1"]
    plain = results["plain"]
    cpu_line = None
    if cpu is not None:
        cpu_line = dict(cpu["config1"])
        cpu_line.update({"rothermel_mcells_s": cpu["rothermel"]["value"], "rothermel_cores": cpu["rothermel"]["cores"],
                         "host": cpu["host"], "large_4096": cpu.get("large_4096"),
                         "ensemble_2048": cpu.get("ensemble_2048")})
    out = {
        "metric": METRIC,
        "value": ens["value"],
        "unit": "scenarios/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ens["wall"] * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"ensemble of {args.batch} 1024x1024 10-h scenarios on the config-1 landscape",
                   "cells_per_scenario": n, "cell_m": 5.0, "max_time_s": 36000.0,
                   "form": "batched (one shared persistent launch per ensemble)",
                   "schedule": "one launch per sweep" if args.mode == 1 else "persistent",
                   "tuning": ens["tuning"]},
        "latency_ms": {"ensemble_p50": float(np.percentile(ens["ens_ms"], 50)),
                       "ensemble_p99": float(np.percentile(ens["ens_ms"], 99)),
                       "plain_p50": float(np.percentile(plain["lat"], 50)),
                       "plain_p99": float(np.percentile(plain["lat"], 99))},
        "plain": {"value": plain["value"], "unit": "scenarios/s", "scenarios": plain["scenarios"],
                  "speedup_batched": ens["value"] / plain["value"], "parity_first_scenarios": bool(ens["parity"]),
                  "config1_ms": c1["ms"], "config1_p50_ms": c1["p50_ms"], "config1_p99_ms": c1["p99_ms"],
                  "config1_roofline": c1["roofline"]},
        "autotune": ens["autotune"],
        "roofline": ens["roofline"],
        "rothermel": rothermel,
        "serving": serving,
        "ensemble_2048": ens2048,
        "large_4096": large,
        "readme_scenario": readme,
        "moisture": moisture,
        "cpu_baseline": cpu_line,
        "propagation": {"ensemble": {k: es[k] for k in ("iterations", "relaxations", "improvements", "dda_steps",
                                                       "tiles_visited", "tile_marks", "idle_tiles", "relax_ms",
                                                       "elapsed_ms", "workgroups", "work_ms", "wait_ms", "crit_ms")},
                        "config1": c1["stats"]},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
