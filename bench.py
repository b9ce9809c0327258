"""Benchmark of the propagation hot path on one MI355X.

Workload (one step = one scenario): the single-scenario landscape,
1024x1024 cells of 5 m, 20 Gaussian hills, 200 Voronoi fuel patches,
constant moisture, smooth wind 6+-2 m/s toward 225+-20 deg, ignition at
(500, 500), max_time 36000 s.  The terrain rasters are resident in HBM
before the timed region; a step resets the arrival-time state on the
device, runs the burn pass (pg_burn) and the propagation (pg_propagate:
boundaries, exact-Jacobi relaxation to convergence, post-pass).  Steps run
one after another, one scenario per call (the plain drop-in form), so the
value is scenarios/s of that form with p50/p99 step latency.

Also reported on the same line:
  * roofline of the dominant kernel (the relaxation kernel), with
    algorithmic bytes = 24 B per cell relaxation + 8 B per committed
    improvement, timed with HIP events around that kernel on its stream;
  * the Rothermel burn pass alone on the 1024x1024 i.i.d. field (Mcells/s,
    49 B/cell);
  * the CPU baseline: the oracle (a serial C transcription) on the same
    scenario, one thread, and its burn pass on all host cores.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
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


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--mode", type=int, default=0, help="0 persistent kernel, 1 one launch per sweep")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--profile-phases", action="store_true",
                    help="time the sweep phases inside the kernel (slower; diagnostics only)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="also run a device copy of this many MiB (rocprof counter scaling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for {args.gpus}); "
              "scenarios are independent and shard one-per-GPU, see DESIGN.md", file=sys.stderr)
        sys.exit(2)

    import torch
    import py_propag as pp
    from py_propag import synth

    dev = torch.device("cuda:0")
    torch.cuda.set_device(0)
    stream = torch.cuda.current_stream()

    # ---------------------------------------------------------- workload
    t0 = time.time()
    sc = synth.single_scenario(seed=1, side=1024)
    W, H = sc["width"], sc["height"]
    n = W * H
    log(f"generated single_scenario_1024 in {time.time() - t0:.1f}s")
    g = pp.GeoReference(W, H, "+proj=utm +zone=30", [0.0, sc["dx"], 0.0, 0.0, 0.0, sc["dy"]])
    settings = pp.Settings(g, float(sc["max_time"]))
    fuel = torch.from_numpy(sc["terrain"]["fuel"]).to(dev)
    cols = [torch.from_numpy(np.ascontiguousarray(sc["terrain"][k], np.float32)).to(dev)
            for k in FIELDS]
    fire = [torch.zeros(n, dtype=torch.float32, device=dev) for _ in range(3)]
    t_init = torch.from_numpy(sc["time"]).to(dev)
    rx_init = torch.from_numpy(sc["refs_x"].view(np.int16)).to(dev)
    ry_init = torch.from_numpy(sc["refs_y"].view(np.int16)).to(dev)
    t_buf = torch.empty_like(t_init)
    rx_buf = torch.empty_like(rx_init)
    ry_buf = torch.empty_like(ry_init)
    bits = torch.zeros((n + 31) // 32, dtype=torch.int32, device=dev)
    tuning = {"mode": args.mode, "profile": 1 if args.profile_phases else 0}

    def step():
        t_buf.copy_(t_init)
        rx_buf.copy_(rx_init)
        ry_buf.copy_(ry_init)
        for f in fire:
            f.zero_()
        pp.burn_device(fuel, *cols, *fire, stream=stream)
        return pp.propagate_device(settings, *fire, t_buf, rx_buf, ry_buf, bits, tuning=tuning,
                                   stream=stream)

    for _ in range(args.warmup):
        st = step()
    torch.cuda.synchronize()
    lat, stats = [], []
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
    torch.cuda.synchronize()
    wall0 = time.perf_counter()
    ev[0].record(stream)
    for k in range(args.steps):
        s0 = time.perf_counter()
        stats.append(step())
        ev[k + 1].record(stream)
        lat.append((time.perf_counter() - s0) * 1e3)
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    burned = int((t_buf < pp.FLT_MAX).sum().item())
    ms_per_step = wall * 1e3 / args.steps
    relax_ms = np.array([s["relax_ms"] for s in stats])
    relax = stats[-1]["relaxations"]
    impr = stats[-1]["improvements"]
    alg_bytes = 24.0 * relax + 8.0 * impr
    achieved = alg_bytes / (relax_ms.mean() * 1e-3) / 1e9
    log(f"propagation: {ms_per_step:.3f} ms/scenario, relax kernel {relax_ms.mean():.3f} ms, "
        f"{stats[-1]['iterations']} sweeps, {relax} relaxations, {burned} burned; per workgroup "
        f"work {stats[-1]['work_ms']:.3f} ms, barrier wait {stats[-1]['wait_ms']:.3f} ms, "
        f"{stats[-1]['workgroups']} workgroups, LDS bitset {stats[-1]['lds_bitset']}, critical path "
        f"{stats[-1]['crit_ms']:.3f} ms, phases (stage/rule/walk/tail) {stats[-1]['phase_ms']}, "
        f"barrier (start/span/gather/book) {stats[-1]['barrier_ms']}")

    # ---------------------------------------------------- Rothermel field
    field = synth.rothermel_field(seed=0, side=1024)
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

    if args.copy_probe_mib:
        src = torch.empty(args.copy_probe_mib << 18, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy probe: {args.copy_probe_mib} MiB read + written")

    # -------------------------------------------------------- CPU baseline
    cpu = None
    if not args.no_cpu_baseline:
        import oracle

        t0 = time.perf_counter()
        cpu_fire = oracle.burn_f32(sc["terrain"], nthreads=1)
        og = oracle.grid(W, H, sc["dx"], sc["dy"], sc["max_time"])
        ref = oracle.propagate(og, *cpu_fire, sc["time"], sc["refs_x"], sc["refs_y"], nthreads=1)
        cpu_s = time.perf_counter() - t0
        cores = min(16, len(os.sched_getaffinity(0)))
        t1 = time.perf_counter()
        oracle.burn_f32(field, nthreads=cores)
        cpu_burn_s = time.perf_counter() - t1
        cpu = {"value": 1.0 / cpu_s, "unit": "scenarios/s", "cores": 1, "kind": "port",
               "sample": "one full single_scenario_1024 (burn + serial exact-Jacobi propagation to "
                         f"convergence, {ref['sweeps']} sweeps) on 1 thread, {cpu_s:.2f} s",
               "rothermel_mcells_s": n / cpu_burn_s / 1e6, "rothermel_cores": cores}
        log(f"cpu baseline: {cpu_s:.2f} s per scenario (1 thread); burn {n / cpu_burn_s / 1e6:.1f} "
            f"Mcells/s on {cores} threads")

    p50, p99 = (float(np.percentile(lat, q)) for q in (50, 99))
    out = {
        "metric": METRIC,
        "value": args.steps / wall,
        "unit": "scenarios/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "single_scenario_1024", "cells": n, "cell_m": 5.0,
                   "max_time_s": 36000.0, "form": "plain (one scenario per call)",
                   "schedule": "persistent" if args.mode == 0 else "one launch per sweep"},
        "latency_ms": {"p50": p50, "p99": p99},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "propagate_persistent" if args.mode == 0 else "propagate_sweep",
                     "kernel_ms": float(relax_ms.mean()), "algorithmic_bytes": alg_bytes},
        "rothermel": {"value": burn_mcells, "unit": "Mcells/s", "us_per_pass": burn_ms * 1e3,
                      "achieved_gbs": burn_gbs, "frac": burn_gbs / HBM_PEAK_GBS,
                      "workload": "rothermel_field_1024"},
        "cpu_baseline": cpu,
        "propagation": {"sweeps": stats[-1]["iterations"], "relaxations": relax,
                        "improvements": impr, "dda_steps": stats[-1]["dda_steps"],
                        "tiles_visited": stats[-1]["tiles_visited"], "tile_marks": stats[-1]["tile_marks"],
                        "burned_cells": burned,
                        "elapsed_ms": float(np.mean([s["elapsed_ms"] for s in stats])),
                        "work_ms": float(np.mean([s["work_ms"] for s in stats])),
                        "wait_ms": float(np.mean([s["wait_ms"] for s in stats])),
                        "workgroups": stats[-1]["workgroups"], "lds_bitset": stats[-1]["lds_bitset"],
                        "crit_ms": float(np.mean([s["crit_ms"] for s in stats])),
                        "phase_ms": stats[-1]["phase_ms"],
                        "barrier_ms": stats[-1]["barrier_ms"]},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
