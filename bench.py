#!/usr/bin/env python3
"""Benchmark of the MI355X time-stepping path on the 12M-cell grid.

One "step" is one full time step (mass + momentum passes, dt reduction, metrics) of
a 4000x3000 lon/lat grid at 2 arc-min spacing built by the seeded generator (20
seamounts, coastline along the east and north edges, four 100x50 km subfaults with
2 m slip applied at t = 0, 64 gauges).  State is resident in HBM before the timed
region starts.  Prints one JSON line (rank 0, N = 1):

  value        Gcell-steps/s over exactly --steps steps after --warmup steps
  roofline     dominant kernel (the step kernel): algorithmic bytes per launch / measured
               launch time (HIP events on the library's stream) vs 8 TB/s
  cpu_baseline FP64 restatement of the reference scheme (the oracle) timed on the
               host cores for a bounded sample of the same workload
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "tsunami-hysea_amd"))

METRIC = "Gcell-steps/s at 12M cells; batched scenarios/hr + p99 latency vs plain"
HBM_PEAK_GBS = 8000.0
B_ALG = {False: 36.0, True: 44.0}  # bytes per cell-step (arrival tracking off / on)
KERNEL_NAME = {"sweep": "k_sweep", "tile": "k_step"}


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def cpu_baseline(prob, steps, threads):
    """Oracle (FP64 restatement of the reference scheme) on the host cores."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    s = O.OracleSim(prob, threads=threads)
    for f in prob.faults:
        s.okada(f)
    dt = s.initial_dt()
    dt = s.step(dt)  # first touch outside the timing
    t0 = time.perf_counter()
    for _ in range(steps):
        dt = s.step(dt)
    sec = time.perf_counter() - t0
    s.close()
    return prob.nx * prob.ny * steps / sec / 1e9, sec


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=480)
    ap.add_argument("--warmup", type=int, default=48)
    ap.add_argument("--nx", type=int, default=4000)
    ap.add_argument("--ny", type=int, default=3000)
    ap.add_argument("--no-arrival", action="store_true", help="disable arrival-time tracking")
    ap.add_argument("--kernel", default="sweep", choices=["sweep", "tile"], help="step kernel")
    ap.add_argument("--strip-rows", type=int, default=0, help="row-sweep kernel: rows per strip (0: default)")
    ap.add_argument("--tile", default="0x0", help="2D-tile kernel: tile shape (0x0: default)")
    ap.add_argument("--sync-every", type=int, default=48, help="steps per captured graph replay")
    ap.add_argument("--cpu-steps", type=int, default=30)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="also run a plain device copy of this size (for counter scaling)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "multi-GPU sharding is design-only", file=sys.stderr)
        sys.exit(2)

    from thysea import native, scenario
    if native.probe() != 0:
        raise SystemExit("bench.py: no gfx950 device visible")
    log("generating workload", args.nx, "x", args.ny)
    prob = scenario.problem("large", nx=args.nx, ny=args.ny)
    # the configured 1800 s is ~500 steps; keep stepping past it when --warmup + --steps asks for more
    prob.t_end = max(prob.t_end, 1e9)
    track = not args.no_arrival
    tile = tuple(int(v) for v in args.tile.split("x"))
    ncell = prob.nx * prob.ny

    copy = None
    if args.copy_probe_mib > 0:
        ms, gbs = native.copy_probe(args.copy_probe_mib << 20, reps=5)
        copy = {"bytes": args.copy_probe_mib << 20, "ms": ms, "GBs_read_plus_write": gbs}
        log(f"copy probe {args.copy_probe_mib} MiB: {ms:.3f} ms, {gbs:.0f} GB/s")

    solver = native.Solver(prob, precision=0, track_arrival=track, tile=tile, kernel=args.kernel,
                           strip_rows=args.strip_rows)
    log("warmup", args.warmup, "steps (includes Okada at t=0)")
    w = solver.run(max_steps=args.warmup, steps_per_sync=args.sync_every)
    log("warmup done: t =", w["t_final"] * prob.T, "s")
    t0 = time.perf_counter()
    r = solver.run_more(args.steps)
    wall = time.perf_counter() - t0
    done = r["steps"] - w["steps"]
    if done != args.steps:
        raise SystemExit(f"bench.py: ran {done} steps, expected {args.steps}")
    gpu_s = r["seconds"]
    value = ncell * args.steps / wall / 1e9
    kern_ms = gpu_s * 1e3 / args.steps
    achieved = B_ALG[track] * ncell / (kern_ms * 1e-3) / 1e9
    log(f"timed {args.steps} steps: wall {wall*1e3:.2f} ms, gpu {gpu_s*1e3:.2f} ms, {value:.1f} Gcell-steps/s")

    traffic = None
    tp = os.path.join(ROOT, "profiles", f"{KERNEL_NAME[args.kernel]}_12M_traffic.json")
    if os.path.exists(tp):
        with open(tp) as f:
            traffic = json.load(f).get("bytes_per_launch")

    cpu = None
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        log("cpu baseline:", args.cpu_steps, "oracle steps on", threads, "threads")
        gc, sec = cpu_baseline(prob, args.cpu_steps, threads)
        cpu = {"value": gc, "unit": "Gcell-steps/s", "cores": threads, "kind": "port",
               "sample": f"{args.cpu_steps} steps of the same {prob.nx}x{prob.ny} grid after Okada "
                         f"({sec:.1f} s), FP64 oracle, OpenMP"}
    solver.close()
    out = {
        "metric": METRIC, "value": value, "unit": "Gcell-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{prob.nx}x{prob.ny} lon/lat grid ({ncell/1e6:.1f}M cells, 2 arc-min), "
                               "4 Okada subfaults, 64 gauges, open borders",
                   "cells": ncell, "track_arrival": track, "steps_per_sync": args.sync_every,
                   "step_kernel": args.kernel,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": KERNEL_NAME[args.kernel], "kernel_ms": kern_ms,
                     "bytes_per_cell_step": B_ALG[track]},
        "cpu_baseline": cpu,
    }
    if copy:
        out["copy_probe"] = copy
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
