#!/usr/bin/env python3
"""Benchmark of the MI355X time-stepping path on the 12M-cell grid.

One "step" is one full time step (mass + momentum passes, dt reduction, metrics) of
a 4000x3000 lon/lat grid at 2 arc-min spacing built by the seeded generator (20
seamounts, coastline along the east and north edges, four 100x50 km subfaults with
2 m slip applied at t = 0, 64 gauges).  State is resident in HBM before the timed
region starts.  Prints one JSON line (rank 0, N = 1):

  value        Gcell-steps/s over exactly --steps steps after --warmup steps: host wall
               time of one thy_run_more call that replays one captured sequence of
               exactly --steps steps (strip heights tuned and the sequence captured and
               uploaded before the clock starts)
  roofline     44 algorithmic bytes per cell-step at that rate vs 8 TB/s (the same time
               as value); kernel_ms is the HIP-event time of the replay per step
  cpu_baseline FP64 restatement of the reference scheme (the oracle) timed on all host
               cores for a bounded sample of the same workload, plus the 240x200 workload
               run in full
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "tsunami-hysea_amd"))

METRIC = "Gcell-steps/s at 12M cells; batched scenarios/hr + p99 latency vs plain"
HBM_PEAK_GBS = 8000.0
B_ALG = {False: 36.0, True: 44.0}  # bytes per cell-step (arrival tracking off / on)
# the launches of one step, and the profile file stem of its measured HBM traffic
KERNEL_NAME = {"sweep": "k_sweep_deep + k_sweep (general, hand-back)", "tile": "k_step"}
TRAFFIC_FILE = {"sweep": "sweep_step_12M_traffic.json", "tile": "k_step_12M_traffic.json"}


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def cpu_model():
    """Model name of the host CPU with its family/model/stepping numbers (Linux), or None."""
    info = {}
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if ":" in line:
                    k, v = line.split(":", 1)
                    info.setdefault(k.strip(), v.strip())
    except OSError:
        return None
    name = info.get("model name")
    if name is None:
        return None
    return (f"{name} (family {info.get('cpu family', '?')}, model {info.get('model', '?')}, "
            f"stepping {info.get('stepping', '?')}, {os.cpu_count()} logical CPUs)")


def cpu_full_run(threads):
    """The 240x200 workload (3600 s, one fault, 16 gauges) run in full by the oracle."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from thysea import scenario
    pr = scenario.problem("small")
    s = O.OracleSim(pr, threads=threads)
    t0 = time.perf_counter()
    r = s.run()
    sec = time.perf_counter() - t0
    s.close()
    return {"workload": f"{pr.nx}x{pr.ny}, {pr.t_end * pr.T:.0f} s, whole driver loop", "steps": int(r["steps"]),
            "seconds": sec, "gcell_steps_per_s": pr.nx * pr.ny * r["steps"] / sec / 1e9, "cores": threads}


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


def batched_phase(n, sample, sync):
    """256 scenarios on one 1000x800 grid: one batch vs the plain solver one at a time
    (timed on ``sample`` of them, scaled to n).  Batched results of the sampled
    scenarios are compared with their plain runs."""
    from thysea import scenario, serving
    probs = scenario.batch_problems(n)
    keep = list(range(min(sample, n)))
    log(f"batched: {n} scenarios on {probs[0].nx}x{probs[0].ny}")
    # the batch's strip height for this grid, chosen before the clock starts (the plain
    # runs keep the default, which is tuned for one scenario alone)
    rows_b, cost_b = serving.tune_batch(probs, steps_per_sync=sync)
    log("batched: strip heights " + ", ".join(f"{k}: {v:.2f}" for k, v in cost_b.items())
        + f" us per scenario-step; {rows_b} rows")
    # the sequential sample first, so that nothing the batch phase leaves behind can
    # slow the denominator of the speedup; each scenario timed on its own
    t0 = time.perf_counter()
    rec_s, res_s = serving.run_sequential([probs[i] for i in keep], steps_per_sync=sync, keep=keep)
    wall_s = time.perf_counter() - t0
    per = np.array([r["finished"] - r["admitted"] for r in rec_s])
    per_s = wall_s / len(keep)
    log(f"sequential: {len(keep)} scenarios in {wall_s:.2f} s ({per_s*1e3:.1f} ms each; "
        f"min {per.min()*1e3:.1f}, median {np.median(per)*1e3:.1f}, max {per.max()*1e3:.1f} ms)")
    t0 = time.perf_counter()
    rec_b, res_b, gpu_b = serving.run_batched(probs, steps_per_sync=sync, keep=keep,
                                              solver_kw=dict(strip_rows=rows_b))
    wall_b = time.perf_counter() - t0
    log(f"batched: {wall_b:.2f} s wall, {gpu_b:.2f} s in step launches")
    err = 0.0
    for i in keep:
        a, b = res_b[i]["max_height"], res_s[i]["max_height"]
        wet = b > -9000
        peak = float(np.max(np.abs(b[wet]))) if wet.any() else 1.0
        err = max(err, float(np.max(np.abs(a[wet] - b[wet]))) / peak)
        assert res_b[i]["steps"] == res_s[i]["steps"]
    cells = probs[0].nx * probs[0].ny
    steps = sum(r["steps"] for r in rec_b)
    st = serving.latency_stats(rec_b)
    return {"workload": f"{n} scenarios, shared {probs[0].nx}x{probs[0].ny} grid, Mw 7.0-8.8, 3600 s each",
            "scenarios_per_hour": n / wall_b * 3600.0,
            "sequential_scenarios_per_hour": 3600.0 / per_s,
            "speedup_vs_sequential": (n / wall_b) / (1.0 / per_s),
            "gcell_steps_per_s": cells * steps / wall_b / 1e9,
            "p50_s": st["p50_s"], "p99_s": st["p99_s"], "wall_s": wall_b,
            "sequential_sample": len(keep),
            "sequential_ms_per_scenario": {"min": float(per.min() * 1e3), "median": float(np.median(per) * 1e3),
                                           "max": float(per.max() * 1e3)},
            "max_height_rel_err_vs_plain": err, "strip_rows": rows_b,
            "strip_rows_tuning_us_per_scenario_step": {str(k): round(v, 3) for k, v in cost_b.items()}}


def serving_phase(rate, horizon, sync, seq):
    """Poisson arrivals of requests on grids of different sizes: batched serving, and
    (when ``seq``) the same requests through the plain solver one at a time."""
    from thysea import scenario, serving
    reqs = scenario.serving_requests(rate, horizon)
    log(f"serving: {len(reqs)} requests over {horizon:.0f} s at {rate}/s")
    probs = [scenario.request_problem(spec) for _, spec in reqs]
    arr = [a for a, _ in reqs]
    keep = list(range(0, len(probs), max(1, len(probs) // 4)))
    rec_b, res_b, gpu_b = serving.run_batched(probs, arrivals=arr, steps_per_sync=sync, keep=keep)
    out = {"workload": f"Poisson {rate}/s for {horizon:.0f} s, grids 300x300-2400x1800, 1800-10800 s, Mw 7-9",
           "batched": serving.latency_stats(rec_b), "gpu_s_in_step_launches": gpu_b,
           "cell_steps": int(sum(p.nx * p.ny * r["steps"] for p, r in zip(probs, rec_b)))}
    if seq:
        rec_s, res_s = serving.run_sequential(probs, arrivals=arr, steps_per_sync=sync, keep=keep)
        out["sequential"] = serving.latency_stats(rec_s)
        err = 0.0
        for i in keep:
            a, b = res_b[i]["max_height"], res_s[i]["max_height"]
            wet = b > -9000
            err = max(err, float(np.max(np.abs(a[wet] - b[wet]))) / max(1e-30, float(np.max(np.abs(b[wet])))))
        out["max_height_rel_err_vs_plain"] = err
    log("serving:", json.dumps(out))
    return out


def precision_phase():
    """Precision study on the 2000x2000 grid (7200 s, Mw 8.5): the FP32 production path
    against the FP64-everywhere run of the same kernels, plus mass drift with walls."""
    from thysea import native, scenario
    prob = scenario.problem("precision")
    log(f"precision: {prob.nx}x{prob.ny}, {prob.t_end * prob.T:.0f} s")
    res = {}
    for prec in (0, 1):
        s = native.Solver(prob, precision=prec)
        t0 = time.perf_counter()
        r = s.run(steps_per_sync=48)
        wall = time.perf_counter() - t0
        h, H, qx, qy = s.state()
        res[prec] = dict(r=r, wall=wall, eta=h - H - prob.hmin, maxh=s.field("max_height"),
                         arr=s.field("arrival_times"))
        s.close()
        log(f"precision: fp{32 if prec == 0 else 64} {r['steps']} steps, {wall:.2f} s")
    a, b = res[0], res[1]
    wet = b["maxh"] > -9000
    peak = float(np.max(np.abs(b["maxh"][wet])))
    eta_peak = float(np.max(np.abs(b["eta"])))
    both = (a["arr"] >= 0) & (b["arr"] >= 0)
    darr = np.abs(a["arr"][both] - b["arr"][both])
    # gauge arrival: first sample with |eta| above 1 cm
    def gauge_arrival(r):
        g, t = r["g_eta"], r["g_time"]
        out = []
        for k in range(g.shape[1]):
            hit = np.nonzero(np.abs(g[:, k]) > 0.01)[0]
            out.append(float(t[hit[0]]) if len(hit) else np.nan)
        return np.array(out)
    ga, gb = gauge_arrival(a["r"]), gauge_arrival(b["r"])
    ok = ~np.isnan(ga) & ~np.isnan(gb)
    walls = scenario.problem("precision", borders=(-1, -1, -1, -1))
    drift = {}
    for prec in (0, 1):
        s = native.Solver(walls, precision=prec)
        s.run(max_steps=1, steps_per_sync=3)  # applies the Okada source, one step
        m0 = s.mass()
        s.run_more(600)
        drift["fp32" if prec == 0 else "fp64"] = abs(s.mass() - m0) / m0
        s.close()
    out = {"workload": f"{prob.nx}x{prob.ny}, {prob.t_end * prob.T:.0f} s, Mw 8.5 (L 300, W 100 km, slip 8 m)",
           "steps": [int(a["r"]["steps"]), int(b["r"]["steps"])],
           "eta_linf_rel_final": float(np.max(np.abs(a["eta"] - b["eta"]))) / eta_peak,
           "max_height_linf_rel": float(np.max(np.abs(a["maxh"][wet] - b["maxh"][wet]))) / peak,
           "wet_mask_equal": bool(np.array_equal(a["maxh"] > -9000, wet)),
           "arrival_time_abs_err_s": {"p50": float(np.percentile(darr, 50)), "p99": float(np.percentile(darr, 99)),
                                      "max": float(darr.max())},
           "gauge_arrival_abs_err_s_max": float(np.max(np.abs(ga[ok] - gb[ok]))) if ok.any() else None,
           "mass_drift_walls_600_steps": drift,
           "fp64_over_fp32_time": b["wall"] / a["wall"]}
    log("precision:", json.dumps(out))
    return out


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
    ap.add_argument("--sync-every", type=int, default=0,
                    help="steps per captured graph replay (0: --steps, up to 960)")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64-everywhere 12M measurement")
    ap.add_argument("--batch-sync", type=int, default=48, help="steps per captured sequence of the batch phases")
    ap.add_argument("--cpu-steps", type=int, default=30)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--cpu-steps-single", type=int, default=2, help="oracle steps timed on one thread (0: skip)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="also run a plain device copy of this size (for counter scaling)")
    ap.add_argument("--batch", type=int, default=256, help="batched scenarios (0: skip)")
    ap.add_argument("--batch-sample", type=int, default=8, help="of them, also run one at a time")
    ap.add_argument("--serve-seconds", type=float, default=30.0, help="serving arrival window (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=2.0)
    ap.add_argument("--no-serve-sequential", action="store_true")
    ap.add_argument("--saturated-rate", type=float, default=16.0,
                    help="arrival rate of the saturated serving run, above what one at a time sustains (0: skip)")
    ap.add_argument("--saturated-seconds", type=float, default=10.0)
    ap.add_argument("--no-precision", action="store_true", help="skip the FP32-vs-FP64 study")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: the time-stepping path runs on one GPU only (asked for --gpus {args.gpus}); "
              "sharding across GPUs is not implemented", file=sys.stderr)
        sys.exit(2)

    from thysea import native, scenario
    if native.probe() != 0:
        raise SystemExit("bench.py: no gfx950 device visible")
    log("generating workload", args.nx, "x", args.ny)
    prob = scenario.problem("large", nx=args.nx, ny=args.ny)
    # the configured 1800 s is ~500 steps; keep stepping past it when --warmup + --steps
    # asks for more (about 4 s per step; the gauge buffers are sized by this end time)
    prob.t_end = max(prob.t_end, (args.warmup + args.steps + 100) * 10.0 / prob.T)
    track = not args.no_arrival
    tile = tuple(int(v) for v in args.tile.split("x"))
    ncell = prob.nx * prob.ny

    copy = None
    if args.copy_probe_mib > 0:
        ms, gbs = native.copy_probe(args.copy_probe_mib << 20, reps=5)
        copy = {"bytes": args.copy_probe_mib << 20, "ms": ms, "GBs_read_plus_write": gbs}
        log(f"copy probe {args.copy_probe_mib} MiB: {ms:.3f} ms, {gbs:.0f} GB/s")

    # one captured sequence of exactly --steps steps (any count: the library keeps one
    # sequence per launch slot it can start at), so no idle launch enters the window
    sync = args.sync_every if args.sync_every > 0 else (args.steps if args.steps <= 960 else 48)
    solver = native.Solver(prob, precision=0, track_arrival=track, tile=tile, kernel=args.kernel,
                           strip_rows=args.strip_rows)
    log("warmup", args.warmup, "steps (includes Okada at t=0)")
    w = solver.run(max_steps=args.warmup, steps_per_sync=sync)
    log("warmup done: t =", w["t_final"] * prob.T, "s")
    tuned = None
    if args.kernel == "sweep" and args.strip_rows == 0:
        rows, tms = solver.autotune(reps=9)  # candidates differ by 1-2 %: a median of 9 replays
        tuned = {"strip_rows": rows, "general_rows": solver.general_rows, "ms_per_step": tms}
        log(f"autotune: {rows}-row strips ({solver.general_rows}-row general strips), {tms:.4f} ms per step")
    # capture and upload the step sequences of the tuned strips before the clock starts
    # (no step is taken: the run is at its step limit)
    solver.run_more(0)
    t0 = time.perf_counter()
    r = solver.run_more(args.steps)
    wall = time.perf_counter() - t0
    done = r["steps"] - w["steps"]
    if done != args.steps:
        raise SystemExit(f"bench.py: ran {done} steps, expected {args.steps}")
    gpu_s = r["seconds"]
    value = ncell * args.steps / wall / 1e9
    kern_ms = gpu_s * 1e3 / args.steps
    achieved = B_ALG[track] * value  # GB/s: algorithmic bytes per cell-step at the measured rate
    log(f"timed {args.steps} steps: wall {wall*1e3:.2f} ms, gpu {gpu_s*1e3:.2f} ms, {value:.1f} Gcell-steps/s")

    traffic = None
    tp = os.path.join(ROOT, "profiles", TRAFFIC_FILE[args.kernel])
    if os.path.exists(tp):
        with open(tp) as f:
            traffic = json.load(f).get("bytes_per_step")

    # regime of the cells at the end of the timed steps: wet, and wet with h <= hpos (the
    # WAF branch, friction and the near-dry filter act there)
    h, _, _, _ = solver.state()
    wet = h > prob.eps_h
    regime = {"wet_fraction": float(np.mean(wet)),
              "wet_h_le_hpos_fraction": float(np.mean(wet & (h <= prob.hpos)))}
    del h, wet
    solver.close()

    # the same workload at the reference's precision (FP64 state and arithmetic)
    fp64 = None
    if not args.no_fp64:
        s64 = native.Solver(prob, precision=1, track_arrival=track)
        w64 = s64.run(max_steps=args.warmup, steps_per_sync=sync)
        tuned64 = None
        if args.kernel == "sweep" and args.strip_rows == 0:
            rows64, tms64 = s64.autotune(reps=5)  # FP64 has its own best strip heights
            tuned64 = {"strip_rows": rows64, "general_rows": s64.general_rows, "ms_per_step": tms64}
        s64.run_more(0)
        t0 = time.perf_counter()
        r64 = s64.run_more(args.steps)
        wall64 = time.perf_counter() - t0
        assert r64["steps"] - w64["steps"] == args.steps
        v64 = ncell * args.steps / wall64 / 1e9
        fp64 = {"value": v64, "unit": "Gcell-steps/s", "ms_per_step": wall64 * 1e3 / args.steps,
                "bytes_per_cell_step": 2 * B_ALG[track], "roofline_frac": 2 * B_ALG[track] * v64 / HBM_PEAK_GBS,
                "kernel": "k_sweep_deep<double> + k_sweep<double> (general, hand-back)", "autotune": tuned64}
        log(f"fp64: {v64:.1f} Gcell-steps/s")
        s64.close()

    cpu = None
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or (os.cpu_count() or 1)
        log("cpu baseline:", args.cpu_steps, "oracle steps on", threads, "threads")
        gc, sec = cpu_baseline(prob, args.cpu_steps, threads)
        cpu = {"value": gc, "unit": "Gcell-steps/s", "cores": threads, "kind": "port",
               "sample": f"{args.cpu_steps} steps of the same {prob.nx}x{prob.ny} grid after Okada "
                         f"({sec:.1f} s), FP64 oracle, OpenMP on every host core",
               "cpu_model": cpu_model()}
        if args.cpu_steps_single > 0:
            g1, s1 = cpu_baseline(prob, args.cpu_steps_single, 1)
            cpu["single_thread"] = {"value": g1, "cores": 1,
                                    "sample": f"{args.cpu_steps_single} steps of the same grid ({s1:.1f} s)"}
        cpu["small_workload_full_run"] = cpu_full_run(threads)
        log("cpu: 240x200 full run", json.dumps(cpu["small_workload_full_run"]))
    batched = batched_phase(args.batch, args.batch_sample, args.batch_sync) if args.batch > 0 else None
    serve = (serving_phase(args.serve_rate, args.serve_seconds, args.batch_sync, not args.no_serve_sequential)
             if args.serve_seconds > 0 else None)
    saturated = (serving_phase(args.saturated_rate, args.saturated_seconds, args.batch_sync,
                               not args.no_serve_sequential)
                 if args.saturated_rate > 0 and args.saturated_seconds > 0 else None)
    prec = None if args.no_precision else precision_phase()
    out = {
        "metric": METRIC, "value": value, "unit": "Gcell-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{prob.nx}x{prob.ny} lon/lat grid ({ncell/1e6:.1f}M cells, 2 arc-min), "
                               "4 Okada subfaults, 64 gauges, open borders",
                   "cells": ncell, "track_arrival": track, "steps_per_sync": sync,
                   "step_kernel": args.kernel, "autotune": tuned, "cell_regime": regime,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": KERNEL_NAME[args.kernel], "kernel_ms": kern_ms,
                     "kernel_ms_basis": "HIP events around the graph replay on the library stream, per step",
                     "bytes_per_cell_step": B_ALG[track]},
        "fp64": fp64,
        "cpu_baseline": cpu,
    }
    if batched:
        out["batched"] = batched
    if prec:
        out["precision_study"] = prec
    if serve:
        out["serving"] = serve
    if saturated:
        out["serving_saturated"] = saturated
    if copy:
        out["copy_probe"] = copy
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
