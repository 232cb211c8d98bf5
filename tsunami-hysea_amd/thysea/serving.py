"""Python driver of the request loop (include/thysea_serve.h, libthysea_serve.so).

The loop itself is C++ (tsunami-hysea_amd/host/serve.cpp): scenarios are admitted as
they arrive and as the HBM budget allows, their state stays resident, and every
running scenario advances in each captured step sequence; finished ones hand back max
height, arrival times and gauges and free their slot.  The comparison serves the same
requests one at a time through thy_run.  This module only turns Problem objects into
request descriptions and reads the results back.

Time-to-result of a request = from its arrival to its max-height, arrival-time and
gauge arrays in host memory (input generation happens before the clock starts, as
reading input files would)."""
import ctypes as C

import numpy as np

from . import native


class _Requests:
    """thy_request array for ``problems`` with the arrays it points to kept alive, and
    result buffers for the requests in ``keep``."""

    def __init__(self, problems, arrivals, keep, steps_per_sync, solver_kw):
        n = len(problems)
        self.n = n
        self.req = (native.Request * max(1, n))()
        self.out = (native.RequestResult * max(1, n))()
        self.alive = []
        self.bufs = {}
        for i, pr in enumerate(problems):
            q = self.req[i]
            q.arrival_s = float(arrivals[i])
            geo = [np.ascontiguousarray(a, np.float64) for a in (pr.lon, pr.lat, pr.area, pr.cosphi, pr.ancho, pr.alto)]
            q.grid = native.Grid(pr.nx, pr.ny, *[a.ctypes.data_as(native._dp) for a in geo])
            q.params = native.params_of(pr, **(solver_kw or {}))
            # the state as four arrays (shared by requests on one grid), paired by the
            # loop's worker threads
            soa = [np.ascontiguousarray(a, np.float64).ravel() for a in (pr.h, pr.H, pr.qx, pr.qy)]
            q.h, q.H, q.qx, q.qy = [a.ctypes.data_as(native._dp) for a in soa]
            faults = sorted(pr.faults, key=lambda f: f["time"])
            farr = (native.Fault * max(1, len(faults)))(*[native.fault_of(f) for f in faults])
            gpos = np.ascontiguousarray(pr.gauge_pos if pr.gauge_pos is not None else np.zeros(0), np.int32)
            c = q.run
            c.t_end, c.save_interval, c.series_interval = pr.t_end, -1.0, pr.series_interval
            c.n_faults, c.faults = len(faults), farr
            c.crop_flag, c.crop_value = pr.crop_flag, pr.crop_value
            c.n_gauges, c.gauge_pos = len(gpos), gpos.ctypes.data_as(native._ip)
            c.max_steps, c.steps_per_sync = 0, int(steps_per_sync)
            self.alive.append((geo, soa, farr, gpos))
            if i in keep:
                ng = len(gpos)
                cap = int(pr.t_end / pr.series_interval) + 8 if (pr.series_interval > 0 and ng) else 0
                b = {"max_height": np.empty((pr.ny, pr.nx), np.float32),
                     "arrival_times": np.empty((pr.ny, pr.nx), np.float32),
                     "g_time": np.zeros(max(1, cap), np.float32),
                     "g_eta": np.zeros((max(1, cap), max(1, ng)), np.float32),
                     "g_u": np.zeros((max(1, cap), max(1, ng)), np.float32),
                     "g_v": np.zeros((max(1, cap), max(1, ng)), np.float32)}
                o = self.out[i]
                for k in b:
                    setattr(o, k, b[k].ctypes.data_as(native._fp))
                o.gauge_capacity = cap
                self.bufs[i] = (b, ng)

    def records(self, arrivals):
        rec = []
        for i in range(self.n):
            o = self.out[i]
            if o.status != 0:
                raise native.ThyError(f"request {i}: {native.STATUS.get(o.status, o.status)}", o.status)
            rec.append(dict(index=i, arrival=float(arrivals[i]), admitted=o.admitted_s, finished=o.finished_s,
                            steps=int(o.steps)))
        return rec

    def results(self):
        out = {}
        for i, (b, ng) in self.bufs.items():
            ns = min(self.out[i].n_samples, self.out[i].gauge_capacity)
            out[i] = {"max_height": b["max_height"], "arrival_times": b["arrival_times"],
                      "steps": int(self.out[i].steps), "n_samples": ns, "g_time": b["g_time"][:ns],
                      "g_eta": b["g_eta"][:ns, :ng], "g_u": b["g_u"][:ns, :ng], "g_v": b["g_v"][:ns, :ng]}
        return out


def _options(capacity=512, hbm_budget=0, steps_per_sync=48, workers=8, burst=8):
    o = native.ServeOptions()
    o.capacity, o.hbm_budget, o.steps_per_sync = int(capacity), int(hbm_budget), int(steps_per_sync)
    o.workers, o.burst = int(workers), int(burst)
    return o


def run_batched(problems, arrivals=None, capacity=512, hbm_budget=0, steps_per_sync=48, keep=None, burst=8,
                workers=8, solver_kw=None):
    """Serve ``problems`` (arrival times in seconds, default all at 0) through one batch
    with the C++ request loop (thy_serve_batched).  Returns per-request records
    (arrival, admitted, finished, steps), the results of the requests in ``keep`` (max
    height, arrival times, gauges) and the GPU seconds of the batch's step launches.
    ``solver_kw`` sets the kernel parameters of every request (strip heights, for one)."""
    n = len(problems)
    arrivals = np.zeros(n) if arrivals is None else np.asarray(arrivals, np.float64)
    rq = _Requests(problems, arrivals, set(keep or ()), steps_per_sync, solver_kw)
    gpu = C.c_double()
    opt = _options(capacity, hbm_budget, steps_per_sync, workers, burst)
    rc = native.serve_lib().thy_serve_batched(rq.req, n, C.byref(opt), rq.out, C.byref(gpu))
    if rc != 0:
        raise native.ThyError(f"thy_serve_batched: {native.STATUS.get(rc, rc)}", rc)
    return rq.records(arrivals), rq.results(), gpu.value


def run_sequential(problems, arrivals=None, steps_per_sync=48, keep=None, solver_kw=None):
    """The same requests one at a time through thy_run, first come first served
    (thy_serve_sequential)."""
    n = len(problems)
    arrivals = np.zeros(n) if arrivals is None else np.asarray(arrivals, np.float64)
    rq = _Requests(problems, arrivals, set(keep or ()), steps_per_sync, solver_kw)
    opt = _options(steps_per_sync=steps_per_sync)
    rc = native.serve_lib().thy_serve_sequential(rq.req, n, C.byref(opt), rq.out)
    if rc != 0:
        raise native.ThyError(f"thy_serve_sequential: {native.STATUS.get(rc, rc)}", rc)
    return rq.records(arrivals), rq.results()


def tune_batch(problems, candidates=(24, 32, 40, 48), n=64, launches=2, steps_per_sync=48):
    """Pick the strip height for batched runs of scenarios like ``problems``: for each
    candidate, a batch of the first ``n`` of them is advanced by one step sequence
    (capture, faults at t = 0), then by ``launches`` more that are timed.  Returns the
    strip height with the least GPU time per scenario-step and {candidate: us per
    scenario-step}.  Many scenarios side by side keep every CU busy, so the batch
    favours taller strips (fewer halo rows) than one scenario alone, whose short
    strips shorten the step's critical path; thy_autotune's choice for one scenario is
    therefore not the batch's.  The tuning batch holds 64 scenarios: nearer the 60-140
    that a large batch runs at once than 32, on which the order of the candidates
    differs (256 scenarios of 1000x800 on one MI355X: 32-row strips cheapest on 64,
    48-row on 32; the whole batch within 3 % at every height from 24 to 48)."""
    cost = {}
    for rows in candidates:
        b = native.Batch(n, steps_per_sync=steps_per_sync)
        solvers = [native.Solver(p, strip_rows=rows) for p in problems[:n]]
        try:
            slots = [b.add(s, steps_per_sync=steps_per_sync) for s in solvers]
            if any(k is None for k in slots):
                raise native.ThyError("the tuning batch does not fit the HBM budget")
            b.advance()

            def totals():
                return b.info()["gpu_seconds"], sum(b.status(k)[1]["steps"] for k in slots)

            g0, n0 = totals()
            for _ in range(launches):
                b.advance()
            g1, n1 = totals()
            if n1 > n0:
                cost[rows] = (g1 - g0) / (n1 - n0) * 1e6
        finally:
            b.close()
            for s in solvers:
                s.close()
    if not cost:
        raise native.ThyError("no candidate strip height advanced a scenario")
    return min(cost, key=cost.get), cost


def latency_stats(rec):
    lat = np.array([r["finished"] - r["arrival"] for r in rec])
    span = max(r["finished"] for r in rec) - min(r["arrival"] for r in rec)
    return {"n": len(rec), "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "makespan_s": float(span),
            "scenarios_per_hour": float(len(rec) / span * 3600.0) if span > 0 else None}
