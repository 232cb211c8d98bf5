"""Request loop over the batched C-ABI (thy_batch_*): scenarios are admitted as they
arrive and as the HBM budget allows, their state stays resident, and every running
scenario advances in each captured step sequence; finished ones hand back max
height, arrival times and gauges and free their slot.  A plain loop running the same
requests one at a time through thy_run is the comparison.

Time-to-result of a request = from its arrival to its max-height, arrival-time and
gauge arrays in host memory (input generation happens before the clock starts, as
reading input files would)."""
import time

import numpy as np

from . import native


def _result(r):
    return {"max_height": r["max_height"], "arrival_times": r["arrival_times"], "g_eta": r["g_eta"],
            "steps": r["steps"]}


def run_batched(problems, arrivals=None, capacity=512, hbm_budget=0, steps_per_sync=48, keep=None, burst=8):
    """Serve ``problems`` (arrival times in seconds, default all at 0) through one
    Batch.  Returns per-request records and, for indices in ``keep``, the results.

    The host work of a request (context set-up, upload, admission; read-out and
    release once finished) runs while the GPU steps the batch: each step sequence is
    launched asynchronously and the host works through finished and arrived requests
    until it has completed (at most ``burst`` admissions while the GPU is idle, so
    that stepping starts early)."""
    n = len(problems)
    arrivals = np.zeros(n) if arrivals is None else np.asarray(arrivals, np.float64)
    order = list(np.argsort(arrivals, kind="stable"))
    keep = set(keep or ())
    rec = [dict(arrival=float(arrivals[i]), admitted=None, finished=None, steps=0) for i in range(n)]
    results = {}
    b = native.Batch(capacity, hbm_budget=hbm_budget, steps_per_sync=steps_per_sync)
    live = {}      # slot -> (request index, solver)
    pending = []   # finished slots not yet read out
    t0 = time.perf_counter()
    nxt = 0
    done = 0
    in_flight = False

    def admit():
        i = order[nxt]
        s = native.Solver(problems[i])
        slot = b.add(s, steps_per_sync=steps_per_sync)
        if slot is None:  # budget or slots full: wait for a finisher
            s.close()
            if not live:
                raise native.ThyError(f"request {i} does not fit the batch's HBM budget")
            return False
        live[slot] = (i, s)
        rec[i]["admitted"] = time.perf_counter() - t0
        return True

    def retire(slot):
        i, s = live.pop(slot)
        r = b.result(slot)
        rec[i]["finished"] = time.perf_counter() - t0
        rec[i]["steps"] = int(r["steps"])
        if i in keep:
            results[i] = _result(r)
        b.remove(slot)
        s.close()

    while done < n:
        k = 0
        while True:
            if in_flight:
                if b.ready():
                    break
            elif k >= burst:
                break
            if pending:
                retire(pending.pop(0))
                done += 1
            elif nxt < n and arrivals[order[nxt]] <= time.perf_counter() - t0:
                if not admit():
                    break
                nxt += 1
                k += 1
            else:
                break
        if in_flight:
            b.wait()
            in_flight = False
            pending += [slot for slot in live if slot not in pending and b.status(slot)[0] == native.SLOT_FINISHED]
        if any(b.status(slot)[0] == native.SLOT_RUNNING for slot in live):
            b.launch()
            in_flight = True
        elif not pending and nxt < n:
            now = time.perf_counter() - t0
            time.sleep(max(0.0, min(0.01, arrivals[order[nxt]] - now)))
    gpu = b.info()["gpu_seconds"]
    b.close()
    return rec, results, gpu


def run_sequential(problems, arrivals=None, steps_per_sync=48, keep=None):
    """The same requests one at a time through thy_run (first come, first served)."""
    n = len(problems)
    arrivals = np.zeros(n) if arrivals is None else np.asarray(arrivals, np.float64)
    keep = set(keep or ())
    rec, results = [], {}
    t0 = time.perf_counter()
    for i in np.argsort(arrivals, kind="stable"):
        now = time.perf_counter() - t0
        if arrivals[i] > now:
            time.sleep(arrivals[i] - now)
        s = native.Solver(problems[i])
        start = time.perf_counter() - t0
        r = s.run(steps_per_sync=steps_per_sync)
        out = {"max_height": s.field("max_height"), "arrival_times": s.field("arrival_times"),
               "g_eta": r["g_eta"], "steps": r["steps"]}
        s.close()
        rec.append(dict(index=int(i), arrival=float(arrivals[i]), admitted=start,
                        finished=time.perf_counter() - t0, steps=int(r["steps"])))
        if i in keep:
            results[int(i)] = out
    rec.sort(key=lambda x: x["index"])
    return rec, results


def latency_stats(rec):
    lat = np.array([r["finished"] - r["arrival"] for r in rec])
    span = max(r["finished"] for r in rec) - min(r["arrival"] for r in rec)
    return {"n": len(rec), "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "makespan_s": float(span),
            "scenarios_per_hour": float(len(rec) / span * 3600.0) if span > 0 else None}
