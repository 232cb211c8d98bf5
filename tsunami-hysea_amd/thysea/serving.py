"""Request loop over the batched C-ABI (thy_batch_*): scenarios are admitted as they
arrive and as the HBM budget allows, their state stays resident, and every running
scenario advances in each captured step sequence; finished ones hand back max
height, arrival times and gauges and free their slot.  A plain loop running the same
requests one at a time through thy_run is the comparison.

Time-to-result of a request = from its arrival to its max-height, arrival-time and
gauge arrays in host memory (input generation happens before the clock starts, as
reading input files would)."""
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import native


def _result(r):
    return {"max_height": r["max_height"], "arrival_times": r["arrival_times"], "g_eta": r["g_eta"],
            "steps": r["steps"]}


def run_batched(problems, arrivals=None, capacity=512, hbm_budget=0, steps_per_sync=48, keep=None, burst=8,
                workers=8, solver_kw=None):
    """Serve ``problems`` (arrival times in seconds, default all at 0) through one
    Batch.  Returns per-request records and, for indices in ``keep``, the results.

    The host work of a request runs while the GPU steps the batch: each step sequence
    is launched asynchronously and the loop admits arrived requests and retires
    finished ones until it has completed (at most ``burst`` admissions while the GPU is
    idle, so that stepping starts early).  The work that touches one context only
    (creating and uploading it, preparing its run with the faults due at t = 0;
    reading out its max height, arrival times and gauges and releasing it) runs on
    ``workers`` threads; the batch itself is driven from this thread.  ``solver_kw`` goes to every native.Solver (strip heights, for one)."""
    n = len(problems)
    arrivals = np.zeros(n) if arrivals is None else np.asarray(arrivals, np.float64)
    order = list(np.argsort(arrivals, kind="stable"))
    keep = set(keep or ())
    rec = [dict(arrival=float(arrivals[i]), admitted=None, finished=None, steps=0) for i in range(n)]
    results = {}
    b = native.Batch(capacity, hbm_budget=hbm_budget, steps_per_sync=steps_per_sync)
    pool = ThreadPoolExecutor(max_workers=max(1, workers))
    live = {}        # slot -> (request index, solver)
    pending = []     # finished slots not yet retired
    prepared = {}    # request index -> future of its created and uploaded solver
    readouts = []    # (request index, future of its results)
    t0 = time.perf_counter()
    state = {"nxt": 0, "nsub": 0, "done": 0}

    def now():
        return time.perf_counter() - t0

    def prepare(i):
        s = native.Solver(problems[i], **(solver_kw or {}))
        s.prepare_batch(steps_per_sync=steps_per_sync)
        return s

    def readout(s, n_samples, steps):
        out = {"max_height": s.field("max_height"), "arrival_times": s.field("arrival_times"), "steps": steps}
        out.update(s.gauges(n_samples))
        s.close()
        return out, now()

    def submit_prepares():
        # keep up to two requests per worker in preparation, in arrival order
        while (state["nsub"] < n and arrivals[order[state["nsub"]]] <= now()
               and state["nsub"] - state["nxt"] < 2 * max(1, workers)):
            i = order[state["nsub"]]
            prepared[i] = pool.submit(prepare, i)
            state["nsub"] += 1

    def admit(block):
        """Admit the next request: True, False (budget or slots full), None (not prepared yet)."""
        i = order[state["nxt"]]
        f = prepared.get(i)
        if f is None or (not block and not f.done()):
            return None
        s = f.result()
        slot = b.add(s, steps_per_sync=steps_per_sync)
        if slot is None:
            if not live:
                s.close()
                raise native.ThyError(f"request {i} does not fit the batch's HBM budget")
            return False
        del prepared[i]
        live[slot] = (i, s)
        rec[i]["admitted"] = now()
        state["nxt"] += 1
        return True

    def retire(slot):
        i, s = live.pop(slot)
        r = b.status(slot)[1]
        b.remove(slot)
        readouts.append((i, pool.submit(readout, s, r["n_samples"], r["steps"])))

    def collect(block=False):
        rest = []
        for i, f in readouts:
            if block or f.done():
                out, t_done = f.result()
                rec[i]["finished"] = t_done
                rec[i]["steps"] = int(out["steps"])
                if i in keep:
                    results[i] = _result(out)
                state["done"] += 1
            else:
                rest.append((i, f))
        readouts[:] = rest

    in_flight = False
    try:
        while state["done"] < n:
            submit_prepares()
            k = 0
            while True:
                collect()
                if in_flight:
                    if b.ready():
                        break
                elif k >= burst:
                    break
                if pending:
                    retire(pending.pop(0))
                    continue
                if state["nxt"] < n and arrivals[order[state["nxt"]]] <= now():
                    ok = admit(block=not in_flight)
                    if ok:
                        k += 1
                        submit_prepares()
                        continue
                    if ok is False:
                        break
                    submit_prepares()   # it arrived after the last submission
                    time.sleep(0.0002)  # or it is still being prepared
                    continue
                break
            if in_flight:
                b.wait()
                in_flight = False
                pending += [slot for slot in live if slot not in pending and b.status(slot)[0] == native.SLOT_FINISHED]
            if any(b.status(slot)[0] == native.SLOT_RUNNING for slot in live):
                b.launch()
                in_flight = True
            elif not pending:
                if readouts:
                    collect(block=True)
                elif state["nxt"] < n:
                    time.sleep(max(0.0, min(0.01, arrivals[order[state["nxt"]]] - now())))
    finally:
        pool.shutdown(wait=True)
    gpu = b.info()["gpu_seconds"]
    b.close()
    return rec, results, gpu


def tune_batch(problems, candidates=(24, 32, 40, 48), n=32, launches=2, steps_per_sync=48):
    """Pick the strip height for batched runs of scenarios like ``problems``: for each
    candidate, a batch of the first ``n`` of them is advanced by one step sequence
    (capture, faults at t = 0), then by ``launches`` more that are timed.  Returns the
    strip height with the least GPU time per scenario-step and {candidate: us per
    scenario-step}.  Many scenarios side by side keep every CU busy, so the batch
    favours taller strips (fewer halo rows) than one scenario alone, whose short
    strips shorten the step's critical path; thy_autotune's choice for one scenario is
    therefore not the batch's."""
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


def run_sequential(problems, arrivals=None, steps_per_sync=48, keep=None, solver_kw=None):
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
        s = native.Solver(problems[i], **(solver_kw or {}))
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
