"""Host logic of the request loop (thysea/serving.py) on the CPU.

The loop is driven against a simulated batch that keeps the contract of
thy_batch_* (include/thysea.h): a scenario admitted while a launch is in flight joins
the next launch; add returns no slot when the slots or the HBM budget are full; the
batch is driven from one thread while contexts are created and read out on worker
threads.  Every request must finish with its own results, in arrival order of
admission, never above the budget.
"""
import threading
import time
import types

import numpy as np
import pytest

from thysea import serving

RUNNING, FINISHED = 1, 2


class FakeError(RuntimeError):
    pass


class FakeSolver:
    made = []
    prepared_on = []

    def __init__(self, prob, strip_rows=0, general_rows=0):
        time.sleep(0.001)  # set-up and upload take a while
        self.prob = prob
        self.rows = strip_rows
        self.steps = 0
        self.closed = False
        self.batch_prepared = False
        FakeSolver.made.append(self)

    def prepare_batch(self, steps_per_sync=0, max_steps=0):
        FakeSolver.prepared_on.append(threading.get_ident())
        self.batch_prepared = True

    def field(self, name):
        off = 0.0 if name == "max_height" else 0.5
        return np.full((3, 2), self.prob["id"] + off, np.float32)

    def gauges(self, n_samples):
        return {"n_samples": n_samples, "g_eta": np.full((n_samples, 2), self.prob["id"], np.float32)}

    def run(self, steps_per_sync=0):
        self.steps = self.prob["steps"]
        return {"steps": self.steps, "g_eta": self.gauges(self.steps // 10)["g_eta"]}

    def close(self):
        assert not self.closed
        self.closed = True


class FakeBatch:
    """Slots advance by K steps in each launch they were admitted before."""

    def __init__(self, capacity, hbm_budget=0, steps_per_sync=48):
        self.cap, self.budget, self.K = capacity, hbm_budget, steps_per_sync
        self.slots = {}
        self.used = 0
        self.max_used = 0
        self.max_live = 0
        self.flight = None
        self.polls = 0
        self.thread = threading.get_ident()
        self.admitted = []
        self.unprepared = 0
        self.gpu = 0.0

    def _main(self):
        assert threading.get_ident() == self.thread, "the batch is driven from one thread"

    def add(self, s, steps_per_sync=0, max_steps=0):
        self._main()
        assert not s.closed
        if len(self.slots) >= self.cap or (self.budget and self.used + s.prob["bytes"] > self.budget):
            return None
        slot = min(set(range(self.cap)) - set(self.slots))
        self.slots[slot] = [s, RUNNING]
        self.unprepared += 0 if s.batch_prepared else 1
        s.batch_prepared = False
        self.used += s.prob["bytes"]
        self.max_used = max(self.max_used, self.used)
        self.max_live = max(self.max_live, len(self.slots))
        self.admitted.append(s.prob["id"])
        return slot

    def launch(self):
        self._main()
        assert self.flight is None
        self.flight = [k for k, v in self.slots.items() if v[1] == RUNNING]
        assert self.flight, "a launch with nothing running"
        self.polls = 2

    def ready(self):
        self._main()
        if self.flight is None:
            return True
        self.polls -= 1
        return self.polls <= 0

    def wait(self):
        self._main()
        for k in self.flight:
            s = self.slots[k][0]
            n0 = s.steps
            s.steps = min(s.prob["steps"], s.steps + self.K)
            self.gpu += (s.steps - n0) * (abs(s.rows - 40) + 1) * 1e-6  # 40-row strips are the fastest
            if s.steps == s.prob["steps"]:
                self.slots[k][1] = FINISHED
        self.flight = None

    def status(self, slot):
        self._main()
        s, st = self.slots[slot]
        return st, {"steps": s.steps, "n_samples": s.steps // 10}

    def remove(self, slot):
        self._main()
        assert self.flight is None or slot not in self.flight
        s, _ = self.slots.pop(slot)
        self.used -= s.prob["bytes"]
        return s

    def advance(self):
        self.launch()
        self.wait()
        return sum(v[1] == RUNNING for v in self.slots.values())

    def info(self):
        return {"gpu_seconds": self.gpu}

    def close(self):
        self._main()


@pytest.fixture
def fake(monkeypatch):
    batches = []

    class Batch(FakeBatch):
        def __init__(self, *a, **k):
            super().__init__(*a, **k)
            batches.append(self)

    ns = types.SimpleNamespace(Batch=Batch, Solver=FakeSolver, ThyError=FakeError, SLOT_FREE=0,
                               SLOT_RUNNING=RUNNING, SLOT_FINISHED=FINISHED)
    monkeypatch.setattr(serving, "native", ns)
    return batches


def _problems(n, seed=3):
    rng = np.random.default_rng(seed)
    return [dict(id=i, steps=int(rng.integers(1, 400)), bytes=int(rng.integers(1, 10))) for i in range(n)]


def _check_results(probs, rec, res):
    for i, p in enumerate(probs):
        r = res[i]
        assert r["steps"] == p["steps"] == rec[i]["steps"]
        assert np.all(r["max_height"] == i) and np.all(r["arrival_times"] == i + 0.5)
        assert r["g_eta"].shape == (p["steps"] // 10, 2) and np.all(r["g_eta"] == i)
        assert rec[i]["arrival"] <= rec[i]["admitted"] <= rec[i]["finished"]


@pytest.mark.parametrize("workers", [1, 4])
def test_every_request_finishes_with_its_own_results(fake, workers):
    FakeSolver.prepared_on.clear()
    probs = _problems(40)
    rec, res, _ = serving.run_batched(probs, capacity=6, steps_per_sync=48, keep=range(40), workers=workers)
    _check_results(probs, rec, res)
    b = fake[0]
    assert b.max_live <= 6 and not b.slots
    assert b.admitted == list(range(40))  # admitted in arrival order
    # runs are prepared on the worker threads, never on the thread that drives the batch
    assert len(FakeSolver.prepared_on) == 40 and b.thread not in FakeSolver.prepared_on
    assert b.unprepared == 0


def test_budget_bounds_the_resident_scenarios(fake):
    probs = [dict(id=i, steps=50 + 7 * i, bytes=4) for i in range(12)]
    rec, res, _ = serving.run_batched(probs, capacity=64, hbm_budget=9, steps_per_sync=24, keep=range(12))
    _check_results(probs, rec, res)
    assert fake[0].max_used <= 9 and fake[0].max_live == 2


def test_arrivals_are_respected(fake):
    probs = _problems(10, seed=5)
    arrivals = np.linspace(0.0, 0.05, 10)[::-1].copy()  # last request first
    rec, res, _ = serving.run_batched(probs, arrivals=arrivals, capacity=4, keep=range(10))
    _check_results(probs, rec, res)
    assert fake[0].admitted == list(range(10))[::-1]
    for i in range(10):
        assert rec[i]["admitted"] >= arrivals[i]


def test_a_request_above_the_budget_is_an_error(fake):
    probs = [dict(id=0, steps=10, bytes=20)]
    with pytest.raises(FakeError, match="does not fit"):
        serving.run_batched(probs, hbm_budget=8)


def test_sequential_loop_returns_the_same_results(fake):
    probs = _problems(6, seed=9)
    rec_b, res_b, _ = serving.run_batched(probs, capacity=3, keep=range(6))
    rec_s, res_s = serving.run_sequential(probs, keep=range(6))
    for i in range(6):
        assert rec_s[i]["steps"] == rec_b[i]["steps"]
        for k in ("max_height", "arrival_times", "g_eta"):
            assert np.array_equal(res_s[i][k], res_b[i][k])


def test_batch_tuner_picks_the_cheapest_strip_height(fake):
    FakeSolver.made.clear()
    probs = [dict(id=i, steps=1000, bytes=1) for i in range(40)]
    rows, cost = serving.tune_batch(probs, candidates=(24, 32, 40, 48), n=8, launches=2, steps_per_sync=12)
    assert rows == 40 and sorted(cost) == [24, 32, 40, 48]
    assert cost[40] == pytest.approx(1.0) and cost[24] == pytest.approx(17.0)
    assert len(FakeSolver.made) == 32 and all(s.closed for s in FakeSolver.made)
