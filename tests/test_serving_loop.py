"""Host logic of the C++ request loop (tsunami-hysea_amd/host/serve.cpp,
include/thysea_serve.h) on the CPU.

The loop is compiled here against a CPU stand-in of libthysea.so
(tests/fake_thysea/fake_thysea.cpp) that keeps the contract of the context and batch
entry points it calls: a scenario admitted while a launch is in flight joins the next
launch; add refuses when the slots or the HBM budget are full; the batch must be driven
from one thread while contexts are set up and read out on worker threads.  Every
request must finish with its own results, admitted in arrival order, never above the
budget; runs are prepared on the worker threads.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

PKG = os.path.join(ROOT, "tsunami-hysea_amd")


@pytest.fixture(scope="module")
def loop(tmp_path_factory):
    d = tmp_path_factory.mktemp("fakeloop")
    fake = d / "libthysea_fake.so"
    serve = d / "libthysea_fakeserve.so"
    subprocess.run(["g++", "-O1", "-std=c++17", "-fPIC", "-shared", "-pthread",
                    os.path.join(ROOT, "tests", "fake_thysea", "fake_thysea.cpp"), "-o", str(fake)], check=True)
    subprocess.run(["g++", "-O1", "-std=c++17", "-fPIC", "-shared", "-pthread", os.path.join(PKG, "host", "serve.cpp"),
                    "-o", str(serve), f"-L{d}", "-lthysea_fake", f"-Wl,-rpath,{d}"], check=True)
    from thysea import native
    F = C.CDLL(str(fake))
    S = C.CDLL(str(serve))
    S.thy_serve_batched.argtypes = [C.POINTER(native.Request), C.c_int, C.POINTER(native.ServeOptions),
                                    C.POINTER(native.RequestResult), C.POINTER(C.c_double)]
    S.thy_serve_sequential.argtypes = [C.POINTER(native.Request), C.c_int, C.POINTER(native.ServeOptions),
                                       C.POINTER(native.RequestResult)]
    F.fake_stats.argtypes = [C.POINTER(C.c_double), C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_size_t),
                             C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
    return native, F, S


class Reqs:
    """Requests of the stand-in: id in lon[0], steps in t_end, footprint nx * ny bytes
    with nx = 2."""

    def __init__(self, native, specs, arrivals=None):
        n = len(specs)
        self.n = n
        self.req = (native.Request * n)()
        self.out = (native.RequestResult * n)()
        self.keep = []
        self.res = []
        self.specs = specs
        for i, (steps, ny) in enumerate(specs):
            q = self.req[i]
            q.arrival_s = 0.0 if arrivals is None else float(arrivals[i])
            lon = np.array([float(i), float(i) + 1.0])
            geo = np.zeros(ny + 1)
            q.grid = native.Grid(2, ny, lon.ctypes.data_as(native._dp), *[geo.ctypes.data_as(native._dp)] * 5)
            st = np.zeros(2 * ny)
            q.h = q.H = q.qx = q.qy = st.ctypes.data_as(native._dp)
            q.run.t_end = float(steps)
            gpos = np.zeros(2, np.int32)
            q.run.n_gauges, q.run.gauge_pos = 2, gpos.ctypes.data_as(native._ip)
            ns = steps // 10 + 1
            b = {"max_height": np.zeros(2 * ny, np.float32), "arrival_times": np.zeros(2 * ny, np.float32),
                 "g_time": np.zeros(ns, np.float32), "g_eta": np.zeros((ns, 2), np.float32),
                 "g_u": np.zeros((ns, 2), np.float32), "g_v": np.zeros((ns, 2), np.float32)}
            for k, v in b.items():
                setattr(self.out[i], k, v.ctypes.data_as(native._fp))
            self.out[i].gauge_capacity = ns
            self.keep.append((lon, geo, gpos, st))
            self.res.append(b)

    def check(self):
        for i, (steps, _) in enumerate(self.specs):
            o = self.out[i]
            assert o.status == 0, (i, o.status)
            assert o.steps == steps and o.n_samples == steps // 10
            b = self.res[i]
            assert np.all(b["max_height"] == i) and np.all(b["arrival_times"] == i + 0.5)
            assert np.all(b["g_eta"][:steps // 10] == i)
            assert self.req[i].arrival_s <= o.admitted_s <= o.finished_s


def _opts(native, capacity=6, budget=0, K=48, workers=4, burst=8):
    o = native.ServeOptions()
    o.capacity, o.hbm_budget, o.steps_per_sync, o.workers, o.burst = capacity, budget, K, workers, burst
    return o


def _stats(F, n):
    adm = (C.c_double * max(1, n))()
    ml, mp, po, un = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    mu = C.c_size_t()
    k = F.fake_stats(adm, n, C.byref(ml), C.byref(mu), C.byref(mp), C.byref(po), C.byref(un))
    return dict(admitted=[int(x) for x in adm[:k]], max_live=ml.value, max_used=mu.value, prepared=mp.value,
                prepared_on_driver=po.value, unprepared=un.value)


def _specs(n, seed=3, max_ny=5):
    rng = np.random.default_rng(seed)
    return [(int(rng.integers(1, 400)), int(rng.integers(1, max_ny + 1))) for _ in range(n)]


@pytest.mark.parametrize("workers", [1, 4])
def test_every_request_finishes_with_its_own_results(loop, workers):
    native, F, S = loop
    F.fake_reset()
    r = Reqs(native, _specs(40))
    gpu = C.c_double()
    assert S.thy_serve_batched(r.req, 40, C.byref(_opts(native, workers=workers)), r.out, C.byref(gpu)) == 0
    r.check()
    st = _stats(F, 40)
    assert st["max_live"] <= 6
    assert st["admitted"] == list(range(40))  # admitted in arrival order
    # runs are prepared on the worker threads, never on the thread that drives the batch
    assert st["prepared"] == 40 and st["prepared_on_driver"] == 0 and st["unprepared"] == 0
    assert gpu.value > 0


def test_budget_bounds_the_resident_scenarios(loop):
    native, F, S = loop
    F.fake_reset()
    specs = [(50 + 7 * i, 2) for i in range(12)]  # 4 bytes each
    r = Reqs(native, specs)
    assert S.thy_serve_batched(r.req, 12, C.byref(_opts(native, capacity=64, budget=9, K=24)), r.out, None) == 0
    r.check()
    st = _stats(F, 12)
    assert st["max_used"] <= 9 and st["max_live"] == 2


def test_arrivals_are_respected(loop):
    native, F, S = loop
    F.fake_reset()
    arrivals = np.linspace(0.0, 0.05, 10)[::-1].copy()  # last request first
    r = Reqs(native, _specs(10, seed=5), arrivals)
    assert S.thy_serve_batched(r.req, 10, C.byref(_opts(native, capacity=4)), r.out, None) == 0
    r.check()
    assert _stats(F, 10)["admitted"] == list(range(10))[::-1]
    for i in range(10):
        assert r.out[i].admitted_s >= arrivals[i]


def test_a_request_above_the_budget_fails_alone(loop):
    """A request whose footprint exceeds the whole budget gets THY_ERR_BUDGET; the
    others are served."""
    native, F, S = loop
    F.fake_reset()
    r = Reqs(native, [(30, 2), (40, 20), (50, 2)])
    assert S.thy_serve_batched(r.req, 3, C.byref(_opts(native, budget=8)), r.out, None) == 0
    assert r.out[1].status == native.STATUS_CODES["BUDGET"]
    assert r.out[0].status == 0 and r.out[2].status == 0 and r.out[2].steps == 50


def test_a_failing_scenario_is_retired_with_its_status(loop):
    """A scenario the health check stops (the stand-in fails a run of 999999 steps) ends
    with THY_ERR_NONFINITE; the batch and the other requests go on."""
    native, F, S = loop
    F.fake_reset()
    r = Reqs(native, [(60, 2), (999999, 2), (70, 2)])
    assert S.thy_serve_batched(r.req, 3, C.byref(_opts(native, capacity=2)), r.out, None) == 0
    assert r.out[1].status == native.STATUS_CODES["NONFINITE"]
    assert [r.out[i].steps for i in (0, 2)] == [60, 70] and r.out[0].status == r.out[2].status == 0


def test_sequential_loop_returns_the_same_results(loop):
    native, F, S = loop
    specs = _specs(6, seed=9)
    rb, rs = Reqs(native, specs), Reqs(native, specs)
    assert S.thy_serve_batched(rb.req, 6, C.byref(_opts(native, capacity=3)), rb.out, None) == 0
    assert S.thy_serve_sequential(rs.req, 6, C.byref(_opts(native)), rs.out) == 0
    rb.check()
    rs.check()
    for i in range(1, 6):  # one at a time, in arrival order
        assert rs.out[i].admitted_s >= rs.out[i - 1].finished_s
