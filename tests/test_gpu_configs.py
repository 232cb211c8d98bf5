"""The workloads at their stated sizes on the GPU.

* 2000x2000, 7200 s, Mw 8.5 (the precision study): the FP32 path against the
  FP64-everywhere run of the same kernels, and ten steps against the oracle with the
  oracle's dt sequence.
* 256 scenarios on one shared 1000x800 grid (the batched workload): 30 steps of all 256
  in one batch; sampled scenarios bit-identical to their runs alone, and one scenario
  against the oracle's own driver loop.
* Serving requests on grids from 300x300 to 2400x1800 (the serving workload): a mix that
  includes a 2400x1800 grid through the C++ request loop, equal to the same requests one
  at a time.
"""
import numpy as np
import pytest

from conftest import rel_linf

pytestmark = pytest.mark.gpu

TOL = 1e-4


def _native():
    from thysea import native
    return native


@pytest.fixture(scope="module")
def precision_problem():
    from thysea import scenario
    return scenario.problem("precision")


def test_precision_workload_fp32_against_fp64(precision_problem):
    """The whole 7200 s run: same step count, final eta and max height within the bar of
    their peaks, identical wet masks, arrival times within one step for 99 % of the
    cells."""
    pr = precision_problem
    assert (pr.nx, pr.ny) == (2000, 2000) and abs(pr.t_end * pr.T - 7200.0) < 1e-6
    nat = _native()
    out = {}
    for prec in (0, 1):
        s = nat.Solver(pr, precision=prec)
        r = s.run(steps_per_sync=48)
        h, H, _, _ = s.state()
        out[prec] = (r, h - H - pr.hmin, s.field("max_height"), s.field("arrival_times"))
        s.close()
    (r32, e32, m32, a32), (r64, e64, m64, a64) = out[0], out[1]
    assert r32["steps"] == r64["steps"] > 1500
    wet = m64 > -9000
    assert np.array_equal(m32 > -9000, wet)
    err_m = rel_linf(m32[wet], m64[wet], float(np.max(np.abs(m64[wet]))))
    err_e = rel_linf(e32, e64, float(np.max(np.abs(e64))))
    both = (a32 >= 0) & (a64 >= 0)
    dt_s = r64["t_final"] * pr.T / r64["steps"]
    d = np.abs(a32[both] - a64[both])
    print(f"2000x2000: max height {err_m:.3e}, final eta {err_e:.3e}, arrival p99 {np.percentile(d, 99):.3g} s")
    assert err_m < TOL and err_e < TOL
    assert np.percentile(d, 99) <= dt_s


def test_precision_workload_against_oracle(precision_problem):
    """Ten steps of the 2000x2000 workload after its Okada source, the oracle's dt
    replayed: eta (wet cells), qx and qy within the bar of the peak eta."""
    import oracle as O
    pr = precision_problem
    s = _native().Solver(pr)
    o = O.OracleSim(pr)
    for f in pr.faults:
        s.okada(f)
        o.okada(f)
    dt = o.initial_dt()
    for _ in range(10):
        dn = o.step(dt)
        s.step(dt)
        dt = dn
    hg, Hg, qxg, qyg = s.state()
    ho, Ho, qxo, qyo = o.state()
    wet = ho > pr.eps_h
    peak = float(np.max(np.abs((ho - Ho - pr.hmin)[wet])))
    errs = (rel_linf((hg - Hg)[wet], (ho - Ho)[wet], peak), rel_linf(qxg, qxo, peak), rel_linf(qyg, qyo, peak))
    print("2000x2000 oracle replay:", errs)
    assert max(errs) < TOL
    s.close()
    o.close()


def test_batch_workload_256_scenarios():
    """All 256 scenarios of the batched workload advance 30 steps in one batch; four
    sampled scenarios end bit-identical to their runs alone, and one of them within the
    bar of the oracle's own 30-step driver loop (own dt sequence)."""
    import oracle as O
    from thysea import scenario
    nat = _native()
    probs = scenario.batch_problems(256)
    assert len(probs) == 256 and (probs[0].nx, probs[0].ny) == (1000, 800)
    sample = (0, 77, 170, 255)
    alone = {}
    for i in sample:
        s = nat.Solver(probs[i])
        s.run(max_steps=30, steps_per_sync=6)
        h, H, qx, qy = s.state()
        alone[i] = (h - H, qx, qy, s.field("max_height"), s.field("arrival_times"))
        s.close()
    b = nat.Batch(capacity=256, steps_per_sync=6)
    solvers = [nat.Solver(p) for p in probs]
    slots = [b.add(s, max_steps=30) for s in solvers]
    assert None not in slots
    while b.advance() > 0:
        pass
    for i in sample:
        assert b.status(slots[i])[1]["steps"] == 30
        h, H, qx, qy = solvers[i].state()
        got = (h - H, qx, qy, solvers[i].field("max_height"), solvers[i].field("arrival_times"))
        for x, y in zip(got, alone[i]):
            assert np.array_equal(x, y), i
    b.close()
    for s in solvers:
        s.close()
    pr = probs[77]
    o = O.OracleSim(pr)
    ro = o.run(max_steps=30)
    assert ro["steps"] == 30
    ho, Ho, qxo, qyo = o.state()
    eg = alone[77][0] - pr.hmin
    wet = ho > pr.eps_h
    peak = float(np.max(np.abs((ho - Ho - pr.hmin)[wet])))
    e = rel_linf(eg[wet], (ho - Ho - pr.hmin)[wet], peak)
    mh = o.output("max_height")
    w = mh > -9000
    em = rel_linf(alone[77][3][w], mh[w], float(np.max(np.abs(mh[w]))))
    print(f"batch scenario 77 vs oracle: eta {e:.3e}, max height {em:.3e}")
    assert e < TOL and em < TOL
    o.close()


def test_serving_mix_with_largest_grid():
    """Requests of the serving workload, including one on a 2400x1800 grid (the top of
    the size range) and one on 300x300, arriving over 0.3 s: the C++ request loop gives
    every request the results of the same request run alone."""
    from thysea import scenario, serving
    specs = [dict(nx=2400, ny=1800, t_sim=600.0, seed=901, mw_range=(8.5, 9.0)),
             dict(nx=300, ny=300, t_sim=900.0, seed=902, mw_range=(7.0, 7.5)),
             dict(nx=1300, ny=700, t_sim=400.0, seed=903, mw_range=(7.5, 8.5)),
             dict(nx=800, ny=1500, t_sim=500.0, seed=904, mw_range=(7.0, 9.0))]
    probs = [scenario.request_problem(s) for s in specs]
    arrivals = [0.0, 0.05, 0.15, 0.3]
    keep = list(range(len(probs)))
    rec_b, res_b, _ = serving.run_batched(probs, arrivals=arrivals, steps_per_sync=12, keep=keep)
    rec_s, res_s = serving.run_sequential(probs, arrivals=arrivals, steps_per_sync=12, keep=keep)
    for i in keep:
        assert rec_b[i]["steps"] == rec_s[i]["steps"] > 50
        for key in ("max_height", "arrival_times", "g_eta", "g_u", "g_v"):
            assert np.array_equal(res_b[i][key], res_s[i][key]), (i, key)
        assert rec_b[i]["finished"] >= rec_b[i]["admitted"] >= rec_b[i]["arrival"]
