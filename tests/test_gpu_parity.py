"""GPU parity: the HIP path (through the C-ABI) against the FP64 oracle.

Tolerances: eta, qx, qy within relative L-inf 1e-4 of the peak |eta|
after the same step count with the oracle's dt replayed; relative mass drift 1e-6
with walls; max-height and gauge series matching.
"""
import numpy as np
import pytest

from conftest import make_problem, rel_linf

pytestmark = pytest.mark.gpu


def _native():
    from thysea import native
    return native


def _oracle(pr, **kw):
    import oracle as O
    return O.OracleSim(pr, **kw)


# step kernels under test: the row sweep with its default strip height and with short
# strips (many strips, ragged last strip), and the 2D LDS-tile kernel
KERNELS = [pytest.param(dict(kernel="sweep"), id="sweep"),
           pytest.param(dict(kernel="sweep", strip_rows=5), id="sweep5"),
           pytest.param(dict(kernel="tile"), id="tile")]


def test_probe_gfx950():
    assert _native().probe() == 0


@pytest.mark.parametrize("kw", KERNELS)
def test_lake_at_rest(kw):
    pr = make_problem(fault=False)
    s = _native().Solver(pr, **kw)
    dt = s.initial_dt()
    h0, H0, qx0, qy0 = s.state()
    for _ in range(20):
        s.step(dt)
    h, H, qx, qy = s.state()
    assert np.max(np.abs(qx)) < 1e-10 and np.max(np.abs(qy)) < 1e-10
    assert np.max(np.abs(h - h0)) < 1e-9


def test_initial_dt_matches_oracle():
    pr = make_problem()
    s = _native().Solver(pr)
    o = _oracle(pr)
    assert abs(s.initial_dt() - o.initial_dt()) / o.initial_dt() < 1e-5


def test_okada_matches_oracle():
    pr = make_problem(nx=96, ny=80)
    s = _native().Solver(pr)
    o = _oracle(pr)
    f = pr.faults[0]
    dg = s.okada(f, want_def=True)
    do = o.okada(f, want_def=True)
    scale = np.max(np.abs(do))
    assert scale > 0
    assert rel_linf(dg, do, scale) < 1e-10


def _replay(pr, kw, steps=40):
    s = _native().Solver(pr, **kw)
    o = _oracle(pr)
    f = pr.faults[0]
    s.okada(f)
    o.okada(f)
    dt = o.initial_dt()
    for _ in range(steps):
        dnext = o.step(dt)
        s.step(dt)
        dt = dnext
    hg, Hg, qxg, qyg = s.state()
    ho, Ho, qxo, qyo = o.state()
    eta_o = ho - Ho - pr.hmin
    eta_g = hg - Hg - pr.hmin
    wet = ho > pr.eps_h   # on dry land eta = -depth is not a model output (masked -9999)
    peak = np.max(np.abs(eta_o[wet]))
    return (rel_linf(eta_g[wet], eta_o[wet], peak), rel_linf(qxg, qxo, peak), rel_linf(qyg, qyo, peak))


@pytest.mark.parametrize("kw", KERNELS)
@pytest.mark.parametrize("borders", [(1, 1, 1, 1), (-1, -1, -1, -1), (1, -1, 2, 2)])
def test_step_replay(borders, kw):
    """Same dt sequence on both sides (the oracle's), compare after N steps; FP32 state
    within 1e-4 of the peak |eta|."""
    pr = make_problem(nx=70, ny=52, borders=borders)
    errs = _replay(pr, kw)
    assert max(errs) < 1e-4, errs


@pytest.mark.parametrize("kw", KERNELS)
@pytest.mark.parametrize("borders", [(1, 1, 1, 1), (-1, 1, 2, 2)])
def test_step_replay_multi_strip(borders, kw):
    """A grid spanning several 56-column sweep segments (ragged last one) and several
    row strips, with sponges on the open borders."""
    pr = make_problem(nx=131, ny=97, borders=borders, seamounts=5, seed=11)
    errs = _replay(pr, kw, steps=30)
    assert max(errs) < 1e-4, errs


@pytest.mark.parametrize("kw", KERNELS)
def test_step_replay_fp64(kw):
    """FP64-everywhere path: the oracle's arithmetic to round-off."""
    pr = make_problem(nx=70, ny=52)
    errs = _replay(pr, dict(kw, precision=1))
    assert max(errs) < 1e-9, errs


@pytest.mark.parametrize("kw", KERNELS)
def test_mass_walls(kw):
    pr = make_problem(nx=64, ny=48, borders=(-1, -1, -1, -1))
    s = _native().Solver(pr, **kw)
    s.okada(pr.faults[0])
    m0 = s.mass()
    dt = s.initial_dt()
    for _ in range(60):
        dt = s.step(dt)
    assert abs(s.mass() - m0) / m0 < 1e-6


@pytest.mark.parametrize("kw", KERNELS)
def test_full_run_small_matches_oracle(kw):
    pr = make_problem(nx=96, ny=80, t_sim=900.0)
    from thysea import scenario as S
    glon, glat = S.gauges_lattice(pr.lon, pr.lat, 3)
    from thysea import problem as P
    pr.gauge_pos = np.array([P.nearest_cell(pr.lon, pr.lat, a, b) for a, b in zip(glon, glat)], np.int32)
    pr.series_interval = 30.0 / pr.T
    s = _native().Solver(pr, **kw)
    rg = s.run()
    o = _oracle(pr)
    ro = o.run()
    assert abs(rg["steps"] - ro["steps"]) <= 1
    mg = s.field("max_height")
    mo = o.output("max_height")
    wet = mo > -9000
    peak = np.max(np.abs(mo[wet]))
    assert np.array_equal(mg > -9000, wet)
    assert rel_linf(mg[wet], mo[wet], peak) < 1e-4
    n = min(len(rg["g_time"]), len(ro["g_time"]))
    assert n > 10
    assert rel_linf(rg["g_eta"][:n], ro["g_eta"][:n], peak) < 1e-4


def _scenario(nx, ny, seed, fault_frac, slip, t_sim):
    pr = make_problem(nx=nx, ny=ny, seed=seed, t_sim=t_sim)
    f = dict(pr.faults[0])
    f["lon_c"] = float(pr.lon[int(nx * fault_frac)])
    f["slip"] = slip
    pr.faults = [f]
    from thysea import scenario as S
    from thysea import problem as P
    glon, glat = S.gauges_lattice(pr.lon, pr.lat, 2)
    pr.gauge_pos = np.array([P.nearest_cell(pr.lon, pr.lat, a, b) for a, b in zip(glon, glat)], np.int32)
    pr.series_interval = 20.0 / pr.T
    return pr


def test_batch_matches_alone():
    """Scenarios of different sizes and durations, advanced together in one batch,
    give the results of each run alone (same kernel code, own dt per scenario): max
    height, arrival times, gauges and step counts identical; one scenario created,
    uploaded and admitted while a launch of the others is in flight."""
    nat = _native()
    probs = [_scenario(96, 80, 3, 0.4, 2.0, 600.0), _scenario(150, 64, 5, 0.5, 4.0, 900.0),
             _scenario(70, 120, 9, 0.3, 1.0, 400.0)]
    alone = []
    for pr in probs:
        s = nat.Solver(pr)
        r = s.run(steps_per_sync=6)
        r["max_height"] = s.field("max_height")
        r["arrival_times"] = s.field("arrival_times")
        alone.append(r)
        s.close()
    b = nat.Batch(capacity=4, steps_per_sync=6)
    solvers = [nat.Solver(pr) for pr in probs[:2]]
    slots = [b.add(solvers[0]), b.add(solvers[1])]
    b.launch()
    solvers.append(nat.Solver(probs[2]))
    slots.append(b.add(solvers[2]))  # admitted while the launch is in flight
    assert b.wait() == 3
    # progress of running slots: the first two stepped one sequence, the third none
    assert [b.status(k)[1]["steps"] for k in slots] == [6, 6, 0]
    for _ in range(10000):
        if b.advance() == 0:
            break
    assert sorted(b.finished()) == sorted(slots)
    for slot, ra in zip(slots, alone):
        rb = b.result(slot)
        assert rb["steps"] == ra["steps"]
        assert np.array_equal(rb["max_height"], ra["max_height"])
        assert np.array_equal(rb["arrival_times"], ra["arrival_times"])
        assert rb["n_samples"] == ra["n_samples"] and rb["n_samples"] > 5
        assert np.array_equal(rb["g_eta"], ra["g_eta"])
        b.remove(slot)
    assert b.info()["running"] == 0 and b.info()["bytes"] == 0
    b.close()
    for s in solvers:
        s.close()


def test_serving_loop_matches_sequential():
    """Requests arriving over time, served by the pipelined loop (admission and
    read-out while the GPU steps the batch, finished slots removed while a launch is
    in flight), give the results of the same requests run one at a time."""
    from thysea import serving
    probs = [_scenario(96, 80, 3, 0.4, 2.0, 300.0), _scenario(150, 64, 5, 0.5, 4.0, 500.0),
             _scenario(70, 120, 9, 0.3, 1.0, 200.0), _scenario(400, 300, 13, 0.4, 3.0, 400.0),
             _scenario(96, 80, 4, 0.6, 2.5, 250.0)]
    arrivals = [0.0, 0.0, 0.02, 0.05, 0.08]
    keep = list(range(len(probs)))
    rec_b, res_b, _ = serving.run_batched(probs, arrivals=arrivals, steps_per_sync=6, keep=keep, burst=2)
    rec_s, res_s = serving.run_sequential(probs, arrivals=arrivals, steps_per_sync=6, keep=keep)
    for i in keep:
        assert rec_b[i]["steps"] == rec_s[i]["steps"] > 0
        for key in ("max_height", "arrival_times", "g_eta"):
            assert np.array_equal(res_b[i][key], res_s[i][key]), (i, key)
        assert rec_b[i]["finished"] >= rec_b[i]["admitted"] >= rec_b[i]["arrival"]


def test_threads_beside_batch_capture():
    """Contexts set up, run (each capturing its own step sequences) and read out on
    worker threads while the main thread records and replays batch step sequences:
    no capture is invalidated, and every result equals a run alone."""
    from concurrent.futures import ThreadPoolExecutor
    nat = _native()
    side = [_scenario(96, 80, 20 + k, 0.4, 2.0, 200.0) for k in range(6)]
    main = [_scenario(150, 64, 40 + k, 0.5, 3.0, 300.0) for k in range(3)]

    def alone(pr):
        s = nat.Solver(pr)
        r = s.run(steps_per_sync=6)
        out = (r["steps"], s.field("max_height"), r["g_eta"])
        s.close()
        return out

    ref_side = [alone(pr) for pr in side]
    ref_main = [alone(pr) for pr in main]
    b = nat.Batch(capacity=8, steps_per_sync=6)
    with ThreadPoolExecutor(max_workers=3) as pool:
        futs = [pool.submit(alone, pr) for pr in side]
        slots = {}
        for k, pr in enumerate(main):
            slots[b.add(nat.Solver(pr), steps_per_sync=6)] = k
            b.advance()  # the set changed: the batch records a new step sequence
        while b.advance() > 0:
            pass
        got_side = [f.result() for f in futs]
    for slot, k in slots.items():
        r = b.result(slot)
        assert r["steps"] == ref_main[k][0]
        assert np.array_equal(r["max_height"], ref_main[k][1]), k
        assert np.array_equal(r["g_eta"], ref_main[k][2]), k
        b.remove(slot).close()
    b.close()
    for k in range(len(side)):
        assert got_side[k][0] == ref_side[k][0]
        assert np.array_equal(got_side[k][1], ref_side[k][1]), k
        assert np.array_equal(got_side[k][2], ref_side[k][2]), k


def test_batch_prepared_on_another_thread_matches_alone():
    """Runs prepared by thy_batch_prepare on a worker thread (faults at t = 0 applied
    there) and admitted with no run configuration give the results of runs alone; a
    context admitted with no configuration and no preparation is refused."""
    from concurrent.futures import ThreadPoolExecutor
    nat = _native()
    probs = [_scenario(96, 80, 61, 0.4, 2.0, 300.0), _scenario(150, 64, 62, 0.5, 3.0, 400.0)]
    alone = []
    for pr in probs:
        s = nat.Solver(pr)
        r = s.run(steps_per_sync=6)
        alone.append((r["steps"], s.field("max_height"), s.field("arrival_times"), r["g_eta"]))
        s.close()

    def prepared(pr):
        s = nat.Solver(pr)
        s.prepare_batch(steps_per_sync=6)
        return s

    b = nat.Batch(capacity=4, steps_per_sync=6)
    raw = nat.Solver(probs[0])
    rc = nat.lib().thy_batch_add(b.h, raw.h, None, nat.C.byref(nat.C.c_int()))
    assert rc == nat.STATUS_CODES["INVALID"]
    raw.close()
    with ThreadPoolExecutor(max_workers=2) as pool:
        solvers = list(pool.map(prepared, probs))
    slots = [b.add(s) for s in solvers]
    while b.advance() > 0:
        pass
    for slot, ra in zip(slots, alone):
        rb = b.result(slot)
        assert rb["steps"] == ra[0]
        assert np.array_equal(rb["max_height"], ra[1])
        assert np.array_equal(rb["arrival_times"], ra[2])
        assert np.array_equal(rb["g_eta"], ra[3])
        b.remove(slot).close()
    b.close()


def test_batch_tuner_times_every_candidate():
    """tune_batch advances a batch at each candidate strip height and returns the
    cheapest, with a positive cost for every candidate."""
    from thysea import serving
    probs = [_scenario(150, 96, 50 + k, 0.5, 2.0, 300.0) for k in range(4)]
    rows, cost = serving.tune_batch(probs, candidates=(8, 16), n=4, launches=1, steps_per_sync=6)
    assert sorted(cost) == [8, 16] and rows in cost
    assert all(v > 0 for v in cost.values())


def test_batch_budget_admission():
    """A scenario beyond the HBM budget is refused (None), and admitted once a
    finished one has been removed."""
    nat = _native()
    pr = _scenario(96, 80, 3, 0.4, 2.0, 120.0)
    need = nat.footprint(pr.nx, pr.ny)
    b = nat.Batch(capacity=4, hbm_budget=int(need * 1.5), steps_per_sync=6)
    s1, s2 = nat.Solver(pr), nat.Solver(pr)
    a = b.add(s1)
    assert a is not None
    assert b.add(s2) is None
    while b.advance() > 0:
        pass
    b.remove(a)
    assert b.add(s2) is not None
    while b.advance() > 0:
        pass
    b.close()
    s1.close()
    s2.close()


def _fields(s):
    h, H, qx, qy = s.state()
    return h - H, qx, qy, s.field("max_height"), s.field("arrival_times")


def test_sweep_results_independent_of_strip_height():
    """Each face is evaluated once with the same arithmetic whatever the strip
    boundaries, so the state after 25 steps is bit-identical for any strip height."""
    pr = make_problem(nx=131, ny=97, seamounts=5, seed=11)
    ref = None
    for rows in (3, 7, 32, 64):
        s = _native().Solver(pr, strip_rows=rows)
        s.run(max_steps=25, steps_per_sync=6)
        got = _fields(s)
        s.close()
        if ref is None:
            ref = got
        else:
            for a, b in zip(ref, got):
                assert np.array_equal(a, b), rows


def test_autotune_mid_run_leaves_the_run_unchanged():
    """thy_autotune between thy_run and thy_run_more picks a strip height and restores
    the state, control blocks and dt slots: the run ends exactly as without it."""
    pr = make_problem(nx=131, ny=97, seamounts=5, seed=11)
    a = _native().Solver(pr)
    a.run(max_steps=30, steps_per_sync=6)
    b = _native().Solver(pr)
    b.run(max_steps=12, steps_per_sync=6)
    rows, ms = b.autotune(reps=2)
    assert rows in (16, 24, 32, 36, 40, 48) and ms > 0  # the candidate strip heights
    r = b.run_more(18)
    assert r["steps"] == 30
    for x, y in zip(_fields(a), _fields(b)):
        assert np.array_equal(x, y)
    a.close()
    b.close()


def test_precision_fp32_against_fp64_run():
    """Precision study on a small grid: the FP32 production path against the
    FP64-everywhere run of the same kernels (max height within 1e-4 of its peak,
    arrival times within one step, same step count within 1), and mass drift with walls
    below 1e-6 for both."""
    nat = _native()
    pr = make_problem(nx=120, ny=100, seamounts=5, seed=4, t_sim=1200.0)
    out = {}
    for prec in (0, 1):
        s = nat.Solver(pr, precision=prec)
        r = s.run(steps_per_sync=6)
        out[prec] = (r, s.field("max_height"), s.field("arrival_times"))
        s.close()
    (r32, m32, a32), (r64, m64, a64) = out[0], out[1]
    assert abs(r32["steps"] - r64["steps"]) <= 1
    wet = m64 > -9000
    assert np.array_equal(m32 > -9000, wet)
    assert rel_linf(m32[wet], m64[wet], np.max(np.abs(m64[wet]))) < 1e-4
    both = (a32 >= 0) & (a64 >= 0)
    dt_s = r64["t_final"] * pr.T / r64["steps"]
    assert np.percentile(np.abs(a32[both] - a64[both]), 99) <= 2.01 * dt_s
    prw = make_problem(nx=96, ny=80, borders=(-1, -1, -1, -1), seed=4)
    for prec in (0, 1):
        s = nat.Solver(prw, precision=prec)
        s.okada(prw.faults[0])
        m0 = s.mass()
        dt = s.initial_dt()
        for _ in range(60):
            dt = s.step(dt)
        assert abs(s.mass() - m0) / m0 < 1e-6
        s.close()


def _split_pair(pr, steps, precision=0):
    out = []
    for kernel in ("sweep", "sweep_general"):
        s = _native().Solver(pr, kernel=kernel, strip_rows=16, precision=precision)
        s.run(max_steps=steps, steps_per_sync=6)
        out.append(_fields(s))
        s.close()
    return out


@pytest.mark.parametrize("precision", [0, 1], ids=["fp32", "fp64"])
@pytest.mark.parametrize("borders", [(1, 1, 1, 1), (-1, -1, -1, -1), (1, -1, 2, 2)],
                         ids=["open", "walls", "periodic_x"])
def test_deep_split_identical_to_general_faces(borders, precision):
    """The deep-water launch (interior blocks far from coast and borders) and the
    general launch evaluate the deep faces with the same functions: the split run is
    bit-identical to the general-faces-only run, with open, wall and periodic borders,
    in FP32 and in FP64."""
    pr = make_problem(nx=400, ny=300, seamounts=6, seed=21, borders=borders)
    a, b = _split_pair(pr, 30, precision)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_deep_split_identical_on_serving_grid():
    """The 1000x800 shared serving grid (open ocean, shelf and coast in one strip
    row): 60 steps of the split run equal the general-faces-only run bit for bit, so
    no formula is rounded differently in the two kernels."""
    from thysea import scenario as S
    pr = S.batch_problems(1)[0]
    a, b = _split_pair(pr, 60)
    for x, y in zip(a, b):
        assert np.array_equal(x, y), float(np.max(np.abs(x - y)))


def test_deep_split_hands_back_failing_blocks():
    """A dry patch carved into deep water (classified deep from the still-water depth,
    but h <= hpos at run time): the deep-water launch hands those blocks back to the
    general launch of the same step; the result is still bit-identical to the
    general-faces-only run (FP32 and FP64) and within tolerance of the oracle."""
    pr = _hole_problem()
    for precision in (0, 1):
        a, b = _split_pair(pr, 20, precision)
        for x, y in zip(a, b):
            assert np.array_equal(x, y), precision
    s = _native().Solver(pr)
    o = _oracle(pr)
    dt = o.initial_dt()
    for _ in range(20):
        dn = o.step(dt)
        s.step(dt)
        dt = dn
    hg, Hg, qxg, qyg = s.state()
    ho, Ho, qxo, qyo = o.state()
    peak = np.max(np.abs(qxo)) + np.max(np.abs(qyo))
    assert rel_linf(qxg, qxo, peak) < 1e-4 and rel_linf(qyg, qyo, peak) < 1e-4
    wet = ho > pr.eps_h
    assert np.array_equal(hg > pr.eps_h, wet)
    # eta = h - H - Hmin on wet cells, relative to its peak (the dry hole fills from its rim)
    eta_o, eta_g = (ho - Ho - pr.hmin)[wet], (hg - Hg - pr.hmin)[wet]
    assert rel_linf(eta_g, eta_o, np.max(np.abs(eta_o))) < 1e-4
    s.close()


def _hole_problem():
    """Flat deep basin with a dry hole carved into it (deep-water strips by the still
    water depth, h <= hpos at run time)."""
    pr = make_problem(nx=400, ny=300, seamounts=0, seed=22, fault=False)
    H = pr.H.reshape(pr.ny, pr.nx)
    h = pr.h.reshape(pr.ny, pr.nx).copy()
    assert (H[140:150, 180:190] > 1.0).all()
    h[140:150, 180:190] = 0.0
    pr.h = h.ravel() if pr.h.ndim == 1 else h
    return pr


def test_batch_hands_back_failing_deep_strips():
    """The dry hole next to a scenario with a fault, in one batch: the batch hands the
    hole's strips back to its hand-back launch, and both scenarios end in the state of
    their runs alone, bit for bit."""
    nat = _native()
    probs = [_hole_problem(), _scenario(400, 300, 13, 0.4, 3.0, 900.0)]
    alone = []
    for pr in probs:
        s = nat.Solver(pr)
        s.run(max_steps=30, steps_per_sync=6)
        alone.append(_fields(s))
        s.close()
    b = nat.Batch(capacity=2, steps_per_sync=6)
    solvers = [nat.Solver(pr) for pr in probs]
    slots = [b.add(s, max_steps=30) for s in solvers]
    while b.advance() > 0:
        pass
    for slot, s, ref in zip(slots, solvers, alone):
        assert b.status(slot)[1]["steps"] == 30
        for x, y in zip(_fields(s), ref):
            assert np.array_equal(x, y), float(np.max(np.abs(x - y)))
    b.close()
    for s in solvers:
        s.close()


def test_batch_matches_alone_with_deep_water_strips():
    """A grid large enough for the deep-water launch: the batch (general faces on every
    strip) gives the same max height, arrival times and gauges as the scenario run
    alone (deep-water launch beside the general one)."""
    nat = _native()
    pr = _scenario(400, 300, 13, 0.4, 3.0, 900.0)
    s = nat.Solver(pr)
    ra = s.run(steps_per_sync=6)
    ma, aa = s.field("max_height"), s.field("arrival_times")
    s.close()
    b = nat.Batch(capacity=2, steps_per_sync=6)
    sb = nat.Solver(pr)
    slot = b.add(sb)
    while b.advance() > 0:
        pass
    rb = b.result(slot)
    assert rb["steps"] == ra["steps"]
    d = np.abs(rb["max_height"] - ma)
    assert np.array_equal(rb["max_height"], ma), float(d.max())
    assert np.array_equal(rb["arrival_times"], aa)
    assert np.array_equal(rb["g_eta"], ra["g_eta"])
    b.close()
    sb.close()


# ---------------------------------------------------------------------------
# The 4000x3000 (12M-cell) workload grid itself: properties that hold at any size.
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def large_open():
    from thysea import scenario as S
    return S.problem("large")


def test_full_size_lake_at_rest(large_open):
    """The resting ocean over the 12M-cell grid's bathymetry (coasts, seamounts, dry
    land, sponges) stays at rest: every flux vanishes to round-off (the initial eta is
    zero to the round-off of recovering the depth from the Hmin-shifted H, as in the
    small-grid test)."""
    s = _native().Solver(large_open)
    h0, H0, _, _ = s.state()
    dt = s.initial_dt()
    for _ in range(12):
        s.step(dt)
    h, H, qx, qy = s.state()
    assert np.max(np.abs(qx)) < 1e-10 and np.max(np.abs(qy)) < 1e-10
    assert np.max(np.abs(h - h0)) < 1e-9 and np.array_equal(H, H0)
    s.close()


def test_full_size_split_and_mass():
    """The 12M-cell grid with walls and its four Okada subfaults: over 48 steps the
    total mass is conserved to 1e-6, and the split run (deep-water launch beside the
    general launch) equals the general-faces-only run bit for bit."""
    from thysea import scenario as S
    pr = S.problem("large", borders=(-1, -1, -1, -1))
    out = []
    for kernel in ("sweep", "sweep_general"):
        s = _native().Solver(pr, kernel=kernel)
        m0 = s.mass()
        r = s.run(max_steps=48, steps_per_sync=24)
        assert r["steps"] == 48
        m1 = s.mass()
        assert abs(m1 - m0) / m0 < 1e-6, (kernel, m0, m1)
        out.append(_fields(s))
        s.close()
    for x, y in zip(*out):
        assert np.array_equal(x, y), float(np.max(np.abs(x - y)))


def test_batch_many_general_strips_matches_alone():
    """Twelve scenarios on the 1000x800 serving grid: enough general strips that the
    batch runs them as whole strips (a run alone cuts them shorter); every sampled
    scenario still ends in the state of its run alone, bit for bit."""
    from thysea import scenario as S
    nat = _native()
    probs = S.batch_problems(12)
    alone = {}
    for i in (0, 5, 11):
        s = nat.Solver(probs[i])
        s.run(max_steps=30, steps_per_sync=6)
        alone[i] = _fields(s)
        s.close()
    b = nat.Batch(capacity=16, steps_per_sync=6)
    solvers = [nat.Solver(p) for p in probs]
    slots = [b.add(s, max_steps=30) for s in solvers]
    while b.advance() > 0:
        pass
    for i, ref in alone.items():
        assert b.status(slots[i])[1]["steps"] == 30
        for x, y in zip(_fields(solvers[i]), ref):
            assert np.array_equal(x, y), (i, float(np.max(np.abs(x - y))))
    b.close()
    for s in solvers:
        s.close()
