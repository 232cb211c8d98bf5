"""GPU parity of what a user reads: the output fields, snapshots, gauge series,
arrival times and deformed bathymetry of the HIP path against the oracle; fault
schedules with later fault times; Okada cropping; a WAF-everywhere dam break; a wave
that runs up a shelf; the health check.

Tolerance: the north-star bar, 1e-4 of the peak of the compared quantity (eta peak for
eta, max height and gauge eta; the peak speed of the compared cells for velocities),
with identical wet/dry masks.  Arrival times are compared as times (a cell whose
amplitude crosses the threshold a step earlier or later moves by one dt).
"""
import math

import numpy as np
import pytest

from conftest import make_problem, rel_linf

pytestmark = pytest.mark.gpu

TOL = 1e-4


def _native():
    from thysea import native
    return native


def _oracle(pr, **kw):
    import oracle as O
    return O.OracleSim(pr, **kw)


def _with_gauges(pr, n=3, every=20.0):
    from thysea import problem as P
    from thysea import scenario as S
    glon, glat = S.gauges_lattice(pr.lon, pr.lat, n)
    pr.gauge_pos = np.array([P.nearest_cell(pr.lon, pr.lat, a, b) for a, b in zip(glon, glat)], np.int32)
    pr.series_interval = every / pr.T
    return pr


def _cmp_masked(a, b, what, tol=TOL):
    """Fields with -9999 masks: identical masks, values within tol of b's peak."""
    ma, mb = a > -9000, b > -9000
    assert np.array_equal(ma, mb), (what, int((ma != mb).sum()))
    peak = float(np.max(np.abs(b[mb])))
    err = rel_linf(a[mb], b[mb], peak)
    print(f"{what}: rel err {err:.3e} (peak {peak:.4g})")
    assert err < tol, (what, err)
    return err


def _cmp_arrival(a, b, dt_s, what="arrival"):
    """Arrival times: the same cells arrived; times within one step (a cell at the
    threshold may cross it one step apart), and most of them equal to float precision."""
    ra, rb = a >= 0, b >= 0
    assert np.array_equal(ra, rb), (what, int((ra != rb).sum()))
    d = np.abs(a[rb] - b[rb])
    print(f"{what}: {rb.sum()} cells, max diff {d.max() if d.size else 0:.3g} s, "
          f"p99 {np.percentile(d, 99) if d.size else 0:.3g} s, dt {dt_s:.3g} s")
    assert d.size == 0 or d.max() <= 1.01 * dt_s
    assert d.size == 0 or np.mean(d <= 1e-3 * dt_s + 1e-6 * np.abs(b[rb])) > 0.99


def test_outputs_snapshots_and_gauges_match_oracle():
    """Full driver loop (Okada at t = 0, coast with run-up, snapshots every 150 s, gauges
    every 20 s): every snapshot's eta, ux, uy, the final max height, arrival times and
    deformed bathymetry, and the gauge eta/u/v series match the oracle's."""
    pr = make_problem(nx=120, ny=96, seamounts=4, seed=5, t_sim=600.0, coast=("east", "north"))
    pr = _with_gauges(pr, n=4)
    pr.save_interval = 150.0 / pr.T
    snaps = {"g": [], "o": []}
    s = _native().Solver(pr)
    o = _oracle(pr)

    def grab(store, getter):
        def cb(i, t):
            store.append((t, getter("eta"), getter("ux"), getter("uy")))
        return cb

    rg = s.run(steps_per_sync=12, on_snapshot=grab(snaps["g"], s.field))
    ro = o.run(on_snapshot=grab(snaps["o"], o.output))
    assert rg["steps"] == ro["steps"] and rg["n_snapshots"] == len(snaps["o"]) == len(snaps["g"]) >= 4
    dt_s = ro["t_final"] * pr.T / ro["steps"]
    for (tg, eg, ug, vg), (to, eo, uo, vo) in zip(snaps["g"], snaps["o"]):
        assert abs(tg - to) * pr.T < 1e-3 * dt_s
        _cmp_masked(eg, eo, f"eta snapshot t={to * pr.T:.0f}")
        wet = eo > -9000
        # velocities where the water is deeper than 1 m (below that, u = q/h amplifies the
        # FP32 rounding of h = e + D)
        h_o = eo + np.where(wet, o.output("deformed_bathy"), 0.0)
        deep = wet & (h_o > 1.0)
        for a, b, name in ((ug, uo, "ux"), (vg, vo, "uy")):
            peak = float(np.max(np.abs(b[deep])))
            err = rel_linf(a[deep], b[deep], peak)
            print(f"{name} snapshot t={to * pr.T:.0f}: rel err {err:.3e} (peak {peak:.3g} m/s)")
            if peak > 1e-3:
                assert err < TOL, (name, err)
    _cmp_masked(s.field("max_height"), o.output("max_height"), "max_height")
    _cmp_masked(s.field("eta"), o.output("eta"), "final eta")
    bg, bo = s.field("deformed_bathy"), o.output("deformed_bathy")
    err_b = float(np.max(np.abs(bg - bo)))
    u_peak = float(np.max(np.abs(bo - (pr.H + pr.hmin) * pr.H_ref)))
    print(f"deformed_bathy: max abs diff {err_b:.3g} m (uplift peak {u_peak:.3g} m)")
    assert err_b < 1e-6 * float(np.max(np.abs(bo))) + 1e-6
    _cmp_arrival(s.field("arrival_times"), o.output("arrival_times"), dt_s)
    n = min(len(rg["g_time"]), len(ro["g_time"]))
    assert n == len(ro["g_time"]) > 20
    np.testing.assert_allclose(rg["g_time"][:n], ro["g_time"][:n], rtol=1e-6)
    peak = float(np.max(np.abs(o.output("max_height")[o.output("max_height") > -9000])))
    e_eta = rel_linf(rg["g_eta"][:n], ro["g_eta"][:n], peak)
    upk = max(float(np.max(np.abs(ro["g_u"][:n]))), float(np.max(np.abs(ro["g_v"][:n]))))
    e_u = rel_linf(rg["g_u"][:n], ro["g_u"][:n], upk)
    e_v = rel_linf(rg["g_v"][:n], ro["g_v"][:n], upk)
    print(f"gauges: eta {e_eta:.3e}, u {e_u:.3e}, v {e_v:.3e} (speed peak {upk:.3g} m/s)")
    assert e_eta < TOL and e_u < TOL and e_v < TOL
    gi = s.gauges(rg["n_samples"])
    np.testing.assert_allclose(gi["g_eta_min"], ro["g_eta_min"], rtol=0, atol=TOL * peak)
    np.testing.assert_allclose(gi["g_eta_max"], ro["g_eta_max"], rtol=0, atol=TOL * peak)
    s.close()
    o.close()


def _schedule(pr):
    from test_oracle_outputs import fault_schedule
    return fault_schedule(pr)


def test_fault_schedule_matches_oracle():
    """Faults at t = 0, two at t1 and one at t2, through thy_run: the same number of
    steps, the dt sequence landing on t1 and t2 at the same iterations as the oracle's
    (per-iteration log), and max height and gauges within the bar."""
    pr = _with_gauges(make_problem(nx=96, ny=80, t_sim=240.0), n=3, every=10.0)
    pr.faults = _schedule(pr)
    T = pr.T
    log = []
    s = _native().Solver(pr)
    rg = s.run(steps_per_sync=9, on_step=lambda it, dt, t: log.append((it, dt, t)))
    o = _oracle(pr)
    pr.save_interval = 0.0
    tl = []
    ro = o.run(on_snapshot=lambda i, t: tl.append(t))
    pr.save_interval = -1.0
    assert rg["steps"] == ro["steps"] == len(log)
    assert [it for it, _, _ in log] == list(range(1, rg["steps"] + 1))
    tg = np.array([0.0] + [t for _, _, t in log[:-1]])  # time at the start of each iteration
    to = np.array(tl)
    assert len(to) == len(tg)
    for tf in (60.0, 150.0):
        kg = int(np.argmin(np.abs(tg * T - tf)))
        ko = int(np.argmin(np.abs(to * T - tf)))
        assert kg == ko, (tf, kg, ko)
        assert abs(tg[kg] * T - tf) < 1e-9 * tf and abs(to[ko] * T - tf) < 1e-9 * tf
    np.testing.assert_allclose(tg, to, rtol=1e-5, atol=0)
    _cmp_masked(s.field("max_height"), o.output("max_height"), "max_height (4 faults)")
    mo = o.output("max_height")
    peak = float(np.max(np.abs(mo[mo > -9000])))
    assert rel_linf(rg["g_eta"], ro["g_eta"], peak) < TOL
    # the same run inside a batch: identical to the run alone
    b = _native().Batch(capacity=2, steps_per_sync=6)
    sb = _native().Solver(pr)
    slot = b.add(sb)
    while b.advance() > 0:
        pass
    r = b.result(slot)
    assert r["steps"] == rg["steps"]
    assert np.array_equal(r["max_height"], s.field("max_height"))
    assert np.array_equal(r["g_eta"], rg["g_eta"])
    b.close()
    s.close()
    o.close()


@pytest.mark.parametrize("crop", [(1, 0.1), (2, 0.05)], ids=["relative", "absolute"])
def test_okada_crop_matches_oracle(crop):
    """Relative crop (cells below 10 % of the float max |U_Z| zeroed) and absolute crop
    (below 5 cm) give the oracle's deformation: the same cropped cells and eta within
    1e-7 of the uplift peak (src/GPU/Deformacion.cu:13-36,132-142)."""
    flag, val = crop
    pr = make_problem(nx=96, ny=80)
    cv = val if flag == 1 else val / pr.H_ref
    s = _native().Solver(pr)
    o = _oracle(pr)
    s.okada(pr.faults[0], crop_flag=flag, crop_value=cv)
    o.okada(pr.faults[0], crop_flag=flag, crop_value=cv)
    eg, eo = s.field("eta"), o.output("eta")
    wet = eo > -9000
    assert np.array_equal(eg > -9000, wet)
    # the state before Okada is at rest (eta is round-off of recovering h from H and
    # Hmin): cropped cells keep that, the others carry the uplift (cm and more)
    zero_o = wet & (np.abs(eo) < 1e-6)
    zero_g = wet & (np.abs(eg) < 1e-6)
    assert np.array_equal(zero_g, zero_o) and zero_o.sum() > 100 and (wet & ~zero_o).sum() > 100
    assert np.min(np.abs(eo[wet & ~zero_o])) > (1e-3 if flag == 1 else 0.05 - 1e-6)
    peak = float(np.max(np.abs(eo[wet])))
    assert rel_linf(eg[wet], eo[wet], peak) < 1e-6
    s.close()
    o.close()


def test_dam_break_waf_everywhere():
    """hpos far above every depth: every face takes the WAF branch with the wet/dry
    treatment and limiters.  A 1-D dam break over a 10 m bed on the GPU, with the
    oracle's dt sequence, matches the oracle within the bar after 200 steps."""
    from thysea import problem as P
    nx, ny, d = 600, 8, 0.0001
    lon = d * np.arange(nx)
    lat = d * (np.arange(ny) - ny // 2)
    z = np.full((ny, nx), -10.0, np.float32)
    pr = P.build("dam", "dam", lon, lat, z, faults=[], borders=(-1, -1, -1, -1), hpos=1000.0,
                 manning=0.0, vmax=0.0, eps_h=1e-5)
    h = pr.h.reshape(ny, nx).copy()
    h[:, : nx // 2] += 1.0 / pr.H_ref
    pr.h = h
    s = _native().Solver(pr)
    o = _oracle(pr)
    dt = o.initial_dt()
    for _ in range(200):
        dn = o.step(dt)
        s.step(dt)
        dt = dn
    hg, Hg, qxg, qyg = s.state()
    ho, Ho, qxo, qyo = o.state()
    eta_o, eta_g = ho - Ho - pr.hmin, hg - Hg - pr.hmin
    peak = float(np.max(np.abs(eta_o)))
    e = rel_linf(eta_g, eta_o, peak)
    eq = rel_linf(qxg, qxo, float(np.max(np.abs(qxo))))
    print(f"dam break: eta {e:.3e}, qx {eq:.3e} (of its peak), qy max {np.max(np.abs(qyg)):.2e}")
    assert e < TOL and eq < TOL
    s.close()
    o.close()


def test_runup_coastal_band():
    """A wave that reaches the shelf and runs up the coast inside the compared window
    (fault next to the shelf, 400 steps, oracle dt replayed): eta on the coastal band
    (still-water depth below 200 m, wet in the oracle) within 1e-4 of that band's own
    peak, the same wet/dry mask everywhere, and the WAF branch exercised (cells wet
    with h <= hpos)."""
    pr = make_problem(nx=160, ny=96, seamounts=2, seed=8, coast=("east",))
    f = dict(pr.faults[0])
    f["lon_c"] = float(pr.lon[int(pr.nx * 0.62)])
    f["slip"] = 6.0
    pr.faults = [f]
    s = _native().Solver(pr)
    o = _oracle(pr)
    s.okada(f)
    o.okada(f)
    dt = o.initial_dt()
    shallow_wet = 0
    D = (pr.H + pr.hmin).reshape(pr.ny, pr.nx)
    band = (D * pr.H_ref < 200.0)
    for k in range(400):
        dn = o.step(dt)
        s.step(dt)
        dt = dn
        if k % 50 == 49:
            ho = o.state()[0]
            shallow_wet = max(shallow_wet, int(((ho > pr.eps_h) & (ho <= pr.hpos)).sum()))
    hg, Hg, qxg, qyg = s.state()
    ho, Ho, qxo, qyo = o.state()
    wet_o, wet_g = ho > pr.eps_h, hg > pr.eps_h
    assert np.array_equal(wet_g, wet_o), int((wet_g != wet_o).sum())
    cb = band & wet_o
    eta_o, eta_g = ho - Ho - pr.hmin, hg - Hg - pr.hmin
    peak = float(np.max(np.abs(eta_o[cb])))
    assert peak * pr.H_ref > 0.05, "the wave has not reached the coastal band"
    e = rel_linf(eta_g[cb], eta_o[cb], peak)
    qpk = float(np.max(np.abs(qxo[cb]))) + float(np.max(np.abs(qyo[cb])))
    eq = max(rel_linf(qxg[cb], qxo[cb], qpk), rel_linf(qyg[cb], qyo[cb], qpk))
    print(f"run-up band: {cb.sum()} cells, eta peak {peak * pr.H_ref:.3g} m, eta err {e:.3e}, q err {eq:.3e}, "
          f"shallow wet cells {shallow_wet}")
    assert shallow_wet > 20
    assert e < TOL and eq < TOL
    s.close()
    o.close()


def test_snapshot_interval_shorter_than_dt():
    """A snapshot interval shorter than dt: at most one snapshot per iteration, as the
    reference (src/GPU/ShallowWater.cu:404-409); the same snapshot times and step count
    as the oracle."""
    pr = make_problem(nx=64, ny=48, t_sim=60.0)
    pr.save_interval = 0.5 / pr.T  # dt is a few seconds
    tg, to = [], []
    s = _native().Solver(pr)
    rg = s.run(steps_per_sync=6, on_snapshot=lambda i, t: tg.append(t))
    ro = _oracle(pr).run(on_snapshot=lambda i, t: to.append(t))
    assert rg["steps"] == ro["steps"]
    assert len(tg) == len(to) == ro["steps"]
    np.testing.assert_allclose(tg, to, rtol=1e-5)
    assert len(set(tg)) == len(tg)
    s.close()


def test_health_check_stops_on_nonfinite():
    """A NaN in the input state is refused at upload; a discharge so large that the
    step overflows to Inf makes thy_step and thy_run return THY_ERR_NONFINITE (7), and
    in a batch that scenario is retired as failed while the other one finishes with the
    results of its run alone."""
    nat = _native()
    code = nat.STATUS_CODES["NONFINITE"]
    pr = make_problem(nx=96, ny=80, t_sim=300.0)
    nan = make_problem(nx=96, ny=80, t_sim=300.0)
    q = np.zeros((nan.ny, nan.nx))
    q[40, 30] = np.nan
    nan.qx = q
    with pytest.raises(nat.ThyError) as ei:
        nat.Solver(nan)
    assert ei.value.code == code and "non-finite input" in str(ei.value)
    bad = make_problem(nx=96, ny=80, t_sim=300.0)
    assert bad.h[40, 30] > 1.0  # deep water
    q = np.zeros((bad.ny, bad.nx))
    q[40, 30] = 3e38  # finite, but the fluxes of the step overflow
    bad.qx = q
    s = nat.Solver(bad)
    with pytest.raises(nat.ThyError) as ei:
        s.step(s.initial_dt())
    assert ei.value.code == code and "non-finite" in str(ei.value)
    s.close()
    s = nat.Solver(bad)
    with pytest.raises(nat.ThyError) as ei:
        s.run(steps_per_sync=6)
    assert ei.value.code == code and "after step 1" in str(ei.value)
    s.close()
    ref = nat.Solver(pr)
    ra = ref.run(steps_per_sync=6)
    ma = ref.field("max_height")
    ref.close()
    b = nat.Batch(capacity=2, steps_per_sync=6)
    sg, sb = nat.Solver(pr), nat.Solver(bad)
    kg, kb = b.add(sg), b.add(sb)
    while b.advance() > 0:
        pass
    assert b.status(kb)[0] == nat.SLOT_FAILED
    assert b.status(kg)[0] == nat.SLOT_FINISHED
    with pytest.raises(nat.ThyError):
        b.result(kb)
    r = b.result(kg)
    assert r["steps"] == ra["steps"] and np.array_equal(r["max_height"], ma)
    b.close()
    sg.close()
    sb.close()


def test_batch_slot_removed_in_flight_and_readmitted():
    """A running scenario removed while a launch is in flight and a fresh scenario
    admitted to the same slot before the launch is collected: the new scenario is not
    taken for the old one; both the survivor and the newcomer end as their runs alone."""
    nat = _native()
    from test_gpu_parity import _scenario
    pa, pb, pc = (_scenario(96, 80, 3, 0.4, 2.0, 300.0), _scenario(150, 64, 5, 0.5, 4.0, 400.0),
                  _scenario(70, 120, 9, 0.3, 1.0, 200.0))
    alone = {}
    for key, pr in (("b", pb), ("c", pc)):
        s = nat.Solver(pr)
        r = s.run(steps_per_sync=6)
        alone[key] = (r["steps"], s.field("max_height"), r["g_eta"])
        s.close()
    b = nat.Batch(capacity=2, steps_per_sync=6)
    sa, sb = nat.Solver(pa), nat.Solver(pb)
    ka, kb = b.add(sa), b.add(sb)
    b.advance()
    b.launch()
    b.remove(ka)  # waits for the launch; the slot's launched mark must go
    sc = nat.Solver(pc)
    kc = b.add(sc)
    assert kc == ka
    b.wait()
    assert b.status(kc)[0] == nat.SLOT_RUNNING and b.status(kc)[1]["steps"] == 0
    while b.advance() > 0:
        pass
    for key, slot in (("b", kb), ("c", kc)):
        r = b.result(slot)
        assert r["steps"] == alone[key][0]
        assert np.array_equal(r["max_height"], alone[key][1]), key
        assert np.array_equal(r["g_eta"], alone[key][2]), key
    b.close()
    for x in (sa, sb, sc):
        x.close()
