"""Pins the oracle's output conversions, metrics, gauge gather, sponge and fault-time
truncation (oracle/thy_oracle.c: orc_output_field, orc_update_metrics, orc_gauges,
the sponge of the mass update, orc_run's truncation) against closed forms on
hand-built states and against brute force over the saved states of a run.

Each check is written from the reference's definition of the output
(src/GPU/netcdf_kernel.cu:6-126, src/GPU/Metrics.cu:6-77,
src/GPU/Volumen_kernel.cu:6-14,96-158,223-254, src/GPU/ShallowWater.cu:69-90,397-436,
src/GPU/Timestep.cu:35-68) applied to numbers the test builds itself, so that a
wrong sign, scale, mask, threshold, index or loop order in the oracle fails.
"""
import math

import numpy as np
import pytest

import oracle as O
from conftest import make_problem


def _hand_state(pr, seed=3):
    """A state with wet, nearly dry (0 < h < eps_h) and dry (h = 0) cells and random
    discharges."""
    rng = np.random.default_rng(seed)
    ny, nx = pr.ny, pr.nx
    H = pr.H.reshape(ny, nx)
    eta = rng.uniform(-2e-3, 2e-3, (ny, nx))
    h = np.maximum(H + pr.hmin + eta, 0.0)
    h[1, 1:6] = pr.eps_h * np.array([0.0, 0.1, 0.5, 0.99, 1.5])
    h[2, 2] = 0.0
    qx = rng.uniform(-1e-3, 1e-3, (ny, nx))
    qy = rng.uniform(-1e-3, 1e-3, (ny, nx))
    return h, H, qx, qy


def test_output_fields_closed_form():
    """eta = (h - H - Hmin) H_ref, -9999 where h < eps_h; u, v = q Q/(h H_ref), or the
    desingularised sqrt(2) h/sqrt(h^4 + eps^4) Q/H_ref below eps_h; deformed bathymetry
    (H + Hmin) H_ref (src/GPU/netcdf_kernel.cu:6-88)."""
    pr = make_problem(nx=24, ny=18, coast=("east",))
    h, H, qx, qy = _hand_state(pr)
    s = O.OracleSim(pr)
    s.set_state(h, H, qx, qy)
    Hr, Q, eps = pr.H_ref, pr.Q, pr.eps_h
    eta = np.where(h < eps, -9999.0, (h - H - pr.hmin) * Hr).astype(np.float32)
    fac = np.where(h < eps, math.sqrt(2.0) * h / np.sqrt(h ** 4 + eps ** 4) * (Q / Hr),
                   Q / (np.maximum(h, 1e-300) * Hr))
    ux = (qx * fac).astype(np.float32)
    uy = (qy * fac).astype(np.float32)
    bat = ((H + pr.hmin) * Hr).astype(np.float32)
    assert (h < eps).sum() >= 5 and (h == 0).sum() >= 2
    np.testing.assert_array_equal(s.output("eta"), eta)
    np.testing.assert_allclose(s.output("ux"), ux, rtol=2e-7, atol=0)
    np.testing.assert_allclose(s.output("uy"), uy, rtol=2e-7, atol=0)
    np.testing.assert_array_equal(s.output("deformed_bathy"), bat)
    s.close()


def test_metrics_outputs_closed_form():
    """max_height = (eta_max - Hmin) H_ref with -9999 where the cell was never wet
    (eta_max starts at -1e30 on dry cells); arrival = t T with -1 where it never came.
    eta_ini is initialised by the same kernel as eta_max (src/GPU/ShallowWater.cu:305-308),
    so it is -1e30 on cells dry at the start, which therefore count as arrived at the
    first update (src/GPU/Metrics.cu:6-77, src/GPU/netcdf_kernel.cu:92-126)."""
    pr = make_problem(nx=24, ny=18, coast=("east",))
    h, H, qx, qy = _hand_state(pr)
    s = O.OracleSim(pr)
    s.set_state(h, H, qx, qy)
    s.init_metrics()
    never_wet = h < pr.eps_h
    mh = s.output("max_height")
    want = np.where(never_wet, -9999.0, (h - H - pr.hmin) * pr.H_ref).astype(np.float32)
    np.testing.assert_array_equal(mh, want)
    np.testing.assert_array_equal(s.output("arrival_times"), np.full(h.shape, -1.0, np.float32))
    # raise eta by more than dif_at in a block of wet cells, update at t = 0.25
    h2 = h.copy()
    blk = (slice(5, 9), slice(3, 10))
    h2[blk] += np.where(h[blk] > pr.eps_h, 5 * pr.dif_at, 0.0)
    s.set_state(h2, H, qx, qy)
    s.update_metrics(0.25)
    arr = s.output("arrival_times")
    eta_ini = np.where(h < pr.eps_h, -1e30, h - H)
    moved = np.abs((h2 - H) - eta_ini) > pr.dif_at
    assert moved[blk].sum() > 20 and (moved & (h < pr.eps_h)).sum() >= 5
    np.testing.assert_array_equal(arr, np.where(moved, np.float32(0.25 * pr.T), np.float32(-1.0)))
    # a second crossing does not move the arrival time
    s.update_metrics(0.5)
    np.testing.assert_array_equal(s.output("arrival_times"), arr)
    s.close()


def test_gauges_closed_form():
    """Gauge eta = (h - H - Hmin) H_ref, not dry-masked; u, v = q/h Q/H_ref (desingularised
    below eps_h); -9999 for a gauge outside the grid (src/GPU/Volumen_kernel.cu:223-254,
    src/GPU/ShallowWater.cu:69-90)."""
    pr = make_problem(nx=24, ny=18, coast=("east",))
    h, H, qx, qy = _hand_state(pr)
    s = O.OracleSim(pr)
    s.set_state(h, H, qx, qy)
    nx = pr.nx
    cells = [(4, 7), (1, 1), (1, 3), (2, 2), (10, 20)]
    pos = np.array([j * nx + i for j, i in cells] + [-1], np.int32)
    e, u, v = s.gauges(pos)
    Hr, Q, eps = pr.H_ref, pr.Q, pr.eps_h
    for g, (j, i) in enumerate(cells):
        hh = h[j, i]
        f = math.sqrt(2.0) * hh / math.sqrt(hh ** 4 + eps ** 4) if hh < eps else 1.0 / hh
        assert e[g] == np.float32((hh - H[j, i] - pr.hmin) * Hr)
        assert u[g] == pytest.approx(np.float32(f * qx[j, i] * Q / Hr), rel=2e-7, abs=0)
        assert v[g] == pytest.approx(np.float32(f * qy[j, i] * Q / Hr), rel=2e-7, abs=0)
    assert (e[-1], u[-1], v[-1]) == (-9999.0, -9999.0, -9999.0)
    s.close()


def test_metrics_brute_force_over_saved_states():
    """After a run, eta_max is the maximum of eta over the wet states *before* each
    step (the final state excluded) and the arrival time the first such state's time
    with |eta - eta_ini| > dif_at, by brute force over the saved states
    (src/GPU/ShallowWater.cu:305-308,423-428)."""
    pr = make_problem(nx=120, ny=96, t_sim=200.0)
    pr.dif_at = 5e-3 / pr.H_ref  # 5 mm: the wave front reaches part of the grid
    s = O.OracleSim(pr)
    ref = O.OracleSim(pr)
    # the run: Okada at t = 0 after the metrics were initialised from the pre-Okada state
    res = s.run()
    n = res["steps"]
    assert n > 10
    # the same run by hand, saving the state before each step
    h0, H0, _, _ = ref.state()
    eta_ini = np.where(h0 < pr.eps_h, -1e30, h0 - H0)  # same initialisation as eta_max
    U = ref.okada(pr.faults[0], want_def=True)
    eta_ini = eta_ini + U  # Okada: H -= U_Z and eta_ini += U_Z (src/GPU/Deformacion.cu:38-53)
    wet0 = h0 >= pr.eps_h
    emax = np.where(wet0, h0 - H0, -1e30)  # initialised from the pre-Okada state
    tarr = np.full(h0.shape, -1.0)
    dt = res["dt_initial"]
    # one fault, at t = 0: the skipping variant finds no later fault time, so no dt is
    # truncated
    t = 0.0
    for k in range(n):
        h, H, _, _ = ref.state()
        eta = h - H
        upd = (eta > emax) & (h > pr.eps_h)
        emax = np.where(upd, eta, emax)
        hit = (tarr < 0) & (np.abs(eta - eta_ini) > pr.dif_at)
        tarr = np.where(hit, t, tarr)
        dn = ref.step(dt)
        t += dt
        dt = dn
    saved_final = ref.state()
    assert abs(t - res["t_final"]) < 1e-12 * max(1.0, t)
    got_max, got_arr, got_ini = s.metrics()
    np.testing.assert_array_equal(got_ini, eta_ini)
    np.testing.assert_array_equal(got_max, emax)
    np.testing.assert_array_equal(got_arr, tarr)
    assert (tarr > 0).sum() > 50 and (tarr < 0).sum() > 50
    # the final state is not included: including it would change eta_max somewhere
    hf, Hf, _, _ = saved_final
    with_final = np.where((hf - Hf > emax) & (hf > pr.eps_h), hf - Hf, emax)
    assert (with_final != emax).sum() > 0
    s.close()
    ref.close()


def test_sponge_relaxation_closed_form():
    """A flat basin at rest with a uniform offset eta0 and open borders: every flux
    vanishes in the mass pass, so one step changes eta only in the 4-cell sponges, by the
    relaxation h -= 0.9 (d/l)^2 (h - h0) with d = (n - i - 0.5) dx from the inner sponge
    edge and l = n dx (src/GPU/Volumen_kernel.cu:6-14,122-154).  The momentum pass then
    sees the relaxed h* only next to the sponges: momentum stays 0 further in."""
    from thysea import problem as P
    nx, ny, d = 40, 60, 1.0 / 30.0
    lon = 140.0 + d * np.arange(nx)
    lat = 10.0 + d * np.arange(ny)
    z = np.full((ny, nx), -3000.0, np.float32)
    pr = P.build("sp", "sp", lon, lat, z, faults=[], borders=(1, 1, 1, 1), hpos=20.0, eps_h=1e-5)
    eta0 = 0.05 / pr.H_ref
    h = pr.h.reshape(ny, nx) + eta0
    s = O.OracleSim(pr)
    s.set_state(h, pr.H, pr.qx, pr.qy)
    dt = s.initial_dt()
    s.step(dt)
    h1, H1, qx1, qy1 = s.state()
    eta1 = h1 - H1 - pr.hmin
    n = pr.sponge_izq
    inner = (slice(n + 2, ny - n - 2), slice(n + 2, nx - n - 2))
    assert np.max(np.abs(qx1[inner])) == 0.0 and np.max(np.abs(qy1[inner])) == 0.0
    assert np.max(np.abs(qx1)) > 0.0
    assert n == 4
    mid = slice(10, ny - 10)  # rows where the x sponges are nearest
    for i in range(n):
        r = (n - i - 0.5) / n
        want = eta0 * (1.0 - 0.9 * r * r)
        np.testing.assert_allclose(eta1[mid, i], want, rtol=0, atol=1e-12 * eta0 + 1e-15)
        np.testing.assert_allclose(eta1[mid, nx - 1 - i], want, rtol=0, atol=1e-12 * eta0 + 1e-15)
    cols = slice(10, nx - 10)
    for j in range(n):
        r = (n - j - 0.5) / n
        want = eta0 * (1.0 - 0.9 * r * r)
        np.testing.assert_allclose(eta1[j, cols], want, rtol=0, atol=1e-12 * eta0 + 1e-15)
        np.testing.assert_allclose(eta1[ny - 1 - j, cols], want, rtol=0, atol=1e-12 * eta0 + 1e-15)
    np.testing.assert_allclose(eta1[n:ny - n, n:nx - n], eta0, rtol=0, atol=1e-15)
    s.close()


def _fault(pr, t, frac, slip):
    f = dict(pr.faults[0])
    f["time"] = float(t)
    f["lon_c"] = float(pr.lon[int(pr.nx * frac)])
    f["slip"] = float(slip)
    return f


def fault_schedule(pr, t1=60.0, t2=150.0):
    """Faults at t = 0, two at t1 and one at t2 (seconds)."""
    return [_fault(pr, 0.0, 0.5, 2.0), _fault(pr, t1, 0.3, 1.5), _fault(pr, t1, 0.7, 1.0),
            _fault(pr, t2, 0.45, 0.8)]


def test_fault_schedule_truncation_and_application():
    """Faults at t = 0, t1, t1 and t2: the dt sequence lands on t1 and on t2 (the step
    before each is shortened to reach it), the two faults at t1 are applied together at
    the first iteration whose t T reaches t1 and the last at the first reaching t2, and
    every other dt is the raw CFL minimum (src/GPU/Timestep.cu:35-68,
    src/GPU/ShallowWater.cu:125-147,378-436)."""
    pr = make_problem(nx=64, ny=48, t_sim=240.0)
    pr.faults = fault_schedule(pr)
    T = pr.T
    t1, t2 = 60.0, 150.0
    # displacement of each fault alone
    U = []
    for f in pr.faults:
        q = O.OracleSim(pr)
        U.append(q.okada(f, want_def=True))
        q.close()
    s = O.OracleSim(pr)
    pr.save_interval = 0.0  # a snapshot callback every iteration: (t, H) after the faults due
    log = []

    def snap(i, t):
        log.append((t, s.state()[1].copy()))

    res = s.run(on_snapshot=snap)
    ts = np.array([t for t, _ in log])
    assert len(ts) == res["steps"] + 0 and res["steps"] > 40
    H0 = pr.H.reshape(pr.ny, pr.nx)
    # H at each iteration: the faults applied so far
    applied = []
    for t, H in log:
        k = 0
        for m in range(len(U) + 1):
            if np.allclose(H, H0 - sum(U[:m]), rtol=0, atol=1e-14):
                k = m
        applied.append(k)
    applied = np.array(applied)
    assert applied[0] == 1  # the t = 0 fault before the first step
    i1 = int(np.argmax(applied >= 3))
    i2 = int(np.argmax(applied >= 4))
    assert applied[i1] == 3 and applied[i2] == 4 and (applied[:i1] == 1).all()  # both t1 faults at once
    for i_land, tf in ((i1, t1), (i2, t2)):
        # the iteration applying the fault is the first with t T >= tf; the dt sequence
        # landed on tf at it or one iteration before (when t T rounded to just below tf)
        assert ts[i_land] * T >= tf and ts[i_land - 1] * T < tf
        land = i_land if abs(ts[i_land] * T - tf) < 1e-9 * tf else i_land - 1
        assert abs(ts[land] * T - tf) < 1e-9 * tf
    s.close()
