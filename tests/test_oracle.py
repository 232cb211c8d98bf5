"""Pins the FP64 oracle (oracle/thy_oracle.c) against what the mathematics fixes.

The reference has no tests, fixtures or golden vectors (its only example input is not
in its checkout), and it is CUDA-only, so it cannot be compiled here.  Each oracle
function is therefore checked against an invariant, a closed form, an independent
implementation of the published formula, or brute force, chosen so that a plausible
mistake (a dropped term, a flipped sign, a transposed index) fails.
"""
import math

import numpy as np
import pytest

import oracle as O
from conftest import make_problem


# ---------------------------------------------------------------------------
# invariants of the scheme
# ---------------------------------------------------------------------------

def test_lake_at_rest():
    """eta = const, q = 0 over arbitrary bathymetry with dry land stays at rest
    (every flux of src/GPU/Arista_kernel.cu vanishes: Delta eta = 0 and q = 0)."""
    pr = make_problem(nx=40, ny=32, fault=False, coast=("east", "north"))
    s = O.OracleSim(pr)
    h0 = pr.h.copy()
    dt = s.initial_dt()
    for _ in range(25):
        dt = s.step(dt)
    h, H, qx, qy = s.state()
    assert np.max(np.abs(qx)) < 1e-12 and np.max(np.abs(qy)) < 1e-12
    assert np.max(np.abs(h - h0)) < 1e-13


@pytest.mark.parametrize("periodic", [False, True])
def test_mass_conserved_with_walls(periodic):
    """With wall (or periodic) borders and no sponge, sum(h*area) is invariant:
    face fluxes are antisymmetric and Okada leaves h unchanged."""
    b = (-1, -1, 2, 2) if periodic else (-1, -1, -1, -1)
    pr = make_problem(nx=48, ny=36, borders=b)
    s = O.OracleSim(pr)
    m0 = s.mass()
    s.okada(pr.faults[0])
    assert abs(s.mass() - m0) / m0 < 1e-15
    dt = s.initial_dt()
    for _ in range(60):
        dt = s.step(dt)
    assert abs(s.mass() - m0) / m0 < 1e-12


def _stoker(hl, hr, g=9.81):
    """Wet-bed dam break (Stoker): intermediate depth and velocity, shock speed."""
    lo, hi = hr, hl
    for _ in range(200):
        hm = 0.5 * (lo + hi)
        ur = 2.0 * (math.sqrt(g * hl) - math.sqrt(g * hm))
        us = (hm - hr) * math.sqrt(g * (hm + hr) / (2.0 * hm * hr))
        if ur > us:
            lo = hm
        else:
            hi = hm
    um = 2.0 * (math.sqrt(g * hl) - math.sqrt(g * hm))
    return hm, um, hm * um / (hm - hr)


def test_dam_break_waf():
    """With hpos far above the depths every face uses the WAF flux (blend weight ~0);
    a 1-D dam break on a flat bed then reproduces Stoker's intermediate state and
    shock speed."""
    from thysea import problem as P
    nx, ny, d = 600, 3, 0.0001
    lon = d * np.arange(nx)
    lat = d * (np.arange(ny) - 1)
    z = np.full((ny, nx), -10.0, np.float32)
    pr = P.build("dam", "dam", lon, lat, z, faults=[], borders=(-1, -1, -1, -1), hpos=1000.0,
                 manning=0.0, vmax=0.0, eps_h=1e-5)
    s = O.OracleSim(pr)
    h = pr.h.copy()
    h[:, : nx // 2] += 1.0 / pr.H_ref
    s.set_state(h, pr.H, pr.qx, pr.qy)
    t, dt = 0.0, s.initial_dt()
    tend = 40.0 / pr.T
    while t < tend:
        dtn = s.step(dt)
        t += dt
        dt = dtn
    hh, HH, qx, _ = s.state()
    hm, um, S = _stoker(11.0, 10.0)
    dx = 6378136.6 * math.radians(d)
    row = hh[1] * pr.H_ref
    x = (np.arange(nx) - nx // 2 + 0.5) * dx
    tt = t * pr.T
    plateau = (x > 0.2 * um * tt) & (x < 0.8 * S * tt)
    assert plateau.sum() > 10
    assert abs(np.median(row[plateau]) - hm) / (hm - 10.0) < 0.03
    u = qx[1] / hh[1] * pr.Q / pr.H_ref
    assert abs(np.median(u[plateau]) - um) / um < 0.03
    xs = x[nx // 2:][np.argmax(row[nx // 2:] < 10.0 + 0.5 * (hm - 10.0))]
    assert abs(xs - S * tt) / (S * tt) < 0.05


def test_initial_dt_closed_form():
    """Flat still basin of depth d: every face speed is sqrt(d) (g = 1 after
    non-dimensionalisation), so dt = CFL*A / (sqrt(d)*(dlat/cos_max + dlon))."""
    nx, ny = 30, 24
    from thysea import problem as P
    lon = 140.0 + np.arange(nx) / 30.0
    lat = 10.0 + np.arange(ny) / 30.0
    z = np.full((ny, nx), -4000.0, np.float32)
    pr = P.build("flat", "flat", lon, lat, z, faults=[], borders=(1, 1, 1, 1))
    s = O.OracleSim(pr)
    d = 4000.0 / pr.H_ref
    A = pr.area[0]
    want = min(pr.cfl * A / (math.sqrt(d) * (pr.alto[j] / pr.cosphi[j] + pr.ancho[j])) for j in range(ny))
    assert abs(s.initial_dt() - want) / want < 1e-13


def test_linear_wave_speed():
    """A small hump on a flat basin spreads at sqrt(g d): the front reaches a point
    20 cells east at t ~ distance/sqrt(d) (non-dimensional)."""
    nx, ny = 90, 31
    from thysea import problem as P
    lon = 140.0 + np.arange(nx) / 30.0
    lat = np.arange(ny) / 30.0 - 0.5
    z = np.full((ny, nx), -4000.0, np.float32)
    pr = P.build("flat", "flat", lon, lat, z, faults=[], borders=(-1, -1, -1, -1))
    s = O.OracleSim(pr)
    x = np.arange(nx) - 20.0
    yy = np.arange(ny) - 15.0
    h = pr.h + 1e-4 * np.exp(-(x[None, :] ** 2 + yy[:, None] ** 2) / 4.0)
    s.set_state(h, pr.H, pr.qx, pr.qy)
    dx = pr.area[15] / pr.alto[15] / pr.cosphi[15] * pr.cosphi[15]  # R*dlon*pi/180 (non-dim) at the equator
    dx = (6378136.6 / pr.L) * (1.0 / 30.0) * math.pi / 180.0 * pr.cosphi[15]
    dist = 30 * dx
    t_expect = dist / math.sqrt(4000.0 / pr.H_ref)
    t, dt = 0.0, s.initial_dt()
    arrived = None
    while t < 2.0 * t_expect:
        dt_next = s.step(dt)
        t += dt
        dt = dt_next
        h_, H_, _, _ = s.state()
        if arrived is None and abs((h_ - H_)[15, 50] - (pr.h - pr.H)[15, 50]) > 2e-6:
            arrived = t
    assert arrived is not None
    assert abs(arrived - t_expect) / t_expect < 0.15


# ---------------------------------------------------------------------------
# reduction
# ---------------------------------------------------------------------------

def test_reduce_min_exact():
    rng = np.random.default_rng(3)
    for n in (1, 2, 7, 1000, 48000, 100003):
        v = rng.uniform(0.1, 10.0, n)
        assert O.reduce_min(v, quirk=0) == v.min()


def test_reduce_min_reference_quirk():
    """The reference's launch arithmetic (src/GPU/Reduccion_kernel.cu:130-163): for
    N >= 65536 both stages run and the result is the global minimum; for the 240x200
    grid the second stage asks for 46 threads, no kernel runs, and the host reads
    block 0's partial: cells [0,512) and [47616, 48000)."""
    rng = np.random.default_rng(5)
    v = rng.uniform(1.0, 2.0, 48000)
    v[30000] = 0.5                  # outside block 0's subset
    assert O.reduce_min(v, quirk=0) == 0.5
    want = min(v[:512].min(), v[47616:].min())
    assert O.reduce_min(v, quirk=1) == want
    w = rng.uniform(1.0, 2.0, 12000000 // 64)
    w[12345] = 0.25
    big = np.concatenate([w] * 64)
    assert O.reduce_min(big, quirk=1) == 0.25


# ---------------------------------------------------------------------------
# scalar closed forms
# ---------------------------------------------------------------------------

def test_limiter_properties():
    L = O.limiter
    # smooth (linear) data: no limiting
    assert abs(L(0.0, 1.0, 2.0, 3.0, 1.0) - 1.0) < 1e-12
    assert abs(L(0.0, 1.0, 2.0, 3.0, -1.0) - 1.0) < 1e-12
    # local extremum: full limiting (monotone variant)
    assert L(0.0, 1.0, 2.0, 0.5, -1.0) == 0.0
    assert L(1.5, 1.0, 2.0, 3.0, 1.0) == 0.0
    # flat left state: chi = 1 by the c0 < eps rule
    assert L(0.0, 1.0, 1.0, 5.0, 1.0) == 1.0
    # bounded, symmetric in the two ratios, increasing with smoothness
    vals = [L(0.0, 1.0, 2.0, 2.0 + r, -1.0) for r in (0.05, 0.2, 0.5, 1.0)]
    assert all(0.0 <= v <= 1.0 for v in vals)
    assert vals == sorted(vals)
    # the upwind side is chosen by the sign of the wave speed: S < 0 looks right
    # (p1p - p1), S > 0 looks left (p0 - p0m)
    assert L(0.0, 1.0, 2.0, 2.05, -1.0) < 0.2 < 0.9 < L(0.0, 1.0, 2.0, 2.05, 1.0)
    assert L(0.95, 1.0, 2.0, 3.0, 1.0) < 0.2 < 0.9 < L(0.95, 1.0, 2.0, 3.0, -1.0)
    # S = 0 uses half the three-interval difference 0.5*(p1p - p0m): same value as the
    # S > 0 branch fed a left slope of that size
    assert L(0.0, 1.0, 2.0, 3.0, 0.0) == L(1.0 - 1.5, 1.0, 2.0, 3.0, 1.0)


def test_desingularisation():
    eps = 1e-8
    for h in (1e-3, 0.1, 4.0):
        assert abs(O.desing(h, 0.37, eps) - 0.37 / h) / (0.37 / h) < 1e-12
    assert O.desing(0.0, 0.37, eps) == 0.0
    assert abs(O.desing(eps, 0.37, eps) - 0.37 / eps) / (0.37 / eps) < 1e-12
    # vanishes smoothly below eps: u ~ sqrt(2) h q / eps^2
    assert abs(O.desing(eps * 1e-3, 1.0, eps) - math.sqrt(2) * eps * 1e-3 / eps ** 2) / (math.sqrt(2) / eps * 1e-3) < 1e-6


def test_manning_semi_implicit():
    """Away from the dry threshold the filter is the identity and the friction is the
    textbook semi-implicit Manning update q' = q*/(1 + dt*mf0*|u^n|/h^(4/3))."""
    hs, dt, mf0, eps = 0.01, 0.004, 0.0883, 1e-8
    qxo, qyo = 2e-4, -1e-4
    qx, qy = 3e-4, 5e-5
    ax, ay = O.friction_filter(hs, qxo, qyo, qx, qy, dt, mf0, 0.0, eps)
    un = math.hypot(qxo, qyo) / hs
    f = 1.0 / (1.0 + dt * mf0 * un / hs ** (4.0 / 3.0))
    assert abs(ax - qx * f) < 1e-15 and abs(ay - qy * f) < 1e-15


def test_velocity_cap_and_dry_filter():
    hs, dt, eps = 1e-3, 0.004, 1e-8
    vmax = 0.5
    ax, ay = O.friction_filter(hs, 0.0, 0.0, 3e-3, 4e-3, dt, 0.0, vmax, eps)  # |u| = 5 > vmax
    assert abs(math.hypot(ax, ay) / hs - vmax) < 1e-12
    assert abs(ax / ay - 0.75) < 1e-12
    # a dry cell (h* = 0) keeps no momentum
    ax, ay = O.friction_filter(0.0, 1e-3, 1e-3, 1e-3, 1e-3, dt, 0.1, vmax, eps)
    assert ax == 0.0 and ay == 0.0


# ---------------------------------------------------------------------------
# Okada and Gauss-Krueger
# ---------------------------------------------------------------------------

def okada85_uz(x, y, depth, dip, L1, L2, W1, W2, u1, u2, mu_over=0.5):
    """Okada (1985, BSSA 75:1135) surface vertical displacement, eqs. (25)-(30), with
    Chinnery's notation over the corners (xi = x - L, eta = p - W).  Written from the
    paper, independently of DC3D (Okada 1992)."""
    d = math.radians(dip)
    sd, cd = math.sin(d), math.cos(d)
    p = y * cd + depth * sd
    q = y * sd - depth * cd
    tot = 0.0
    for sx, Lc in ((1, L1), (-1, L2)):
        for se, Wc in ((1, W1), (-1, W2)):
            xi, eta = x - Lc, p - Wc
            R = math.sqrt(xi * xi + eta * eta + q * q)
            dt_ = eta * sd - q * cd
            X = math.sqrt(xi * xi + q * q)
            I4 = mu_over / cd * (math.log(R + dt_) - sd * math.log(R + eta))
            I5 = mu_over * 2.0 / cd * math.atan((eta * (X + q * cd) + X * (R + X) * sd) / (xi * (R + X) * cd))
            ss = dt_ * q / (R * (R + eta)) + q * sd / (R + eta) + I4 * sd
            ds = dt_ * q / (R * (R + xi)) + sd * math.atan(xi * eta / (q * R)) - I5 * sd * cd
            tot += sx * se * (-u1 / (2 * math.pi) * ss - u2 / (2 * math.pi) * ds)
    return tot


def test_dc3d_surface_vs_okada85():
    rng = np.random.default_rng(11)
    worst = 0.0
    for _ in range(400):
        dip = rng.uniform(10.0, 80.0)
        L, W = rng.uniform(20, 200), rng.uniform(10, 80)
        depth = W / 2 * math.sin(math.radians(dip)) + rng.uniform(2, 30)
        x, y = rng.uniform(-300, 300), rng.uniform(-300, 300)
        u1, u2 = rng.uniform(-5, 5), rng.uniform(-5, 5)
        _, _, uz, ir = O.dc3d_surface(2.0 / 3.0, x, y, depth, dip, -L / 2, L / 2, -W / 2, W / 2, u1, u2, 0.0)
        assert ir == 0
        ref = okada85_uz(x, y, depth, dip, -L / 2, L / 2, -W / 2, W / 2, u1, u2)
        worst = max(worst, abs(uz - ref) / max(abs(ref), 1e-6 * (abs(u1) + abs(u2))))
    assert worst < 1e-8


def test_okada_thrust_sign_and_decay():
    """A thrust (rake 90) uplifts more than it subsides; displacement decays with
    distance (far field ~ r^-2 or faster)."""
    up = [O.dc3d_surface(2 / 3, 0.0, y, 22.5, 30.0, -50, 50, -25, 25, 0.0, 3.0, 0.0)[2] for y in np.linspace(-150, 150, 301)]
    assert max(up) > 0.3 and max(up) > -min(up)
    far = [abs(O.dc3d_surface(2 / 3, r, r, 22.5, 30.0, -50, 50, -25, 25, 0.0, 3.0, 0.0)[2]) for r in (500.0, 1000.0, 2000.0)]
    assert far[1] < far[0] / 3.5 and far[2] < far[1] / 3.5


def test_dc3d_singular_edge_is_zero():
    # a point on the fault's surface trace when the fault reaches z = 0 (q = 0 on an edge)
    ux, uy, uz, ir = O.dc3d_surface(2 / 3, 0.0, 0.0, 0.0, 90.0, -10, 10, -5, 0, 1.0, 1.0, 0.0)
    assert ir == 1 and uz == 0.0


def test_gauss_krueger_round_trip_and_limit():
    lat0, lon0 = 35.7, 22.3
    for dlat, dlon in ((0.5, 0.5), (-1.2, 2.0), (3.0, -4.0)):
        x, y = O.prtoxy(lat0 + dlat, lon0 + dlon, lat0, lon0, 0)
        la, lo = O.prtoxy(0.0, 0.0, lat0, lon0, 1, x, y)
        assert abs(la - (lat0 + dlat)) < 2e-3 and abs(lo - (lon0 + dlon)) < 2e-3
    # small offsets: y ~ M dphi, x ~ N cos(phi) dlambda (Bessel-like ellipsoid a=6378.16 km)
    a, e2 = 6378.16, 0.0066946053
    phi = math.radians(lat0)
    M = a * (1 - e2) / (1 - e2 * math.sin(phi) ** 2) ** 1.5
    N = a / math.sqrt(1 - e2 * math.sin(phi) ** 2)
    x, y = O.prtoxy(lat0 + 0.001, lon0, lat0, lon0, 0)
    assert abs(y - M * math.radians(0.001)) / (M * math.radians(0.001)) < 1e-4
    x, y = O.prtoxy(lat0, lon0 + 0.001, lat0, lon0, 0)
    assert abs(x - N * math.cos(phi) * math.radians(0.001)) / (N * math.cos(phi) * math.radians(0.001)) < 1e-4


def test_okada_field_on_grid_conserves_h():
    pr = make_problem(nx=40, ny=30)
    s = O.OracleSim(pr)
    h0, H0, _, _ = s.state()
    d = s.okada(pr.faults[0], want_def=True)
    h1, H1, _, _ = s.state()
    assert np.array_equal(h0, h1)
    assert np.allclose(H0 - H1, d, rtol=0, atol=1e-15)
    assert np.max(d) > 0
