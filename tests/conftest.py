"""Shared test setup.

Markers: ``gpu`` tests need a gfx950 device and call the HIP library through the
C-ABI; everything else runs on the CPU.  The oracle (oracle/) is imported here only
as the checker.
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "tsunami-hysea_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def small_problem():
    from thysea import scenario
    return scenario.problem("small")


def make_problem(nx=48, ny=40, seed=7, borders=(1, 1, 1, 1), coast=("east",), seamounts=3,
                 fault=True, t_sim=600.0, eps_h=1e-5, lon0=140.0, lat0=30.0, noise=50.0):
    """A small seeded problem built in memory (no files)."""
    from thysea import problem as P
    from thysea import scenario as S
    lon, lat, z = S.bathymetry(nx, ny, lon0, lat0, seamounts, seed, coast=coast, noise=noise)
    faults = []
    if fault:
        faults = [dict(time=0.0, lon_c=float(lon[nx // 2]), lat_c=float(lat[ny // 2]), depth_c=22.5,
                       length=60.0, width=30.0, strike=90.0, dip=30.0, rake=90.0, slip=3.0)]
    pr = P.build("t", "t", lon, lat, z, faults=faults, borders=borders, t_tot=t_sim, eps_h=eps_h,
                 hpos=20.0, stability=0.2, manning=0.03, vmax=100.0, dif_at=1e-6)
    return pr


def rel_linf(a, b, scale):
    return float(np.max(np.abs(np.asarray(a, np.float64) - np.asarray(b, np.float64))) / scale)
