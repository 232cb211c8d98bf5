"""End to end on the GPU: the TsunamiHySEA executable (C++ host over the C-ABI) on a
parameter file, against the oracle's driver loop on the same file."""
import os
import subprocess

import numpy as np
import pytest
from scipy.io import netcdf_file

from conftest import ROOT, rel_linf

pytestmark = pytest.mark.gpu

EXE = os.path.join(ROOT, "tsunami-hysea_amd", "TsunamiHySEA")


def test_cli_outputs_match_oracle(tmp_path):
    """small workload, 600 s, snapshots every 300 s: the max height written to
    <prefix>.nc and the gauge series in <prefix>_ts.nc match the oracle (wet/dry mask
    identical, values within 1e-4 of the peak), and the run prints the
    reference's per-iteration lines."""
    import oracle as O
    from thysea import problem, scenario
    par = scenario.make("small", str(tmp_path), t_sim=600.0)
    txt = open(par).read().replace("-1.0  # Saving time of NetCDF files", "300  # Saving time of NetCDF files")
    open(par, "w").write(txt)
    r = subprocess.run([EXE, par], cwd=tmp_path, capture_output=True, text=True, timeout=100)
    assert r.returncode == 0, r.stderr
    assert "Iteration" in r.stdout or "iteration" in r.stdout.lower()
    pr = problem.load(par)
    o = O.OracleSim(pr)
    ro = o.run()
    mo = o.output("max_height")
    o.close()
    with netcdf_file(str(tmp_path / "small.nc"), "r", mmap=False) as nc:
        mh = np.array(nc.variables["max_height"][:])
        assert nc.variables["eta"].shape[0] >= 2
    wet = mo > -9000
    assert np.array_equal(mh > -9000, wet)
    peak = float(np.max(np.abs(mo[wet])))
    assert rel_linf(mh[wet], mo[wet], peak) < 1e-4
    with netcdf_file(str(tmp_path / "small_ts.nc"), "r", mmap=False) as nc:
        ge = np.array(nc.variables["eta"][:])
    n = min(ge.shape[0], ro["g_eta"].shape[0])
    assert n > 5
    assert rel_linf(ge[:n], ro["g_eta"][:n], peak) < 1e-4


def test_cli_initial_state_from_file_matches_oracle(tmp_path):
    """okada_flag = 0: the initial sea surface is read from a NetCDF grid
    (src/Problema.cxx:233-240,458-473) instead of an Okada source.  The executable's max
    height, arrival times (written to <prefix>.nc with the bathymetry variable of that
    mode) and gauge series match the oracle's run of the same parameter file."""
    import oracle as O
    from thysea import problem, scenario
    par = scenario.make("small", str(tmp_path), t_sim=400.0)
    lines = open(par).read().splitlines()
    k0 = next(i for i, l in enumerate(lines) if "Initialization type" in l)
    k1 = next(i for i, l in enumerate(lines) if "Crop deformations" in l)
    lines[k0:k1 + 1] = ["0  # Initialization type (0: From file, 1: Standard Okada)", "eta0.grd  # Initial state"]
    lines = [l.replace("-1.0  # Saving time of NetCDF files", "200  # Saving time of NetCDF files") for l in lines]
    open(par, "w").write("\n".join(lines) + "\n")
    lon, lat, z = problem.read_grd(str(tmp_path / "small.grd"))
    X, Y = np.meshgrid(lon, lat)
    cx, cy = lon[len(lon) // 3], lat[len(lat) // 2]
    eta = (1.5 * np.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / (2 * 0.3 ** 2))).astype(np.float32)
    eta = np.where(z < 0, eta, 0.0).astype(np.float32)
    scenario.write_grd(str(tmp_path / "eta0.grd"), lon, lat, eta)
    r = subprocess.run([EXE, par], cwd=tmp_path, capture_output=True, text=True, timeout=100)
    assert r.returncode == 0, r.stderr
    pr = problem.load(par)
    assert pr.okada_flag == 0 and not pr.faults
    o = O.OracleSim(pr)
    ro = o.run()
    mo, ao = o.output("max_height"), o.output("arrival_times")
    o.close()
    with netcdf_file(str(tmp_path / "small.nc"), "r", mmap=False) as nc:
        assert "bathymetry" in nc.variables and "deformed_bathy" not in nc.variables
        mh = np.array(nc.variables["max_height"][:])
        at = np.array(nc.variables["arrival_times"][:])
        assert nc.variables["eta"].shape[0] == 2
        e0 = np.array(nc.variables["eta"][0])
    wet = mo > -9000
    assert np.array_equal(mh > -9000, wet)
    peak = float(np.max(np.abs(mo[wet])))
    assert peak > 1.0
    assert rel_linf(mh[wet], mo[wet], peak) < 1e-4
    # the first snapshot is the initial state read from the file
    w0 = e0 > -9000
    assert rel_linf(e0[w0], eta[w0], 1.5) < 1e-6
    arrived = ao >= 0
    assert np.array_equal(at >= 0, arrived)
    dt_s = ro["t_final"] * pr.T / ro["steps"]
    assert np.max(np.abs(at[arrived] - ao[arrived])) <= 1.01 * dt_s
    with netcdf_file(str(tmp_path / "small_ts.nc"), "r", mmap=False) as nc:
        ge = np.array(nc.variables["eta"][:])
    n = min(ge.shape[0], ro["g_eta"].shape[0])
    assert n > 5
    assert rel_linf(ge[:n], ro["g_eta"][:n], peak) < 1e-4
