"""Host-side checks that need no GPU: C-ABI exports, the C++ parameter-file parser and
NetCDF codec (TsunamiHySEA --check) against the Python mirror, output file structure,
and the reference's error messages for bad parameter files."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest
from scipy.io import netcdf_file

from conftest import ROOT

PKG = os.path.join(ROOT, "tsunami-hysea_amd")
EXE = os.path.join(PKG, "TsunamiHySEA")
HEADER = os.path.join(ROOT, "include", "thysea.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(thy_[a-z_]+)\s*\(", src)) - {"thy_snapshot_fn", "thy_okada_fn", "thy_step_fn"})


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(os.path.join(PKG, "libthysea.so"))
    names = declared_functions()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n


def test_shallow_water_library_exports_the_reference_boundary():
    """libthysea_sw.so exports shallowWater, declared in include/thysea_sw.h with the
    reference's 53 parameters (src/GPU/ShallowWater.cu:182-190)."""
    src = re.sub(r"//[^\n]*", "", open(os.path.join(ROOT, "include", "thysea_sw.h")).read())
    decl = re.search(r'extern "C" int shallowWater\((.*?)\);', src, flags=re.S).group(1)
    assert len([a for a in decl.split(",") if a.strip()]) == 53
    lib = ctypes.CDLL(os.path.join(PKG, "libthysea_sw.so"))
    assert hasattr(lib, "shallowWater")


def _scenario(tmp_path, **kw):
    from thysea import scenario
    return scenario.make("small", str(tmp_path), **kw)


def test_check_mode_matches_python_setup(tmp_path):
    from thysea import problem
    par = _scenario(tmp_path)
    dump = str(tmp_path / "dump.nc")
    r = subprocess.run([EXE, "--check", par, dump], cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    pr = problem.load(par)
    with netcdf_file(dump, "r", mmap=False) as nc:
        v = nc.variables
        assert np.array_equal(np.array(v["h"][:]), pr.h)
        assert np.array_equal(np.array(v["H"][:]), pr.H)
        for k in ("area", "cosphi", "ancho", "alto"):
            assert np.array_equal(np.array(v[k][:]), getattr(pr, k)), k
        assert np.array_equal(np.array(v["gauge_pos"][:]), pr.gauge_pos)
        g = nc._attributes
        for k, want in (("hmin", pr.hmin), ("mf0", pr.mf0), ("vmax", pr.vmax), ("eps_h", pr.eps_h),
                        ("hpos", pr.hpos), ("cvis", pr.cvis), ("dif_at", pr.dif_at), ("t_end", pr.t_end),
                        ("t_series", pr.series_interval), ("T", pr.T), ("Q", pr.Q)):
            assert float(g[k]) == want, k
        f = np.array(v["faults"][:])[0]
        assert list(f) == [pr.faults[0][k] for k in ("time", "lon_c", "lat_c", "depth_c", "length", "width",
                                                     "strike", "dip", "rake", "slip")]


def test_output_files_structure(tmp_path):
    """Names, dimensions, types and attributes of <prefix>.nc and <prefix>_ts.nc as
    the reference defines them (src/GPU/netcdfTH.cu:118-428, netcdfSeries.cu:51-321)."""
    from thysea import scenario
    par = scenario.make("small", str(tmp_path))
    txt = open(par).read().replace("-1.0  # Saving time of NetCDF files", "600  # Saving time of NetCDF files")
    open(par, "w").write(txt)
    r = subprocess.run([EXE, "--check", par, str(tmp_path / "d.nc")], cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    with netcdf_file(str(tmp_path / "small.nc"), "r", mmap=False) as nc:
        assert set(nc.dimensions) == {"lon", "lat", "grid_lon", "grid_lat", "time"}
        assert nc.dimensions["time"] is None
        want = {"lon": "d", "lat": "d", "grid_lon": "d", "grid_lat": "d", "original_bathy": "f",
                "deformed_bathy": "f", "time": "f", "max_height": "f", "arrival_times": "f",
                "eta": "f", "ux": "f", "uy": "f"}
        assert set(nc.variables) == set(want)
        for k, t in want.items():
            assert nc.variables[k].typecode() == t, k
        assert nc.variables["eta"].dimensions == ("time", "lat", "lon")
        assert nc.variables["max_height"].dimensions == ("lat", "lon")
        assert nc.variables["original_bathy"].dimensions == ("grid_lat", "grid_lon")
        assert nc.variables["eta"].long_name == b"Wave amplitude"
        assert float(nc.variables["max_height"]._FillValue) == -9999.0
        assert nc.variables["time"].units == b"seconds since 1970-01-01"
        g = nc._attributes
        assert g["Conventions"] == b"CF-1.0" and g["initialization_mode"] == b"okada_standard"
        assert int(g["num_faults"]) == 1 and g["crop_deformations"] == b"no"
        assert g["fault_1"].startswith(b"time: 0.0000, lon: ")
        assert g["left_border"] == b"open"
        z = np.array(nc.variables["original_bathy"][:])
        from thysea import problem
        pr = problem.load(par)
        assert np.allclose(z, ((pr.H + pr.hmin) * pr.H_ref).astype(np.float32), rtol=0, atol=1e-3)
        assert nc.variables["eta"].shape == (1, pr.ny, pr.nx)
    with netcdf_file(str(tmp_path / "small_ts.nc"), "r", mmap=False) as nc:
        assert set(nc.dimensions) == {"grid_npoints", "time"}
        assert set(nc.variables) == {"deformed_bathy", "time", "longitude", "latitude", "min_height",
                                     "max_height", "eta", "ux", "uy"}
        assert nc.variables["eta"].dimensions == ("time", "grid_npoints")
        assert nc.variables["longitude"].typecode() == "d"
        assert nc._attributes["title"] == b"Time series output of TsunamiHySEA model"


@pytest.mark.parametrize("edit,msg", [
    (("1  # Initialization type", "3  # Initialization type"), "Error: The initialization flag should be 0 or 1"),
    (("1  # Number of faults", "0  # Number of faults"), "Error: The number of faults should be between 1 and 5000"),
    (("0  # Crop deformations", "7  # Crop deformations"), "Error: The cropping flag should be 0, 1 or 2"),
    (("1  # Number of levels", "2  # Number of levels"), "Error: The number of levels should be 1"),
    (("1  # border", "5  # border"), "Error: The boundary conditions should be -1, 1 or 2"),
    (("1  # border", "2  # border"), "Error: Periodic boundary conditions are only supported for left and right borders"),
])
def test_parameter_file_errors(tmp_path, edit, msg):
    from thysea import problem
    par = _scenario(tmp_path)
    txt = open(par).read().replace(edit[0], edit[1], 1)
    open(par, "w").write(txt)
    r = subprocess.run([EXE, "--check", par, str(tmp_path / "d.nc")], cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode != 0
    assert msg in r.stderr
    with pytest.raises(problem.ProblemError, match=re.escape(msg)):
        problem.load(par)


def test_missing_files(tmp_path):
    r = subprocess.run([EXE, str(tmp_path / "nope.txt")], capture_output=True, text=True)
    assert r.returncode != 0 and "Error: File '" in r.stderr
    par = _scenario(tmp_path)
    os.remove(str(tmp_path / "small.grd"))
    r = subprocess.run([EXE, "--check", par, str(tmp_path / "d.nc")], capture_output=True, text=True)
    assert r.returncode != 0 and "small.grd' not found" in r.stderr


def test_usage_without_arguments():
    r = subprocess.run([EXE], capture_output=True, text=True)
    assert r.returncode != 0 and "dataFile format:" in r.stderr


def test_grd_reader_variants(tmp_path):
    """x/y variable names (README.md:55) and CDF-1 vs CDF-2 containers."""
    from thysea import problem, scenario
    lon, lat, z = scenario.bathymetry(30, 20, 140.0, 30.0, 1, 3)
    for ver, (xn, yn) in ((1, ("x", "y")), (2, ("lon", "lat"))):
        path = str(tmp_path / f"g{ver}.grd")
        with netcdf_file(path, "w", version=ver) as nc:
            nc.createDimension(xn, len(lon))
            nc.createDimension(yn, len(lat))
            nc.createVariable(xn, "d", (xn,))[:] = lon
            nc.createVariable(yn, "d", (yn,))[:] = lat
            nc.createVariable("z", "f", (yn, xn))[:] = z
        par = str(tmp_path / f"p{ver}.txt")
        scenario.write_param_file(par, "t", path, [dict(time=0, lon_c=141, lat_c=30.3, depth_c=20, length=10,
                                                         width=5, strike=0, dip=30, rake=90, slip=1)], "t", 60.0)
        r = subprocess.run([EXE, "--check", par, str(tmp_path / f"d{ver}.nc")], cwd=tmp_path,
                           capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
        with netcdf_file(str(tmp_path / f"d{ver}.nc"), "r", mmap=False) as nc:
            assert np.array_equal(np.array(nc.variables["H"][:]), problem.load(par).H)
