"""Seeded synthetic scenarios: bathymetry, Okada faults and gauges.

The reference ships no runnable input (its only example grid is missing from the
checkout), so every workload here is generated.  Randomness comes from SplitMix64
on integer counters, evaluated in numpy with wrap-around uint64 arithmetic, so a
seed gives the same grid on every platform.

Workloads (all at 2 arc-minute spacing):
  * ``small``  240x200, 140.0E/30.0N origin, 5 seamounts, east shelf; one fault
    (L 100 km, W 50 km, top edge 10 km deep -> centroid 22.5 km, strike 90,
    dip 30, rake 90, slip 3 m); 3600 s; 16 gauges.
  * ``large``  4000x3000, 100.0E/50.0S origin, 20 seamounts, coastline along the
    east and north edges; four 100x50 km subfaults with 2 m slip; 1800 s; 64 gauges.
  * ``precision`` 2000x2000; Mw 8.5 source (L 300, W 100 km, slip 8 m); 7200 s.
  * ``batch``  1000x800 grid shared by many faults drawn from magnitude scaling.
"""
import math
import os

import numpy as np

from . import problem as P

ARCMIN2 = 1.0 / 30.0
KM_PER_DEG = 111.19492664455873  # 2*pi*6371/360


def splitmix64(seed, idx):
    """SplitMix64 output for counters ``idx`` under ``seed`` (uint64 arrays)."""
    with np.errstate(over="ignore"):
        z = (np.uint64(seed) + (np.asarray(idx, dtype=np.uint64) + np.uint64(1)) *
             np.uint64(0x9E3779B97F4A7C15))
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniform(seed, n, stream=0):
    """n doubles in [0, 1) from 53-bit SplitMix64 outputs."""
    idx = np.arange(n, dtype=np.uint64) + np.uint64(stream) * np.uint64(1 << 40)
    return (splitmix64(seed, idx) >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


def normal(seed, n, stream=0):
    """Box-Muller standard normals."""
    m = (n + 1) // 2
    u1 = uniform(seed, m, stream=2 * stream + 1)
    u2 = uniform(seed, m, stream=2 * stream + 2)
    r = np.sqrt(-2.0 * np.log(1.0 - u1))
    z = np.concatenate([r * np.cos(2 * math.pi * u2), r * np.sin(2 * math.pi * u2)])
    return z[:n]


def _profile(frac, start=0.75, coast=0.95, depth=4000.0, land=200.0):
    """Basin floor, linear shelf from ``start`` to the coastline, land beyond."""
    z = np.full_like(frac, -depth)
    shelf = (frac > start) & (frac <= coast)
    z[shelf] = -depth + (frac[shelf] - start) / (coast - start) * depth
    up = frac > coast
    z[up] = (frac[up] - coast) / (1.0 - coast) * land
    return z


def bathymetry(nx, ny, lon0, lat0, n_seamounts, seed, coast=("east",), noise=50.0,
               spacing=ARCMIN2, basin_depth=4000.0):
    """Seeded FP32 topobathymetry z (positive up, metres), shape (ny, nx)."""
    lon = lon0 + spacing * np.arange(nx, dtype=np.float64)
    lat = lat0 + spacing * np.arange(ny, dtype=np.float64)
    fx = np.arange(nx, dtype=np.float64) / (nx - 1)
    fy = np.arange(ny, dtype=np.float64) / (ny - 1)
    z = np.full((ny, nx), -basin_depth)
    if "east" in coast:
        z = np.maximum(z, _profile(fx, depth=basin_depth)[None, :])
    if "north" in coast:
        z = np.maximum(z, _profile(fy, depth=basin_depth)[:, None])
    u = uniform(seed, 5 * n_seamounts, stream=7).reshape(n_seamounts, 5)
    for k in range(n_seamounts):
        cx = lon[0] + (0.05 + 0.6 * u[k, 0]) * (lon[-1] - lon[0])
        cy = lat[0] + (0.05 + 0.6 * u[k, 1]) * (lat[-1] - lat[0])
        height = 1000.0 + 2000.0 * u[k, 2]
        sigma = 5.0 + 10.0 * u[k, 3]
        kx = KM_PER_DEG * math.cos(math.radians(cy))
        ix0 = max(0, int((cx - lon[0] - 5 * sigma / kx) / spacing))
        ix1 = min(nx, int((cx - lon[0] + 5 * sigma / kx) / spacing) + 2)
        iy0 = max(0, int((cy - lat[0] - 5 * sigma / KM_PER_DEG) / spacing))
        iy1 = min(ny, int((cy - lat[0] + 5 * sigma / KM_PER_DEG) / spacing) + 2)
        dx = (lon[ix0:ix1] - cx) * kx
        dy = (lat[iy0:iy1] - cy) * KM_PER_DEG
        g = height * np.exp(-(dy[:, None] ** 2 + dx[None, :] ** 2) / (2.0 * sigma * sigma))
        z[iy0:iy1, ix0:ix1] += g
    if noise > 0:
        z += noise * normal(seed, nx * ny, stream=3).reshape(ny, nx)
    return lon, lat, z.astype(np.float32)


def write_grd(path, lon, lat, z):
    """Classic NetCDF grid with lon/lat double and z float (README.md:53-57)."""
    from scipy.io import netcdf_file
    ver = 2 if z.nbytes > (1 << 30) else 1
    with netcdf_file(path, "w", version=ver) as nc:
        nc.createDimension("lon", len(lon))
        nc.createDimension("lat", len(lat))
        vlon = nc.createVariable("lon", "d", ("lon",))
        vlat = nc.createVariable("lat", "d", ("lat",))
        vz = nc.createVariable("z", "f", ("lat", "lon"))
        vlon[:] = lon
        vlat[:] = lat
        vz[:] = z


def write_points(path, lons, lats):
    with open(path, "w") as f:
        f.write(f"{len(lons)}\n")
        for a, b in zip(lons, lats):
            f.write(f"{a:.6f} {b:.6f}\n")


def write_param_file(path, name, grd, faults, prefix, t_sim, t_save=-1.0, points=None,
                     t_series=30.0, borders=(1, 1, 1, 1), cfl=0.5, eps_h=1e-5, hpos=20.0,
                     stability=0.2, manning=0.03, vmax=100.0, L=1e5, H=1e3, dif_at=1e-6,
                     crop_flag=0, crop_value=0.0, seed=None):
    """Parameter file in the reference's positional format (Mediterranean/mediterranean.txt)."""
    L_ = []
    if seed is not None:
        L_.append(f"# generated by thysea.scenario, seed {seed}")
    L_ += [f"{name}  # Problem name", f"{os.path.basename(grd)}  # Bathymetry file",
           "1  # Initialization type (0: From file, 1: Standard Okada)",
           f"{len(faults)}  # Number of faults"]
    for f in faults:
        L_.append("{time} {lon_c:.6f} {lat_c:.6f} {depth_c:.4f} {length:.4f} {width:.4f} "
                  "{strike:.4f} {dip:.4f} {rake:.4f} {slip:.4f}".format(**f))
    L_.append(f"{crop_flag}  # Crop deformations")
    if crop_flag in (1, 2):
        L_.append(f"{crop_value}")
    L_ += [f"{prefix}  # NetCDF file prefix", "1  # Number of levels"]
    L_ += [f"{b}  # border" for b in borders]
    L_ += [f"{t_sim}  # Simulation time (sec)", f"{t_save}  # Saving time of NetCDF files"]
    if points:
        L_ += ["1  # Read points from file", f"{os.path.basename(points)}", f"{t_series}"]
    else:
        L_ += ["0  # Read points from file"]
    L_ += [f"{cfl}  # CFL", f"{eps_h}  # Epsilon h (m)", f"{hpos}  # Threshold for the 2s+WAF scheme (m)",
           f"{stability}  # Stability coefficient", f"{manning}  # Manning", f"{vmax}  # vmax",
           f"{L}  # L", f"{H}  # H", f"{dif_at}  # Threshold for arrival times (m)"]
    with open(path, "w") as fh:
        fh.write("\n".join(L_) + "\n")


def centroid_depth(top_km, width_km, dip_deg):
    """Centroid depth of a rectangle whose up-dip edge is ``top_km`` deep."""
    return top_km + 0.5 * width_km * math.sin(math.radians(dip_deg))


def gauges_lattice(lon, lat, n_side, frac=(0.1, 0.9)):
    a = np.linspace(lon[0] + frac[0] * (lon[-1] - lon[0]), lon[0] + frac[1] * (lon[-1] - lon[0]), n_side)
    b = np.linspace(lat[0] + frac[0] * (lat[-1] - lat[0]), lat[0] + frac[1] * (lat[-1] - lat[0]), n_side)
    gl, gt = np.meshgrid(a, b)
    return gl.ravel(), gt.ravel()


def magnitude_fault(mw, lon_c, lat_c, strike, dip, rake=90.0, mu=3e10, time=0.0):
    """Fault size from magnitude: Strasser et al. (2010) interface scaling
    (log10 L = -2.477 + 0.585 Mw, log10 W = -0.882 + 0.351 Mw, km), slip closed by
    M0 = 10^(1.5 Mw + 9.1) = mu L W D; centroid deep enough for the top edge to sit
    at 5 km."""
    L = 10 ** (-2.477 + 0.585 * mw)
    W = 10 ** (-0.882 + 0.351 * mw)
    m0 = 10 ** (1.5 * mw + 9.1)
    slip = m0 / (mu * L * 1e3 * W * 1e3)
    return dict(time=time, lon_c=lon_c, lat_c=lat_c, depth_c=centroid_depth(5.0, W, dip),
                length=L, width=W, strike=strike, dip=dip, rake=rake, slip=slip)


def random_faults(lon, lat, n, seed, mw_range=(7.0, 8.8), stream=21):
    """n magnitude-scaled thrust faults: Mw uniform over ``mw_range`` (log-uniform
    seismic moment), epicentre in the basin part of the grid, strike uniform, dip in
    [10, 30] degrees."""
    u = uniform(seed, 5 * n, stream=stream).reshape(n, 5)
    out = []
    for k in range(n):
        mw = mw_range[0] + (mw_range[1] - mw_range[0]) * u[k, 0]
        lon_c = float(lon[0] + (0.15 + 0.45 * u[k, 1]) * (lon[-1] - lon[0]))
        lat_c = float(lat[0] + (0.25 + 0.5 * u[k, 2]) * (lat[-1] - lat[0]))
        out.append(magnitude_fault(mw, lon_c, lat_c, 360.0 * u[k, 3], 10.0 + 20.0 * u[k, 4]))
    return out


def _in_memory(name, lon, lat, z, faults, t_sim, n_side, series_s):
    pr = P.build(name, name, lon, lat, z, faults=faults, t_tot=t_sim, t_series=series_s, eps_h=1e-5,
                 dif_at=1e-6)
    glon, glat = gauges_lattice(lon, lat, n_side)
    pr.gauge_pos = np.array([P.nearest_cell(lon, lat, a, b) for a, b in zip(glon, glat)], np.int32)
    return pr


def batch_problems(n, seed=20240601, nx=1000, ny=800, t_sim=3600.0, series_s=60.0):
    """The batched-scenario workload: n scenarios on one shared 1000x800 grid (east
    shelf and coast, 10 seamounts), each with its own magnitude-scaled fault
    (Mw 7.0-8.8) and 16 gauges, 3600 s each."""
    import dataclasses
    lon, lat, z = bathymetry(nx, ny, 140.0, 10.0, 10, seed, coast=("east",))
    faults = random_faults(lon, lat, n, seed)
    base = _in_memory("batch", lon, lat, z, faults[:1], t_sim, 4, series_s)
    # the scenarios differ only in their fault: share the grid arrays
    return [dataclasses.replace(base, faults=[dict(f)]) for f in faults]


def serving_requests(rate, horizon_s, seed=20240601, nx_range=(300, 2400), ny_range=(300, 1800),
                     t_range=(1800.0, 10800.0), mw_range=(7.0, 9.0)):
    """Poisson arrivals (``rate`` per second over ``horizon_s``) of scenario requests:
    grid size uniform in the ranges, simulated duration uniform in ``t_range``, one
    Mw source.  Returns [(arrival_s, spec)], spec = dict(nx, ny, t_sim, seed)."""
    n_max = int(rate * horizon_s * 3 + 20)
    u = uniform(seed, 5 * n_max, stream=31).reshape(n_max, 5)
    out, t = [], 0.0
    for k in range(n_max):
        t += -math.log(1.0 - u[k, 0]) / rate
        if t > horizon_s:
            break
        out.append((t, dict(nx=int(nx_range[0] + (nx_range[1] - nx_range[0]) * u[k, 1]),
                            ny=int(ny_range[0] + (ny_range[1] - ny_range[0]) * u[k, 2]),
                            t_sim=t_range[0] + (t_range[1] - t_range[0]) * u[k, 3],
                            seed=int(seed + 1000 + k), mw_range=mw_range)))
    return out


def request_problem(spec, series_s=60.0):
    """The problem of one serving request (its own grid, one random fault)."""
    lon, lat, z = bathymetry(spec["nx"], spec["ny"], 140.0, 10.0, 10, spec["seed"], coast=("east",))
    f = random_faults(lon, lat, 1, spec["seed"], mw_range=spec["mw_range"])
    return _in_memory("request", lon, lat, z, f, spec["t_sim"], 4, series_s)


def make(workload, outdir, seed=20240601, nx=None, ny=None, t_sim=None, gauges=True, borders=None):
    """Write the inputs of a named workload to ``outdir``; return the parameter-file path."""
    os.makedirs(outdir, exist_ok=True)
    if workload == "small":
        nx, ny = nx or 240, ny or 200
        lon, lat, z = bathymetry(nx, ny, 140.0, 30.0, 5, seed, coast=("east",))
        dep = centroid_depth(10.0, 50.0, 30.0)
        faults = [dict(time=0.0, lon_c=float(lon[nx // 2]), lat_c=float(lat[ny // 2]), depth_c=dep,
                       length=100.0, width=50.0, strike=90.0, dip=30.0, rake=90.0, slip=3.0)]
        t_sim = 3600.0 if t_sim is None else t_sim
        ng = 4
    elif workload == "large":
        nx, ny = nx or 4000, ny or 3000
        lon, lat, z = bathymetry(nx, ny, 100.0, -50.0, 20, seed, coast=("east", "north"))
        cx, cy = float(lon[int(nx * 0.45)]), float(lat[int(ny * 0.45)])
        faults = []
        for k in range(4):
            faults.append(dict(time=0.0, lon_c=cx, lat_c=cy + (k - 1.5) * 100.0 / KM_PER_DEG,
                               depth_c=25.0, length=100.0, width=50.0, strike=0.0, dip=15.0,
                               rake=90.0, slip=2.0))
        t_sim = 1800.0 if t_sim is None else t_sim
        ng = 8
    elif workload == "precision":
        nx, ny = nx or 2000, ny or 2000
        lon, lat, z = bathymetry(nx, ny, 130.0, -33.3, 20, seed, coast=("east", "north"))
        u = uniform(seed, 3, stream=11)
        strike, dip = 360.0 * u[0], 10.0 + 20.0 * u[1]
        faults = [dict(time=0.0, lon_c=float(lon[int(nx * 0.4)]), lat_c=float(lat[int(ny * 0.4)]),
                       depth_c=50.0 * math.sin(math.radians(dip)) + 5.0, length=300.0, width=100.0,
                       strike=strike, dip=dip, rake=90.0, slip=8.0)]
        t_sim = 7200.0 if t_sim is None else t_sim
        ng = 8
    elif workload == "batch":
        nx, ny = nx or 1000, ny or 800
        lon, lat, z = bathymetry(nx, ny, 140.0, 10.0, 10, seed, coast=("east",))
        faults = [magnitude_fault(8.0, float(lon[nx // 3]), float(lat[ny // 2]), 20.0, 15.0)]
        t_sim = 3600.0 if t_sim is None else t_sim
        ng = 4
    else:
        raise ValueError(workload)
    grd = os.path.join(outdir, f"{workload}.grd")
    write_grd(grd, lon, lat, z)
    pts = None
    if gauges:
        glon, glat = gauges_lattice(lon, lat, ng)
        pts = os.path.join(outdir, f"{workload}_gauges.txt")
        write_points(pts, glon, glat)
    par = os.path.join(outdir, f"{workload}.txt")
    write_param_file(par, workload, grd, faults, workload, t_sim, points=pts,
                     borders=borders or (1, 1, 1, 1), seed=seed)
    return par


def problem(workload, seed=20240601, nx=None, ny=None, t_sim=None, borders=(1, 1, 1, 1),
            gauges=True, eps_h=1e-5):
    """Build a workload in memory (same numbers as writing and re-reading its files,
    except that no NetCDF round trip happens)."""
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        par = make(workload, d, seed=seed, nx=nx, ny=ny, t_sim=t_sim, gauges=gauges, borders=borders)
        return P.load(par)
