"""Problem setup: parameter file, NetCDF grid, non-dimensionalisation, geometry, gauges.

Python mirror of the reference's host setup ``cargarDatosProblema``
(src/Problema.cxx:176-477).  The same numbers feed the HIP library (through the
C-ABI in include/thysea.h) and the CPU oracle, so both consume identical inputs.
The C++ executable has its own implementation of the same rules
(tsunami-hysea_amd/host/problem.cpp); tests cross-check the two.
"""
import math
import os
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

EARTH_RADIUS = 6378136.6      # src/Constantes.hxx:12
MAX_FAULTS = 5000             # src/Constantes.hxx:13
SPONGE_SIZE = 4               # src/Constantes.hxx:14
DBL_EPS = np.finfo(np.float64).eps
SEA_SURFACE_FROM_FILE, OKADA_STANDARD = 0, 1
NO_CROP, CROP_RELATIVE, CROP_ABSOLUTE = 0, 1, 2


class ProblemError(RuntimeError):
    """A parameter-file error; the message matches the reference's stderr text."""


@dataclass
class Problem:
    name: str
    prefix: str
    nx: int
    ny: int
    lon: np.ndarray
    lat: np.ndarray
    h: np.ndarray          # (ny, nx) float64 non-dimensional total depth
    H: np.ndarray          # (ny, nx) float64 shifted still-water depth (>= 0)
    qx: np.ndarray
    qy: np.ndarray
    hmin: float
    area: np.ndarray
    cosphi: np.ndarray
    ancho: np.ndarray
    alto: np.ndarray
    borde_sup: float = 1.0
    borde_inf: float = 1.0
    borde_izq: float = 1.0
    borde_der: float = 1.0
    cfl: float = 0.5
    eps_h: float = 0.0
    hpos: float = 0.0
    cvis: float = 0.8
    mf0: float = 0.0
    vmax: float = 0.0
    dif_at: float = 0.0
    L: float = 1e5
    H_ref: float = 1e3
    Q: float = 0.0
    T: float = 0.0
    t_end: float = 0.0
    save_interval: float = -1.0
    series_interval: float = -1.0
    okada_flag: int = OKADA_STANDARD
    faults: List[dict] = field(default_factory=list)
    crop_flag: int = NO_CROP
    crop_value: float = 0.0
    gauge_lon: Optional[np.ndarray] = None
    gauge_lat: Optional[np.ndarray] = None
    gauge_pos: Optional[np.ndarray] = None
    bathy_file: str = ""

    @property
    def periodic(self):
        return abs(self.borde_izq - 2.0) < DBL_EPS and abs(self.borde_der - 2.0) < DBL_EPS

    def sponge(self, b):
        return SPONGE_SIZE if abs(b - 1.0) < DBL_EPS else 0

    @property
    def sponge_sup(self):
        return self.sponge(self.borde_sup)

    @property
    def sponge_inf(self):
        return self.sponge(self.borde_inf)

    @property
    def sponge_izq(self):
        return self.sponge(self.borde_izq)

    @property
    def sponge_der(self):
        return self.sponge(self.borde_der)

    # uniform coordinates used by Okada (src/GPU/ShallowWater.cu:215-218)
    @property
    def lon0(self):
        return float(self.lon[0])

    @property
    def lat0(self):
        return float(self.lat[0])

    @property
    def dlon(self):
        return float((self.lon[self.nx - 1] - self.lon[0]) / (self.nx - 1))

    @property
    def dlat(self):
        return float((self.lat[self.ny - 1] - self.lat[0]) / (self.ny - 1))

    @property
    def n_cells(self):
        return self.nx * self.ny


def geometry(lon, lat, L):
    """Per-row area, cos(lat), cell width/height (src/Problema.cxx:431-448)."""
    ny = len(lat)
    R = EARTH_RADIUS / L
    g2r = math.pi / 180.0
    dlon = abs(float(lon[1]) - float(lon[0]))
    ancho = np.empty(ny + 1)
    alto = np.empty(ny)
    area = np.empty(ny)
    cosphi = np.empty(ny)
    dlat = 0.0
    for i in range(ny - 1):
        dlat = abs(float(lat[i + 1]) - float(lat[i]))
        ancho[i] = dlon
        alto[i] = dlat
        area[i] = R * dlon * dlat * g2r
        cosphi[i] = math.cos(float(lat[i]) * g2r)
    ancho[ny - 1] = dlon
    alto[ny - 1] = alto[ny - 2]
    ancho[ny] = dlon
    area[ny - 1] = R * dlon * dlat * g2r
    cosphi[ny - 1] = math.cos(float(lat[ny - 1]) * g2r)
    return area, cosphi, ancho, alto


def nearest_cell(lon, lat, plon, plat):
    """src/Problema.cxx:54-95 (first grid value >= point, then the closer neighbour)."""
    nx, ny = len(lon), len(lat)
    if plon < lon[0]:
        return -1
    i = 0
    while i < nx and not lon[i] >= plon:
        i += 1
    if i >= nx:
        return -1
    if abs(plon - lon[max(i - 1, 0)]) < abs(plon - lon[i]):
        i = max(i - 1, 0)
    if plat < lat[0]:
        return -1
    j = 0
    while j < ny and not lat[j] >= plat:
        j += 1
    if j >= ny:
        return -1
    if abs(plat - lat[max(j - 1, 0)]) < abs(plat - lat[j]):
        j = max(j - 1, 0)
    return j * nx + i


def read_points(path, lon, lat):
    """Gauge file (count, then lon lat pairs); duplicates within 1e-6 dropped
    (src/Problema.cxx:39-52,400-429)."""
    with open(path) as f:
        toks = f.read().split()
    n = int(toks[0])
    glon, glat, gpos = [], [], []
    for k in range(n):
        plon, plat = float(toks[1 + 2 * k]), float(toks[2 + 2 * k])
        dup = any(abs(a - plon) < 1e-6 and abs(b - plat) < 1e-6 for a, b in zip(glon, glat))
        if not dup:
            glon.append(plon)
            glat.append(plat)
            gpos.append(nearest_cell(lon, lat, plon, plat))
    return np.array(glon), np.array(glat), np.array(gpos, dtype=np.int32)


def read_grd(path):
    """lon/lat (or x/y) double and z float from a classic NetCDF grid
    (src/GPU/netcdfTH.cu:13-91)."""
    from scipy.io import netcdf_file
    with netcdf_file(path, "r", mmap=False) as nc:
        lon = np.array(nc.variables["lon" if "lon" in nc.variables else "x"][:], dtype=np.float64)
        lat = np.array(nc.variables["lat" if "lat" in nc.variables else "y"][:], dtype=np.float64)
        z = np.array(nc.variables["z"][:], dtype=np.float32)
    return lon, lat, z.reshape(len(lat), len(lon))


class _Lines:
    """One datum per non-blank line not starting with '#'
    (src/Problema.cxx:150-170); a datum is the first whitespace token."""

    def __init__(self, path):
        with open(path) as f:
            self.lines = f.read().splitlines()
        self.k = 0

    def next_line(self):
        while self.k < len(self.lines):
            s = self.lines[self.k].lstrip(" \t\r\n")
            self.k += 1
            if s and not s.startswith("#"):
                return s
        return None

    def token(self, conv, default=None):
        s = self.next_line()
        if s is None or not s.split():
            return default
        return conv(s.split()[0])


def nondim_friction(n, L, H):
    """src/Problema.cxx:335: mf0 *= 9.81*mf0*L/H^(4/3)"""
    return n * (9.81 * n * L / math.pow(H, 4.0 / 3.0))


def load(param_path):
    """Read a reference parameter file and its inputs (src/Problema.cxx:176-477)."""
    d = os.path.dirname(param_path)
    d = d + "/" if d else ""
    f = _Lines(param_path)
    name = f.token(str)
    topo = f.token(str)
    if not os.path.exists(d + topo):
        raise ProblemError(f"Error: File '{d + topo}' not found")
    okada_flag = f.token(int)
    if okada_flag not in (SEA_SURFACE_FROM_FILE, OKADA_STANDARD):
        raise ProblemError("Error: The initialization flag should be 0 or 1")
    faults, crop_flag, crop_value, est = [], NO_CROP, 0.0, None
    if okada_flag == SEA_SURFACE_FROM_FILE:
        est = f.token(str)
        if not os.path.exists(d + est):
            raise ProblemError(f"Error: File '{d + est}' not found")
    else:
        nf = f.token(int)
        if nf < 1 or nf > MAX_FAULTS:
            raise ProblemError(f"Error: The number of faults should be between 1 and {MAX_FAULTS}")
        keys = ("time", "lon_c", "lat_c", "depth_c", "length", "width", "strike", "dip", "rake", "slip")
        for _ in range(nf):
            vals = [float(x) for x in f.next_line().split()[:10]]
            faults.append(dict(zip(keys, vals)))
        # selection sort by time, as src/Problema.cxx:97-126 (stable for ties)
        for i in range(len(faults) - 1):
            m = i
            for j in range(i + 1, len(faults)):
                if faults[j]["time"] < faults[m]["time"]:
                    m = j
            if m != i:
                faults[i], faults[m] = faults[m], faults[i]
        crop_flag = f.token(int)
        if crop_flag not in (NO_CROP, CROP_RELATIVE, CROP_ABSOLUTE):
            raise ProblemError("Error: The cropping flag should be 0, 1 or 2")
        if crop_flag in (CROP_RELATIVE, CROP_ABSOLUTE):
            crop_value = f.token(float)
    prefix = f.token(str)
    lon, lat, z = read_grd(d + topo)
    nx, ny = len(lon), len(lat)
    if nx < 2 or ny < 2:
        raise ProblemError("Error: Mesh size too small. The number of rows and columns should be >= 2")
    levels = f.token(int)
    if levels != 1:
        raise ProblemError("Error: The number of levels should be 1")
    b = [f.token(float) for _ in range(4)]
    for v in b:
        if abs(v - 1.0) > DBL_EPS and abs(v - 2.0) > DBL_EPS and abs(v + 1.0) > DBL_EPS:
            raise ProblemError("Error: The boundary conditions should be -1, 1 or 2")
    if abs(b[0] - 2.0) < DBL_EPS or abs(b[1] - 2.0) < DBL_EPS:
        raise ProblemError("Error: Periodic boundary conditions are only supported for left and right borders")
    if (abs(b[2] - 2.0) < DBL_EPS) != (abs(b[3] - 2.0) < DBL_EPS):
        raise ProblemError("Error: Both left and right borders should have periodic boundary conditions")
    t_tot = f.token(float)
    t_save = f.token(float)
    read_pts = f.token(int)
    t_series = -1.0
    pts = None
    if read_pts == 1:
        pts = f.token(str)
        t_series = f.token(float)
        if not os.path.exists(d + pts):
            raise ProblemError(f"Error: File '{d + pts}' not found")
    cfl = f.token(float)
    eps_h = f.token(float)
    hpos = f.token(float)
    stab = f.token(float)
    manning = f.token(float)
    vmax = f.token(float)
    L = f.token(float)
    Href = f.token(float)
    dif_at = f.token(float, default=eps_h)
    prob = build(name, prefix, lon, lat, z, okada_flag=okada_flag, faults=faults,
                 crop_flag=crop_flag, crop_value=crop_value, borders=b, t_tot=t_tot,
                 t_save=t_save, t_series=t_series, cfl=cfl, eps_h=eps_h, hpos=hpos,
                 stability=stab, manning=manning, vmax=vmax, L=L, H=Href, dif_at=dif_at,
                 eta_file=(d + est) if est else None)
    prob.bathy_file = d + topo
    if pts is not None:
        prob.gauge_lon, prob.gauge_lat, prob.gauge_pos = read_points(d + pts, lon, lat)
    return prob


def build(name, prefix, lon, lat, z, okada_flag=OKADA_STANDARD, faults=(), crop_flag=NO_CROP,
          crop_value=0.0, borders=(1, 1, 1, 1), t_tot=3600.0, t_save=-1.0, t_series=-1.0,
          cfl=0.5, eps_h=5e-3, hpos=20.0, stability=0.2, manning=0.03, vmax=100.0, L=1e5,
          H=1e3, dif_at=None, eta_file=None, eta=None):
    """Non-dimensionalise a problem given in physical units (src/Problema.cxx:327-473)."""
    lon = np.asarray(lon, np.float64)
    lat = np.asarray(lat, np.float64)
    z = np.asarray(z, np.float32).reshape(len(lat), len(lon))
    if dif_at is None:
        dif_at = eps_h
    Q = math.sqrt(9.81 * math.pow(H, 3.0))
    T = L * H / Q
    # H = -z/H_ref, h = max(H, 0) for Okada; Hmin shift (src/Problema.cxx:383-398,450-456)
    val = (-1.0 * z.astype(np.float64)) / H
    Hs = val.copy()
    hmin = float(val.min())
    h = np.where(val > 0.0, val, 0.0) if okada_flag == OKADA_STANDARD else np.zeros_like(val)
    if hmin < 0.0:
        Hs -= hmin
    else:
        hmin = 0.0
    if okada_flag == SEA_SURFACE_FROM_FILE:
        if eta is None:
            _, _, eta = read_grd(eta_file)
        e = np.asarray(eta, np.float32).astype(np.float64)
        h = np.maximum(e / H + Hs + hmin, 0.0)
    area, cosphi, ancho, alto = geometry(lon, lat, L)
    cv = crop_value / H if crop_flag == CROP_ABSOLUTE else crop_value
    return Problem(
        name=name, prefix=prefix, nx=len(lon), ny=len(lat), lon=lon, lat=lat,
        h=h, H=Hs, qx=np.zeros_like(h), qy=np.zeros_like(h), hmin=hmin,
        area=area, cosphi=cosphi, ancho=ancho, alto=alto,
        borde_sup=float(borders[0]), borde_inf=float(borders[1]),
        borde_izq=float(borders[2]), borde_der=float(borders[3]),
        cfl=cfl, eps_h=eps_h / H, hpos=hpos / H, cvis=1.0 - stability,
        mf0=nondim_friction(manning, L, H), vmax=vmax / (Q / H), dif_at=dif_at / H,
        L=L, H_ref=H, Q=Q, T=T, t_end=t_tot / T, save_interval=t_save / T,
        series_interval=t_series / T, okada_flag=okada_flag,
        faults=[dict(f) for f in faults], crop_flag=crop_flag, crop_value=cv)
