"""ctypes wrapper of the FP64 CPU oracle (oracle/thy_oracle.c).

TEST INFRASTRUCTURE ONLY.  This module is the parity checker and the CPU baseline.
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import it; the
product path (libthysea.so, the TsunamiHySEA executable and the ``thysea`` package)
never does.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)


class Params(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int),
        ("borde_sup", C.c_double), ("borde_inf", C.c_double),
        ("borde_izq", C.c_double), ("borde_der", C.c_double),
        ("sponge_sup", C.c_int), ("sponge_inf", C.c_int),
        ("sponge_izq", C.c_int), ("sponge_der", C.c_int),
        ("cfl", C.c_double), ("mf0", C.c_double), ("vmax", C.c_double),
        ("eps_h", C.c_double), ("hpos", C.c_double), ("cvis", C.c_double),
        ("dif_at", C.c_double), ("hmin", C.c_double), ("sea_level", C.c_double),
        ("quirk_reduction", C.c_int),
    ]


class Fault(C.Structure):
    _fields_ = [(k, C.c_double) for k in
                ("time", "lon_c", "lat_c", "depth_c", "length", "width", "strike", "dip", "rake", "slip")]


SNAPCB = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.c_double)


class RunCfg(C.Structure):
    _fields_ = [
        ("t_end", C.c_double), ("save_interval", C.c_double), ("series_interval", C.c_double),
        ("T", C.c_double), ("H", C.c_double), ("Q", C.c_double),
        ("n_faults", C.c_int), ("faults", C.POINTER(Fault)),
        ("crop_flag", C.c_int), ("crop_value", C.c_double),
        ("lon0", C.c_double), ("dlon", C.c_double), ("lat0", C.c_double), ("dlat", C.c_double),
        ("n_gauges", C.c_int), ("gauge_pos", _ip),
        ("max_samples", C.c_int),
        ("g_time", _fp), ("g_eta", _fp), ("g_u", _fp), ("g_v", _fp),
        ("g_eta_min", _fp), ("g_eta_max", _fp),
        ("max_steps", C.c_int64),
        ("on_snapshot", SNAPCB), ("user", C.c_void_p),
    ]


class RunOut(C.Structure):
    _fields_ = [("steps", C.c_int64), ("n_samples", C.c_int), ("n_snapshots", C.c_int),
                ("t_final", C.c_double), ("dt_initial", C.c_double), ("dt_last", C.c_double)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libthy_oracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle library not built: run `make -C oracle`")
        L = C.CDLL(path)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.POINTER(Params), _dp, _dp, _dp, _dp]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_set_state.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
        L.orc_get_state.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
        L.orc_init_metrics.argtypes = [C.c_void_p]
        L.orc_update_metrics.argtypes = [C.c_void_p, C.c_double]
        L.orc_get_metrics.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.orc_initial_dt.argtypes = [C.c_void_p]
        L.orc_initial_dt.restype = C.c_double
        L.orc_get_cell_dt.argtypes = [C.c_void_p, _dp]
        L.orc_step.argtypes = [C.c_void_p, C.c_double]
        L.orc_step.restype = C.c_double
        L.orc_mass_pass.argtypes = [C.c_void_p, C.c_double]
        L.orc_momentum_pass.argtypes = [C.c_void_p, C.c_double]
        L.orc_momentum_pass.restype = C.c_double
        L.orc_reduce_min.argtypes = [_dp, C.c_int64, C.c_int]
        L.orc_reduce_min.restype = C.c_double
        L.orc_okada.argtypes = [C.c_void_p] + [C.c_double] * 13 + [C.c_int, C.c_double, C.c_double, _dp]
        L.orc_prtoxy.argtypes = [_dp, _dp, C.c_double, C.c_double, _dp, _dp, C.c_int]
        L.orc_dc3d_surface.argtypes = [C.c_double] * 12 + [_dp, _dp, _dp, _ip]
        L.orc_limiter.argtypes = [C.c_double] * 5
        L.orc_limiter.restype = C.c_double
        L.orc_desing.argtypes = [C.c_double] * 3
        L.orc_desing.restype = C.c_double
        L.orc_friction_filter.argtypes = [C.c_double, C.c_double, C.c_double, _dp, _dp,
                                          C.c_double, C.c_double, C.c_double, C.c_double]
        L.orc_output_field.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_double, _fp]
        L.orc_gauges.argtypes = [C.c_void_p, _ip, C.c_int, C.c_double, C.c_double, _fp, _fp, _fp]
        L.orc_run.argtypes = [C.c_void_p, C.POINTER(RunCfg), C.POINTER(RunOut)]
        L.orc_mass.argtypes = [C.c_void_p]
        L.orc_mass.restype = C.c_double
        _LIB = L
    return _LIB


def _d(a):
    return a.ctypes.data_as(_dp)


FIELDS = {"eta": 0, "ux": 1, "uy": 2, "deformed_bathy": 3, "max_height": 4, "arrival_times": 5}


class OracleSim:
    """FP64 restatement of one level-0 grid; mirrors the reference's device state."""

    def __init__(self, prob, quirk_reduction=0, threads=0):
        L = lib()
        if threads:
            L.orc_set_threads(int(threads))
        self.prob = prob
        p = Params()
        p.nx, p.ny = prob.nx, prob.ny
        p.borde_sup, p.borde_inf = prob.borde_sup, prob.borde_inf
        p.borde_izq, p.borde_der = prob.borde_izq, prob.borde_der
        p.sponge_sup, p.sponge_inf = prob.sponge_sup, prob.sponge_inf
        p.sponge_izq, p.sponge_der = prob.sponge_izq, prob.sponge_der
        p.cfl, p.mf0, p.vmax = prob.cfl, prob.mf0, prob.vmax
        p.eps_h, p.hpos, p.cvis, p.dif_at = prob.eps_h, prob.hpos, prob.cvis, prob.dif_at
        p.hmin, p.sea_level = prob.hmin, 0.0
        p.quirk_reduction = int(quirk_reduction)
        self._geo = [np.ascontiguousarray(a, dtype=np.float64)
                     for a in (prob.area, prob.cosphi, prob.ancho, prob.alto)]
        self.h = L.orc_create(C.byref(p), *[_d(a) for a in self._geo])
        self.n = prob.nx * prob.ny
        self.set_state(prob.h, prob.H, prob.qx, prob.qy)

    def close(self):
        if self.h:
            lib().orc_destroy(self.h)
            self.h = None

    __del__ = close

    def set_state(self, h, H, qx, qy):
        arrs = [np.ascontiguousarray(a, dtype=np.float64).ravel() for a in (h, H, qx, qy)]
        lib().orc_set_state(self.h, *[_d(a) for a in arrs])

    def state(self):
        out = [np.empty(self.n) for _ in range(4)]
        lib().orc_get_state(self.h, *[_d(a) for a in out])
        ny, nx = self.prob.ny, self.prob.nx
        return tuple(a.reshape(ny, nx) for a in out)

    def metrics(self):
        out = [np.empty(self.n) for _ in range(3)]
        lib().orc_get_metrics(self.h, *[_d(a) for a in out])
        ny, nx = self.prob.ny, self.prob.nx
        return tuple(a.reshape(ny, nx) for a in out)

    def init_metrics(self):
        lib().orc_init_metrics(self.h)

    def update_metrics(self, t):
        lib().orc_update_metrics(self.h, float(t))

    def initial_dt(self):
        return lib().orc_initial_dt(self.h)

    def cell_dt(self):
        out = np.empty(self.n)
        lib().orc_get_cell_dt(self.h, _d(out))
        return out.reshape(self.prob.ny, self.prob.nx)

    def step(self, dt):
        return lib().orc_step(self.h, float(dt))

    def mass_pass(self, dt):
        lib().orc_mass_pass(self.h, float(dt))

    def momentum_pass(self, dt):
        return lib().orc_momentum_pass(self.h, float(dt))

    def okada(self, f, crop_flag=0, crop_value=0.0, want_def=False):
        pr = self.prob
        d = np.empty(self.n) if want_def else None
        lib().orc_okada(self.h, pr.lon0, pr.dlon, pr.lat0, pr.dlat,
                        f["lon_c"], f["lat_c"], f["depth_c"], f["length"], f["width"],
                        f["strike"], f["dip"], f["rake"], f["slip"],
                        int(crop_flag), float(crop_value), pr.H_ref, _d(d) if want_def else None)
        return None if d is None else d.reshape(pr.ny, pr.nx)

    def output(self, which):
        out = np.empty(self.n, dtype=np.float32)
        pr = self.prob
        lib().orc_output_field(self.h, FIELDS[which], pr.H_ref, pr.Q, pr.T, out.ctypes.data_as(_fp))
        return out.reshape(pr.ny, pr.nx)

    def mass(self):
        return lib().orc_mass(self.h)

    def gauges(self, pos):
        pos = np.ascontiguousarray(pos, dtype=np.int32)
        n = len(pos)
        e, u, v = (np.empty(n, np.float32) for _ in range(3))
        pr = self.prob
        lib().orc_gauges(self.h, pos.ctypes.data_as(_ip), n, pr.H_ref, pr.Q,
                         e.ctypes.data_as(_fp), u.ctypes.data_as(_fp), v.ctypes.data_as(_fp))
        return e, u, v

    def run(self, max_steps=0, max_samples=4096, on_snapshot=None):
        """Full driver loop (src/GPU/ShallowWater.cu:374-453) on this problem."""
        pr = self.prob
        faults = sorted(pr.faults, key=lambda f: f["time"])
        farr = (Fault * max(1, len(faults)))()
        for k, f in enumerate(faults):
            for name, _ in Fault._fields_:
                setattr(farr[k], name, float(f[name]))
        gpos = np.ascontiguousarray(pr.gauge_pos if pr.gauge_pos is not None else np.zeros(0), np.int32)
        ng = len(gpos)
        bufs = {k: np.zeros(max(1, max_samples * max(ng, 1)), np.float32) for k in ("eta", "u", "v")}
        gt = np.zeros(max(1, max_samples), np.float32)
        gmin = np.zeros(max(1, ng), np.float32)
        gmax = np.zeros(max(1, ng), np.float32)
        c = RunCfg()
        c.t_end, c.save_interval, c.series_interval = pr.t_end, pr.save_interval, pr.series_interval
        c.T, c.H, c.Q = pr.T, pr.H_ref, pr.Q
        c.n_faults, c.faults = len(faults), farr
        c.crop_flag, c.crop_value = pr.crop_flag, pr.crop_value
        c.lon0, c.dlon, c.lat0, c.dlat = pr.lon0, pr.dlon, pr.lat0, pr.dlat
        c.n_gauges, c.gauge_pos = ng, gpos.ctypes.data_as(_ip)
        c.max_samples = max_samples
        c.g_time = gt.ctypes.data_as(_fp)
        c.g_eta = bufs["eta"].ctypes.data_as(_fp)
        c.g_u = bufs["u"].ctypes.data_as(_fp)
        c.g_v = bufs["v"].ctypes.data_as(_fp)
        c.g_eta_min = gmin.ctypes.data_as(_fp)
        c.g_eta_max = gmax.ctypes.data_as(_fp)
        c.max_steps = int(max_steps)
        cb = SNAPCB(lambda u, i, t: on_snapshot(i, t)) if on_snapshot else SNAPCB()
        c.on_snapshot = cb
        o = RunOut()
        lib().orc_run(self.h, C.byref(c), C.byref(o))
        ns = min(o.n_samples, max_samples)
        res = {
            "steps": o.steps, "t_final": o.t_final, "dt_initial": o.dt_initial, "dt_last": o.dt_last,
            "n_samples": o.n_samples, "n_snapshots": o.n_snapshots,
            "g_time": gt[:ns].copy(),
            "g_eta": bufs["eta"][:ns * ng].reshape(ns, ng).copy(),
            "g_u": bufs["u"][:ns * ng].reshape(ns, ng).copy(),
            "g_v": bufs["v"][:ns * ng].reshape(ns, ng).copy(),
            "g_eta_min": gmin[:ng].copy(), "g_eta_max": gmax[:ng].copy(),
        }
        return res


def reduce_min(v, quirk=0):
    v = np.ascontiguousarray(v, dtype=np.float64).ravel()
    return lib().orc_reduce_min(_d(v), len(v), int(quirk))


def dc3d_surface(alpha, x, y, depth, dip, al1, al2, aw1, aw2, d1, d2, d3):
    ux, uy, uz = C.c_double(), C.c_double(), C.c_double()
    ir = C.c_int()
    lib().orc_dc3d_surface(alpha, x, y, depth, dip, al1, al2, aw1, aw2, d1, d2, d3,
                           C.byref(ux), C.byref(uy), C.byref(uz), C.byref(ir))
    return ux.value, uy.value, uz.value, ir.value


def prtoxy(lat, lon, lat0, lon0, ind=0, x=0.0, y=0.0):
    a, b = C.c_double(lat), C.c_double(lon)
    xx, yy = C.c_double(x), C.c_double(y)
    lib().orc_prtoxy(C.byref(a), C.byref(b), lat0, lon0, C.byref(xx), C.byref(yy), ind)
    return (xx.value, yy.value) if ind == 0 else (a.value, b.value)


def limiter(p0m, p0, p1, p1p, S):
    return lib().orc_limiter(p0m, p0, p1, p1p, S)


def desing(h, q, eps_h):
    return lib().orc_desing(h, q, eps_h)


def friction_filter(hs, qxo, qyo, qx, qy, dt, mf0, vmax, eps_h):
    a, b = C.c_double(qx), C.c_double(qy)
    lib().orc_friction_filter(hs, qxo, qyo, C.byref(a), C.byref(b), dt, mf0, vmax, eps_h)
    return a.value, b.value
