"""ctypes binding of libthysea.so (C-ABI in include/thysea.h).

The HIP library is the compute path.  There is no fallback: if the library is not
built, or no gfx950 device is present, the calls raise.
"""
import ctypes as C
import os

import numpy as np

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG, "libthysea.so")
_LIB = None

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)

FIELDS = {"eta": 0, "ux": 1, "uy": 2, "deformed_bathy": 3, "max_height": 4, "arrival_times": 5}
STATUS = {0: "ok", 1: "device memory", 2: "host memory", 3: "invalid argument", 4: "no device",
          5: "HIP error", 6: "memory budget", 7: "non-finite state"}
STATUS_CODES = {"OK": 0, "DEVICE_MEMORY": 1, "HOST_MEMORY": 2, "INVALID": 3, "NO_DEVICE": 4, "HIP": 5, "BUDGET": 6,
                "NONFINITE": 7}


class Grid(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("lon", _dp), ("lat", _dp), ("area", _dp),
                ("cosphi", _dp), ("ancho", _dp), ("alto", _dp)]


class Params(C.Structure):
    _fields_ = [("borde_sup", C.c_double), ("borde_inf", C.c_double), ("borde_izq", C.c_double),
                ("borde_der", C.c_double), ("sponge_sup", C.c_int), ("sponge_inf", C.c_int),
                ("sponge_izq", C.c_int), ("sponge_der", C.c_int), ("cfl", C.c_double), ("mf0", C.c_double),
                ("vmax", C.c_double), ("eps_h", C.c_double), ("hpos", C.c_double), ("cvis", C.c_double),
                ("dif_at", C.c_double), ("hmin", C.c_double), ("sea_level", C.c_double), ("L", C.c_double),
                ("H", C.c_double), ("Q", C.c_double), ("T", C.c_double), ("precision", C.c_int),
                ("track_arrival", C.c_int), ("tile_x", C.c_int), ("tile_y", C.c_int), ("kernel", C.c_int),
                ("strip_rows", C.c_int), ("general_rows", C.c_int)]

KERNELS = {"sweep": 0, "tile": 1, "sweep_general": 2}


class Fault(C.Structure):
    _fields_ = [(k, C.c_double) for k in
                ("time", "lon_c", "lat_c", "depth_c", "length", "width", "strike", "dip", "rake", "slip")]


SNAPFN = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.c_double, C.c_void_p)
OKADAFN = C.CFUNCTYPE(None, C.c_void_p, C.c_int)
STEPFN = C.CFUNCTYPE(None, C.c_void_p, C.c_int64, C.c_double, C.c_double)


class RunConfig(C.Structure):
    _fields_ = [("t_end", C.c_double), ("save_interval", C.c_double), ("series_interval", C.c_double),
                ("n_faults", C.c_int), ("faults", C.POINTER(Fault)), ("crop_flag", C.c_int),
                ("crop_value", C.c_double), ("n_gauges", C.c_int), ("gauge_pos", _ip),
                ("max_steps", C.c_int64), ("steps_per_sync", C.c_int), ("on_snapshot", SNAPFN),
                ("on_okada", OKADAFN), ("on_step", STEPFN), ("user", C.c_void_p)]


class RunResult(C.Structure):
    _fields_ = [("steps", C.c_int64), ("n_samples", C.c_int), ("n_snapshots", C.c_int),
                ("t_final", C.c_double), ("dt_initial", C.c_double), ("seconds", C.c_double)]


class Request(C.Structure):
    """thy_request (include/thysea_serve.h)."""
    _fields_ = [("arrival_s", C.c_double), ("grid", Grid), ("params", Params), ("state1", _dp), ("state2", _dp),
                ("h", _dp), ("H", _dp), ("qx", _dp), ("qy", _dp), ("run", RunConfig)]


class RequestResult(C.Structure):
    """thy_request_result (include/thysea_serve.h)."""
    _fields_ = [("status", C.c_int), ("admitted_s", C.c_double), ("finished_s", C.c_double), ("steps", C.c_int64),
                ("n_samples", C.c_int), ("max_height", _fp), ("arrival_times", _fp), ("g_time", _fp),
                ("g_eta", _fp), ("g_u", _fp), ("g_v", _fp), ("gauge_capacity", C.c_int)]


class ServeOptions(C.Structure):
    """thy_serve_options (include/thysea_serve.h)."""
    _fields_ = [("capacity", C.c_int), ("hbm_budget", C.c_size_t), ("steps_per_sync", C.c_int),
                ("workers", C.c_int), ("burst", C.c_int)]


SERVE_LIB_PATH = os.path.join(_PKG, "libthysea_serve.so")
_SERVE = None


def serve_lib():
    """libthysea_serve.so: the C++ request loop over the batched C-ABI."""
    global _SERVE
    if _SERVE is None:
        lib()  # libthysea.so first, from this package
        if not os.path.exists(SERVE_LIB_PATH):
            raise ThyError(f"{SERVE_LIB_PATH} is not built (run __graft_entry__.build())")
        L = C.CDLL(SERVE_LIB_PATH)
        L.thy_serve_batched.argtypes = [C.POINTER(Request), C.c_int, C.POINTER(ServeOptions),
                                        C.POINTER(RequestResult), _dp]
        L.thy_serve_sequential.argtypes = [C.POINTER(Request), C.c_int, C.POINTER(ServeOptions),
                                           C.POINTER(RequestResult)]
        _SERVE = L
    return _SERVE


class ThyError(RuntimeError):
    def __init__(self, msg, code=None):
        super().__init__(msg)
        self.code = code


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ThyError(f"{LIB_PATH} is not built (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        L.thy_probe.restype = C.c_int
        L.thy_create.argtypes = [C.POINTER(Grid), C.POINTER(Params), C.POINTER(C.c_void_p)]
        L.thy_destroy.argtypes = [C.c_void_p]
        L.thy_last_error.argtypes = [C.c_void_p]
        L.thy_last_error.restype = C.c_char_p
        L.thy_upload_state.argtypes = [C.c_void_p, _dp, _dp]
        L.thy_download_state.argtypes = [C.c_void_p, _dp, _dp]
        L.thy_init_metrics.argtypes = [C.c_void_p]
        L.thy_update_metrics.argtypes = [C.c_void_p, C.c_double]
        L.thy_apply_okada.argtypes = [C.c_void_p, C.POINTER(Fault), C.c_int, C.c_double, _dp]
        L.thy_initial_dt.argtypes = [C.c_void_p, _dp]
        L.thy_step.argtypes = [C.c_void_p, C.c_double, _dp]
        L.thy_run.argtypes = [C.c_void_p, C.POINTER(RunConfig), C.POINTER(RunResult)]
        L.thy_run_more.argtypes = [C.c_void_p, C.c_int64, C.POINTER(RunResult)]
        L.thy_get_field.argtypes = [C.c_void_p, C.c_int, _fp]
        L.thy_get_gauges.argtypes = [C.c_void_p, _fp, _fp, _fp, _fp, _fp, _fp, _fp]
        L.thy_mass.argtypes = [C.c_void_p, _dp]
        L.thy_mass_noalloc.argtypes = [C.c_void_p, _dp]
        L.thy_footprint.argtypes = [C.c_int, C.c_int, C.POINTER(Params)]
        L.thy_footprint.restype = C.c_size_t
        L.thy_last_timing.argtypes = [C.c_void_p, _dp]
        L.thy_copy_probe.argtypes = [C.c_size_t, C.c_int, _dp]
        L.thy_autotune.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), _dp]
        L.thy_batch_create.argtypes = [C.c_int, C.c_size_t, C.c_int, C.POINTER(C.c_void_p)]
        L.thy_batch_prepare.argtypes = [C.c_void_p, C.POINTER(RunConfig)]
        L.thy_batch_destroy.argtypes = [C.c_void_p]
        L.thy_batch_last_error.argtypes = [C.c_void_p]
        L.thy_batch_last_error.restype = C.c_char_p
        L.thy_batch_add.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(RunConfig), C.POINTER(C.c_int)]
        L.thy_batch_advance.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
        L.thy_batch_launch.argtypes = [C.c_void_p]
        L.thy_batch_ready.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
        L.thy_batch_wait.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
        L.thy_release_cached.argtypes = []
        L.thy_release_cached.restype = None
        L.thy_batch_status.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(RunResult)]
        L.thy_batch_remove.argtypes = [C.c_void_p, C.c_int]
        L.thy_batch_info.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_size_t), C.POINTER(C.c_double)]
        _LIB = L
    return _LIB


def probe():
    return lib().thy_probe()


def release_cached():
    """Release the device memory, pinned memory and streams kept in the library's caches
    (call with no Solver or Batch open)."""
    lib().thy_release_cached()


def params_of(prob, precision=0, track_arrival=True, tile=(0, 0), kernel="sweep", strip_rows=0, general_rows=0):
    p = Params()
    p.borde_sup, p.borde_inf, p.borde_izq, p.borde_der = prob.borde_sup, prob.borde_inf, prob.borde_izq, prob.borde_der
    p.sponge_sup, p.sponge_inf = prob.sponge_sup, prob.sponge_inf
    p.sponge_izq, p.sponge_der = prob.sponge_izq, prob.sponge_der
    p.cfl, p.mf0, p.vmax = prob.cfl, prob.mf0, prob.vmax
    p.eps_h, p.hpos, p.cvis, p.dif_at = prob.eps_h, prob.hpos, prob.cvis, prob.dif_at
    p.hmin, p.sea_level = prob.hmin, 0.0
    p.L, p.H, p.Q, p.T = prob.L, prob.H_ref, prob.Q, prob.T
    p.precision = int(precision)
    p.track_arrival = 1 if track_arrival else 0
    p.tile_x, p.tile_y = int(tile[0]), int(tile[1])
    p.kernel = KERNELS[kernel]
    p.strip_rows = int(strip_rows)
    p.general_rows = int(general_rows)
    return p


def fault_of(f):
    out = Fault()
    for k, _ in Fault._fields_:
        setattr(out, k, float(f[k]))
    return out


class Solver:
    """One grid on the GPU, driven through the C-ABI."""

    def __init__(self, prob, precision=0, track_arrival=True, tile=(0, 0), upload=True, kernel="sweep",
                 strip_rows=0, general_rows=0):
        self.prob = prob
        self.batch_prepared = False
        self._keep = [np.ascontiguousarray(a, np.float64) for a in
                      (prob.lon, prob.lat, prob.area, prob.cosphi, prob.ancho, prob.alto)]
        g = Grid(prob.nx, prob.ny, *[a.ctypes.data_as(_dp) for a in self._keep])
        self.params = params_of(prob, precision, track_arrival, tile, kernel, strip_rows, general_rows)
        h = C.c_void_p()
        rc = lib().thy_create(C.byref(g), C.byref(self.params), C.byref(h))
        self.h = h
        if rc != 0:
            msg = lib().thy_last_error(h).decode() if h.value else ""
            self.close()
            raise ThyError(f"thy_create: {STATUS.get(rc, rc)} {msg}")
        self.n = prob.nx * prob.ny
        if upload:
            self.upload(prob.h, prob.H, prob.qx, prob.qy)

    def _chk(self, rc, what):
        if rc != 0:
            raise ThyError(f"{what}: {STATUS.get(rc, rc)}: {lib().thy_last_error(self.h).decode()}", rc)

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            lib().thy_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, h, H, qx, qy):
        s1 = np.empty((self.n, 2))
        s2 = np.empty((self.n, 2))
        s1[:, 0] = np.ravel(h)
        s1[:, 1] = np.ravel(H)
        s2[:, 0] = np.ravel(qx)
        s2[:, 1] = np.ravel(qy)
        self.batch_prepared = False
        self._chk(lib().thy_upload_state(self.h, s1.ctypes.data_as(_dp), s2.ctypes.data_as(_dp)), "upload")

    def state(self):
        s1 = np.empty((self.n, 2))
        s2 = np.empty((self.n, 2))
        self._chk(lib().thy_download_state(self.h, s1.ctypes.data_as(_dp), s2.ctypes.data_as(_dp)), "download")
        ny, nx = self.prob.ny, self.prob.nx
        return (s1[:, 0].reshape(ny, nx), s1[:, 1].reshape(ny, nx),
                s2[:, 0].reshape(ny, nx), s2[:, 1].reshape(ny, nx))

    def init_metrics(self):
        self._chk(lib().thy_init_metrics(self.h), "init_metrics")

    def update_metrics(self, t):
        self._chk(lib().thy_update_metrics(self.h, float(t)), "update_metrics")

    def okada(self, f, crop_flag=0, crop_value=0.0, want_def=False):
        d = np.empty(self.n) if want_def else None
        ff = fault_of(f)
        self._chk(lib().thy_apply_okada(self.h, C.byref(ff), int(crop_flag), float(crop_value),
                                        d.ctypes.data_as(_dp) if want_def else None), "okada")
        return None if d is None else d.reshape(self.prob.ny, self.prob.nx)

    def initial_dt(self):
        v = C.c_double()
        self._chk(lib().thy_initial_dt(self.h, C.byref(v)), "initial_dt")
        return v.value

    def step(self, dt):
        v = C.c_double()
        self._chk(lib().thy_step(self.h, float(dt), C.byref(v)), "step")
        return v.value

    def field(self, name):
        out = np.empty(self.n, np.float32)
        self._chk(lib().thy_get_field(self.h, FIELDS[name], out.ctypes.data_as(_fp)), "get_field")
        return out.reshape(self.prob.ny, self.prob.nx)

    def mass(self):
        v = C.c_double()
        self._chk(lib().thy_mass(self.h, C.byref(v)), "mass")
        return v.value

    def mass_noalloc(self):
        """thy_mass_noalloc: the same sum through the library's file-scope scratch."""
        v = C.c_double()
        self._chk(lib().thy_mass_noalloc(self.h, C.byref(v)), "mass_noalloc")
        return v.value

    def autotune(self, reps=5):
        """Pick the fastest strip height and general strip height of the row-sweep
        kernel (thy_autotune); returns (strip_rows, ms per step) and keeps the general
        strip height in self.general_rows."""
        rows = C.c_int()
        grows = C.c_int()
        ms = C.c_double()
        self._chk(lib().thy_autotune(self.h, int(reps), C.byref(rows), C.byref(grows), C.byref(ms)), "autotune")
        self.general_rows = grows.value
        return rows.value, ms.value

    def run_more(self, n_steps):
        """Continue the last run for n_steps more steps; returns (steps, t, gpu_seconds)."""
        r = RunResult()
        self._chk(lib().thy_run_more(self.h, int(n_steps), C.byref(r)), "run_more")
        return {"steps": r.steps, "t_final": r.t_final, "seconds": r.seconds, "n_samples": r.n_samples}

    def run_config(self, max_steps=0, steps_per_sync=0, on_snapshot=None, on_okada=None, faults=None,
                   on_step=None, save_interval=None):
        """thy_run_config of this problem; the objects it points to are kept on self."""
        pr = self.prob
        faults = sorted(pr.faults if faults is None else faults, key=lambda f: f["time"])
        farr = (Fault * max(1, len(faults)))(*[fault_of(f) for f in faults])
        gpos = np.ascontiguousarray(pr.gauge_pos if pr.gauge_pos is not None else np.zeros(0), np.int32)
        c = RunConfig()
        c.t_end, c.series_interval = pr.t_end, pr.series_interval
        c.save_interval = pr.save_interval if save_interval is None else save_interval
        c.n_faults, c.faults = len(faults), farr
        c.crop_flag, c.crop_value = pr.crop_flag, pr.crop_value
        c.n_gauges, c.gauge_pos = len(gpos), gpos.ctypes.data_as(_ip)
        c.max_steps, c.steps_per_sync = int(max_steps), int(steps_per_sync)
        snap = SNAPFN(lambda u, i, t, ctx: on_snapshot(i, t)) if on_snapshot else SNAPFN()
        oka = OKADAFN(lambda u, k: on_okada(k)) if on_okada else OKADAFN()
        stp = STEPFN(lambda u, it, dt, t: on_step(it, dt, t)) if on_step else STEPFN()
        c.on_snapshot, c.on_okada, c.on_step = snap, oka, stp
        self._cb = (snap, oka, stp, farr, gpos, c)  # must outlive a later run_more / batch advance
        return c

    def gauges(self, n_samples):
        ng = len(self._cb[4])
        ns = n_samples
        bufs = [np.zeros(max(1, ns), np.float32)] + [np.zeros(max(1, ns * ng), np.float32) for _ in range(3)] + \
               [np.zeros(max(1, ng), np.float32) for _ in range(3)]
        self._chk(lib().thy_get_gauges(self.h, *[b.ctypes.data_as(_fp) for b in bufs]), "get_gauges")
        return {"n_samples": ns, "g_time": bufs[0][:ns],
                "g_eta": bufs[1][:ns * ng].reshape(ns, ng), "g_u": bufs[2][:ns * ng].reshape(ns, ng),
                "g_v": bufs[3][:ns * ng].reshape(ns, ng), "g_eta_min": bufs[4][:ng], "g_eta_max": bufs[5][:ng],
                "g_bathy": bufs[6][:ng]}

    def prepare_batch(self, steps_per_sync=0, max_steps=0):
        """The context-local part of admission to a batch (thy_batch_prepare): run
        configuration, metrics, initial dt, faults due at t = 0.  May run on another
        thread than the batch's; Batch.add then admits the solver as prepared."""
        c = self.run_config(max_steps=max_steps, steps_per_sync=steps_per_sync, save_interval=-1.0)
        self.batch_prepared = False
        self._chk(lib().thy_batch_prepare(self.h, C.byref(c)), "batch_prepare")
        self.batch_prepared = True

    def run(self, max_steps=0, steps_per_sync=0, on_snapshot=None, on_okada=None, faults=None, on_step=None):
        c = self.run_config(max_steps, steps_per_sync, on_snapshot, on_okada, faults, on_step)
        self.batch_prepared = False
        r = RunResult()
        self._chk(lib().thy_run(self.h, C.byref(c), C.byref(r)), "run")
        out = {"steps": r.steps, "t_final": r.t_final, "dt_initial": r.dt_initial, "seconds": r.seconds,
               "n_snapshots": r.n_snapshots}
        out.update(self.gauges(r.n_samples))
        return out


SLOT_FREE, SLOT_RUNNING, SLOT_FINISHED, SLOT_FAILED = 0, 1, 2, 3


class Batch:
    """Many scenarios advanced together (thy_batch_* in include/thysea.h).

    add() admits a Solver (created with upload) with its run configuration; advance()
    runs one captured step sequence for every running scenario; finished scenarios
    are read through their Solver and released with remove()."""

    def __init__(self, capacity, hbm_budget=0, steps_per_sync=0):
        h = C.c_void_p()
        rc = lib().thy_batch_create(int(capacity), int(hbm_budget), int(steps_per_sync), C.byref(h))
        self.h = h
        if rc != 0:
            raise ThyError(f"thy_batch_create: {STATUS.get(rc, rc)}")
        self.solvers = {}

    def _chk(self, rc, what):
        if rc != 0:
            msg = lib().thy_batch_last_error(self.h).decode(errors="replace")
            raise ThyError(f"{what}: {STATUS.get(rc, rc)}: {msg}", rc)

    def add(self, solver, steps_per_sync=0, max_steps=0):
        """Admit a scenario; returns its slot, or None when the budget or the slots are
        full.  A solver prepared by Solver.prepare_batch is admitted as prepared (the
        arguments here are then unused)."""
        slot = C.c_int()
        if solver.batch_prepared:
            rc = lib().thy_batch_add(self.h, solver.h, None, C.byref(slot))
        else:
            c = solver.run_config(max_steps=max_steps, steps_per_sync=steps_per_sync, save_interval=-1.0)
            rc = lib().thy_batch_add(self.h, solver.h, C.byref(c), C.byref(slot))
        if rc == STATUS_CODES["BUDGET"]:
            return None
        self._chk(rc, "batch_add")
        solver.batch_prepared = False
        self.solvers[slot.value] = solver
        return slot.value

    def advance(self):
        """One step sequence for every running scenario; returns how many still run."""
        n = C.c_int()
        self._chk(lib().thy_batch_advance(self.h, C.byref(n)), "batch_advance")
        return n.value

    def launch(self):
        """Queue one step sequence and return at once (collect it with wait())."""
        self._chk(lib().thy_batch_launch(self.h), "batch_launch")

    def ready(self):
        """True when no launch is in flight or it has finished."""
        d = C.c_int()
        self._chk(lib().thy_batch_ready(self.h, C.byref(d)), "batch_ready")
        return bool(d.value)

    def wait(self):
        """Collect the launch in flight; returns how many scenarios still run."""
        n = C.c_int()
        self._chk(lib().thy_batch_wait(self.h, C.byref(n)), "batch_wait")
        return n.value

    def status(self, slot):
        st = C.c_int()
        r = RunResult()
        self._chk(lib().thy_batch_status(self.h, int(slot), C.byref(st), C.byref(r)), "batch_status")
        return st.value, {"steps": r.steps, "t_final": r.t_final, "dt_initial": r.dt_initial, "n_samples": r.n_samples}

    def finished(self):
        return [s for s in self.solvers if self.status(s)[0] == SLOT_FINISHED]

    def result(self, slot):
        """Max height, arrival times and gauges of a finished scenario."""
        st, r = self.status(slot)
        if st == SLOT_FAILED:
            msg = lib().thy_last_error(self.solvers[slot].h).decode()
            raise ThyError(f"slot {slot} failed: {msg}", STATUS_CODES["NONFINITE"])
        if st != SLOT_FINISHED:
            raise ThyError(f"slot {slot} is not finished")
        s = self.solvers[slot]
        out = dict(r)
        out["max_height"] = s.field("max_height")
        out["arrival_times"] = s.field("arrival_times")
        out.update(s.gauges(r["n_samples"]))
        return out

    def remove(self, slot):
        self._chk(lib().thy_batch_remove(self.h, int(slot)), "batch_remove")
        return self.solvers.pop(slot)

    def info(self):
        n = C.c_int()
        b = C.c_size_t()
        sec = C.c_double()
        self._chk(lib().thy_batch_info(self.h, C.byref(n), C.byref(b), C.byref(sec)), "batch_info")
        return {"running": n.value, "bytes": b.value, "gpu_seconds": sec.value}

    def close(self):
        if self.h:
            lib().thy_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def copy_probe(nbytes, reps=10):
    """Time a plain device copy of nbytes; returns (ms per copy, GB/s of read+write)."""
    ms = C.c_double()
    rc = lib().thy_copy_probe(int(nbytes), int(reps), C.byref(ms))
    if rc != 0:
        raise ThyError(f"copy probe: {STATUS.get(rc, rc)}")
    return ms.value, 2.0 * nbytes / (ms.value * 1e-3) / 1e9


def footprint(nx, ny, precision=0, track_arrival=True):
    p = Params()
    p.precision = precision
    p.track_arrival = 1 if track_arrival else 0
    return lib().thy_footprint(nx, ny, C.byref(p))
