"""ctypes binding of libisph.so (include/isph_capi.h).

``IsphSim`` mirrors the reference's step interface: one method per
subroutine of ``main.f90:41-66`` (projection, cellshift, resetid,
neighbour, effpor, eddyvis, int_vel, ppesolve, comp_vel, comp_pos,
cellshiftalt, freesurf, opt2_shift, timestep), each a thin call into the
gfx950 library.  Errors come back as :class:`IsphError` carrying the C
status code and ``isph_last_error()`` text, the counterpart of the
reference's silent fall-through (``solver.f90:330-343``).

The binding is generic over the symbol prefix so that the CPU oracle,
which exposes the same API with an ``ora_`` prefix, can reuse it from the
tests; this module itself never loads the oracle.
"""
import ctypes
import os

import numpy as np

from .params import ERRORS, FIELD, IsphParams

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libisph.so")

_c_dp = ctypes.POINTER(ctypes.c_double)
_c_ip = ctypes.POINTER(ctypes.c_int32)

SUBROUTINES = [
    "projection", "cellshift", "resetid", "neighbour", "effpor", "eddyvis",
    "int_vel", "ppe_assemble", "comp_vel", "comp_pos", "cellshiftalt",
    "freesurf", "opt2_shift", "timestep", "scalart", "densityupdate", "eddyvis_lagged",
]


OBJ_MAX = 8

# stage ids of isph_adjoint_stage (include/isph_capi.h)
ADJ_STAGE = {
    "projection": 1, "neighbour": 3, "effpor": 4, "eddyvis": 5, "varts": 6, "vstar": 7,
    "ppesolve": 8, "comp_vel": 10, "comp_pos": 11, "shift": 16, "objective": 20,
}


class IsphObjective(ctypes.Structure):
    """isph_objective: Shepard pressure probes and water-depth gauges."""
    _fields_ = [("nprobe", ctypes.c_int32), ("ngauge", ctypes.c_int32),
                ("probe_x", ctypes.c_double * OBJ_MAX), ("probe_y", ctypes.c_double * OBJ_MAX),
                ("gauge_x", ctypes.c_double * OBJ_MAX)]


def make_objective(probes=(), gauges=()):
    """probes: sequence of (x, y); gauges: sequence of x."""
    o = IsphObjective()
    probes, gauges = list(probes), list(gauges)
    if len(probes) > OBJ_MAX or len(gauges) > OBJ_MAX:
        raise ValueError(f"at most {OBJ_MAX} probes and {OBJ_MAX} gauges")
    o.nprobe, o.ngauge = len(probes), len(gauges)
    for k, (x, y) in enumerate(probes):
        o.probe_x[k], o.probe_y[k] = x, y
    for k, x in enumerate(gauges):
        o.gauge_x[k] = x
    return o


MAX_ZONES = 4


class IsphBufferZone(ctypes.Structure):
    """isph_buffer_zone: a buffer zone's template, entry cells and schedule."""
    _fields_ = [("ntemplate", ctypes.c_int32), ("ncells", ctypes.c_int32),
                ("x", _c_dp), ("y", _c_dp), ("vx", _c_dp), ("vy", _c_dp), ("pressure", _c_dp),
                ("flags", _c_ip), ("cells", _c_ip),
                ("mass", ctypes.c_double), ("density", ctypes.c_double),
                ("ovol", ctypes.c_double), ("con", ctypes.c_double),
                ("period", ctypes.c_double), ("xcut", ctypes.c_double),
                ("remove_right", ctypes.c_int32), ("reserved", ctypes.c_int32)]


def make_zones(zones):
    """ctypes zones from dicts (x, y, vx, vy, pressure, flags, cells (k, 2),
    mass, density, ovol, con, period, xcut, remove_right); returns the array
    and the numpy buffers it points into (keep both alive for the call)."""
    arr = (IsphBufferZone * max(len(zones), 1))()
    keep = []
    for k, z in enumerate(zones):
        a = {f: np.ascontiguousarray(z[f], dtype=np.float64) for f in ("x", "y", "vx", "vy",
                                                                      "pressure")}
        fl = np.ascontiguousarray(z["flags"], dtype=np.int32)
        cells = np.ascontiguousarray(np.asarray(z["cells"], dtype=np.int32).reshape(-1, 2))
        keep += list(a.values()) + [fl, cells]
        q = arr[k]
        q.ntemplate, q.ncells = len(a["x"]), len(cells)
        q.x, q.y, q.vx, q.vy, q.pressure = (_dp(a[f]) for f in ("x", "y", "vx", "vy", "pressure"))
        q.flags, q.cells = _ip(fl), _ip(cells)
        for f in ("mass", "density", "ovol", "con", "period", "xcut"):
            setattr(q, f, float(z[f]))
        q.remove_right, q.reserved = int(z["remove_right"]), 0
    return arr, keep


class IsphError(RuntimeError):
    def __init__(self, code, where, msg):
        self.code = code
        super().__init__(f"{where}: {ERRORS.get(code, code)} ({msg})")


def _dp(a):
    return None if a is None else a.ctypes.data_as(_c_dp)


def _ip(a):
    return None if a is None else a.ctypes.data_as(_c_ip)


def declare(lib, prefix):
    """Declare argument/return types of every entry point under ``prefix``."""
    P = ctypes.c_void_p
    i32 = ctypes.c_int32
    sig = {
        "create": [ctypes.POINTER(IsphParams), i32, i32, ctypes.POINTER(P)],
        "destroy": [P],
        "set_params": [P, ctypes.POINTER(IsphParams)],
        "set_particles": [P, i32] + [_c_dp] * 8 + [_c_ip] * 3,
        "get_particles": [P, _c_ip] + [_c_dp] * 8 + [_c_ip] * 3,
        "set_soil": [P, i32] + [_c_dp] * 4,
        "get_field": [P, i32, _c_dp],
        "set_field": [P, i32, _c_dp],
        "get_neighbours": [P, _c_ip, _c_ip],
        "get_csr": [P, _c_ip, _c_ip, _c_dp, _c_dp],
        "get_scalars": [P, _c_dp, _c_dp, _c_dp, _c_ip, _c_ip, _c_dp],
        "set_scalars": [P, ctypes.c_double, ctypes.c_double, ctypes.c_double],
        "ppe_solve": [P, _c_ip, _c_dp],
        "ppesolve": [P, _c_ip, _c_dp],
        "step": [P, i32],
        "scalarupdate": [P, ctypes.c_double],
        "salt_init": [P, ctypes.c_double, ctypes.c_double, ctypes.c_double],
        "get_substeps": [P, _c_ip, _c_dp],
        "probe": [P, i32, _c_dp, _c_dp, _c_dp],
        "set_buffer_zones": [P, i32, ctypes.POINTER(IsphBufferZone), ctypes.c_double],
        "remove_buffer": [P, i32],
        "insert_buffer": [P, i32],
    }
    for name in SUBROUTINES:
        sig[name] = [P]
    for name, args in sig.items():
        f = getattr(lib, prefix + name)
        f.argtypes = args
        f.restype = ctypes.c_int
    f = getattr(lib, prefix + "last_error")
    f.argtypes = [P]
    f.restype = ctypes.c_char_p
    return lib


_LIB = None


def load_library(path=LIB_PATH):
    """Load libisph.so.  There is no fallback: a missing library is an error."""
    global _LIB
    if _LIB is None:
        # torch ships its own libamdhip64.so.7 (same soname as /opt/rocm's);
        # loading torch first makes the process use one HIP runtime, shared
        # by torch and this library.
        import torch  # noqa: F401
        if not os.path.exists(path):
            raise ImportError(
                f"{path} is not built; run __graft_entry__.build() (hipcc --offload-arch=gfx950)")
        lib = ctypes.CDLL(path)
        declare(lib, "isph_")
        for name, args, res in [
            ("get_params", [ctypes.c_void_p, ctypes.POINTER(IsphParams)], ctypes.c_int),
            ("profile_enable", [ctypes.c_void_p, ctypes.c_int32], ctypes.c_int),
            ("get_kernel_time", [ctypes.c_void_p, ctypes.c_int32,
                                 ctypes.POINTER(ctypes.c_int64), _c_dp], ctypes.c_int),
            ("copy_bench", [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, _c_dp], ctypes.c_int),
            ("checkpoint", [ctypes.c_void_p, ctypes.c_int32], ctypes.c_int),
            ("restore", [ctypes.c_void_p, ctypes.c_int32], ctypes.c_int),
            ("adjoint_forward", [ctypes.c_void_p, ctypes.POINTER(IsphObjective), ctypes.c_int32,
                                 ctypes.c_int32, _c_dp], ctypes.c_int),
            ("adjoint_backward", [ctypes.c_void_p, _c_dp, _c_dp, ctypes.POINTER(ctypes.c_int64)],
             ctypes.c_int),
            ("adjoint_stage", [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(IsphObjective)],
             ctypes.c_int),
            ("adjoint_zero", [ctypes.c_void_p], ctypes.c_int),
            ("adjoint_get", [ctypes.c_void_p, ctypes.c_int32, _c_dp], ctypes.c_int),
            ("adjoint_set", [ctypes.c_void_p, ctypes.c_int32, _c_dp], ctypes.c_int),
            ("adjoint_params", [ctypes.c_void_p, _c_dp, _c_dp], ctypes.c_int),
            ("objective_eval", [ctypes.c_void_p, ctypes.POINTER(IsphObjective), _c_dp],
             ctypes.c_int),
            ("combined", [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int32], ctypes.c_int),
            ("format_f", [ctypes.c_double, ctypes.c_int32, ctypes.c_int32, ctypes.c_char_p],
             ctypes.c_int),
            ("format_es", [ctypes.c_double, ctypes.c_int32, ctypes.c_int32, ctypes.c_char_p],
             ctypes.c_int),
            ("snapshot_record", [ctypes.c_double] * 6 + [ctypes.c_char_p], ctypes.c_int),
            ("write_probe_series", [ctypes.c_char_p, ctypes.c_int32, _c_dp, _c_dp],
             ctypes.c_int),
            ("probe_noalloc", [ctypes.c_void_p, ctypes.c_int32, _c_dp, _c_dp, _c_dp],
             ctypes.c_int),
            ("set_progress", [ctypes.c_void_p, PROGRESS_FN, ctypes.c_void_p], ctypes.c_int),
        ]:
            f = getattr(lib, "isph_" + name)
            f.argtypes = args
            f.restype = res
        _LIB = lib
    return _LIB


# void fn(void* user, const char* what, int32_t done, int32_t total)
PROGRESS_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int32,
                               ctypes.c_int32)


class Sim:
    """One simulation context behind an ISPH C API (prefix ``isph_`` for the
    gfx950 library)."""

    def __init__(self, lib, prefix, params, cap, cap_soil=0):
        self._lib = lib
        self._pre = prefix
        self.params = params.copy()
        self.cap = int(cap)
        self.cap_soil = int(cap_soil)
        h = ctypes.c_void_p()
        rc = self._fn("create")(ctypes.byref(self.params), self.cap, self.cap_soil, ctypes.byref(h))
        if rc != 0:
            raise IsphError(rc, prefix + "create", "context creation failed")
        self._h = h
        self.n = 0

    # -- plumbing ---------------------------------------------------------
    def _fn(self, name):
        return getattr(self._lib, self._pre + name)

    def _check(self, rc, where):
        if rc != 0:
            msg = self._fn("last_error")(self._h)
            raise IsphError(rc, self._pre + where, msg.decode() if msg else "")
        return rc

    def close(self):
        if getattr(self, "_h", None):
            self._fn("destroy")(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- state ------------------------------------------------------------
    def set_params(self, params):
        self._check(self._fn("set_params")(self._h, ctypes.byref(params)), "set_params")
        self.params = params.copy()

    def set_particles(self, parts):
        n = len(parts["x"])
        a = {k: np.ascontiguousarray(parts[k], dtype=np.float64)
             for k in ("x", "y", "vx", "vy", "mass", "density", "oden", "pressure")}
        t = {k: np.ascontiguousarray(parts[k], dtype=np.int32) for k in ("tid", "pid")}
        fl = parts.get("flags")
        fl = None if fl is None else np.ascontiguousarray(fl, dtype=np.int32)
        self._check(self._fn("set_particles")(
            self._h, n, _dp(a["x"]), _dp(a["y"]), _dp(a["vx"]), _dp(a["vy"]),
            _dp(a["mass"]), _dp(a["density"]), _dp(a["oden"]), _dp(a["pressure"]),
            _ip(t["tid"]), _ip(fl), _ip(t["pid"])), "set_particles")
        self.n = self.count()

    def count(self):
        n = ctypes.c_int32()
        self._check(self._fn("get_particles")(self._h, ctypes.byref(n), *([None] * 11)),
                    "get_particles")
        return n.value

    def get_particles(self):
        n = self.count()
        d = {k: np.empty(n, np.float64) for k in
             ("x", "y", "vx", "vy", "mass", "density", "oden", "pressure")}
        i = {k: np.empty(n, np.int32) for k in ("tid", "flags", "pid")}
        nn = ctypes.c_int32()
        self._check(self._fn("get_particles")(
            self._h, ctypes.byref(nn), _dp(d["x"]), _dp(d["y"]), _dp(d["vx"]), _dp(d["vy"]),
            _dp(d["mass"]), _dp(d["density"]), _dp(d["oden"]), _dp(d["pressure"]),
            _ip(i["tid"]), _ip(i["flags"]), _ip(i["pid"])), "get_particles")
        d.update(i)
        self.n = n
        return d

    def set_soil(self, soil):
        ns = len(soil["x"])
        a = {k: np.ascontiguousarray(soil[k], dtype=np.float64)
             for k in ("x", "y", "mass", "density")}
        self._check(self._fn("set_soil")(self._h, ns, _dp(a["x"]), _dp(a["y"]),
                                         _dp(a["mass"]), _dp(a["density"])), "set_soil")

    def get_field(self, name):
        n = self.count()
        w = 4 if name == "coff" else 1
        out = np.empty(n * w, np.float64)
        self._check(self._fn("get_field")(self._h, FIELD[name], _dp(out)), "get_field")
        return out.reshape(n, 4) if w == 4 else out

    def set_field(self, name, values):
        v = np.ascontiguousarray(values, dtype=np.float64).ravel()
        self._check(self._fn("set_field")(self._h, FIELD[name], _dp(v)), "set_field")

    def get_neighbours(self):
        n = self.count()
        ptr = np.empty(n + 1, np.int32)
        self._check(self._fn("get_neighbours")(self._h, _ip(ptr), None), "get_neighbours")
        idx = np.empty(int(ptr[-1]), np.int32)
        self._check(self._fn("get_neighbours")(self._h, _ip(ptr), _ip(idx)), "get_neighbours")
        return ptr, idx

    def get_csr(self):
        n = self.count()
        rp = np.empty(n + 1, np.int32)
        self._check(self._fn("get_csr")(self._h, _ip(rp), None, None, None), "get_csr")
        nnz = int(rp[-1])
        col = np.empty(nnz, np.int32)
        val = np.empty(nnz, np.float64)
        rhs = np.empty(n, np.float64)
        self._check(self._fn("get_csr")(self._h, _ip(rp), _ip(col), _dp(val), _dp(rhs)),
                    "get_csr")
        return rp, col, val, rhs

    def scalars(self):
        u, dt, t, rel = (ctypes.c_double() for _ in range(4))
        it, outer = ctypes.c_int32(), ctypes.c_int32()
        self._check(self._fn("get_scalars")(
            self._h, ctypes.byref(u), ctypes.byref(dt), ctypes.byref(t), ctypes.byref(it),
            ctypes.byref(outer), ctypes.byref(rel)), "get_scalars")
        return dict(umax=u.value, dt=dt.value, t=t.value, iters=it.value,
                    outer=outer.value, relres=rel.value)

    def set_scalars(self, umax, dt, t):
        self._check(self._fn("set_scalars")(self._h, umax, dt, t), "set_scalars")

    # -- the reference's subroutines (main.f90:41-66) ----------------------
    def call(self, name):
        return self._check(self._fn(name)(self._h), name)

    def projection(self): return self.call("projection")
    def cellshift(self): return self.call("cellshift")
    def resetid(self): return self.call("resetid")
    def neighbour(self): return self.call("neighbour")
    def effpor(self): return self.call("effpor")
    def eddyvis(self): return self.call("eddyvis")
    def int_vel(self): return self.call("int_vel")
    def ppe_assemble(self): return self.call("ppe_assemble")
    def comp_vel(self): return self.call("comp_vel")
    def comp_pos(self): return self.call("comp_pos")
    def cellshiftalt(self): return self.call("cellshiftalt")
    def freesurf(self): return self.call("freesurf")
    def opt2_shift(self): return self.call("opt2_shift")
    def timestep(self): return self.call("timestep")

    def ppe_solve(self):
        it, rel = ctypes.c_int32(), ctypes.c_double()
        self._check(self._fn("ppe_solve")(self._h, ctypes.byref(it), ctypes.byref(rel)),
                    "ppe_solve")
        return it.value, rel.value

    def ppesolve(self):
        it, rel = ctypes.c_int32(), ctypes.c_double()
        self._check(self._fn("ppesolve")(self._h, ctypes.byref(it), ctypes.byref(rel)),
                    "ppesolve")
        return it.value, rel.value

    def step(self, nsteps=1):
        self._check(self._fn("step")(self._h, int(nsteps)), "step")

    # -- salinity transport (scalar.f90) -------------------------------------
    def scalart(self): return self.call("scalart")
    def eddyvis_lagged(self): return self.call("eddyvis_lagged")
    def densityupdate(self): return self.call("densityupdate")

    def scalarupdate(self, stime):
        self._check(self._fn("scalarupdate")(self._h, float(stime)), "scalarupdate")

    def salt_init(self, xl, yl, line_grad):
        """Salt-water start of main.f90:160-195."""
        self._check(self._fn("salt_init")(self._h, float(xl), float(yl), float(line_grad)),
                    "salt_init")

    def probe(self, points):
        """Shepard pressure at (x, y) points (probe.f90:86-130)."""
        pts = np.asarray(points, dtype=np.float64).reshape(-1, 2)
        px = np.ascontiguousarray(pts[:, 0])
        py = np.ascontiguousarray(pts[:, 1])
        out = np.zeros(len(px))
        self._check(self._fn("probe")(self._h, len(px), _dp(px), _dp(py), _dp(out)), "probe")
        return out

    # -- buffer zones (internal.f90:475-614) -----------------------------------
    def set_buffer_zones(self, zones, time_shift=0.0):
        arr, keep = make_zones(zones)
        self._check(self._fn("set_buffer_zones")(self._h, len(zones), arr, float(time_shift)),
                    "set_buffer_zones")
        del keep

    def remove_buffer(self, zone):
        self._check(self._fn("remove_buffer")(self._h, int(zone)), "remove_buffer")

    def insert_buffer(self, zone):
        self._check(self._fn("insert_buffer")(self._h, int(zone)), "insert_buffer")

    def substeps(self):
        """(solsteps, dtsol) of the next step's transport."""
        n, d = ctypes.c_int32(), ctypes.c_double()
        self._check(self._fn("get_substeps")(self._h, ctypes.byref(n), ctypes.byref(d)),
                    "get_substeps")
        return n.value, d.value

    def init_state(self):
        """Start-up as main.f90:35-37 (effpor) plus one timestep() so that
        umax is the initial max |v|/n rather than 0."""
        self.effpor()
        self.timestep()


class IsphSim(Sim):
    """Context of the gfx950 library."""

    def __init__(self, params, cap, cap_soil=0, lib=None):
        super().__init__(lib or load_library(), "isph_", params, cap, cap_soil)

    def checkpoint(self, slot):
        self._check(self._lib.isph_checkpoint(self._h, int(slot)), "checkpoint")

    def restore(self, slot):
        self._check(self._lib.isph_restore(self._h, int(slot)), "restore")
        self.n = self.count()

    def set_progress(self, fn):
        """fn(what, done, total) after each checkpoint segment of the adjoint
        forward and backward sweeps; None turns it off."""
        if fn is None:
            self._progress = None
            cb = ctypes.cast(None, PROGRESS_FN)
        else:
            self._progress = PROGRESS_FN(lambda user, what, done, total:
                                         fn(what.decode(), done, total))
            cb = self._progress
        self._check(self._lib.isph_set_progress(self._h, cb, None), "set_progress")

    def profile_enable(self, stride=1):
        """Time every `stride`-th SpMV launch with HIP events (0: off)."""
        self._check(self._lib.isph_profile_enable(self._h, int(stride)), "profile_enable")

    def kernel_time(self, kernel):
        n = ctypes.c_int64()
        ms = ctypes.c_double()
        self._check(self._lib.isph_get_kernel_time(self._h, kernel, ctypes.byref(n),
                                                   ctypes.byref(ms)), "get_kernel_time")
        return n.value, ms.value

    def copy_bench(self, nbytes, reps):
        ms = ctypes.c_double()
        self._check(self._lib.isph_copy_bench(self._h, int(nbytes), int(reps), ctypes.byref(ms)),
                    "copy_bench")
        return ms.value

    # -- discrete adjoint ---------------------------------------------------
    def adjoint_forward(self, obj, nsteps, every=10):
        """Porous closure + timestep at the current state, then nsteps steps
        with checkpoints every `every` steps; returns the objective J."""
        J = ctypes.c_double()
        self._check(self._lib.isph_adjoint_forward(self._h, ctypes.byref(obj), int(nsteps),
                                                   int(every), ctypes.byref(J)), "adjoint_forward")
        return J.value

    def adjoint_backward(self):
        """dJ/d(por, alpha, beta) of the last adjoint_forward, J, and the
        BiCGSTAB iterations of the transposed solves."""
        g = np.zeros(3)
        J = ctypes.c_double()
        it = ctypes.c_int64()
        self._check(self._lib.isph_adjoint_backward(self._h, _dp(g), ctypes.byref(J),
                                                    ctypes.byref(it)), "adjoint_backward")
        return g, J.value, it.value

    def adjoint_zero(self):
        self._check(self._lib.isph_adjoint_zero(self._h), "adjoint_zero")

    def adjoint_stage(self, stage, obj=None):
        sid = ADJ_STAGE[stage] if isinstance(stage, str) else int(stage)
        self._check(self._lib.isph_adjoint_stage(self._h, sid,
                                                 ctypes.byref(obj) if obj is not None else None),
                    "adjoint_stage")

    def adjoint_get(self, name):
        n = self.count()
        w = 4 if name == "coff" else 1
        out = np.empty(n * w, np.float64)
        self._check(self._lib.isph_adjoint_get(self._h, FIELD[name], _dp(out)), "adjoint_get")
        return out.reshape(n, 4) if w == 4 else out

    def adjoint_set(self, name, values):
        v = np.ascontiguousarray(values, dtype=np.float64).ravel()
        self._check(self._lib.isph_adjoint_set(self._h, FIELD[name], _dp(v)), "adjoint_set")

    def adjoint_params(self):
        """(dJ/d(por, alpha, beta), dJ/d umax) accumulated by stage adjoints."""
        g = np.zeros(3)
        u = ctypes.c_double()
        self._check(self._lib.isph_adjoint_params(self._h, _dp(g), ctypes.byref(u)),
                    "adjoint_params")
        return g, u.value

    def combined(self, directory, index):
        """Snapshot files fluid/fluid_<index>.txt, buffer/buffer_<index>.txt
        and fluid/dead_<index>.txt under `directory` (output.f90:188-234)."""
        self._check(self._lib.isph_combined(self._h, os.fsencode(directory), int(index)),
                    "combined")

    def objective_eval(self, obj):
        J = ctypes.c_double()
        self._check(self._lib.isph_objective_eval(self._h, ctypes.byref(obj), ctypes.byref(J)),
                    "objective_eval")
        return J.value


def run_step_stages(sim, nsteps=1):
    """The step as separate calls in main.f90:41-66 order (t advances first,
    :42-43); used to compare stage by stage."""
    for _ in range(nsteps):
        s = sim.scalars()
        sim.set_scalars(s["umax"], s["dt"], s["t"] + s["dt"])
        sim.projection()
        sim.cellshift()
        sim.resetid()
        sim.neighbour()
        sim.effpor()
        sim.eddyvis()
        sim.int_vel()
        sim.ppesolve()
        sim.comp_vel()
        sim.comp_pos()
        sim.cellshiftalt()
        sim.neighbour()
        sim.freesurf()
        sim.effpor()
        sim.opt2_shift()
        sim.timestep()


def format_f(v, w, d):
    """Fortran Fw.d of v, as the library writes it."""
    buf = ctypes.create_string_buffer(w + 1)
    if load_library().isph_format_f(float(v), w, d, buf) != 0:
        raise ValueError("bad edit descriptor")
    return buf.value.decode()


def format_es(v, w, d):
    """Fortran ESw.d of v, as the library writes it."""
    buf = ctypes.create_string_buffer(w + 1)
    if load_library().isph_format_es(float(v), w, d, buf) != 0:
        raise ValueError("bad edit descriptor")
    return buf.value.decode()


def snapshot_record(x, y, p, vx, vy, con):
    """One record of output.f90:206-209 (without the newline)."""
    buf = ctypes.create_string_buffer(80)
    load_library().isph_snapshot_record(x, y, p, vx, vy, con, buf)
    return buf.value.decode()


def write_probe_series(path, t, p, g=-9.81, wc=0.3, rho=1000.0):
    """probe.txt of output.f90:150-166: records (t sqrt(|g|/wc),
    p / (rho |g| wc)) (probe.f90:126-127) as (F10.3, 1X, F10.3)."""
    a = np.ascontiguousarray(np.asarray(t, np.float64) * np.sqrt(abs(g) / wc))
    b = np.ascontiguousarray(np.asarray(p, np.float64) / (rho * abs(g) * wc))
    rc = load_library().isph_write_probe_series(os.fsencode(path), len(a), _dp(a), _dp(b))
    if rc != 0:
        raise OSError(f"cannot write {path}")
