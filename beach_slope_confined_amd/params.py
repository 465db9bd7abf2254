"""ctypes mirror of ``isph_params`` (include/isph_capi.h) and named presets.

The reference keeps its run parameters as module globals with compiled-in
defaults (``initialize.f90:19-79``) plus constants scattered through the
physics files (``kernel.f90:30``, ``porous.f90:111``, ``part_shift.f90:71``,
``turbulence.f90:16``); ``setup.f90:47-82`` derives ``h1``, ``dl`` and
``lam``.  Here they are one struct passed across the C ABI.
"""
import ctypes

KERNEL_HYBRID = 0
KERNEL_WENDLAND_C2 = 1
DRAG_REFERENCE = 0
DRAG_ERGUN = 1
PREC_FP64 = 0
PREC_MIXED = 1
PREC_MIXED_V64 = 2
PREC_MIXED_MG = 3
PREC_FGMRES = 4
PREC_IR = 5
PREC_FP64_MG = 6
VISC_REFERENCE = 0
VISC_CONSTANT = 1
ORDER_CELL = 0
ORDER_REFERENCE = 1

FLAG_DEAD = 1
FLAG_BUFFER = 2
FLAG_FREE = 4
FLAG_VICINITY = 8

FIELD = {
    "xs": 0, "ys": 1, "vxs": 2, "vys": 3, "porosity": 4, "resist": 5,
    "coff": 6, "divr": 7, "nut": 8, "tke": 9, "varts": 10, "vel": 11, "rhs": 12,
    "x": 13, "y": 14, "vx": 15, "vy": 16, "density": 17, "pressure": 18,
    "con": 19, "ovol": 20, "cdiff": 21, "mass": 22, "oden": 23,
}

OK = 0
ERRORS = {
    -1: "ISPH_E_ARG", -2: "ISPH_E_HIP", -3: "ISPH_E_CAPACITY", -4: "ISPH_E_NAN",
    -5: "ISPH_E_NOCONV", -6: "ISPH_E_DOMAIN", -7: "ISPH_E_SINGULAR", -8: "ISPH_E_STATE",
}

_DOUBLES = [
    "h", "cell_size", "xmin", "ymin", "xmax", "ymax", "dl", "lam", "dt", "rho0",
    "rel_den", "g", "nu", "csh", "maxshift", "lamfs", "maxdivr", "por", "alpha",
    "beta", "d50", "tol", "inner_tol", "xlcutoff", "ytcutoff", "sig1", "mu",
    "rhomin", "tschmidt", "salt_xbuf", "xrelease_lo", "xrelease_hi",
]
_INTS = [
    "kernel", "drag_law", "precision", "maxit", "inner_maxit", "max_outer",
    "max_nbr", "shifting", "adaptive_dt", "sink", "visc_law", "salinity",
    "cell_order", "reserved",
]


class IsphParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in _DOUBLES] + [(n, ctypes.c_int32) for n in _INTS]

    def as_dict(self):
        return {n: getattr(self, n) for n in _DOUBLES + _INTS}

    def copy(self, **changes):
        p = IsphParams()
        ctypes.pointer(p)[0] = self
        for k, v in changes.items():
            if not hasattr(p, k):
                raise KeyError(k)
            setattr(p, k, v)
        return p


assert ctypes.sizeof(IsphParams) == 8 * len(_DOUBLES) + 4 * len(_INTS)


def make_params(**kw):
    """Parameters with the reference's constants where the configs do not set
    them (``initialize.f90:21-43``: csh 0.5, maxshift 0.1, lamfs 0.8,
    maxdivr 2, rel_den 1.025, sig1 0.2, mu 1e-3, rhomin 1000, tschmidt 1;
    the buffer reset line x < 1.6 of ``scalar.f90:252``)."""
    p = IsphParams()
    defaults = dict(
        kernel=KERNEL_WENDLAND_C2, drag_law=DRAG_ERGUN, visc_law=VISC_CONSTANT,
        precision=PREC_FP64, rho0=1000.0, rel_den=1.025, g=-9.81, nu=1e-6,
        csh=0.5, maxshift=0.1, lamfs=0.8, maxdivr=2.0, por=0.4, alpha=200.0,
        beta=1.1, d50=0.01, tol=1e-8, inner_tol=0.0, maxit=50000,
        inner_maxit=0, max_outer=200, max_nbr=64, shifting=1, adaptive_dt=0,
        sink=0, xlcutoff=-1e300, ytcutoff=1e300, sig1=0.2, mu=1e-3, salinity=0,
        cell_order=ORDER_CELL, reserved=0,
        rhomin=1000.0, tschmidt=1.0, salt_xbuf=1.6, xrelease_lo=1e300, xrelease_hi=1e300,
    )
    defaults.update(kw)
    for k, v in defaults.items():
        setattr(p, k, v)
    if p.lam == 0.0:
        p.lam = 1e-3 * p.h * p.h  # setup.f90:82
    if p.cell_size == 0.0:
        p.cell_size = 2.0 * p.h
    return p
