"""ctypes wrapper of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker of the gfx950 library.  It
exposes the same methods as beach_slope_confined_amd.capi.IsphSim.
"""
import ctypes
import os

from beach_slope_confined_amd.capi import Sim, declare

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
_LIB = None


def load():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} missing: run `make -C oracle`")
        lib = ctypes.CDLL(LIB_PATH)
        declare(lib, "ora_")
        lib.ora_set_threads.argtypes = [ctypes.c_int32]
        lib.ora_set_threads.restype = ctypes.c_int32
        lib.ora_kernel_w.argtypes = [ctypes.c_void_p, ctypes.c_double]
        lib.ora_kernel_w.restype = ctypes.c_double
        lib.ora_kernel_w0.argtypes = [ctypes.c_void_p]
        lib.ora_kernel_w0.restype = ctypes.c_double
        dp = ctypes.POINTER(ctypes.c_double)
        lib.ora_kernel_gradw.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_double, dp, dp]
        lib.ora_kernel_gradw.restype = None
        lib.ora_invert2.argtypes = [dp]
        lib.ora_invert2.restype = None
        lib.ora_dot_tree.argtypes = [dp, dp, ctypes.c_int64]
        lib.ora_dot_tree.restype = ctypes.c_double
        ip = ctypes.POINTER(ctypes.c_int32)
        lib.ora_bicgstab_csr.argtypes = [ctypes.c_int32, ip, ip, dp, dp, dp, ctypes.c_double,
                                         ctypes.c_int32, ip, dp]
        lib.ora_bicgstab_csr.restype = ctypes.c_int
        lib.ora_fgmres_csr.argtypes = [ctypes.c_int32, ip, ip, dp, dp, dp, ctypes.c_double,
                                       ctypes.c_int32, ctypes.c_int32, ip, dp, ip]
        lib.ora_fgmres_csr.restype = ctypes.c_int
        _LIB = lib
    return _LIB


def set_threads(n):
    return load().ora_set_threads(int(n))


class OracleSim(Sim):
    def __init__(self, params, cap, cap_soil=0):
        super().__init__(load(), "ora_", params, cap, cap_soil)


def kernel_w(params, r):
    return load().ora_kernel_w(ctypes.byref(params), r)


def kernel_w0(params):
    return load().ora_kernel_w0(ctypes.byref(params))


def kernel_gradw(params, r, dx, dy):
    wx, wy = ctypes.c_double(), ctypes.c_double()
    load().ora_kernel_gradw(ctypes.byref(params), r, dx, dy, ctypes.byref(wx), ctypes.byref(wy))
    return wx.value, wy.value


def invert2(a):
    import numpy as np
    arr = np.ascontiguousarray(a, dtype=np.float64).copy()
    load().ora_invert2(arr.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return arr


def dot_tree(a, b):
    import numpy as np
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    dp = ctypes.POINTER(ctypes.c_double)
    return load().ora_dot_tree(a.ctypes.data_as(dp), b.ctypes.data_as(dp), len(a))


def bicgstab_csr(rowptr, col, val, b, x0, tol=1e-10, maxit=1000):
    import numpy as np
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int32)
    rowptr = np.ascontiguousarray(rowptr, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    val = np.ascontiguousarray(val, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    x = np.ascontiguousarray(x0, np.float64).copy()
    it = ctypes.c_int32()
    rel = ctypes.c_double()
    rc = load().ora_bicgstab_csr(len(b), rowptr.ctypes.data_as(ip), col.ctypes.data_as(ip),
                                 val.ctypes.data_as(dp), b.ctypes.data_as(dp),
                                 x.ctypes.data_as(dp), tol, maxit, ctypes.byref(it),
                                 ctypes.byref(rel))
    return rc, x, it.value, rel.value


def fgmres_csr(rowptr, col, val, b, x0, tol, maxit, restart=0):
    """The oracle's GMRES of the reference's live solve on a CSR matrix:
    returns (rc, x, iterations, residual estimate / ||r0||, converged).  x is
    the warm start x0 where the iteration limit came first within the first
    restart cycle."""
    import numpy as np
    ip = ctypes.POINTER(ctypes.c_int32)
    dp = ctypes.POINTER(ctypes.c_double)
    rowptr = np.ascontiguousarray(rowptr, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    val = np.ascontiguousarray(val, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    x = np.ascontiguousarray(x0, np.float64).copy()
    it = ctypes.c_int32()
    rel = ctypes.c_double()
    conv = ctypes.c_int32()
    rc = load().ora_fgmres_csr(len(b), rowptr.ctypes.data_as(ip), col.ctypes.data_as(ip),
                               val.ctypes.data_as(dp), b.ctypes.data_as(dp), x.ctypes.data_as(dp),
                               tol, maxit, restart, ctypes.byref(it), ctypes.byref(rel),
                               ctypes.byref(conv))
    return rc, x, it.value, rel.value, bool(conv.value)
