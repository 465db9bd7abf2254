"""Case and result files of the Fortran driver (fortran/isph_driver.f90).

case:   isph_params (raw struct bytes), int32 n, int32 ns, float64 x, y, vx,
        vy, mass, density, oden, pressure, int32 tid, flags, pid, float64
        soil x, y, mass, density (little-endian, stream access)
result: int32 n, float64 t, float64 x, y, vx, vy, pressure, int32 pid
"""
import os
import subprocess

import numpy as np

DRIVER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fortran", "isph_driver")


def write_case(path, cfg):
    p, s = cfg["parts"], cfg["soil"]
    n, ns = len(p["x"]), len(s["x"])
    with open(path, "wb") as f:
        f.write(bytes(cfg["params"]))
        f.write(np.array([n, ns], "<i4").tobytes())
        for k in ("x", "y", "vx", "vy", "mass", "density", "oden", "pressure"):
            f.write(np.ascontiguousarray(p[k], "<f8").tobytes())
        for k in ("tid", "flags", "pid"):
            f.write(np.ascontiguousarray(p[k], "<i4").tobytes())
        for k in ("x", "y", "mass", "density"):
            f.write(np.ascontiguousarray(s[k], "<f8").tobytes())


def read_result(path):
    with open(path, "rb") as f:
        n = int(np.frombuffer(f.read(4), "<i4")[0])
        t = float(np.frombuffer(f.read(8), "<f8")[0])
        out = {"t": t}
        for k in ("x", "y", "vx", "vy", "pressure"):
            out[k] = np.frombuffer(f.read(8 * n), "<f8").copy()
        out["pid"] = np.frombuffer(f.read(4 * n), "<i4").copy()
    return out


def run_driver(case, result, nsteps, timeout=300):
    if not os.path.exists(DRIVER):
        raise FileNotFoundError(f"{DRIVER} is not built (make -C beach_slope_confined_amd/fortran)")
    return subprocess.run([DRIVER, case, result, str(nsteps)], check=True, capture_output=True,
                          text=True, timeout=timeout).stdout
