"""Helpers for comparing the gfx950 library with the CPU oracle."""
import numpy as np

from beach_slope_confined_amd import capi, configs
from oracle.pyoracle import OracleSim

PART_KEYS = ("x", "y", "vx", "vy", "mass", "density", "oden", "pressure", "tid", "flags", "pid")
FIELDS = ("xs", "ys", "vxs", "vys", "porosity", "resist", "coff", "divr", "nut", "tke",
          "varts", "vel")

STAGES = ["projection", "cellshift", "resetid", "neighbour", "effpor", "eddyvis", "int_vel",
          "ppesolve", "comp_vel", "comp_pos", "cellshiftalt", "neighbour", "freesurf",
          "effpor", "opt2_shift", "timestep"]


def pair(cfg, **prm_changes):
    if prm_changes:
        cfg = dict(cfg)
        cfg["params"] = cfg["params"].copy(**prm_changes)
    gpu = configs.make_sim(cfg, capi.IsphSim)
    ora = configs.make_sim(cfg, OracleSim)
    return gpu, ora


def diff_state(a, b, fields=FIELDS):
    """Names of state arrays that differ bitwise (NaN == NaN)."""
    out = []
    pa, pb = a.get_particles(), b.get_particles()
    for k in PART_KEYS:
        if pa[k].shape != pb[k].shape or not np.array_equal(pa[k], pb[k], equal_nan=True):
            out.append(k)
    for f in fields:
        fa, fb = a.get_field(f), b.get_field(f)
        if not np.array_equal(fa, fb, equal_nan=True):
            out.append(f)
    return out


def max_rel(a, b):
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    s = np.maximum(np.abs(b).max(), 1e-300)
    return float(np.abs(a - b).max() / s) if len(a) else 0.0


def run_stage(sim, name):
    if name == "ppesolve":
        return sim.ppesolve()
    if name == "projection":
        s = sim.scalars()
        sim.set_scalars(s["umax"], s["dt"], s["t"] + s["dt"])
    return getattr(sim, name)()
