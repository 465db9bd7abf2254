"""Limits and error codes, forced from the host.

- Column offsets beyond 16 bits: the solvers read 16-bit offsets j - i when
  every offset of the lists fits, else the int32 columns.  A tall water
  column (12 x 7000 particles, 4h cells) puts neighbours 35,000 rows apart,
  and a whole step through the int32 path must stay bit-identical to the
  oracle, for the FP64 and the mixed-precision solve.
- Neighbour-list capacity: a list longer than max_nbr is ISPH_E_CAPACITY on
  the library and on the oracle alike (the reference leaves it unchecked,
  memory.f90:25-27).
- Non-finite velocity: comp_vel reports ISPH_E_NAN.  The reference and its
  restatement carry the NaN on silently; the library checks it.
"""
import re

import numpy as np
import pytest

from beach_slope_confined_amd import capi, configs
from beach_slope_confined_amd.configs import _assemble, _walls, uniform
from beach_slope_confined_amd.params import (DRAG_ERGUN, KERNEL_WENDLAND_C2, PREC_FP64,
                                             PREC_MIXED_V64, VISC_CONSTANT, make_params)
from oracle.pyoracle import OracleSim

from parity_util import STAGES, diff_state, pair, run_stage

pytestmark = pytest.mark.gpu


def tall_column(nx=12, ny=7000, dx=1e-3, seed=7):
    ii, jj = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    ii = ii.ravel().astype(np.float64)
    jj = jj.ravel().astype(np.float64)
    nf = len(ii)
    fx = (ii + 0.5) * dx + (2.0 * uniform(nf, 1, seed) - 1.0) * 0.05 * dx
    fy = (jj + 0.5) * dx + (2.0 * uniform(nf, 2, seed) - 1.0) * 0.05 * dx
    wx, wy, wtid = _walls(nx, ny, dx, 0.0, 0.0, top=False)
    H = ny * dx
    fp = 1000.0 * 9.81 * np.maximum(H - fy, 0.0)
    parts = _assemble(fx, fy, np.zeros(nf), np.zeros(nf), wx, wy, wtid, 1000.0, dx, fp,
                      np.zeros(len(wx)))
    m = 3.5 * dx
    prm = make_params(kernel=KERNEL_WENDLAND_C2, drag_law=DRAG_ERGUN, visc_law=VISC_CONSTANT,
                      h=1.3 * dx, dl=dx, dt=1e-4, cell_size=5.2 * dx, xmin=-m, ymin=-m,
                      xmax=nx * dx + m, ymax=H + 10 * dx, tol=1e-2, maxit=200000)
    soil = dict(x=np.zeros(0), y=np.zeros(0), mass=np.zeros(0), density=np.zeros(0))
    return dict(name="tall_column", parts=parts, soil=soil, params=prm, n_fluid=nf)


@pytest.mark.parametrize("precision", [PREC_FP64, PREC_MIXED_V64])
def test_int32_columns_step_bitwise(require_gpu, precision):
    gpu, ora = pair(tall_column(), precision=precision)
    for name in STAGES:
        print(f"tall column precision {precision}: {name}", flush=True)
        rg = run_stage(gpu, name)
        ro = run_stage(ora, name)
        if name == "neighbour":
            ptr, idx = gpu.get_neighbours()
            rows = np.repeat(np.arange(len(ptr) - 1), np.diff(ptr))
            assert np.abs(idx.astype(np.int64) - rows).max() > 32767
        if name == "ppesolve":
            assert rg == ro, f"iterations/relres {rg} vs {ro}"
        bad = diff_state(gpu, ora)
        assert bad == [], f"{name}: differs in {bad}"
    gpu.close()
    ora.close()


def code(err):
    m = re.search(r"ISPH_E_\w+", str(err))
    return m.group(0) if m else str(err)


def test_neighbour_capacity_overflow_is_an_error(require_gpu):
    cfg = configs.dam_break(0.01)
    cfg["params"] = cfg["params"].copy(max_nbr=8)
    got = []
    for cls in (capi.IsphSim, OracleSim):
        sim = configs.make_sim(cfg, cls)
        with pytest.raises((capi.IsphError, RuntimeError)) as e:
            sim.neighbour()
        got.append(code(e.value))
        sim.close()
    assert got == ["ISPH_E_CAPACITY", "ISPH_E_CAPACITY"]


def test_non_finite_velocity_is_reported(require_gpu):
    gpu, ora = pair(configs.dam_break(0.01))
    for name in STAGES[:STAGES.index("ppesolve") + 1]:
        run_stage(gpu, name)
        run_stage(ora, name)
    p = gpu.get_particles()
    k = int(np.flatnonzero(p["tid"] == 3)[len(np.flatnonzero(p["tid"] == 3)) // 2])
    for sim in (gpu, ora):
        pr = sim.get_field("pressure")
        pr[k] = np.nan
        sim.set_field("pressure", pr)
    with pytest.raises(capi.IsphError, match="ISPH_E_NAN"):
        gpu.comp_vel()
    ora.comp_vel()  # the restatement, like the reference, carries the NaN on
    assert np.isnan(ora.get_particles()["vx"]).any()
    gpu.close()
    ora.close()
