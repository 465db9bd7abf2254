"""The Fortran driver (iso_c_binding, main.f90 call order) and the ctypes
path drive the same library to bit-identical states."""
import os

import numpy as np
import pytest

from beach_slope_confined_amd import capi, configs, fortran_io

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("make", [lambda: configs.dam_break(0.01),
                                  lambda: configs.dam_break_reference(0.01)])
def test_fortran_driver_matches_ctypes(require_gpu, tmp_path, make):
    cfg = make()
    case, res = str(tmp_path / "case.bin"), str(tmp_path / "res.bin")
    fortran_io.write_case(case, cfg)
    out = fortran_io.run_driver(case, res, 3)
    assert "step      3" in out
    f = fortran_io.read_result(res)
    sim = configs.make_sim(cfg, capi.IsphSim)
    capi.run_step_stages(sim, 3)
    P = sim.get_particles()
    assert f["t"] == sim.scalars()["t"]
    for k in ("x", "y", "vx", "vy", "pressure", "pid"):
        assert np.array_equal(f[k], P[k]), k
