"""Parity of the gfx950 salinity transport with the CPU oracle.

The salt-water step (main.f90:199-263: transport substeps on the previous
step's lists, the usual step, densityupdate after the shift, eddyvis after
timestep) and its stages scalart / scalarupdate / densityupdate
(scalar.f90:114-305) must give bit-identical particle state, salinity and
substep counts on the GPU and in the oracle.
"""
import numpy as np
import pytest

from beach_slope_confined_amd import capi, configs
from oracle.pyoracle import OracleSim

from parity_util import FIELDS, diff_state, pair

pytestmark = pytest.mark.gpu

SALT_FIELDS = FIELDS + ("con", "ovol", "cdiff", "oden", "mass")
CASES = {
    "dam_break": lambda: configs.dam_break(0.01),
    "dam_break_reference": lambda: configs.dam_break_reference(0.01),
}
LINE = (0.2, 0.1, 0.3)   # xl, yl, line_grad of the salt wedge


def _salt_pair(case, **prm):
    gpu, ora = pair(CASES[case](), salinity=1, **prm)
    for s in (gpu, ora):
        s.salt_init(*LINE)
    assert diff_state(gpu, ora, SALT_FIELDS) == [], "after salt_init"
    return gpu, ora


@pytest.mark.parametrize("case", list(CASES))
def test_salt_steps_bitwise(require_gpu, case):
    gpu, ora = _salt_pair(case)
    for step in range(4):
        print(f"salt step {step} {case}", flush=True)
        gpu.step(1)
        ora.step(1)
        assert gpu.substeps() == ora.substeps(), f"step {step} substeps"
        assert gpu.scalars() == ora.scalars(), f"step {step} scalars"
        bad = diff_state(gpu, ora, SALT_FIELDS)
        assert bad == [], f"step {step}: differs in {bad}"
    con = gpu.get_field("con")[gpu.get_particles()["tid"] == 3]
    assert np.unique(con).size > 2, "no transport happened"


def test_salt_several_substeps_bitwise(require_gpu):
    """A small turbulent Schmidt number forces dtsol < dt and several
    transport substeps per step (integrator.f90:964-972)."""
    gpu, ora = _salt_pair("dam_break", tschmidt=1e-6)
    seen = 0
    for step in range(4):
        gpu.step(1)
        ora.step(1)
        k = gpu.substeps()
        assert k == ora.substeps(), f"step {step} substeps"
        seen = max(seen, k[0])
        bad = diff_state(gpu, ora, SALT_FIELDS)
        assert bad == [], f"step {step}: differs in {bad}"
    assert seen > 1


@pytest.mark.parametrize("case", list(CASES))
def test_salt_stages_bitwise(require_gpu, case):
    gpu, ora = _salt_pair(case)
    for s in (gpu, ora):
        s.step(2)
    rough = None
    for name in ("scalart", "scalarupdate", "densityupdate", "scalart", "eddyvis_lagged",
                 "eddyvis"):
        for s in (gpu, ora):
            if name == "scalarupdate":
                s.scalarupdate(3.7e-3)
            elif name == "scalart" and rough is None:
                x = s.get_field("x")
                s.set_field("con", s.get_field("con") + 0.2 * np.sin(50 * x))
                s.scalart()
            else:
                getattr(s, name)()
        rough = True
        bad = diff_state(gpu, ora, SALT_FIELDS)
        assert bad == [], f"stage {name}: differs in {bad}"


def test_salt_transport_full_size(require_gpu):
    """scalart / scalarupdate / densityupdate on the 1024 x 1024 tank
    (1,057,810 particles) with pores in the lower half, patches of eddy
    viscosity on both sides of the 1e-6 threshold and a rough salinity
    field: bit-identical with the oracle."""
    cfg = configs.tank_1m(1024)
    cfg["params"] = cfg["params"].copy(salinity=1)
    n = len(cfg["parts"]["x"])
    sims = [capi.IsphSim(cfg["params"], n, 1), OracleSim(cfg["params"], n, 1)]
    for s in sims:
        print(f"full-size setup {type(s).__name__}", flush=True)
        s.set_particles(cfg["parts"])
        s.neighbour()
        q = s.get_particles()
        x, y = q["x"], q["y"]
        s.set_field("porosity", np.where(y < 0.5, 0.4 + 0.2 * x, 1.0))
        s.set_field("nut", np.where(np.sin(9 * x) > 0, 3e-6, 4e-7) * (x > 0.3))
        s.set_field("con", np.clip(y - 0.5 + 0.1 * np.sin(30 * x), -0.5, 0.5))
        # the shift of a step moves particles after the search
        s.set_field("x", x + 1e-5 * np.sin(97 * y))
    for name in ("scalart", "scalarupdate", "densityupdate"):
        for s in sims:
            print(f"full-size {name} {type(s).__name__}", flush=True)
            if name == "scalarupdate":
                s.scalarupdate(1e-4)
            else:
                getattr(s, name)()
        for f in ("cdiff", "con", "oden", "mass", "density"):
            a, b = sims[0].get_field(f), sims[1].get_field(f)
            assert np.array_equal(a, b), f"{name}: {f} differs"
    cd = sims[0].get_field("cdiff")
    assert np.count_nonzero(cd) > n // 2
