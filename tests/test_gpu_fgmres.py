"""The reference's live PPE solve (precision = PREC_FGMRES: restarted GMRES
as solver.f90:306-347 drives MKL's dfgmres) on the gfx950 library:
bit-identical with the oracle's restatement in the reference's settings
(tol 1e-3, at most 150 iterations, warm start from fsol), across restart
cycles, at the iteration limit (the pressure keeps the warm start), and
through a checkpoint."""
import numpy as np
import pytest

from beach_slope_confined_amd import configs
from beach_slope_confined_amd.params import PREC_FGMRES

from parity_util import FIELDS, STAGES, diff_state, pair, run_stage

pytestmark = pytest.mark.gpu

REFERENCE = dict(precision=PREC_FGMRES, tol=1e-3, maxit=150, inner_maxit=0)


def test_fgmres_steps_bitwise(require_gpu):
    """full steps in the reference's settings: the warm start of every solve
    is the previous solution by row"""
    gpu, ora = pair(configs.dam_break(0.01), **REFERENCE)
    assert diff_state(gpu, ora) == []
    for step in range(4):
        print(f"fgmres step {step}", flush=True)
        gpu.step(1)
        ora.step(1)
        bad = diff_state(gpu, ora, FIELDS + ("pressure",))
        assert bad == [], f"step {step}: differs in {bad}"
        assert gpu.scalars() == ora.scalars(), f"step {step}"
        assert 0 < gpu.scalars()["iters"] <= 150


def test_fgmres_stage_by_stage_reference_order(require_gpu):
    cfg = configs.with_reference_cells(configs.dam_break_reference(0.01))
    gpu, ora = pair(cfg, **REFERENCE)
    for step in range(2):
        for name in STAGES:
            print(f"fgmres reference order step {step} {name}", flush=True)
            rg = run_stage(gpu, name)
            ro = run_stage(ora, name)
            if name == "ppesolve":
                assert rg == ro, f"step {step}: {rg} vs {ro}"
            bad = diff_state(gpu, ora)
            assert bad == [], f"step {step} stage {name}: differs in {bad}"


@pytest.mark.parametrize("restart,maxit,tol", [(10, 600, 1e-4), (150, 6, 1e-12)])
def test_fgmres_restart_cycles_and_iteration_limit(require_gpu, restart, maxit, tol):
    """restart cycles of 10 columns to 1e-4 (13 cycles), and a limit of 6
    iterations that leaves the pressure at the warm start (the uploaded
    pressures)"""
    gpu, ora = pair(configs.dam_break(0.01), precision=PREC_FGMRES, tol=tol, maxit=maxit,
                    inner_maxit=restart)
    p0 = gpu.get_particles()["pressure"]
    for name in STAGES[:STAGES.index("ppesolve") + 1]:
        print(f"fgmres restart {restart} maxit {maxit} {name}", flush=True)
        rg = run_stage(gpu, name)
        ro = run_stage(ora, name)
    assert rg == ro
    assert diff_state(gpu, ora, FIELDS + ("pressure",)) == []
    it, rel = rg
    if maxit == 6:
        assert it == 6 and rel > tol
        # rows are particles in the first row order here (no particle has
        # changed cell), so fsol is the uploaded pressure particle by particle
        assert np.array_equal(gpu.get_particles()["pressure"], p0)
    else:
        assert it > restart and rel <= tol


def test_fgmres_checkpoint_restores_the_warm_start(require_gpu):
    cfg = configs.dam_break(0.01)
    cfg["params"] = cfg["params"].copy(**REFERENCE)
    from beach_slope_confined_amd import capi
    sim = configs.make_sim(cfg, capi.IsphSim)
    sim.step(2)
    sim.checkpoint(0)
    sim.step(2)
    a = sim.get_particles()
    sim.restore(0)
    sim.step(2)
    b = sim.get_particles()
    for k in ("x", "y", "vx", "vy", "pressure"):
        assert np.array_equal(a[k], b[k]), k
