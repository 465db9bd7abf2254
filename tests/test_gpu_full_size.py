"""Full-size cases of the configurations:

- the mixed-precision PPE with FP64 Krylov vectors (ISPH_PREC_MIXED_V64) on
  the 1024 x 1024 tank (1,057,810 particles), one whole step on the GPU and
  on the oracle, bit for bit, with its pressure error against the FP64 step;
- the small porous dam break (dx = 0.01, 2,556 particles) for its full 200
  steps, GPU against oracle bit for bit;
- the large beach (8192 x 512 column, 4,261,924 particles), one multigrid
  step: FP64 true residual on the host, pressure against a tight FP64
  multigrid solve.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from beach_slope_confined_amd import capi, configs
from beach_slope_confined_amd.params import PREC_FP64_MG, PREC_MIXED_MG, PREC_MIXED_V64

from parity_util import diff_state, pair

pytestmark = pytest.mark.gpu


def test_mixed_v64_full_size_step_bitwise(require_gpu):
    from oracle import pyoracle
    pyoracle.set_threads(16)
    cfg = configs.tank_1m(1024)
    gpu, ora = pair(cfg, precision=PREC_MIXED_V64)
    print("1M tank, mixed precision (FP64 vectors): GPU step", flush=True)
    gpu.step(1)
    print("oracle step", flush=True)
    ora.step(1)
    sg, so = gpu.scalars(), ora.scalars()
    print(f"iterations {sg['iters']} / {so['iters']}, reliable updates {sg['outer']}", flush=True)
    assert sg["iters"] == so["iters"] and sg["outer"] == so["outer"]
    assert diff_state(gpu, ora) == []
    pm = gpu.get_particles()
    gpu.close()
    ora.close()
    f64 = configs.make_sim(cfg, capi.IsphSim)
    f64.step(1)
    p64 = f64.get_particles()
    f64.close()
    pm_p = pm["pressure"][np.argsort(pm["pid"])]
    p64_p = p64["pressure"][np.argsort(p64["pid"])]
    err = np.abs(pm_p - p64_p).max() / np.abs(p64_p).max()
    print(f"pressure error against the FP64 step {err:.2e}", flush=True)
    assert err < 1e-6


def test_small_dam_break_200_steps_bitwise(require_gpu):
    gpu, ora = pair(configs.dam_break(0.01))
    for k in range(4):
        gpu.step(50)
        ora.step(50)
        print(f"dam break: {50 * (k + 1)} steps, t = {gpu.scalars()['t']:.4f}", flush=True)
        assert gpu.scalars() == ora.scalars()
        assert diff_state(gpu, ora) == [], f"after {50 * (k + 1)} steps"


def _large_solve(prec, tol):
    cfg = configs.beach_wide(dt=1.25e-5, jitter=0.0)
    cfg["params"] = cfg["params"].copy(precision=prec, tol=tol, maxit=40000)
    sim = configs.make_sim(cfg, capi.IsphSim)
    s = sim.scalars()
    sim.set_scalars(s["umax"], s["dt"], s["t"] + s["dt"])
    for st in ("projection", "cellshift", "resetid", "neighbour", "effpor", "eddyvis", "int_vel"):
        getattr(sim, st)()
    sim.ppe_assemble()
    it, rel = sim.ppe_solve()
    return sim, it, rel


def test_large_beach_multigrid_solve(require_gpu):
    print("large beach: mixed multigrid solve", flush=True)
    sim, it, rel = _large_solve(PREC_MIXED_MG, 1e-9)
    assert sim.count() == 4261924
    x = sim.get_field("pressure")
    rp, col, val, rhs = sim.get_csr()
    sim.close()
    A = sp.csr_matrix((val, col, rp), shape=(len(rhs), len(rhs)))
    true_rel = np.linalg.norm(rhs - A @ x) / np.linalg.norm(rhs)
    del A, col, val
    print("large beach: tight FP64 multigrid solve", flush=True)
    tight, it_t, _ = _large_solve(PREC_FP64_MG, 1e-11)
    xt = tight.get_field("pressure")
    tight.close()
    err = np.abs(x - xt).max() / np.abs(xt).max()
    print(f"mixed multigrid {it} iterations, true residual {true_rel:.2e}, error against the "
          f"tight solve ({it_t} iterations) {err:.2e}", flush=True)
    assert true_rel <= 1e-9 * (1 + 1e-6)
    assert err < 1e-6
    assert it < 120
