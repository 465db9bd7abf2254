"""Two PPE forms that isolate what FP32 buys:

- ISPH_PREC_FP64_MG, the multigrid-preconditioned BiCGSTAB of
  ISPH_PREC_MIXED_MG with every value FP64 (FP64 matrix in the outer
  products, FP64 V-cycle on FP64 Galerkin operators).  It has no oracle
  counterpart; it must meet the FP64 true residual, agree with a tight FP64
  solve, and take about as many iterations as the mixed form (the same
  algorithm, so the iteration counts differ only by rounding).
- ISPH_PREC_IR, iterative refinement with an FP32 inner BiCGSTAB (FP32
  values and vectors, FP64 dots) restarted from zero on every FP64 true
  residual: the oracle restates it, so GPU and oracle agree bit for bit.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from beach_slope_confined_amd import capi, configs
from beach_slope_confined_amd.params import PREC_FP64, PREC_FP64_MG, PREC_IR, PREC_MIXED_MG

from parity_util import diff_state, pair

pytestmark = pytest.mark.gpu

PRE = ("projection", "cellshift", "resetid", "neighbour", "effpor", "eddyvis", "int_vel")


def _to_solve(cfg, precision, tol):
    c = dict(cfg)
    c["params"] = cfg["params"].copy(precision=precision, tol=tol, maxit=200000)
    sim = configs.make_sim(c, capi.IsphSim)
    s = sim.scalars()
    sim.set_scalars(s["umax"], s["dt"], s["t"] + s["dt"])
    for st in PRE:
        getattr(sim, st)()
    sim.ppe_assemble()
    return sim


@pytest.mark.parametrize("case", ["tank_256", "dam_break"])
def test_fp64_multigrid_solve(require_gpu, case):
    cfg = configs.tank_1m(256) if case == "tank_256" else configs.dam_break(0.004)
    tol = 1e-9
    out = {}
    for name, prec, t in (("fp64_mg", PREC_FP64_MG, tol), ("mixed_mg", PREC_MIXED_MG, tol),
                          ("tight", PREC_FP64, 1e-13)):
        print(f"{case}: {name} solve", flush=True)
        sim = _to_solve(cfg, prec, t)
        it, rel = sim.ppe_solve()
        out[name] = (it, sim.get_field("pressure"), sim.get_csr())
        sim.close()
    it64, x64, (rp, col, val, rhs) = out["fp64_mg"]
    A = sp.csr_matrix((val, col, rp), shape=(len(rhs), len(rhs)))
    true_rel = np.linalg.norm(rhs - A @ x64) / np.linalg.norm(rhs)
    x_ex = out["tight"][1]
    err64 = np.abs(x64 - x_ex).max() / np.abs(x_ex).max()
    itmx, xmx = out["mixed_mg"][0], out["mixed_mg"][1]
    errmx = np.abs(xmx - x_ex).max() / np.abs(x_ex).max()
    print(f"{case}: FP64 multigrid {it64} iterations, true residual {true_rel:.2e}, error "
          f"{err64:.2e}; mixed multigrid {itmx} iterations, error {errmx:.2e}; tight FP64 "
          f"{out['tight'][0]}", flush=True)
    assert true_rel <= tol * (1 + 1e-6)
    assert err64 < 1e-6
    assert abs(it64 - itmx) <= max(3, itmx // 4)
    assert it64 * 5 < out["tight"][0]


@pytest.mark.parametrize("case", ["tank_64", "dam_break"])
def test_iterative_refinement_bitwise(require_gpu, case):
    cfg = configs.tank_1m(64) if case == "tank_64" else configs.dam_break(0.01)
    gpu, ora = pair(cfg, precision=PREC_IR, max_outer=1000)
    for step in range(2):
        print(f"{case}: IR step {step + 1}", flush=True)
        gpu.step(1)
        ora.step(1)
        sg, so = gpu.scalars(), ora.scalars()
        assert sg["iters"] == so["iters"] and sg["outer"] == so["outer"], (sg, so)
        assert sg["relres"] <= 1e-8
        assert diff_state(gpu, ora) == [], f"step {step + 1}"
    # restarts: more corrections than the reliable-update form needs
    assert gpu.scalars()["outer"] >= 4
