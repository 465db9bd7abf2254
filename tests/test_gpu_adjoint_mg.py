"""The discrete adjoint with the multigrid-preconditioned PPE: the forward
solves A p = b and the reverse sweep solves A^T lambda = pbar, both with
BiCGSTAB preconditioned by an FP32 V-cycle on their own Galerkin hierarchy
(A^T's coarse operators are the transposes of A's), to a tight FP64 true
residual.  The gradient dJ/d(por, alpha, beta) must equal the FP64
adjoint's to the solves' accuracy, and central differences of the
multigrid forward."""
import numpy as np
import pytest
import torch

from beach_slope_confined_amd import adjoint, capi, configs
from beach_slope_confined_amd.params import PREC_FP64, PREC_MIXED_MG

pytestmark = pytest.mark.gpu

DX = 0.02
SLOT = 0


def developed(precision, tol):
    """the dam break 40 steps in (a moving flow: the drag terms matter)"""
    cfg = configs.dam_break(DX)
    cfg["params"] = cfg["params"].copy(precision=precision, tol=tol, maxit=20000)
    sim = configs.make_sim(cfg, capi.IsphSim)
    sim.step(40)
    sim.checkpoint(SLOT)
    return sim


def gradient(sim, obj, nsteps, every):
    th = torch.tensor(adjoint.theta_of(sim.params), dtype=torch.float64, requires_grad=True)
    J = adjoint.objective(th, sim, SLOT, obj, nsteps, every)
    J.backward()
    return J.item(), th.grad.numpy().copy()


def test_multigrid_adjoint_equals_fp64_adjoint(require_gpu):
    obj = adjoint.slope_objective(gauges=(1.2, 0.3))
    s64 = developed(PREC_FP64, 1e-13)
    J64, g64 = gradient(s64, obj, 10, 5)
    it64 = s64.last_adjoint_iters
    s64.close()
    smg = developed(PREC_MIXED_MG, 1e-12)
    Jmg, gmg = gradient(smg, obj, 10, 5)
    itmg = smg.last_adjoint_iters
    smg.close()
    err = np.abs(gmg - g64) / np.abs(g64)
    print(f"J {J64:.12e} {Jmg:.12e}  grad fp64 {g64} mg {gmg} rel.err {err}  "
          f"adjoint iterations {it64} {itmg}", flush=True)
    assert abs(Jmg - J64) <= 1e-9 * abs(J64)
    assert np.all(err <= 1e-6), err
    assert itmg * 5 < it64


def test_multigrid_adjoint_matches_central_fd(require_gpu):
    obj = adjoint.slope_objective(gauges=(1.2, 0.3))
    s = developed(PREC_MIXED_MG, 1e-12)
    theta0 = adjoint.theta_of(s.params)
    J, g = gradient(s, obj, 7, 3)
    g_fd = adjoint.central_fd(s, SLOT, obj, theta0, 7, 3, rel=1e-4)
    err = np.abs(g - g_fd) / np.abs(g_fd)
    print(f"J {J:.12e} adjoint {g} fd {g_fd} rel.err {err}", flush=True)
    assert np.all(err <= 1e-5), err
    s.close()
