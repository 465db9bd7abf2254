"""The discrete adjoint at the size of the adjoint configuration: the porous
dam break at dx = 0.002 (30,000 fluid, 36,636 particles, 81,597 soil), 50
steps of dt = 1e-4 with shifting from the column at rest, checkpoints every
10 steps, objective = the time-mean Shepard pressure at 3 probes on the slope
face plus the water depth at x = 1.2 m and x = 0.3 m.  dJ/d(por, alpha,
beta) from the reverse sweep (through torch.autograd) against central
differences with a 1e-6 relative step, agreement 1e-5 relative.

The PPE is solved tightly in both runs so that the finite differences
resolve the drag coefficients (a 1e-6 step in alpha moves J by about 3e-9 of
itself): FP64 BiCGSTAB to 1e-13, and the multigrid-preconditioned solve to
1e-12.  Larger steps do not help: a 1e-5 or 1e-4 step in porosity crosses
the kinks of the porous closure (the clamp n >= por) and the FD of por is
then off by 1e-2."""
import numpy as np
import pytest
import torch

from beach_slope_confined_amd import adjoint, capi, configs
from beach_slope_confined_amd.params import PREC_FP64, PREC_MIXED_MG

pytestmark = pytest.mark.gpu

DX, NSTEPS, EVERY, SLOT = 0.002, 50, 10, 0


@pytest.mark.parametrize("precision,tol", [(PREC_MIXED_MG, 1e-12), (PREC_FP64, 1e-13)],
                         ids=["multigrid", "fp64"])
def test_c3_adjoint_matches_central_fd(require_gpu, precision, tol):
    cfg = configs.dam_break(DX)
    cfg["params"] = cfg["params"].copy(precision=precision, tol=tol, maxit=100000)
    sim = configs.make_sim(cfg, capi.IsphSim)
    assert len(cfg["parts"]["x"]) == 36636 and cfg["n_fluid"] == 30000
    sim.checkpoint(SLOT)
    obj = adjoint.slope_objective(gauges=(1.2, 0.3))
    theta0 = adjoint.theta_of(sim.params)
    th = torch.tensor(theta0, dtype=torch.float64, requires_grad=True)
    print(f"C3 adjoint, precision {precision}, tol {tol:g}", flush=True)
    J = adjoint.objective(th, sim, SLOT, obj, NSTEPS, EVERY)
    J.backward()
    g = th.grad.numpy().copy()
    g_fd = adjoint.central_fd(sim, SLOT, obj, theta0, NSTEPS, EVERY, rel=1e-6)
    err = np.abs(g - g_fd) / np.abs(g_fd)
    print(f"J {J.item():.12e} adjoint {g} fd {g_fd} rel.err {err} "
          f"transposed iterations {sim.last_adjoint_iters}", flush=True)
    assert np.all(np.abs(g) > 0)
    assert np.all(err <= 1e-5), err
    sim.close()
