"""Gradients of probe pressure and water depth with respect to the porous
parameters (porosity, Darcy coefficient alpha, Forchheimer coefficient
beta), through the discrete adjoint of the gfx950 ISPH step.

The reference has no derivatives; its porous closure (``porous.f90:107-224``)
is where the three parameters enter.  ``PorousObjective`` wraps the
library's forward run (``isph_adjoint_forward``: steps with checkpoints) and
reverse sweep (``isph_adjoint_backward``: recompute + stage adjoints +
transposed PPE solves) as a ``torch.autograd.Function``; ``central_fd`` is
the finite-difference check it is validated against.
"""
import numpy as np
import torch

from . import capi


def slope_objective(toe=0.2, probe_x=(0.25, 0.30, 0.35), gauges=(1.2,)):
    """Probes on the face of the 1:3 beach y = (x - toe)/3 and depth gauges
    at the given x (the dam-break geometry of ``configs.dam_break``)."""
    probes = [(x, (x - toe) / 3.0) for x in probe_x]
    return capi.make_objective(probes, gauges)


def theta_of(params):
    return np.array([params.por, params.alpha, params.beta], dtype=np.float64)


def _run(sim, slot, obj, theta, nsteps, every):
    por, alpha, beta = (float(v) for v in theta)
    sim.set_params(sim.params.copy(por=por, alpha=alpha, beta=beta))
    sim.restore(slot)
    return sim.adjoint_forward(obj, nsteps, every)


class PorousObjective(torch.autograd.Function):
    """J(theta) for theta = (por, alpha, beta), float64 tensor of 3.

    forward: sets the parameters, restores the initial state from
    checkpoint `slot`, runs `nsteps` steps (checkpoints every `every`).
    backward: the library's reverse sweep; one call per forward."""

    @staticmethod
    def forward(ctx, theta, sim, slot, obj, nsteps, every):
        J = _run(sim, slot, obj, theta.detach().cpu().numpy(), nsteps, every)
        ctx.sim = sim
        return torch.tensor(J, dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad_out):
        g, _, iters = ctx.sim.adjoint_backward()
        ctx.sim.last_adjoint_iters = iters
        return (torch.from_numpy(g) * grad_out, None, None, None, None, None)


def objective(theta, sim, slot, obj, nsteps, every=10):
    return PorousObjective.apply(theta, sim, slot, obj, nsteps, every)


def central_fd(sim, slot, obj, theta, nsteps, every=10, rel=1e-6):
    """Central differences of J with a relative step `rel` per parameter."""
    theta = np.asarray(theta, dtype=np.float64)
    g = np.zeros(3)
    for k in range(3):
        h = rel * abs(theta[k]) if theta[k] != 0.0 else rel
        tp, tm = theta.copy(), theta.copy()
        tp[k] += h
        tm[k] -= h
        g[k] = (_run(sim, slot, obj, tp, nsteps, every) -
                _run(sim, slot, obj, tm, nsteps, every)) / (tp[k] - tm[k])
    _run(sim, slot, obj, theta, nsteps, every)  # leave the context at theta
    return g
