"""Discrete adjoint of the ISPH step on the gfx950 library.

Stage tests: for each stage S with inputs u and outputs y = S(u), random
output adjoints ybar and a random input direction d must satisfy the
dot-product identity <ybar, (S(u + eps d) - S(u - eps d)) / (2 eps)> =
<S^T ybar, d>, where S^T ybar is what the library's stage adjoint returns.
Whole run: dJ/d(por, alpha, beta) from the reverse sweep (through
torch.autograd) against central differences: a 1e-6 relative step to 1e-5
relative, and a 1e-4 step to 1e-6.
"""
import numpy as np
import pytest
import torch

from beach_slope_confined_amd import adjoint, capi, configs

pytestmark = pytest.mark.gpu

DX = 0.02
TOL = 1e-13      # PPE tolerance of the finite-difference runs
SLOTS = {"start": 0, "pre_asm": 1, "post_vel": 2, "pre_shift": 3}
INIT_SLOT = 10


def make_case(steps=3):
    cfg = configs.dam_break(DX)
    cfg["params"] = cfg["params"].copy(tol=TOL, maxit=20000)
    sim = configs.make_sim(cfg, capi.IsphSim)
    sim.checkpoint(INIT_SLOT)
    sim.step(steps)  # let the column start to move
    # states at the inputs of the stages of one step (main.f90:41-66)
    sim.checkpoint(SLOTS["start"])
    s = sim.scalars()
    sim.set_scalars(s["umax"], s["dt"], s["t"] + s["dt"])
    for name in ("projection", "cellshift", "resetid", "neighbour", "effpor", "eddyvis", "int_vel"):
        getattr(sim, name)()
    sim.checkpoint(SLOTS["pre_asm"])
    sim.ppesolve()
    sim.comp_vel()
    sim.checkpoint(SLOTS["post_vel"])
    for name in ("comp_pos", "cellshiftalt", "neighbour", "freesurf", "effpor"):
        getattr(sim, name)()
    sim.checkpoint(SLOTS["pre_shift"])
    return sim


@pytest.fixture(scope="module")
def sim(require_gpu):
    s = make_case()
    yield s
    s.close()


# stage: (slot, rebuild lists, forward calls, adjoint stages, inputs, outputs)
CASES = {
    "projection": ("start", False, ["projection"], ["projection"],
                   ["x", "y", "vx", "vy", "porosity"], ["x", "y", "xs", "ys"]),
    "neighbour": ("pre_asm", True, ["neighbour"], ["neighbour"],
                  ["x", "y", "density"], ["coff"]),
    "effpor": ("pre_asm", False, ["effpor"], ["effpor"],
               ["x", "y", "vx", "vy"], ["porosity", "density", "resist"]),
    "eddyvis": ("pre_asm", True, ["eddyvis"], ["eddyvis"],
                ["x", "y", "vx", "vy", "porosity", "density", "coff"], ["nut", "tke"]),
    # int_vel = varts then v*; its adjoint is v* then varts
    "int_vel": ("pre_asm", True, ["int_vel"], ["vstar", "varts"],
                ["x", "y", "vx", "vy", "porosity", "density", "resist", "nut", "tke", "coff"],
                ["vxs", "vys", "varts"]),
    "ppesolve": ("pre_asm", True, ["ppesolve"], ["ppesolve"],
                 ["x", "y", "vxs", "vys", "porosity", "density", "coff"], ["pressure"]),
    "comp_vel": ("post_vel", True, ["comp_vel"], ["comp_vel"],
                 ["x", "y", "vxs", "vys", "porosity", "density", "pressure", "coff"],
                 ["vx", "vy", "vel"]),
    "comp_pos": ("post_vel", False, ["comp_pos"], ["comp_pos"],
                 ["x", "y", "xs", "ys", "vx", "vy", "porosity"], ["x", "y"]),
    "shift": ("pre_shift", True, ["opt2_shift"], ["shift"],
              ["x", "y", "vx", "vy", "density", "coff"], ["x", "y", "vx", "vy"]),
}
RTOL = {"ppesolve": 1e-4}


def prep(sim, slot, lists):
    sim.restore(SLOTS[slot])
    if lists:
        sim.neighbour()


def scale_of(sim, name, v):
    if name in ("x", "y", "xs", "ys"):
        return sim.params.dl
    s = float(np.sqrt(np.mean(v * v))) if v.size else 0.0
    return s if s > 0.0 else 1.0


def forward_dot(sim, case, base, d, eps, ybar):
    slot, lists, fwd, _, inp, out = case
    prep(sim, slot, lists)
    for f in inp:
        sim.set_field(f, base[f] + eps * d[f])
    for name in fwd:
        if name == "ppesolve":
            sim.ppesolve()
        else:
            getattr(sim, name)()
    return sum(float(np.dot(ybar[f].ravel(), sim.get_field(f).ravel())) for f in out)


@pytest.mark.parametrize("stage", list(CASES))
def test_stage_adjoint_dot_product(sim, stage):
    case = CASES[stage]
    slot, lists, fwd, adj, inp, out = case
    rng = np.random.default_rng(20261202)
    prep(sim, slot, lists)
    base = {f: sim.get_field(f).copy() for f in inp}
    if stage == "ppesolve":
        sim.ppesolve()  # the adjoint uses the solved pressure
    ybar = {f: rng.standard_normal(sim.get_field(f).shape) for f in out}
    sim.adjoint_zero()
    for f in out:
        sim.adjoint_set(f, ybar[f])
    for a in adj:
        sim.adjoint_stage(a)
    ubar = {f: sim.adjoint_get(f) for f in inp}
    d = {f: rng.standard_normal(base[f].shape) * scale_of(sim, f, base[f]) for f in inp}
    eps = 1e-6 if stage != "ppesolve" else 1e-5
    fd = (forward_dot(sim, case, base, d, eps, ybar) -
          forward_dot(sim, case, base, d, -eps, ybar)) / (2.0 * eps)
    ad = sum(float(np.dot(ubar[f].ravel(), d[f].ravel())) for f in inp)
    rtol = RTOL.get(stage, 1e-6)
    assert abs(fd - ad) <= rtol * max(abs(fd), abs(ad)), (stage, fd, ad)


def test_effpor_parameter_adjoint(sim):
    """dJ/d(por, alpha, beta) of the porous closure (porous.f90:107-224)."""
    rng = np.random.default_rng(7)
    prep(sim, "pre_asm", False)
    out = ["porosity", "density", "resist"]
    ybar = {f: rng.standard_normal(sim.get_field(f).shape) for f in out}
    sim.adjoint_zero()
    for f in out:
        sim.adjoint_set(f, ybar[f])
    sim.adjoint_stage("effpor")
    g, _ = sim.adjoint_params()
    p0 = sim.params.copy()
    for k, name in enumerate(("por", "alpha", "beta")):
        h = 1e-6 * abs(getattr(p0, name))
        vals = []
        for sgn in (1.0, -1.0):
            sim.set_params(p0.copy(**{name: getattr(p0, name) + sgn * h}))
            prep(sim, "pre_asm", False)
            sim.effpor()
            vals.append(sum(float(np.dot(ybar[f], sim.get_field(f))) for f in out))
        sim.set_params(p0)
        fd = (vals[0] - vals[1]) / (2.0 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(abs(fd), abs(g[k])), (name, fd, g[k])


def test_varts_umax_adjoint(sim):
    """varts depends on the previous step's umax (integrator.f90:495)."""
    rng = np.random.default_rng(11)
    prep(sim, "pre_asm", True)
    ybar = rng.standard_normal(sim.get_field("varts").shape)
    sim.adjoint_zero()
    sim.adjoint_set("varts", ybar)
    sim.adjoint_stage("varts")
    _, ub = sim.adjoint_params()
    s = sim.scalars()
    h = 1e-6 * s["umax"]
    vals = []
    for sgn in (1.0, -1.0):
        prep(sim, "pre_asm", True)
        sim.set_scalars(s["umax"] + sgn * h, s["dt"], s["t"])
        sim.int_vel()
        vals.append(float(np.dot(ybar, sim.get_field("varts"))))
    fd = (vals[0] - vals[1]) / (2.0 * h)
    assert np.count_nonzero(ybar) and abs(fd - ub) <= 1e-6 * max(abs(fd), abs(ub)), (fd, ub)


def test_objective_adjoint(sim):
    obj = adjoint.slope_objective(gauges=(1.2, 0.3))
    rng = np.random.default_rng(3)
    inp = ["x", "y", "density", "pressure"]
    prep(sim, "start", False)
    base = {f: sim.get_field(f).copy() for f in inp}
    J0 = sim.objective_eval(obj)
    assert J0 > 0.0
    sim.adjoint_zero()
    sim.adjoint_stage("objective", obj)
    ubar = {f: sim.adjoint_get(f) for f in inp}
    d = {f: rng.standard_normal(base[f].shape) * scale_of(sim, f, base[f]) for f in inp}
    eps = 1e-6
    vals = []
    for sgn in (1.0, -1.0):
        prep(sim, "start", False)
        for f in inp:
            sim.set_field(f, base[f] + sgn * eps * d[f])
        vals.append(sim.objective_eval(obj))
    fd = (vals[0] - vals[1]) / (2.0 * eps)
    ad = sum(float(np.dot(ubar[f], d[f])) for f in inp)
    assert abs(fd - ad) <= 1e-6 * max(abs(fd), abs(ad)), (fd, ad)


@pytest.mark.parametrize("nsteps,every", [(10, 5), (7, 3)])
def test_adjoint_gradient_matches_central_fd(require_gpu, nsteps, every):
    """The reverse sweep through torch.autograd against central differences
    for porosity, alpha and beta, over checkpoint segments of `every` steps
    (7 steps in segments of 3 leaves a short last segment).

    The run starts from a developed flow (40 steps in): on a column at rest
    the drag coefficients move J (about 3e3 Pa) by ~1e-9 per 1e-6 relative
    step, below what a 1e-13 PPE tolerance resolves.  Two bars: a 1e-6
    relative step to 1e-5, and a 1e-4 step (truncation ~1e-8, solver noise
    ~1e-9) to 1e-6."""
    s = make_case(steps=40)
    obj = adjoint.slope_objective(gauges=(1.2, 0.3))
    theta0 = adjoint.theta_of(s.params)
    th = torch.tensor(theta0, dtype=torch.float64, requires_grad=True)
    J = adjoint.objective(th, s, SLOTS["start"], obj, nsteps, every)
    J.backward()
    g_adj = th.grad.numpy().copy()
    for rel, tol in ((1e-6, 1e-5), (1e-4, 1e-6)):
        g_fd = adjoint.central_fd(s, SLOTS["start"], obj, theta0, nsteps, every, rel=rel)
        err = np.abs(g_adj - g_fd) / np.abs(g_fd)
        print(f"J {J.item():.12e} adjoint {g_adj} fd({rel:g}) {g_fd} rel.err {err} "
              f"adjoint iters {s.last_adjoint_iters}")
        assert np.all(err <= tol), (rel, g_adj, g_fd, err)
    s.close()
