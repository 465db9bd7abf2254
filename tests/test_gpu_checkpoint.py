"""Checkpoints carry what a restored state needs for the next stage: a
neighbour search right after isph_restore uses the bin starts of the
restored state's own sort, not those of whatever the context sorted last;
and a context whose solve failed steps on from a restored checkpoint as a
fresh one does."""
import numpy as np
import pytest

from beach_slope_confined_amd import capi, configs

pytestmark = pytest.mark.gpu


def test_neighbour_search_after_restore(require_gpu):
    sim = configs.make_sim(configs.dam_break(0.01), capi.IsphSim)
    sim.step(40)  # a moving flow: particles change cells between sorts
    s = sim.scalars()
    sim.set_scalars(s["umax"], s["dt"], s["t"] + s["dt"])
    sim.projection()
    sim.cellshift()
    sim.checkpoint(1)
    sim.neighbour()
    ptr0, idx0 = sim.get_neighbours()
    sim.step(30)
    sim.restore(1)
    sim.neighbour()
    ptr1, idx1 = sim.get_neighbours()
    sim.close()
    assert np.array_equal(ptr0, ptr1)
    assert np.array_equal(idx0, idx1)


PRE = ("projection", "cellshift", "resetid", "neighbour", "effpor", "eddyvis", "int_vel")


@pytest.mark.parametrize("precision,tol", [(0, 1e-8), (2, 1e-8), (3, 1e-9), (4, 1e-3)])
def test_restore_after_a_failed_solve_steps_as_a_fresh_context(require_gpu, precision, tol):
    """A solve that fails (NaN right-hand side: ISPH_E_NOCONV, in every
    solver mode) leaves the context usable: isph_restore and five steps give
    bit for bit the state of a fresh context's five steps."""
    def make():
        cfg = configs.dam_break(0.01)
        cfg["params"] = cfg["params"].copy(precision=precision, tol=tol,
                                           maxit=150 if precision == 4 else 200000)
        return configs.make_sim(cfg, capi.IsphSim)

    fresh = make()
    fresh.step(5)
    want, ws = fresh.get_particles(), fresh.scalars()
    fresh.close()
    sim = make()
    sim.checkpoint(0)
    s = sim.scalars()
    sim.set_scalars(s["umax"], s["dt"], s["t"] + s["dt"])
    for st in PRE:
        getattr(sim, st)()
    sim.ppe_assemble()
    sim.set_field("rhs", np.full(sim.count(), np.nan))
    with pytest.raises(capi.IsphError, match="NOCONV"):
        sim.ppe_solve()
    sim.restore(0)
    sim.step(5)
    got, gs = sim.get_particles(), sim.scalars()
    sim.close()
    for k in want:
        assert np.array_equal(got[k], want[k]), k
    for k in ("umax", "dt", "t", "iters", "relres"):
        assert gs[k] == ws[k], k
