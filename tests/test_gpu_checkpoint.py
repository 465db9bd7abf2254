"""Checkpoints carry what a restored state needs for the next stage: a
neighbour search right after isph_restore uses the bin starts of the
restored state's own sort, not those of whatever the context sorted last."""
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
