"""Contexts are independent of each other and of their capacity.

- Two contexts stepped alternately (multigrid and FP64 solves, so their
  kernels interleave on the GPU) end bit for bit where each ends when run
  alone.
- Spare capacity changes only the stride of the per-neighbour arrays (ELL,
  leading dimension = capacity), never the results: the same run with
  777 spare slots is bit-identical.
"""
import numpy as np
import pytest

from beach_slope_confined_amd import capi, configs
from beach_slope_confined_amd.params import PREC_FP64, PREC_MIXED_MG

pytestmark = pytest.mark.gpu

STEPS = 4


def make(precision, cap_extra=0):
    cfg = configs.dam_break(0.01)
    tol = 1e-9 if precision == PREC_MIXED_MG else 1e-8
    cfg["params"] = cfg["params"].copy(precision=precision, tol=tol, maxit=200000)
    return configs.make_sim(cfg, capi.IsphSim, cap_extra=cap_extra)


def final(sim):
    p = sim.get_particles()
    s = sim.scalars()
    return p, (s["umax"], s["dt"], s["t"], s["iters"], s["relres"])


def same(a, b):
    pa, sa = a
    pb, sb = b
    return [k for k in pa if not np.array_equal(pa[k], pb[k])] + (["scalars"] if sa != sb else [])


def test_interleaved_contexts_match_solo_runs(require_gpu):
    solo = {}
    for prec in (PREC_MIXED_MG, PREC_FP64):
        sim = make(prec)
        sim.step(STEPS)
        solo[prec] = final(sim)
        sim.close()
    a, b = make(PREC_MIXED_MG), make(PREC_FP64)
    for _ in range(STEPS):
        a.step(1)
        b.step(1)
    fa, fb = final(a), final(b)
    a.close()
    b.close()
    assert same(fa, solo[PREC_MIXED_MG]) == []
    assert same(fb, solo[PREC_FP64]) == []


@pytest.mark.parametrize("precision", [PREC_FP64, PREC_MIXED_MG])
def test_spare_capacity_changes_nothing(require_gpu, precision):
    out = []
    for extra in (0, 777):
        sim = make(precision, extra)
        sim.step(STEPS)
        out.append(final(sim))
        sim.close()
    assert same(out[0], out[1]) == []
