"""Parity of the gfx950 library with the CPU oracle on degenerate inputs:
no particles at all, walls without fluid, and a single fluid particle in a
walled tank.  Stage by stage, both must leave bit-identical state after
every stage that succeeds, and a stage that fails must fail on both with
the same error code (a lone free particle makes the PPE singular: the
solve ends in ISPH_E_NOCONV on both)."""
import re

import numpy as np
import pytest

from beach_slope_confined_amd import capi, configs
from oracle.pyoracle import OracleSim

from parity_util import STAGES, diff_state, run_stage

pytestmark = pytest.mark.gpu


def subset(cfg, keep):
    out = dict(cfg)
    out["parts"] = {k: v[keep] for k, v in cfg["parts"].items()}
    out["parts"]["pid"] = np.arange(int(keep.sum()), dtype=np.int32)
    return out


def case(name):
    base = configs.dam_break(0.02)
    tid = base["parts"]["tid"]
    keep = {"no_particles": np.zeros(len(tid), bool),
            "walls_only": tid != 3,
            "one_fluid": (tid != 3) | (np.cumsum(tid == 3) == 1)}[name]
    return subset(base, keep)


def code(err):
    m = re.search(r"ISPH_E_\w+", str(err))
    return m.group(0) if m else str(err)


def attempt(sim, name):
    try:
        return run_stage(sim, name), None
    except (capi.IsphError, RuntimeError) as e:
        return None, code(e)


@pytest.mark.parametrize("name", ["no_particles", "walls_only", "one_fluid"])
def test_degenerate_inputs_stage_by_stage(require_gpu, name):
    cfg = case(name)
    gpu = configs.make_sim(cfg, capi.IsphSim, cap_extra=1)
    ora = configs.make_sim(cfg, OracleSim, cap_extra=1)
    assert gpu.count() == ora.count()
    assert diff_state(gpu, ora) == [], "after init"
    failed = None
    for step in range(2):
        for st in STAGES:
            print(f"{name} step {step} {st}", flush=True)
            rg, eg = attempt(gpu, st)
            ro, eo = attempt(ora, st)
            assert eg == eo, f"step {step} {st}: GPU {eg}, oracle {eo}"
            if eg:
                failed = (st, eg)
                break
            if st == "ppesolve":
                assert rg == ro, f"step {step}: iterations/relres {rg} vs {ro}"
            bad = diff_state(gpu, ora)
            assert bad == [], f"step {step} {st}: differs in {bad}"
        if failed:
            break
    gpu.close()
    ora.close()
    if name == "one_fluid":
        assert failed == ("ppesolve", "ISPH_E_NOCONV")
    else:
        assert failed is None
