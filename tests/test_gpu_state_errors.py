"""Misuse of the stage interface is an error on the library exactly as on
the oracle: a stage that needs the neighbour lists before the first
search, a solve before an assembly.  Both return ISPH_E_STATE and leave
the state untouched, and the context works normally afterwards.  (A
restore from an empty checkpoint slot, which the oracle has no slots for,
is ISPH_E_STATE on the library.)"""
import re

import numpy as np
import pytest

from beach_slope_confined_amd import capi, configs
from oracle.pyoracle import OracleSim

from parity_util import diff_state

pytestmark = pytest.mark.gpu

MISUSE = [
    ("eddyvis", lambda s: s.eddyvis()),
    ("int_vel", lambda s: s.int_vel()),
    ("ppe_assemble", lambda s: s.ppe_assemble()),
    ("ppe_solve", lambda s: s.ppe_solve()),
    ("comp_vel", lambda s: s.comp_vel()),
    ("freesurf", lambda s: s.freesurf()),
    ("opt2_shift", lambda s: s.opt2_shift()),
    ("get_csr", lambda s: s.get_csr()),
]


def code(err):
    m = re.search(r"ISPH_E_\w+", str(err))
    return m.group(0) if m else str(err)


def test_stage_misuse_is_a_state_error_on_both(require_gpu):
    cfg = configs.dam_break(0.01)
    gpu = configs.make_sim(cfg, capi.IsphSim)
    ora = configs.make_sim(cfg, OracleSim)
    for name, call in MISUSE:
        got = []
        for sim in (gpu, ora):
            with pytest.raises((capi.IsphError, RuntimeError)) as e:
                call(sim)
            got.append(code(e.value))
        assert got == ["ISPH_E_STATE", "ISPH_E_STATE"], (name, got)
        assert diff_state(gpu, ora) == [], name
    with pytest.raises(capi.IsphError, match="ISPH_E_STATE"):
        gpu.restore(5)
    gpu.step(2)
    ora.step(2)
    assert diff_state(gpu, ora) == []
    gpu.close()
    ora.close()
