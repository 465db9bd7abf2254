"""The reference particle order (cell_order = ORDER_REFERENCE: cellshift's
ten-class cell lists, resetid numbering, porsearch's 2x2-cell soil order) on
the gfx950 library: bit-identical with the oracle stage by stage, with the
outflow sink deleting particles, and in the salt-water step."""
import numpy as np
import pytest

from beach_slope_confined_amd import configs

from parity_util import FIELDS, STAGES, diff_state, pair, run_stage

pytestmark = pytest.mark.gpu

CASES = {
    "dam_break": lambda: configs.with_reference_cells(configs.dam_break(0.01)),
    "dam_break_reference": lambda: configs.with_reference_cells(configs.dam_break_reference(0.01)),
}


@pytest.mark.parametrize("case", list(CASES))
def test_reference_order_stage_by_stage(require_gpu, case):
    gpu, ora = pair(CASES[case]())
    assert diff_state(gpu, ora) == [], "after init"
    for step in range(2):
        for name in STAGES:
            print(f"order {case} step {step} {name}", flush=True)
            rg = run_stage(gpu, name)
            ro = run_stage(ora, name)
            if name == "ppesolve":
                assert rg == ro, f"step {step} {name}: {rg} vs {ro}"
            bad = diff_state(gpu, ora)
            assert bad == [], f"step {step} stage {name}: differs in {bad}"
            if name == "neighbour":
                pg, ig = gpu.get_neighbours()
                po, io = ora.get_neighbours()
                assert np.array_equal(pg, po) and np.array_equal(ig, io), f"step {step} lists"


def test_reference_order_with_sink_and_salt(require_gpu):
    """Steps with the outflow sink of cellshiftalt (integrator.f90:60-61)
    deleting the column's upper left corner, and the salt-water step."""
    cfg = CASES["dam_break_reference"]()
    gpu, ora = pair(cfg, sink=1, xlcutoff=0.05, ytcutoff=0.2, salinity=1)
    for s in (gpu, ora):
        s.salt_init(0.2, 0.1, 0.3)
    n0 = gpu.count()
    for step in range(4):
        print(f"order sink step {step}", flush=True)
        gpu.step(1)
        ora.step(1)
        bad = diff_state(gpu, ora, FIELDS + ("con", "oden"))
        assert bad == [], f"step {step}: differs in {bad}"
        assert gpu.scalars() == ora.scalars()
    assert gpu.count() < n0
