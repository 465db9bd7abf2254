"""The cell sort (an LSD radix sort of the cell-hash keys, stable) on
particles handed over in random order.  The first binning of a shuffled
set must give the oracle's order bit for bit: bin key ascending, ties in
input order.  The property is also checked on the GPU's own output for the
(cell, fluid bit) keys, and the steps after it must stay bitwise."""
import numpy as np
import pytest

from beach_slope_confined_amd import capi, configs
from oracle.pyoracle import OracleSim

from parity_util import STAGES, diff_state, run_stage

pytestmark = pytest.mark.gpu

CASES = {
    "dam_break": lambda: configs.dam_break(0.01),
    "dam_break_refcells": lambda: configs.with_reference_cells(configs.dam_break(0.01)),
    "tank_256": lambda: configs.tank_1m(256),
}


def shuffled(cfg, seed=7):
    perm = np.random.default_rng(seed).permutation(len(cfg["parts"]["x"]))
    out = dict(cfg)
    out["parts"] = {k: v[perm] for k, v in cfg["parts"].items()}
    return out


def cell_keys(p, prm):
    """(cell, fluid bit) keys of the cell order, as the grid bins positions:
    interior cells 1-based, columns of ny cells."""
    nx = int(np.floor((prm.xmax - prm.xmin) / prm.cell_size)) + 3
    ny = int(np.floor((prm.ymax - prm.ymin) / prm.cell_size)) + 3
    cx = np.floor((p["x"] - prm.xmin) / prm.cell_size).astype(np.int64) + 1
    cy = np.floor((p["y"] - prm.ymin) / prm.cell_size).astype(np.int64) + 1
    assert cx.min() >= 1 and cx.max() <= nx - 2 and cy.min() >= 1 and cy.max() <= ny - 2
    return 2 * (cx * ny + cy) + (p["tid"] == 3)


@pytest.mark.parametrize("name", list(CASES))
def test_shuffled_input_sorts_like_the_oracle(require_gpu, name):
    cfg = shuffled(CASES[name]())
    print(f"{name}: {len(cfg['parts']['x'])} particles, shuffled", flush=True)
    gpu = configs.make_sim(cfg, capi.IsphSim)
    ora = configs.make_sim(cfg, OracleSim)
    assert diff_state(gpu, ora) == [], "after the first binning"
    if cfg["params"].cell_order == 0:
        p = gpu.get_particles()
        key = cell_keys(p, cfg["params"])
        # input position of every particle, by its id
        where = np.empty(len(key), np.int64)
        where[cfg["parts"]["pid"]] = np.arange(len(key))
        pos = where[p["pid"]]
        assert np.all(np.diff(key) >= 0), "keys ascend"
        same = np.diff(key) == 0
        assert np.all(np.diff(pos)[same] > 0), "ties in input order"
    for st in STAGES:
        run_stage(gpu, st)
        run_stage(ora, st)
        bad = diff_state(gpu, ora)
        assert bad == [], f"stage {st}: differs in {bad}"
