"""The reference's text output and probes through the gfx950 library:
isph_probe (probe.f90:86-130) bit-identical with the oracle, and the
snapshot files of isph_combined (output.f90:188-234) holding one record
per live fluid, buffer and dead particle."""
import numpy as np
import pytest

from beach_slope_confined_amd import capi, configs
from beach_slope_confined_amd.params import FLAG_BUFFER, FLAG_DEAD

from parity_util import pair

pytestmark = pytest.mark.gpu

PROBES = [(0.25, (0.25 - 0.2) / 3.0), (0.30, (0.30 - 0.2) / 3.0), (0.35, (0.35 - 0.2) / 3.0),
          (0.1, 0.05), (0.39, 0.29), (1.5, 0.5)]


@pytest.mark.parametrize("case", ["dam_break", "dam_break_reference"])
def test_probe_bitwise(require_gpu, case):
    cfg = configs.dam_break(0.01) if case == "dam_break" else configs.dam_break_reference(0.01)
    gpu, ora = pair(cfg)
    for step in range(3):
        a, b = gpu.probe(PROBES), ora.probe(PROBES)
        assert np.array_equal(a, b), (step, a, b)
        assert a[-1] == 0.0 and np.all(a[:5] != 0.0)
        gpu.step(1)
        ora.step(1)


def test_combined_files(require_gpu, tmp_path):
    cfg = configs.dam_break(0.01)
    parts = cfg["parts"]
    fl = parts["flags"].copy()
    fluid = np.where(parts["tid"] == 3)[0]
    fl[fluid[::17]] |= FLAG_BUFFER
    fl[fluid[5::23]] |= FLAG_DEAD
    parts["flags"] = fl
    sim = configs.make_sim(cfg, capi.IsphSim)
    sim.set_field("con", np.linspace(-0.5, 0.5, sim.count()))
    for d in ("fluid", "buffer"):
        (tmp_path / d).mkdir()
    sim.combined(str(tmp_path), 7)
    q = sim.get_particles()
    con = sim.get_field("con")
    recs = [capi.snapshot_record(q["x"][i], q["y"][i], q["pressure"][i], q["vx"][i], q["vy"][i],
                                 con[i]) for i in range(len(con))]
    buf = (q["flags"] & FLAG_BUFFER) != 0
    dead = (q["flags"] & FLAG_DEAD) != 0
    f3 = q["tid"] == 3
    expect = {"fluid/fluid_7.txt": f3 & ~buf & ~dead, "buffer/buffer_7.txt": f3 & buf & ~dead,
              "fluid/dead_7.txt": f3 & dead & ~buf}
    for name, sel in expect.items():
        lines = (tmp_path / name).read_text().splitlines()
        assert lines == [recs[i] for i in np.where(sel)[0]], name
        assert len(lines) == sel.sum() > 0
    assert all(len(r) == 65 for r in recs)
