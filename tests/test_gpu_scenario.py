"""The reference scenario on the gfx950 library: buffer zone removal and
insertion (internal.f90:475-614) bit-identical with the oracle, and main.f90's
three phases (start-up, buffered, salt-water) with buffer insertions inside
the steps, bit-identical after every phase."""
import numpy as np
import pytest

from beach_slope_confined_amd import capi, scenario
from oracle.pyoracle import OracleSim

from parity_util import FIELDS, diff_state

pytestmark = pytest.mark.gpu

SALT = FIELDS + ("con", "ovol", "oden", "mass")


def _pair(sc, **prm):
    p = sc["params"].copy(**prm) if prm else sc["params"]
    cap = len(sc["parts"]["x"]) + 4000
    sims = [capi.IsphSim(p, cap, len(sc["soil"]["x"])), OracleSim(p, cap, len(sc["soil"]["x"]))]
    for s in sims:
        s.set_particles(sc["parts"])
        s.set_soil(sc["soil"])
    return sims


def test_buffer_ops_bitwise(require_gpu):
    sc = scenario.reference_scenario()
    gpu, ora = _pair(sc)
    zones = scenario.phase_zones(sc, sc["setup"]["rel_den"])
    for s in (gpu, ora):
        s.set_buffer_zones(zones)
    n0 = gpu.count()
    for op, zi in (("insert", 0), ("insert", 1), ("remove", 0), ("remove", 1), ("insert", 1),
                   ("cellshift", None), ("remove", 1)):
        print(f"buffer op {op} {zi}", flush=True)
        for s in (gpu, ora):
            if op == "cellshift":
                s.cellshift()
            else:
                getattr(s, op + "_buffer")(zi)
        assert gpu.count() == ora.count()
        bad = diff_state(gpu, ora, SALT)
        assert bad == [], f"{op} {zi}: differs in {bad}"
    assert gpu.count() != n0


def _snapshot(sim):
    q = sim.get_particles()
    q.update({f: sim.get_field(f) for f in SALT})
    q["scalars"] = sim.scalars()
    return q


def test_reference_run_bitwise(require_gpu):
    """main.f90's phases, a few steps each, with insertion periods shortened
    so that zone 1 acts inside phase 2 (at 2.5 dt and 5 dt) and zone 0
    inside phase 3 (6 dt, past the 5 dt time shift)."""
    sc = scenario.reference_scenario()
    dt = sc["setup"]["dt"]
    sc["zones"][0]["period"] = 6.0 * dt
    sc["zones"][1]["period"] = 2.5 * dt
    gpu, ora = _pair(sc)
    snaps = {}
    for tag, s in (("gpu", gpu), ("ora", ora)):
        print(f"reference run {tag}", flush=True)
        got = []
        scenario.run_reference(s, sc, steps1=2, steps2=3, steps3=3,
                               on_phase=lambda sim, name: got.append((name, _snapshot(sim))))
        snaps[tag] = got
    top = []
    for (name, a), (_, b) in zip(snaps["gpu"], snaps["ora"]):
        assert a["scalars"] == b["scalars"], name
        for k in a:
            if k == "scalars":
                continue
            assert a[k].shape == b[k].shape and np.array_equal(a[k], b[k], equal_nan=True), (name, k)
        top.append(int(a["pid"].max()))
    # inserted particles get fresh pids in both buffered phases
    assert top[0] < top[1] < top[2], top
