"""The reference scenario (setup.f90, boundary.f90, internal.f90, main.f90)
on the oracle: the generated geometry against what the setup fixes, the
initial particle order against the reference's per-cell lists, and the
buffer zones against a literal emulation of insert_buffer/remove_buffer."""
import math

import numpy as np
import pytest

from beach_slope_confined_amd import scenario
from beach_slope_confined_amd.params import FLAG_BUFFER, FLAG_DEAD
from oracle.pyoracle import OracleSim


@pytest.fixture(scope="module")
def sc():
    return scenario.reference_scenario()


def _sim(sc, extra=4000, **prm):
    p = sc["params"].copy(**prm) if prm else sc["params"]
    o = OracleSim(p, len(sc["parts"]["x"]) + extra, len(sc["soil"]["x"]))
    o.set_particles(sc["parts"])
    o.set_soil(sc["soil"])
    return o


def test_geometry_invariants(sc):
    s, p = sc["setup"], sc["parts"]
    sq = math.sqrt(s["por"])
    # the first lattice count fixes the fluid mass: fmass fpx0 fpy0 = wc wl rho
    fpx0 = math.floor(s["wl"] / (2 * s["pr"]))
    fpy0 = math.floor(s["wc"] / (2 * s["pr"])) + 1
    assert s["fmass"] * fpx0 * fpy0 == pytest.approx(s["wc"] * s["wl"] * s["rho"], rel=1e-14)
    # cells of 4 h1; every particle inside the placed cells j = 2..cellx-1, i = 2..celly-1
    cs = s["cs"]
    assert cs == 4 * s["h1"]
    assert p["x"].min() >= 0.0 and p["x"].max() < (s["cellx"] - 2) * cs
    assert p["y"].min() >= 0.0 and p["y"].max() < (s["celly"] - 2) * cs
    # the fluid column is a lattice of spacing 2 pr / sqrt(por) (internal.f90:41-44)
    f = (p["tid"] == 3) & ((p["flags"] & FLAG_BUFFER) == 0) & (p["y"] < s["set_ht"] + 0.05)
    xs = np.unique(np.round(p["x"][f], 12))
    gaps = np.diff(xs)
    assert np.allclose(gaps[gaps < 1.5 * 2 * s["prrealx"] / sq], 2 * s["prrealx"] / sq, rtol=1e-9)
    # three wall layers on the left, innermost fixed (tid 1)
    left = (p["tid"] < 3) & (p["x"] < s["domain_shift"] + 5 * s["brrealx"]) & (p["y"] > 0.05)
    lx = np.unique(np.round(p["x"][left], 12))
    assert len(lx) == 3
    assert np.all(p["tid"][left & np.isclose(p["x"], lx[-1])] == 1)
    assert np.all(p["tid"][left & ~np.isclose(p["x"], lx[-1])] == 2)
    # nothing of the column or the soil above the slope line (boundary.f90:548, internal.f90:72)
    above = lambda x, y: ((y - s["yl"]) - s["line_grad"] * (x - s["xl"])) > 0.0  # noqa: E731
    soil = sc["soil"]
    assert not np.any(above(soil["x"], soil["y"]))
    # buffers: 2 template columns each side, their dead band 2 rows deep
    for z in sc["zones"]:
        assert np.all(z["flags"] & FLAG_BUFFER)
    assert ((sc["zones"][0]["flags"] & FLAG_DEAD) != 0).sum() == 2 * 2
    # insertion period = the time to cross 2 columns at the pore velocity
    assert s["ins_1"] * abs(s["entry_vel"]) / 0.387 == pytest.approx(4 * s["prrealx"] / sq, rel=1e-14)


def test_initial_order_is_the_reference_cell_lists(sc):
    """The reference appends each placed particle to its cell's list, walls
    (btot) before fluid, segment by segment (boundary.f90, internal.f90);
    resetid then numbers the cells column by column, top row first."""
    s, p = sc["setup"], sc["parts"]
    cells = {}
    for wall in (True, False):
        for k in range(len(p["x"])):
            if (p["tid"][k] != 3) == wall:
                cells.setdefault(scenario._cell(s, p["x"][k], p["y"][k]), []).append(k)
    order = []
    for j in range(1, s["cellx"] + 1):
        for i in range(1, s["celly"] + 1):
            order += cells.get((i, j), [])
    o = _sim(sc)
    assert list(o.get_particles()["pid"]) == order


def test_buffer_zones_match_the_reference_lists(sc):
    """remove_buffer then insert_buffer on both zones, then a cellshift: the
    order equals per-cell lists where remove_buffer1/2 compacted each entry
    cell's list (internal.f90:544-614) and insert_buffer1/2 appended the
    template particles lying in that cell (:475-542)."""
    s = sc["setup"]
    o = _sim(sc)
    o.set_buffer_zones(scenario.phase_zones(sc, s["rel_den"]))
    q0 = o.get_particles()
    lists = {}
    for k, pid in enumerate(q0["pid"]):
        lists.setdefault(scenario._cell(s, q0["x"][k], q0["y"][k]), []).append(
            ("old", int(pid), q0["x"][k], int(q0["flags"][k])))
    new_pid = int(q0["pid"].max()) + 1
    for zi, z in enumerate(sc["zones"]):
        entry = [(s["celly"] + 1 - cy, cx) for (cx, cy) in z["cells"]]
        for ij in entry:
            lists[ij] = [e for e in lists.get(ij, []) if not (e[3] & FLAG_BUFFER) or
                         (e[2] < z["xcut"] if z["remove_right"] else e[2] > z["xcut"])]
        for ij in entry:
            for t in range(len(z["x"])):
                if scenario._cell(s, z["x"][t], z["y"][t]) == ij:
                    lists[ij].append(("new", None, z["x"][t], int(z["flags"][t])))
        o.remove_buffer(zi)
        o.insert_buffer(zi)
    o.cellshift()
    expect = []
    for j in range(1, s["cellx"] + 1):
        for i in range(1, s["celly"] + 1):
            expect += lists.get((i, j), [])
    q = o.get_particles()
    assert len(q["pid"]) == len(expect)
    got_old = [int(pid) if pid < new_pid else None for pid in q["pid"]]
    assert got_old == [e[1] for e in expect]
    fresh = q["pid"] >= new_pid
    assert np.all(q["tid"][fresh] == 3) and np.all(q["flags"][fresh] & FLAG_BUFFER)
    con = o.get_field("con")[fresh]
    assert set(np.unique(con)) == {-0.5, 0.5}
    assert np.allclose(o.get_field("mass")[fresh][con > 0], s["fmass"] * s["rel_den"], rtol=1e-15)


def test_buffer_schedule_fires_on_period_crossings(sc):
    """main.f90:103-121: zone k acts in the step whose advance crosses
    m_k * period_k, then m_k counts on."""
    s = sc["setup"]
    o = _sim(sc, adaptive_dt=0)
    dt = s["dt"]
    zones = scenario.phase_zones(sc)
    zones[0]["period"] = 2.5 * dt
    zones[1]["period"] = 1e9
    o.set_buffer_zones(zones)
    top = [int(o.get_particles()["pid"].max())]
    for _ in range(6):
        o.step(1)
        top.append(int(o.get_particles()["pid"].max()))
    # 2.5 dt and 5 dt are crossed by the advances of steps 3 and 5 (k = 2, 4):
    # only those steps insert new particles (fresh pids)
    fired = [k for k in range(6) if top[k + 1] > top[k]]
    assert fired == [2, 4]
