"""Text output of the reference (output.f90) and its probe (probe.f90),
checked on the host.

The library's Fortran edit descriptors must print what a Fortran run-time
library prints: tests/golden/fortran_records.json holds records written by
the image's flang with the snapshot format of output.f90:206-209 and the
probe-series format of output.f90:162 (tests/golden/make_fortran_records.py).
The oracle's Shepard probe (probe.f90:86-130) is pinned by what Shepard
interpolation reproduces exactly: constants everywhere, linear fields at
points of lattice symmetry.
"""
import json
import os
import struct

import numpy as np
import pytest

from beach_slope_confined_amd import capi
from beach_slope_confined_amd.params import KERNEL_HYBRID, KERNEL_WENDLAND_C2

from oracle_util import DX, lattice, lattice_params, sim_from

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _records():
    with open(os.path.join(GOLDEN, "fortran_records.json")) as f:
        d = json.load(f)
    return [([struct.unpack("<d", bytes.fromhex(h))[0] for h in r["v"]], r["snapshot"], r["probe"])
            for r in d["records"]]


def test_snapshot_records_match_fortran_runtime():
    recs = _records()
    assert len(recs) > 200
    bad = []
    for v, snap, probe in recs:
        x, y, p, vx, vy, c = v
        # the record prints con + 0.5; pass con so that con + 0.5 == v[5]
        got = capi.snapshot_record(x, y, p, vx, vy, c - 0.5)
        if c - 0.5 + 0.5 == c and got != snap:
            bad.append((v, got, snap))
        fields = [capi.format_f(x, 10, 3), capi.format_f(y, 10, 3), capi.format_f(p, 10, 3),
                  capi.format_es(vx, 10, 3), capi.format_es(vy, 10, 3), capi.format_f(c, 10, 5)]
        if " ".join(fields) != snap:
            bad.append((v, " ".join(fields), snap))
        got_p = capi.format_f(x, 10, 3) + " " + capi.format_f(y, 10, 3)
        if got_p != probe:
            bad.append((v, got_p, probe))
    assert bad == [], bad[:5]


@pytest.mark.parametrize("w,d", [(10, 3), (10, 5), (6, 3), (5, 3), (4, 3)])
def test_format_f_widths(w, d):
    """Fw.d: right-justified, the leading zero of |v| < 1 dropped when only it
    does not fit, asterisks when the digits do not fit."""
    for v in (0.5, -0.5, 12.25, -123.456, 0.0, 1e7):
        s = capi.format_f(v, w, d)
        assert len(s) == w
        full = f"{v:.{d}f}"
        if len(full) <= w:
            assert s == full.rjust(w)
        elif full.lstrip("-").startswith("0.") and len(full) == w + 1:
            assert s == full.replace("0.", ".", 1).rjust(w)
        else:
            assert s == "*" * w


def test_format_es_exponent_forms():
    assert capi.format_es(0.0, 10, 3) == " 0.000E+00"
    assert capi.format_es(-1.5e-7, 10, 3) == "-1.500E-07"
    assert capi.format_es(2.5e123, 10, 3) == " 2.500+123"
    assert capi.format_es(9.9996, 10, 3) == " 1.000E+01"
    assert capi.format_es(1e-5, 6, 3) == "******"


def _probe_lattice(kernel):
    p = lattice_params(kernel=kernel, hfac=1.3)
    parts = lattice(21)
    return p, parts


@pytest.mark.parametrize("kernel", [KERNEL_HYBRID, KERNEL_WENDLAND_C2])
def test_probe_reproduces_constant_and_linear_fields(kernel):
    """Shepard interpolation is exact for constants anywhere and for linear
    fields at a point of lattice symmetry (a lattice site or the centre of
    a lattice square)."""
    p, parts = _probe_lattice(kernel)
    x0 = 0.25 + 10 * DX
    pts = [(x0, x0), (x0 + 0.5 * DX, x0 + 0.5 * DX), (x0 + 0.3 * DX, x0 - 0.7 * DX)]
    parts["pressure"] = np.full(len(parts["x"]), 1234.5)
    s = sim_from(parts, p)
    assert np.abs(s.probe(pts) - 1234.5).max() < 1e-9
    parts["pressure"] = 1000.0 + 3.0e4 * (parts["x"] - 0.3) - 2.0e4 * (parts["y"] - 0.4)
    s = sim_from(parts, p)
    got = s.probe(pts[:2])
    expect = [1000.0 + 3.0e4 * (px - 0.3) - 2.0e4 * (py - 0.4) for px, py in pts[:2]]
    assert np.abs(got - expect).max() < 1e-8


def test_probe_ignores_walls_and_far_fluid():
    """Only fluid (tid > 2) within 2h counts (probe.f90:68); with none in
    range the probe reads 0 (probe.f90:102)."""
    p, parts = _probe_lattice(KERNEL_WENDLAND_C2)
    n = len(parts["x"])
    parts["pressure"] = np.full(n, 500.0)
    ii = parts["pid"] // 21
    parts["tid"] = np.where(ii < 10, 1, 3).astype(np.int32)     # left half wall
    parts["pressure"][ii < 10] = 1e9
    s = sim_from(parts, p)
    x_in = 0.25 + 15 * DX
    far = 0.25 + 30 * DX
    got = s.probe([(x_in, 0.25 + 10 * DX), (far, far)])
    assert abs(got[0] - 500.0) < 1e-9
    assert got[1] == 0.0


def test_probe_series_file(tmp_path):
    t = np.array([0.0, 0.5, 1.25])
    pr = np.array([0.0, 2943.0, -15.0])
    f = tmp_path / "probe.txt"
    capi.write_probe_series(str(f), t, pr, g=-9.81, wc=0.3, rho=1000.0)
    lines = f.read_text().splitlines()
    a = t * np.sqrt(9.81 / 0.3)
    b = pr / (1000.0 * 9.81 * 0.3)
    assert lines == [capi.format_f(x, 10, 3) + " " + capi.format_f(y, 10, 3) for x, y in zip(a, b)]
    assert lines[1] == "     2.859      1.000"
