"""CPU-side checks of the boundary: the gfx950 library loads and exports
every entry point include/isph_capi.h declares; the parameter struct has
the same layout in C and in the ctypes mirror; the workload generators are
deterministic and produce the stated sizes."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from beach_slope_confined_amd import capi, configs
from beach_slope_confined_amd.params import IsphParams

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "isph_capi.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(isph_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_entry_point():
    lib = capi.load_library()
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert missing == []


def test_params_layout_matches_c(tmp_path):
    src = tmp_path / "sz.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "isph_capi.h"\n'
                   'int main(void){printf("%zu %zu %zu %zu\\n", sizeof(isph_params),'
                   ' offsetof(isph_params, kernel), offsetof(isph_params, mu),'
                   ' offsetof(isph_params, salinity));return 0;}\n')
    exe = tmp_path / "sz"
    subprocess.run(["gcc", "-I", os.path.dirname(HEADER), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    size, off_kernel, off_mu, off_res = map(int, out)
    assert size == ctypes.sizeof(IsphParams)
    assert off_kernel == IsphParams.kernel.offset
    assert off_mu == IsphParams.mu.offset
    assert off_res == IsphParams.salinity.offset


def test_generators_sizes_and_determinism():
    c = configs.tank_1m(1024)
    assert c["n_fluid"] == 1048576
    assert len(c["parts"]["x"]) == 1048576 + 3 * 1024 * 2 + 3 * 1030
    d = configs.dam_break(0.01)
    assert d["n_fluid"] == 1200
    d2 = configs.dam_break(0.01)
    for k in ("x", "y", "mass", "tid"):
        assert np.array_equal(d["parts"][k], d2["parts"][k])
    # jitter bound +-0.1 dx on the tank lattice
    p = c["parts"]
    f = p["tid"] == 3
    dx = 1.0 / 1024
    fx = (p["x"][f] / dx - 0.5)
    assert np.abs(fx - np.round(fx)).max() <= 0.1 + 1e-9


def test_splitmix_reference_values():
    # first outputs of splitmix64 with state seed + k*phi (published algorithm)
    u = configs.uniform(3, 0, seed=0)
    expect = [0xe220a8397b1dcdaf, 0x6e789e6aa1b965f4, 0x06c45d188009454f]
    assert np.array_equal(u, np.array([e >> 11 for e in expect], dtype=np.float64) / 2.0 ** 53)


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(capi, "_LIB", None)
    with pytest.raises(ImportError):
        capi.load_library(str(tmp_path / "absent.so"))


@pytest.mark.parametrize("bad", [-0.1, 1.0, float("nan")])
def test_inner_tol_outside_its_range_is_rejected(bad):
    """inner_tol lies in [0, 1) (0: the precision's default); the parameter
    check runs before any device call, so this needs no GPU."""
    lib = capi.load_library()
    prm = configs.dam_break(0.01)["params"].copy(inner_tol=bad)
    ctx = ctypes.c_void_p()
    rc = lib.isph_create(ctypes.byref(prm), 1000, 0, ctypes.byref(ctx))
    assert rc == -1 and not ctx.value
