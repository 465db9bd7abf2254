"""Parity at the full size of the single-step-at-scale workload: the
1024x1024 tank (1,057,810 particles), one whole ISPH step stage by stage on
the gfx950 library and on the CPU oracle (OpenMP on the host cores).  The
neighbour lists (both searches), the PPE matrix, the iteration count and
every particle array must be bit-identical."""
import hashlib
import json
import os
import time

import numpy as np
import pytest

from beach_slope_confined_amd import configs

from parity_util import STAGES, diff_state, pair, run_stage

pytestmark = pytest.mark.gpu


def digest(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def test_full_size_step_bitwise(require_gpu):
    from oracle import pyoracle
    pyoracle.set_threads(16)
    with open(os.path.join(os.path.dirname(__file__), "golden", "tank1024_digests.json")) as f:
        golden = json.load(f)
    gpu, ora = pair(configs.tank_1m(1024))
    assert gpu.count() == 1057810
    t = {"gpu": 0.0, "oracle": 0.0}
    for k, name in enumerate(STAGES):
        a = time.perf_counter()
        rg = run_stage(gpu, name)
        b = time.perf_counter()
        ro = run_stage(ora, name)
        t["gpu"] += b - a
        t["oracle"] += time.perf_counter() - b
        if name == "neighbour":
            dg, do = digest(*gpu.get_neighbours()), digest(*ora.get_neighbours())
            print(f"stage {k} neighbour lists sha256 {dg[:16]} (oracle {do[:16]})", flush=True)
            assert dg == do, f"stage {k}: neighbour lists differ"
            if k == 3:
                assert dg == golden["neighbour_lists"]
        if name == "ppesolve":
            assert rg == ro, f"iterations/relres {rg} vs {ro}"
            dg, do = digest(*gpu.get_csr()), digest(*ora.get_csr())
            assert dg == do, "PPE matrix differs"
            assert dg == golden["ppe_csr"]
            print(f"PPE solve {rg}, matrix sha256 {dg[:16]}", flush=True)
    bad = diff_state(gpu, ora)
    print(f"one step: GPU {t['gpu']:.2f} s, oracle {t['oracle']:.2f} s", flush=True)
    assert bad == [], f"after one full step: differs in {bad}"
    gpu.close()
    ora.close()
