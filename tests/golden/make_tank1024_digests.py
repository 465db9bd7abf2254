"""Writes tank1024_digests.json: SHA-256 digests of the CPU oracle's first
neighbour lists and assembled PPE matrix for the 1024x1024 tank
(1,057,810 particles), the single-step-at-scale workload.  The GPU test
tests/test_gpu_parity_full.py checks the gfx950 library against the oracle
bit for bit at this size; these digests pin the oracle (and the seeded
generator) on the CPU so that a change to either is caught without a GPU.

    python tests/golden/make_tank1024_digests.py
"""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))


def digest(*arrays):
    import numpy as np
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def compute():
    from beach_slope_confined_amd import configs
    from oracle.pyoracle import OracleSim
    sim = configs.make_sim(configs.tank_1m(1024), OracleSim)
    s = sim.scalars()
    sim.set_scalars(s["umax"], s["dt"], s["t"] + s["dt"])
    for name in ("projection", "cellshift", "resetid", "neighbour"):
        getattr(sim, name)()
    out = {"n": sim.count(), "neighbour_lists": digest(*sim.get_neighbours())}
    for name in ("effpor", "eddyvis", "int_vel", "ppe_assemble"):
        getattr(sim, name)()
    out["ppe_csr"] = digest(*sim.get_csr())
    sim.close()
    return out


if __name__ == "__main__":
    d = compute()
    with open(os.path.join(HERE, "tank1024_digests.json"), "w") as f:
        json.dump(d, f, indent=1)
    print(d)
