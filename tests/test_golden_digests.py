"""The CPU oracle's neighbour lists and PPE matrix for the 1024x1024 tank
(1,057,810 particles) against the committed digests
(tests/golden/make_tank1024_digests.py).  The GPU library matches the
oracle bit for bit at this size (tests/test_gpu_parity_full.py)."""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))


def test_tank1024_digests():
    import make_tank1024_digests as m
    with open(os.path.join(HERE, "golden", "tank1024_digests.json")) as f:
        want = json.load(f)
    assert m.compute() == want
