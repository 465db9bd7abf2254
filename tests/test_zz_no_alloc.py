"""isph_probe_noalloc, the probe entry point that allocates no device memory,
against isph_probe on the same state: bit-identical values on two calls in a
row (the second smaller, so stale scratch contents would show), from two
contexts on one GPU (they share the scratch array), on an empty context,
and a clear error beyond ISPH_PROBE_NOALLOC_MAX."""
import numpy as np
import pytest

from beach_slope_confined_amd import capi, configs

pytestmark = pytest.mark.gpu

MAX = 4096  # ISPH_PROBE_NOALLOC_MAX


def call(sim, name, px, py):
    px = np.ascontiguousarray(px, np.float64)
    py = np.ascontiguousarray(py, np.float64)
    out = np.full(len(px), np.nan)
    rc = getattr(sim._lib, "isph_" + name)(sim._h, len(px), capi._dp(px), capi._dp(py),
                                           capi._dp(out))
    return rc, out


def points(rng, n, parts):
    """Half the points on particles (values from many neighbours), half
    anywhere in the tank (some with no fluid in range)."""
    k = rng.integers(0, len(parts["x"]), n // 2)
    px = np.concatenate([parts["x"][k], rng.uniform(0.0, 1.6, n - n // 2)])
    py = np.concatenate([parts["y"][k], rng.uniform(0.0, 0.6, n - n // 2)])
    return px, py


def test_probe_noalloc_equals_probe(require_gpu):
    cfg = configs.dam_break(0.01)
    sims = [configs.make_sim(cfg, capi.IsphSim) for _ in range(2)]
    sims[0].step(3)
    sims[1].step(5)
    rng = np.random.default_rng(11)
    for n in (MAX, 700):             # two calls in a row, the second smaller
        for s in sims:
            print(f"probe_noalloc: {n} points", flush=True)
            px, py = points(rng, n, s.get_particles())
            rc0, a = call(s, "probe", px, py)
            rc1, b = call(s, "probe_noalloc", px, py)
            assert rc0 == 0 and rc1 == 0
            assert np.array_equal(a, b)
            assert np.count_nonzero(a) > n // 5
    # beyond the maximum: a clear error before any work
    z = np.zeros(MAX + 1)
    rc, _ = call(sims[0], "probe_noalloc", z, z)
    assert rc == -1
    assert b"ISPH_PROBE_NOALLOC_MAX" in sims[0]._lib.isph_last_error(sims[0]._h)
    rc, _ = call(sims[0], "probe_noalloc", z[:0], z[:0])
    assert rc == 0
    for s in sims:
        s.close()
    # a context without particles: every value is 0, as isph_probe gives
    e = capi.IsphSim(cfg["params"], 16, 1)
    px, py = np.full(5, 0.3), np.full(5, 0.2)
    rc0, a = call(e, "probe", px, py)
    rc1, b = call(e, "probe_noalloc", px, py)
    assert rc0 == 0 and rc1 == 0 and np.array_equal(a, np.zeros(5)) and np.array_equal(b, a)
    e.close()
