"""Pins the oracle's restatement of the reference's live PPE solve
(solver.f90:306-347: MKL's RCI dfgmres without a preconditioner, relative
tolerance 1e-3, zero-norm check, restart and iteration limit min(150, n))
against what fixes GMRES independently of how it is coded: scipy's GMRES
on the same systems and warm starts, the minimal-residual property of the
k-th iterate over x0 + K_k(A, r0), exact convergence within n steps on the
reference's own 5x5 solvertest system, and the RCI loop's behaviour when
the iteration limit comes first (the solution is never fetched, so x keeps
the warm start).  MKL is absent, so MKL itself is not run; its documented
parameters are what the restatement follows.

A wrong Givens sign, a dropped reorthogonalisation, a residual taken
against ||b|| instead of ||r0||, or a restart that forgets x fail one of
these checks; none re-types the oracle's recurrences.
"""
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from beach_slope_confined_amd import configs
from beach_slope_confined_amd.params import PREC_FGMRES
from oracle import pyoracle
from oracle.pyoracle import OracleSim

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def system(n, seed, shift=4.0, density=0.03):
    rng = np.random.default_rng(seed)
    A = sp.random(n, n, density=density, random_state=seed + 1, format="csr")
    A = (A + sp.eye(n) * shift).tocsr()
    A.sort_indices()
    return A, rng.normal(size=n), rng.normal(size=n)


def scipy_gmres(A, b, x0, atol, restart, cycles):
    """scipy's restarted GMRES with the stop test ||r|| <= atol and the
    number of inner iterations it took"""
    count = [0]

    def cb(_):
        count[0] += 1

    x, info = spla.gmres(A, b, x0=x0.copy(), rtol=0.0, atol=atol, restart=restart,
                         maxiter=cycles, callback=cb, callback_type="pr_norm")
    return x, info, count[0]


@pytest.mark.parametrize("n,seed,tol", [(200, 0, 1e-3), (300, 3, 1e-8), (120, 7, 1e-6)])
def test_matches_scipy_gmres_from_zero(n, seed, tol):
    A, b, _ = system(n, seed)
    rc, x, it, rel, conv = pyoracle.fgmres_csr(A.indptr, A.indices, A.data, b, np.zeros(n),
                                               tol, 150)
    assert rc == 0 and conv
    xs, info, its = scipy_gmres(A, b, np.zeros(n), tol * np.linalg.norm(b), 150, 1)
    assert info == 0
    assert it == its
    assert np.abs(x - xs).max() <= 1e-10 * np.abs(xs).max()
    true = np.linalg.norm(b - A @ x) / np.linalg.norm(b)
    assert true <= tol * (1 + 1e-8)
    assert rel == pytest.approx(true, rel=1e-6)


def test_warm_start_tolerance_is_relative_to_the_initial_residual():
    """dpar(4) = dpar(1) * dpar(3) with dpar(3) = ||b - A x0||, not ||b||"""
    A, b, x0 = system(250, 5)
    x0 = np.linalg.solve(A.toarray(), b) + 1e-6 * x0
    r0 = np.linalg.norm(b - A @ x0)
    rc, x, it, rel, conv = pyoracle.fgmres_csr(A.indptr, A.indices, A.data, b, x0, 1e-3, 150)
    assert rc == 0 and conv
    xs, info, its = scipy_gmres(A, b, x0, 1e-3 * r0, 150, 1)
    assert it == its
    assert np.abs(x - xs).max() <= 1e-10 * np.abs(xs).max()
    assert np.linalg.norm(b - A @ x) <= 1e-3 * r0 * (1 + 1e-8)
    # a test against ||b|| would have stopped at once: r0 < 1e-3 ||b|| here
    assert r0 < 1e-3 * np.linalg.norm(b) and it > 1


@pytest.mark.parametrize("restart", [3, 7])
def test_restart_cycles_match_scipy(restart):
    A, b, _ = system(150, 9, shift=2.5)
    cycles = 6
    rc, x, it, rel, conv = pyoracle.fgmres_csr(A.indptr, A.indices, A.data, b, np.zeros(150),
                                               1e-300, restart * cycles + 1, restart)
    # the limit comes one iteration into the next cycle: x is the iterate
    # after `cycles` full cycles
    assert rc == 0 and not conv and it == restart * cycles + 1
    xs, info, _ = scipy_gmres(A, b, np.zeros(150), 0.0, restart, cycles)
    assert np.abs(x - xs).max() <= 1e-9 * np.abs(xs).max()


@pytest.mark.parametrize("k", [1, 2, 5, 8])
def test_kth_iterate_minimises_the_residual_over_the_krylov_space(k):
    A, b, x0 = system(60, 13, shift=3.0, density=0.1)
    Ad = A.toarray()
    rc, x, it, rel, conv = pyoracle.fgmres_csr(A.indptr, A.indices, A.data, b, x0, 1e-300,
                                               k + 1, k)
    assert rc == 0 and it == k + 1
    r0 = b - Ad @ x0
    K = np.empty((60, k))
    v = r0.copy()
    for j in range(k):
        K[:, j] = v
        v = Ad @ v
    Q, _ = np.linalg.qr(K)
    y, *_ = np.linalg.lstsq(Ad @ Q, r0, rcond=None)
    best = np.linalg.norm(r0 - Ad @ Q @ y)
    got = np.linalg.norm(b - Ad @ x)
    assert got == pytest.approx(best, rel=1e-9, abs=1e-13)
    assert np.abs(x - (x0 + Q @ y)).max() <= 1e-8 * np.abs(x).max()


def test_known_answer_solvertest():
    """the reference's solvertest system (x = 1..5): exact within 5 steps"""
    with open(os.path.join(GOLDEN, "solvertest_5x5.json")) as f:
        d = json.load(f)
    rc, x, it, rel, conv = pyoracle.fgmres_csr(d["rowptr"], d["col"], d["val"], d["b"],
                                               d["guess"], 1e-14, 150)
    assert rc == 0 and conv and it <= 5
    assert np.allclose(x, d["x"], rtol=1e-12)


def test_iteration_limit_keeps_the_warm_start():
    """dfgmres_get runs only for rci_request = 0 (solver.f90:330-333)"""
    A, b, x0 = system(200, 21)
    rc, x, it, rel, conv = pyoracle.fgmres_csr(A.indptr, A.indices, A.data, b, x0, 1e-12, 4)
    assert rc == 0 and not conv and it == 4
    assert np.array_equal(x, x0)
    assert rel > 1e-12


def test_zero_norm_stop_on_an_invariant_subspace():
    """b in a 3-dimensional invariant subspace: the 4th Arnoldi vector has
    zero norm (ipar(12) = 1) and the 3-step solution is exact"""
    n = 40
    D = np.diag(np.arange(1.0, n + 1))
    A = sp.csr_matrix(D)
    b = np.zeros(n)
    b[[2, 10, 30]] = [1.0, -2.0, 0.5]
    rc, x, it, rel, conv = pyoracle.fgmres_csr(A.indptr, A.indices, A.data, b, np.zeros(n),
                                               1e-300, 150)
    assert rc == 0 and conv and it == 3
    assert np.allclose(x, b / np.arange(1.0, n + 1), rtol=1e-13, atol=1e-15)


def test_zero_initial_residual_does_nothing():
    A, b, _ = system(50, 2)
    x0 = np.linalg.solve(A.toarray(), b)
    b = A @ x0
    rc, x, it, rel, conv = pyoracle.fgmres_csr(A.indptr, A.indices, A.data, b, x0, 1e-3, 150)
    assert rc == 0 and conv
    assert it <= 1
    assert np.linalg.norm(b - A @ x) <= 1e-3 * max(np.linalg.norm(b - A @ x0), 1e-300) + 1e-12


def test_ppe_step_warm_start_is_fsol_in_row_order():
    """Two FGMRES solves on the PPE: the second starts from the first's
    solution vector by row (fsol), which after the re-binning belongs to
    other particles; with the limit reached first the pressure is fsol."""
    cfg = configs.dam_break(0.02)
    cfg["params"] = cfg["params"].copy(precision=PREC_FGMRES, tol=1e-3, maxit=150)
    sim = configs.make_sim(cfg, OracleSim)
    p_up = sim.get_particles()["pressure"]
    s = sim.scalars()
    sim.set_scalars(s["umax"], s["dt"], s["t"] + s["dt"])
    for st in ("projection", "cellshift", "resetid", "neighbour", "effpor", "eddyvis", "int_vel"):
        getattr(sim, st)()
    sim.ppe_assemble()
    rp, col, val, rhs = sim.get_csr()
    it, rel = sim.ppe_solve()
    p1 = sim.get_particles()["pressure"]
    A = sp.csr_matrix((val, col, rp))
    # the warm start was the uploaded pressures in the first row order
    rc, x, it2, rel2, conv = pyoracle.fgmres_csr(rp, col, val, rhs, p_up, 1e-3, 150)
    assert conv and it2 == it and np.array_equal(x, p1)
    r0 = np.linalg.norm(rhs - A @ p_up)
    assert np.linalg.norm(rhs - A @ p1) <= 1e-3 * r0 * (1 + 1e-8)
    # the same matrix again with a limit of 2 iterations from fsol = p1
    cfg2 = dict(cfg)
    cfg2["params"] = cfg["params"].copy(maxit=2, tol=1e-14)
    rc, x, it3, rel3, conv3 = pyoracle.fgmres_csr(rp, col, val, rhs, p1, 1e-14, 2)
    assert not conv3 and np.array_equal(x, p1)


def test_second_gram_schmidt_pass_keeps_the_basis_orthogonal():
    """On a clustered spectrum (I + 0.3 R) GMRES's minimal residual keeps
    falling after the solution is reached to rounding: the Arnoldi vectors
    stay orthonormal only with the second Gram-Schmidt pass, and the residual
    estimate |g_(j+1)| then follows the exact minimum far below eps.  One
    classical pass leaves components of size eps along the basis and floors
    the estimate near 1e-16."""
    n = 200
    R = sp.random(n, n, density=0.05, random_state=5, format="csr")
    R = R / abs(R).sum(axis=1).max()
    A = (sp.eye(n) + 0.3 * R).tocsr()
    A.sort_indices()
    b = np.random.default_rng(4).normal(size=n)
    rc, x, it, rel, conv = pyoracle.fgmres_csr(A.indptr, A.indices, A.data, b, np.zeros(n),
                                               1e-300, 20, 20)
    assert rc == 0 and it == 20 and not conv
    assert rel < 1e-20
    rc, x, it, rel, conv = pyoracle.fgmres_csr(A.indptr, A.indices, A.data, b, np.zeros(n),
                                               1e-300, 21, 20)
    assert np.linalg.norm(b - A @ x) <= 1e-14 * np.linalg.norm(b)
