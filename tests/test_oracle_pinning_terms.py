"""Pins three terms of the CPU oracle (the checker of the gfx950 library)
that the rest of the suite leaves open, each by what the mathematics or the
physics fixes rather than by re-typing the oracle's formula:

- opt2_shift's velocity correction uses the shift *before* its clamp
  (part_shift.f90:112-113 clamp only the stored xs, ys; :137,159 use the
  local, unclamped xs, ys);
- comp_vel's pair pressure is the density-weighted interface pressure
  (integrator.f90:742-752), and its wall branch (:754-760) uses the
  particle's own density and porosity;
- int_vel's semi-implicit drag divides by 1 - dt R n (integrator.f90:680-686).

Each check fails for the plausible slips listed in its docstring; they were
planted one at a time in a copy of the oracle.
"""
import math

import numpy as np
import pytest

from beach_slope_confined_amd.params import DRAG_ERGUN

from oracle_util import DX, interior, lattice, lattice_params, sim_from


def jittered_block(nside, seed, amp=0.15):
    parts = lattice(nside)
    rng = np.random.default_rng(seed)
    parts["x"] = parts["x"] + rng.uniform(-amp, amp, parts["x"].size) * DX
    parts["y"] = parts["y"] + rng.uniform(-amp, amp, parts["y"].size) * DX
    return parts


def by_pid(sim, *names):
    """Fields in particle-id order (the sorted row order moves particles)."""
    order = np.argsort(sim.get_particles()["pid"])
    return [sim.get_field(n)[order] for n in names]


# ------------------------------------------------------------- shifting ---
def test_shift_velocity_correction_uses_the_unclamped_shift():
    """The shift's Taylor correction is dv = (grad v) ds_u with the corrected
    gradient and the shift ds_u before the 0.1 dl clamp.  For a linear field
    v = (a x + c y, d x + b y) + const the corrected gradient is exact at any
    arrangement (a radial kernel's moment matrix is symmetric, so L inverts
    it on either side), hence dv = G ds_u with G = [[a, c], [d, b]] and
    ds_u = G^-1 dv.  On a jittered block with one particle pushed well off
    its site: two different fields recover the same ds_u; every component
    below the cap equals the applied shift; every component at the cap was
    beyond it before the clamp, with the same sign.  Fails for: the clamped
    shift in the correction, d(vx)/dy and d(vy)/dx swapped, the uncorrected
    gradient, a dropped cross term."""
    nside = 21
    centre = (nside // 2) * nside + nside // 2
    prm = lattice_params()
    cap = prm.maxshift * prm.dl
    fields = ((0.7, -0.4, 1.3, -2.1), (-1.1, 0.9, 0.35, 2.4))
    recovered = []
    for a, b, c, d in fields:
        parts = jittered_block(nside, 7)
        parts["x"][centre] += 0.55 * DX
        parts["y"][centre] -= 0.6 * DX
        parts["vx"] = 0.3 + a * parts["x"] + c * parts["y"]
        parts["vy"] = -0.1 + d * parts["x"] + b * parts["y"]
        sim = sim_from(parts, prm)
        sim.neighbour()
        sim.opt2_shift()
        xs, ys, vxs, vys = by_pid(sim, "xs", "ys", "vxs", "vys")
        G = np.array([[a, c], [d, b]])
        du = np.linalg.solve(G, np.stack([vxs, vys]))
        recovered.append((np.stack([xs, ys]), du))
    (applied, du0), (applied1, du1) = recovered
    assert np.array_equal(applied, applied1)  # the shift itself ignores v
    assert np.allclose(du0, du1, rtol=0.0, atol=1e-9 * cap)
    at_cap = np.abs(applied) >= cap * (1.0 - 1e-12)
    assert at_cap[:, centre].all()            # the pushed particle, both axes
    assert np.allclose(du0[~at_cap], applied[~at_cap], rtol=0.0, atol=1e-9 * cap)
    assert np.all(np.abs(du0[at_cap]) > (1.0 + 1e-6) * cap)
    assert np.all(np.abs(du0[:, centre]) > 1.1 * cap)
    assert np.array_equal(np.sign(du0[at_cap]), np.sign(applied[at_cap]))


# ---------------------------------------------------------- comp_vel ------
def test_comp_vel_pair_pressure_is_density_weighted_across_an_interface():
    """Two fluids of densities rho_a < rho_b meet at a horizontal interface
    midway between two lattice rows, in hydrostatic balance with a body
    acceleration G: p = p* + rho_k G (y - y*) on each side.  The interface
    pressure that keeps the acceleration continuous across the jump,
    (P - p_i)/rho_i = (p_j - P)/rho_j, is P = (rho_i p_j + rho_j p_i) /
    (rho_i + rho_j) (Hu & Adams' multi-phase pair pressure).  With it, for
    every pair straddling the interface at its midpoint
    2 (p_j - p_i)/(rho_i + rho_j) = G (y_j - y_i), as within each layer; with
    equal volumes the p_i part cancels on a lattice, and the corrected
    gradient of the linear G y is exact, so every interior particle,
    including the two rows at the interface, gets dv = -dt n^2 (0, G).  A
    support of 1.5 dx keeps every crossing pair in adjacent rows.  Fails
    for: p_j rho_j + p_i rho_i in the pair pressure (momentum-conserving as
    well, so a pairwise momentum check misses it), rho_j for rho_i in the
    denominator, a missing n_i^2."""
    nside, hfac = 20, 0.75
    parts = lattice(nside)
    prm = lattice_params(hfac=hfac)
    rho_a, rho_b, G, n, dt = 1000.0, 1250.0, -9.81, 0.6, prm.dt
    jrow = np.arange(parts["x"].size) % nside
    upper = jrow >= nside // 2
    y_star = 0.25 + (nside // 2 - 0.5) * DX
    rho = np.where(upper, rho_b, rho_a)
    parts["density"] = rho
    parts["oden"] = rho / n
    parts["mass"] = rho * DX * DX          # equal volumes m/rho = dx^2
    parts["pressure"] = 2.0e4 + rho * G * (parts["y"] - y_star)
    sim = sim_from(parts, prm)
    sim.neighbour()
    m = sim.count()
    sim.set_field("porosity", np.full(m, n))
    sim.set_field("vxs", np.zeros(m))
    sim.set_field("vys", np.zeros(m))
    sim.comp_vel()
    P = sim.get_particles()
    inner = interior(P, nside, margin=2)
    jr = P["pid"] % nside
    at_interface = inner & ((jr == nside // 2) | (jr == nside // 2 - 1))
    assert at_interface.sum() == 2 * (nside - 4)
    want = -dt * n * n * G
    assert np.allclose(P["vy"][inner], want, rtol=1e-9, atol=0.0)
    assert np.all(np.abs(P["vx"][inner]) <= 1e-9 * abs(want))


def test_comp_vel_pair_forces_conserve_momentum():
    """Three fluid particles on a line (their moment matrices are singular,
    so L = I, functions.f90:75-87) with unequal densities, masses and
    pressures at one porosity: every pair force is equal and opposite, so
    sum m_i dv_i = 0 in both components.  Fails for an asymmetric pair
    pressure (p_j rho_i + p_i rho_i, or rho_i for rho_j in the
    denominator's sum)."""
    prm = lattice_params()
    d = 0.8 * DX
    x = 0.5 + np.array([0.0, d, 2.0 * d]) * 0.6
    y = 0.5 + np.array([0.0, d, 2.0 * d]) * 0.8
    rho = np.array([1000.0, 1100.0, 1250.0])
    parts = dict(x=x, y=y, vx=np.zeros(3), vy=np.zeros(3), mass=rho * DX * DX * [1.0, 0.9, 1.15],
                 density=rho, oden=rho, pressure=np.array([300.0, 1700.0, 650.0]),
                 tid=np.full(3, 3, np.int32), flags=np.zeros(3, np.int32),
                 pid=np.arange(3, dtype=np.int32))
    sim = sim_from(parts, prm)
    sim.neighbour()
    L = sim.get_field("coff").reshape(-1, 4)
    assert np.array_equal(L, np.tile([1.0, 0.0, 0.0, 1.0], (3, 1)))
    sim.set_field("porosity", np.full(3, 0.7))
    sim.set_field("vxs", np.zeros(3))
    sim.set_field("vys", np.zeros(3))
    sim.comp_vel()
    P = sim.get_particles()
    px, py = np.sum(P["mass"] * P["vx"]), np.sum(P["mass"] * P["vy"])
    scale = np.max(np.abs(P["mass"] * P["vx"]))
    assert scale > 0.0
    assert abs(px) <= 1e-13 * scale and abs(py) <= 1e-13 * scale


def test_comp_vel_wall_branch_gives_the_exact_gradient_beside_a_wall():
    """A fluid block over three rows of dummy wall particles, both on one
    lattice with equal volumes, the wall twice as dense and at another
    porosity; the pressure p = p0 + gx x + gy y is linear through both.  For
    two fluid particles at one density the density-weighted pair pressure is
    (p_i + p_j)/2, so with the wall branch V_j n_i^2 (p_i + p_j) grad/rho_i
    every term has the form V_j n_i^2 (p_i + p_j) L grad W / rho_i; the p_i
    part cancels on the lattice and the rest is the exact corrected gradient:
    dv = -dt n^2 (gx, gy) / rho for every fluid particle, the ones beside the
    wall included.  Fails for: the wall's density for the particle's, a
    factor 2 in the wall branch, the wall's porosity for the particle's, the
    wall pressure dropped."""
    nside, nwall = 21, 3
    parts = lattice(nside)
    jrow = np.arange(parts["x"].size) % nside
    wall = jrow < nwall
    rho_f, rho_w, n, gx, gy = 1000.0, 2000.0, 0.6, 850.0, -9810.0
    parts["tid"][wall] = 2
    rho = np.where(wall, rho_w, rho_f)
    parts["density"] = rho
    parts["oden"] = rho
    parts["mass"] = rho * DX * DX
    parts["pressure"] = 3.0e4 + gx * parts["x"] + gy * parts["y"]
    prm = lattice_params()
    sim = sim_from(parts, prm)
    sim.neighbour()
    P = sim.get_particles()
    m = sim.count()
    sim.set_field("porosity", np.where(P["tid"] == 3, n, 1.0))
    sim.set_field("vxs", np.zeros(m))
    sim.set_field("vys", np.zeros(m))
    sim.comp_vel()
    P = sim.get_particles()
    ii, jj = np.divmod(P["pid"], nside)
    fluid = (P["tid"] == 3) & (ii >= 4) & (ii < nside - 4) & (jj < nside - 4)
    beside = fluid & (jj < nwall + 2)         # within 2h of a wall particle
    assert beside.sum() == 2 * (nside - 8)
    dt = prm.dt
    assert np.allclose(P["vx"][fluid], -dt * n * n * gx / rho_f, rtol=1e-9, atol=0.0)
    assert np.allclose(P["vy"][fluid], -dt * n * n * gy / rho_f, rtol=1e-9, atol=0.0)


# -------------------------------------------------------------- drag ------
def porous_block(por):
    """A fluid lattice half inside a soil lattice of the same spacing; the
    fluid deep inside the soil sits at porosity n = por exactly."""
    nside = 25
    parts = lattice(nside)
    sx, sy = np.meshgrid(np.arange(0.25, 0.25 + 12.5 * DX, DX),
                         np.arange(0.25, 0.25 + 25 * DX, DX), indexing="ij")
    soil = dict(x=sx.ravel(), y=sy.ravel(), mass=np.full(sx.size, 1800.0 * DX * DX),
                density=np.full(sx.size, 1800.0))
    return parts, soil


@pytest.mark.parametrize("por", [0.4, 0.55])
def test_drag_update_is_stationary_at_the_ergun_terminal_velocity(por):
    """Uniform flow falling through a porous medium under gravity reaches the
    Ergun terminal (Darcy) velocity U, where the resistance balances the
    weight: alpha nu (1-n)^2 U/(n^3 d50^2) + beta (1-n) U^2/(n^3 d50) = |g|
    (Ergun's law with dp/dl = rho g).  int_vel integrates
    dv/dt = n (g + R v) with the drag implicit, so U is its stationary point
    to first order in dt: one step from v = U changes v by O(dt^2),
    shrinking fourfold when dt halves, and by far less than the dt n g that
    gravity alone adds.  Fails for: 1 - dt R for 1 - dt R n (stationary at
    n |g| instead), gravity without n, the drag explicit."""
    alpha, beta, nu, d50, g = 200.0, 1.1, 1e-6, 0.01, -9.81
    a = alpha * nu * (1.0 - por) ** 2 / (por ** 3 * d50 ** 2)
    b = beta * (1.0 - por) / (por ** 3 * d50)
    U = (-a + math.sqrt(a * a + 4.0 * b * abs(g))) / (2.0 * b)
    deltas = []
    for dt in (1e-4, 5e-5):
        parts, soil = porous_block(por)
        parts["vy"][:] = -U
        prm = lattice_params(drag_law=DRAG_ERGUN, por=por, alpha=alpha, beta=beta, nu=nu,
                             d50=d50, g=g, dt=dt)
        sim = sim_from(parts, prm, soil)
        sim.neighbour()
        sim.effpor()     # nut = tke = 0: no eddy viscosity in a uniform flow
        sim.int_vel()
        P = sim.get_particles()
        x, y = P["x"], P["y"]
        deep = (x > 0.25 + 2.5 * DX) & (x < 0.25 + 9.5 * DX) & (y > 0.25 + 2.5 * DX) & \
            (y < 0.25 + 21.5 * DX)
        assert deep.sum() >= 100
        assert np.array_equal(sim.get_field("porosity")[deep], np.full(deep.sum(), por))
        vys = sim.get_field("vys")[deep]
        assert np.ptp(vys) == 0.0
        assert np.all(sim.get_field("vxs")[deep] == 0.0)
        deltas.append(vys[0] + U)
        assert abs(deltas[-1]) < 0.01 * abs(g) * por * dt
    assert 3.8 < deltas[0] / deltas[1] < 4.2


# ------------------------------------------------------- buffer release ---
def test_comp_pos_releases_buffer_particles_inside_the_window():
    """integrator.f90:888-897: after the position update a live buffer fluid
    particle with xrelease_lo < x < xrelease_hi (strict) stops being a buffer
    particle; one outside the window, on its edge, dead, or not fluid keeps
    its flags.  Positions are set so that the update leaves them in place
    (xs = x, v = 0): x_new = (x + xs + dt v/n)/2 = x exactly."""
    from beach_slope_confined_amd.params import FLAG_BUFFER, FLAG_DEAD
    lo, hi = 0.40625, 0.59375            # binary fractions: x == lo is exact
    x = np.array([0.39, 0.41, 0.5, 0.59, 0.61, lo, hi, 0.5, 0.5, 0.5])
    flags = np.full(x.size, FLAG_BUFFER, np.int32)
    flags[7] = FLAG_BUFFER | FLAG_DEAD   # dead: comp_pos skips it
    flags[8] = 0                         # plain fluid
    tid = np.full(x.size, 3, np.int32)
    tid[9] = 2                           # a wall particle with the flag
    n = x.size
    parts = dict(x=x, y=np.linspace(0.3, 0.7, n), vx=np.zeros(n), vy=np.zeros(n),
                 mass=np.full(n, 1000.0 * DX * DX), density=np.full(n, 1000.0),
                 oden=np.full(n, 1000.0), pressure=np.zeros(n), tid=tid, flags=flags,
                 pid=np.arange(n, dtype=np.int32))
    sim = sim_from(parts, lattice_params(xrelease_lo=lo, xrelease_hi=hi))
    sim.set_field("xs", sim.get_field("x"))
    sim.set_field("ys", sim.get_field("y"))
    sim.comp_pos()
    P = sim.get_particles()
    order = np.argsort(P["pid"])
    assert np.array_equal(P["x"][order], x)
    buf = (P["flags"][order] & FLAG_BUFFER) != 0
    assert buf.tolist() == [True, False, False, False, True, True, True, True, False, True]
