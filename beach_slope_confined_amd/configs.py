"""Deterministic synthetic inputs for the benchmark and test workloads.

Random numbers come from a counter-based splitmix64 (seed stated, default
0x5EED2026): draw ``i`` of stream ``s`` mixes ``seed + (i+1)*phi + s*c``.
Positions use only ``+ - *`` on exactly representable values, so they are
bit-reproducible on any host; velocities use sin/cos and are generated once
and passed to every consumer.

Geometry conventions follow the reference (``boundary.f90:50-83``): walls
are three particle layers outside the fluid region, the innermost layer is
tid 1 (fixed wall), the outer layers tid 2 (dummy wall), fluid is tid 3.
Soil skeleton particles (the reference's tid 4 ``porlist``) fill the porous
wedge and are returned separately.

Workloads
  dam_break(dx)      water column 0.4 x 0.3 m in a closed 1.6 x 0.6 m tank,
                     porous 1:3 wedge with its toe at x = 0.2 m (under the
                     column, so pore water is present from the start; pore
                     water has m = por rho0 dx^2, i.e. V = dx^2 at n = por),
                     Wendland C2, h = 1.3 dx, Ergun drag (por 0.4, d50 0.01,
                     alpha 200, beta 1.1), nu 1e-6, dt 1e-4.
                     dx = 0.01 gives 1,200 fluid particles; dx = 0.002 the
                     adjoint case; dx = 0.0005 the large case.
  dam_break_reference(dx)  the same geometry with the reference's own
                     constants (hybrid kernel, h = 1.27 dx, lamk 450,
                     d50 1.02e-3, por 0.38, vis(oden), adaptive dt).
  tank_1m(nside)     nside^2 fluid particles on a jittered lattice in the
                     unit square with 3 wall layers on the left, right and
                     bottom (free surface on top), velocity = 16 random
                     Fourier modes per component (amplitude 0.1/16 m/s each)
                     plus U(-1e-3, 1e-3), pressure 0.  nside = 1024 gives
                     1,048,576 fluid particles.
"""
import math

import numpy as np

from .params import (DRAG_ERGUN, DRAG_REFERENCE, KERNEL_HYBRID, KERNEL_WENDLAND_C2,
                     VISC_CONSTANT, VISC_REFERENCE, make_params)

SEED = 0x5EED2026
_PHI = np.uint64(0x9E3779B97F4A7C15)
_STREAM = np.uint64(0xD1B54A32D192ED03)


def uniform(n, stream, seed=SEED):
    """n uniforms in [0, 1) from 53 bits of splitmix64."""
    with np.errstate(over="ignore"):
        i = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed) + i * _PHI + np.uint64(stream) * _STREAM
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def _walls(nx, ny, dx, x0, y0, top, layers=3):
    """Wall particles on the lattice cells outside [0,nx)x[0,ny) (cell units),
    `layers` deep, on the left, right, bottom (and top if `top`)."""
    ii, jj = np.meshgrid(np.arange(-layers, nx + layers), np.arange(-layers, ny + (layers if top else 0)),
                         indexing="ij")
    ii = ii.ravel()
    jj = jj.ravel()
    out_x = np.maximum(-1 - ii, ii - nx)       # >= 0 when outside in x
    out_y = np.maximum(-1 - jj, (jj - ny) if top else -1)
    outside = (out_x >= 0) | (out_y >= 0)
    ii, jj = ii[outside], jj[outside]
    layer = np.maximum(np.maximum(-1 - ii, ii - nx), np.maximum(-1 - jj, (jj - ny) if top else -1))
    x = x0 + (ii + 0.5) * dx
    y = y0 + (jj + 0.5) * dx
    tid = np.where(layer == 0, 1, 2).astype(np.int32)
    return x, y, tid


def _assemble(fx, fy, fvx, fvy, wx, wy, wtid, rho0, dx, fp, wp):
    nf, nw = len(fx), len(wx)
    n = nf + nw
    m = rho0 * dx * dx
    parts = dict(
        x=np.concatenate([wx, fx]), y=np.concatenate([wy, fy]),
        vx=np.concatenate([np.zeros(nw), fvx]), vy=np.concatenate([np.zeros(nw), fvy]),
        mass=np.full(n, m), density=np.full(n, rho0), oden=np.full(n, rho0),
        pressure=np.concatenate([wp, fp]),
        tid=np.concatenate([wtid, np.full(nf, 3, np.int32)]).astype(np.int32),
        flags=np.zeros(n, np.int32), pid=np.arange(n, dtype=np.int32),
    )
    return parts


def dam_break(dx=0.01, seed=SEED, jitter=0.05, reference_physics=False, toe=0.2):
    """Dam break onto a porous 1:3 beach in a closed 1.6 x 0.6 m tank."""
    L, H = 1.6, 0.6
    wl, wc = 0.4, 0.3
    nx, ny = int(round(L / dx)), int(round(H / dx))
    nfx, nfy = int(round(wl / dx)), int(round(wc / dx))
    ii, jj = np.meshgrid(np.arange(nfx), np.arange(nfy), indexing="ij")
    ii = ii.ravel().astype(np.float64)
    jj = jj.ravel().astype(np.float64)
    nf = len(ii)
    jit = jitter * dx
    fx = (ii + 0.5) * dx + (2.0 * uniform(nf, 1, seed) - 1.0) * jit
    fy = (jj + 0.5) * dx + (2.0 * uniform(nf, 2, seed) - 1.0) * jit
    wx, wy, wtid = _walls(nx, ny, dx, 0.0, 0.0, top=True)
    rho0, g = 1000.0, -9.81
    fp = rho0 * abs(g) * np.maximum(wc - fy, 0.0)   # hydrostatic (internal.f90:50)
    wp = np.zeros(len(wx))
    parts = _assemble(fx, fy, np.zeros(nf), np.zeros(nf), wx, wy, wtid, rho0, dx, fp, wp)
    por = 0.38 if reference_physics else 0.4
    # pore water carries m = por*rho0*dx^2 so that V = m/(n*oden) = dx^2 where n = por
    in_pores = np.concatenate([np.zeros(len(wx), bool), fy < (fx - toe) / 3.0])
    parts["mass"][in_pores] = por * rho0 * dx * dx
    # soil skeleton lattice under the slope y < (x - toe)/3 (boundary.f90:523-593)
    si, sj = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    sx = (si.ravel() + 0.5) * dx
    sy = (sj.ravel() + 0.5) * dx
    keep = sy < (sx - toe) / 3.0
    sx, sy = sx[keep], sy[keep]
    bulk = 1800.0                               # bulkden, initialize.f90:24
    soil = dict(x=sx, y=sy, mass=np.full(len(sx), bulk * dx * dx), density=np.full(len(sx), bulk))
    m = 3.5 * dx
    if reference_physics:
        h = 1.27 * dx                           # setup.f90:47 with hfac = 5.08
        prm = make_params(kernel=KERNEL_HYBRID, drag_law=DRAG_REFERENCE, visc_law=VISC_REFERENCE,
                          h=h, dl=dx, dt=0.0010000000474974513, por=0.38, alpha=450.0,
                          d50=0.00102, adaptive_dt=1,
                          xmin=-m, ymin=-m, xmax=L + m, ymax=H + m)
    else:
        h = 1.3 * dx
        prm = make_params(kernel=KERNEL_WENDLAND_C2, drag_law=DRAG_ERGUN, visc_law=VISC_CONSTANT,
                          h=h, dl=dx, dt=1e-4, por=0.4, alpha=200.0, beta=1.1, d50=0.01,
                          xmin=-m, ymin=-m, xmax=L + m, ymax=H + m)
    return dict(name=f"dam_break_dx{dx:g}" + ("_ref" if reference_physics else ""),
                parts=parts, soil=soil, params=prm, n_fluid=nf)


def dam_break_reference(dx=0.01, seed=SEED):
    return dam_break(dx=dx, seed=seed, reference_physics=True)


def beach_large(side=2048, dx=0.0005, seed=SEED, jitter=0.05):
    """The large beach run: a square water column of side x side particles
    (4,194,304 at the default) at spacing dx, in a tank 4 column widths long
    and 2 high, with the dam break's 1:3 porous slope (toe at half a column
    width), porosity 0.4, d50 0.01, alpha 200, beta 1.1."""
    w = side * dx
    L, H = 4.0 * w, 2.0 * w
    toe = 0.5 * w
    nx, ny = int(round(L / dx)), int(round(H / dx))
    ii, jj = np.meshgrid(np.arange(side), np.arange(side), indexing="ij")
    ii = ii.ravel().astype(np.float64)
    jj = jj.ravel().astype(np.float64)
    nf = len(ii)
    jit = jitter * dx
    fx = (ii + 0.5) * dx + (2.0 * uniform(nf, 1, seed) - 1.0) * jit
    fy = (jj + 0.5) * dx + (2.0 * uniform(nf, 2, seed) - 1.0) * jit
    wx, wy, wtid = _walls(nx, ny, dx, 0.0, 0.0, top=True)
    rho0, g = 1000.0, -9.81
    fp = rho0 * abs(g) * np.maximum(w - fy, 0.0)
    parts = _assemble(fx, fy, np.zeros(nf), np.zeros(nf), wx, wy, wtid, rho0, dx, fp,
                      np.zeros(len(wx)))
    por = 0.4
    in_pores = np.concatenate([np.zeros(len(wx), bool), fy < (fx - toe) / 3.0])
    parts["mass"][in_pores] = por * rho0 * dx * dx
    si, sj = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    sx = (si.ravel() + 0.5) * dx
    sy = (sj.ravel() + 0.5) * dx
    keep = sy < (sx - toe) / 3.0
    sx, sy = sx[keep], sy[keep]
    soil = dict(x=sx, y=sy, mass=np.full(len(sx), 1800.0 * dx * dx),
                density=np.full(len(sx), 1800.0))
    m = 3.5 * dx
    prm = make_params(kernel=KERNEL_WENDLAND_C2, drag_law=DRAG_ERGUN, visc_law=VISC_CONSTANT,
                      h=1.3 * dx, dl=dx, dt=1e-4, por=por, alpha=200.0, beta=1.1, d50=0.01,
                      xmin=-m, ymin=-m, xmax=L + m, ymax=H + m)
    return dict(name=f"beach_large_{side}x{side}_dx{dx:g}", parts=parts, soil=soil, params=prm,
                n_fluid=nf)


def beach_wide(ncol=8192, nrow=512, dx=0.0005, seed=SEED, jitter=0.05, dt=1.25e-5):
    """The large beach run as a wide, shallow dam break: a water column of
    ncol x nrow particles (4,194,304 at the default) at spacing dx, the dam
    break's proportions about the column's height Hc = nrow dx (tank 4 Hc
    longer than the column and 2 Hc high, the toe of the 1:3 porous slope
    2 Hc / 3 inside the column, so the slope stands 2 Hc / 9 high at the
    column's end, as in the dam break), porosity 0.4, d50 0.01, alpha 200,
    beta 1.1.  A depth of 512 dx keeps H/dx near the dam break's at
    dx = 0.001 (300), the finest that runs 500 steps, where the square
    2048 x 2048 column (H/dx = 2048) leaves the grid in its first steps."""
    w, hc = ncol * dx, nrow * dx
    L, H = w + 4.0 * hc, 2.0 * hc
    toe = w - 2.0 * hc / 3.0
    nx, ny = int(round(L / dx)), int(round(H / dx))
    ii, jj = np.meshgrid(np.arange(ncol), np.arange(nrow), indexing="ij")
    ii = ii.ravel().astype(np.float64)
    jj = jj.ravel().astype(np.float64)
    nf = len(ii)
    jit = jitter * dx
    fx = (ii + 0.5) * dx + (2.0 * uniform(nf, 1, seed) - 1.0) * jit
    fy = (jj + 0.5) * dx + (2.0 * uniform(nf, 2, seed) - 1.0) * jit
    wx, wy, wtid = _walls(nx, ny, dx, 0.0, 0.0, top=True)
    rho0, g = 1000.0, -9.81
    fp = rho0 * abs(g) * np.maximum(hc - fy, 0.0)
    parts = _assemble(fx, fy, np.zeros(nf), np.zeros(nf), wx, wy, wtid, rho0, dx, fp,
                      np.zeros(len(wx)))
    por = 0.4
    in_pores = np.concatenate([np.zeros(len(wx), bool), fy < (fx - toe) / 3.0])
    parts["mass"][in_pores] = por * rho0 * dx * dx
    # soil lattice under the slope only where the slope is (x > toe)
    i0 = int(toe / dx)
    si, sj = np.meshgrid(np.arange(i0, nx), np.arange(ny), indexing="ij")
    sx = (si.ravel() + 0.5) * dx
    sy = (sj.ravel() + 0.5) * dx
    keep = sy < (sx - toe) / 3.0
    sx, sy = sx[keep], sy[keep]
    soil = dict(x=sx, y=sy, mass=np.full(len(sx), 1800.0 * dx * dx),
                density=np.full(len(sx), 1800.0))
    m = 3.5 * dx
    prm = make_params(kernel=KERNEL_WENDLAND_C2, drag_law=DRAG_ERGUN, visc_law=VISC_CONSTANT,
                      h=1.3 * dx, dl=dx, dt=dt, por=por, alpha=200.0, beta=1.1, d50=0.01,
                      xmin=-m, ymin=-m, xmax=L + m, ymax=H + m)
    return dict(name=f"beach_wide_{ncol}x{nrow}_dx{dx:g}", parts=parts, soil=soil, params=prm,
                n_fluid=nf, toe=toe, column=(w, hc))


def fourier_velocity(x, y, seed=SEED, modes=16, amp=0.1, noise=1e-3):
    """vx = sum_k (amp/modes) cos(2 pi (kx x + ky y) + phi_k), vy likewise with
    sin, integer wavevectors in [1, 8]^2, plus U(-noise, noise)."""
    n = len(x)
    vx = np.zeros(n)
    vy = np.zeros(n)
    a = amp / modes
    kk = np.floor(uniform(4 * modes, 10, seed) * 8.0).astype(np.int64) + 1
    ph = uniform(2 * modes, 11, seed) * (2.0 * math.pi)
    for k in range(modes):
        vx += a * np.cos(2.0 * math.pi * (kk[4 * k] * x + kk[4 * k + 1] * y) + ph[2 * k])
        vy += a * np.sin(2.0 * math.pi * (kk[4 * k + 2] * x + kk[4 * k + 3] * y) + ph[2 * k + 1])
    vx += (2.0 * uniform(n, 12, seed) - 1.0) * noise
    vy += (2.0 * uniform(n, 13, seed) - 1.0) * noise
    return vx, vy


def tank_1m(nside=1024, seed=SEED, jitter=0.1):
    """nside^2 jittered fluid particles in the unit square, walls on three sides."""
    dx = 1.0 / nside
    ii, jj = np.meshgrid(np.arange(nside), np.arange(nside), indexing="ij")
    ii = ii.ravel().astype(np.float64)
    jj = jj.ravel().astype(np.float64)
    nf = len(ii)
    jit = jitter * dx
    fx = (ii + 0.5) * dx + (2.0 * uniform(nf, 1, seed) - 1.0) * jit
    fy = (jj + 0.5) * dx + (2.0 * uniform(nf, 2, seed) - 1.0) * jit
    fvx, fvy = fourier_velocity(fx, fy, seed)
    wx, wy, wtid = _walls(nside, nside, dx, 0.0, 0.0, top=False)
    rho0 = 1000.0
    parts = _assemble(fx, fy, fvx, fvy, wx, wy, wtid, rho0, dx, np.zeros(nf), np.zeros(len(wx)))
    m = 3.5 * dx
    prm = make_params(kernel=KERNEL_WENDLAND_C2, drag_law=DRAG_ERGUN, visc_law=VISC_CONSTANT,
                      h=1.3 * dx, dl=dx, dt=1e-4, xmin=-m, ymin=-m, xmax=1.0 + m, ymax=1.25)
    soil = dict(x=np.zeros(0), y=np.zeros(0), mass=np.zeros(0), density=np.zeros(0))
    return dict(name=f"tank_{nside}x{nside}", parts=parts, soil=soil, params=prm, n_fluid=nf)


def with_reference_cells(cfg):
    """The config on the reference's cell grid and particle order: cells of
    4h (incr * h1, setup.f90:120-123) with boundaries at multiples of 4h,
    the lowest cells starting at -4h as the reference's j = 1 and i = celly,
    and cell_order = ORDER_REFERENCE."""
    from .params import ORDER_REFERENCE
    p = cfg["params"]
    cs = 4.0 * p.h
    out = dict(cfg)
    out["params"] = p.copy(cell_size=cs, xmin=-cs, ymin=-cs, cell_order=ORDER_REFERENCE)
    out["name"] = cfg["name"] + "_refcells"
    return out


def make_sim(cfg, sim_cls, cap_extra=0, **kw):
    """Create a context of `sim_cls` loaded with a config and initialised
    (effpor + timestep, main.f90:35-37)."""
    parts, soil = cfg["parts"], cfg["soil"]
    n = len(parts["x"])
    sim = sim_cls(cfg["params"], n + cap_extra, max(len(soil["x"]), 1), **kw)
    sim.set_particles(parts)
    sim.set_soil(soil)
    sim.init_state()
    return sim
