"""The reference's own scenario: geometry, buffers and the three-phase run.

`reference_setup()` restates the derived constants of init() (setup.f90:10-136)
from the globals of initialize.f90:19-56; `reference_scenario()` restates the
particle placement of gen_boundary (boundary.f90:12-605) and gen_fluid
(internal.f90:10-473): wall segments, the soil lattice under the slope, the
fluid column, the coastal fluid, and the two buffer zones with their insertion
templates, entry cells and schedule (ins_1, ins_2, rv_buf_r, rv_buf_l).
Expressions keep Fortran's evaluation order (left to right, integer products
before their conversion), so positions, pressures and constants come out with
the reference's bits.

Particles reach the library in the order gen_boundary/gen_fluid place them:
segment after segment, each in its placement-loop order.  A stable binning by
(cell, walls before fluid) of that sequence is the reference's initial
per-cell list (dpcell(i,j)%plist), so with cell_order = ORDER_REFERENCE the
library starts from the reference's particle order.  Particles the reference
never places (outside the cells j = 2..cellx-1, i = 2..celly-1, or above the
slope line) are left out here too.

`run_reference()` drives a context through main.f90's phases: the start-up
steps (:41-66), the buffered steps (:94-147) and the salt-water steps
(:198-263), with step counts chosen by the caller.
"""
import math

import numpy as np

from .params import (DRAG_REFERENCE, FLAG_BUFFER, FLAG_DEAD, KERNEL_HYBRID, ORDER_REFERENCE,
                     VISC_REFERENCE, make_params)

# initialize.f90:19-56 (the double-precision globals; default-real literals
# such as 2.0 and 0.5 in the formulas below are exact)
DEFAULTS = dict(wc=0.650, L=3.11, H=0.80, pr=0.0050 * 0.90, br=0.00490 * 0.90, wl=3.10,
                rho=1000.0, g=-9.810, soill=3.230, soilh=0.80, set_ht=0.10, coastal_ht=0.550,
                por=0.380, bulkden=1800.0, entry_vel=-2.5 / (3600 * 24), lhs_btm=0.20,
                rel_den=1.0250, dt=0.001, hfac=5.080, line_grad=15.0, distfac=1.0,
                bl=3, open_lhs=43, incr=4)


def _floor(v):
    return int(math.floor(v))


def _ceil(v):
    return int(math.ceil(v))


def reference_setup(**over):
    """Derived constants of init() (setup.f90:17-136) and gen_fluid's schedule
    constants (internal.f90:19-32)."""
    c = dict(DEFAULTS)
    c.update(over)
    wc, wl, pr, br, por = c["wc"], c["wl"], c["pr"], c["br"], c["por"]
    bl, incr = c["bl"], c["incr"]
    s = dict(c)
    fpy = _floor(wc / (2 * pr)) + 1
    fpx = _floor(wl / (2 * pr))
    prrealx = (wl / fpx) / 2.0
    prrealy = (wc / (fpy - 1)) / 2.0
    fmass = ((wc * wl * 1.0) * c["rho"]) / (fpx * fpy)
    spy = _floor(c["soilh"] / (2 * prrealy)) + 4
    spx = _floor(c["soill"] / (2 * prrealx))
    solidx = (c["soill"] / spx) / 2.0
    solidy = (c["soilh"] / (spy - 1)) / 2.0
    L = c["L"]
    H = c["H"] + 2 * prrealy
    bny = _floor(H / (2 * br))
    bnx = _floor(L / (2 * br))
    brrealx = (L / bnx) / 2.0
    brrealy = (H / bny) / 2.0
    h1 = (max(prrealx, prrealy) * c["hfac"]) / (2.0 * math.sqrt(por))
    cellx = _ceil(L / (incr * h1)) + 4
    celly = _ceil(H / (incr * h1)) + 3
    dl = min(2 * prrealx, 2 * prrealy)
    bny = _floor(c["lhs_btm"] / (2 * brrealy))
    domain_shift = 5 * prrealx
    bnx = bnx + 2 * bl
    bny = bny + 2 * bl
    xl = (2 * bl - 1) * brrealx + brrealx
    yl = (2 * bl - 1) * brrealy + solidy + c["set_ht"]
    line_grad = (c["line_grad"] * 22) / (180.0 * 7)
    fpy = _floor(wc / (2 * prrealy / math.sqrt(por)))
    fpx = _floor(wl / (2 * prrealx / math.sqrt(por)))
    lam = 0.0010 * (h1 * h1)
    xlcutoff = (brrealx * (2 * bl - 1)) + brrealx
    s.update(prrealx=prrealx, prrealy=prrealy, fmass=fmass, spy=spy, spx=spx, solidx=solidx,
             solidy=solidy, H=H, bny=bny, bnx=bnx, brrealx=brrealx, brrealy=brrealy, h1=h1,
             cellx=cellx, celly=celly, dl=dl, domain_shift=domain_shift, xl=xl, yl=yl,
             line_grad=line_grad, fpy=fpy, fpx=fpx, lam=lam, xlcutoff=xlcutoff,
             cs=incr * h1)
    # gen_fluid's schedule constants (internal.f90:19-32)
    sq = math.sqrt(por)
    bounlen = fpy * 2 * prrealy
    bounlen2 = c["open_lhs"] * 2 * prrealy
    ev = c["entry_vel"]
    s.update(dead_ht=_ceil(0.3 / (2 * prrealy / sq)), bounlen=bounlen, bounlen2=bounlen2,
             ins_1=(4 * prrealx / sq) / abs(ev / 0.387),
             ins_2=4 * prrealx / ((bounlen / bounlen2) * abs(ev)),
             rv_buf_r=(brrealx * (2 * bl - 1)) + (fpx - 1) * 2 * prrealx / sq + prrealx / sq
             + domain_shift + 3 * 2 * prrealx / sq + 0.5 * prrealx / sq,
             rv_buf_l=2 * prrealx + brrealx + 0.5 * prrealx)
    # the buffer release window of comp_pos (integrator.f90:888-889), with
    # xrcutoff from setup.f90:76 (fpx of :72, domain_shift of :56)
    s.update(xrcutoff=(((brrealx * (2 * bl - 1)) + ((s["fpx"] - 1) * 2) * prrealx / sq)
                       + (3 * prrealx) / sq + domain_shift) - 0.001,
             xrelease_lo=((8 * prrealx) + brrealx) + 0.001)
    return s


def _cell(s, x, y):
    """(i, j) of the reference cell holding (x, y) by its bound tests
    (setup.f90:120-123, placement loops of boundary.f90/internal.f90), or
    None outside the cells j = 2..cellx-1, i = 2..celly-1."""
    cs = s["incr"] * s["h1"]
    j0 = _floor(x / cs) + 2
    for j in (j0 - 1, j0, j0 + 1):
        if not 2 <= j <= s["cellx"] - 1:
            continue
        if (j - 2) * cs <= x < (j - 1) * cs:
            i0 = s["celly"] - 1 - _floor(y / cs)
            for i in (i0 - 1, i0, i0 + 1):
                if 2 <= i <= s["celly"] - 1 and (s["celly"] - 1 - i) * cs <= y < (s["celly"] - i) * cs:
                    return i, j
    return None


class _Acc:
    def __init__(self):
        self.rows = []

    def add(self, x, y, tid, mass, dens, oden=None, vx=0.0, vy=0.0, p=0.0, flags=0):
        self.rows.append((x, y, vx, vy, mass, dens, dens if oden is None else oden, p, tid, flags))


def reference_scenario(**over):
    """Particles, soil, buffer zones and parameters of the reference run."""
    s = reference_setup(**over)
    bl, sq = s["bl"], math.sqrt(s["por"])
    brx, bry, prx, pry = s["brrealx"], s["brrealy"], s["prrealx"], s["prrealy"]
    ds, fmass, rho = s["domain_shift"], s["fmass"], s["rho"]
    walls, fluid = _Acc(), _Acc()

    def place(acc, pts):
        # pts in placement-loop order; cells outside j = 2..cellx-1,
        # i = 2..celly-1 receive nothing
        cells = []
        for pt in pts:
            ij = _cell(s, pt[0], pt[1])
            if ij is not None:
                acc.add(*pt)
                cells.append(ij)
        return cells

    # left wall, lower part (boundary.f90:23-80, placed :130-166)
    bny = s["bny"]
    pts = []
    for m in range(1, bl + 1):
        for k in range(1, bny + 1):
            tid = 1 if (m == bl and k <= bny - bl + 1) else 2
            pts.append(((bny - k) * bry * 2.0, tid))
    place(walls, [((m - 1) * brx * 2 + ds, y, t, fmass, rho) for m in range(1, bl + 1)
                  for (y, t) in pts[(m - 1) * bny:m * bny]])
    # left wall above the inflow opening (boundary.f90:167-205)
    remain_len = s["wc"] - s["lhs_btm"] - s["open_lhs"] * 2 * pry
    bn1 = _floor(remain_len / (2 * bry)) - 2
    place(walls, [((m - 1) * brx * 2 + ds,
                   (bny * 2 * bry) + (s["open_lhs"] * 2 * pry) + (bn1 - k) * bry * 2,
                   1 if m == bl else 2, fmass, rho)
                  for m in range(1, bl + 1) for k in range(1, bn1 + 1)])
    # bottom (boundary.f90:206-250)
    nb = s["bnx"] - 2 * bl + 7
    place(walls, [((2 * bl - 1) * brx + brx + (m - 1) * brx * 2 + ds, (bl - k) * 2 * bry,
                   1 if k == 1 else 2, fmass, rho)
                  for m in range(1, nb + 1) for k in range(1, bl + 1)])
    # right wall above the outflow buffer (boundary.f90:251-300)
    buy1 = s["fpy"] - 1
    place(walls, [(((brx * ((2 * bl) - 1)) + (s["fpx"] - 1) * 2 * prx / sq + 3 * prx / sq)
                   + (m - 1) * 2 * brx + ds,
                   ((bry * ((2 * bl) - 1)) + buy1 * 2 * pry / sq + 2 * pry / sq)
                   + bry + (20 - k) * 2 * bry,
                   1 if m == 1 else 2, fmass, rho)
                  for m in range(1, bl + 1) for k in range(1, 21)])
    # soil skeleton under the slope (boundary.f90:530-598)
    spx, spy, sx_, sy_ = s["spx"], s["spy"], s["solidx"], s["solidy"]
    soil = []
    for m in range(1, spx + 1):
        for k in range(1, spy + 1):
            y = (spy - k) * 2 * sy_ + sy_ + (bry * ((2 * bl) - 1)) - 6 * sy_
            x = (brx * ((2 * bl) - 1)) + (m - 1) * 2 * sx_ + brx - 6 * sx_ + ds
            if ((y - s["yl"]) - s["line_grad"] * (x - s["xl"])) > 0.0:
                continue
            if _cell(s, x, y) is not None:
                soil.append((x, y))
    smass = s["soill"] * s["soilh"] * s["bulkden"] / (spx * spy)

    # fluid column (internal.f90:34-108)
    fpx, fpy, dead_ht, ev, g, wc = s["fpx"], s["fpy"], s["dead_ht"], s["entry_vel"], s["g"], s["wc"]
    ptop = (bry * s["distfac"]) * ((2 * bl) - 1)
    col = {}
    for j in range(1, fpx + 1):
        for dcount, i in enumerate(range(fpy, 0, -1), start=1):
            y = (bry * ((2 * bl) - 1)) + (fpy - i) * 2 * pry / sq + pry / sq
            x = (brx * ((2 * bl) - 1)) + (j - 1) * 2 * prx / sq + prx / sq + ds
            vx = ev * fpy / fpy
            fl = 0
            if dead_ht < dcount <= dead_ht + 2:
                fl, vx = FLAG_DEAD, 0.0
            p = (-y + ptop + wc + pry / sq) * rho * abs(g)
            tid = 0 if ((y - s["yl"]) - s["line_grad"] * (x - s["xl"])) > 0.0 else 3
            col[i, j] = (x, y, tid, fmass, rho, rho, vx, 0.0, p, fl)
    place(fluid, [col[k, l1] for l1 in range(1, fpx + 1) for k in range(1, fpy + 1)
                  if col[k, l1][2] != 0])
    # outflow (right) buffer and its template buffer1 (internal.f90:115-249)
    bfy1 = fpy
    rb, templ1 = {}, []
    for j in range(1, 6):
        for dcount, i in enumerate(range(bfy1, 0, -1), start=1):
            y = (bry * ((2 * bl) - 1)) + (bfy1 - i) * 2 * pry / sq + pry / sq
            x = (brx * ((2 * bl) - 1)) + (fpx - 1) * 2 * prx / sq + prx / sq + ds + j * 2 * prx / sq
            vx, fl = ev * fpy / (fpy - 2), FLAG_BUFFER
            if dead_ht < dcount <= dead_ht + 2:
                fl, vx = FLAG_BUFFER | FLAG_DEAD, 0.0
            p = (-y + ptop + wc + pry / sq) * rho * abs(g)
            rb[i, j] = (x, y, 3, fmass, rho, rho, vx, 0.0, p, fl)
    for jj in (4, 5):       # buffer1(:, 1:2) = flist(:, 4:5), same construction
        for i in range(1, bfy1 + 1):
            templ1.append(rb[i, jj])
    cells1 = place(fluid, [rb[k, l1] for l1 in range(1, 6) for k in range(1, bfy1 + 1)])
    entry1 = sorted(set(cells1), key=lambda ij: (ij[1], ij[0]))
    # coastal fluid on the slope (internal.f90:253-318)
    fpy2 = _floor(s["coastal_ht"] / (2 * pry)) + 1
    fpx2 = _floor(2.0 / (2 * prx))
    cf = {}
    for j in range(1, fpx2 + 1):
        for i in range(fpy2, 0, -1):
            y = (bry * ((2 * bl) - 1)) + (fpy2 - i) * 2 * pry + pry + s["set_ht"]
            x = (brx * ((2 * bl) - 1)) + (j - 1) * 2 * prx + prx + ds
            p = 0.0 if i == 1 else (-y + ptop + s["coastal_ht"] + pry + s["set_ht"]) * rho * abs(g)
            if ((y - s["yl"] + pry) - s["line_grad"] * (x - s["xl"])) < 0.0:
                tid = 0
            else:
                tid, y = 3, y + 2 * s["dl"]
            cf[i, j] = (x, y, tid, fmass * 1.0, rho * 1.0, rho * 1.0, 0.0, 0.0, p, 0)
    place(fluid, [cf[k, l1] for l1 in range(1, fpx2 + 1) for k in range(1, fpy2 + 1)
                  if cf[k, l1][2] != 0])
    # inflow (left) buffer and its template buffer2 (internal.f90:320-470)
    ol = s["open_lhs"]
    lb, templ2 = {}, []
    for j in range(1, 6):
        for i in range(ol, 0, -1):
            y = 2 * s["bny"] * bry - bry + pry + (ol - i) * 2 * pry
            x = (j - 1) * 2 * prx + brx
            vx = (-s["bounlen"] / s["bounlen2"]) * ev
            p = (-y + ptop + wc + pry) * rho * abs(g)
            lb[i, j] = (x, y, 3, fmass * 1.0, rho * 1.0, rho * 1.0, vx, 0.0, p, FLAG_BUFFER)
    for jj in (1, 2):
        for i in range(1, ol + 1):
            t = lb[i, jj]
            templ2.append(t[:8] + (t[8] * s["rel_den"], t[9]))
    cells2 = place(fluid, [lb[k, l1] for l1 in range(1, 6) for k in range(1, ol + 1)])
    entry2 = sorted({ij for ij in cells2 if ij[1] <= s["cellx"] // 2},
                    key=lambda ij: (ij[1], ij[0]))

    rows = walls.rows + fluid.rows
    n = len(rows)
    a = np.array([r[:8] for r in rows], dtype=np.float64)
    parts = dict(x=a[:, 0], y=a[:, 1], vx=a[:, 2], vy=a[:, 3], mass=a[:, 4], density=a[:, 5],
                 oden=a[:, 6], pressure=a[:, 7],
                 tid=np.array([r[8] for r in rows], np.int32),
                 flags=np.array([r[9] for r in rows], np.int32),
                 pid=np.arange(n, dtype=np.int32))
    sa = np.array(soil, dtype=np.float64).reshape(-1, 2)
    soil_d = dict(x=sa[:, 0].copy(), y=sa[:, 1].copy(), mass=np.full(len(sa), smass),
                  density=np.full(len(sa), s["bulkden"]))
    cs = s["cs"]
    prm = make_params(kernel=KERNEL_HYBRID, drag_law=DRAG_REFERENCE, visc_law=VISC_REFERENCE,
                      h=s["h1"], dl=s["dl"], lam=s["lam"], dt=s["dt"], rho0=rho,
                      rel_den=s["rel_den"], g=s["g"], por=s["por"], alpha=450.0, d50=0.00102,
                      adaptive_dt=1, cell_size=cs, xmin=-cs, ymin=-cs,
                      xmax=(s["cellx"] - 1) * cs - 1e-12, ymax=(s["celly"] - 1) * cs - 1e-12,
                      cell_order=ORDER_REFERENCE, sink=0, xlcutoff=s["xlcutoff"],
                      salt_xbuf=1.6, rhomin=1000.0, tschmidt=1.0, max_nbr=96,
                      xrelease_lo=s["xrelease_lo"], xrelease_hi=s["xrcutoff"])

    def zone(templ, entry, period, xcut, remove_right, con):
        # template rows are (x, y, tid, mass, density, oden, vx, vy, pressure, flags);
        # the library's cell of reference cell (i, j) is cx = j, cy = celly + 1 - i
        # when xmin = ymin = -cs (cx = floor(x/cs) + 2, cy = floor(y/cs) + 2)
        col = lambda k, dt=np.float64: np.array([r[k] for r in templ], dtype=dt)  # noqa: E731
        return dict(x=col(0), y=col(1), vx=col(6), vy=col(7), pressure=col(8),
                    flags=col(9, np.int32),
                    cells=np.array([(j, s["celly"] + 1 - i) for (i, j) in entry], np.int32),
                    period=period, xcut=xcut, remove_right=remove_right, con=con,
                    mass=fmass, density=rho, ovol=fmass / rho)

    zones = [zone(templ1, entry1, s["ins_1"], s["rv_buf_r"], 1, -0.5),
             zone(templ2, entry2, s["ins_2"], s["rv_buf_l"], 0, 0.5)]
    return dict(name="reference_scenario", parts=parts, soil=soil_d, params=prm,
                n_fluid=int((parts["tid"] == 3).sum()), setup=s, zones=zones)


def phase_zones(sc, relden=1.0, time_shift=0.0):
    """The two zones as main.f90 inserts them: zone 0 (insert_buffer1, fresh
    water), zone 1 (insert_buffer2(relden): mass fmass relden, density
    rho relden, con 0.5)."""
    z0, z1 = dict(sc["zones"][0]), dict(sc["zones"][1])
    z1["mass"] = z1["mass"] * relden
    z1["density"] = z1["density"] * relden
    return [z0, z1]


def run_reference(sim, sc, steps1=50, steps2=0, steps3=0, output=None, every=None,
                  on_phase=None):
    """main.f90 on a context loaded with reference_scenario(): effpor
    (:35-37); `steps1` start-up steps (:39-67); fluid velocities zeroed
    (:69-91); `steps2` steps with the buffer zones (:96-147); the salt-water
    start (:149-192) and `steps3` salt-water steps with the zones, t
    restarted from 0 (:194-263).  `output(sim, k)` is called every `every`
    steps of the last phase (the reference's combined(), :266-269);
    `on_phase(sim, name)` after each phase ("start", "buffered", "salt")."""
    s = sc["setup"]
    hook = on_phase or (lambda sim_, name: None)
    sim.effpor()
    sim.step(steps1)
    hook(sim, "start")
    q = sim.get_particles()
    still = (q["tid"] == 3) & ((q["flags"] & FLAG_BUFFER) == 0)
    sim.set_field("vx", np.where(still, 0.0, q["vx"]))
    sim.set_field("vy", np.where(still, 0.0, q["vy"]))
    sim.set_buffer_zones(phase_zones(sc, 1.0), 0.0)
    sim.step(steps2)
    hook(sim, "buffered")
    sim.salt_init(s["xl"], s["yl"], s["line_grad"])
    sc_ = sim.scalars()
    sim.set_params(sim.params.copy(salinity=1))
    sim.set_buffer_zones(phase_zones(sc, s["rel_den"]), sc_["t"])
    sim.set_scalars(sc_["umax"], sc_["dt"], 0.0)
    for k in range(steps3):
        sim.step(1)
        if output is not None and every and (k + 1) % every == 0:
            output(sim, (k + 1) // every)
    hook(sim, "salt")
    return sim
