#!/usr/bin/env python3
"""ISPH step throughput on one MI355X.

Workload (the single-step-at-scale configuration): 1,048,576 fluid
particles on a jittered 1024x1024 lattice in the unit square with three
wall layers on the left, right and bottom, 16-mode Fourier velocity field,
Wendland C2 kernel, h = 1.3 dx, BiCGSTAB to a true relative residual of
1e-8.  A step is the full ISPH step of the reference (main.f90:41-66):
projection, re-binning, neighbour search + KGC, porous drag, eddy
viscosity, v*, PPE assembly, PPE solve, pressure-gradient correction,
position update, re-binning, neighbour search, free surface, porous drag,
shifting, timestep reduction.

Prints one JSON line.  `value` is fluid particle-steps/s of the FP64 step;
the mixed-precision PPE step on the same input is reported beside it
(`mixed`), with its pressure and velocity error against the FP64 step.
`roofline` is the dominant kernel (a BiCGSTAB SpMV kernel) timed with HIP
events on the library's stream during the timed steps; `cpu_baseline` is
the CPU oracle (this repository's C restatement of the reference step,
OpenMP) on the host cores of the same machine.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "ISPH particle-steps/s @1M (FP64 vs mixed-prec PPE); adjoint/forward cost"
HBM_PEAK_GBS = 8000.0
PROFILE_STRIDE = 16  # HIP events around every 16th SpMV launch of the timed steps


def log(msg):
    """Progress on stderr (stdout carries only the result line)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nside", type=int, default=1024)
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-salinity", action="store_true")
    ap.add_argument("--salinity-only", action="store_true",
                    help="run only the salt-water step leg (for profiling)")
    ap.add_argument("--salt-dx", type=float, default=0.001)
    ap.add_argument("--no-reference-run", action="store_true")
    ap.add_argument("--no-large", action="store_true",
                    help="skip the 4M-particle beach run")
    ap.add_argument("--large-steps", type=int, default=500,
                    help="steps of the 4M-particle beach run")
    ap.add_argument("--no-large-adjoint", dest="large_adjoint", action="store_false",
                    help="skip the adjoint of the 4M-particle beach run")
    ap.add_argument("--mg-only", action="store_true",
                    help="time only the multigrid-preconditioned step (for profiling)")
    ap.add_argument("--fp64mg-only", action="store_true",
                    help="time only the FP64-valued multigrid step (for profiling)")
    ap.add_argument("--fgmres-only", action="store_true",
                    help="time only the step with the reference's FGMRES solve (for profiling)")
    ap.add_argument("--reference-run-only", action="store_true")
    ap.add_argument("--adjoint-dx", type=float, default=0.002)
    ap.add_argument("--adjoint-steps", type=int, default=50)
    ap.add_argument("--cpu-runs", type=int, default=3,
                    help="runs of each CPU-oracle workload (the median is reported)")
    ap.add_argument("--pmc-run", action="store_true",
                    help="short run for hardware-counter profiles: one step whose FP64 solve is "
                         "capped at 40 iterations, then 5 copies of 1 GiB with the library's "
                         "copy kernel (counter calibration)")
    return ap.parse_args()


def model_spmv_bytes(n, nnz_off):
    """Algorithmic bytes of one FP64 SpMV in the per-row model of the ISPH
    step: 32 + 12 K bytes per row (int32 column + FP64 value per entry; row
    pointer, x and y), K = off-diagonal entries per row."""
    return 32 * n + 12 * nnz_off


def model_step_bytes(n, nnz_off, iters):
    """Algorithmic bytes of one FP64 step in the same model: per row the
    fixed stages 1284 + 60 K and per BiCGSTAB iteration 248 + 24 K (two
    SpMVs, five dots, a norm, three axpys)."""
    k = nnz_off / n
    return n * ((1284 + 60 * k) + iters * (248 + 24 * k))


def kernel_bytes(kind, nnz_off, n, val_bytes, vec_bytes, idx_bytes=4):
    """Algorithmic bytes of one launch of a BiCGSTAB SpMV kernel: every
    neighbour index (4-byte column or 2-byte offset) and matrix value once,
    per row cnt and dpos, each gathered vector counted once per row, own-row
    reads and writes once."""
    idx = idx_bytes * nnz_off + 8 * n    # nbr + cnt + dpos
    val = val_bytes * nnz_off
    if kind == "p_spmv":                 # gathers (r, q); reads rhat; writes p and (r, v)
        vec = vec_bytes * (2 + 1 + 1 + 2) * n
    else:                                # gathers r, v; writes s, t
        vec = vec_bytes * (2 + 2) * n
    return idx + val + vec


def run_gpu(cfg, steps, warmup, precision, profile, tol=None, maxit=200000, lists=True):
    from beach_slope_confined_amd import capi, configs
    from beach_slope_confined_amd.params import PREC_FP64
    c = dict(cfg)
    c["params"] = cfg["params"].copy(precision=precision, tol=tol or cfg["params"].tol,
                                     maxit=maxit)
    log(f"{cfg['name']}: precision {precision}, {warmup} + {steps} steps")
    sim = configs.make_sim(c, capi.IsphSim)
    # every step starts from the same device-resident input (the workload is
    # one ISPH step at scale; the snapshot copy is device-to-device)
    sim.checkpoint(0)
    iters = []
    first = None
    for _ in range(warmup):
        sim.restore(0)
        sim.step(1)
        iters.append(sim.scalars()["iters"])
        if first is None:
            first = sim.get_particles()
    if profile:
        sim.profile_enable(PROFILE_STRIDE)
    import torch
    # each step is timed on its own, bracketed by device synchronisations;
    # the restore of the input (a device-to-device copy, synchronous) lies
    # outside the timed region and is reported apart
    seconds = restore_s = 0.0
    for _ in range(steps):
        r0 = time.perf_counter()
        sim.restore(0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sim.step(1)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        seconds += t1 - t0
        restore_s += t0 - r0
        iters.append(sim.scalars()["iters"])
    s = sim.scalars()
    out = dict(sim=sim, seconds=seconds, restore_s=restore_s, iters=iters, relres=s["relres"],
               outer=s["outer"], first=first)
    if lists:
        ptr, idx = sim.get_neighbours()
        rows = np.repeat(np.arange(len(ptr) - 1), np.diff(ptr))
        # the solver reads 16-bit column offsets when every offset fits
        narrow = bool(len(idx) == 0 or np.abs(idx.astype(np.int64) - rows).max() <= 32767)
        out.update(nnz_off=int(ptr[-1]), n=len(ptr) - 1, idx_bytes=2 if narrow else 4)
    if profile:
        k0, k1 = (0, 1) if precision == PREC_FP64 else (2, 3)
        out["ktime"] = {"p_spmv": sim.kernel_time(k0), "s_spmv": sim.kernel_time(k1)}
    return out


def step_errors(a, b):
    """max |p_b - p_a| / max |p_a| and the velocity error over the fluid,
    relative to the largest fluid speed of a"""
    if not np.array_equal(a["pid"], b["pid"]):  # align by particle id
        ia, ib = np.argsort(a["pid"]), np.argsort(b["pid"])
        a = {k: v[ia] for k, v in a.items()}
        b = {k: v[ib] for k, v in b.items()}
    f = a["tid"] == 3
    perr = float(np.abs(b["pressure"] - a["pressure"]).max() / np.abs(a["pressure"]).max())
    vmag = np.hypot(a["vx"][f], a["vy"][f]).max()
    verr = float(max(np.abs(b["vx"] - a["vx"])[f].max(), np.abs(b["vy"] - a["vy"])[f].max()) / vmag)
    return perr, verr


def scalart_bytes(ptr, tid):
    """Algorithmic bytes of one scalart launch: per live fluid row its own
    state (tid, flags, con, (Dx, Dy), n, nut, ovol, position, search
    position, cnt: 96 B), its list (4 B per neighbour) and the cdiff write
    (8 B); of the other particles tid and flags (8 B).  Neighbour data is
    counted once, as the particle's own."""
    cnt = np.diff(ptr)
    fl = tid == 3
    return int(104 * fl.sum() + 4 * cnt[fl].sum() + 8 * (~fl).sum())


SALT_DT = 2.5e-5     # the dam break at dx = 0.001 is stable at this dt (not at 1e-4)
SALT_WARMUP, SALT_STEPS = 2, 10


def run_salinity(dx, steps=SALT_STEPS, warmup=SALT_WARMUP, dt=SALT_DT):
    """The salt-water step (main.f90:199-263) on the porous dam break at
    spacing dx and time step dt: salt water above the line
    y - 0.1 = 0.3 (x - 0.2) in the column, fresh below; transport substeps
    on the previous step's lists, densityupdate, the lagged eddyvis.  Steps
    run back to back (the transport reads the lists of the previous step),
    so their number is fixed here, independent of --steps: the run must
    stay physical for all of them."""
    import torch
    from beach_slope_confined_amd import capi, configs
    cfg = configs.dam_break(dx)
    cfg["params"] = cfg["params"].copy(salinity=1, dt=dt)
    sim = configs.make_sim(cfg, capi.IsphSim)
    sim.salt_init(0.2, 0.1, 0.3)
    for _ in range(max(warmup, 1)):   # the first step builds the lists
        sim.step(1)
    sim.profile_enable(1)
    sub = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        sub.append(sim.substeps()[0])
        sim.step(1)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    iters = sim.scalars()["iters"]
    launches, tot = sim.kernel_time(4)
    ptr, _ = sim.get_neighbours()
    tid = sim.get_particles()["tid"]
    nb = scalart_bytes(ptr, tid)
    avg_s = tot / max(launches, 1) / 1e3
    gbs = nb / avg_s / 1e9 if launches else 0.0
    con = sim.get_field("con")[tid == 3]
    pid = sim.get_particles()["pid"][tid == 3]
    sim.close()
    n_fluid = cfg["n_fluid"]
    n = len(cfg["parts"]["x"])
    # the same steps with the multigrid-preconditioned PPE (to 1e-9), no
    # profiling events; con compared particle by particle with the FP64 run
    from beach_slope_confined_amd.params import PREC_MIXED_MG
    cmg = configs.dam_break(dx)
    cmg["params"] = cmg["params"].copy(salinity=1, precision=PREC_MIXED_MG, tol=1e-9, dt=dt)
    sim = configs.make_sim(cmg, capi.IsphSim)
    sim.salt_init(0.2, 0.1, 0.3)
    for _ in range(max(warmup, 1)):
        sim.step(1)
    torch.cuda.synchronize()
    m0 = time.perf_counter()
    for _ in range(steps):
        sim.step(1)
    torch.cuda.synchronize()
    m1 = time.perf_counter()
    pm = sim.get_particles()
    fm = pm["tid"] == 3
    con_mg = dict(zip(pm["pid"][fm].tolist(), sim.get_field("con")[fm].tolist()))
    dcon = max(abs(con_mg[k] - v) for k, v in zip(pid.tolist(), con.tolist()))
    mg = {"ms_per_step": round(1e3 * (m1 - m0) / steps, 3),
          "value": round(n_fluid * steps / (m1 - m0), 1),
          "ppe_iters_last_step": sim.scalars()["iters"], "tol": 1e-9,
          "max_con_diff_vs_fp64": dcon}
    sim.close()
    return {"workload": f"dam_break dx={dx:g}: {n_fluid} fluid, {n} particles, "
                        f"{len(cfg['soil']['x'])} soil; salt-water step at dt={dt:g}, salt "
                        f"above y - 0.1 = 0.3 (x - 0.2); {warmup} + {steps} steps",
            "ppe_iters_last_step": iters,
            "value": round(n_fluid * steps / (t1 - t0), 1),
            "ms_per_step": round(1e3 * (t1 - t0) / steps, 3), "substeps_per_step": sub,
            "con_range": [float(con.min()), float(con.max())],
            "k_scalart": {"launches": launches, "avg_us": round(avg_s * 1e6, 2),
                          "bytes_per_launch": nb, "achieved_GBs": round(gbs, 1),
                          "hbm_frac": round(gbs / HBM_PEAK_GBS, 4)},
            "multigrid": mg}


def run_scalart_at_scale(cfg, reps=20):
    """scalart alone on the bench tank (1,057,810 particles), with pores in
    the lower half, eddy viscosity on both sides of the 1e-6 threshold and
    a rough salinity field, on the lists of one neighbour search: the
    transport kernel at the bench's particle count (the open tank itself
    cannot run two salt-water steps in a row)."""
    from beach_slope_confined_amd import capi
    n = len(cfg["parts"]["x"])
    sim = capi.IsphSim(cfg["params"].copy(salinity=1), n, 1)
    sim.set_particles(cfg["parts"])
    sim.neighbour()
    q = sim.get_particles()
    x, y = q["x"], q["y"]
    sim.set_field("porosity", np.where(y < 0.5, 0.4 + 0.2 * x, 1.0))
    sim.set_field("nut", np.where(np.sin(9 * x) > 0, 3e-6, 4e-7) * (x > 0.3))
    sim.set_field("con", np.clip(y - 0.5 + 0.1 * np.sin(30 * x), -0.5, 0.5))
    sim.scalart()
    sim.profile_enable(1)
    for _ in range(reps):
        sim.scalart()
    launches, tot = sim.kernel_time(4)
    ptr, _ = sim.get_neighbours()
    nb = scalart_bytes(ptr, q["tid"])
    sim.close()
    avg_s = tot / max(launches, 1) / 1e3
    gbs = nb / avg_s / 1e9 if launches else 0.0
    return {"particles": n, "pairs": int(ptr[-1]), "launches": launches,
            "avg_us": round(avg_s * 1e6, 2), "bytes_per_launch": nb,
            "achieved_GBs": round(gbs, 1), "hbm_frac": round(gbs / HBM_PEAK_GBS, 4)}


def run_reference_scenario(steps, cpu_steps):
    """The reference's own run (setup.f90, boundary.f90, internal.f90,
    main.f90): its geometry (14,193 fluid, 1,233 wall, 21,795 soil
    particles, 4h cells in the reference particle order, hybrid kernel,
    reference drag, adaptive dt), `steps` steps of each of main.f90's three
    phases (start-up, buffered, salt-water with transport), on the GPU and
    `cpu_steps` per phase on the oracle (OpenMP, this machine's cores)."""
    import torch
    from beach_slope_confined_amd import capi, scenario
    sc = scenario.reference_scenario()
    cap = len(sc["parts"]["x"]) + 4000

    def timed(sim, k):
        sim.set_particles(sc["parts"])
        sim.set_soil(sc["soil"])
        marks = []

        def hook(sim_, name):
            if isinstance(sim_, capi.IsphSim):
                torch.cuda.synchronize()
            marks.append((name, time.perf_counter()))
        if isinstance(sim, capi.IsphSim):
            torch.cuda.synchronize()
        marks.append(("init", time.perf_counter()))
        scenario.run_reference(sim, sc, steps1=k, steps2=k, steps3=k, on_phase=hook)
        out, last = {}, marks[0][1]
        for name, tm in marks[1:]:
            out[name] = round(1e3 * (tm - last) / k, 3)
            last = tm
        return out, sim.scalars()["iters"]
    gpu = capi.IsphSim(sc["params"], cap, len(sc["soil"]["x"]))
    timed(gpu, 2)                      # warm-up
    gpu.close()
    gpu = capi.IsphSim(sc["params"], cap, len(sc["soil"]["x"]))
    g_ms, g_it = timed(gpu, steps)
    gpu.close()
    # the same run with the multigrid-preconditioned PPE (to 1e-9)
    from beach_slope_confined_amd.params import PREC_MIXED_MG
    mg = capi.IsphSim(sc["params"].copy(precision=PREC_MIXED_MG, tol=1e-9), cap,
                      len(sc["soil"]["x"]))
    m_ms, m_it = timed(mg, steps)
    mg.close()
    out = {"workload": "reference scenario (main.f90 phases 1-3, buffers, salinity): "
                       f"{sc['n_fluid']} fluid, {len(sc['parts']['x'])} particles, "
                       f"{len(sc['soil']['x'])} soil", "steps_per_phase": steps,
           "gpu_ms_per_step": g_ms, "gpu_ppe_iters_last_step": g_it,
           "gpu_multigrid_ms_per_step": m_ms, "gpu_multigrid_ppe_iters_last_step": m_it}
    if cpu_steps:
        from oracle import pyoracle
        from oracle.pyoracle import OracleSim
        cores = os.cpu_count() or 1
        pyoracle.set_threads(min(16, cores))
        ora = OracleSim(sc["params"], cap, len(sc["soil"]["x"]))
        c_ms, _ = timed(ora, cpu_steps)
        ora.close()
        out.update(oracle_ms_per_step=c_ms, oracle_threads=min(16, cores),
                   oracle_steps_per_phase=cpu_steps)
    return out


def run_adjoint(dx, nsteps, every, multigrid=False, dt=None):
    """Adjoint/forward cost on the porous beach-slope dam break: nsteps FP64
    steps with checkpoints every `every` steps, objective = time-mean
    Shepard pressure at 3 probes on the slope face + water depth at
    x = 1.2 m and x = 0.3 m; gradient with respect to (por, alpha, beta).
    With `multigrid` both the forward and the transposed solves are
    multigrid-preconditioned (to 1e-9)."""
    import torch
    from beach_slope_confined_amd import adjoint, capi, configs
    from beach_slope_confined_amd.params import PREC_MIXED_MG
    cfg = configs.dam_break(dx)
    if multigrid:
        cfg["params"] = cfg["params"].copy(precision=PREC_MIXED_MG, tol=1e-9)
    if dt is not None:
        cfg["params"] = cfg["params"].copy(dt=dt)
    sim = configs.make_sim(cfg, capi.IsphSim)
    sim.checkpoint(0)
    obj = adjoint.slope_objective(gauges=(1.2, 0.3))
    # warm-up: one short forward/backward (allocations, code paths)
    sim.adjoint_forward(obj, 2, 1)
    sim.adjoint_backward()
    sim.restore(0)
    torch.cuda.synchronize()
    a = time.perf_counter()
    sim.step(nsteps)  # plain forward, no checkpoints or objective
    torch.cuda.synchronize()
    plain = time.perf_counter() - a
    sim.restore(0)
    a = time.perf_counter()
    J = sim.adjoint_forward(obj, nsteps, every)
    torch.cuda.synchronize()
    fwd = time.perf_counter() - a
    a = time.perf_counter()
    g, _, adj_iters = sim.adjoint_backward()
    torch.cuda.synchronize()
    bwd = time.perf_counter() - a
    n = len(cfg["parts"]["x"])
    sim.close()
    kind = "multigrid" if multigrid else "FP64"
    return {"workload": f"dam_break dx={dx:g}: {cfg['n_fluid']} fluid, {n} particles, "
                        f"{len(cfg['soil']['x'])} soil; {nsteps} {kind} steps of "
                        f"dt={cfg['params'].dt:g} with shifting, checkpoints every {every}",
            "particle_steps_per_s": round(cfg["n_fluid"] * nsteps / plain, 1),
            "forward_s": round(fwd, 4), "plain_forward_s": round(plain, 4),
            "backward_s": round(bwd, 4), "adjoint_over_forward": round(bwd / fwd, 3),
            "J": J, "grad_por_alpha_beta": [float(v) for v in g],
            "adjoint_bicgstab_iters": int(adj_iters)}


def host_cores():
    """Physical cores of the host (unique core/socket pairs from lscpu), the
    CPUs this process may run on, and the OpenMP threads the oracle uses."""
    import subprocess
    phys = None
    try:
        out = subprocess.run(["lscpu", "-p=CORE,SOCKET"], capture_output=True, text=True,
                             check=True).stdout
        phys = len({ln for ln in out.splitlines() if ln and not ln.startswith("#")})
    except Exception:
        pass
    allowed = len(os.sched_getaffinity(0))
    threads = int(os.environ.get("OMP_NUM_THREADS", allowed))
    return phys, allowed, min(threads, allowed)


def model_mg_step_bytes(n, nnz_off, iters, val_bytes=4):
    """Algorithmic bytes of one multigrid-preconditioned step, per row: the
    fixed stages 1284 + 60 K (the FP64 step's model) plus, per BiCGSTAB
    iteration, with v = val_bytes (4 for FP32 values, 8 for FP64):
    - two V-cycles, each: the point right-hand side (a 16-byte pair read,
      8 + v written), its restriction to the cells (v), the cell levels
      (about 23 v / 4: nine stencil values and three vector values per cell
      and level, down and up, on a fifth as many cells as rows) and the
      particle sweep (nine cell-offset sums, the cell, b and the output:
      9 v + 20);
    - the two outer products A p^ and A s^, the only passes over the
      matrix: K (v + 2) + 60 and K (v + 2) + 36;
    - the update pass (96).
    With FP32 values that is 12 K + 414 bytes a row per iteration."""
    k = nnz_off / n
    v = val_bytes
    vcycle = (24 + v) + v + 23 * v / 4 + (9 * v + 20)
    per_iter = 2 * vcycle + 2 * k * (v + 2) + 60 + 36 + 96
    return n * ((1284 + 60 * k) + iters * per_iter)


def run_mixed_errors():
    """Pressure and velocity error of the mixed-precision forms on the first
    step of the small dam break (dx = 0.01) and of the adjoint configuration
    (dx = 0.002), against a tightly solved FP64 step (1e-13) and against the
    FP64 drop-in (1e-8), from the same input."""
    from beach_slope_confined_amd import capi, configs
    from beach_slope_confined_amd.params import (PREC_FP64, PREC_FP64_MG, PREC_MIXED_MG,
                                                 PREC_MIXED_V64)
    out = {}
    for name, dx in (("dam_break_dx0.01", 0.01), ("dam_break_dx0.002", 0.002)):
        cfg = configs.dam_break(dx)
        q = {}
        for key, prec, tol in (("tight", PREC_FP64, 1e-13), ("fp64", PREC_FP64, 1e-8),
                               ("mixed_v64", PREC_MIXED_V64, 1e-8),
                               ("mixed_mg", PREC_MIXED_MG, 1e-9), ("fp64_mg", PREC_FP64_MG, 1e-9)):
            c = dict(cfg)
            c["params"] = cfg["params"].copy(precision=prec, tol=tol, maxit=200000)
            sim = configs.make_sim(c, capi.IsphSim)
            sim.step(1)
            q[key] = (sim.get_particles(), sim.scalars()["iters"])
            sim.close()
        r = {"n_total": len(cfg["parts"]["x"])}
        for key in ("fp64", "mixed_v64", "mixed_mg", "fp64_mg"):
            pe, ve = step_errors(q["tight"][0], q[key][0])
            pf, vf = step_errors(q["fp64"][0], q[key][0])
            r[key] = {"iters": q[key][1], "pressure_err_vs_tight": pe, "velocity_err_vs_tight": ve,
                      "pressure_err_vs_fp64": pf, "velocity_err_vs_fp64": vf}
        out[name] = r
    return out


def run_large(steps, with_adjoint):
    """The large beach run: 4,194,304 fluid particles (an 8192 x 512 column
    at dx = 0.0005 with the dam break's 1:3 porous slope, porosity 0.4,
    d50 0.01, alpha 200, beta 1.1; configs.beach_wide), `steps` multigrid
    steps with shifting at dt = 1.25e-5 from rest, then the adjoint of the
    whole run (checkpoints every 10 steps) with respect to (por, alpha,
    beta).  The column starts on a perfect lattice: with the +-0.05 dx
    jitter of the small dam break, the reference's symmetric pressure
    gradient is unstable at a depth of 512 particles (the run leaves the
    grid at step 4, at dx = 0.0005 and at dx = 0.001 alike; 256 is stable).
    Beside it, from the same input: one step with the FP64-valued
    multigrid solve (the same algorithm, FP64 values), the error of the
    first step's pressure against a tight FP64 solve (FP64 multigrid to
    1e-11), and one plain FP64 BiCGSTAB step (the drop-in) within 40,000
    iterations."""
    import torch
    from beach_slope_confined_amd import adjoint as adj
    from beach_slope_confined_amd import capi, configs
    from beach_slope_confined_amd.params import PREC_FP64, PREC_FP64_MG, PREC_MIXED_MG
    big = configs.beach_wide(dt=1.25e-5, jitter=0.0)
    nf, nt, ns = big["n_fluid"], len(big["parts"]["x"]), len(big["soil"]["x"])
    out = {"workload": f"{big['name']}: {nf} fluid, {nt} particles, {ns} soil; multigrid PPE "
                       f"(tol 1e-9), dt 1.25e-5, shifting, from rest on a perfect lattice",
           "n_fluid": nf, "n_total": nt, "n_soil": ns, "steps": steps, "tol": 1e-9}

    def first_step(prec, tol, maxit=40000, warm=True):
        c = dict(big)
        c["params"] = big["params"].copy(precision=prec, tol=tol, maxit=maxit)
        sim = configs.make_sim(c, capi.IsphSim)
        if warm:
            sim.checkpoint(0)
            sim.step(1)        # warm-up (allocations, the multigrid hierarchy)
            sim.restore(0)
        torch.cuda.synchronize()
        a = time.perf_counter()
        sim.step(1)
        torch.cuda.synchronize()
        ms = 1e3 * (time.perf_counter() - a)
        s = sim.scalars()
        q = sim.get_particles()
        sim.close()
        return ms, s, q

    log(f"{big['name']}: first steps (mixed and FP64 multigrid, tight, plain FP64)")
    ms_mx, s_mx, q_mx = first_step(PREC_MIXED_MG, 1e-9)
    ms_64, s_64, q_64 = first_step(PREC_FP64_MG, 1e-9)
    # 1e-11: the FP64 true residual of 4.26M rows does not reach 1e-12
    _, s_t, q_t = first_step(PREC_FP64_MG, 1e-11)
    pe, ve = step_errors(q_t, q_mx)
    pe64, ve64 = step_errors(q_t, q_64)
    out["first_step"] = {
        "mixed_mg_ms": round(ms_mx, 3), "mixed_mg_iters": s_mx["iters"],
        "fp64_mg_ms": round(ms_64, 3), "fp64_mg_iters": s_64["iters"],
        "mixed_mg_speedup_vs_fp64_mg": round(ms_64 / ms_mx, 3),
        "tight": "FP64 multigrid to 1e-11, %d iterations" % s_t["iters"],
        "mixed_mg_pressure_err_vs_tight": pe, "mixed_mg_velocity_err_vs_tight": ve,
        "fp64_mg_pressure_err_vs_tight": pe64, "fp64_mg_velocity_err_vs_tight": ve64}
    try:
        ms_p, s_p, q_p = first_step(PREC_FP64, 1e-8, warm=False)
        pp, vp = step_errors(q_t, q_p)
        out["first_step"]["fp64_plain"] = {"ms": round(ms_p, 3), "iters": s_p["iters"],
                                           "pressure_err_vs_tight": pp,
                                           "velocity_err_vs_tight": vp}
    except Exception as e:  # reported, not hidden
        out["first_step"]["fp64_plain"] = {"error": str(e), "maxit": 40000}
    del q_mx, q_64, q_t

    c = dict(big)
    c["params"] = big["params"].copy(precision=PREC_MIXED_MG, tol=1e-9, maxit=40000)
    sim = configs.make_sim(c, capi.IsphSim)
    sim.checkpoint(0)
    its = []
    torch.cuda.synchronize()
    a = time.perf_counter()
    for k in range(steps):
        sim.step(1)
        its.append(sim.scalars()["iters"])
        if (k + 1) % 100 == 0:
            log(f"{big['name']}: {k + 1} of {steps} steps")
    torch.cuda.synchronize()
    sec = time.perf_counter() - a
    s = sim.scalars()
    ptr, _ = sim.get_neighbours()
    nnz = int(ptr[-1])
    n = len(ptr) - 1
    sb = model_mg_step_bytes(n, nnz, float(np.mean(its)))
    out["forward"] = {"seconds": round(sec, 3), "ms_per_step": round(1e3 * sec / steps, 3),
                      "value": round(nf * steps / sec, 1),
                      "iters_mean": round(float(np.mean(its)), 2), "iters_max": int(max(its)),
                      "umax_at_end": s["umax"], "t_at_end": s["t"]}
    out["step_model"] = {"bytes_per_step": round(sb), "achieved": round(sb / (sec / steps) / 1e9, 1),
                         "unit": "GB/s", "frac": round(sb / (sec / steps) / 1e9 / HBM_PEAK_GBS, 4),
                         "model": "rows x (1284 + 60K + iterations x (12K + 414)) bytes, "
                                  "K = %.3f" % (nnz / n)}
    if with_adjoint:
        w, _hc = big["column"]
        toe = big["toe"]
        obj = adj.slope_objective(toe=toe, probe_x=(toe + 0.01, toe + 0.02, toe + 0.03),
                                  gauges=(w - 0.05, w - 0.5))
        sim.set_progress(lambda what, done, total:
                         log(f"{big['name']}: {what} {done} of {total} segments")
                         if done % 10 == 0 or done == total else None)
        sim.restore(0)
        torch.cuda.synchronize()
        a0 = time.perf_counter()
        J = sim.adjoint_forward(obj, steps, 10)
        torch.cuda.synchronize()
        a1 = time.perf_counter()
        g, _, it = sim.adjoint_backward()
        torch.cuda.synchronize()
        a2 = time.perf_counter()
        out["adjoint"] = {"steps": steps, "checkpoint_every": 10,
                          "objective": "mean over steps of 3 slope-face probes + 2 depth gauges",
                          "forward_s": round(a1 - a0, 3), "backward_s": round(a2 - a1, 3),
                          "adjoint_over_forward": round((a2 - a1) / (a1 - a0), 3), "J": J,
                          "grad_por_alpha_beta": [float(v) for v in g],
                          "transposed_iters": int(it)}
    sim.close()
    return out


def cpu_baseline(cfg, runs):
    """The CPU oracle (this repository's C/OpenMP restatement of the
    reference step) on the host cores: the full single step at scale (every
    stage, BiCGSTAB to the same 1e-8 true residual; its iterates are
    bit-identical to the GPU's), `runs` times from the same input, median;
    and the small dam break (C0 geometry, 2,556 particles) for 200 steps,
    `runs` times, median."""
    from beach_slope_confined_amd import configs
    from oracle import pyoracle
    from oracle.pyoracle import OracleSim
    phys, allowed, want = host_cores()
    threads = pyoracle.set_threads(want)
    big = []
    for r in range(runs):
        sim = configs.make_sim(cfg, OracleSim)
        a = time.perf_counter()
        sim.step(1)
        big.append(time.perf_counter() - a)
        log(f"oracle full step {r + 1} of {runs}: {big[-1]:.2f} s")
        it = sim.scalars()["iters"]
        sim.close()
    small = []
    c0 = configs.dam_break(0.01)
    for _ in range(runs):
        sim = configs.make_sim(c0, OracleSim)
        a = time.perf_counter()
        sim.step(200)
        small.append(time.perf_counter() - a)
        sim.close()
    step = float(np.median(big))
    s200 = float(np.median(small))
    sample = (f"oracle: the full step at {len(cfg['parts']['x'])} particles ({it} BiCGSTAB "
              f"iterations), {runs} runs from the same input, median {step:.2f} s (runs "
              + ", ".join(f"{v:.2f}" for v in big) + f" s); dam break dx=0.01 "
              f"({len(c0['parts']['x'])} particles) for 200 steps, {runs} runs, median "
              f"{s200:.3f} s")
    return dict(value=cfg["n_fluid"] / step, unit="particle-steps/s", cores=threads,
                kind="port", sample=sample, seconds_per_step=step, runs_s=big,
                host_physical_cores=phys, host_cpus_allowed=allowed,
                c0_200_steps={"seconds": s200, "runs_s": small,
                              "particle_steps_per_s": c0["n_fluid"] * 200 / s200,
                              "n_fluid": c0["n_fluid"], "n_total": len(c0["parts"]["x"])})


def main():
    global args
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "the ISPH library runs on one GPU per process; --gpus must be "
                          "1 (the step shards by x-slabs of cells, each GPU a slab with a "
                          "one-cell halo, but no multi-GPU path is built)"}))
        sys.exit(2)
    import torch
    torch.cuda.init()
    from beach_slope_confined_amd import configs
    from beach_slope_confined_amd.params import (PREC_FGMRES, PREC_FP64, PREC_FP64_MG, PREC_IR,
                                                 PREC_MIXED, PREC_MIXED_MG, PREC_MIXED_V64)

    cfg = configs.tank_1m(args.nside)
    if args.pmc_run:
        from beach_slope_confined_amd import capi
        c = dict(cfg)
        c["params"] = cfg["params"].copy(maxit=40)
        sim = configs.make_sim(c, capi.IsphSim)
        try:
            sim.step(1)
        except capi.IsphError as e:
            print("solve capped as intended:", e)
        ms = sim.copy_bench(1 << 30, 5)
        print(json.dumps({"pmc_run": True, "copy_bytes": 1 << 30, "copy_ms": ms,
                          "copy_GBs": 2 * (1 << 30) / ms / 1e6}))
        return
    n_fluid = cfg["n_fluid"]
    n_total = len(cfg["parts"]["x"])
    if args.reference_run_only:
        print(json.dumps({"reference_run": run_reference_scenario(20, 0 if args.no_cpu else 2)}),
              flush=True)
        return
    if args.salinity_only:
        sal = run_salinity(args.salt_dx)
        sal["k_scalart_tank"] = run_scalart_at_scale(cfg)
        print(json.dumps({"salinity": sal}), flush=True)
        return

    if args.fgmres_only:
        rgm = run_gpu(cfg, args.steps, args.warmup, PREC_FGMRES, False, tol=1e-3, maxit=150)
        rgm["sim"].close()
        print(json.dumps({"fgmres_reference": {
            "value": round(n_fluid * args.steps / rgm["seconds"], 1),
            "ms_per_step": round(1e3 * rgm["seconds"] / args.steps, 3), "tol": 1e-3,
            "iters_per_step": rgm["iters"][args.warmup:], "relres": rgm["relres"]}}), flush=True)
        return

    if args.mg_only or args.fp64mg_only:
        key = "mixed_mg" if args.mg_only else "fp64_mg"
        rmg = run_gpu(cfg, args.steps, args.warmup,
                      PREC_MIXED_MG if args.mg_only else PREC_FP64_MG, False, tol=1e-9)
        rmg["sim"].close()
        print(json.dumps({key: {
            "value": round(n_fluid * args.steps / rmg["seconds"], 1),
            "ms_per_step": round(1e3 * rmg["seconds"] / args.steps, 3), "tol": 1e-9,
            "iters_per_step": rmg["iters"][args.warmup:], "relres": rmg["relres"]}}), flush=True)
        return

    r64 = run_gpu(cfg, args.steps, args.warmup, PREC_FP64, True)
    ms = 1e3 * r64["seconds"] / args.steps
    value = n_fluid * args.steps / r64["seconds"]
    timed_iters = r64["iters"][args.warmup:]

    # roofline of the dominant kernel (larger total time of the two SpMV kernels)
    kt = r64["ktime"]
    dom = max(kt, key=lambda k: kt[k][1])
    launches, tot = kt[dom]
    avg_s = tot / max(launches, 1) / 1e3
    # algorithmic bytes: the per-row model (32 + 12 K per SpMV row); the
    # kernel's own count (16-bit offsets, vectors once) is reported beside it
    nbytes = model_spmv_bytes(r64["n"], r64["nnz_off"])
    own = kernel_bytes(dom, r64["nnz_off"], r64["n"], 8, 8, r64["idx_bytes"])
    achieved = nbytes / avg_s / 1e9 if launches else 0.0
    ti = "short" if r64["idx_bytes"] == 2 else "int"
    kname = (f"k_krylov_1t<double, double, {ti}, false>" if dom == "p_spmv"
             else f"k_krylov_2<double, double, {ti}>")
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": kname, "launches": launches, "avg_us": round(avg_s * 1e6, 2),
                "bytes_per_launch": nbytes,
                "bytes_model": "32 + 12K per row, K = %.3f off-diagonal entries per row"
                               % (r64["nnz_off"] / r64["n"]),
                "own_bytes_per_launch": own,
                "own_frac": round(own / avg_s / 1e9 / HBM_PEAK_GBS, 4) if launches else None}
    # attainable streaming bandwidth on this GPU: a 1 GiB copy (read + write)
    try:
        cms = r64["sim"].copy_bench(1 << 30, 10)
        copy_gbs = 2 * (1 << 30) / cms / 1e6
        roofline["copy_GBs"] = round(copy_gbs, 1)
        roofline["frac_of_copy"] = round(achieved / copy_gbs, 4)
    except Exception as e:  # reported, not hidden
        roofline["copy_GBs"] = {"error": str(e)}
    sb = model_step_bytes(r64["n"], r64["nnz_off"], timed_iters[-1])
    step_model = {"bytes_per_step": round(sb), "achieved": round(sb / (ms / 1e3) / 1e9, 1),
                  "unit": "GB/s", "frac": round(sb / (ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
                  "model": "rows x (1284 + 60K + iterations x (248 + 24K)) bytes"}
    pmc = os.path.join(ROOT, "profiles", "bicg_spmv_traffic.json")
    if os.path.exists(pmc):
        with open(pmc) as f:
            d = json.load(f)
        roofline["traffic"] = d.get(roofline["kernel"], {}).get("hbm_bytes_per_launch")
        if roofline["traffic"]:
            roofline["traffic_source"] = ("profiles/bicg_spmv_traffic.json: rocprofv3 --pmc, "
                                          "read bytes by request size, WRITE_SIZE")
            roofline["traffic_over_model"] = round(roofline["traffic"] / nbytes, 3)

    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "particle-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"tank_{args.nside}x{args.nside}: full ISPH step, FP64 PPE "
                               f"(BiCGSTAB, true residual 1e-8)",
                   "n_fluid": n_fluid, "n_total": n_total, "global_batch": n_fluid,
                   "seq_len": 1, "parallelism": "single GPU",
                   "kernel": "wendland_c2", "h_over_dx": 1.3, "tol": 1e-8},
        "ppe_iters_per_step": timed_iters, "relres": r64["relres"],
        "restore_ms_per_step_untimed": round(1e3 * r64["restore_s"] / max(args.steps, 1), 3),
        "particle_steps_per_s_total": round(n_total * args.steps / r64["seconds"], 1),
        "roofline": roofline, "step_model": step_model,
    }

    if not args.no_mixed:
        for key, prec, vb in (("mixed", PREC_MIXED_V64, 8), ("mixed_f32vec", PREC_MIXED, 4)):
            try:
                rmx = run_gpu(cfg, args.steps, args.warmup, prec, True)
            except Exception as e:  # reported, not hidden
                line[key] = {"error": str(e)}
                if prec == PREC_MIXED:
                    line[key]["note"] = ("FP32 Krylov vectors stagnate on this system; a "
                                         "measured negative result (the FP64-vector and "
                                         "multigrid forms are the mixed-precision solves)")
                continue
            vmx = n_fluid * args.steps / rmx["seconds"]
            a, b = r64["first"], rmx["first"]
            f = a["tid"] == 3
            perr = float(np.abs(b["pressure"] - a["pressure"]).max() / np.abs(a["pressure"]).max())
            vmag = np.hypot(a["vx"][f], a["vy"][f]).max()
            verr = float(max(np.abs(b["vx"] - a["vx"])[f].max(),
                             np.abs(b["vy"] - a["vy"])[f].max()) / vmag)
            ktm = rmx["ktime"]
            domm = max(ktm, key=lambda k: ktm[k][1])
            lm, tm = ktm[domm]
            avgm = tm / max(lm, 1) / 1e3
            # FP32 values, 16-bit column offsets, vectors of vb bytes
            bm = kernel_bytes(domm, rmx["nnz_off"], rmx["n"], 4, vb, rmx["idx_bytes"])
            line[key] = {
                "value": round(vmx, 1), "ms_per_step": round(1e3 * rmx["seconds"] / args.steps, 3),
                "speedup_vs_fp64": round(vmx / value, 3),
                "krylov_vectors": "fp64" if vb == 8 else "fp32",
                "iters_per_step": rmx["iters"][args.warmup:], "reliable_updates": rmx["outer"],
                "relres": rmx["relres"], "pressure_err_vs_fp64": perr,
                "velocity_err_vs_fp64": verr,
                "dominant_kernel": {"kernel": f"k_krylov_{'1t' if domm == 'p_spmv' else 2}<float, "
                                              f"{'double' if vb == 8 else 'float'}, "
                                              f"{'short' if rmx['idx_bytes'] == 2 else 'int'}>",
                                    "avg_us": round(avgm * 1e6, 2),
                                    "achieved_GBs": round(bm / avgm / 1e9, 1) if lm else None},
            }
            rmx["sim"].close()

    if not args.no_mixed:
        # multigrid-preconditioned mixed PPE, solved to 1e-9; errors against
        # the FP64 step and against a tightly solved FP64 step (1e-12: FP64
        # BiCGSTAB's recurrence stalls above 1e-13 at a million rows)
        try:
            rmg = run_gpu(cfg, args.steps, args.warmup, PREC_MIXED_MG, False, tol=1e-9)
            rmg["sim"].close()
            vmg = n_fluid * args.steps / rmg["seconds"]
            pe, ve = step_errors(r64["first"], rmg["first"])
            line["mixed_mg"] = {
                "value": round(vmg, 1), "ms_per_step": round(1e3 * rmg["seconds"] / args.steps, 3),
                "speedup_vs_fp64": round(vmg / value, 3), "tol": 1e-9,
                "iters_per_step": rmg["iters"][args.warmup:], "reliable_updates": rmg["outer"],
                "relres": rmg["relres"], "pressure_err_vs_fp64": pe, "velocity_err_vs_fp64": ve,
            }
            # whole-step byte model of the multigrid step (as for the large beach)
            sb = model_mg_step_bytes(r64["n"], r64["nnz_off"],
                                     float(np.mean(rmg["iters"][args.warmup:])))
            ach = sb / (rmg["seconds"] / args.steps) / 1e9
            line["mixed_mg"]["step_model"] = {
                "bytes_per_step": round(sb), "achieved": round(ach, 1), "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4),
                "model": "rows x (1284 + 60K + iterations x (12K + 414)) bytes, K = %.3f"
                         % (r64["nnz_off"] / r64["n"])}
        except Exception as e:  # reported, not hidden
            line["mixed_mg"] = {"error": str(e)}
        if "value" in line["mixed_mg"]:
            try:
                rtight = run_gpu(cfg, 0, 1, PREC_FP64, False, tol=1e-12)
                rtight["sim"].close()
                pt, vt = step_errors(rtight["first"], rmg["first"])
                p64, v64e = step_errors(rtight["first"], r64["first"])
                line["mixed_mg"].update({
                    "tight_tol": 1e-12, "tight_iters": rtight["iters"][0],
                    "pressure_err_vs_tight": pt, "velocity_err_vs_tight": vt,
                    "fp64_pressure_err_vs_tight": p64, "fp64_velocity_err_vs_tight": v64e,
                })
            except Exception as e:
                line["mixed_mg"]["tight_error"] = str(e)

    if not args.no_mixed and "value" in line.get("mixed_mg", {}):
        # the same multigrid algorithm with every value FP64: mixed_mg over
        # fp64_mg is what FP32 buys at an equal true residual (1e-9)
        try:
            rf = run_gpu(cfg, args.steps, args.warmup, PREC_FP64_MG, False, tol=1e-9)
            rf["sim"].close()
            vf = n_fluid * args.steps / rf["seconds"]
            pe, ve = step_errors(r64["first"], rf["first"])
            mg = line["mixed_mg"]
            line["fp64_mg"] = {
                "value": round(vf, 1), "ms_per_step": round(1e3 * rf["seconds"] / args.steps, 3),
                "speedup_vs_fp64": round(vf / value, 3), "tol": 1e-9,
                "iters_per_step": rf["iters"][args.warmup:], "relres": rf["relres"],
                "pressure_err_vs_fp64": pe, "velocity_err_vs_fp64": ve,
            }
            if "tight_tol" in mg:
                pt, vt = step_errors(rtight["first"], rf["first"])
                line["fp64_mg"].update(pressure_err_vs_tight=pt, velocity_err_vs_tight=vt)
            pm, vm = step_errors(rf["first"], rmg["first"])
            mg["speedup_vs_fp64_mg"] = round(mg["value"] / vf, 3)
            mg["pressure_err_vs_fp64_mg"] = pm
            mg["velocity_err_vs_fp64_mg"] = vm
        except Exception as e:  # reported, not hidden
            line["fp64_mg"] = {"error": str(e)}
        # iterative refinement with an FP32 inner BiCGSTAB restarted from
        # zero on every FP64 true residual, to the FP64 bar (1e-8)
        try:
            rir = run_gpu(cfg, args.steps, args.warmup, PREC_IR, False)
            rir["sim"].close()
            vir = n_fluid * args.steps / rir["seconds"]
            pe, ve = step_errors(r64["first"], rir["first"])
            line["ir"] = {
                "value": round(vir, 1), "ms_per_step": round(1e3 * rir["seconds"] / args.steps, 3),
                "speedup_vs_fp64": round(vir / value, 3), "tol": 1e-8,
                "iters_per_step": rir["iters"][args.warmup:], "corrections": rir["outer"],
                "relres": rir["relres"], "pressure_err_vs_fp64": pe, "velocity_err_vs_fp64": ve,
            }
        except Exception as e:  # reported, not hidden
            line["ir"] = {"error": str(e),
                          "note": "iterative refinement with an unpreconditioned FP32 inner "
                                  "BiCGSTAB restarted from zero: a measured negative result "
                                  "when it fails"}

    if not args.no_mixed:
        # the reference's live solve (FGMRES, tol 1e-3 on the residual
        # estimate, at most 150 iterations, warm start fsol) on the same input
        try:
            rgm = run_gpu(cfg, args.steps, args.warmup, PREC_FGMRES, False, tol=1e-3, maxit=150)
            rgm["sim"].close()
            vgm = n_fluid * args.steps / rgm["seconds"]
            pe, ve = step_errors(r64["first"], rgm["first"])
            line["fgmres_reference"] = {
                "value": round(vgm, 1), "ms_per_step": round(1e3 * rgm["seconds"] / args.steps, 3),
                "speedup_vs_fp64": round(vgm / value, 3), "tol": 1e-3, "maxit": 150,
                "iters_per_step": rgm["iters"][args.warmup:], "relres": rgm["relres"],
                "converged": bool(rgm["relres"] <= 1e-3),
                "pressure_err_vs_fp64": pe, "velocity_err_vs_fp64": ve,
            }
        except Exception as e:  # reported, not hidden
            line["fgmres_reference"] = {"error": str(e)}

    if not args.no_mixed:
        try:
            log("mixed-precision errors on the small configurations")
            line["mixed_errors"] = run_mixed_errors()
        except Exception as e:  # reported, not hidden
            line["mixed_errors"] = {"error": str(e)}

    if not args.no_mixed and not args.no_large:
        try:
            line["large_mg"] = run_large(args.large_steps, args.large_adjoint)
        except Exception as e:  # reported, not hidden
            line["large_mg"] = {"error": str(e)}

    if not args.no_salinity:
        try:
            log("salinity leg")
            line["salinity"] = run_salinity(args.salt_dx)
            line["salinity"]["k_scalart_tank"] = run_scalart_at_scale(cfg)
        except Exception as e:  # reported, not hidden
            line["salinity"] = {"error": str(e)}

    if not args.no_reference_run:
        try:
            log("reference scenario leg")
            line["reference_run"] = run_reference_scenario(20, 0 if args.no_cpu else 2)
        except Exception as e:  # reported, not hidden
            line["reference_run"] = {"error": str(e)}

    if not args.no_adjoint:
        try:
            log("adjoint leg")
            line["adjoint"] = run_adjoint(args.adjoint_dx, args.adjoint_steps, 10)
            mg = run_adjoint(args.adjoint_dx, args.adjoint_steps, 10, multigrid=True)
            g64 = np.array(line["adjoint"]["grad_por_alpha_beta"])
            mg["grad_rel_diff_vs_fp64"] = [float(v) for v in
                                           np.abs(np.array(mg["grad_por_alpha_beta"]) - g64) /
                                           np.abs(g64)]
            line["adjoint"]["multigrid"] = mg
        except Exception as e:  # reported, not hidden
            line["adjoint"] = {"error": str(e)}
        # 500 steps with shifting and the multigrid PPE, and the adjoint of
        # the whole run, on the finest dam break that stays stable for that
        # long (dx = 0.001 at dt = 2.5e-5; the 4M beach does not)
        try:
            log("long-run leg")
            line["long_run"] = run_adjoint(0.001, 500, 10, multigrid=True, dt=2.5e-5)
        except Exception as e:  # reported, not hidden
            line["long_run"] = {"error": str(e)}

    if not args.no_cpu:
        log("CPU baseline (oracle)")
        line["cpu_baseline"] = cpu_baseline(cfg, args.cpu_runs)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
