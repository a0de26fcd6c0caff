#!/usr/bin/env python3
"""World-step benchmark: grid cell-steps per second on one MI355X.

Workload (the headline configuration): a 2048 x 2048 grid world with a
1024 x 1024 fluid, FP32 fluid and multigrid pressure, exact u32 monomer
transport.  Inputs are synthetic and seeded: 16-mode sinusoidal velocity
at 0.5 fluid cells per step, ~15 % Gaussian-blob rock plus a rubble shell,
Poisson(4) monomers in 8 channels, a 32-grid-row (16 fluid rows) vent band.
Parameters are the reference defaults with the parity profile (fluid and
transport every step, no UV, no solar heating).  A step is one full world
step: fluid (temperature, forces, viscosity, advection, vorticity,
divergence, multigrid pressure, projection, net-flux, walls, thermal
activation) then transport (aggregation, scatter, commit) with the exact
per-channel ledger checked on the device every step.

Prints one JSON line.  The timed region covers exactly --steps steps with the
world's stream drained on both sides; inputs are resident in HBM before it.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ribossome-v2_amd"))

import numpy as np  # noqa: E402

import ribossome_amd as ra  # noqa: E402

GRID, FLUID = 2048, 1024
SEED = 2
METRIC = "world cell-steps/s at 2048^2 (mixed vs FP32, HBM %)"

# Algorithmic bytes of each kernel: every array it must touch read or written
# once (stencil halos and bilinear gathers are cache traffic; an atomic add
# counts as a read and a write).  Units: per fluid cell (F), per grid cell (G).
BYTES_PER_UNIT = {
    "update_temperature": (56, "F"),
    "buoyancy_gather": (12, "F"),
    "add_forces": (52, "F"),
    "diffuse_velocity": (20, "F"),
    "advect_velocity": (20, "F"),
    "vorticity": (20, "F"),
    "divergence": (16, "F"),
    "subtract_gradient": (24, "F"),
    "net_flux": (12, "F"),
    "enforce_boundaries": (20, "F"),
    # temperature 1 B per grid cell (4 B per fluid cell, read once); a hot
    # cell (temperature above TEMP_ACT_THRESHOLD) also reads and writes its
    # 16 B chem word: kernel_bytes adds 32 B per hot cell
    "thermal_activation": (1, "G"),
    "agg": (20, "G"),
    "clear_next": (16, "G"),
    # chem 16 + gamma 4 + aggregation 4 + light 1 + velocity 2 + temperature 1
    # read, chem_next 16 read + 16 written by the atomics
    "transport_scatter": (60, "G"),
    "ledger": (16, "G"),
    # light (per grid cell; a light cell covers 4): gamma 16 + chem 64 read, both
    # bands written (+ the rows above, 8, for a propagation pass) per light cell
    "light_transmit": (22, "G"),
    "light_propagate": (24, "G"),
    # the sequential sweep reads and writes both bands once
    "light_sweep": (4, "G"),
    # terrain grain dynamics, per pass: the decision reads gamma and writes a
    # move byte, the move reads gamma and the codes and writes gamma, the
    # evictions read the codes (the monomers of the few cells that received a
    # grain are not counted); the slope reads gamma and writes 8 B
    "relax_gamma_a": (15, "G"),
    "relax_gamma_b": (15, "G"),
    "compute_gamma_slope": (12, "G"),
    # agents' coupling points: the smoothed velocity reads the velocity once
    # (the neighbours from cache) and writes 8 B; the force gather with agent
    # words reads and clears 8 B more than buoyancy_gather
    "smooth_velocity": (16, "F"),
    "buoyancy_gather_agents": (28, "F"),
    # tolerance-mode solver, level 0 (S = bytes of the preconditioner's storage)
    "pcg_init": (33, "F"),
    "pcg_final": (9, "F"),
}
MG_LDS_LIMIT = 163840  # LDS per workgroup on gfx950
MG_BLOCK_LEVELS = 8
MG_SYM_COARSE_MAX = 1024


def mg_block_lds(sizes):
    """LDS bytes of the one-workgroup multigrid kernel over these levels:
    unknowns, right-hand sides and masks, the staged coarsest right-hand side
    and res / W of the top level."""
    cells = sum(n * n for n in sizes)
    return ((9 * cells + 3) & ~3) + 4 * MG_SYM_COARSE_MAX + 4 * sizes[0] * sizes[0]


def block_start(sizes):
    """First multigrid level solved inside one workgroup (the library's rule)."""
    b = len(sizes)
    while b > 0 and len(sizes) - (b - 1) <= MG_BLOCK_LEVELS and mg_block_lds(sizes[b - 1:]) <= MG_LDS_LIMIT:
        b -= 1
    return b


# transport.wgsl:49: thermal activation runs in cells hotter than this
TEMP_ACT_THRESHOLD = 2.0


def hot_grid_cells(w, cfg):
    """Grid cells whose fluid cell is above TEMP_ACT_THRESHOLD (the cells
    thermal_activation reads and writes the chem of)."""
    t = w.download(ra.TEMP)
    return int(np.count_nonzero(t > TEMP_ACT_THRESHOLD)) * (cfg.cells() // cfg.fluid_cells())


def kernel_bytes(name, cfg, sizes, mode, hot=0):
    """Algorithmic bytes of one launch (the mean over one V-cycle's launches
    for the per-level multigrid kernels); `hot` is the hot grid cells of
    thermal_activation."""
    S = 2 if MODES[mode]["precision"] == 1 else 4
    tol = MODES[mode]["mg_tol"] > 0
    b = max(block_start(sizes), 1) if tol else block_start(sizes)
    lo = 1 if tol else 0
    F, G = cfg.fluid_cells(), cfg.cells()
    if name in BYTES_PER_UNIT:
        per, unit = BYTES_PER_UNIT[name]
        return per * (G if unit == "G" else F) + (32 * hot if name == "thermal_activation" else 0)
    if name == "mg_down":  # fluid 1, p and rhs read, pre-smoothed p written; coarse rhs, p
        levels = sizes[lo:b]
        return sum((1 + 3 * S + S / 2) * n * n for n in levels) / max(len(levels), 1)
    if name == "mg_up":  # fluid 1, pre-smoothed p and rhs read, coarse p, p written
        levels = sizes[lo:b]
        return sum((1 + 3 * S + S / 4) * n * n for n in levels) / max(len(levels), 1)
    if name == "mg_block":  # top level's p, rhs, fluid in, p out; coarser masks
        n = sizes[b]
        return (1 + 3 * S) * n * n + sum(m * m for m in sizes[b + 1:])
    if name == "cg_down0":  # fluid, r, s, w (halo), u, p, x; r, s, p, x, pre-smoothed u, bf16 r; coarse rhs, p
        return (1 + 12 + S + 8 + 16 + S + (S if S == 2 else 0) + S / 2) * F
    if name == "cg_up0":  # fluid, pre-smoothed u, rhs, coarse u, s, p (and r when bf16); u, w
        return (1 + S + (S if S == 2 else 4) + S / 4 + 8 + (4 if S == 2 else 0) + S + 4) * F
    return 0.0


def world_inputs(cfg):
    gamma = ra.gen_gamma(cfg, SEED, 0.15, True)
    vel = ra.gen_velocity(cfg, SEED, 0.5 / 0.017, gamma)
    chem = ra.gen_chem(cfg, SEED, 4.0, gamma)
    heat = ra.gen_vent_heat(cfg, 16, 5000.0 * 0.3)
    return gamma, vel, chem, heat


MODES = {
    # the headline: FP32 multigrid-preconditioned CG to relative residual 1e-5
    "fp32": dict(mg_tol=1e-5, precision=0),
    # mixed precision: the preconditioner stored in bfloat16, same tolerance
    "mixed": dict(mg_tol=1e-5, precision=1),
    # the reference's own schedule: one V-cycle per step, no residual target
    "reference": dict(mg_tol=0.0, precision=0),
    # the plain FP32 path as the benchmark configuration states it: the
    # reference V-cycle repeated until the relative residual reaches 1e-5
    "vcycle": dict(mg_tol=1e-5, precision=0, stationary=1),
}


def configure(p, s, args, mode="fp32"):
    p.transport_every = 1
    p.uv_strength = 0.0
    p.sun_heat = 0.0
    s.fluid_substep = 1
    s.terrain_enabled = 0  # the benchmark's profile: frozen terrain and light field
    s.light_mode = 0
    s.mg_cycles = args.mg_cycles
    s.mg_max_cycles = args.mg_max_cycles
    s.mg_tol = MODES[mode]["mg_tol"]
    s.precision = MODES[mode]["precision"]
    s.mg_stationary = MODES[mode].get("stationary", 0)
    if s.mg_stationary:
        s.mg_max_cycles = args.vcycle_max_cycles
    s.ledger_check = 1
    s.no_graphs = 1 if getattr(args, "no_graphs", False) else 0
    return p, s


def build_world(cfg, inputs, args, mode="fp32"):
    gamma, vel, chem, heat = inputs
    w = ra.World(cfg, SEED)
    p, s = configure(ra.SimParams.default(SEED), ra.WorldSettings.default(), args, mode)
    w.params = p
    w.upload(ra.GAMMA, gamma)
    w.upload(ra.VELOCITY, vel)
    w.upload(ra.CHEM, chem)
    w.upload(ra.HEAT_SRC, heat)
    w.settings = s
    if getattr(args, "no_overlap", False):
        w.set_pipeline(False)  # every kernel on its own, as the HIP-event pass times them
    if not getattr(args, "no_autotune", False):
        w.autotune()
    return w


def timed_run(cfg, inputs, args, mode):
    """Warm-up, then exactly args.steps steps with the stream drained on both
    sides; returns (seconds, world)."""
    w = build_world(cfg, inputs, args, mode)
    seeded = ra.Ledger.from_cells(inputs[2])
    # the flow settles first: the solver's iterations grow over the first
    # few hundred steps as the vent heats the water (8-10 a solve before
    # step 200, 12-17 after 300), so a short timed window right after the
    # start would be faster than the run it stands for
    w.run(args.settle_steps)
    # warm-up, including one replay of every (chunk length, buffer parity)
    # graph the timed loop will use, so no capture falls in the timed region
    w.run(args.warmup)
    for n in sorted({64 if args.steps >= 64 else 0, args.steps % 64} - {0}):
        w.encode_steps(n)
        w.encode_steps(n)
    w.sync()
    before = w.solver_stats()
    w.solver_residuals(reset=True)
    t0 = time.perf_counter()
    w.run(args.steps)
    w.sync()
    dt = time.perf_counter() - t0
    after = w.solver_stats()
    w.timed_solver_stats = tuple(a - b for a, b in zip(after, before))
    w.timed_residuals = w.solver_residuals()
    bad = w.ledger_status()
    if bad is not None:
        raise SystemExit(f"{mode}: per-channel monomer totals changed at epoch {bad}")
    if w.ledger_blocking().channels() != seeded.channels():
        raise SystemExit(f"{mode}: ledger differs from the seeded totals")
    return dt, w


def fp64_error(cfg, inputs, args):
    """Error of the binary32 and mixed-precision solves against the binary64
    oracle on a bounded sample: one step from the same inputs, the binary64
    solve run to relative residual 1e-10 (the exact answer for this
    purpose), the GPU runs to 1e-5.  Relative L2 of velocity and pressure."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_ctypes as oc

    gamma, vel, chem, heat = inputs
    ow = oc.OracleWorld(cfg, SEED, fp64=True)
    p, s = configure(ow.params, ra.WorldSettings.default(), args, "fp32")
    s.mg_tol = 1e-10
    s.mg_max_cycles = 400
    ow.params = p
    ow.set_settings(s)
    for which, a in ((ra.GAMMA, gamma), (ra.VELOCITY, vel), (ra.CHEM, chem), (ra.HEAT_SRC, heat)):
        ow.upload(which, a)
    t0 = time.perf_counter()
    ow.encode_steps(1)
    dt = time.perf_counter() - t0
    ref = {which: ow.download_f64(which) for which in (ra.VELOCITY, ra.PRESSURE)}
    it64, rel64 = ow.mg_stats()
    out = {"sample": f"1 step from the inputs; binary64 oracle to relative residual {rel64:.1e} in {it64} iterations "
                     f"({dt:.1f} s on the host)"}
    for mode in ("fp32", "mixed"):
        w = build_world(cfg, inputs, args, mode)
        w.encode_steps(1)
        w.sync()
        it, rel = w.mg_stats()
        e = {}
        for which, name in ((ra.VELOCITY, "velocity"), (ra.PRESSURE, "pressure")):
            a, b = w.download(which).astype(np.float64).reshape(-1), ref[which].reshape(-1)
            e[name] = float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
        e["iterations"] = it
        e["true_rel_residual"] = rel
        out[mode] = e
        w.close()
    return out


def light_measure(cfg, inputs, args):
    """The light step on the same 2048^2 world (light_mode 1: the reference's
    light schedule, one propagation pass per step, constant sun): HIP-event
    times of the propagation pass and of the full sweep with its pre-pass,
    and the propagation pass against the HBM roofline."""
    w = build_world(cfg, inputs, args, "reference")
    p = w.params
    p.uv_strength = 3.0
    p.uv_depth = 4.0
    p.day_period = 0.0
    w.params = p
    s = w.settings
    s.light_mode = 1
    s.light_rows_per_step = 1
    w.settings = s
    w.run(4)
    w.set_timing(True)
    w.reset_kernel_times()
    w.run(args.profile_steps)
    for _ in range(4):
        w.invalidate_light()
        w.encode_steps(1)
    times = w.kernel_times()
    w.set_timing(False)
    w.close()
    out = {}
    for k in ("light_propagate", "light_transmit", "light_sweep"):
        tot, n = times.get(k, (0.0, 0))
        if not n:
            continue
        us = tot / n * 1000.0
        b = kernel_bytes(k, cfg, [], "reference")
        out[k] = {"avg_launch_us": us, "launches": n, "algorithmic_bytes_per_launch": b,
                  "achieved_GB_per_s": b / (us * 1e-6) / 1e9, "frac_of_8TBps": b / (us * 1e-6) / 8e12}
    return out


def terrain_measure(cfg, inputs, args):
    """Terrain grain dynamics on the same 2048^2 world (terrain_enabled 1,
    reference V-cycle, grains settling so grain_fall runs on its clock):
    HIP-event times of each terrain stage against the HBM roofline, and the
    whole step's rate with the terrain moving."""
    w = build_world(cfg, inputs, args, "reference")
    p = w.params
    p.sediment_settle = 1.0
    w.params = p
    s = w.settings
    s.terrain_enabled = 1
    w.settings = s
    w.run(8)
    w.set_timing(True)
    w.reset_kernel_times()
    w.run(args.profile_steps)
    times = w.kernel_times()
    w.set_timing(False)
    w.run(8)
    w.sync()
    n = max(args.terrain_steps, 1)
    t0 = time.perf_counter()
    w.run(n)
    w.sync()
    dt = time.perf_counter() - t0
    moved = int(np.count_nonzero(w.download(ra.GAMMA) != inputs[0]))
    w.close()
    out = {"steps": n, "ms_per_step": dt / n * 1000.0, "cell_steps_per_s": cfg.cells() * n / dt,
           "grid_cells_changed": moved}
    for k in ("relax_gamma_a", "relax_gamma_b", "grain_fall", "compute_gamma_slope", "build_solid_mask"):
        tot, cnt = times.get(k, (0.0, 0))
        if not cnt:
            continue
        us = tot / cnt * 1000.0
        b = kernel_bytes(k, cfg, [], "reference")
        e = {"avg_launch_us": us, "launches": cnt}
        if b:
            e.update({"algorithmic_bytes_per_launch": b, "achieved_GB_per_s": b / (us * 1e-6) / 1e9,
                      "frac_of_8TBps": b / (us * 1e-6) / 8e12})
        out[k] = e
    return out


def agents_measure(cfg, inputs, args):
    """The agents' coupling points on the same 2048^2 world: the force
    gather of a fluid step that consumes uploaded force words (5 % of the
    fluid cells pushed), the smoothed-velocity download's kernel, and one
    million monomer exchanges on distinct words (spend, activate, take, add),
    with HIP-event times; the exchanges' bytes are the record read, the
    count written and the 4-byte word read and written (in 64-byte sectors,
    one per exchange)."""
    w = build_world(cfg, inputs, args, "reference")
    rng = np.random.default_rng(77)
    F = cfg.fluid_cells()
    words = np.zeros(2 * F, np.int32)
    cells = rng.choice(F, F // 20, replace=False)
    words[2 * cells] = rng.integers(-4000, 4000, cells.size)
    words[2 * cells + 1] = rng.integers(-4000, 4000, cells.size)
    n_ops = 1 << 20
    slots = rng.choice(4 * cfg.cells(), n_ops, replace=False)
    ops = np.stack([slots // 4, slots % 4, rng.integers(0, 6, n_ops), rng.integers(1, 4, n_ops)], 1).astype(np.uint32)
    w.run(4)
    w.set_timing(True)
    w.reset_kernel_times()
    sub = max(1, min(4, w.settings.fluid_substep))
    for _ in range(4):
        w.upload(ra.FORCE_VECTORS, words)
        w.run(sub)  # holds exactly one fluid step, which consumes the words
    agent_gather = w.kernel_times().get("buoyancy_gather", (0.0, 0))
    w.reset_kernel_times()
    w.run(8)
    plain_gather = w.kernel_times().get("buoyancy_gather", (0.0, 0))
    w.reset_kernel_times()
    for _ in range(4):
        w.download(ra.VELOCITY_SMOOTH)
        w.chem_ops(ops)
    times = w.kernel_times()
    w.set_timing(False)
    w.close()
    out = {}

    def entry(us, b):
        return {"avg_launch_us": us, "algorithmic_bytes_per_launch": b, "achieved_GB_per_s": b / (us * 1e-6) / 1e9,
                "frac_of_8TBps": b / (us * 1e-6) / 8e12}

    for key, (tot, cnt), name in (("buoyancy_gather_with_agent_words", agent_gather, "buoyancy_gather_agents"),
                                  ("buoyancy_gather_plain", plain_gather, "buoyancy_gather")):
        if cnt:
            out[key] = entry(tot / cnt * 1000.0, kernel_bytes(name, cfg, [], "reference"))
    tot, cnt = times.get("smooth_velocity", (0.0, 0))
    if cnt:
        out["smooth_velocity"] = entry(tot / cnt * 1000.0, kernel_bytes("smooth_velocity", cfg, [], "reference"))
    tot, cnt = times.get("chem_ops", (0.0, 0))
    if cnt:
        # the library splits a call into launches of at most F^2 / 2
        # exchanges (its staging buffer): 4 calls of n_ops make `cnt` launches
        us = tot / cnt * 1000.0
        per_launch = 4 * n_ops / cnt
        out["chem_ops"] = dict(entry(us, per_launch * (16 + 4 + 64)), exchanges_per_launch=per_launch,
                               exchanges_per_call=n_ops, exchanges_per_s=4 * n_ops / (tot * 1e-3))
    return out


def default_profile_measure(cfg, inputs, args):
    """The reference's default profile on the same 2048^2 world:
    SimParams::default() and WorldSettings::default() (fluid every second
    step, transport every third, one V-cycle), with the terrain moving, the
    light step (one propagation pass per step) under the day cycle's moving
    sun, solar heating and the vent cycle; the steps replay as HIP graphs,
    the sun of each epoch read from the device table.  The timed window
    starts in the day (epoch 2000)."""
    gamma, vel, chem, heat = inputs
    w = ra.World(cfg, SEED)
    p = ra.SimParams.default(SEED)
    p.epoch = 2000
    w.params = p
    s = ra.WorldSettings.default()
    s.terrain_enabled = 1
    s.light_mode = 1
    s.ledger_check = 1
    w.settings = s
    for which, a in ((ra.GAMMA, gamma), (ra.VELOCITY, vel), (ra.CHEM, chem), (ra.HEAT_SRC, heat)):
        w.upload(which, a)
    seeded = ra.Ledger.from_cells(chem)
    w.run(args.warmup)
    for n in sorted({64 if args.default_steps >= 64 else 0, args.default_steps % 64} - {0}):
        w.encode_steps(n)
        w.encode_steps(n)
    w.sync()
    t0 = time.perf_counter()
    w.run(args.default_steps)
    w.sync()
    dt = time.perf_counter() - t0
    bad = w.ledger_status()
    ok = bad is None and w.ledger_blocking().channels() == seeded.channels()
    w.close()
    if not ok:
        raise SystemExit(f"default profile: per-channel monomer totals changed (epoch {bad})")
    n = args.default_steps
    return {"steps": n, "ms_per_step": dt / n * 1000.0, "cell_steps_per_s": cfg.cells() * n / dt,
            "profile": "SimParams::default(), WorldSettings::default() + terrain moving + light step; "
                       "fluid every 2nd step, transport every 3rd, 1 V-cycle, day cycle from epoch 2000"}


def cpu_model():
    """The host CPU's model name (Linux /proc/cpuinfo), or None."""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def oracle_rate(oc, cfg, seed, params, settings, uploads, steps, threads, warm=1):
    """Cell-steps/s of the CPU oracle on `threads` threads: `warm` untimed
    steps, then `steps` timed ones.  Returns (rate, seconds)."""
    oc.lib().oracle_set_threads(threads)
    ow = oc.OracleWorld(cfg, seed)
    ow.params = params
    ow.set_settings(settings)
    for which, a in uploads:
        ow.upload(which, a)
    if warm:
        ow.encode_steps(warm)
    t0 = time.perf_counter()
    ow.run(steps)
    dt = time.perf_counter() - t0
    ow.close()
    return cfg.cells() * steps / dt, dt


def cpu_baseline(cfg, inputs, args):
    """The CPU oracle (C++ restatement of the reference WGSL, OpenMP) on the
    host's cores: a bounded sample of the same 2048^2 workload on every
    thread it is given (the value) and on one thread, and the 256^2 plain
    method (configuration 1: ~10 % rock, Poisson(4), 2 V-cycles a step to a
    16^2 coarsest level, 100 steps) in full, on every thread and on one."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_ctypes as oc

    gamma, vel, chem, heat = inputs
    all_threads = oc.lib().oracle_get_threads()
    ow = oc.OracleWorld(cfg, SEED)
    p, s = configure(ow.params, ra.WorldSettings.default(), args, args.cpu_mode)
    ow.close()
    uploads = ((ra.GAMMA, gamma), (ra.VELOCITY, vel), (ra.CHEM, chem), (ra.HEAT_SRC, heat))
    rate, dt = oracle_rate(oc, cfg, SEED, p, s, uploads, args.cpu_steps, all_threads)
    rate1, dt1 = oracle_rate(oc, cfg, SEED, p, s, uploads, 1, 1)
    # configuration 1 in full
    c1 = ra.WorldConfig(256, 128)
    g1 = ra.gen_gamma(c1, 1, 0.10, False)
    v1 = ra.gen_velocity(c1, 1, 0.5 / 0.017, g1)
    m1 = ra.gen_chem(c1, 1, 4.0, g1)
    p1 = ra.SimParams.default(1)
    p1.transport_every = 1
    p1.uv_strength = 0.0
    p1.sun_heat = 0.0
    s1 = ra.WorldSettings.default()
    s1.fluid_substep = 1
    s1.terrain_enabled = 0
    s1.light_mode = 0
    s1.mg_cycles = 2
    s1.mg_min_size = 16
    s1.ledger_check = 1
    up1 = ((ra.GAMMA, g1), (ra.VELOCITY, v1), (ra.CHEM, m1))
    r1, t1 = oracle_rate(oc, c1, 1, p1, s1, up1, 100, all_threads, warm=0)
    r11, t11 = oracle_rate(oc, c1, 1, p1, s1, up1, 100, 1, warm=0)
    oc.lib().oracle_set_threads(all_threads)
    return {
        "value": rate,
        "unit": "cell-steps/s",
        "cores": int(all_threads),
        "kind": "port",
        "sample": f"{args.cpu_steps} steps of the same 2048^2 world ({args.cpu_mode} mode) after 1 warm-up step, "
                  f"C++ oracle, {all_threads} OpenMP threads, {dt:.2f} s",
        "cpu_model": cpu_model(),
        "one_thread": {"value": rate1, "sample": f"1 step of the same world after 1 warm-up step, {dt1:.2f} s"},
        "config1_plain_256": {"value": r1, "one_thread_value": r11, "steps": 100,
                              "seconds": t1, "one_thread_seconds": t11,
                              "sample": "256^2 grid / 128^2 fluid, ~10 % rock, Poisson(4), 16-mode velocity at "
                                        "0.5 fluid cells per step, 2 V-cycles per step to 16^2, ledger every step"},
    }


# HIP kernel names of the timing names, for the committed rocprofv3 summaries
HIP_NAMES = {
    "transport_scatter": "k_scatter", "mg_down": "k_mg_down", "mg_up": "k_mg_up", "mg_block": "k_mg_block",
    "cg_down0": "k_cg_down0", "cg_up0": "k_cg_up0", "net_flux": "k_net_flux", "vorticity": "k_vorticity",
    "agg": "k_agg", "ledger": "k_ledger", "clear_next": "k_fill_u32", "add_forces": "k_add_forces",
    "diffuse_velocity": "k_diffuse", "advect_velocity": "k_advect", "update_temperature": "k_temperature",
    "divergence": "k_divergence", "subtract_gradient": "k_subgrad", "thermal_activation": "k_thermal",
    "enforce_boundaries": "k_boundaries", "buoyancy_gather": "k_forces",
}
PMC_SUMMARY = os.path.join(ROOT, "profiles", "world_step_2048_hbm_counters.md")
SQ_SUMMARY = os.path.join(ROOT, "profiles", "transport_scatter_2048_sq_counters.md")
STEP_SUMMARY = os.path.join(ROOT, "profiles", "world_step_2048_fp32_per_step.md")
# one wave-level VALU instruction per CU per cycle: 256 CUs at 2.4 GHz
VALU_PEAK_PER_S = 256 * 2.4e9


def sq_valu_per_launch(kernel="k_scatter"):
    """VALU instructions (wave level) per launch of `kernel` from the
    committed SQ counter summary of this workload, or None."""
    if not os.path.exists(SQ_SUMMARY):
        return None
    for line in open(SQ_SUMMARY):
        cols = [c.strip().strip("`") for c in line.strip().strip("|").split("|")]
        if len(cols) > 3 and cols[0] == kernel:
            try:
                return float(cols[3])
            except ValueError:
                return None
    return None


def trace_dominant():
    """The timing name of the kernel with the most kernel time per step of the
    headline mode in the committed kernel trace (the first row of its
    table), or None.  HIP events around each eager launch add a few
    microseconds to every launch, which ranks the solver's short, numerous
    launches above the kernel trace's ranking; the trace decides."""
    if not os.path.exists(STEP_SUMMARY):
        return None
    rows = 0
    for line in open(STEP_SUMMARY):
        cols = [c.strip().strip("`") for c in line.strip().strip("|").split("|")]
        if not line.startswith("|") or len(cols) < 6:
            continue
        rows += 1
        if rows == 3:  # header and separator first
            return cols[0]
    return None


def pmc_traffic(kernel, mode):
    """HBM bytes per launch of `kernel` from the committed FETCH_SIZE /
    WRITE_SIZE summary of this workload (KiB per dispatch, averaged over the
    dispatches; both counters read 1.0x the bytes of a fill and of a cold
    read of known size on MI355X), or None when it has no row for it."""
    name = HIP_NAMES.get(kernel)
    if not name or not os.path.exists(PMC_SUMMARY):
        return None
    want = "true" if MODES[mode]["mg_tol"] > 0 else "false"
    kib, calls = 0.0, 0
    for line in open(PMC_SUMMARY):
        cols = [c.strip().strip("`") for c in line.strip().strip("|").split("|")]
        if len(cols) < 5 or not (cols[0] == name or cols[0].startswith(name + "<")):
            continue
        if "<" in cols[0] and name.startswith("k_mg_") and want not in cols[0]:
            continue
        try:
            k, n = float(cols[2]) + float(cols[3]), int(cols[4].split()[0])  # "kept of all" dispatches
        except ValueError:
            continue
        kib += k * n
        calls += n
    return kib * 1024.0 / calls if calls else None


SOLVER_KERNELS = ("cg_down0", "cg_up0", "mg_block", "mg_down", "mg_up")


def working_launches(times, sizes, mode, iterations, launch_ms=None):
    """{kernel: (ms, launches)} over working launches: every launch of the
    per-step kernels; for the solver's per-iteration kernels, `iterations`
    (summed over the solves of the pass) times the launches per iteration,
    taken as the longest of the kernel's launches (`launch_ms(name)` gives
    every launch's milliseconds): a launch in an iteration slot the solve did
    not use returns at once and is the shortest."""
    tol = MODES[mode]["mg_tol"] > 0
    out = {}
    for k, (ms, n) in times.items():
        if tol and k in SOLVER_KERNELS:
            b = max(block_start(sizes), 1)
            per_it = (b - 1) if k in ("mg_down", "mg_up") else 1
            busy = min(n, iterations * per_it)
            if busy <= 0:
                continue
            each = np.sort(launch_ms(k))[::-1] if launch_ms else None
            if each is not None and len(each) == n:
                ms = float(each[:busy].sum())
            n = busy
        out[k] = (ms, n)
    return out


def step_model_bytes(cfg, sizes, mode, res, hot=0):
    """Algorithmic bytes of one world step: the per-step kernels once, the
    solver's kernels once per V-cycle it ran."""
    F, G = cfg.fluid_cells(), cfg.cells()
    once = ["update_temperature", "buoyancy_gather", "add_forces", "diffuse_velocity", "advect_velocity",
            "vorticity", "divergence", "subtract_gradient", "net_flux", "enforce_boundaries", "thermal_activation",
            "agg", "clear_next", "transport_scatter", "ledger"]
    total = sum(kernel_bytes(k, cfg, sizes, mode, hot) for k in once)
    tol = MODES[mode]["mg_tol"] > 0
    b = max(block_start(sizes), 1) if tol else block_start(sizes)
    nl = b - (1 if tol else 0)
    vcycle = nl * (kernel_bytes("mg_down", cfg, sizes, mode) + kernel_bytes("mg_up", cfg, sizes, mode))
    vcycle += kernel_bytes("mg_block", cfg, sizes, mode) if b < len(sizes) else 0.0
    if tol:
        vcycle += kernel_bytes("cg_down0", cfg, sizes, mode) + kernel_bytes("cg_up0", cfg, sizes, mode)
        v = (res.get("mean_iterations") or 0.0) + 1.0
        total += kernel_bytes("pcg_init", cfg, sizes, mode) + kernel_bytes("pcg_final", cfg, sizes, mode)
    else:
        v = 1.0
    return total + v * vcycle


def batched(args):
    """Independent worlds stepped together (the batched configuration):
    sides from {128, ..., 1024} grid cells, rock U(5, 20) %, monomers
    Poisson(U(2, 8)), the reference's one V-cycle per step, the ledger
    checked every step.  Timed twice over the same worlds: one launch per
    kernel for the whole batch (ribo_batch_*, the value), and one stream per
    world replaying its own graphs (per_stream).  Returns the JSON record."""
    sides = [128, 192, 256, 384, 512, 768, 1024]
    worlds = []
    for i in range(args.worlds):
        r = np.random.default_rng(1000 + i)
        side = int(r.choice(sides))
        rock, lam = float(r.uniform(0.05, 0.20)), float(r.uniform(2.0, 8.0))
        cfg = ra.WorldConfig(side, side // 2)
        gamma = ra.gen_gamma(cfg, 1000 + i, rock, True)
        vel = ra.gen_velocity(cfg, 1000 + i, 0.5 / 0.017, gamma)
        chem = ra.gen_chem(cfg, 1000 + i, lam, gamma)
        w = ra.World(cfg, 1000 + i)
        p, s = configure(ra.SimParams.default(1000 + i), ra.WorldSettings.default(), args, "reference")
        w.params = p
        w.upload(ra.GAMMA, gamma)
        w.upload(ra.VELOCITY, vel)
        w.upload(ra.CHEM, chem)
        w.settings = s
        w.set_pipeline(False)  # one stream per world
        w.autotune()
        worlds.append((w, cfg, ra.Ledger.from_cells(chem)))
    steps = args.batch_steps
    cells = sum(c.cells() for _, c, _ in worlds)

    def per_stream():
        for w, _, _ in worlds:  # warm-up: every graph the timed loop replays
            w.encode_steps(64)
            w.encode_steps(steps % 64 or 64)
        for w, _, _ in worlds:
            w.sync()
        left = [steps] * len(worlds)
        t0 = time.perf_counter()
        while any(left):
            for i, (w, _, _) in enumerate(worlds):
                if left[i]:
                    left[i] -= w.encode_steps(min(64, left[i]))
        for w, _, _ in worlds:
            w.sync()
        return time.perf_counter() - t0

    def one_launch():
        b = ra.Batch([w for w, _, _ in worlds])
        b.encode_steps(8)  # warm-up
        b.sync()
        n0 = b.launches()
        t0 = time.perf_counter()
        b.run(steps)
        b.sync()
        dt = time.perf_counter() - t0
        launches = b.launches() - n0
        b.close()
        return dt, launches

    dt_streams = per_stream()
    print(f"batched: per-stream {cells * steps / dt_streams / 1e9:.3f} G cell-steps/s", file=sys.stderr, flush=True)
    dt_batch, launches = one_launch()
    print(f"batched: one launch per kernel {cells * steps / dt_batch / 1e9:.3f} G cell-steps/s", file=sys.stderr,
          flush=True)
    for w, _, led in worlds:
        if w.ledger_status() is not None or w.ledger_blocking().channels() != led.channels():
            raise SystemExit("batched: a world's per-channel monomer totals changed")
    tiles = {c.grid_size: w.mg_tiles() for w, c, _ in worlds}
    for w, _, _ in worlds:
        w.close()
    return {"metric": "batched world cell-steps/s", "value": cells * steps / dt_batch, "unit": "cell-steps/s",
            "n_gpus": 1, "steps": steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32", "data": "synthetic", "ms_per_step": 1000.0 * dt_batch / steps,
            "per_stream": {"value": cells * steps / dt_streams, "ms_per_step": 1000.0 * dt_streams / steps,
                           "streams": len(worlds)},
            "one_launch_vs_per_stream": dt_streams / dt_batch,
            "launches_per_step": launches / steps,
            "config": {"workload": f"{len(worlds)} independent worlds, grid sides from {sides}, reference solver "
                                   "schedule (one V-cycle per step), ledger checked every step; value: one launch "
                                   "per kernel over every world; per_stream: one stream per world, multigrid tiles "
                                   "from the autotuner",
                       "worlds": len(worlds), "grid_cells": cells, "tiles_by_side": tiles}}


def serving_scenes(args):
    """Scene requests of the serving configuration: sides log-uniform in
    [128, 1024] grid cells (multiples of 16), U(50, steps_max) steps, the
    benchmark generators, the mixed-precision tolerance-mode solver."""
    rng = np.random.default_rng(5000)
    scenes = []
    for i in range(args.scenes):
        side = int(round(np.exp(rng.uniform(np.log(128), np.log(1024))) / 16.0)) * 16
        steps = int(rng.integers(50, args.serve_steps_max + 1))
        cfg = ra.WorldConfig(side, side // 2)
        gamma = ra.gen_gamma(cfg, 7000 + i, 0.15, True)
        inp = {ra.GAMMA: gamma, ra.VELOCITY: ra.gen_velocity(cfg, 7000 + i, 0.5 / 0.017, gamma),
               ra.CHEM: ra.gen_chem(cfg, 7000 + i, 4.0, gamma),
               ra.HEAT_SRC: ra.gen_vent_heat(cfg, max(1, side // 128), 5000.0 * 0.3)}
        p, s = configure(ra.SimParams.default(7000 + i), ra.WorldSettings.default(), args, "mixed")
        scenes.append((cfg, 7000 + i, steps, p, s, inp))
    return scenes


def serve(scenes, budget, max_running=0, arrivals=None, threads=0, batching=True):
    """Runs the scenes through the library's server; arrivals (seconds from
    the start) or all at once.  Returns (span, latencies, cell-steps)."""
    srv = ra.Server(budget, max_running, threads, batching)
    n = len(scenes)
    at = arrivals if arrivals is not None else np.zeros(n)
    tickets, nxt = [], 0
    t0 = time.perf_counter()
    last_note = 0.0
    while True:
        pending = sum(srv.status(t).state in (ra.SCENE_QUEUED, ra.SCENE_RUNNING) for t in tickets)
        if nxt >= n and pending == 0:
            break
        now = time.perf_counter() - t0
        if now - last_note >= 20.0:  # a progress line on stderr while a long serving run goes on
            print(f"serving: {now:.0f} s, {nxt} of {n} submitted, {nxt - pending} done", file=sys.stderr, flush=True)
            last_note = now
        while nxt < n and at[nxt] <= now:
            cfg, seed, steps, p, s, inp = scenes[nxt]
            tickets.append(srv.submit(cfg, seed, steps, p, s, inp))
            nxt += 1
        time.sleep(0.0005)
    span = time.perf_counter() - t0
    lat, work = [], 0
    for t, (cfg, _, steps, _, _, _) in zip(tickets, scenes):
        st = srv.status(t)
        if st.state != ra.SCENE_DONE:
            raise SystemExit(f"serving: scene {t} ended in state {st.state} (error {st.error})")
        lat.append(st.t_done - st.t_submit)
        work += cfg.cells() * steps
        srv.release(t)
    srv.close()
    return span, np.array(lat), work


def serving(args):
    """The serving configuration: the same scene set (1) one scene at a time
    on its own stream (the plain per-scene path), (2) all submitted at once,
    every scene resident on its own stream as soon as the budget admits it,
    (3) all at once with batching (the scenes stepped together, one launch
    per kernel, the batch rebuilt as scenes join and finish), (4) Poisson
    arrivals at --qps with batching.  Latency = finish - submission, from
    the server's clock."""
    scenes = serving_scenes(args)
    budget = int(args.hbm_budget_gb * (1 << 30))
    plain_span, _, work = serve(scenes, budget, max_running=1, threads=1, batching=False)
    streams_span, _, _ = serve(scenes, budget, threads=args.serve_threads, batching=False)
    batch_span, _, _ = serve(scenes, budget, threads=args.serve_threads)
    rng = np.random.default_rng(5001)
    arrivals = np.cumsum(rng.exponential(1.0 / args.qps, len(scenes)))
    span, lat, _ = serve(scenes, budget, arrivals=arrivals, threads=args.serve_threads)
    p50, p99 = float(np.percentile(lat, 50)), float(np.percentile(lat, 99))
    plain, batch = work / plain_span, work / batch_span
    return {"metric": "scene latency p50 (s)", "value": p50, "unit": "s", "p99": p99, "n_gpus": 1,
            "higher_is_better": False, "scaling": "strong", "vs_baseline": None,
            "dtype": "f32 fluid, bf16 preconditioner", "data": "synthetic",
            "throughput_cell_steps_per_s": work / span,
            "saturated_cell_steps_per_s": batch, "plain_per_scene_cell_steps_per_s": plain,
            "saturated_vs_plain": batch / plain,
            "saturated_streams_cell_steps_per_s": work / streams_span,
            "batched_vs_streams": streams_span / batch_span,
            "config": {"workload": f"{len(scenes)} scenes, sides log-uniform in [128, 1024], U(50, "
                                   f"{args.serve_steps_max}) steps each, mixed-precision tolerance-mode solver to 1e-5; "
                                   f"Poisson arrivals at {args.qps} per s; HBM budget {args.hbm_budget_gb} GiB",
                       "scenes": len(scenes), "qps": args.qps, "span_s": span, "plain_span_s": plain_span,
                       "saturated_span_s": batch_span, "saturated_streams_span_s": streams_span,
                       "hw_queues": os.environ.get("GPU_MAX_HW_QUEUES", "default"),
                       "scheduler_threads": args.serve_threads or "library default"}}


def note(msg):
    """A progress line on stderr as each stage of the run ends."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def device_ceiling(nbytes=1 << 30, reps=20):
    """The device's name, CU count, peak engine clock and LDS per workgroup
    (hipDeviceGetAttribute), and a measured HBM copy ceiling: a 1 GiB
    device-to-device hipMemcpyAsync (a copy kernel) timed with HIP events
    over `reps` copies, counting the bytes read and written.  Measurement
    plumbing only; the world step never calls it."""
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so")
    c_int, c_void_p, c_size_t = ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t

    def chk(rc, what):
        if rc != 0:
            raise RuntimeError(f"{what}: HIP error {rc}")

    def attr(a):
        v = c_int(0)
        chk(hip.hipDeviceGetAttribute(ctypes.byref(v), c_int(a), c_int(0)), "hipDeviceGetAttribute")
        return v.value

    # hipDeviceAttribute_t values of ROCm 7.2's hip_runtime_api.h:
    # ClockRate 5 (kHz), MultiprocessorCount 63, MaxSharedMemoryPerBlock 74
    name = ctypes.create_string_buffer(256)
    chk(hip.hipDeviceGetName(name, c_int(256), c_int(0)), "hipDeviceGetName")
    info = {"name": name.value.decode(errors="replace"), "cus": attr(63), "clock_mhz": attr(5) / 1000.0,
            "lds_per_workgroup_bytes": attr(74)}
    a, b = c_void_p(), c_void_p()
    chk(hip.hipMalloc(ctypes.byref(a), c_size_t(nbytes)), "hipMalloc")
    chk(hip.hipMalloc(ctypes.byref(b), c_size_t(nbytes)), "hipMalloc")
    e0, e1 = c_void_p(), c_void_p()
    chk(hip.hipEventCreate(ctypes.byref(e0)), "hipEventCreate")
    chk(hip.hipEventCreate(ctypes.byref(e1)), "hipEventCreate")
    d2d = 3  # hipMemcpyDeviceToDevice
    try:
        chk(hip.hipMemset(a, c_int(1), c_size_t(nbytes)), "hipMemset")
        for _ in range(3):
            chk(hip.hipMemcpyAsync(b, a, c_size_t(nbytes), c_int(d2d), c_void_p(0)), "hipMemcpyAsync")
        chk(hip.hipEventRecord(e0, c_void_p(0)), "hipEventRecord")
        for _ in range(reps):
            chk(hip.hipMemcpyAsync(b, a, c_size_t(nbytes), c_int(d2d), c_void_p(0)), "hipMemcpyAsync")
        chk(hip.hipEventRecord(e1, c_void_p(0)), "hipEventRecord")
        chk(hip.hipEventSynchronize(e1), "hipEventSynchronize")
        ms = ctypes.c_float(0.0)
        chk(hip.hipEventElapsedTime(ctypes.byref(ms), e0, e1), "hipEventElapsedTime")
        info["copy_GBps"] = 2.0 * nbytes * reps / (ms.value / 1000.0) / 1e9
        info["copy_frac_of_8TBps"] = info["copy_GBps"] / 8000.0
        info["copy_sample"] = f"{reps} device-to-device copies of 1 GiB (read + write bytes), HIP events"
    finally:
        hip.hipEventDestroy(e0)
        hip.hipEventDestroy(e1)
        hip.hipFree(a)
        hip.hipFree(b)
    return info


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=64)
    ap.add_argument("--settle-steps", type=int, default=400,
                    help="untimed steps before the warm-up, so the timed window is past the flow's start-up")
    ap.add_argument("--mg-cycles", type=int, default=1)
    ap.add_argument("--mg-max-cycles", type=int, default=80, help="iteration cap of the tolerance-mode solve")
    ap.add_argument("--modes", default="fp32,mixed,reference",
                    help="comma list of fp32, mixed, reference; the first is the headline value")
    ap.add_argument("--cpu-mode", default="fp32")
    ap.add_argument("--profile-steps", type=int, default=32)
    ap.add_argument("--cpu-steps", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true", help="skip the binary64-oracle error sample")
    ap.add_argument("--no-light", action="store_true", help="skip the light-step kernel timing")
    ap.add_argument("--no-agents", action="store_true", help="skip the agents' coupling-point timing")
    ap.add_argument("--no-device", action="store_true", help="skip the device properties and the copy-ceiling measurement")
    ap.add_argument("--no-overlap", action="store_true",
                    help="run each step's transport after its fluid, not beside the next step's (kernel traces of the "
                         "separate kernels)")
    ap.add_argument("--terrain-steps", type=int, default=64,
                    help="steps of the moving-terrain measurement (0 skips it)")
    ap.add_argument("--default-steps", type=int, default=256,
                    help="steps of the reference's default profile (0 skips it)")
    ap.add_argument("--vcycle-steps", type=int, default=100,
                    help="steps of the stationary V-cycle mode (the plain FP32 path to 1e-5); 0 skips it")
    ap.add_argument("--vcycle-max-cycles", type=int, default=400, help="V-cycle cap per solve of the stationary mode")
    ap.add_argument("--config", default="world", choices=["world", "batched", "serving"],
                    help="world: the 2048^2 headline; batched: many independent worlds; serving: Poisson scene arrivals")
    ap.add_argument("--worlds", type=int, default=64)
    ap.add_argument("--line-batched-steps", type=int, default=100,
                    help="steps of the shortened batched configuration in the headline line (0 skips it)")
    ap.add_argument("--line-serving-scenes", type=int, default=24,
                    help="scenes of the serving configuration in the headline line (0 skips it)")
    ap.add_argument("--batch-steps", type=int, default=500)
    ap.add_argument("--scenes", type=int, default=24)
    ap.add_argument("--qps", type=float, default=1.0)
    ap.add_argument("--serve-steps-max", type=int, default=2000)
    ap.add_argument("--hbm-budget-gb", type=float, default=256.0, help="device memory the serving form may admit")
    ap.add_argument("--serve-threads", type=int, default=0, help="scheduler threads of the server (0: its default)")
    ap.add_argument("--hw-queues", type=int, default=0,
                    help="GPU_MAX_HW_QUEUES for this process (0: the runtime's default; at most 32)")
    ap.add_argument("--no-autotune", action="store_true", help="keep the multigrid tile size rule")
    ap.add_argument("--no-graphs", action="store_true",
                    help="launch every kernel eagerly instead of replaying captured HIP graphs (for tracing tools)")
    args = ap.parse_args()
    if args.hw_queues:
        # read by the HIP runtime when it initialises, at the first world
        os.environ["GPU_MAX_HW_QUEUES"] = str(min(max(args.hw_queues, 1), 32))
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the world step is a single-device path here; --gpus {args.gpus} "
              "is not supported (independent worlds would shard one per GPU).", file=sys.stderr)
        sys.exit(2)

    if args.config == "batched":
        print(json.dumps(batched(args)))
        return
    if args.config == "serving":
        print(json.dumps(serving(args)))
        return
    cfg = ra.WorldConfig(GRID, FLUID)
    inputs = world_inputs(cfg)
    results = {}
    for mode in args.modes.split(","):
        dt, w = timed_run(cfg, inputs, args, mode)
        it, rel = w.mg_stats()
        solves, iters, missed = w.timed_solver_stats
        res = w.timed_residuals
        tol_mode = MODES[mode]["mg_tol"] > 0
        note(f"{mode}: {1000.0 * dt / args.steps:.3f} ms/step")
        results[mode] = {"value": cfg.cells() * args.steps / dt, "ms_per_step": 1000.0 * dt / args.steps,
                         "last_solve_iterations": it, "last_true_rel_residual": rel,
                         # over every timed solve: the refined iterate's true
                         # residual (binary64), the stored binary32 pressure's,
                         # and the restarts from the true residual
                         "max_true_rel_residual": res["max"] if tol_mode else None,
                         "max_stored_pressure_rel_residual": res["max_rounded"] if tol_mode else None,
                         "restarts": res["restarts"] if tol_mode else None,
                         "mean_iterations": (iters / solves) if solves else None, "solves_missing_tol": missed}
        if mode == args.modes.split(",")[0]:
            head_world = w
        else:
            w.close()
    head = args.modes.split(",")[0]
    value = results[head]["value"]
    ms = results[head]["ms_per_step"]
    w = head_world

    # Per-kernel HIP-event timing pass (events on the world's stream).
    w.set_timing(True)
    w.reset_kernel_times()
    s0 = w.solver_stats()
    w.run(args.profile_steps)
    s1 = w.solver_stats()
    times = w.kernel_times()
    w.set_timing(False)
    sizes = [lv[0] for lv in w.mg_levels()]
    hot = hot_grid_cells(w, cfg)
    # The dominant kernel by working launches: the solver's per-iteration
    # kernels also take early-exit launches in the iteration slots a solve
    # did not use; only as many launches as the solves ran iterations (a
    # solve applies the preconditioner once before its first iteration) count
    # as working launches of those kernels, the longest ones.
    work = working_launches(times, sizes, head, (s1[1] - s0[1]) + (s1[0] - s0[0]), w.kernel_launch_ms)
    dom_events = max(work, key=lambda k: work[k][0])
    dom = trace_dominant()
    if dom not in work:
        dom = dom_events
    work_ms, launches = work[dom]
    avg_s = work_ms / launches / 1000.0
    per_launch = kernel_bytes(dom, cfg, sizes, head, hot)
    achieved = per_launch / avg_s / 1e9 if per_launch else None
    traffic = pmc_traffic(dom, head)
    roofline = {
        "kernel": dom,
        "bound": "hbm",
        "achieved": achieved,
        "peak": 8000.0,
        "unit": "GB/s",
        "frac": (achieved / 8000.0) if achieved else None,
        "traffic": traffic,
        "avg_launch_us": avg_s * 1e6,
        "launches": launches,
        "algorithmic_bytes_per_launch": per_launch,
        "rule": "the most kernel time per step of the headline mode in the committed kernel trace with the "
                "kernels separate (profiles/world_step_2048_fp32_per_step.md, --no-overlap); its average over "
                "working launches in an eager HIP-event pass of --profile-steps steps (kernels timed one by one)",
        "dominant_by_events": dom_events,
    }
    # the transport scatter, the largest single kernel of every step, beside it
    sc_ms, sc_n = times.get("transport_scatter", (0.0, 0))
    sc_s = sc_ms / max(sc_n, 1) / 1000.0
    sc_bytes = kernel_bytes("transport_scatter", cfg, sizes, head, hot)
    roofline_scatter = {"kernel": "transport_scatter", "bound": "hbm", "achieved": sc_bytes / sc_s / 1e9 if sc_s else None,
                        "peak": 8000.0, "unit": "GB/s", "frac": sc_bytes / sc_s / 8e12 if sc_s else None,
                        "traffic": pmc_traffic("transport_scatter", head), "avg_launch_us": sc_s * 1e6,
                        "launches": sc_n, "algorithmic_bytes_per_launch": sc_bytes,
                        "note": "VALU-issue bound (PCG4D per monomer); HBM is not its limit"}
    # the scatter against its real bound: wave-level VALU instructions per
    # launch (committed SQ_INSTS_VALU) over the live launch time
    valu = sq_valu_per_launch()
    roofline_valu = None
    if valu and sc_s:
        roofline_valu = {"kernel": "transport_scatter", "bound": "valu", "achieved": valu / sc_s / 1e9,
                         "peak": VALU_PEAK_PER_S / 1e9, "unit": "G wave-instructions/s",
                         "frac": valu / sc_s / VALU_PEAK_PER_S, "instructions_per_launch": valu,
                         "avg_launch_us": sc_s * 1e6}
    kernels = {k: [round(v[0] / max(v[1], 1) * 1000.0, 2), v[1]] for k, v in sorted(times.items(), key=lambda kv: -kv[1][0])}
    step_bytes = {m: step_model_bytes(cfg, sizes, m, results[m], hot) for m in results}
    hbm = {m: {"algorithmic_bytes_per_step": step_bytes[m],
               "GB_per_s": step_bytes[m] / (results[m]["ms_per_step"] / 1000.0) / 1e9,
               "frac_of_8TBps": step_bytes[m] / (results[m]["ms_per_step"] / 1000.0) / 8e12} for m in results}

    note("kernel timing pass done")
    light = None if args.no_light else light_measure(cfg, inputs, args)
    note("light done")
    terrain = terrain_measure(cfg, inputs, args) if args.terrain_steps > 0 else None
    note("terrain done")
    agents = None if args.no_agents else agents_measure(cfg, inputs, args)
    note("agents done")
    default_profile = default_profile_measure(cfg, inputs, args) if args.default_steps > 0 else None
    note("default profile done")
    batched_rec = serving_rec = None
    if args.line_batched_steps > 0:
        # the batched configuration, shortened: 64 worlds, one launch per
        # kernel against one stream per world
        sub = argparse.Namespace(**vars(args))
        sub.worlds, sub.batch_steps = 64, args.line_batched_steps
        r = batched(sub)
        batched_rec = {"value": r["value"], "unit": r["unit"], "ms_per_step": r["ms_per_step"], "steps": r["steps"],
                       "per_stream": r["per_stream"], "one_launch_vs_per_stream": r["one_launch_vs_per_stream"],
                       "launches_per_step": r["launches_per_step"], "workload": r["config"]["workload"],
                       "grid_cells": r["config"]["grid_cells"]}
        note("batched done")
    if args.line_serving_scenes > 0:
        # the serving configuration as --config serving runs it (its scene
        # count, step range and arrival rate)
        sub = argparse.Namespace(**vars(args))
        sub.scenes = args.line_serving_scenes
        r = serving(sub)
        serving_rec = {k: r[k] for k in ("p99", "throughput_cell_steps_per_s", "saturated_cell_steps_per_s",
                                          "plain_per_scene_cell_steps_per_s", "saturated_vs_plain",
                                          "saturated_streams_cell_steps_per_s", "batched_vs_streams")}
        serving_rec["p50"] = r["value"]
        serving_rec["workload"] = r["config"]["workload"]
        note("serving done")
    plain = {}
    if args.vcycle_steps > 0:
        # the plain FP32 path (the stationary V-cycle to 1e-5) against the two
        # conjugate-gradient forms over the same window: the first
        # --vcycle-steps steps from the inputs.  Past ~300 steps the binary32
        # pressure's representation floor (~1.2-1.4e-5 on this world) keeps
        # the V-cycle, which iterates on that pressure, from reaching 1e-5 at
        # all; the conjugate-gradient forms refine a double-float iterate.
        sub = argparse.Namespace(**vars(args))
        sub.steps, sub.warmup, sub.settle_steps = args.vcycle_steps, 16, 0
        for mode in ("vcycle", "fp32", "mixed"):
            if mode != "vcycle" and mode not in results:
                continue
            dt_v, wv = timed_run(cfg, inputs, sub, mode)
            note(f"{mode} over the first steps: {1000.0 * dt_v / sub.steps:.3f} ms/step")
            sv, iv, mv = wv.timed_solver_stats
            it_v, rel_v = wv.mg_stats()
            res_v = wv.timed_residuals
            plain[mode] = {"value": cfg.cells() * sub.steps / dt_v, "ms_per_step": 1000.0 * dt_v / sub.steps,
                           "steps": sub.steps, "mg_tol": MODES[mode]["mg_tol"],
                           "mean_iterations": (iv / sv) if sv else None, "solves_missing_tol": mv,
                           "max_true_rel_residual": res_v["max"] if mode != "vcycle" else rel_v}
            if mode == "vcycle":
                plain[mode]["cap"] = args.vcycle_max_cycles
            wv.close()
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(cfg, inputs, args)
        note("cpu baseline done")
    f64 = None
    if not args.no_cpu and not args.no_fp64 and "fp32" in results and "mixed" in results:
        f64 = fp64_error(cfg, inputs, args)

    device = None
    if not args.no_device:
        try:
            device = device_ceiling()
        except (OSError, RuntimeError, AttributeError) as e:
            device = {"error": str(e)}
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "world step, 2048^2 grid / 1024^2 fluid, FP32 fluid, 8-level multigrid-preconditioned "
                        "CG to relative residual 1e-5, 15% rock + rubble, 16-row vent, Poisson(4) monomers in 8 "
                        "channels, exact u32 transport with the ledger checked every step",
            "grid": GRID,
            "fluid": FLUID,
            "mg_levels": len(w.mg_levels()),
            "mg_tiles": w.mg_tiles(),
            "mode": head,
            "mg_tol": MODES[head]["mg_tol"],
            "settle_steps": args.settle_steps,
            "overlap": "off (--no-overlap)" if args.no_overlap else
                       "each step's transport beside the next step's fluid: the scatter's tiles inside the "
                       "conjugate-gradient V-cycles' one-workgroup launches (fp32, mixed), a second stream "
                       "(reference)",
        },
        "modes": results,
        "mixed_vs_fp32": (results["mixed"]["value"] / results["fp32"]["value"]) if "mixed" in results and "fp32" in results else None,
        "plain_vcycle": plain or None,
        # ratios against the stationary V-cycle at the configuration's 1e-5
        # (every mode stops on a true relative residual <= 1e-5)
        # over the same first steps, every solve at a true residual <= 1e-5
        "mixed_vs_plain_vcycle": plain["mixed"]["value"] / plain["vcycle"]["value"]
        if "mixed" in plain and "vcycle" in plain else None,
        "fp32_vs_plain_vcycle": plain["fp32"]["value"] / plain["vcycle"]["value"]
        if "fp32" in plain and "vcycle" in plain else None,
        "error_vs_fp64": f64,
        "roofline": roofline,
        "device": device,
        "roofline_transport_scatter": roofline_scatter,
        "roofline_transport_scatter_valu": roofline_valu,
        "hbm_step_model": hbm,
        "light": light,
        "terrain": terrain,
        "agents": agents,
        "reference_default_profile": default_profile,
        "batched": batched_rec,
        "serving": serving_rec,
        "cpu_baseline": cpu,
        "kernel_avg_us": kernels,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
