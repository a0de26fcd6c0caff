"""GPU parity: the HIP world against the CPU oracle on the same seeded inputs.

Every kernel computes each value with the oracle's operation sequence, so
the expected result is bit-identical buffers: integer monomer grids always,
and the binary32 fluid fields too.  A mismatch prints, per buffer, how many
elements differ and by how much before the assertion fails.
"""
import sys

import numpy as np
import pytest

import ribossome_amd as ra
import oracle_ctypes as oc
from worlds import FIELDS, bit_diff, load, make_inputs, parity_params, parity_settings

pytestmark = pytest.mark.gpu

CASES = {
    # 256^2 / 128^2, ~10 % rock, Poisson(4), 2 V-cycles to a 16^2 coarsest level
    "cfg1": dict(cfg=ra.WorldConfig(256, 128), seed=1, inputs=dict(rock=0.10),
                 params={}, settings=dict(mg_cycles=2, mg_min_size=16)),
    # rubble, vent band, strong diffusion, settle, osmotic pressure, cohesion,
    # thermal activation: every branch of transport and of the fluid forcing
    "rubble_vent": dict(cfg=ra.WorldConfig(128, 64), seed=3, inputs=dict(rock=0.15, rubble=True, vent_rows=6),
                        params=dict(diffusion=30.0, settle=1.0, thermal_activation=1.0, monomer_pressure=0.5,
                                    activation_decay=0.01, reactivation_rate=0.05), settings={}),
    # reference schedule defaults (substep 2, transport every 3, 1 V-cycle)
    "ref_schedule": dict(cfg=ra.WorldConfig(128, 64), seed=5, inputs=dict(rock=0.12, rubble=True, vent_rows=4),
                         params=dict(transport_every=3), settings=dict(fluid_substep=2)),
    # Jacobi pressure instead of multigrid
    "jacobi": dict(cfg=ra.WorldConfig(64, 32), seed=7, inputs=dict(rock=0.08, vent_rows=3),
                   params={}, settings=dict(multigrid=0, jacobi_iters=16)),
    # grid == fluid resolution, levels 96 48 24 12 6 on a 3*2^k side
    "odd_levels": dict(cfg=ra.WorldConfig(96, 96), seed=11, inputs=dict(rock=0.05, vent_rows=5),
                       params={}, settings=dict(mg_min_size=8)),
    # UV light (frozen field) and solar heating on
    "light": dict(cfg=ra.WorldConfig(64, 32), seed=13, inputs=dict(rock=0.1, vent_rows=2),
                  params=dict(uv_strength=3.0, sun_heat=1.0, direct_photoactivation=0.5), settings={}, light=True),
    # the reference's light step: a sweep on the first step, then one
    # propagation pass per step under the day cycle's slanted sun (eager
    # recording: the sun changes every step), water attenuation, solar heating
    "light_propagate": dict(cfg=ra.WorldConfig(128, 64), seed=15, inputs=dict(rock=0.12, rubble=True, vent_rows=3),
                            params=dict(uv_strength=3.0, uv_depth=6.0, sun_heat=1.0, direct_photoactivation=0.5,
                                        epoch=2000),
                            settings=dict(light_mode=1)),
    # constant sun (no day cycle): graph replay with three ping-pong passes per
    # step, and periodic full sweeps every third step
    "light_graph_rows": dict(cfg=ra.WorldConfig(128, 64), seed=17, inputs=dict(rock=0.1, vent_rows=2),
                             params=dict(uv_strength=3.0, uv_depth=4.0, day_period=0.0, direct_photoactivation=0.5),
                             settings=dict(light_mode=1, light_rows_per_step=3)),
    "light_graph_sweeps": dict(cfg=ra.WorldConfig(64, 32), seed=19, inputs=dict(rock=0.15, rubble=True),
                               params=dict(uv_strength=3.0, uv_depth=4.0, day_period=0.0, direct_photoactivation=0.5),
                               settings=dict(light_mode=1, light_rows_per_step=0, light_interval=3)),
    # the reference's default profile (SimParams::default(), WorldSettings::
    # default()): fluid every second step, transport every third, the light
    # step under the day cycle's moving sun with solar heating, the vent
    # cycle, and the terrain moving -- replayed as graphs with the sun of
    # each epoch from the device table
    "reference_default": dict(cfg=ra.WorldConfig(128, 64), seed=21, inputs=dict(rock=0.12, rubble=True, vent_rows=4),
                              params=dict(transport_every=3, uv_strength=3.0, sun_heat=1.0, epoch=2000),
                              settings=dict(fluid_substep=2, terrain_enabled=1, light_mode=1)),
}


def build(case, steps_light_seed=99):
    c = CASES[case]
    cfg = c["cfg"]
    inp = make_inputs(cfg, c["seed"], **c["inputs"])
    if c.get("light"):
        rng = np.random.default_rng(steps_light_seed)
        inp[ra.LIGHT] = rng.random(ra.buffer_bytes(cfg, ra.LIGHT) // 4).astype(np.float32)
    p = parity_params(c["seed"], **c["params"])
    s = parity_settings(**c["settings"])
    gw = ra.World(cfg, c["seed"])
    load(gw, inp, p, s)
    ow = oc.OracleWorld(cfg, c["seed"])
    load(ow, inp, p, s)
    return gw, ow


def compare(gw, ow, tag):
    report = []
    ok = True
    a, b = gw.download(ra.CHEM), ow.download(ra.CHEM)
    nchem = int(np.count_nonzero(a != b))
    report.append(f"chem words differing: {nchem}")
    ok &= nchem == 0
    for which in FIELDS + (ra.SOLID_MASK, ra.SLOPE, ra.LIGHT):
        a, b = gw.download(which), ow.download(which)
        nd = bit_diff(a, b)
        if nd:
            af, bf = a.astype(np.float64), b.astype(np.float64)
            rel = np.linalg.norm(af - bf) / max(np.linalg.norm(bf), 1e-30)
            report.append(f"{ra.BUFFER_NAMES[which]}: {nd} differ, max|d|={np.abs(af - bf).max():.3e}, relL2={rel:.3e}")
            ok = False
    print(f"[{tag}] " + "; ".join(report), flush=True)
    return ok


@pytest.mark.parametrize("case", list(CASES))
def test_world_step_bit_exact(case):
    print(f"case {case}", flush=True)
    gw, ow = build(case)
    assert compare(gw, ow, f"{case} initial")
    for k, n in enumerate((1, 1, 3, 7)):
        assert gw.encode_steps(n) == n and ow.encode_steps(n) == n
        assert compare(gw, ow, f"{case} after chunk {k}"), f"{case}: GPU differs from the oracle"
    assert gw.params.epoch == ow.params.epoch


def test_transport_only_bit_exact():
    """Transport with a frozen uploaded velocity: fluid_substep 4 and odd
    epochs, so the steps run transport but not the fluid."""
    print("case transport_only", flush=True)
    cfg = ra.WorldConfig(128, 64)
    inp = make_inputs(cfg, 21, rock=0.12, rubble=True, vcells_per_step=1.7)
    p = parity_params(21, epoch=1, diffusion=20.0, settle=1.0, monomer_pressure=0.3)
    s = parity_settings(fluid_substep=4)
    gw, ow = ra.World(cfg, 21), oc.OracleWorld(cfg, 21)
    load(gw, inp, p, s)
    load(ow, inp, p, s)
    gw.encode_steps(3)
    ow.encode_steps(3)
    a, b = gw.download(ra.CHEM), ow.download(ra.CHEM)
    assert np.array_equal(a, b), f"{np.count_nonzero(a != b)} chem words differ"
    assert ra.Ledger.from_cells(a).channels() == ra.Ledger.from_cells(inp[ra.CHEM]).channels()


def test_transport_extreme_displacements_bit_exact():
    """Transport under a frozen velocity with fast, huge and non-finite cells:
    displacements of tens of cells leave the LDS window, ones above 2^22
    cells take the saturating float-to-int bins, and NaN or infinite
    velocities are cleared by the sampler's sanitising, all as in the
    oracle."""
    print("case transport_extreme", flush=True)
    cfg = ra.WorldConfig(128, 64)
    inp = make_inputs(cfg, 23, rock=0.1, rubble=True, vcells_per_step=1.0)
    vel = inp[ra.VELOCITY].reshape(-1, 2).copy()
    rng = np.random.default_rng(23)
    n = vel.shape[0]
    fast = rng.choice(n, n // 8, replace=False)
    vel[fast] *= rng.uniform(20.0, 60.0, (fast.size, 1)).astype(np.float32)
    huge = rng.choice(n, n // 40, replace=False)
    vel[huge] = (rng.choice([-1.0, 1.0], (huge.size, 2)) * 10.0 ** rng.uniform(7.0, 19.5, (huge.size, 2))).astype(np.float32)
    bad = rng.choice(n, 12, replace=False)
    vel[bad[:4], 0] = np.nan
    vel[bad[4:8], 1] = np.inf
    vel[bad[8:], 0] = -np.inf
    inp[ra.VELOCITY] = vel.reshape(-1)
    p = parity_params(23, epoch=1, diffusion=10.0, settle=1.0, monomer_pressure=0.3, cohesion=0.2)
    s = parity_settings(fluid_substep=4)
    gw, ow = ra.World(cfg, 23), oc.OracleWorld(cfg, 23)
    load(gw, inp, p, s)
    load(ow, inp, p, s)
    for k in range(3):
        gw.encode_steps(1)
        ow.encode_steps(1)
        a, b = gw.download(ra.CHEM), ow.download(ra.CHEM)
        assert np.array_equal(a, b), f"step {k}: {np.count_nonzero(a != b)} chem words differ"
    assert ra.Ledger.from_cells(a).channels() == ra.Ledger.from_cells(inp[ra.CHEM]).channels()


def test_conservation_and_gpu_ledger():
    """tests/conservation.rs:30-69 on the HIP world: 600 steps with strong
    diffusion, settle, thermal activation and aggregation; exact per-channel
    totals, device-side check every step, GPU ledger == CPU count."""
    print("case conservation", flush=True)
    cfg = ra.WorldConfig(256, 128)
    inp = make_inputs(cfg, 7, rock=0.1, rubble=True, vent_rows=8)
    p = parity_params(7, diffusion=30.0, settle=1.0, thermal_activation=1.0, aggregation=0.02)
    s = parity_settings(ledger_check=1)
    gw = ra.World(cfg, 7)
    gw.params = p
    gw.upload(ra.GAMMA, inp[ra.GAMMA])
    for k in (ra.VELOCITY, ra.CHEM, ra.HEAT_SRC):
        gw.upload(k, inp[k])
    gw.settings = s
    seeded = ra.Ledger.from_cells(inp[ra.CHEM])
    assert gw.ledger_blocking() == seeded
    gw.run(600)
    cells = gw.read_cells_blocking()
    after = ra.Ledger.from_cells(cells)
    moved = int(np.count_nonzero(cells != inp[ra.CHEM]))
    print(f"moved {moved} of {cells.size}; act {sum(seeded.act)} -> {sum(after.act)}", flush=True)
    assert moved > cells.size // 10
    assert sum(seeded.act) != sum(after.act)
    assert after.channels() == seeded.channels()
    assert gw.ledger_blocking() == after
    assert gw.ledger_status() is None


def test_batch_invariance():
    """tests/conservation.rs:120-141: bit-identical results for 64-step and
    1-step batches and for two identical runs."""
    print("case batch_invariance", flush=True)
    cfg = ra.WorldConfig(128, 64)
    inp = make_inputs(cfg, 9, rock=0.1, rubble=True, vent_rows=4)

    def go(batch):
        w = ra.World(cfg, 9)
        load(w, inp, parity_params(9), parity_settings())
        done = 0
        while done < 128:
            done += w.encode_steps(min(batch, 128 - done))
        w.sync()
        return w.read_cells_blocking(), w.download(ra.VELOCITY)

    a = go(64)
    b = go(64)
    c = go(1)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[0], c[0])
    assert bit_diff(a[1], b[1]) == 0 and bit_diff(a[1], c[1]) == 0


def rel_l2(a, b):
    a, b = a.astype(np.float64), b.astype(np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.parametrize("precision", [0, 1])
def test_tolerance_mode_matches_oracle(precision):
    """Tolerance mode (preconditioned conjugate gradients with the symmetric
    V-cycle; precision 1 stores the preconditioner in bfloat16): the solve
    reaches the relative residual on the GPU, takes about as many iterations
    as the oracle, and the fields agree with the oracle's to a stated
    relative L2 (dot products are summed in a different order, so the bits
    differ; monomer totals stay exact)."""
    print(f"case tolerance precision={precision}", flush=True)
    cfg = ra.WorldConfig(256, 128)
    inp = make_inputs(cfg, 4, rock=0.15, rubble=True, vent_rows=4)
    s = parity_settings(mg_tol=1e-5, mg_max_cycles=120, precision=precision)
    gw = ra.World(cfg, 4)
    load(gw, inp, parity_params(4), s)
    ow = oc.OracleWorld(cfg, 4)
    load(ow, inp, parity_params(4), s)
    for k in range(4):
        gw.encode_steps(1)
        ow.encode_steps(1)
        cyc, rel = gw.mg_stats()
        ocyc, orel = ow.mg_stats()
        print(f"step {k}: GPU {cyc} iterations, true rel {rel:.3e}; oracle {ocyc} iterations, rel {orel:.3e}", flush=True)
        # the solve stops on the true residual of its double-float iterate,
        # evaluated in binary64: both sides reach the tolerance
        assert rel <= 1e-5 and orel <= 1e-5 and cyc < 120
        assert abs(cyc - ocyc) <= max(3, ocyc // 5)
        res, ores = gw.solver_residuals(), ow.solver_residuals()
        assert res["last"] == rel and res["last_rounded"] < 1e-4 and ores["last_rounded"] < 1e-4
    # Pressure is fixed only up to the residual: two solves that each stop
    # near relative residual 1e-5 differ by up to cond(A) times that in p,
    # so it gets a looser bound than the projected velocity.
    for which, tol in ((ra.VELOCITY, 1e-4), (ra.PRESSURE, 1e-3), (ra.TEMP, 1e-5)):
        e = rel_l2(gw.download(which), ow.download(which))
        print(f"{ra.BUFFER_NAMES[which]} rel L2 vs oracle {e:.2e}", flush=True)
        assert e < tol
    assert ra.Ledger.from_cells(gw.read_cells_blocking()).channels() == ra.Ledger.from_cells(inp[ra.CHEM]).channels()


def test_stationary_vcycle_mode_bit_exact():
    """The stationary tolerance mode repeats the reference V-cycle until the
    relative residual reaches the tolerance.  Its V-cycles are the reference
    mode's kernels, so GPU and oracle run the same number of cycles per solve
    and give bit-identical fields (the stopping residual is summed in another
    order, which moves it by ~1e-16 relative)."""
    print("case stationary", flush=True)
    cfg = ra.WorldConfig(128, 64)
    inp = make_inputs(cfg, 23, rock=0.12, rubble=True, vent_rows=3)
    s = parity_settings(mg_tol=1e-4, mg_max_cycles=300, mg_stationary=1)
    gw, ow = ra.World(cfg, 23), oc.OracleWorld(cfg, 23)
    load(gw, inp, parity_params(23), s)
    load(ow, inp, parity_params(23), s)
    for k in range(4):
        gw.encode_steps(1)
        ow.encode_steps(1)
        (cg, rg), (co, ro) = gw.mg_stats(), ow.mg_stats()
        print(f"step {k}: GPU {cg} V-cycles, rel {rg:.6e}; oracle {co} V-cycles, rel {ro:.6e}", flush=True)
        assert cg == co and 1 < cg < 300 and rg <= 1e-4
        assert abs(rg - ro) <= 1e-6 * ro
    assert compare(gw, ow, "stationary"), "stationary mode: GPU differs from the oracle"


@pytest.mark.parametrize("tol", [0.0, 1e-5])
def test_multigrid_tiles_do_not_change_results(tol):
    """The fused multigrid kernels give the same bits with 16- and 32-cell
    tiles on every level and with the autotuner's choice (a 512^2 world has
    two tile levels in the reference cycle, one in the symmetric one)."""
    print(f"case tiles tol={tol}", flush=True)
    cfg = ra.WorldConfig(512, 256)
    inp = make_inputs(cfg, 29, rock=0.12, rubble=True, vent_rows=4)
    extra = dict(mg_tol=tol, mg_max_cycles=40) if tol else {}
    out = []
    for force in (16, 32, 0):
        w = ra.World(cfg, 29)
        load(w, inp, parity_params(29), parity_settings(**extra))
        tiles = w.autotune(force)
        print(f"force {force}: tiles {tiles}", flush=True)
        assert all(t in (0, 16, 32) for t in tiles)
        w.encode_steps(3)
        out.append((w.read_cells_blocking(), w.download(ra.VELOCITY), w.download(ra.PRESSURE)))
    for o in out[1:]:
        assert np.array_equal(o[0], out[0][0])
        assert bit_diff(o[1], out[0][1]) == 0 and bit_diff(o[2], out[0][2]) == 0


@pytest.mark.parametrize("iters", [1, 2])
def test_tolerance_mode_preconditioner_matches_oracle(iters):
    """One or two conjugate-gradient iterations of the binary32 tolerance
    mode: the pressure is then x0 + alpha M r0 (+ a second direction), so it
    follows the preconditioner M -- the bilinear V-cycle over tile, block and
    level-0 kernels -- cell by cell.  GPU and oracle differ only in the
    summation order of the dot products behind alpha and beta (a few ulp of
    each scalar); a wrong transfer weight or halo shows up as an error many
    orders of magnitude larger."""
    print(f"case preconditioner iters={iters}", flush=True)
    cfg = ra.WorldConfig(256, 128)
    inp = make_inputs(cfg, 5, rock=0.15, rubble=True, vent_rows=4)
    s = parity_settings(mg_tol=1e-9, mg_max_cycles=iters, precision=0)
    gw = ra.World(cfg, 5)
    load(gw, inp, parity_params(5), s)
    ow = oc.OracleWorld(cfg, 5)
    load(ow, inp, parity_params(5), s)
    gw.encode_steps(1)
    ow.encode_steps(1)
    assert gw.mg_stats()[0] == iters and ow.mg_stats()[0] == iters
    for which in (ra.PRESSURE, ra.VELOCITY):
        e = rel_l2(gw.download(which), ow.download(which))
        print(f"{ra.BUFFER_NAMES[which]} rel L2 vs oracle after {iters} iterations: {e:.2e}", flush=True)
        assert e < 1e-5


@pytest.mark.parametrize("tol", [0.0, 1e-4])
def test_graph_replays_match_eager_recording(tol):
    """Captured HIP graphs, replayed over several chunks with host work and
    synchronisation in between, give the same bits as eager recording; the
    device-side epoch base advances like the host epoch."""
    print(f"case graphs tol={tol}", flush=True)
    cfg = ra.WorldConfig(128, 64)
    inp = make_inputs(cfg, 17, rock=0.12, rubble=True, vent_rows=4)
    extra = dict(mg_tol=tol, mg_max_cycles=12) if tol else {}
    worlds = []
    for no_graphs in (1, 0):
        w = ra.World(cfg, 17)
        load(w, inp, parity_params(17, aggregation=0.02, thermal_activation=1.0),
             parity_settings(no_graphs=no_graphs, ledger_check=1, **extra))
        worlds.append(w)
    for chunk in (5, 5, 17, 64, 64, 3):
        for w in worlds:
            assert w.encode_steps(chunk) == chunk
        worlds[1].sync()
        a, b = worlds[0].read_cells_blocking(), worlds[1].read_cells_blocking()
        assert np.array_equal(a, b), f"chunk {chunk}: chem differs"
        for which in (ra.VELOCITY, ra.PRESSURE, ra.TEMP):
            assert bit_diff(worlds[0].download(which), worlds[1].download(which)) == 0, ra.BUFFER_NAMES[which]
    assert worlds[0].params.epoch == worlds[1].params.epoch == 158
    assert worlds[0].ledger_status() is None and worlds[1].ledger_status() is None


@pytest.mark.parametrize("schedule", [(1, 1), (2, 3)])
def test_pipelined_steps_match_one_stream(schedule):
    """The reference schedule with the transport chain on a second stream
    beside the next fluid step (the default) gives the same bits as one
    stream, through graph capture and replay and eager recording, with steps
    that run only the fluid or only the transport (substep 2, transport
    every 3), and with velocity uploads between chunks."""
    substep, every = schedule
    print(f"case pipeline substep={substep} transport_every={every}", flush=True)
    cfg = ra.WorldConfig(128, 64)
    inp = make_inputs(cfg, 19, rock=0.12, rubble=True, vent_rows=4, vcells_per_step=1.2)
    worlds = []
    for pipeline, no_graphs in ((False, 0), (True, 0), (True, 1)):
        w = ra.World(cfg, 19)
        load(w, inp, parity_params(19, transport_every=every, thermal_activation=1.0, diffusion=5.0),
             parity_settings(fluid_substep=substep, no_graphs=no_graphs, ledger_check=1))
        w.set_pipeline(pipeline)
        worlds.append(w)
    rng = np.random.default_rng(19)
    for k, chunk in enumerate((7, 64, 13, 64, 5)):
        if k == 2:  # an upload between chunks: the transport's velocity copy follows it
            vel = (inp[ra.VELOCITY] * rng.uniform(0.5, 1.5)).astype(np.float32)
            for w in worlds:
                w.upload(ra.VELOCITY, vel)
        for w in worlds:
            assert w.encode_steps(chunk) == chunk
        ref = worlds[0]
        ref.sync()
        for w in worlds[1:]:
            w.sync()
            assert np.array_equal(ref.download(ra.CHEM), w.download(ra.CHEM)), f"chunk {k}: chem differs"
            for which in FIELDS:
                assert bit_diff(ref.download(which), w.download(which)) == 0, f"chunk {k}: {ra.BUFFER_NAMES[which]}"
    for w in worlds:
        assert w.ledger_status() is None


@pytest.mark.parametrize("precision,grid,cycles", [(0, 512, 60), (1, 512, 60), (0, 1024, 3)])
def test_fused_scatter_steps_match_one_stream(precision, grid, cycles):
    """The tolerance mode with each step's transport scatter run inside the
    next step's V-cycles, beside the one-workgroup coarse levels (the
    default), gives the same bits as the scatter on its own after each step,
    through graph capture and replay and eager recording, with chunks that
    end mid-graph, a 512^2 grid whose tiles take two V-cycles, a 1024^2 grid
    whose tiles need five V-cycles but get four (the rest run after the
    solve), and a velocity upload between chunks; the ledger is checked
    every step."""
    print(f"case fused scatter precision={precision} grid={grid} cycles={cycles}", flush=True)
    cfg = ra.WorldConfig(grid, grid // 2)
    inp = make_inputs(cfg, 23, rock=0.12, rubble=True, vent_rows=4, vcells_per_step=1.2)
    worlds = []
    for fused, no_graphs in ((False, 0), (True, 0), (True, 1)):
        w = ra.World(cfg, 23)
        load(w, inp, parity_params(23, thermal_activation=1.0, diffusion=5.0),
             parity_settings(mg_tol=1e-5, mg_max_cycles=cycles, precision=precision, no_graphs=no_graphs, ledger_check=1))
        w.set_pipeline(fused)
        worlds.append(w)
    rng = np.random.default_rng(23)
    for k, chunk in enumerate((7, 12, 9, 1, 16)):
        if k == 2:  # an upload between chunks
            vel = (inp[ra.VELOCITY] * rng.uniform(0.5, 1.5)).astype(np.float32)
            for w in worlds:
                w.upload(ra.VELOCITY, vel)
        for w in worlds:
            assert w.encode_steps(chunk) == chunk
        ref = worlds[0]
        ref.sync()
        for w in worlds[1:]:
            w.sync()
            assert np.array_equal(ref.download(ra.CHEM), w.download(ra.CHEM)), f"chunk {k}: chem differs"
            for which in FIELDS:
                assert bit_diff(ref.download(which), w.download(which)) == 0, f"chunk {k}: {ra.BUFFER_NAMES[which]}"
    for w in worlds:
        assert w.ledger_status() is None


def test_reference_default_profile_graphs_match_eager():
    """The reference's default profile (day cycle, solar heating, light
    propagation, moving terrain, substep 2, transport every 3) replays as
    graphs, the moving sun read per epoch from the device table, with the
    same bits as eager recording over chunks that cross day and night."""
    print("case reference default graphs", flush=True)
    c = CASES["reference_default"]
    cfg = c["cfg"]
    inp = make_inputs(cfg, c["seed"], **c["inputs"])
    worlds = []
    for no_graphs in (1, 0):
        w = ra.World(cfg, c["seed"])
        load(w, inp, parity_params(c["seed"], **dict(c["params"], epoch=5900)),
             parity_settings(no_graphs=no_graphs, ledger_check=1, **c["settings"]))
        worlds.append(w)
    for chunk in (5, 17, 64, 64, 64, 3):  # epochs 5900..6117: sunset of the 8000-epoch day at 6000
        for w in worlds:
            assert w.encode_steps(chunk) == chunk
        for which in (ra.CHEM, ra.GAMMA, ra.VELOCITY, ra.TEMP, ra.REDOX, ra.LIGHT):
            nd = bit_diff(worlds[0].download(which), worlds[1].download(which))
            assert nd == 0, f"chunk {chunk}: {ra.BUFFER_NAMES[which]} differs in {nd}"
    assert worlds[0].ledger_status() is None and worlds[1].ledger_status() is None


def test_idle_reports_finished_work():
    """ribo_world_idle: false while recorded steps run, true once they are
    done (bounded polling; sync() must agree)."""
    import time

    print("case idle", flush=True)
    cfg = ra.WorldConfig(256, 128)
    inp = make_inputs(cfg, 21, rock=0.1)
    w = ra.World(cfg, 21)
    load(w, inp, parity_params(21), parity_settings())
    w.sync()
    assert w.idle()
    w.encode_steps(64)
    t0 = time.time()
    while not w.idle():
        assert time.time() - t0 < 60.0
        time.sleep(0.001)
    w.sync()
    assert w.idle() and w.params.epoch == 64
