"""GPU: the batch (ribo_batch_*) steps many worlds with one launch per
kernel, and each world computes exactly what it computes alone.

The batched configuration is 64 independent worlds with sides drawn from
{128, 192, 256, 384, 512, 768, 1024}, rock U(5, 20) %, monomers
Poisson(U(2, 8)), the reference's V-cycle and the ledger checked every step.
Every buffer of every world is compared bit for bit with the same world
stepped on its own (World.run: its own stream, graphs, autotuned tiles).
"""
import numpy as np
import pytest

import ribossome_amd as ra

pytestmark = pytest.mark.gpu

SIDES = [128, 192, 256, 384, 512, 768, 1024]
COMPARED = (ra.CHEM, ra.VELOCITY, ra.PRESSURE, ra.TEMP, ra.REDOX, ra.DIVERGENCE, ra.NET_FLUX)


def batched_inputs(i: int, side: int):
    r = np.random.default_rng(1000 + i)
    rock, lam = float(r.uniform(0.05, 0.20)), float(r.uniform(2.0, 8.0))
    cfg = ra.WorldConfig(side, side // 2)
    gamma = ra.gen_gamma(cfg, 1000 + i, rock, True)
    vel = ra.gen_velocity(cfg, 1000 + i, 0.5 / 0.017, gamma)
    chem = ra.gen_chem(cfg, 1000 + i, lam, gamma)
    return cfg, {ra.GAMMA: gamma, ra.VELOCITY: vel, ra.CHEM: chem}


def make_world(cfg, seed, inputs, params, settings):
    w = ra.World(cfg, seed)
    w.params = params
    w.upload(ra.GAMMA, inputs[ra.GAMMA])
    for which, a in inputs.items():
        if which != ra.GAMMA:
            w.upload(which, a)
    w.settings = settings
    return w


def batched_profile(seed: int):
    p = ra.SimParams.default(seed)
    p.transport_every = 1
    p.uv_strength = 0.0
    p.sun_heat = 0.0
    s = ra.WorldSettings.default()
    s.fluid_substep = 1
    s.terrain_enabled = 0
    s.light_mode = 0
    s.mg_cycles = 1
    s.mg_tol = 0.0
    s.ledger_check = 1
    return p, s


def compare(tag, a: ra.World, b: ra.World, cfg: ra.WorldConfig):
    bad = []
    for which in COMPARED:
        x, y = a.download(which), b.download(which)
        n = int(np.count_nonzero(x.view(np.uint32) != y.view(np.uint32)))
        if n:
            bad.append(f"{ra.BUFFER_NAMES[which]}: {n} of {x.size} differ")
    assert not bad, f"{tag} (side {cfg.grid_size}): " + "; ".join(bad)


def test_batch_of_64_worlds_matches_each_world_alone():
    steps = 24
    print(f"case batch64 steps={steps}", flush=True)
    batch_worlds, alone, cfgs = [], [], []
    for i in range(64):
        side = SIDES[i % len(SIDES)]  # every side, 192, 384 and 768 included
        cfg, inp = batched_inputs(i, side)
        p, s = batched_profile(1000 + i)
        batch_worlds.append(make_world(cfg, 1000 + i, inp, p, s))
        a = make_world(cfg, 1000 + i, inp, p, s)
        if i < len(SIDES):
            a.autotune()
        alone.append(a)
        cfgs.append(cfg)
    b = ra.Batch(batch_worlds)
    assert b.encode_steps(steps) == steps
    b.sync()
    for a in alone:
        a.run(steps)
    for a in alone:
        a.sync()
    # one launch per kernel for all 64 worlds: about as many launches as
    # one world stepped alone records (some 20-30 a step)
    print(f"batch launches {b.launches()} for {steps} steps of 64 worlds", flush=True)
    assert b.launches() <= 40 * steps + 1
    for i, (w, a, cfg) in enumerate(zip(batch_worlds, alone, cfgs)):
        assert w.ledger_status() is None, f"world {i}: ledger mismatch in the batch"
        assert a.ledger_status() is None, f"world {i}: ledger mismatch alone"
        compare(f"world {i}", w, a, cfg)
    b.close()
    for w in batch_worlds + alone:
        w.close()


def test_batch_schedule_and_handoff_to_single_steps():
    """The reference schedule (fluid every second step, transport every
    third), two V-cycles, aggregation and solar heating under a constant
    sun; batch steps, then each world on its own, then the batch again,
    against the same worlds stepped only on their own."""
    print("case batch_handoff", flush=True)
    sides = [128, 192, 256, 384, 128, 192]
    bw, al, cfgs = [], [], []
    for i, side in enumerate(sides):
        cfg, inp = batched_inputs(100 + i, side)
        inp[ra.HEAT_SRC] = ra.gen_vent_heat(cfg, 3, 1500.0)
        p = ra.SimParams.default(200 + i)
        p.transport_every = 3
        p.uv_strength = 3.0
        p.sun_heat = 1.0
        p.day_period = 0.0
        p.aggregation = 0.5
        s = ra.WorldSettings.default()
        s.fluid_substep = 2
        s.mg_cycles = 2
        s.mg_tol = 0.0
        s.light_mode = 0
        s.terrain_enabled = 0
        s.ledger_check = 1
        bw.append(make_world(cfg, 200 + i, inp, p, s))
        al.append(make_world(cfg, 200 + i, inp, p, s))
        cfgs.append(cfg)
    b = ra.Batch(bw)
    b.encode_steps(7)
    for w in bw:
        w.encode_steps(5)
    b.encode_steps(4)
    b.sync()
    for w in bw:
        w.sync()
    for a in al:
        a.run(16)
        a.sync()
    for i, (w, a, cfg) in enumerate(zip(bw, al, cfgs)):
        assert w.params.epoch == a.params.epoch == 16
        assert w.ledger_status() is None and a.ledger_status() is None
        compare(f"world {i}", w, a, cfg)
    b.close()
    for w in bw + al:
        w.close()


def tolerance_profile(seed: int, precision: int):
    p, s = batched_profile(seed)
    s.mg_tol = 1e-5
    s.mg_max_cycles = 60
    s.precision = precision
    return p, s


@pytest.mark.parametrize("precision", [0, 1])
def test_batch_tolerance_mode_matches_each_world_alone(precision):
    """The tolerance mode's solve over a batch: each world takes its own
    iteration slots (chosen every 8 steps as it would alone) and the batch
    launches up to the largest; every field of every world is bit-identical
    to the world stepped alone."""
    steps = 20
    print(f"case batch tolerance precision={precision}", flush=True)
    sides = [128, 192, 256, 384, 128, 192, 256, 512]
    bw, al, cfgs = [], [], []
    for i, side in enumerate(sides):
        cfg, inp = batched_inputs(300 + i, side)
        p, s = tolerance_profile(300 + i, precision)
        bw.append(make_world(cfg, 300 + i, inp, p, s))
        al.append(make_world(cfg, 300 + i, inp, p, s))
        cfgs.append(cfg)
    b = ra.Batch(bw)
    assert b.encode_steps(steps) == steps
    b.sync()
    for a in al:
        a.run(steps)
        a.sync()
    for i, (w, a, cfg) in enumerate(zip(bw, al, cfgs)):
        (it, rel), (ait, arel) = w.mg_stats(), a.mg_stats()
        print(f"world {i} (side {cfg.grid_size}): batch {it} iterations rel {rel:.3e}; alone {ait} rel {arel:.3e}",
              flush=True)
        assert rel <= 1e-5 and w.ledger_status() is None
        compare(f"world {i}", w, a, cfg)
    b.close()
    for w in bw + al:
        w.close()


def test_batch_worlds_at_different_epochs():
    """Worlds admitted at different times step together (fluid and
    transport every step): a world that has run 5 steps alone joins a fresh
    one; each ends bit-identical to itself stepped alone with the same calls."""
    print("case batch different epochs", flush=True)
    cfg_a, inp_a = batched_inputs(401, 192)
    cfg_b, inp_b = batched_inputs(402, 256)
    pa, sa = tolerance_profile(401, 1)
    pb, sb = tolerance_profile(402, 1)
    a, a2 = make_world(cfg_a, 401, inp_a, pa, sa), make_world(cfg_a, 401, inp_a, pa, sa)
    c, c2 = make_world(cfg_b, 402, inp_b, pb, sb), make_world(cfg_b, 402, inp_b, pb, sb)
    a.run(5)
    a2.run(5)
    b = ra.Batch([a, c])
    assert b.encode_steps(11) == 11
    b.sync()
    a2.run(11)
    c2.run(11)
    for w in (a2, c2):
        w.sync()
    assert a.params.epoch == 16 and c.params.epoch == 11
    compare("joined late", a, a2, cfg_a)
    compare("fresh", c, c2, cfg_b)
    b.close()
    for w in (a, a2, c, c2):
        w.close()


@pytest.mark.parametrize("tol", [False, True])
def test_batch_graph_replays_match_eager(tol):
    """A batch captures each chunk (its length, slots, parities and schedule
    phase) as a graph the first time and replays it after: over several
    calls with host work in between, every world matches the same batch
    recorded eagerly (worlds with no_graphs)."""
    print(f"case batch graphs tol={tol}", flush=True)
    sides = [128, 192, 256, 128, 384]
    runs = []
    for no_graphs in (0, 1):
        ws = []
        for i, side in enumerate(sides):
            cfg, inp = batched_inputs(500 + i, side)
            p, s = tolerance_profile(500 + i, 1) if tol else batched_profile(500 + i)
            s.no_graphs = no_graphs
            ws.append((make_world(cfg, 500 + i, inp, p, s), cfg))
        b = ra.Batch([w for w, _ in ws])
        for _ in range(3):
            assert b.encode_steps(16) == 16
            b.sync()
            _ = ws[0][0].download(ra.PRESSURE)  # host work between replays
        runs.append((b, ws))
    for (wa, cfg), (wb, _) in zip(runs[0][1], runs[1][1]):
        compare("graphs vs eager", wa, wb, cfg)
        assert wa.ledger_status() is None
    for b, ws in runs:
        b.close()
        for w, _ in ws:
            w.close()


def test_batch_refuses_worlds_outside_its_profile():
    print("case batch_refusals", flush=True)
    cfg, inp = batched_inputs(7, 128)
    p, s = batched_profile(7)
    ok = make_world(cfg, 7, inp, p, s)
    tol = make_world(cfg, 7, inp, p, s)
    st = tol.settings
    st.mg_tol = 1e-5
    st.mg_stationary = 1
    st.mg_max_cycles = 40
    tol.settings = st
    sub = make_world(cfg, 7, inp, p, s)
    st = sub.settings
    st.fluid_substep = 2
    sub.settings = st
    sun = make_world(cfg, 7, inp, p, s)
    ps = sun.params
    ps.sun_heat = 1.0
    ps.day_period = 8000.0
    sun.params = ps
    for bad, what in ((tol, "stationary tolerance mode"), (sun, "moving sun")):
        with pytest.raises(ra.RiboError):
            ra.Batch([ok, bad])
        print(f"refused: {what}", flush=True)
    with pytest.raises(ra.RiboError):
        ra.Batch([ok, sub])
    with pytest.raises(ra.RiboError):
        ra.Batch([ok, ok])
    # a parameter changed while in a batch
    b = ra.Batch([ok])
    ps = ok.params
    ps.dt = ps.dt * 0.5
    ok.params = ps
    with pytest.raises(ra.RiboError):
        b.encode_steps(1)
    b.close()
    for w in (ok, tol, sub, sun):
        w.close()
