"""Invariants and closed forms of the oracle's exact monomer transport
(transport.wgsl:119-415, chem.wgsl, ledger.wgsl)."""
import numpy as np

import ribossome_amd as ra
import oracle_ctypes as oc

G, FL = 64, 32
CFG = ra.WorldConfig(G, FL)


def quiet_params(**over):
    """No hops, no settling, no reactions."""
    p = ra.SimParams.default(1)
    p.diffusion = 0.0
    p.settle = 0.0
    p.reactivation_rate = 0.0
    p.activation_decay = 0.0
    p.uv_strength = 0.0
    p.aggregation = 0.0
    p.monomer_pressure = 0.0
    p.transport_every = 1
    p.fluid_enabled = 0
    for k, v in over.items():
        setattr(p, k, v)
    return p


def world(p, gamma=None, chem=None, vel=None, seed=3):
    w = oc.OracleWorld(CFG, 1)
    w.params = p
    w.upload(ra.GAMMA, np.zeros(CFG.cells(), np.uint32) if gamma is None else gamma)
    if chem is None:
        chem = ra.gen_chem(CFG, seed, 2.0, gamma)
    w.upload(ra.CHEM, chem)
    if vel is not None:
        w.upload(ra.VELOCITY, vel)
    return w, chem


def test_still_world_is_identity():
    w, chem = world(quiet_params())
    w.run_kernel("transport")
    assert np.array_equal(w.download(ra.CHEM), chem)


def test_integer_displacement_is_an_exact_shift_with_edge_clamp():
    # uniform v = 8 fluid cells/s, grid/fluid = 2, dt = 1/16: disp = 1 grid
    # cell exactly; every monomer moves one cell right, the last column keeps
    # what it had and receives its left neighbour (walls clamp, :368-369)
    p = quiet_params(dt=0.0625, fluid_enabled=1)
    vel = np.zeros(2 * CFG.fluid_cells(), np.float32)
    vel[0::2] = 8.0
    w, chem = world(p, vel=vel)
    w.run_kernel("transport")
    out = w.download(ra.CHEM).reshape(G, G, 4)
    c = chem.reshape(G, G, 4)
    assert np.array_equal(out[:, 1:-1], c[:, :-2])
    assert np.all(out[:, 0] == 0)
    assert np.array_equal(out[:, -1], c[:, -2] + c[:, -1])


def test_rock_blocks_entry_and_monomers_stay():
    p = quiet_params(dt=0.0625, fluid_enabled=1)
    vel = np.zeros(2 * CFG.fluid_cells(), np.float32)
    vel[0::2] = 8.0
    gamma = np.zeros((G, G), np.uint32)
    gamma[:, 40:42] = 6  # a rock wall two cells wide
    chem = ra.gen_chem(CFG, 4, 2.0, gamma.reshape(-1))
    w, chem = world(p, gamma=gamma.reshape(-1), chem=chem, vel=vel)
    w.run_kernel("transport")
    out = w.download(ra.CHEM).reshape(G, G, 4)
    c = chem.reshape(G, G, 4)
    # column 39 keeps its own monomers (blocked) and receives column 38
    assert np.array_equal(out[:, 39], c[:, 39] + c[:, 38])
    assert np.all(out[:, 40:42] == 0)
    assert ra.Ledger.from_cells(out).channels() == ra.Ledger.from_cells(chem).channels()


def test_more_than_256_movers_stay():
    p = quiet_params(dt=0.0625, fluid_enabled=1)
    vel = np.zeros(2 * CFG.fluid_cells(), np.float32)
    vel[0::2] = 8.0
    chem = np.zeros(4 * CFG.cells(), np.uint32)
    i = 10 * G + 10
    chem[4 * i + 2] = 300 | (5 << 16)  # 300 activated + 5 spent of base G
    w, _ = world(p, chem=chem, vel=vel)
    w.run_kernel("transport")
    out = w.download(ra.CHEM)
    # only the first 256 monomers move (activated first, :313-315): the 5
    # spent ones and 44 activated stay; the cell is over capacity, so the
    # squeeze hop probability 0.25 applies (:175-177) and movers that hop
    # back left also land at home
    home = out[4 * i + 2]
    assert home >> 16 == 5
    back = (home & 0xFFFF) - 44
    right = out[4 * (i + 1) + 2]
    assert 0 <= back < 256 and right + back + sum(int(out[4 * j + 2]) for j in (i + 2, i + 1 + G, i + 1 - G)) == 256
    assert abs(back - 256 * 0.25 * 0.25) < 5 * np.sqrt(256 * 0.0625 * 0.9375)
    assert ra.Ledger.from_cells(out).channels() == ra.Ledger.from_cells(chem).channels()


def test_hop_fraction_matches_p_diff():
    # still water: p_diff = 0.01 * mix(0.15, 1, 0) * diffusion = 0.15 at
    # diffusion 100; hops go to the 4 neighbours with equal weight
    p = quiet_params(diffusion=100.0)
    chem = np.zeros(4 * CFG.cells(), np.uint32)
    cells = [(y, x) for y in range(8, 56, 4) for x in range(8, 56, 4)]
    for (y, x) in cells:
        chem[4 * (y * G + x)] = 40
    w, _ = world(p, chem=chem)
    w.run_kernel("transport")
    out = w.download(ra.CHEM).reshape(G, G, 4)[..., 0].astype(np.int64)
    stayed = sum(out[y, x] for (y, x) in cells)
    total = 40 * len(cells)
    moved = np.array([sum(out[y + dy, x + dx] for (y, x) in cells) for dy, dx in ((0, 1), (0, -1), (1, 0), (-1, 0))])
    frac = (total - stayed) / total
    sd = np.sqrt(0.15 * 0.85 / total)
    assert abs(frac - 0.15) < 5 * sd
    assert moved.sum() == total - stayed
    assert np.all(np.abs(moved / moved.sum() - 0.25) < 5 * np.sqrt(0.25 * 0.75 / moved.sum()))


def test_full_physics_conserves_each_channel_exactly():
    cfg = ra.WorldConfig(128, 64)
    gamma = ra.gen_gamma(cfg, 5, 0.15, True)
    w = oc.OracleWorld(cfg, 5)
    p = ra.SimParams.default(5)
    p.transport_every = 1
    p.diffusion, p.settle, p.thermal_activation, p.aggregation = 30.0, 1.0, 1.0, 0.02
    p.monomer_pressure, p.activation_decay = 0.5, 0.01
    w.params = p
    s = ra.WorldSettings.default()
    s.fluid_substep = 1
    s.terrain_enabled = 0
    s.light_mode = 0
    w.set_settings(s)
    w.upload(ra.GAMMA, gamma)
    chem = ra.gen_chem(cfg, 5, 4.0, gamma)
    w.upload(ra.CHEM, chem)
    w.upload(ra.VELOCITY, ra.gen_velocity(cfg, 5, 30.0, gamma))
    w.upload(ra.HEAT_SRC, ra.gen_vent_heat(cfg, 4, 1500.0))
    before = ra.Ledger.from_cells(chem)
    w.run(60)
    out = w.download(ra.CHEM)
    after = ra.Ledger.from_cells(out)
    assert after.channels() == before.channels()
    assert after.act != before.act  # reactions ran
    assert np.count_nonzero(out != chem) > out.size // 10
    assert w.ledger() == after  # ledger_reduce == Ledger::from_cells


def test_thermal_activation_turns_spent_into_activated_in_place():
    p = quiet_params(thermal_activation=1.0)
    w, chem = world(p)
    temp = np.full(CFG.fluid_cells(), 8.0, np.float32)  # p_act = clamp(0.15 * 6 / 2) = 0.45
    w.upload(ra.TEMP, temp)
    w.run_kernel("thermal_activation")
    out = w.download(ra.CHEM)
    a0, a1 = chem & 0xFFFF, out & 0xFFFF
    s0, s1 = chem >> 16, out >> 16
    assert np.array_equal(a0 + s0, a1 + s1)
    d = a1.astype(np.int64) - a0
    assert d.min() >= 0 and d.max() <= 1
    eligible = s0 > 0
    assert abs(d[eligible].mean() - 0.45) < 0.02
