"""Statistics and closed forms for the oracle's transport decisions that the
invariants in test_oracle_transport.py leave open (transport.wgsl:95-117,
185-300): the Kawasaki acceptance of activated and spent hops, the bound
fraction p_bound with its shared smooth start, photoactivation with
sensitisation and uniform reactivation, shielded decay, osmotic pressure and
cohesion.  Each test sets up isolated patterns whose expected counts follow
from the WGSL formulas alone and checks them against several thousand
random draws, with tolerances of a few standard deviations."""
import numpy as np

import ribossome_amd as ra
import oracle_ctypes as oc

G, FL = 64, 32
CFG = ra.WorldConfig(G, FL)
DT = 0.0625


def params(**over):
    p = ra.SimParams.default(1)
    p.dt = p.fluid_dt = DT
    p.diffusion = 0.0
    p.settle = 0.0
    p.reactivation_rate = 0.0
    p.activation_decay = 0.0
    p.uv_strength = 0.0
    p.aggregation = 0.0
    p.monomer_pressure = 0.0
    p.cohesion = 0.0
    p.transport_every = 1
    p.fluid_enabled = 0
    for k, v in over.items():
        setattr(p, k, v)
    return p


def step(chem, p, epoch, vel=None, temp=None, light=None):
    w = oc.OracleWorld(CFG, 1)
    p.epoch = epoch
    w.params = p
    w.upload(ra.GAMMA, np.zeros(CFG.cells(), np.uint32))
    w.upload(ra.CHEM, chem.reshape(-1))
    if vel is not None:
        w.upload(ra.VELOCITY, vel)
    if temp is not None:
        w.upload(ra.TEMP, temp)
    if light is not None:
        w.upload(ra.LIGHT, light)
    w.run_kernel("transport")
    return w.download(ra.CHEM).reshape(G, G, 4)


def grid():
    return np.zeros((G, G, 4), np.uint32)


def act(c):
    return (c & 0xFFFF).astype(np.int64)


def spent(c):
    return (c >> 16).astype(np.int64)


def sites(spacing, lo=4, hi=G - 4):
    return [(y, x) for y in range(lo, hi, spacing) for x in range(lo, hi, spacing)]


def still():
    return np.zeros(2 * CFG.fluid_cells(), np.float32)


# ---- Kawasaki acceptance (transport.wgsl:222-260, 330-350) ---------------------
def test_activated_hops_to_fewer_neighbours_are_accepted_with_exp_minus_k_drop():
    # cell c: n activated; its lower-left diagonal: m activated.  The 3 x 3
    # block of c holds n + m, the blocks of c's right and upper neighbours
    # only n: drop m, acceptance exp(-k m), k = aggregation (room 1 below
    # 80 % of capacity, T = 0).  The right and upper neighbours receive only
    # from c.  Still water: p_diff = 0.01 mix(0.15, 1, 0) diffusion.
    n, m, a, diff = 12, 6, 0.1, 200.0
    p_diff = 0.01 * 0.15 * diff
    got = {0: 0, m: 0}
    cs = sites(6)
    for mm in (0, m):
        for epoch in range(4):
            c = grid()
            for (y, x) in cs:
                c[y, x, 0] = n
                c[y - 1, x - 1, 0] = mm
            out = step(c, params(aggregation=a, diffusion=diff), epoch)
            got[mm] += sum(act(out[y, x + 1, 0]) + act(out[y + 1, x, 0]) for (y, x) in cs)
    trials = 4 * len(cs) * n * 2  # two directions
    for mm in (0, m):
        want = trials * p_diff / 4.0 * np.exp(-a * mm)
        assert abs(got[mm] - want) < 4.0 * np.sqrt(want), (mm, got[mm], want)
    assert abs(got[m] / got[0] - np.exp(-a * m)) < 0.15


def test_spent_hops_into_more_activated_are_accepted_with_exp_minus_k_rise():
    # c: n spent; m activated two cells to the right (outside c's block,
    # inside its right neighbour's): rise m, acceptance exp(-k m) for spent
    # movers; spent arrivals at the right neighbour come only from c
    n, m, a, diff = 12, 8, 0.1, 200.0
    p_diff = 0.01 * 0.15 * diff
    cs = sites(6)
    got = 0
    for epoch in range(4):
        c = grid()
        for (y, x) in cs:
            c[y, x, 0] = n << 16
            c[y, x + 2, 0] = m
        out = step(c, params(aggregation=a, diffusion=diff), epoch)
        got += sum(spent(out[y, x + 1, 0]) for (y, x) in cs)
    want = 4 * len(cs) * n * p_diff / 4.0 * np.exp(-a * m)
    assert abs(got - want) < 4.0 * np.sqrt(want), (got, want)


# ---- p_bound and the shared smooth start (transport.wgsl:95-117, 252-260, 322-328) ----
def split_fraction(a, temp=None, epochs=4):
    # pairs of activated monomers, 4 cells apart (e0 - 1 = 1 companion), in a
    # uniform flow of half a grid cell per step (v = 4 fluid cells/s, grid/fluid
    # = 2, dt 1/16); no hops.  A free monomer lands right with probability 1/2
    # from its own start; bound ones share the cell's start, so a pair splits
    # with probability (1 - p^2) / 2, p = 1 - exp(-k min(e0 - 1, 48)).
    vel = np.zeros(2 * CFG.fluid_cells(), np.float32)
    vel[0::2] = 4.0
    cs = sites(4)
    split = 0
    for epoch in range(epochs):
        c = grid()
        for (y, x) in cs:
            c[y, x, 0] = 2
        out = step(c, params(aggregation=a, fluid_enabled=1), epoch, vel=vel, temp=temp)
        split += sum(int(act(out[y, x, 0]) == 1) for (y, x) in cs)
        assert act(out).sum() == 2 * len(cs)
    return split / (epochs * len(cs)), epochs * len(cs)


def test_bound_monomers_share_their_start_with_p_bound_from_the_companions():
    for a, temp, k in ((0.0, None, 0.0), (1.0, None, 1.0), (30.0, None, 30.0),
                       (1.0, np.full(CFG.fluid_cells(), 2.0, np.float32), 0.5)):  # T = 2: t_rel = 2
        p = 1.0 - np.exp(-k * 1.0) if a > 0 else 0.0
        want = 0.5 * (1.0 - p * p)
        frac, N = split_fraction(a, temp)
        assert abs(frac - want) < 4.0 * np.sqrt(max(want * (1 - want), 0.01) / N), (a, frac, want)


def test_spent_monomers_are_never_bound():
    # two spent monomers per cell: always free, split with probability 1/2
    vel = np.zeros(2 * CFG.fluid_cells(), np.float32)
    vel[0::2] = 4.0
    cs = sites(4)
    split = 0
    for epoch in range(4):
        c = grid()
        for (y, x) in cs:
            c[y, x, 0] = 2 << 16
        out = step(c, params(aggregation=30.0, fluid_enabled=1), epoch, vel=vel)
        split += sum(int(spent(out[y, x, 0]) == 1) for (y, x) in cs)
    N = 4 * len(cs)
    assert abs(split / N - 0.5) < 4.0 * np.sqrt(0.25 / N)


# ---- photoactivation and decay (transport.wgsl:271-300) ------------------------
def reaction_gain(chem_word, p, epochs=6, light=None):
    cs = sites(2, 2, G - 2)
    gain = []
    for epoch in range(epochs):
        c = grid()
        for (y, x) in cs:
            c[y, x, 1] = chem_word
        out = step(c, p, epoch, light=light)
        gain += [int(act(out[y, x, 1])) - int(chem_word & 0xFFFF) for (y, x) in cs]
        assert np.array_equal(act(out) + spent(out), act(c) + spent(c))  # in place, conserved
    return np.array(gain)


def test_uniform_reactivation_expected_count():
    # exp_act = min(spent, 8) tn r; na = floor + Bernoulli(fraction), at most min(spent, 8)
    r = 0.3
    for s, want in ((2, 0.6), (10, 2.4)):
        g = reaction_gain(s << 16, params(reactivation_rate=r))
        assert abs(g.mean() - want) < 5.0 * 0.5 / np.sqrt(g.size), (s, g.mean())
        assert set(np.unique(g)) <= {int(np.floor(want)), int(np.floor(want)) + 1}


def test_photoactivation_with_sensitisation():
    # exp_act = min(s, 8) (0.006 uv dp light sens + r), sens = 1 + 0.5 min(act, 8)
    L = 50.0
    light = np.full(ra.buffer_bytes(CFG, ra.LIGHT) // 4, L, np.float32)
    p = params(uv_strength=1.0, direct_photoactivation=1.0)
    for a0, want in ((0, 0.3), (2, 0.6)):
        g = reaction_gain((1 << 16) | a0, p, light=light)
        assert abs(g.mean() - want) < 5.0 * 0.5 / np.sqrt(g.size), (a0, g.mean(), want)


def test_shielded_decay():
    # one activated decays with probability min(a, 8) d / (1 + 0.5 (min(a, 8) - 1))
    d = 0.05
    for a0 in (1, 4):
        want = a0 * d / (1.0 + 0.5 * (a0 - 1))
        g = -reaction_gain(a0, params(activation_decay=d))
        assert set(np.unique(g)) <= {0, 1}
        assert abs(g.mean() - want) < 5.0 * np.sqrt(want * (1 - want) / g.size), (a0, g.mean(), want)


# ---- osmotic pressure and cohesion (transport.wgsl:185-206, 302-312, 336-344) ----
def arrivals(c, p, cs, epochs=4, chan=0, state=act):
    tot = np.zeros(4)
    for epoch in range(epochs):
        out = step(c, p, epoch)
        for (y, x) in cs:
            tot += [state(out[y, x + 1, chan]), state(out[y, x - 1, chan]), state(out[y + 1, x, chan]),
                    state(out[y - 1, x, chan])]
    return tot


def test_osmotic_pressure_weights_the_hop_directions_and_adds_hops():
    # c: 24 spent (half full) in channel 0; its right neighbour: 24 spent in
    # channel 1.  press = max(0.05, 1 + mp (f_self - f_n)): right 1, the three
    # empty sides 1 + 2 x 0.5 = 2; p_diff gains 0.01 mp max_drop
    mp, diff, n = 2.0, 100.0, 24
    p_diff = 0.01 * 0.15 * diff + 0.01 * mp * 0.5
    cs = sites(6)
    c = grid()
    for (y, x) in cs:
        c[y, x, 0] = n << 16
        c[y, x + 1, 1] = n << 16
    tot = arrivals(c, params(monomer_pressure=mp, diffusion=diff), cs, state=spent)
    N = 4 * len(cs) * n
    for k, wgt in enumerate((1.0, 2.0, 2.0, 2.0)):  # +x, -x, +y, -y
        want = N * p_diff * wgt / 7.0
        assert abs(tot[k] - want) < 4.0 * np.sqrt(want), (k, tot[k], want)


def test_cohesion_pulls_activated_towards_their_kind():
    # c: 12 activated and 12 spent of channel 0; right neighbour: 8 activated
    # of channel 0.  Activated movers weigh the right by 1 + 0.25 x 8 = 3
    # (the others 1); spent movers go all four ways alike
    diff, n = 100.0, 12
    p_diff = 0.01 * 0.15 * diff
    cs = sites(6)
    c = grid()
    for (y, x) in cs:
        c[y, x, 0] = n | (n << 16)
        c[y, x + 1, 0] = 8
    p = params(cohesion=0.25, diffusion=diff)
    epochs = 16
    ta = arrivals(c, p, cs, epochs=epochs)
    ts = arrivals(c, p, cs, epochs=epochs, state=spent)
    N = epochs * len(cs) * n
    for k in (1, 2, 3):  # -x, +y, -y: weight 1 of 6 for activated, 1 of 4 for spent
        wa, ws = N * p_diff / 6.0, N * p_diff / 4.0
        assert abs(ta[k] - wa) < 4.0 * np.sqrt(wa), (k, ta[k], wa)
        assert abs(ts[k] - ws) < 4.0 * np.sqrt(ws), (k, ts[k], ws)
    # over the three sides together: activated / spent = (1/6) / (1/4)
    ratio = ta[1:].sum() / ts[1:].sum()
    assert abs(ratio - 2.0 / 3.0) < 0.07, ratio
