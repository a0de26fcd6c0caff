"""The coupling points with the agents on the HIP world against the oracle:
force words uploaded between steps (consumed by the next fluid step, with
graph replays before and after), the smoothed velocity the agents read, and
the monomer exchanges (bit-exact where no two exchanges compete for a word;
where they do, the totals that no order can change)."""
import numpy as np
import pytest

import ribossome_amd as ra
import oracle_ctypes as oc
from test_gpu_parity import build, compare
from worlds import bit_diff, load, make_inputs, parity_params, parity_settings

pytestmark = pytest.mark.gpu


def agent_words(cfg, seed, frac=0.05, amp=4000):
    rng = np.random.default_rng(seed)
    F = cfg.fluid_cells()
    words = np.zeros(2 * F, np.int32)
    cells = rng.choice(F, int(frac * F), replace=False)
    words[2 * cells] = rng.integers(-amp, amp, cells.size)
    words[2 * cells + 1] = rng.integers(-amp, amp, cells.size)
    return words


def compare_agent_buffers(gw, ow, tag):
    ok = True
    for which in (ra.FORCE_VECTORS, ra.VELOCITY_SMOOTH):
        nd = bit_diff(gw.download(which), ow.download(which))
        print(f"[{tag}] {ra.BUFFER_NAMES[which]}: {nd} differ", flush=True)
        ok &= nd == 0
    return ok


def test_force_words_and_smoothed_velocity_bit_exact():
    """The reference schedule (fluid every second step, transport every
    third): words uploaded on an epoch without a fluid step wait for the next
    one; the steps around them replay graphs."""
    print("case agent forces", flush=True)
    gw, ow = build("ref_schedule")
    cfg = ra.WorldConfig(128, 64)
    gw.encode_steps(4)
    ow.encode_steps(4)
    assert compare(gw, ow, "before") and compare_agent_buffers(gw, ow, "before")
    for k, n in enumerate((1, 2, 5, 8)):
        words = agent_words(cfg, 40 + k)
        gw.upload(ra.FORCE_VECTORS, words)
        ow.upload(ra.FORCE_VECTORS, words)
        assert compare_agent_buffers(gw, ow, f"uploaded {k}")
        assert gw.encode_steps(n) == n and ow.encode_steps(n) == n
        assert compare(gw, ow, f"after {k}") and compare_agent_buffers(gw, ow, f"after {k}")
    assert not gw.download(ra.FORCE_VECTORS).any()


def exchange_world(seed, ledger=True):
    cfg = ra.WorldConfig(128, 64)
    inp = make_inputs(cfg, seed, rock=0.12, rubble=True, vent_rows=3)
    p = parity_params(seed, diffusion=10.0, thermal_activation=1.0)
    s = parity_settings(ledger_check=1 if ledger else 0)
    gw, ow = ra.World(cfg, seed), oc.OracleWorld(cfg, seed)
    load(gw, inp, p, s)
    load(ow, inp, p, s)
    return cfg, inp, gw, ow


def test_exchanges_bit_exact_and_ledger():
    """Exchanges on distinct words, all kinds, then steps: the same grid and
    moved counts as the oracle, and the device's per-step ledger check
    follows the matter that left and entered."""
    print("case exchanges", flush=True)
    cfg, inp, gw, ow = exchange_world(31)
    rng = np.random.default_rng(31)
    words = rng.choice(4 * cfg.cells(), 20_000, replace=False)
    ops = np.stack([words // 4, words % 4, rng.integers(0, 8, words.size), rng.integers(0, 6, words.size)], 1)
    # deposits into rock move to an open cell: keep them off the other words
    dep = ops[:, 2] >= ra.CHEM_DEPOSIT_ACT
    ops[dep, 2] -= 2  # plain adds there instead
    ops = ops.astype(np.uint32)
    dg, do = gw.chem_ops(ops), ow.chem_ops(ops)
    assert np.array_equal(dg, do)
    assert np.array_equal(gw.download(ra.CHEM), ow.download(ra.CHEM))
    assert dg.sum() > 0
    before = ra.Ledger.from_cells(inp[ra.CHEM]).channels()
    after = ra.Ledger.from_cells(gw.download(ra.CHEM)).channels()
    assert after != before
    for n in (1, 3, 6):
        gw.encode_steps(n)
        ow.encode_steps(n)
        assert compare(gw, ow, f"exchanges, then {n} steps")
    assert gw.ledger_status() is None
    assert ra.Ledger.from_cells(gw.download(ra.CHEM)).channels() == after


def test_competing_exchanges_keep_the_totals():
    """Takes competing for a few words: whatever the order, the word gives
    min(held, asked) in total; spends turn exactly that many activated into
    spent; deposits into rock land in the same open cell as in the oracle."""
    print("case competing exchanges", flush=True)
    cfg, inp, gw, _ = exchange_world(33)
    G = cfg.grid_size
    chem = inp[ra.CHEM].copy()
    hot = np.array([4 * (10 * G + 10) + 1, 4 * (50 * G + 70) + 2, 4 * (90 * G + 33)])
    chem[hot] = [(3 << 16) | 40, (0 << 16) | 7, (9 << 16) | 0]
    gw.upload(ra.CHEM, chem)
    ops = []
    for _ in range(300):
        ops.append((10 * G + 10, 1, ra.CHEM_TAKE_ACT, 1))
        ops.append((50 * G + 70, 2, ra.CHEM_SPEND, 2))
        ops.append((90 * G + 33, 0, ra.CHEM_ACTIVATE, 1))
    gamma = inp[ra.GAMMA]
    rock = np.flatnonzero(gamma >= 3)[:50]
    for c in rock:
        ops.append((int(c), 3, ra.CHEM_DEPOSIT_SPENT, 2))
    ops = np.array(ops, np.uint32)
    done = gw.chem_ops(ops)
    c = gw.download(ra.CHEM)
    takes, spends, acts = done[0:900:3], done[1:900:3], done[2:900:3]
    assert takes.sum() == 40 and set(takes.tolist()) <= {0, 1}
    assert spends.sum() == 7 and acts.sum() == 9
    assert c[hot[0]] == (3 << 16) and c[hot[1]] == (7 << 16) and c[hot[2]] == 9
    # the deposits alone, one after another on the oracle: same cells
    base = c.copy()
    o = oc.OracleWorld(cfg, 33)
    load(o, {ra.GAMMA: gamma, ra.CHEM: chem}, parity_params(33), parity_settings())
    o.chem_ops(ops[900:])
    want = o.download(ra.CHEM) - chem  # the deposits' increments
    got = base - chem
    got[hot] = 0
    assert np.array_equal(got, want)
    gw.encode_steps(3)
    assert gw.ledger_status() is None


def test_coupling_points_reject_bad_input():
    """Exchanges out of range are refused before anything changes; the
    smoothed velocity cannot be uploaded; force words must be 2F words."""
    print("case agent errors", flush=True)
    cfg, inp, gw, _ = exchange_world(35, ledger=False)
    before = gw.download(ra.CHEM)
    assert gw.chem_ops(np.zeros((0, 4), np.uint32)).size == 0
    for bad in ([(cfg.cells(), 0, ra.CHEM_ADD_ACT, 1)], [(0, 4, ra.CHEM_ADD_ACT, 1)], [(0, 0, 8, 1)],
                [(0, 0, ra.CHEM_ADD_SPENT, 65536)], [(0, 0, ra.CHEM_DEPOSIT_ACT, 70000)]):
        with pytest.raises(ra.RiboError):
            gw.chem_ops([(5, 1, ra.CHEM_ADD_ACT, 3)] + bad)
    assert np.array_equal(gw.download(ra.CHEM), before)
    with pytest.raises(ra.RiboError):
        gw.upload(ra.VELOCITY_SMOOTH, np.zeros(2 * cfg.fluid_cells(), np.float32))
    with pytest.raises(ra.RiboError):
        gw.upload(ra.FORCE_VECTORS, np.zeros(cfg.fluid_cells(), np.int32))


def test_redox_eaten_words_bit_exact_across_graph_replay():
    """The oxidant the agents ate (RIBO_BUF_REDOX_EATEN, fluid.wgsl:360,371,
    fed by lifecycle.wgsl:821): words uploaded between steps, on epochs with
    and without a fluid step of the reference schedule (fluid every second
    step), are subtracted by the next fluid step and then cleared; the steps
    around them replay graphs.  Reductant and every other field stay
    bit-identical to the oracle's."""
    print("case redox eaten", flush=True)
    gw, ow = build("ref_schedule")
    _, twin = build("ref_schedule")  # the same world with no words eaten
    cfg = ra.WorldConfig(128, 64)
    F = cfg.fluid_cells()
    # reductant to eat: a uniform field of 3.0 plus what the vents add
    for w in (gw, ow, twin):
        w.upload(ra.REDOX, np.full(F, 3.0, np.float32))
        w.encode_steps(4)
    assert compare(gw, ow, "eaten before")
    rng = np.random.default_rng(77)
    for k, n in enumerate((1, 2, 5, 8)):
        words = np.zeros(F, np.uint32)
        cells = rng.choice(F, F // 8, replace=False)
        words[cells] = rng.integers(1, 40000, cells.size)
        gw.upload(ra.REDOX_EATEN, words)
        ow.upload(ra.REDOX_EATEN, words)
        assert np.array_equal(gw.download(ra.REDOX_EATEN), words)
        assert gw.encode_steps(n) == n and ow.encode_steps(n) == n and twin.encode_steps(n) == n
        assert compare(gw, ow, f"eaten {k}, after {n} steps"), f"eaten {k}: GPU differs from the oracle"
        assert not gw.download(ra.REDOX_EATEN).any() and not ow.download(ra.REDOX_EATEN).any()
    # the words took reductant away: the twin that ate nothing holds more
    assert ow.download(ra.REDOX).astype(np.float64).sum() < twin.download(ra.REDOX).astype(np.float64).sum()
