"""The oracle's coupling points with the agents against closed forms:

- force_vectors (fluid.wgsl:280-296,392-398,410-432): a word of a*256 at one
  cell of a still, open world reaches the velocity as f_user * dt, with the
  9-point blur mixed in by a quarter: 8a/9 at the cell, a/36 at its four
  neighbours, a/144 at the diagonals; the words are cleared once used, and
  an agent word of -q cancels the cell's own buoyancy word q exactly.
- smooth_velocity (fluid.wgsl:680-693): constant fields stay, an impulse
  spreads 0.2 to the cell and its neighbours, edges count the cell twice,
  NaN is sanitised to 0.
- the monomer exchanges (chem.wgsl:40-98): each kind's effect on a word,
  stopping when the state runs out, ledger totals, and chem_open_cell's
  ring order on a rock patch.
"""
import numpy as np
import pytest

import ribossome_amd as ra
import oracle_ctypes as oc
from worlds import parity_params, parity_settings

CFG = ra.WorldConfig(64, 32)
F = 32
DT = np.float32(0.034)


def still_world(temp=None):
    w = oc.OracleWorld(CFG, 5)
    w.params = parity_params(5)
    w.set_settings(parity_settings())
    w.upload(ra.GAMMA, np.zeros(CFG.cells(), np.uint32))
    w.upload(ra.CHEM, np.zeros(4 * CFG.cells(), np.uint32))
    if temp is not None:
        w.upload(ra.TEMP, temp)
    return w


def test_agent_force_words_reach_the_velocity():
    w = still_world()
    x, y, a = 10, 12, 3.5
    words = np.zeros(2 * F * F, np.int32)
    words[2 * (y * F + x)] = int(a * 256)
    w.upload(ra.FORCE_VECTORS, words)
    w.run_kernel("forces")
    v = w.download(ra.VELOCITY).reshape(F, F, 2)
    expect = {(0, 0): 8 * a / 9, (1, 0): a / 36, (-1, 0): a / 36, (0, 1): a / 36, (0, -1): a / 36,
              (1, 1): a / 144, (-1, -1): a / 144, (1, -1): a / 144, (-1, 1): a / 144}
    for (dx, dy), f in expect.items():
        assert v[y + dy, x + dx, 0] == pytest.approx(f * DT, rel=2e-6), (dx, dy)
    assert np.count_nonzero(v[..., 0]) == 9 and not v[..., 1].any()
    assert not w.download(ra.FORCE_VECTORS).any()  # cleared after the gather
    w.run_kernel("forces")  # nothing left to add: the velocity stays
    assert np.array_equal(w.download(ra.VELOCITY).reshape(F, F, 2), v)


def test_agent_words_add_to_buoyancy_as_integers():
    """A cell 0.5 above the ambient (0) pushes up with 12 * 0.5 = 6, the word
    6 * 256 = 1536; an agent word of -1536 cancels it to zero exactly."""
    x, y = 9, 20
    temp = np.zeros(F * F, np.float32)
    temp[y * F + x] = 0.5
    plain = still_world(temp)
    plain.run_kernel("forces")
    vp = plain.download(ra.VELOCITY).reshape(F, F, 2)
    assert vp[y, x, 1] == pytest.approx(6.0 * 8 / 9 * DT, rel=2e-6)
    w = still_world(temp)
    words = np.zeros(2 * F * F, np.int32)
    words[2 * (y * F + x) + 1] = -1536
    w.upload(ra.FORCE_VECTORS, words)
    w.run_kernel("forces")
    v = w.download(ra.VELOCITY).reshape(F, F, 2)
    assert not v[..., 1].any()
    assert np.array_equal(v[..., 0], vp[..., 0])  # the sideways part of the buoyancy is untouched


def test_force_word_quantisation():
    f = np.array([1.0, -2.5, 1e6, -1e6, 0.001, 0.003], np.float32)
    assert ra.force_words(f).tolist() == [256, -640, 25_600_000, -25_600_000, 0, 1]


def test_smooth_velocity_closed_forms():
    w = still_world()
    v = np.zeros((F, F, 2), np.float32)
    v[..., 0] = 1.0
    v[..., 1] = -3.0
    w.upload(ra.VELOCITY, v.ravel())
    s = w.download(ra.VELOCITY_SMOOTH).reshape(F, F, 2)
    assert np.all(s[..., 0] == np.float32(1.0)) and np.all(s[..., 1] == np.float32(-3.0))
    v[:] = 0
    v[7, 11, 0] = 5.0  # interior impulse
    v[0, 0, 1] = 5.0   # corner impulse: the cell is its own left and lower neighbour
    v[20, 3, 0] = np.nan
    v[20, 5, 1] = 1e30  # beyond 1e20: sanitised
    w.upload(ra.VELOCITY, v.ravel())
    s = w.download(ra.VELOCITY_SMOOTH).reshape(F, F, 2)
    for yy, xx in ((7, 11), (6, 11), (8, 11), (7, 10), (7, 12)):
        assert s[yy, xx, 0] == np.float32(5.0) * np.float32(0.2)
    assert np.count_nonzero(s[..., 0]) == 5
    assert s[0, 0, 1] == np.float32(15.0) * np.float32(0.2)
    assert s[0, 1, 1] == s[1, 0, 1] == np.float32(5.0) * np.float32(0.2)
    assert np.count_nonzero(s[..., 1]) == 3


def word(act, spent):
    return (spent << 16) | act


def test_chem_exchanges_closed_forms():
    w = still_world()
    chem = np.zeros(4 * CFG.cells(), np.uint32)
    chem[4 * 100 + 1] = word(5, 2)
    chem[4 * 200 + 3] = word(0, 4)
    chem[4 * 300 + 0] = word(7, 0)
    w.upload(ra.CHEM, chem)
    before = w.ledger()
    ops = [(100, 1, ra.CHEM_SPEND, 3),        # 5/2 -> 2/5
           (100, 1, ra.CHEM_TAKE_ACT, 9),     # takes the last 2 activated: 0/5
           (200, 3, ra.CHEM_ACTIVATE, 10),    # 0/4 -> 4/0, stops after 4
           (200, 3, ra.CHEM_TAKE_SPENT, 1),   # none spent left: 0 moved
           (300, 0, ra.CHEM_ADD_SPENT, 6),    # 7/0 -> 7/6
           (300, 0, ra.CHEM_ADD_ACT, 0),
           (400, 2, ra.CHEM_DEPOSIT_ACT, 11)]  # open water: the cell itself
    done = w.chem_ops(ops)
    assert done.tolist() == [3, 2, 4, 0, 6, 0, 11]
    c = w.download(ra.CHEM)
    assert c[4 * 100 + 1] == word(0, 5) and c[4 * 200 + 3] == word(4, 0) and c[4 * 300] == word(7, 6)
    assert c[4 * 400 + 2] == word(11, 0)
    after = w.ledger()
    assert after.act[1] + after.spent[1] == before.act[1] + before.spent[1] - 2
    assert after.act[3] + after.spent[3] == before.act[3] + before.spent[3]
    assert after.act[0] + after.spent[0] == before.act[0] + before.spent[0] + 6
    assert after.act[2] == before.act[2] + 11
    with pytest.raises(RuntimeError):
        w.chem_ops([(CFG.cells(), 0, ra.CHEM_ADD_ACT, 1)])
    with pytest.raises(RuntimeError):
        w.chem_ops([(0, 4, ra.CHEM_ADD_ACT, 1)])
    with pytest.raises(RuntimeError):
        w.chem_ops([(0, 0, 8, 1)])
    assert np.array_equal(w.download(ra.CHEM), c)  # a rejected batch changes nothing


def test_deposit_ring_order():
    """A rock cell at (30, 30) inside a 5 x 5 rock patch with two open cells
    on the radius-2 ring: (29, 28) (row y - 2, found first) and (28, 32); a
    fully buried cell keeps its deposit."""
    G = CFG.grid_size
    gamma = np.zeros(CFG.cells(), np.uint32)
    gamma.reshape(G, G)[28:33, 28:33] = 3
    gamma.reshape(G, G)[28, 29] = 0
    gamma.reshape(G, G)[32, 28] = 1  # rubble holds monomers too
    gamma.reshape(G, G)[40:, :] = 5  # a buried band below
    w = still_world()
    w.upload(ra.GAMMA, gamma)
    done = w.chem_ops([(30 * G + 30, 1, ra.CHEM_DEPOSIT_SPENT, 2), (29 * G + 30, 0, ra.CHEM_DEPOSIT_ACT, 1),
                       (60 * G + 10, 2, ra.CHEM_DEPOSIT_ACT, 4), (30 * G + 30, 3, ra.CHEM_ADD_ACT, 1)])
    assert done.tolist() == [2, 1, 4, 1]
    c = w.download(ra.CHEM).reshape(G, G, 4)
    assert c[28, 29, 1] == word(0, 2)
    assert c[28, 29, 0] == word(1, 0)   # from (30, 29): the first cell of ring 1, (29, 28), is open
    assert c[60, 10, 2] == word(4, 0)   # 20 rows into the band: nothing open within 12 cells
    assert c[30, 30, 3] == word(1, 0)   # a plain add ignores the rock
    assert int(np.count_nonzero(c)) == 4


def test_redox_eaten_is_subtracted_before_decay_and_cleared():
    """redox_eaten (fluid.wgsl:360,371; fed by lifecycle.wgsl:821): on a
    uniform reductant field with no flow, no vent and no rock, advection and
    diffusion leave r unchanged, then r' = max(r - e / 10^4, 0) x
    exp(-redox_decay dt); copy_temperature clears the words (:365-372), so a
    second fluid step only decays."""
    dt = np.float32(0.0625)
    p = parity_params(5, vent_cycle_amp=0.0)
    p.dt = p.fluid_dt = float(dt)
    w = oc.OracleWorld(CFG, 5)
    w.params = p
    w.set_settings(parity_settings())
    w.upload(ra.GAMMA, np.zeros(CFG.cells(), np.uint32))
    w.upload(ra.REDOX, np.full(F * F, 4.0, np.float32))
    eaten = np.zeros(F * F, np.uint32)
    eaten[: F * F // 2] = 15000          # 1.5: r = 2.5
    eaten[F * F // 2:] = 60000           # 6.0 > r: clamps at 0
    w.upload(ra.REDOX_EATEN, eaten)
    w.run_kernel("temperature")
    r = w.download(ra.REDOX)
    decay = np.float32(oc.expf(float(-np.float32(p.redox_decay) * dt)))
    # the first row and column see the clamped neighbour of the other half,
    # which is equal inside each half: compare rows away from the seam
    assert np.all(r[: F * F // 2 - F] == np.float32(2.5) * decay)
    assert np.all(r[F * F // 2 + F:] == 0.0)
    assert not w.download(ra.REDOX_EATEN).any()
    # a plausible slip fails here: subtracting after the decay would give
    # 4 decay - 1.5 instead of (4 - 1.5) decay
    assert np.float32(2.5) * decay != np.float32(4.0) * decay - np.float32(1.5)
    w.run_kernel("temperature")
    r2 = w.download(ra.REDOX)
    assert np.allclose(r2[: F * F // 2 - 2 * F], np.float32(2.5) * decay * decay, rtol=1e-6)
