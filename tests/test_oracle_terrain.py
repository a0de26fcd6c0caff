"""The oracle's terrain grain dynamics (terrain.wgsl:38-102,159-437) against
closed forms and statistics the rules fix:

- grain_fall: on a tick of the fixed-point clock every loose run without
  support drops one cell as a block; rock and the floor support; off-tick
  epochs move nothing.  The monomers that no longer fit the cell a grain
  lands in leave for the grain's source first, spent before activated.
- the sand pile: a tower 3 or more above its lowest 4-neighbour sheds one
  grain with probability 0.12 * max(settle, crumble), always to the first
  lowest neighbour in the order x-1, x+1, y-1, y+1.
- loose grains: an isolated grain with no current explores one of its 8
  neighbours with probability 0.05 (mobility 0.25^0), uniformly; a bonded
  pair without current stays.
- grains and monomers are conserved over long runs with every process on.
"""
import numpy as np
import pytest

import ribossome_amd as ra
import oracle_ctypes as oc
from worlds import load, make_inputs, parity_params, parity_settings


def terrain_world(cfg, gamma, chem=None, seed=3, **params):
    w = oc.OracleWorld(cfg, seed)
    p = parity_params(seed, fluid_enabled=0, sediment_settle=0.0, sediment_compaction=0.0, rock_crumble=0.0)
    for k, v in params.items():
        setattr(p, k, v)
    w.params = p
    w.set_settings(parity_settings(terrain_enabled=1))
    w.upload(ra.GAMMA, gamma)
    w.upload(ra.CHEM, np.zeros(4 * cfg.cells(), np.uint32) if chem is None else chem)
    return w


def test_grain_fall_drops_unsupported_runs_one_cell():
    cfg = ra.WorldConfig(64, 32)
    G, x = 64, 10
    gamma = np.zeros(G * G, np.uint32)
    col = {0: 0, 1: 3, 2: 0, 3: 2, 4: 1, 5: 0, 6: 2, 9: 1, 10: 5}
    for y, n in col.items():
        gamma[y * G + x] = n
    chem = np.zeros(4 * G * G, np.uint32)
    chem[4 * (2 * G + x)] = (2 << 16) | 18  # 2 spent + 18 activated where two grains land
    # settle 1: rate = 0.5 * (64 / 32) * 0.017 = 0.017, clock word 1114 per epoch: epoch 58 ticks, 0 does not
    w = terrain_world(cfg, gamma, chem, sediment_settle=1.0, epoch=0)
    w.run_kernel("grain_fall")
    assert np.array_equal(w.download(ra.GAMMA), gamma)
    w.set_param(epoch=58)
    w.run_kernel("grain_fall")
    g = w.download(ra.GAMMA).reshape(G, G)[:, x]
    want = {1: 3, 2: 2, 3: 1, 5: 2, 8: 1, 10: 5}  # rows 3-4, 6 and 9 drop one cell; rock stays
    assert {y: int(v) for y, v in enumerate(g) if v} == want
    c = w.download(ra.CHEM).reshape(G, G, 4)
    # the second grain into row 2 leaves room for 16: 4 monomers go to the source (row 3), spent first
    assert c[2, x, 0] == 16 and c[3, x, 0] == (2 << 16) | 2
    assert int(np.count_nonzero(c)) == 2


def test_grain_fall_needs_settling():
    cfg = ra.WorldConfig(64, 32)
    gamma = np.zeros(64 * 64, np.uint32)
    gamma[20 * 64 + 5] = 1
    w = terrain_world(cfg, gamma, sediment_settle=0.0, epoch=58)
    w.run_kernel("grain_fall")
    assert np.array_equal(w.download(ra.GAMMA), gamma)


def spaced(G, step, lo=2):
    pos = [(x, y) for y in range(lo, G - lo, step) for x in range(lo, G - lo, step)]
    return pos


def test_sand_pile_sheds_to_the_first_lowest_neighbour():
    cfg = ra.WorldConfig(256, 128)
    G = 256
    pos = spaced(G, 4)
    gamma = np.zeros(G * G, np.uint32)
    for x, y in pos:
        gamma[y * G + x] = 9
    w = terrain_world(cfg, gamma, rock_crumble=1.0, epoch=77)
    w.run_kernel("relax_a")
    g = w.download(ra.GAMMA).reshape(G, G)
    moved = 0
    for x, y in pos:
        if g[y, x] == 8:
            moved += 1
            assert g[y, x - 1] == 1 and g[y, x + 1] == 0 and g[y - 1, x] == 0 and g[y + 1, x] == 0
        else:
            assert g[y, x] == 9 and g[y, x - 1] == 0
    assert int(g.sum()) == 9 * len(pos)
    n = len(pos)
    assert abs(moved / n - 0.12) < 4 * np.sqrt(0.12 * 0.88 / n), moved
    # without settling or crumbling the pile is frozen
    w0 = terrain_world(cfg, gamma, epoch=77)
    w0.run_kernel("relax_a")
    assert np.array_equal(w0.download(ra.GAMMA), gamma)


def test_lone_grains_explore_at_five_percent():
    cfg = ra.WorldConfig(256, 128)
    G = 256
    pos = spaced(G, 4)
    gamma = np.zeros(G * G, np.uint32)
    for x, y in pos:
        gamma[y * G + x] = 1
    w = terrain_world(cfg, gamma, epoch=5)
    w.run_kernel("relax_b")
    g = w.download(ra.GAMMA).reshape(G, G)
    dirs = {}
    for x, y in pos:
        if g[y, x] == 1:
            continue
        nb = [(dx, dy) for dy in (-1, 0, 1) for dx in (-1, 0, 1) if (dx or dy) and g[y + dy, x + dx] == 1]
        assert len(nb) == 1
        dirs[nb[0]] = dirs.get(nb[0], 0) + 1
    moved, n = sum(dirs.values()), len(pos)
    assert int(g.sum()) == n
    assert abs(moved / n - 0.05) < 4 * np.sqrt(0.05 * 0.95 / n), moved
    assert len(dirs) == 8 and all(moved / 8 * 0.4 < c < moved / 8 * 1.6 for c in dirs.values()), dirs


def test_bonded_pairs_without_current_stay():
    cfg = ra.WorldConfig(128, 64)
    G = 128
    gamma = np.zeros(G * G, np.uint32)
    for x, y in spaced(G, 5):
        gamma[y * G + x] = 1
        gamma[y * G + x + 1] = 2
    for ph, e in (("relax_a", 1), ("relax_b", 2), ("relax_a", 3)):
        w = terrain_world(cfg, gamma, epoch=e)
        w.run_kernel(ph)
        assert np.array_equal(w.download(ra.GAMMA), gamma)


def test_tolerance_mode_with_moving_terrain():
    """The tolerance mode follows the moving grains: the coarsest inverse is
    rebuilt from the current mask at every solve, every solve reaches the
    tolerance on its true residual, and grains and monomers are conserved."""
    cfg = ra.WorldConfig(64, 32)
    inp = make_inputs(cfg, 6, rock=0.12, rubble=True, vent_rows=3, vcells_per_step=2.0)
    p = parity_params(6, sediment_settle=1.0, sediment_compaction=0.3, rock_crumble=1.0, sediment_transport=3.0,
                      sediment_threshold=0.05)
    w = oc.OracleWorld(cfg, 6)
    load(w, inp, p, parity_settings(terrain_enabled=1, mg_tol=1e-5, mg_max_cycles=60))
    for _ in range(6):
        w.encode_steps(1)
        it, rel = w.mg_stats()
        assert 0 < it < 60 and rel <= 1e-5
    g = w.download(ra.GAMMA)
    assert int(g.sum()) == int(inp[ra.GAMMA].sum()) and int(np.count_nonzero(g != inp[ra.GAMMA])) > 0
    assert ra.Ledger.from_cells(w.download(ra.CHEM)).channels() == ra.Ledger.from_cells(inp[ra.CHEM]).channels()


@pytest.mark.parametrize("seed", [4, 9])
def test_terrain_conserves_grains_and_monomers(seed):
    cfg = ra.WorldConfig(128, 64)
    inp = make_inputs(cfg, seed, rock=0.12, rubble=True, vent_rows=4, vcells_per_step=2.0)
    p = parity_params(seed, sediment_settle=1.0, sediment_compaction=0.3, rock_crumble=1.0, sediment_transport=3.0,
                      sediment_threshold=0.05, diffusion=5.0)
    s = parity_settings(terrain_enabled=1)
    w = oc.OracleWorld(cfg, seed)
    load(w, inp, p, s)
    w.run(120)
    g = w.download(ra.GAMMA)
    assert int(g.sum()) == int(inp[ra.GAMMA].sum())
    assert int(np.count_nonzero(g != inp[ra.GAMMA])) > 100
    assert w.ledger().channels() == ra.Ledger.from_cells(inp[ra.CHEM]).channels()
