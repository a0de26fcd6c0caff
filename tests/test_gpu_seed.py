"""The world-level forms of the reference's host preparation on the HIP
world: World::upload_heat from fumaroles and images, restart_keeping_terrain
(seed_cells + fit_matter_to_terrain on the world's terrain) and
apply_terrain_live (refit_matter_to_terrain, then both grids uploaded),
each equal to the host functions and followed by steps that keep the exact
per-channel ledger."""
import numpy as np
import pytest

import ribossome_amd as ra
from worlds import load, make_inputs, parity_params, parity_settings

pytestmark = pytest.mark.gpu

CFG = ra.WorldConfig(256, 128)


def test_vents_restart_and_live_terrain():
    print("case seed and vents", flush=True)
    inp = make_inputs(CFG, 41, rock=0.12, rubble=True)
    w = ra.World(CFG, 41)
    load(w, inp, parity_params(41), parity_settings(ledger_check=1))
    # vents: two fumaroles and a heat image, default gain
    fs = [(0.3, 0.05, 40.0, 600.0), (0.7, 0.08, 25.0, 400.0)]
    img = np.zeros(CFG.cells(), np.float32)
    img[:256 * 4] = 0.5  # a hot bottom band of 4 grid rows
    w.set_vents(fs, 0.3, heat_image=img)
    r = ra.rasterize_fumaroles(CFG, fs)
    want = ra.vent_heat_src(CFG, r + img, r + img, 0.3)
    assert np.array_equal(w.download(ra.HEAT_SRC), want)
    # restart the matter on the current terrain
    led = w.restart_matter(9, 0.83, 0.5)
    cells = w.read_cells_blocking()
    host = ra.fit_matter_to_terrain(ra.seed_cells(CFG, 9, 0.83, 0.5), inp[ra.GAMMA], 9)
    assert np.array_equal(cells, host)
    assert led == ra.Ledger.from_cells(cells) and w.params.epoch == 0
    w.run(16)
    w.sync()
    assert w.ledger_status() is None and w.ledger_blocking().channels() == led.channels()
    # a new terrain under the live matter: refitted exactly, then stepped
    before = w.read_cells_blocking()
    g2 = inp[ra.GAMMA].copy()
    g2[CFG.cells() // 2: CFG.cells() // 2 + 256 * 8] = 6  # a rock band across the middle
    w.apply_terrain_live(g2)
    after = w.read_cells_blocking()
    assert np.array_equal(after, ra.refit_matter_to_terrain(before, g2))
    assert np.array_equal(w.download(ra.GAMMA), g2)
    assert ra.Ledger.from_cells(after) == ra.Ledger.from_cells(before)
    assert not after.reshape(-1, 4)[g2 >= 3].any()
    w.run(16)
    w.sync()
    assert w.ledger_status() is None
    assert ra.Ledger.from_cells(w.read_cells_blocking()).channels() == ra.Ledger.from_cells(before).channels()
