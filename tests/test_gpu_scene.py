"""Saved scenes on the GPU (tests/scene.rs:20-81 for the world step): a world
saved and loaded into a new world has the same grids and parameters, and
both continue bit for bit while conserving matter."""
import time

import numpy as np
import pytest

import ribossome_amd as ra
from worlds import bit_diff, load, make_inputs, parity_params, parity_settings

pytestmark = pytest.mark.gpu


def test_scene_round_trip_continues_identically(tmp_path):
    print("case scene round trip", flush=True)
    cfg = ra.WorldConfig(128, 64)
    inp = make_inputs(cfg, 3, rock=0.12, rubble=True, vent_rows=3)
    a = ra.World(cfg, 3)
    load(a, inp, parity_params(3, diffusion=20.0, thermal_activation=1.0), parity_settings(ledger_check=1))
    a.run(40)
    seeded = ra.Ledger.from_cells(inp[ra.CHEM])
    path = str(tmp_path / "cena.ribo")
    a.save_scene(path)
    b = ra.World(cfg, 99)
    b.settings = parity_settings(ledger_check=1)
    b.load_scene(path)
    assert b.params.epoch == a.params.epoch == 40 and b.params.seed == 3
    assert b.params.diffusion == a.params.diffusion
    for which in (ra.CHEM, ra.GAMMA, ra.VELOCITY, ra.PRESSURE, ra.TEMP, ra.REDOX, ra.SOLID_MASK, ra.SLOPE):
        assert bit_diff(a.download(which), b.download(which)) == 0, ra.BUFFER_NAMES[which]
    a.run(50)
    b.run(50)
    assert np.array_equal(a.read_cells_blocking(), b.read_cells_blocking())
    for which in (ra.VELOCITY, ra.PRESSURE, ra.TEMP):
        assert bit_diff(a.download(which), b.download(which)) == 0, ra.BUFFER_NAMES[which]
    assert b.ledger_blocking().channels() == seeded.channels()


def test_scene_of_the_headline_world(tmp_path):
    """Save and load the 2048^2 world (times printed; host side: LZ4 and the
    file, plus the downloads and uploads)."""
    print("case scene 2048", flush=True)
    cfg = ra.WorldConfig(2048, 1024)
    inp = make_inputs(cfg, 2, rock=0.15, rubble=True, vent_rows=16)
    a = ra.World(cfg, 2)
    load(a, inp, parity_params(2), parity_settings())
    a.run(4)
    path = str(tmp_path / "big.ribo")
    t0 = time.perf_counter()
    a.save_scene(path)
    t1 = time.perf_counter()
    b = ra.World(cfg, 5)
    b.settings = parity_settings()
    t2 = time.perf_counter()
    b.load_scene(path)
    t3 = time.perf_counter()
    import os
    size = os.path.getsize(path)
    print(f"2048^2 scene: {size / 1e6:.1f} MB on disk, save {t1 - t0:.2f} s, load {t3 - t2:.2f} s", flush=True)
    assert np.array_equal(a.read_cells_blocking(), b.read_cells_blocking())
    assert bit_diff(a.download(ra.VELOCITY), b.download(ra.VELOCITY)) == 0
