"""The reference's host-side input preparation, restated in the library
(ribo_rasterize_fumaroles, ribo_vent_heat_src, ribo_seed_cells,
ribo_fit_matter_to_terrain, ribo_refit_matter_to_terrain), against the
numpy restatement in oracle/ribo_seed_oracle.py (bit for bit), against
closed forms, and against the reference's own unit test of the refit
(src/world/mod.rs:2091-2126)."""
import numpy as np
import pytest

import ribossome_amd as ra
import ribo_seed_oracle as so

CFG = ra.WorldConfig(64, 32)


def test_fumarole_raster_matches_the_oracle_and_its_profile():
    fs = [(0.5, 0.1, 40.0, 300.0, 1), (0.02, 0.95, 25.0, 90.0, 1), (0.7, 0.7, 10.0, 10.0, 1),  # r clamps to 1
          (0.3, 0.3, 99.0, 200.0, 0), (0.4, 0.6, 0.0, 200.0, 1),  # disabled, no strength: ignored
          (0.52, 0.12, 15.0, 150.0, 1)]  # overlaps the first
    got = ra.rasterize_fumaroles(CFG, fs)
    want = so.rasterize_fumaroles(64, 30, fs)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    # the profile of one fumarole: strength (1 - d / r)^2 / 5000 inside r cells
    one = ra.rasterize_fumaroles(CFG, [(0.5, 0.5, 50.0, 300.0)]).reshape(64, 64).astype(np.float64)
    y, x = np.mgrid[0:64, 0:64]
    d = np.hypot(x + 0.5 - 32.0, y + 0.5 - 32.0)
    r = 10.0
    prof = np.where(d < r, 50.0 * (1.0 - d / r) ** 2 / 5000.0, 0.0)
    assert np.allclose(one, prof, rtol=1e-5, atol=1e-9)
    assert np.count_nonzero(one) == np.count_nonzero(d < r)


def test_vent_source_is_the_box_sum_times_strength_and_gain():
    rng = np.random.default_rng(4)
    heat = rng.random(CFG.cells()).astype(np.float32)
    chem = rng.random(CFG.cells()).astype(np.float32)
    got = ra.vent_heat_src(CFG, heat, chem, 0.3)
    assert np.array_equal(got.view(np.uint32), so.vent_heat_src(64, 32, heat, chem, 0.3).view(np.uint32))
    # uniform grids: k^2 samples x 5000 gain / k^2 = 5000 gain h
    u = ra.vent_heat_src(CFG, np.full(CFG.cells(), 0.25, np.float32), np.full(CFG.cells(), 0.5, np.float32), 0.3)
    assert np.allclose(u[0::2], 0.25 * 5000 * 0.3, rtol=1e-6) and np.allclose(u[1::2], 0.5 * 5000 * 0.3, rtol=1e-6)
    # box sums of a random grid, in binary64
    b = heat.reshape(32, 2, 32, 2).astype(np.float64).sum(axis=(1, 3)).reshape(-1) * 5000 * 0.3 / 4
    assert np.allclose(got[0::2], b, rtol=1e-6)
    assert not ra.vent_heat_src(CFG, heat, chem, -1.0).any()  # gain clamps at 0


@pytest.mark.parametrize("side,seed,density,active", [(64, 7, 0.83, 1.0), (48, 99, 0.5, 0.3), (32, 1 << 40, 1.0, 0.5)])
def test_seed_cells_matches_the_oracle(side, seed, density, active):
    cfg = ra.WorldConfig(side, side // 2)
    got = ra.seed_cells(cfg, seed, density, active)
    want = so.seed_cells(side, seed, density, active)
    assert np.array_equal(got, want)
    a, s = got & 0xFFFF, got >> 16
    assert (a + s).max() <= 12  # CHEM_CELL_CAP / 4 per channel
    if active == 1.0:
        assert not s.any()
    else:
        assert abs(a.sum() / max((a + s).sum(), 1) - active) < 0.05
    # the density follows the fBm field: its mean count is about
    # density x 10 x E[d] + 1/2 with E[d] near 1/2
    assert abs((a + s).mean() - (density * 10 * 0.5 + 0.5)) < 1.5


def test_fit_matter_to_terrain():
    gamma = np.zeros(64 * 64, np.uint32)
    gamma[100:400] = 1
    gamma[500:700] = 2
    gamma[900:1300] = 6
    cells = ra.seed_cells(CFG, 3)
    got = ra.fit_matter_to_terrain(cells, gamma, 3)
    assert np.array_equal(got, so.fit_matter_to_terrain(cells, gamma, 3))
    c = got.reshape(-1, 4)
    tot = (c & 0xFFFF) + (c >> 16)
    assert np.all(tot[100:400] == 8) and np.all(tot[500:700] == 4) and not tot[900:1300].any()
    assert np.array_equal(c[gamma == 0], cells.reshape(-1, 4)[gamma == 0])
    rub = c[(gamma == 1) | (gamma == 2)]
    assert abs((rub & 0xFFFF).sum() / ((rub & 0xFFFF) + (rub >> 16)).sum() - 0.5) < 0.03


def test_refit_reference_unit_test():
    # src/world/mod.rs:2091-2126: a rock in the middle of a row; its matter
    # is spread over every cell with room, with the same mix of kinds
    n = 400
    cells = np.full(n * 4, 2 | (3 << 16), np.uint32)
    gamma = np.zeros(n, np.uint32)
    gamma[100:140] = 3
    for out in (ra.refit_matter_to_terrain(cells, gamma), so.refit_matter_to_terrain(cells, gamma)):
        assert ra.Ledger.from_cells(out) == ra.Ledger.from_cells(cells)
        c = out.reshape(n, 4)
        per = (c & 0xFFFF) + (c >> 16)
        tot = per.sum(axis=1)
        assert np.all(tot[100:140] == 0)
        ok = np.ones(n, bool)
        ok[100:140] = False
        assert np.all((tot[ok] >= 21) & (tot[ok] <= 24)), tot[ok]
        assert np.all((per[ok] >= 5) & (per[ok] <= 7))
    assert np.array_equal(ra.refit_matter_to_terrain(cells, gamma), so.refit_matter_to_terrain(cells, gamma))


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_refit_random_terrains_exact_and_within_capacity(seed):
    rng = np.random.default_rng(seed)
    n = 2048
    # ~12 a cell on average against ~26 of room: everything fits somewhere
    cells = rng.integers(0, 3, 4 * n).astype(np.uint32) | (rng.integers(0, 3, 4 * n).astype(np.uint32) << 16)
    gamma = rng.choice([0, 0, 0, 1, 2, 3, 6], n).astype(np.uint32)
    got = ra.refit_matter_to_terrain(cells, gamma)
    assert np.array_equal(got, so.refit_matter_to_terrain(cells, gamma))
    assert ra.Ledger.from_cells(got) == ra.Ledger.from_cells(cells)
    cap = np.where(gamma >= 3, 0, 48 * (3 - gamma.astype(np.int64)) // 3)
    c = got.reshape(n, 4).astype(np.int64)
    tot = ((c & 0xFFFF) + (c >> 16)).sum(axis=1)
    assert np.all(tot <= cap) and not tot[gamma >= 3].any()
    # no room anywhere: everything lands in the first non-rock cell, nothing lost
    full = np.full(4 * 16, 12 | (12 << 16), np.uint32)
    g2 = np.full(16, 3, np.uint32)
    g2[5] = 0
    out = ra.refit_matter_to_terrain(full, g2)
    assert np.array_equal(out, so.refit_matter_to_terrain(full, g2))
    assert ra.Ledger.from_cells(out) == ra.Ledger.from_cells(full)


def test_splitmix64_published_vectors():
    """The generator under seed_cells and fit_matter_to_terrain is Vigna's
    splitmix64 (src/world/mod.rs:1842-1855): its published outputs for
    seeds 0 and 1234567."""
    r = so.SplitMix(0)
    assert [r.next() for _ in range(3)] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    r = so.SplitMix(1234567)
    assert [r.next() for _ in range(5)] == [6457827717110365317, 3203168211198807973, 9817491932198370423,
                                            4593380528125082431, 16408922859458223821]


def _lattice_point_value(seed: int, pts) -> float:
    """fbm_field at a pixel whose coordinate is a lattice point of every
    octave: the amplitude-weighted lattice values (amplitudes 0.5 x 0.65^o,
    six octaves), normalised by the amplitudes' sum; each lattice value is a
    splitmix64 draw keyed by octave and lattice coordinates
    (src/world/mod.rs:2046-2078)."""
    M = (1 << 64) - 1
    num, norm, amp = 0.0, 0.0, 0.5
    for o in range(6):
        ix, iy = pts(o)
        r = so.SplitMix(seed ^ ((o * 0x9E3779B97F4A7C15) & M) ^ ((ix * 0xD6E8FEB86659FD93) & M)
                        ^ ((iy * 0xCA5A826395121157) & M))
        num += amp * float(r.f32())
        norm += amp
        amp *= 0.65
    return num / norm


def test_fbm_field_at_lattice_points():
    """At x = 0 every octave samples its lattice point (0, 0); at x = m/4
    octave o samples (2^o, 0), and at y = m/4 (0, 2^o): the field there is
    the weighted mean of those lattice draws, which pins the octave count,
    the persistence, the normalisation and each hash term."""
    m, seed = 32, 0x1234567
    f = so.fbm_field(m, seed)
    assert abs(float(f[0]) - _lattice_point_value(seed, lambda o: (0, 0))) < 1e-6
    assert abs(float(f[m // 4]) - _lattice_point_value(seed, lambda o: (1 << o, 0))) < 1e-6
    assert abs(float(f[(m // 4) * m]) - _lattice_point_value(seed, lambda o: (0, 1 << o))) < 1e-6
    assert abs(float(f[(m // 4) * m + m // 4]) - _lattice_point_value(seed, lambda o: (1 << o, 1 << o))) < 1e-6


def test_seed_cells_density_zero_is_empty():
    """With density 0 the expected count is 0 and the stochastic rounding
    never adds a monomer."""
    assert not so.seed_cells(32, 5, 0.0, 1.0).any()
    assert not ra.seed_cells(ra.WorldConfig(32, 16), 5, 0.0, 1.0).any()
