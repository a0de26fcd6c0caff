"""The oracle's light step (light.wgsl:44-183, scheduled by
src/world/mod.rs:1506-1525) against closed forms and its own fixed point.

- An empty world (no grains, no monomers) transmits everything, so the
  sweep leaves sun * w^(LS - y) in light row y, w = exp(-uv_depth * depth
  factor * LIGHT_DIV / grid) per band.
- The top row is sun * w * T, with T = exp(-0.6 * grains / LIGHT_DIV -
  monomer_uv_absorb * density * LIGHT_DIV / grid * depth factor) over the
  light cell's LIGHT_DIV^2 grid cells: a wrong sum, divisor or band factor
  shows up at once.
- Under a slanted sun a beam through a gap drifts by sun_slope cells per
  light row (light at x comes from x - slope one row up).
- Propagation from a dark field reaches the sweep's field exactly after
  LS passes (the sweep is its fixed point), and light_mode 0 keeps an
  uploaded field frozen.
"""
import math

import numpy as np

import ribossome_amd as ra
import oracle_ctypes as oc
from worlds import parity_params, parity_settings


def light_world(cfg, gamma=None, chem=None, light=None, epoch=0, settings=None, **params):
    w = oc.OracleWorld(cfg, 3)
    p = parity_params(3, **{"epoch": epoch, "uv_strength": 3.0, "day_period": 0.0, **params})
    w.params = p
    s = parity_settings(light_mode=1)
    for k, v in (settings or {}).items():
        setattr(s, k, v)
    w.set_settings(s)
    G = cfg.grid_size
    w.upload(ra.GAMMA, np.zeros(G * G, np.uint32) if gamma is None else gamma)
    w.upload(ra.CHEM, np.zeros(4 * G * G, np.uint32) if chem is None else chem)
    if light is not None:
        w.upload(ra.LIGHT, light)
    return w


def bands(w, cfg):
    LS = cfg.grid_size // 2
    return w.download(ra.LIGHT).reshape(2, LS, LS)


def test_empty_world_sweep_is_water_attenuation():
    cfg = ra.WorldConfig(64, 32)
    LS, depth = 32, 20.0
    w = light_world(cfg, uv_depth=depth)
    w.encode_steps(1)
    L = bands(w, cfg).astype(np.float64)
    for band, f in ((0, 1.0), (1, 0.03)):
        rw = math.exp(-depth * f * 2.0 / 64.0)
        want = rw ** (LS - np.arange(LS, dtype=np.float64))  # row y, top row LS - 1
        got = L[band]
        assert np.allclose(got, want[:, None], rtol=2e-5, atol=0), band
        assert np.all(got[:, 1:] == got[:, :1])  # no sideways structure under a vertical sun


def test_top_row_transmission_closed_form():
    cfg = ra.WorldConfig(32, 16)
    G, LS = 32, 16
    rng = np.random.default_rng(5)
    gamma = rng.integers(0, 3, G * G).astype(np.uint32)
    chem = np.zeros((G * G, 4), np.uint32)
    act = rng.integers(0, 9, (G * G, 4)).astype(np.uint32)
    spent = rng.integers(0, 9, (G * G, 4)).astype(np.uint32)
    chem[:] = act | (spent << 16)
    mua = 0.8
    w = light_world(cfg, gamma, chem.reshape(-1), uv_depth=0.0, monomer_uv_absorb=mua)
    w.encode_steps(1)
    L = bands(w, cfg).astype(np.float64)
    g2 = gamma.reshape(G, G).astype(np.float64)
    m2 = (act + spent).sum(1).reshape(G, G).astype(np.float64)
    y0 = 2 * (LS - 1)
    for lx in range(LS):
        gs = g2[y0:y0 + 2, 2 * lx:2 * lx + 2].sum()
        dens = m2[y0:y0 + 2, 2 * lx:2 * lx + 2].sum() / 4.0
        for band, depth in ((0, 1.0), (1, 0.35)):
            t = math.exp(-0.6 * gs / 2.0 - mua * dens * 2.0 / G * depth)
            # sun 1, no water: the top row is the transmission itself
            assert abs(L[band, LS - 1, lx] - t) <= 4e-7 * t, (band, lx)


def sun_slope(epoch, day_period, sun_angle, day_fraction):
    """SimParams::sun_slope_at (src/params.rs:507-518), binary64 then f32."""
    ph = epoch / day_period
    ph -= math.floor(ph)
    frac = min(max(day_fraction, 0.05), 1.0)
    phase = ph / frac if ph < frac else 1.0
    angle = (phase - 0.5) * 2.0 * math.radians(min(max(sun_angle, 0.0), 85.0))
    return float(np.float32(math.tan(angle)))


def test_slanted_beam_drifts_by_sun_slope():
    cfg = ra.WorldConfig(128, 64)
    G, LS = 128, 64
    gamma = np.zeros((G, G), np.uint32)
    gamma[G - 8:, :] = 6  # an opaque roof four light rows thick ...
    x_gap = 16
    gamma[G - 8:, 2 * x_gap:2 * x_gap + 2] = 0  # ... with a one-light-cell gap
    epoch, period = 75, 100.0
    slope = sun_slope(epoch, period, 45.0, 1.0)
    assert 0.4 < slope < 0.42
    w = light_world(cfg, gamma.reshape(-1), epoch=epoch, day_period=period, sun_angle=45.0, day_fraction=1.0)
    w.encode_steps(1)
    L = bands(w, cfg)[0].astype(np.float64)
    y_gap = LS - 4  # lowest roof row
    for k in range(4, 25, 4):
        row = L[y_gap - k]
        c = (row * np.arange(LS)).sum() / row.sum()
        assert abs(c - (x_gap + slope * k)) < 0.5, (k, c)


def test_propagation_reaches_the_sweep_exactly():
    cfg = ra.WorldConfig(128, 64)
    G, LS = 128, 64
    rng = np.random.default_rng(11)
    gamma = np.where(rng.random(G * G) < 0.2, 6, rng.integers(0, 3, G * G)).astype(np.uint32)
    params = dict(uv_depth=5.0)
    swept = light_world(cfg, gamma, settings=dict(light_rows_per_step=0, light_interval=1), **params)
    swept.encode_steps(1)
    want = swept.download(ra.LIGHT)
    dark = np.zeros(2 * LS * LS, np.float32)
    prop = light_world(cfg, gamma, light=dark, settings=dict(light_rows_per_step=16), **params)
    prop.encode_steps(LS // 16 - 1)
    assert not np.array_equal(prop.download(ra.LIGHT), want)  # the front has not reached the bottom yet
    prop.encode_steps(1)
    assert np.array_equal(prop.download(ra.LIGHT).view(np.uint32), want.view(np.uint32))
    prop.encode_steps(2)  # the sweep's field is a fixed point
    assert np.array_equal(prop.download(ra.LIGHT).view(np.uint32), want.view(np.uint32))
    # invalidate_light: the next step sweeps again (same field: nothing changed)
    prop.invalidate_light()
    prop.encode_steps(1)
    assert np.array_equal(prop.download(ra.LIGHT).view(np.uint32), want.view(np.uint32))
    frozen = light_world(cfg, gamma, light=dark, settings=dict(light_mode=0), **params)
    frozen.encode_steps(3)
    assert np.all(frozen.download(ra.LIGHT) == 0.0)
