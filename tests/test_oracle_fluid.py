"""Closed forms for the oracle's fluid kernels (fluid.wgsl), one kernel at a
time on hand-built fields (exact binary32 values: powers of two)."""
import numpy as np

import ribossome_amd as ra
import oracle_ctypes as oc

F = 32
CFG = ra.WorldConfig(64, F)


def world(dt=0.0625, **over):
    w = oc.OracleWorld(CFG, 1)
    p = ra.SimParams.default(1)
    p.dt = dt
    p.fluid_dt = dt
    for k, v in over.items():
        setattr(p, k, v)
    w.params = p
    w.upload(ra.GAMMA, np.zeros(CFG.cells(), np.uint32))
    return w


def vel_field(fx, fy):
    y, x = np.mgrid[0:F, 0:F]
    v = np.zeros((F, F, 2), np.float32)
    v[..., 0] = fx(x, y)
    v[..., 1] = fy(x, y)
    return v.reshape(-1)


def test_advect_uniform_flow_shifts_by_one_cell():
    # v = 16 cells/s, dt = 1/16: the backtrace lands exactly on the
    # neighbour's centre, so the sample is that cell's velocity times the
    # decay pow(0.99, 60 dt) (fluid.wgsl:458-477); cells whose backtrace
    # leaves the box sample zero
    decay = np.float32(0.99 ** (0.0625 * 60))
    w = world()
    w.upload(ra.VELOCITY, vel_field(lambda x, y: 16.0 + 0 * x, lambda x, y: 0 * x))
    w.run_kernel("advect")
    v = w.download(ra.VELOCITY).reshape(F, F, 2)
    assert np.all(v[:, 0, :] == 0)
    assert np.all(v[:, 1:, 0] == np.float32(16.0) * decay) and np.all(v[..., 1] == 0)
    w = world()
    w.upload(ra.VELOCITY, vel_field(lambda x, y: 0 * x, lambda x, y: -16.0 + 0 * x))
    w.run_kernel("advect")
    v = w.download(ra.VELOCITY).reshape(F, F, 2)
    assert np.all(v[-1, :, :] == 0)
    assert np.all(v[:-1, :, 1] == np.float32(-16.0) * decay) and np.all(v[..., 0] == 0)


def test_divergence_of_linear_field_is_constant_with_wall_mirroring():
    # v = (a (x + 1/2), b (y + 1/2)): central differences give a + b inside,
    # and the mirrored ghost cells at the walls give the same (fluid.wgsl:525-551)
    a, b = 0.25, -0.125
    w = world()
    w.upload(ra.VELOCITY, vel_field(lambda x, y: a * (x + 0.5), lambda x, y: b * (y + 0.5)))
    w.run_kernel("divergence")
    d = w.download(ra.DIVERGENCE).reshape(F, F)
    assert np.all(d[:-1, :-1] == np.float32(a + b))  # interior and the x = 0 / y = 0 walls
    # at the far walls the ghost mirrors v_c: 0.5 (-a (F - 1/2) - a (F - 3/2)) = -a (F - 1)
    assert np.all(d[:-1, -1] == np.float32(-a * (F - 1) + b))
    assert np.all(d[-1, :-1] == np.float32(a - b * (F - 1)))


def test_gradient_of_linear_pressure():
    # p = c x: grad = c inside, c / 2 at the side walls where the missing
    # neighbour is replaced by p_c (fluid.wgsl:610-630)
    c = 0.5
    w = world()
    y, x = np.mgrid[0:F, 0:F]
    w.upload(ra.PRESSURE, (c * x).astype(np.float32).reshape(-1))
    w.upload(ra.VELOCITY, np.zeros(2 * F * F, np.float32))
    w.run_kernel("subtract_gradient")
    v = w.download(ra.VELOCITY).reshape(F, F, 2)
    assert np.all(v[:, 1:-1, 0] == -c)
    assert np.all(v[:, 0, 0] == -c / 2) and np.all(v[:, -1, 0] == -c / 2)
    assert np.all(v[..., 1] == 0)


def test_net_flux_then_boundaries_zero_the_row_and_column_means():
    rng = np.random.default_rng(8)
    w = world()
    w.upload(ra.VELOCITY, rng.standard_normal(2 * F * F).astype(np.float32))
    w.run_kernel("net_flux")
    nf = w.download(ra.NET_FLUX)
    v0 = w.download(ra.VELOCITY).reshape(F, F, 2).astype(np.float64)
    assert np.allclose(nf[:F], v0[..., 1].mean(axis=1), atol=1e-6)  # rows: mean v_y
    assert np.allclose(nf[F:], v0[..., 0].mean(axis=0), atol=1e-6)  # columns: mean v_x
    w.run_kernel("boundaries")
    v = w.download(ra.VELOCITY).reshape(F, F, 2).astype(np.float64)
    assert np.all(v[:, 0, 0] == 0) and np.all(v[:, -1, 0] == 0)  # no normal flow at the walls
    assert np.all(v[0, :, 1] == 0) and np.all(v[-1, :, 1] == 0)
    assert np.abs(v[1:-1, :, 1].mean(axis=1)).max() < 1e-6  # tests/conservation.rs:97-104
    assert np.abs(v[:, 1:-1, 0].mean(axis=0)).max() < 1e-6


def test_forces_blur_weights_sum_to_40_over_36():
    # uniform temperature 1 above ambient: buoyancy 12 up on interior cells;
    # the 9-point blur weights (4 axis x 4 + 4 diagonal + 20 centre) / 36 sum
    # to 40/36 and the smoothed force mixes 1/4 of it (fluid.wgsl:410-432)
    dt = 0.0625
    w = world(dt=dt, slope_steer_rate=0.0)
    w.upload(ra.TEMP, np.ones(F * F, np.float32))
    w.upload(ra.VELOCITY, np.zeros(2 * F * F, np.float32))
    w.run_kernel("forces")
    v = w.download(ra.VELOCITY).reshape(F, F, 2)
    f_user = np.float32(12.0) * np.float32(0.75) + np.float32(12.0 * 40.0 / 36.0) * np.float32(0.25)
    inner = v[2:-2, 2:-2, 1].astype(np.float64)
    assert np.allclose(inner, float(f_user) * dt, rtol=1e-6)
    # border row: no buoyancy of its own (f_c = 0); the clamped blur sees the
    # row above: (4 * 12 + 2 * 12) / 36 = 2, mixed at 1/4 -> 0.5
    assert np.allclose(v[0, 2:-2, 1].astype(np.float64), 0.5 * dt, rtol=1e-6)


def test_temperature_cools_towards_ambient():
    # uniform T, no flow, no vent, no rock: T' = T exp(-0.12 dt) (ambient 0)
    dt = 0.0625
    w = world(dt=dt, vent_cycle_amp=0.0, sun_heat=0.0)
    w.upload(ra.TEMP, np.full(F * F, 4.0, np.float32))
    w.upload(ra.VELOCITY, np.zeros(2 * F * F, np.float32))
    w.run_kernel("temperature")
    t = w.download(ra.TEMP).astype(np.float64)
    assert np.allclose(t, 4.0 * np.exp(-0.12 * dt), rtol=3e-7)


def test_solid_mask_needs_two_rock_samples_of_four():
    # build_solid_mask (fluid.wgsl:185-189): a fluid cell is a wall when at
    # least 2 of its 4 grid sub-samples are rock
    w = world()
    g = np.zeros((64, 64), np.uint32)
    g[0, 0] = 6              # 1 of 4 -> fluid
    g[0, 2:4] = 6            # 2 of 4 -> wall
    g[2:4, 2:4] = 2          # rubble only -> fluid
    g[10:12, 10:12] = 6      # 4 of 4 -> wall
    w.upload(ra.GAMMA, g.reshape(-1))
    m = w.download(ra.SOLID_MASK).reshape(F, F)
    assert m[0, 0] == 0 and m[0, 1] == 1 and m[1, 1] == 0 and m[5, 5] == 1
    assert m.sum() == 2
