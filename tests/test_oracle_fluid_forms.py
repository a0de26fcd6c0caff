"""Closed forms and statistics for the oracle's fluid kernels that the
basic cases in test_oracle_fluid.py leave open: viscosity, vorticity
confinement, rubble drag, slope steering, the wall reflection, advection
against a wall, the temperature and reductant update (advection, diffusion,
rock conduction, vents, clamp) and the buoyancy's sideways wobble
(fluid.wgsl:46-54,150-224,226-255,300-363,376-389,435-518).

Each expectation is written from the WGSL formula, not from the oracle's
code, and is chosen so that a plausible slip (a wrong weight, sign, constant
or operand) moves the result far beyond the tolerance."""
import numpy as np

import ribossome_amd as ra
import oracle_ctypes as oc

F = 32
CFG = ra.WorldConfig(64, F)
DT = 0.0625


def world(gamma=None, **over):
    w = oc.OracleWorld(CFG, 1)
    p = ra.SimParams.default(1)
    p.dt = p.fluid_dt = DT
    p.sun_heat = 0.0
    p.vent_cycle_amp = 0.0
    for k, v in over.items():
        setattr(p, k, v)
    w.params = p
    w.upload(ra.GAMMA, np.zeros(CFG.cells(), np.uint32) if gamma is None else gamma.reshape(-1))
    return w


def field(vx, vy):
    v = np.zeros((F, F, 2), np.float32)
    v[..., 0], v[..., 1] = vx, vy
    return v.reshape(-1)


def vel(w):
    return w.download(ra.VELOCITY).reshape(F, F, 2).astype(np.float64)


YX = np.mgrid[0:F, 0:F]


# ---- diffuse_velocity (fluid.wgsl:435-456) ----------------------------------
def test_viscosity_keeps_uniform_fields_and_spreads_an_impulse_with_9_point_weights():
    w = world()
    w.upload(ra.VELOCITY, field(3.25, -1.5))
    w.run_kernel("diffuse")
    v = vel(w)
    assert np.all(v[..., 0] == 3.25) and np.all(v[..., 1] == -1.5)  # (16 + 4 - 20) u / 6 = 0 exactly
    # an impulse V at one interior cell: lap = (4 axis + diag - 20 c) / 6
    V, x0, y0 = 6.0, 10, 12
    vx = np.zeros((F, F))
    vx[y0, x0] = V
    w = world()
    w.upload(ra.VELOCITY, field(vx, 0.0))
    w.run_kernel("diffuse")
    v = vel(w)[..., 0]
    a = 3.7 * DT  # max(viscosity, 0) * fluid_dt, viscosity default 3.7
    assert np.isclose(v[y0, x0], V + a * (-20.0 * V / 6.0), rtol=1e-6)
    for dy, dx in ((0, 1), (0, -1), (1, 0), (-1, 0)):
        assert np.isclose(v[y0 + dy, x0 + dx], a * 4.0 * V / 6.0, rtol=1e-6)
    for dy, dx in ((1, 1), (1, -1), (-1, 1), (-1, -1)):
        assert np.isclose(v[y0 + dy, x0 + dx], a * V / 6.0, rtol=1e-6)
    m = np.ones((F, F), bool)
    m[y0 - 1:y0 + 2, x0 - 1:x0 + 2] = False
    assert np.all(v[m] == 0.0)
    # border cells are copied, not diffused
    vx = np.zeros((F, F))
    vx[0, 7] = V
    w = world()
    w.upload(ra.VELOCITY, field(vx, 0.0))
    w.run_kernel("diffuse")
    assert vel(w)[0, 7, 0] == V


# ---- vorticity_confinement (fluid.wgsl:479-518) -----------------------------
def test_vorticity_leaves_irrotational_flow_and_pushes_across_the_curl_gradient():
    y, x = YX
    # potential flow (x, -y) / 8: every curl is 0 -> unchanged
    w = world()
    w.upload(ra.VELOCITY, field(x / 8.0, -y / 8.0))
    w.run_kernel("vorticity")
    v = vel(w)
    assert np.array_equal(v[..., 0], (x / 8.0).astype(np.float32)) and np.array_equal(v[..., 1], (-y / 8.0).astype(np.float32))
    # v = (0, c x^2): curl = 0.5 ((x+1)^2 - (x-1)^2) c = 2 c x, grad |curl| = (2c, 0),
    # n = (1, 0), f = (n.y, -n.x) w eps = (0, -2 c x eps): v_y -= 2 c x eps dt
    eps = 7.7  # clamp(fluid_vorticity, 0, 10), default 7.7
    for c, clamp in ((1.0 / 64.0, False), (1.0, True)):
        w = world()
        w.upload(ra.VELOCITY, field(0.0, c * x * x))
        w.run_kernel("vorticity")
        v = vel(w)
        core = (slice(2, F - 2), slice(2, F - 2))
        xs = x[core].astype(np.float64)
        dv = -2.0 * c * xs * eps * DT
        if clamp:  # |dv| <= 2 per step
            dv = np.maximum(dv, -2.0)
        assert np.allclose(v[core][..., 1], c * xs * xs + dv, rtol=1e-5, atol=1e-6)
        assert np.all(v[core][..., 0] == 0.0)
        # the 2-cell border is copied
        assert np.array_equal(v[:, :2, 1], (c * x[:, :2] ** 2).astype(np.float32))


# ---- add_forces: rubble_drag and slope_steer_velocity (fluid.wgsl:150-178) ----
def test_rubble_drag_and_slope_steering():
    # rubble of 1 grain from grid column 32 on: fluid column 16 sits on
    # rubble (4 samples of 1 grain: g = 4 / (4 * 3) = 1/3, drag 1 / (1 + 300
    # g dt)) with zero slope; fluid column 15's centre (grid column 31) sees
    # the step: slope = ((0.25 - 0) 0.5 + (0.25 + 0.25) / (4 sqrt 2), 0) / 30
    # (terrain.wgsl:116-133), downslope -x, and no rubble of its own
    g = np.zeros((64, 64), np.uint32)
    g[:, 32:] = 1
    rate = 8.0
    w = world(gamma=g, slope_steer_rate=rate)
    V = 10.0
    w.upload(ra.VELOCITY, field(0.0, V))
    w.run_kernel("forces")  # temperature 0: no buoyancy, f_user = 0
    v = vel(w)
    # drag only
    assert np.allclose(v[:, 16, 1], V / (1.0 + 300.0 * (1.0 / 3.0) * DT), rtol=1e-6) and np.all(v[:, 16, 0] == 0)
    assert np.all(v[:, 17:, 1] == np.float32(V / (1.0 + 100.0 * DT))) or np.allclose(v[:, 17:, 1], V / (1.0 + 100.0 * DT), rtol=1e-6)
    # untouched where there is neither rubble nor slope
    assert np.all(v[:, :15, 1] == V) and np.all(v[:, :15, 0] == 0)
    # steering: v_dir = (0, 1), s_dir = (-1, 0), misalign = 1 - dot = 1,
    # t = 1 - exp(-rate * 1 * dt), dir = normalize(v_dir + (s_dir - v_dir) t), |v| kept
    t = 1.0 - np.exp(-rate * DT)
    d = np.array([-t, 1.0 - t])
    want = d / np.linalg.norm(d) * V
    assert np.allclose(v[:, 15, 0], want[0], rtol=1e-5) and np.allclose(v[:, 15, 1], want[1], rtol=1e-5)
    # velocity along the slope (misalign 0) is not turned
    w = world(gamma=g, slope_steer_rate=rate)
    w.upload(ra.VELOCITY, field(-V, 0.0))
    w.run_kernel("forces")
    assert np.all(vel(w)[:, 15, 0] == -V) and np.all(vel(w)[:, 15, 1] == 0)


# ---- reflect_if_into_solid (fluid.wgsl:197-224) ------------------------------
def wall_world(**over):
    g = np.zeros((64, 64), np.uint32)
    g[:, :8] = 6  # rock: fluid columns 0..3 are walls
    return world(gamma=g, **over)


def test_reflection_removes_only_the_into_wall_component():
    # subtract_gradient with p = 0 leaves v = reflect(v): at fluid column 4
    # the outward normal is (1, 0) (left 1, diagonals (1 + 1) - 0 weighted
    # 0.7071, vertical parts cancel), so (-1.5, 0.5) -> (0, 0.5); moving
    # away, (1.5, 0.5) is kept; columns further out are untouched
    for vx, want in ((-1.5, 0.0), (1.5, 1.5)):
        w = wall_world()
        w.upload(ra.PRESSURE, np.zeros(F * F, np.float32))
        w.upload(ra.VELOCITY, field(vx, 0.5))
        w.run_kernel("subtract_gradient")
        v = vel(w)
        assert np.all(v[1:-1, 4, 0] == want) and np.all(v[1:-1, 4, 1] == 0.5)
        assert np.all(v[:, 5:, 0] == vx) and np.all(v[:, 5:, 1] == 0.5)
        assert np.all(v[:, :4] == 0.0)
    # the corner cell (4, 0): s_l = 1, s_tl = 1 (s_b, s_bl out of the box):
    # n ~ (1 + 0.7071, -0.7071), v - n (v . n) when v . n < 0
    n = np.array([1.0 + 0.70710678, -0.70710678])
    n /= np.linalg.norm(n)
    v0 = np.array([-1.5, 0.5])
    want = v0 - n * np.dot(v0, n)
    w = wall_world()
    w.upload(ra.PRESSURE, np.zeros(F * F, np.float32))
    w.upload(ra.VELOCITY, field(-1.5, 0.5))
    w.run_kernel("subtract_gradient")
    assert np.allclose(vel(w)[0, 4], want, rtol=1e-5, atol=1e-6)


# ---- advect_velocity against a wall (fluid.wgsl:226-255,458-477) -------------
def test_advection_from_a_wall_samples_zero():
    # uniform (16, 0), dt 1/16: column 4's backtrace lands on the centre of
    # wall column 3 (is_solid_at_pos; the re-trace with the reflected
    # velocity is the same, v . n > 0), whose raw velocity is 0: sample 0.
    # Column 5 samples column 4: 16 x the decay.
    w = wall_world()
    w.upload(ra.VELOCITY, field(16.0, 0.0))
    w.run_kernel("advect")
    v = vel(w)
    decay = np.float32(0.99 ** (DT * 60))
    assert np.all(v[:, 4] == 0.0)
    assert np.all(v[:, 5:, 0] == np.float32(16.0) * decay) and np.all(v[:, 5:, 1] == 0)


# ---- update_temperature (fluid.wgsl:46-54,300-363) ---------------------------
def test_temperature_advects_diffuses_with_rock_conduction_and_cools():
    # T = x / 16 + 1, flow (16, 0), dt 1/16: the bilinear backtrace lands on
    # x - 1 (fract 0), the 4-neighbour mean of a linear field is T(x), so
    # t = T(x-1) (1 - k) + T(x) k with k = 0.08 mix(1, 4, fill), then
    # t exp(-0.12 (1 - fill) dt) (ambient 0)
    y, x = YX
    T = (x / 16.0 + 1.0).astype(np.float32)
    for grains, fill in ((0, 0.0), (2, 2.0 / 3.0)):
        g = np.full((64, 64), grains, np.uint32)
        w = world(gamma=g)
        w.upload(ra.VELOCITY, field(16.0, 0.0))
        w.upload(ra.TEMP, T.reshape(-1))
        w.run_kernel("temperature")
        t = w.download(ra.TEMP).reshape(F, F).astype(np.float64)
        k = 0.08 * (1.0 + 3.0 * fill)
        Td = T.astype(np.float64)
        want = (Td[:, :-2] * (1 - k) + Td[:, 1:-1] * k) * np.exp(-0.12 * (1.0 - fill) * DT)
        assert np.allclose(t[:, 1:-1], want, rtol=2e-6), grains
    # the clamp at TEMP_MAX = 12
    w = world()
    w.upload(ra.TEMP, np.full(F * F, 20.0, np.float32))
    w.run_kernel("temperature")
    assert np.all(w.download(ra.TEMP) == 12.0)


def test_vents_heat_and_release_reductant_with_the_vent_cycle():
    # T = r = 0, no flow: t = H m(x) 0.01 dt, cooled by exp(-0.12 dt);
    # r = R m(x) 0.01 dt times exp(-redox_decay dt), with
    # m(x) = max(1 + amp sin(2 pi (t_c + 1.5 x / F)), 0), t_c = (epoch % period) / period
    H, Rv, amp, period, epoch = 40.0, 25.0, 0.91, 100.0, 25
    w = world(vent_cycle_amp=amp, vent_cycle_period=period, epoch=epoch, redox_decay=0.02)
    heat = np.zeros((F, F, 2), np.float32)
    heat[..., 0], heat[..., 1] = H, Rv
    w.upload(ra.HEAT_SRC, heat.reshape(-1))
    w.run_kernel("temperature")
    t = w.download(ra.TEMP).reshape(F, F).astype(np.float64)
    r = w.download(ra.REDOX).reshape(F, F).astype(np.float64)
    xs = np.arange(F)
    m = np.maximum(1.0 + amp * np.sin(6.2831853 * ((epoch % period) / period + 1.5 * xs / F)), 0.0)
    assert m.min() < 0.2 and m.max() > 1.8  # the cycle really moves along x
    assert np.allclose(t, H * m * 0.01 * DT * np.exp(-0.12 * DT), rtol=1e-5, atol=1e-9)
    assert np.allclose(r, Rv * m * 0.01 * DT * np.exp(-0.02 * DT), rtol=1e-5, atol=1e-9)


# ---- buoyancy's sideways wobble (fluid.wgsl:376-389) -------------------------
def wobble(T, epoch):
    w = world(epoch=epoch)
    w.upload(ra.TEMP, np.full(F * F, T, np.float32))
    w.upload(ra.VELOCITY, np.zeros(2 * F * F, np.float32))
    w.run_kernel("forces")
    return vel(w)[..., 0] / DT


def test_buoyancy_side_wobble_statistics_and_time_bucket():
    # dev = 1: f.x = side * 12 |dev|, side = (u 2 - 1) 0.25 |dev| / max(|dev|, 1),
    # u uniform: f.x ~ U(-3, 3), variance 3.  After the 9-point blur mixed in
    # by a quarter, f_user.x = (3/4 + 20/144) f_c + sum(axis) / 36 + sum(diag) / 144.
    fx = wobble(1.0, 0)
    core = fx[2:-2, 2:-2].reshape(-1)
    var = 3.0 * ((0.75 + 20.0 / 144.0) ** 2 + 4.0 / 36.0 ** 2 + 4.0 / 144.0 ** 2)
    assert abs(core.std() / np.sqrt(var) - 1.0) < 0.12
    assert abs(core.mean()) < 5.0 * np.sqrt(var / core.size)
    # the same noise for every epoch of a 16-epoch bucket, new noise after
    assert np.array_equal(fx, wobble(1.0, 15))
    assert not np.allclose(fx, wobble(1.0, 16))
    # dev = 0.5: side scales by |dev| / max(|dev|, 1) = 0.5 and f.x by |dev|:
    # a quarter of the dev = 1 wobble, cell for cell, up to the force words'
    # 1/256 quantisation (fluid.wgsl:281-284: at most 1/512 per word, and the
    # blur's weights sum to 3/4 + 1/4 40/36 < 1.03)
    q = 1.03 / 512.0
    assert np.allclose(wobble(0.5, 0)[2:-2, 2:-2], fx[2:-2, 2:-2] / 4.0, rtol=0, atol=q + q / 4.0)
    assert np.abs(wobble(0.5, 0)[2:-2, 2:-2] - fx[2:-2, 2:-2] / 2.0).max() > 0.1  # not a half


# ---- jacobi_pressure (fluid.wgsl:553-608) ------------------------------------
def test_jacobi_sweep_keeps_harmonic_fields_and_mirrors_at_walls():
    # p_new = (p_l + p_r + p_b + p_t - div) / 4 with walls and the box edge
    # replaced by p_c (Neumann); walls get 0.  A linear p with div = 0 is a
    # fixed point inside; at x = 0 the missing left neighbour is p_c:
    # (p_c + p(1) + 2 p_c) / 4 for p = x / 4 + 1 (uniform in y)
    y, x = YX
    pr = (x / 4.0 + 1.0).astype(np.float32)
    w = wall_world()
    w.upload(ra.VELOCITY, np.zeros(2 * F * F, np.float32))
    w.run_kernel("divergence")  # zero velocity: div = 0
    w.upload(ra.PRESSURE, pr.reshape(-1))
    w.run_kernel("jacobi")
    p = w.download(ra.PRESSURE).reshape(F, F).astype(np.float64)
    assert np.all(p[:, :4] == 0.0)  # walls
    assert np.allclose(p[1:-1, 5:-1], pr[1:-1, 5:-1], rtol=1e-7)
    # fluid column 4: the wall on its left mirrors p_c
    assert np.allclose(p[1:-1, 4], (3.0 * pr[1:-1, 4] + pr[1:-1, 5]) / 4.0, rtol=1e-7)
    # top and bottom rows: the missing neighbour is p_c too
    assert np.allclose(p[0, 5:-1], pr[0, 5:-1], rtol=1e-7)
    # the divergence enters as -div / 4
    w = world()
    w.upload(ra.VELOCITY, field(YX[1] * 0.5, 0.0))  # div = 0.5 inside (central differences)
    w.run_kernel("divergence")
    w.upload(ra.PRESSURE, np.zeros(F * F, np.float32))
    w.run_kernel("jacobi")
    p = w.download(ra.PRESSURE).reshape(F, F)
    assert np.allclose(p[1:-1, 1:-2], -0.5 / 4.0, rtol=1e-7)
