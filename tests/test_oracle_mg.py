"""The oracle's multigrid (multigrid.wgsl, schedule src/world/mod.rs:1545-1577).

Pinned by: a scipy sparse direct solve of the same eps-anchored 5-point
operator (manufactured solutions), closed forms on isolated pockets,
linearity, and the known convergence behaviour of one V-cycle.
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle_ctypes as oc


def operator(n, fluid, eps):
    """A p = sum over fluid 4-neighbours (p_n - p_c) - eps p_c, on fluid cells."""
    f = fluid.astype(bool)
    idx = -np.ones(n * n, dtype=np.int64)
    idx[f] = np.arange(int(f.sum()))
    rows, cols, vals = [], [], []
    for y in range(n):
        for x in range(n):
            i = y * n + x
            if not f[i]:
                continue
            k = 0
            for dx, dy in ((-1, 0), (1, 0), (0, -1), (0, 1)):
                X, Y = x + dx, y + dy
                if 0 <= X < n and 0 <= Y < n and f[Y * n + X]:
                    rows.append(idx[i]); cols.append(idx[Y * n + X]); vals.append(1.0); k += 1
            rows.append(idx[i]); cols.append(idx[i]); vals.append(-(k + eps))
    m = int(f.sum())
    return sp.csr_matrix((vals, (rows, cols)), shape=(m, m)), f


def test_vcycles_converge_to_direct_solution_all_fluid():
    n = 32
    fluid = np.ones(n * n, dtype=np.uint32)
    A, f = operator(n, fluid, 1e-4)
    rng = np.random.default_rng(3)
    p_star = rng.standard_normal(n * n)
    rhs = np.zeros(n * n)
    rhs[f] = A @ p_star[f]
    p = oc.mg_solve(n, 8, fluid, rhs, np.zeros(n * n), 50, fp64=True)
    # the residual converges; the solution is recovered up to its constant
    # mode, whose eigenvalue is -eps and which the V-cycle reduces slowly
    # (it does not change the gradient that projects the velocity)
    assert np.linalg.norm(rhs[f] - A @ p[f]) / np.linalg.norm(rhs[f]) < 1e-6
    dp, ds = p - p.mean(), p_star - p_star.mean()
    assert np.linalg.norm(dp - ds) / np.linalg.norm(ds) < 1e-5
    assert abs((p - p_star).mean()) < 0.02
    direct = spla.spsolve(A.tocsc(), rhs[f])
    assert np.allclose(direct, p_star, atol=1e-8)


def test_one_vcycle_reduces_residual_with_obstacles():
    n = 64
    rng = np.random.default_rng(4)
    fluid = (rng.random(n * n) > 0.1).astype(np.uint32)
    A, f = operator(n, fluid, 1e-4)
    rhs = rng.standard_normal(n * n) * fluid
    r0 = np.linalg.norm(rhs[f])
    p1 = oc.mg_solve(n, 8, fluid, rhs, np.zeros(n * n), 1, fp64=True)
    r1 = np.linalg.norm(rhs[f] - A @ p1[f])
    p4 = oc.mg_solve(n, 8, fluid, rhs, np.zeros(n * n), 4, fp64=True)
    r4 = np.linalg.norm(rhs[f] - A @ p4[f])
    assert r1 < 0.5 * r0 and r4 < r1


def test_isolated_cell_is_solved_exactly():
    # a fluid cell with no fluid neighbour: k = 0, so one sweep gives
    # p = (0 - rhs) / (0 + eps) exactly (multigrid.wgsl:66)
    n = 16
    fluid = np.zeros(n * n, dtype=np.uint32)
    fluid[5 * n + 5] = 1
    rhs = np.zeros(n * n, dtype=np.float32)
    rhs[5 * n + 5] = 3.0e-4
    p = oc.mg_solve(n, 8, fluid, rhs, np.zeros(n * n, dtype=np.float32), 1)
    assert p[5 * n + 5] == np.float32(-3.0e-4) / np.float32(1e-4)
    assert np.count_nonzero(p) == 1


def test_two_cell_pocket_matches_closed_form():
    # two adjacent fluid cells: [[-(1+e), 1], [1, -(1+e)]] p = b.  Both lie
    # under one coarse cell, which has no fluid neighbour: the coarse level
    # solves the near-null mode (eigenvalue -e) exactly, so repeated V-cycles
    # converge to the 2 x 2 solution (a V-cycle that lost the coarse
    # correction would need ~1e5 Gauss-Seidel sweeps for that mode).
    n, e = 16, float(np.float32(1e-4))  # MG_EPS is a binary32 constant (src/world/mod.rs:202)
    fluid = np.zeros(n * n, dtype=np.uint32)
    a, b = 3 * n + 4, 3 * n + 5
    fluid[a] = fluid[b] = 1
    rhs = np.zeros(n * n)
    rhs[a], rhs[b] = 1.0, -0.5
    M = np.array([[-(1 + e), 1.0], [1.0, -(1 + e)]])
    want = np.linalg.solve(M, [1.0, -0.5])
    assert abs(want[0] + want[1]) > 1e3  # the near-null (symmetric) mode dominates: (b_a + b_b) / (-e)
    p = oc.mg_solve(n, 8, fluid, rhs, np.zeros(n * n), 10, fp64=True)
    assert np.allclose([p[a], p[b]], want, rtol=1e-12)
    assert np.count_nonzero(p) == 2


def test_restriction_is_twice_the_mean_of_the_fluid_childrens_residual():
    # mg_restrict (multigrid.wgsl:82-105): rhs_c = 2 sum(res) / nf over fluid
    # children (restrict(constant r) = 2r), p = 0 so res = rhs; a coarse cell
    # with no fluid child gets 0 and is masked out
    n = 16
    fluid = np.ones(n * n, np.uint32)
    rc, fc = oc.mg_restrict(n, fluid, np.full(n * n, 0.75, np.float32), np.zeros(n * n, np.float32))
    assert np.all(rc == 1.5) and np.all(fc == 1)
    # half the children fluid: still twice their mean; none fluid: 0
    f2 = np.ones((n, n), np.uint32)
    f2[:, 0::2] = 0  # even columns solid: every coarse cell keeps 2 of 4
    f2[:4, :4] = 0   # the first 2 x 2 coarse cells lose all
    r = np.arange(n * n, dtype=np.float32).reshape(n, n) / 64.0
    rc, fc = oc.mg_restrict(n, f2.reshape(-1), r.reshape(-1), np.zeros(n * n, np.float32))
    rc, fc = rc.reshape(n // 2, n // 2), fc.reshape(n // 2, n // 2)
    for cy in range(n // 2):
        for cx in range(n // 2):
            kids = [(2 * cy + dy, 2 * cx + dx) for dy in (0, 1) for dx in (0, 1) if f2[2 * cy + dy, 2 * cx + dx]]
            if not kids:
                assert rc[cy, cx] == 0 and fc[cy, cx] == 0
            else:
                assert fc[cy, cx] == 1
                assert np.isclose(rc[cy, cx], 2.0 * np.mean([r[k] for k in kids]), rtol=1e-6)
    # with p: the residual rhs - A p enters (a linear p on an all-fluid level
    # has A p = -eps p inside, so res = rhs + eps p there)
    y, x = np.mgrid[0:n, 0:n]
    pl = (x + 2.0 * y).astype(np.float32)
    rc, _ = oc.mg_restrict(n, fluid, np.zeros(n * n, np.float32), pl.reshape(-1))
    rc = rc.reshape(n // 2, n // 2)
    inner = rc[1:-1, 1:-1]
    want = 2.0 * 1e-4 * np.array([[pl[2 * cy:2 * cy + 2, 2 * cx:2 * cx + 2].mean() for cx in range(1, n // 2 - 1)]
                                  for cy in range(1, n // 2 - 1)])
    assert np.allclose(inner, want, rtol=0, atol=3e-5)  # binary32 cancellation in s - (k + eps) p, |p| ~ 20


def test_linearity_in_binary64():
    n = 32
    rng = np.random.default_rng(5)
    fluid = (rng.random(n * n) > 0.15).astype(np.uint32)
    r1 = rng.standard_normal(n * n)
    r2 = rng.standard_normal(n * n)
    z = np.zeros(n * n)
    p1 = oc.mg_solve(n, 8, fluid, r1, z, 2, fp64=True)
    p2 = oc.mg_solve(n, 8, fluid, r2, z, 2, fp64=True)
    p12 = oc.mg_solve(n, 8, fluid, 2.0 * r1 - 3.0 * r2, z, 2, fp64=True)
    assert np.allclose(p12, 2.0 * p1 - 3.0 * p2, rtol=1e-9, atol=1e-9)


def test_binary32_tracks_binary64():
    n = 64
    rng = np.random.default_rng(6)
    fluid = (rng.random(n * n) > 0.1).astype(np.uint32)
    rhs = rng.standard_normal(n * n) * fluid
    p32 = oc.mg_solve(n, 8, fluid, rhs.astype(np.float32), np.zeros(n * n, np.float32), 3)
    p64 = oc.mg_solve(n, 8, fluid, rhs, np.zeros(n * n), 3, fp64=True)
    assert np.linalg.norm(p32 - p64) / np.linalg.norm(p64) < 1e-5


def _world_after_solve(precision, tol, seed=2, fp64=False):
    import ribossome_amd as ra
    from worlds import load, make_inputs, parity_params, parity_settings

    cfg = ra.WorldConfig(128, 64)
    inp = make_inputs(cfg, seed, rock=0.15, rubble=True, vent_rows=2)
    w = oc.OracleWorld(cfg, seed, fp64=fp64)
    load(w, inp, parity_params(seed), parity_settings(mg_tol=tol, mg_max_cycles=200, precision=precision))
    w.encode_steps(2)
    return w, cfg


def test_tolerance_mode_reaches_the_residual_by_an_independent_check():
    """PCG with the symmetric V-cycle (tolerance mode): recompute the true
    residual of the final pressure with the numpy operator above.  The solve
    stops on the recursively updated residual; in binary32 the true residual
    sits above it by up to the rounding floor of the pressure (~1e-5 relative
    on rock-laden worlds), so it is bounded by 2.5e-5 here."""
    import ribossome_amd as ra

    for precision in (0, 1):
        w, cfg = _world_after_solve(precision, 1e-5)
        n = cfg.fluid_size
        fluid = (w.download(ra.SOLID_MASK) == 0).astype(np.uint32)
        A, f = operator(n, fluid, 1e-4)
        p = w.download_f64(ra.PRESSURE)
        b = w.download_f64(ra.DIVERGENCE)
        rel = np.linalg.norm(b[f] - A @ p[f]) / np.linalg.norm(b[f])
        it, rel_reported = w.mg_stats()
        assert 0 < it < 200
        assert rel < 2.5e-5, (precision, rel, it)
        assert abs(rel - rel_reported) < 0.2 * rel_reported + 1e-9


def test_binary64_oracle_bounds_the_binary32_fluid_error():
    """The same tolerance-mode world run with binary32 and binary64 fields at
    a tolerance binary32 cannot reach.  The binary64 solve reaches it by the
    numpy residual; the binary32 one stops on its recursive residual with a
    true residual at its floor (1-2e-5 here: the rounding of a pressure of
    magnitude ~100).  Velocity -- the field the world uses -- differs by
    under 1e-4 relative L2; pressure, fixed only up to cond(A) times the
    residual, by a few 1e-4 (bound 1e-3)."""
    import ribossome_amd as ra

    w32, cfg = _world_after_solve(0, 1e-7)
    w64, _ = _world_after_solve(0, 1e-7, fp64=True)
    n = cfg.fluid_size
    fluid = (w64.download(ra.SOLID_MASK) == 0).astype(np.uint32)
    A, f = operator(n, fluid, 1e-4)
    for w, bound in ((w64, 1.5e-7), (w32, 2.5e-5)):
        p = w.download_f64(ra.PRESSURE)
        b = w.download_f64(ra.DIVERGENCE)
        rel = np.linalg.norm(b[f] - A @ p[f]) / np.linalg.norm(b[f])
        assert rel < bound, (rel, bound)
    for which, tol in ((ra.VELOCITY, 1e-4), (ra.PRESSURE, 1e-3)):
        a, b = w32.download_f64(which), w64.download_f64(which)
        e = np.linalg.norm(a - b) / np.linalg.norm(b)
        assert e < tol, (which, e)


def test_stationary_vcycle_mode_reaches_the_residual():
    """The stationary tolerance mode (the reference V-cycle repeated until the
    relative residual reaches the tolerance): the final pressure's residual,
    recomputed with the numpy operator above, is within the tolerance, and the
    reported residual is that one."""
    import ribossome_amd as ra
    from worlds import load, make_inputs, parity_params, parity_settings

    cfg = ra.WorldConfig(128, 64)
    inp = make_inputs(cfg, 2, rock=0.15, rubble=True, vent_rows=2)
    w = oc.OracleWorld(cfg, 2)
    load(w, inp, parity_params(2), parity_settings(mg_tol=1e-4, mg_max_cycles=400, mg_stationary=1))
    w.encode_steps(2)
    n = cfg.fluid_size
    fluid = (w.download(ra.SOLID_MASK) == 0).astype(np.uint32)
    A, f = operator(n, fluid, 1e-4)
    p = w.download_f64(ra.PRESSURE)
    b = w.download_f64(ra.DIVERGENCE)
    rel = np.linalg.norm(b[f] - A @ p[f]) / np.linalg.norm(b[f])
    it, rel_reported = w.mg_stats()
    assert 1 < it < 400
    assert rel <= 1.0001e-4 and abs(rel - rel_reported) < 1e-3 * rel_reported


def test_prolongation_is_injection_on_fluid_cells():
    # mg_prolong (multigrid.wgsl:108-114): every fluid fine cell adds its
    # parent's coarse value; solid cells are left alone
    n = 16
    rng = np.random.default_rng(3)
    fluid = (rng.random(n * n) > 0.2).astype(np.uint32)
    p = rng.standard_normal(n * n).astype(np.float32)
    pc = rng.standard_normal((n // 2) ** 2).astype(np.float32)
    out = oc.mg_prolong(n, fluid, p, pc)
    y, x = np.divmod(np.arange(n * n), n)
    want = np.where(fluid > 0, p + pc[(y // 2) * (n // 2) + x // 2], p)
    assert np.array_equal(out, want.astype(np.float32))


def test_bilinear_transfer_is_adjoint_and_interpolates_constants():
    # the tolerance mode's transfer: P c interpolates with weights over fluid
    # coarse cells renormalised by their sum W, and R is its exact transpose,
    # (R r, c) = (r, P c), which the conjugate-gradient preconditioner needs
    n = 32
    rng = np.random.default_rng(11)
    fl = (rng.random((n, n)) > 0.25).astype(np.uint32)
    flc = np.array([[fl[2 * cy:2 * cy + 2, 2 * cx:2 * cx + 2].any() for cx in range(n // 2)] for cy in range(n // 2)],
                   np.uint32)
    fl2 = fl.copy()
    fl2[np.repeat(np.repeat(flc, 2, 0), 2, 1) == 0] = 0  # a fluid fine cell has a fluid parent
    c = rng.standard_normal((n // 2) ** 2)
    r = rng.standard_normal(n * n) * fl2.reshape(-1)
    Pc, Rr = oc.bl_transfer(n, fl2.reshape(-1), flc.reshape(-1), c * flc.reshape(-1), r)
    lhs = float(np.dot(Rr * flc.reshape(-1), c * flc.reshape(-1)))
    rhs = float(np.dot(r, Pc * fl2.reshape(-1)))
    assert abs(lhs - rhs) <= 1e-12 * (abs(lhs) + abs(rhs) + 1.0)
    # a constant coarse field interpolates to the same constant on every
    # fluid fine cell (the weights are renormalised by W)
    Pc1, _ = oc.bl_transfer(n, fl2.reshape(-1), flc.reshape(-1), np.full((n // 2) ** 2, 2.5) * flc.reshape(-1),
                            np.zeros(n * n))
    assert np.allclose(Pc1[fl2.reshape(-1) > 0], 2.5, rtol=1e-14)
    # on an all-fluid level the interior weights are the 3/4, 1/4 tensor product
    ones = np.ones(n * n, np.uint32)
    lin = np.add.outer(np.arange(n // 2), np.zeros(n // 2)).reshape(-1).astype(np.float64)  # c = cy
    Pl, _ = oc.bl_transfer(n, ones, np.ones((n // 2) ** 2, np.uint32), lin, np.zeros(n * n))
    Pl = Pl.reshape(n, n)
    fy = np.arange(2, n - 2)
    want = np.where(fy % 2 == 0, 0.75 * (fy // 2) + 0.25 * (fy // 2 - 1), 0.75 * (fy // 2) + 0.25 * (fy // 2 + 1))
    assert np.allclose(Pl[2:-2, 5], want, rtol=1e-14)


def test_exact_coarse_solve_inverts_the_coarse_operator():
    # the tolerance mode's coarsest level: the dense binary64 inverse of the
    # eps-anchored operator over its fluid unknowns satisfies A inv = I
    rng = np.random.default_rng(12)
    for n, frac in ((8, 0.0), (16, 0.2), (32, 0.35)):
        fluid = (rng.random(n * n) >= frac).astype(np.uint32)
        err = oc.coarse_inverse_error(n, fluid, 1e-4 * 2 ** 6)
        assert 0.0 <= err < 1e-9, (n, err)
    assert oc.coarse_inverse_error(8, np.zeros(64, np.uint32), 1e-4) == -1.0
