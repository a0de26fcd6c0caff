"""Terrain grain dynamics on the HIP world against the oracle: both follow
the same admissible order of the reference's concurrent grain moves, so
grains, monomers, the slope and every fluid field are bit-identical, with
graph replay and with eager recording; grains and monomers are conserved
at a larger size under the device's per-step ledger check."""
import numpy as np
import pytest

import ribossome_amd as ra
import oracle_ctypes as oc
from test_gpu_parity import compare
from worlds import bit_diff, load, make_inputs, parity_params, parity_settings

pytestmark = pytest.mark.gpu

ACTIVE = dict(sediment_settle=1.0, sediment_compaction=0.3, rock_crumble=1.0, sediment_transport=3.0,
              sediment_threshold=0.05, diffusion=5.0)

CASES = {
    "parity_profile": dict(cfg=ra.WorldConfig(128, 64), seed=41, params={}, settings={}),
    "ref_schedule": dict(cfg=ra.WorldConfig(128, 64), seed=43, params=dict(transport_every=3),
                         settings=dict(fluid_substep=2)),
    "jacobi_eager": dict(cfg=ra.WorldConfig(96, 48), seed=45, params={}, settings=dict(multigrid=0, jacobi_iters=16,
                                                                                       no_graphs=1)),
    # the light step reads the grains the terrain stage has just moved
    "light_schedule": dict(cfg=ra.WorldConfig(128, 64), seed=49,
                           params=dict(uv_strength=3.0, uv_depth=4.0, day_period=0.0, direct_photoactivation=0.5,
                                       sun_heat=1.0),
                           settings=dict(light_mode=1, light_rows_per_step=2)),
}


def build(case):
    c = CASES[case]
    cfg = c["cfg"]
    inp = make_inputs(cfg, c["seed"], rock=0.12, rubble=True, vent_rows=4, vcells_per_step=2.0)
    p = parity_params(c["seed"], **{**ACTIVE, **c["params"]})
    s = parity_settings(terrain_enabled=1, **c["settings"])
    gw, ow = ra.World(cfg, c["seed"]), oc.OracleWorld(cfg, c["seed"])
    load(gw, inp, p, s)
    load(ow, inp, p, s)
    return inp, gw, ow


@pytest.mark.parametrize("case", list(CASES))
def test_terrain_dynamics_bit_exact(case):
    print(f"case terrain {case}", flush=True)
    inp, gw, ow = build(case)
    for k, n in enumerate((1, 1, 3, 7, 20)):
        assert gw.encode_steps(n) == n and ow.encode_steps(n) == n
        nd = bit_diff(gw.download(ra.GAMMA), ow.download(ra.GAMMA))
        print(f"[{case} chunk {k}] gamma words differing: {nd}", flush=True)
        assert nd == 0 and compare(gw, ow, f"{case} chunk {k}"), f"{case}: GPU differs from the oracle"
    g = gw.download(ra.GAMMA)
    assert int(np.count_nonzero(g != inp[ra.GAMMA])) > 20
    assert int(g.sum()) == int(inp[ra.GAMMA].sum())


def test_terrain_conserves_at_size():
    print("case terrain conservation 512", flush=True)
    cfg = ra.WorldConfig(512, 256)
    inp = make_inputs(cfg, 47, rock=0.12, rubble=True, vent_rows=8, vcells_per_step=2.0)
    p = parity_params(47, **ACTIVE)
    s = parity_settings(terrain_enabled=1, ledger_check=1, fluid_substep=2)
    gw = ra.World(cfg, 47)
    load(gw, inp, p, s)
    gw.run(300)
    g = gw.download(ra.GAMMA)
    assert int(g.sum()) == int(inp[ra.GAMMA].sum())
    assert int(np.count_nonzero(g != inp[ra.GAMMA])) > 1000
    assert gw.ledger_status() is None
    assert ra.Ledger.from_cells(gw.download(ra.CHEM)).channels() == ra.Ledger.from_cells(inp[ra.CHEM]).channels()


STILL = dict(sediment_settle=0.0, sediment_compaction=0.0, rock_crumble=0.0, sediment_transport=0.0)


@pytest.mark.parametrize("precision", [0, 1])
def test_device_built_coarse_inverse_matches_host(precision):
    """Under a moving terrain the tolerance mode builds the coarsest level's
    inverse on the device (and again whenever the coarsest fluid mask
    changes).  On a terrain of whole rock that nothing moves, the world with
    the terrain on runs the device-built inverse and the world with it off
    the host-built one: the same pressures, bit for bit."""
    print(f"case device coarse inverse precision={precision}", flush=True)
    cfg = ra.WorldConfig(128, 64)
    inp = make_inputs(cfg, 51, rock=0.12, rubble=False, vent_rows=4)
    p = parity_params(51, **STILL)
    out = []
    for terrain in (1, 0):
        w = ra.World(cfg, 51)
        load(w, inp, p, parity_settings(terrain_enabled=terrain, mg_tol=1e-5, mg_max_cycles=40, precision=precision))
        w.run(12)
        it, rel = w.mg_stats()
        print(f"terrain {terrain}: {it} iterations, true rel {rel:.3e}", flush=True)
        assert rel <= 1e-5
        out.append({k: w.download(k) for k in (ra.GAMMA, ra.PRESSURE, ra.VELOCITY, ra.CHEM)})
        w.close()
    assert np.array_equal(out[0][ra.GAMMA], inp[ra.GAMMA])  # nothing moved
    for k in (ra.PRESSURE, ra.VELOCITY, ra.CHEM):
        assert bit_diff(out[0][k], out[1][k]) == 0, ra.BUFFER_NAMES[k]


@pytest.mark.parametrize("precision", [0, 1])
def test_tolerance_mode_with_moving_terrain(precision):
    """The tolerance mode with the grains moving: every solve reaches the
    tolerance on its true residual, the fields agree with the oracle's to
    the tolerance-mode bounds, and grains and monomers are conserved."""
    print(f"case moving terrain tolerance precision={precision}", flush=True)
    cfg = ra.WorldConfig(128, 64)
    inp = make_inputs(cfg, 53, rock=0.12, rubble=True, vent_rows=4, vcells_per_step=2.0)
    p = parity_params(53, **ACTIVE)
    s = parity_settings(terrain_enabled=1, mg_tol=1e-5, mg_max_cycles=60, precision=precision, ledger_check=1)
    gw, ow = ra.World(cfg, 53), oc.OracleWorld(cfg, 53)
    load(gw, inp, p, s)
    load(ow, inp, p, s)
    for k in range(6):
        gw.encode_steps(1)
        ow.encode_steps(1)
        (cyc, rel), (ocyc, orel) = gw.mg_stats(), ow.mg_stats()
        print(f"step {k}: GPU {cyc} iterations, true rel {rel:.3e}; oracle {ocyc}, {orel:.3e}", flush=True)
        assert rel <= 1e-5 and orel <= 1e-5
        assert abs(cyc - ocyc) <= max(3, ocyc // 5)
    for which, tol in ((ra.VELOCITY, 1e-4), (ra.PRESSURE, 1e-3)):
        a, b = gw.download(which).astype(np.float64), ow.download(which).astype(np.float64)
        e = float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
        print(f"{ra.BUFFER_NAMES[which]} rel L2 vs oracle {e:.2e}", flush=True)
        assert e < tol
    g = gw.download(ra.GAMMA)
    assert int(g.sum()) == int(inp[ra.GAMMA].sum()) and int(np.count_nonzero(g != inp[ra.GAMMA])) > 0
    assert gw.ledger_status() is None
    assert ra.Ledger.from_cells(gw.download(ra.CHEM)).channels() == ra.Ledger.from_cells(inp[ra.CHEM]).channels()
