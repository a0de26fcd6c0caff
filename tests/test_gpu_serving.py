"""The serving form on the GPU: the admission figure against the world's
real allocations, scenes identical alone and in the mix, and admission by
the device-memory budget."""
import time

import numpy as np
import pytest

import ribossome_amd as ra
from worlds import bit_diff, load, make_inputs, parity_params, parity_settings

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("tol", [0.0, 1e-5])
def test_footprint_matches_allocations(tol):
    print(f"case footprint tol={tol}", flush=True)
    for cfg in (ra.WorldConfig(64, 32), ra.WorldConfig(96, 96), ra.WorldConfig(256, 128), ra.WorldConfig(2048, 1024)):
        s = parity_settings(mg_tol=tol, mg_max_cycles=8 if tol else 1)
        inp = make_inputs(cfg, 3, rock=0.1)
        w = ra.World(cfg, 3)
        load(w, inp, parity_params(3), s)
        w.encode_steps(1)
        w.sync()
        have, fp = w.device_bytes(), ra.world_footprint(cfg, s)
        print(f"{cfg.grid_size}^2: world holds {have} B, footprint {fp} B", flush=True)
        if tol:
            # the exact coarse solve is charged at its largest (every cell fluid)
            n = w.mg_levels()[-1][0]
            assert have <= fp <= have + (n * n) ** 2 * 8 + n * n * 4
        else:
            assert have == fp
        w.close()
        if tol:
            # under a moving terrain every buffer of the coarse solve takes its capacity
            s.terrain_enabled = 1
            w = ra.World(cfg, 3)
            load(w, inp, parity_params(3), s)
            w.encode_steps(1)
            w.sync()
            have, fp = w.device_bytes(), ra.world_footprint(cfg, s)
            print(f"{cfg.grid_size}^2 moving terrain: world holds {have} B, footprint {fp} B", flush=True)
            assert have == fp
            w.close()


SCENES = [  # (grid side, seed, steps, tolerance-mode precision or None for the reference schedule)
    (128, 31, 150, 1), (256, 32, 90, None), (192, 33, 130, 0), (384, 34, 70, 1), (128, 35, 200, None),
    (512, 36, 60, 0),
]


def scene_request(side, seed, steps, prec):
    cfg = ra.WorldConfig(side, side // 2)
    inp = make_inputs(cfg, seed, rock=0.12, rubble=True, vent_rows=max(1, side // 64))
    p = parity_params(seed, aggregation=0.02, thermal_activation=1.0)
    s = parity_settings(ledger_check=1)
    if prec is not None:
        s.mg_tol, s.mg_max_cycles, s.precision = 1e-5, 60, prec
    return cfg, p, s, inp


def run_server(srv, reqs):
    tickets = [srv.submit(cfg, seed, steps, p, s, inp) for (cfg, p, s, inp), (_, seed, steps, _) in zip(reqs, SCENES)]
    t0 = time.time()
    peak = 0
    while any(srv.status(t).state in (ra.SCENE_QUEUED, ra.SCENE_RUNNING) for t in tickets):
        assert time.time() - t0 < 120, "scenes did not finish"
        peak = max(peak, srv.stats()[1])
        time.sleep(0.002)
    out = []
    for t, (cfg, _, _, _) in zip(tickets, reqs):
        st = srv.status(t)
        assert st.state == ra.SCENE_DONE, (t, st.state, st.error, srv.error(t))
        out.append({k: srv.download(t, cfg, k) for k in (ra.CHEM, ra.VELOCITY, ra.PRESSURE, ra.TEMP)})
        srv.release(t)
    return out, peak


def test_scenes_identical_alone_and_in_the_mix():
    """Each scene's buffers after its steps are bit-identical whether it ran
    alone (one scene admitted at a time, on its own stream) or overlapped
    with five others of other sizes and solver modes -- on their own streams,
    or batched (three batches: the reference V-cycle and the tolerance mode
    in each precision, rebuilt as scenes finish) -- and equal to a plain
    World run with the same chunking."""
    print("case serving alone vs mix", flush=True)
    reqs = [scene_request(*sc) for sc in SCENES]
    budget = 64 << 30
    alone, peak1 = run_server(ra.Server(budget, max_running=1, threads=1, batching=False), reqs)
    streams, peaks = run_server(ra.Server(budget, threads=4, batching=False), reqs)
    batched, peakb = run_server(ra.Server(budget, threads=4), reqs)
    print(f"peak running: alone {peak1}, streams {peaks}, batched {peakb}", flush=True)
    assert peak1 == 1 and peaks >= 3 and peakb >= 3
    for mixed, how in ((streams, "streams"), (batched, "batched")):
        for k, (a, b) in enumerate(zip(alone, mixed)):
            assert np.array_equal(a[ra.CHEM], b[ra.CHEM]), f"scene {k} ({how}): chem differs"
            for which in (ra.VELOCITY, ra.PRESSURE, ra.TEMP):
                assert bit_diff(a[which], b[which]) == 0, f"scene {k} ({how}): {ra.BUFFER_NAMES[which]} differs"
            cfg, p, s, inp = reqs[k]
            assert ra.Ledger.from_cells(b[ra.CHEM]).channels() == ra.Ledger.from_cells(inp[ra.CHEM]).channels()
    cfg, p, s, inp = reqs[2]
    w = ra.World(cfg, SCENES[2][1])
    load(w, inp, p, s)
    left = SCENES[2][2]
    while left:
        left -= w.encode_steps(min(64, left))
    assert np.array_equal(w.read_cells_blocking(), alone[2][ra.CHEM])
    assert bit_diff(w.download(ra.PRESSURE), alone[2][ra.PRESSURE]) == 0


def test_admission_by_budget():
    """With room for two scenes' footprints, never more than two run; the
    charged bytes stay within the budget; release returns them."""
    print("case serving budget", flush=True)
    reqs = [scene_request(256, 40 + i, 80, None) for i in range(5)]
    fp = ra.world_footprint(reqs[0][0], reqs[0][2])
    srv = ra.Server(2 * fp + fp // 2)
    tickets = [srv.submit(cfg, 40 + i, 80, p, s, inp) for i, (cfg, p, s, inp) in enumerate(reqs)]
    t0 = time.time()
    most = 0
    while any(srv.status(t).state in (ra.SCENE_QUEUED, ra.SCENE_RUNNING) for t in tickets):
        assert time.time() - t0 < 120
        q, r, f, used = srv.stats()
        most = max(most, r)
        assert used <= 2 * fp + fp // 2
        # finished scenes keep their memory until released
        for t in tickets:
            if srv.status(t).state == ra.SCENE_DONE:
                srv.release(t)
                tickets = [u for u in tickets if u != t]
                break
        time.sleep(0.001)
    assert 1 <= most <= 2
    for t in tickets:
        srv.release(t)
    assert srv.stats() == (0, 0, 0, 0)
    srv.close()
