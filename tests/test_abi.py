"""The C-ABI library loads without a GPU, exports every declared symbol and
keeps the reference struct layouts."""
import ctypes
import os
import re

import numpy as np

import ribossome_amd as ra

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    inc = os.path.join(ROOT, "include")
    for f in os.listdir(inc):
        if f.endswith(".h"):
            names |= set(re.findall(r"RIBO_API\s+[\w\s\*]+?\b(ribo_\w+)\s*\(", open(os.path.join(inc, f)).read()))
    return names


def test_library_exports_every_declared_symbol():
    L = ra.lib()
    syms = declared_symbols()
    assert len(syms) >= 25
    for s in sorted(syms):
        assert hasattr(L, s), s
    assert L.ribo_abi_version() == 5


def test_struct_layouts_match_reference():
    # SimParams: 116 four-byte fields = 464 bytes (src/params.rs:86-487,
    # size checked against WGSL by tests/shaders.rs:59-77)
    assert ctypes.sizeof(ra.SimParams) == 464
    assert len(ra.SimParams._fields_) == 116
    assert ctypes.sizeof(ra.MgLevel) == 32
    p = ra.SimParams.default(5)
    # spot values of SimParams::default() (src/params.rs:553-690)
    assert p.seed == 5 and p.epoch == 0 and p.transport_every == 3
    assert np.float32(p.dt) == np.float32(0.017) and np.float32(p.fluid_dt) == np.float32(0.034)
    assert np.float32(p.fluid_decay) == np.float32(0.99) and np.float32(p.fluid_viscosity) == np.float32(3.7)
    assert np.float32(p.aggregation) == np.float32(0.03) and np.float32(p.reactivation_rate) == np.float32(1e-5)
    assert np.float32(p.vent_cycle_amp) == np.float32(0.91) and p.vent_cycle_period == 7000.0
    s = ra.WorldSettings.default()
    assert (s.fluid_enabled, s.fluid_substep, s.multigrid, s.mg_cycles, s.jacobi_iters) == (1, 2, 1, 1, 128)
    # the reference's WorldSettings::default() (src/world/mod.rs:121-134):
    # terrain on, and its light step runs every step (light_mode 1)
    assert s.mg_min_size == 8 and s.terrain_enabled == 1 and s.light_mode == 1


def test_buffer_sizes():
    cfg = ra.WorldConfig.DEFAULT
    G, F = 2048 * 2048, 1024 * 1024
    assert ra.buffer_bytes(cfg, ra.CHEM) == 16 * G
    assert ra.buffer_bytes(cfg, ra.GAMMA) == 4 * G
    assert ra.buffer_bytes(cfg, ra.VELOCITY) == 8 * F
    assert ra.buffer_bytes(cfg, ra.LIGHT) == 8 * 1024 * 1024
    assert ra.buffer_bytes(cfg, ra.NET_FLUX) == 8 * 1024
    assert ra.buffer_bytes(cfg, ra.FORCE_VECTORS) == 8 * F and ra.buffer_bytes(cfg, ra.VELOCITY_SMOOTH) == 8 * F
    assert ra.BUF_COUNT == len(ra.BUFFER_NAMES) == 16
    assert ra.buffer_bytes(cfg, 99) == 0


def test_world_create_fails_cleanly_without_device():
    if ra.device_available():
        return
    h = ctypes.c_void_p()
    c = ra.WorldConfig(64, 32).c()
    rc = ra.lib().ribo_world_create(ctypes.byref(c), 1, ctypes.byref(h))
    assert rc == -3 and not h.value
    assert b"device" in ra.lib().ribo_last_error()


def test_bad_config_rejected():
    h = ctypes.c_void_p()
    c = ra.WorldConfig(100, 30).c()
    assert ra.lib().ribo_world_create(ctypes.byref(c), 1, ctypes.byref(h)) == -1


def test_generators_deterministic_and_shaped():
    cfg = ra.WorldConfig(128, 64)
    g1, g2 = ra.gen_gamma(cfg, 3, 0.15, True), ra.gen_gamma(cfg, 3, 0.15, True)
    assert np.array_equal(g1, g2)
    assert abs((g1 >= 3).mean() - 0.15) < 1e-3
    assert set(np.unique(g1)) <= {0, 1, 2, 6}
    ch = ra.gen_chem(cfg, 3, 4.0, g1)
    c = ch.reshape(-1, 4)
    assert np.all(c[g1 >= 3] == 0)
    open_ = g1 < 3
    counts = np.concatenate([(c[open_] & 0xFFFF).ravel(), (c[open_] >> 16).ravel()])
    assert abs(counts.mean() - 4.0) < 0.05 and abs(counts.var() - 4.0) < 0.15
    v = ra.gen_velocity(cfg, 3, 10.0, g1).reshape(-1, 2)
    assert abs(np.sqrt((v.astype(np.float64) ** 2).sum(1)).max() - 10.0) < 1e-4


def test_scene_struct_layout_matches_header(tmp_path):
    """RiboScene / RiboSceneStatus as the C compiler lays them out."""
    import subprocess

    src = tmp_path / "sz.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "ribo_server.h"\n'
                   'int main(void){printf("%zu %zu %zu %zu %zu %zu\\n", sizeof(RiboScene), offsetof(RiboScene, params),'
                   ' offsetof(RiboScene, settings), offsetof(RiboScene, input), offsetof(RiboScene, input_bytes),'
                   ' sizeof(RiboSceneStatus)); return 0;}\n')
    exe = tmp_path / "sz"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    S = ra._CScene
    assert got == [ctypes.sizeof(S), S.params.offset, S.settings.offset, S.input.offset, S.input_bytes.offset,
                   ctypes.sizeof(ra.SceneStatus)]


def test_footprint_model():
    """The admission figure grows with the world and counts the exact coarse
    solve only in the tolerance mode."""
    s = ra.WorldSettings.default()
    small, big = ra.world_footprint(ra.WorldConfig(256, 128), s), ra.world_footprint(ra.WorldConfig(2048, 1024), s)
    assert 60 * 256 * 256 < small < 120 * 256 * 256
    assert 60 * 2048 * 2048 < big < 120 * 2048 * 2048
    t = ra.WorldSettings.default()
    t.mg_tol = 1e-5
    t.terrain_enabled = 0
    assert ra.world_footprint(ra.WorldConfig(256, 128), t) == small + 64 * 64 * 8 + 64 * 4
    # under a moving terrain: the mask snapshot, the elimination work and the count
    t.terrain_enabled = 1
    assert ra.world_footprint(ra.WorldConfig(256, 128), t) == small + 64 * 64 * 8 + 64 * 4 + 64 + 2 * 64 * 64 * 8 + 4


def test_server_host_logic_without_device():
    """Submission copies and checks the inputs, refuses a scene larger than
    the budget, and a scene whose world cannot be created ends FAILED with
    the library's error; release frees it."""
    if ra.device_available():
        return
    cfg = ra.WorldConfig(64, 32)
    fp = ra.world_footprint(cfg, ra.WorldSettings.default())
    srv = ra.Server(hbm_budget=3 * fp)
    p, s = ra.SimParams.default(1), ra.WorldSettings.default()
    with np.testing.assert_raises(ra.RiboError):
        srv.submit(ra.WorldConfig(2048, 1024), 1, 10, p, s, {})
    gamma = np.zeros(cfg.cells(), np.uint32)
    t = srv.submit(cfg, 1, 10, p, s, {ra.GAMMA: gamma})
    import time

    t0 = time.time()
    while srv.status(t).state in (ra.SCENE_QUEUED, ra.SCENE_RUNNING):
        assert time.time() - t0 < 10
        time.sleep(0.01)
    st = srv.status(t)
    assert st.state == ra.SCENE_FAILED and st.error == -3
    assert srv.stats() == (0, 0, 1, 0)
    srv.release(t)
    assert srv.stats() == (0, 0, 0, 0)
    srv.close()


def test_batch_argument_errors_need_no_gpu():
    L = ra.lib()
    out = ctypes.c_void_p()
    assert L.ribo_batch_create(None, 0, ctypes.byref(out)) == -1
    assert b"empty" in L.ribo_last_error()
    assert L.ribo_batch_destroy(None) == -1
    assert L.ribo_batch_encode_steps(None, 4) == -1
    n = ctypes.c_uint64()
    assert L.ribo_batch_launches(None, ctypes.byref(n)) == -1
