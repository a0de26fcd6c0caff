""".ribo scene files (src/world/snapshot.rs) on the host: the LZ4 block codec
against the block format's own rules and an independent decoder written
here, the file layout, and the reading rules of World::load_scene."""
import json
import struct

import numpy as np
import pytest

import ribossome_amd as ra

REF_SETTINGS = ["fluid_enabled", "fluid_substep", "multigrid", "mg_cycles", "jacobi_iters", "light_rows_per_step",
                "light_interval", "terrain_enabled", "contact_enabled"]
AGENT_BLOCKS = ["agents", "genomes", "bodies", "organs", "rna_tails4", "bitten", "body_pos", "signals4", "joint_angle",
                "joint_base", "joint_state", "joint_active", "sensor_mem", "bonds2", "pair_mask"]


def py_lz4_decode(block: bytes) -> bytes:
    """A plain restatement of the LZ4 block format, to check the library's
    compressor independently of its decompressor."""
    out = bytearray()
    i = 0
    while i < len(block):
        token = block[i]
        i += 1
        lit = token >> 4
        if lit == 15:
            while True:
                b = block[i]
                i += 1
                lit += b
                if b != 255:
                    break
        out += block[i:i + lit]
        i += lit
        if i >= len(block):
            break
        off = block[i] | (block[i + 1] << 8)
        i += 2
        ml = (token & 15) + 4
        if token & 15 == 15:
            while True:
                b = block[i]
                i += 1
                ml += b
                if b != 255:
                    break
        assert 0 < off <= len(out)
        for _ in range(ml):
            out.append(out[-off])
    return bytes(out)


def check_block_rules(block: bytes, n: int):
    """The encoder's side of the format: the last 5 bytes are literals and no
    match starts in the last 12 bytes of the output."""
    i, pos = 0, 0
    while i < len(block):
        token = block[i]
        i += 1
        lit = token >> 4
        if lit == 15:
            while True:
                b = block[i]
                i += 1
                lit += b
                if b != 255:
                    break
        i += lit
        pos += lit
        if i >= len(block):
            break
        i += 2
        ml = (token & 15) + 4
        if token & 15 == 15:
            while True:
                b = block[i]
                i += 1
                ml += b
                if b != 255:
                    break
        assert pos + 12 <= n, "a match starts in the last 12 bytes"
        pos += ml
        assert pos <= n - 5, "a match covers the last 5 bytes"
    assert pos == n


def test_lz4_decoder_on_format_vectors():
    assert ra.lz4_decompress(b"\x50hello", 5) == b"hello"
    # one literal, then a match at offset 1 of length 10 (overlapping copy), then 5 literals
    assert ra.lz4_decompress(b"\x16a\x01\x00\x50bcdef", 16) == b"a" * 11 + b"bcdef"
    # 20 literals: 15 in the token plus one more length byte of 5
    lit = bytes(range(20))
    assert ra.lz4_decompress(b"\xf0\x05" + lit, 20) == lit
    # a match of 4 + 15 + 255 + 3 = 277 bytes at offset 2
    blk = b"\x2fab\x02\x00\xff\x03" + b"\x50vwxyz"
    assert ra.lz4_decompress(blk, 284) == (b"ab" * 140)[:279] + b"vwxyz"
    with pytest.raises(ra.RiboError):
        ra.lz4_decompress(b"\x16a\x05\x00\x50bcdef", 16)  # offset beyond the output
    with pytest.raises(ra.RiboError):
        ra.lz4_decompress(b"\x16a\x01", 16)  # truncated
    with pytest.raises(ra.RiboError):
        ra.lz4_decompress(b"\x50hello", 4)  # destination too small


@pytest.mark.parametrize("n", [0, 1, 5, 12, 13, 64, 1000, 100_000])
def test_lz4_round_trip(n):
    rng = np.random.default_rng(n)
    for data in (rng.integers(0, 256, n, dtype=np.uint8).tobytes(),
                 rng.integers(0, 3, n, dtype=np.uint8).tobytes(),
                 (b"ribossome" * (n // 9 + 1))[:n]):
        c = ra.lz4_compress(data)
        assert py_lz4_decode(c) == data
        assert ra.lz4_decompress(c, n) == data
        check_block_rules(c, n)
    z = ra.lz4_compress(bytes(1 << 20))
    assert len(z) < 5000 and py_lz4_decode(z) == bytes(1 << 20)


def grids(cfg, seed):
    rng = np.random.default_rng(seed)
    G, F = cfg.cells(), cfg.fluid_cells()
    return {ra.CHEM: rng.integers(0, 1 << 20, 4 * G, dtype=np.uint32), ra.GAMMA: rng.integers(0, 7, G, dtype=np.uint32),
            ra.VELOCITY: rng.standard_normal(2 * F).astype(np.float32),
            ra.PRESSURE: rng.standard_normal(F).astype(np.float32),
            ra.TEMP: rng.random(F).astype(np.float32), ra.REDOX: rng.random(F).astype(np.float32),
            ra.HEAT_SRC: rng.random(2 * F).astype(np.float32)}


def parse_file(path):
    raw = open(path, "rb").read()
    assert raw[:8] == b"RIBOSCN1"
    (hl,) = struct.unpack("<I", raw[8:12])
    header = json.loads(raw[12:12 + hl])
    blocks, p = {}, 12 + hl
    while p < len(raw):
        (nl,) = struct.unpack("<H", raw[p:p + 2])
        name = raw[p + 2:p + 2 + nl].decode()
        p += 2 + nl
        (cl,) = struct.unpack("<Q", raw[p:p + 8])
        p += 8
        (size,) = struct.unpack("<I", raw[p:p + 4])
        data = py_lz4_decode(raw[p + 4:p + cl])
        assert len(data) == size
        blocks[name] = data
        p += cl
    return header, blocks


def test_scene_write_layout_and_round_trip(tmp_path):
    cfg = ra.WorldConfig(64, 32)
    g = grids(cfg, 3)
    p = ra.SimParams.default(77)
    p.epoch, p.diffusion, p.transport_every = 1234, 2.5, 3
    s = ra.WorldSettings.default()
    s.fluid_substep, s.multigrid, s.light_interval = 3, 0, 7
    path = str(tmp_path / "cena.ribo")
    ra.scene_write(path, cfg, p, s, g)
    header, blocks = parse_file(path)
    assert header["versao"] == 2 and header["agentes"] == 0
    assert header["cfg"] == {"grid_size": 64, "fluid_size": 32, "world_units_per_cell": 30, "max_agents": 0}
    assert list(header["params"]) == [f for f, _ in ra.SimParams._fields_]
    assert header["params"]["epoch"] == 1234 and header["params"]["seed"] == 77
    assert np.float32(header["params"]["diffusion"]) == np.float32(2.5)
    assert list(header["settings"]) == REF_SETTINGS
    assert header["settings"]["multigrid"] is False and header["settings"]["fluid_substep"] == 3
    for name in AGENT_BLOCKS + ["slots"]:
        assert blocks[name] == b""
    assert len(blocks["life_counters"]) == 32
    for name, which in ra.SCENE_GRIDS:
        assert blocks[name] == g[which].tobytes(), name
    # read back: every field of SimParams and the reference's settings fields
    q, t, h = ra.scene_read(path, cfg, seed=5)
    assert bytes(q) == bytes(p)
    for k in REF_SETTINGS:
        assert getattr(t, k) == getattr(s, k), k
    for which in g:
        assert np.array_equal(h[which].view(np.uint32), g[which].view(np.uint32))


def write_raw(path, header, blocks):
    h = json.dumps(header).encode()
    out = b"RIBOSCN1" + struct.pack("<I", len(h)) + h
    for name, data in blocks.items():
        c = struct.pack("<I", len(data)) + ra.lz4_compress(data)
        out += struct.pack("<H", len(name)) + name.encode() + struct.pack("<Q", len(c)) + c
    open(path, "wb").write(out)


def test_scene_read_rules(tmp_path):
    """World::load_scene: parameters by name over the defaults (the caller's
    seed kept), unknown names ignored, serde defaults for missing settings, a
    smaller max_agents accepted, another world size refused, redox optional."""
    cfg = ra.WorldConfig(64, 32, 30, 1000)
    g = grids(cfg, 4)
    blocks = {name: g[which].tobytes() for name, which in ra.SCENE_GRIDS if name not in ("redox", "heat_src")}
    header = {"versao": 2, "cfg": {"grid_size": 64, "fluid_size": 32, "world_units_per_cell": 30, "max_agents": 500},
              "params": {"epoch": 99.0, "diffusion": 4.0, "no_such_param": 1.0, "transport_every": 5.0},
              "settings": {"fluid_substep": 4, "multigrid": False}}
    path = str(tmp_path / "a.ribo")
    write_raw(path, header, blocks)
    q, t, h = ra.scene_read(path, cfg, seed=11)
    d = ra.SimParams.default(11)
    assert q.epoch == 99 and q.diffusion == 4.0 and q.transport_every == 5 and q.seed == 11
    assert q.dt == d.dt and q.fluid_decay == d.fluid_decay
    assert (t.fluid_substep, t.multigrid, t.fluid_enabled, t.jacobi_iters, t.light_interval) == (4, 0, 1, 128, 10)
    assert not h[ra.REDOX].any() and not h[ra.HEAT_SRC].any()  # no such blocks: the buffers stay as they were
    assert np.array_equal(h[ra.CHEM], g[ra.CHEM])
    with pytest.raises(ra.RiboError):
        ra.scene_read(path, ra.WorldConfig(64, 32, 30, 100))  # more agents in the file than the world holds
    with pytest.raises(ra.RiboError):
        ra.scene_read(path, ra.WorldConfig(128, 64, 30, 1000))  # another world size
    bad = str(tmp_path / "b.ribo")
    open(bad, "wb").write(b"NOTASCENE" + bytes(20))
    with pytest.raises(ra.RiboError):
        ra.scene_read(bad, cfg)
    short = dict(blocks)
    short["pressure"] = short["pressure"][:-4]
    write_raw(bad, header, short)
    with pytest.raises(ra.RiboError):
        ra.scene_read(bad, cfg)


def test_agent_blocks_pass_through(tmp_path):
    """A caller with agents writes the reference's agent blocks and count into
    the scene and reads them back; blocks not given are written empty; only
    the reference's block names are accepted."""
    cfg = ra.WorldConfig(64, 32, 30, 100)
    g = grids(cfg, 8)
    p, s = ra.SimParams.default(3), ra.WorldSettings.default()
    rng = np.random.default_rng(8)
    agents = rng.integers(0, 1 << 32, 40 * 16, dtype=np.uint64).astype(np.uint32).tobytes()  # 64-byte Agents
    slots = np.arange(40, dtype=np.uint32).tobytes()
    genomes = bytes(rng.integers(0, 256, 5000, dtype=np.uint8))
    counters = np.arange(8, dtype=np.uint32).tobytes()
    path = str(tmp_path / "vida.ribo")
    ra.scene_write(path, cfg, p, s, g, agents=40,
                   blocks={"agents": agents, "slots": slots, "genomes": genomes, "life_counters": counters,
                           "pair_mask": b""})
    header, blocks = parse_file(path)
    assert header["agentes"] == 40 and ra.scene_agents(path) == 40
    assert blocks["agents"] == agents and blocks["genomes"] == genomes and blocks["life_counters"] == counters
    assert blocks["slots"] == slots
    for name in set(AGENT_BLOCKS) - {"agents", "genomes"}:
        assert blocks[name] == b"", name
    assert ra.scene_read_block(path, "genomes") == genomes
    assert ra.scene_read_block(path, "agents") == agents
    assert ra.scene_read_block(path, "bodies") == b""
    assert ra.scene_read_block(path, "chem") == g[ra.CHEM].tobytes()
    with pytest.raises(ra.RiboError):
        ra.scene_read_block(path, "no_such_block")
    q, t, h = ra.scene_read(path, cfg, seed=3)  # the world part reads as before
    assert np.array_equal(h[ra.GAMMA], g[ra.GAMMA])
    for bad in ({"chem": b"x"}, {"no_such": b"x"}):
        with pytest.raises(ra.RiboError):
            ra.scene_write(str(tmp_path / "bad.ribo"), cfg, p, s, g, agents=1, blocks=bad)
    with pytest.raises(ra.RiboError):
        ra.scene_write(str(tmp_path / "bad.ribo"), cfg, p, s, g, agents=101)  # more than max_agents


def test_agent_counts_must_match_the_blocks(tmp_path):
    """The writer refuses what the reference's load_scene refuses
    (src/world/snapshot.rs:403-405): a count without matching slot and Agent
    blocks, Agents of the wrong size, or life counters other than 8 words."""
    cfg = ra.WorldConfig(64, 32, 30, 100)
    g = grids(cfg, 9)
    p, s = ra.SimParams.default(3), ra.WorldSettings.default()
    ag = bytes(64 * 3)
    sl = bytes(4 * 3)
    path = str(tmp_path / "x.ribo")
    ra.scene_write(path, cfg, p, s, g, agents=3, blocks={"agents": ag, "slots": sl})  # consistent
    for agents, blocks in ((3, {"agents": ag}),                                # no slots
                           (3, {"slots": sl}),                                 # no Agents
                           (3, {"agents": bytes(68 * 3), "slots": sl}),        # 68-byte Agents
                           (2, {"agents": ag, "slots": sl}),                   # count differs
                           (0, {"life_counters": bytes(28)})):                 # 7 counters
        with pytest.raises(ra.RiboError):
            ra.scene_write(str(tmp_path / "bad.ribo"), cfg, p, s, g, agents=agents, blocks=blocks)


def test_crafted_block_size_is_refused_without_allocating(tmp_path):
    """A block whose prepended size claims far more than its compressed bytes
    can expand to (here 4 GiB - 16 from a few bytes) is a malformed file: an
    error, not an allocation of that size."""
    cfg = ra.WorldConfig(64, 32, 30, 0)
    path = str(tmp_path / "w.ribo")
    ra.scene_write(path, cfg, ra.SimParams.default(1), ra.WorldSettings.default(), grids(cfg, 10))
    raw = bytearray(open(path, "rb").read())
    hl = int.from_bytes(raw[8:12], "little")
    p = 12 + hl
    while p < len(raw):  # find the "slots" block (empty: 4 + a few bytes)
        nl = int.from_bytes(raw[p:p + 2], "little")
        name = raw[p + 2:p + 2 + nl].decode()
        cl = int.from_bytes(raw[p + 2 + nl:p + 10 + nl], "little")
        if name == "slots":
            raw[p + 10 + nl:p + 14 + nl] = (0xFFFFFFF0).to_bytes(4, "little")
            break
        p += 10 + nl + cl
    bad = str(tmp_path / "bad.ribo")
    open(bad, "wb").write(bytes(raw))
    with pytest.raises(ra.RiboError):
        ra.scene_read_block(bad, "slots")
    assert ra.scene_read_block(path, "slots") == b""
