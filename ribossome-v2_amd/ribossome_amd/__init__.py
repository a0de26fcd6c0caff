"""ribossome_amd - Python mirror of the reference World API over libribo_world.so.

Names follow the reference (src/world/mod.rs, src/params.rs): ``WorldConfig``,
``SimParams``, ``WorldSettings``, ``Ledger``, ``World`` with ``encode_steps``,
``ledger_blocking``, ``read_cells_blocking`` and ``read_f32_blocking``.  Every
call goes through the C ABI declared in include/ribo_world.h; there is no
fallback: if the HIP library is missing or no gfx950 device is visible, the
calls raise.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libribo_world.so")

_SIMPARAMS_FIELDS = (
    ('epoch', 'u32'),
    ('seed', 'u32'),
    ('dt', 'f32'),
    ('diffusion', 'f32'),
    ('settle', 'f32'),
    ('fluid_dt', 'f32'),
    ('fluid_decay', 'f32'),
    ('fluid_vorticity', 'f32'),
    ('fluid_viscosity', 'f32'),
    ('fumarole_count', 'u32'),
    ('uv_strength', 'f32'),
    ('uv_depth', 'f32'),
    ('fluid_enabled', 'u32'),
    ('fluid_obstacle_strength', 'f32'),
    ('slope_steer_rate', 'f32'),
    ('cohesion', 'f32'),
    ('spawn_count', 'u32'),
    ('max_agents', 'u32'),
    ('death_probability', 'f32'),
    ('spawn_energy', 'f32'),
    ('food_power', 'f32'),
    ('maintenance_cost', 'f32'),
    ('pairing_rate', 'f32'),
    ('mutation_rate', 'f32'),
    ('uv_damage', 'f32'),
    ('uptake_rate', 'f32'),
    ('brownian', 'f32'),
    ('phoretic_gain', 'f32'),
    ('chain_stiffness', 'f32'),
    ('thermal_kt', 'f32'),
    ('motor_amplitude', 'f32'),
    ('rft_enabled', 'u32'),
    ('require_start', 'u32'),
    ('joint_coupling', 'f32'),
    ('hunger_regulation', 'u32'),
    ('pairing_cost', 'f32'),
    ('reactivation_rate', 'f32'),
    ('swim_gain', 'f32'),
    ('bioturbation', 'f32'),
    ('bioturbation_cost', 'f32'),
    ('monomer_uv_absorb', 'f32'),
    ('sedimentation', 'f32'),
    ('direct_photoactivation', 'f32'),
    ('activation_decay', 'f32'),
    ('monomer_pressure', 'f32'),
    ('agent_fluid_push', 'f32'),
    ('swim_wobble', 'f32'),
    ('fluid_swim_only', 'u32'),
    ('motion_cost', 'f32'),
    ('bond_rate', 'f32'),
    ('bond_energy_share', 'f32'),
    ('bond_signal', 'f32'),
    ('aggregation', 'f32'),
    ('photo_yield', 'f32'),
    ('heat_kill', 'f32'),
    ('flow_coupling', 'f32'),
    ('inertia', 'f32'),
    ('death_energy_cap', 'f32'),
    ('chemo_yield', 'f32'),
    ('redox_decay', 'f32'),
    ('thermal_activation', 'f32'),
    ('denature_temp', 'f32'),
    ('day_period', 'f32'),
    ('sun_slope', 'f32'),
    ('sun_now', 'f32'),
    ('sun_angle', 'f32'),
    ('day_fraction', 'f32'),
    ('metabolic_q10', 'f32'),
    ('metabolic_span', 'f32'),
    ('metabolic_ref', 'f32'),
    ('sun_heat', 'f32'),
    ('sediment_transport', 'f32'),
    ('sediment_threshold', 'f32'),
    ('sediment_settle', 'f32'),
    ('signal_mode', 'f32'),
    ('clock_mute', 'f32'),
    ('paint_x', 'f32'),
    ('paint_y', 'f32'),
    ('paint_radius', 'f32'),
    ('paint_grains', 'f32'),
    ('rest_angle_mult', 'f32'),
    ('protease_power', 'f32'),
    ('lysis_yield', 'f32'),
    ('protease_direct', 'f32'),
    ('joint_load', 'f32'),
    ('swim_grip', 'f32'),
    ('swim_memory', 'f32'),
    ('mouth_leak', 'f32'),
    ('leak_base', 'f32'),
    ('skin_uptake', 'f32'),
    ('death_metab', 'f32'),
    ('salvage', 'f32'),
    ('bond_matter_share', 'f32'),
    ('brownian_rot', 'f32'),
    ('transport_every', 'u32'),
    ('signal_crosstalk', 'f32'),
    ('flow_mass', 'f32'),
    ('chemo_take', 'f32'),
    ('vent_cycle_period', 'f32'),
    ('vent_cycle_amp', 'f32'),
    ('sediment_cohesion', 'f32'),
    ('sediment_compaction', 'f32'),
    ('rock_crumble', 'f32'),
    ('lysis_matter', 'f32'),
    ('scramble_conduction', 'u32'),
    ('scramble_angles', 'u32'),
    ('scramble_response', 'u32'),
    ('scramble_amount', 'f32'),
    ('pairing_energy_half', 'f32'),
    ('pairing_min_energy', 'f32'),
    ('uv_mutation', 'f32'),
    ('heat_mutation', 'f32'),
    ('birth_crowding', 'f32'),
    ('photo_damage', 'f32'),
    ('naked_decay', 'f32'),
    ('store_loss', 'f32'),
)


_CT = {"u32": ctypes.c_uint32, "f32": ctypes.c_float}


class SimParams(ctypes.Structure):
    """SimParams (src/params.rs:86-487): 116 four-byte fields, 464 bytes."""

    _fields_ = [(n, _CT[t]) for n, t in _SIMPARAMS_FIELDS]

    @classmethod
    def default(cls, seed: int = 1) -> "SimParams":
        p = cls()
        lib().ribo_sim_params_default(ctypes.byref(p))
        p.seed = seed
        return p

    def copy(self) -> "SimParams":
        q = SimParams()
        ctypes.memmove(ctypes.byref(q), ctypes.byref(self), ctypes.sizeof(self))
        return q


class WorldSettings(ctypes.Structure):
    """WorldSettings (src/world/mod.rs:91-134) plus the solver extensions."""

    _fields_ = [
        ("fluid_enabled", ctypes.c_uint32),
        ("fluid_substep", ctypes.c_uint32),
        ("multigrid", ctypes.c_uint32),
        ("mg_cycles", ctypes.c_uint32),
        ("jacobi_iters", ctypes.c_uint32),
        ("light_rows_per_step", ctypes.c_uint32),
        ("light_interval", ctypes.c_uint32),
        ("terrain_enabled", ctypes.c_uint32),
        ("contact_enabled", ctypes.c_uint32),
        ("mg_min_size", ctypes.c_uint32),
        ("mg_tol", ctypes.c_float),
        ("mg_max_cycles", ctypes.c_uint32),
        ("precision", ctypes.c_uint32),
        ("ledger_check", ctypes.c_uint32),
        ("no_graphs", ctypes.c_uint32),
        ("light_mode", ctypes.c_uint32),
        ("mg_stationary", ctypes.c_uint32),
    ]

    @classmethod
    def default(cls) -> "WorldSettings":
        s = cls()
        lib().ribo_world_settings_default(ctypes.byref(s))
        return s

    def copy(self) -> "WorldSettings":
        q = WorldSettings()
        ctypes.memmove(ctypes.byref(q), ctypes.byref(self), ctypes.sizeof(self))
        return q


class _CWorldConfig(ctypes.Structure):
    _fields_ = [("grid_size", ctypes.c_uint32), ("fluid_size", ctypes.c_uint32),
                ("world_units_per_cell", ctypes.c_uint32), ("max_agents", ctypes.c_uint32)]


class MgLevel(ctypes.Structure):
    """MgLevel (src/params.rs:727-740)."""

    _fields_ = [("n", ctypes.c_uint32), ("off", ctypes.c_uint32), ("n_c", ctypes.c_uint32),
                ("off_c", ctypes.c_uint32), ("eps", ctypes.c_float), ("_pad0", ctypes.c_uint32),
                ("_pad1", ctypes.c_uint32), ("_pad2", ctypes.c_uint32)]


@dataclass(frozen=True)
class WorldConfig:
    """WorldConfig (src/params.rs:1003-1034)."""

    grid_size: int
    fluid_size: int
    world_units_per_cell: int = 30
    max_agents: int = 0

    def c(self) -> _CWorldConfig:
        return _CWorldConfig(self.grid_size, self.fluid_size, self.world_units_per_cell, self.max_agents)

    def cells(self) -> int:
        return self.grid_size * self.grid_size

    def fluid_cells(self) -> int:
        return self.fluid_size * self.fluid_size

    def sim_size(self) -> float:
        return float(self.grid_size * self.world_units_per_cell)


WorldConfig.DEFAULT = WorldConfig(2048, 1024, 30, 400_000)
WorldConfig.TEST = WorldConfig(256, 128, 30, 4096)

# Buffer ids (RiboBuffer)
CHEM, GAMMA, LIGHT, SLOPE, VELOCITY, PRESSURE, TEMP, REDOX, HEAT_SRC, DIVERGENCE, SOLID_MASK, NET_FLUX, \
    REDOX_EATEN, AGG, FORCE_VECTORS, VELOCITY_SMOOTH = range(16)
BUF_COUNT = 16
_BUF_DTYPE = {CHEM: np.uint32, GAMMA: np.uint32, LIGHT: np.float32, SLOPE: np.float32, VELOCITY: np.float32,
              PRESSURE: np.float32, TEMP: np.float32, REDOX: np.float32, HEAT_SRC: np.float32,
              DIVERGENCE: np.float32, SOLID_MASK: np.uint32, NET_FLUX: np.float32, REDOX_EATEN: np.uint32,
              AGG: np.uint32, FORCE_VECTORS: np.int32, VELOCITY_SMOOTH: np.float32}
BUFFER_NAMES = {CHEM: "chem", GAMMA: "gamma", LIGHT: "light", SLOPE: "slope", VELOCITY: "velocity",
                PRESSURE: "pressure", TEMP: "temp", REDOX: "redox", HEAT_SRC: "heat_src",
                DIVERGENCE: "divergence", SOLID_MASK: "solid_mask", NET_FLUX: "net_flux",
                REDOX_EATEN: "redox_eaten", AGG: "agg", FORCE_VECTORS: "force_vectors",
                VELOCITY_SMOOTH: "velocity_smooth"}
# Exchanges between the grid and the agents (RiboChemOp.kind, chem.wgsl:40-98)
CHEM_SPEND, CHEM_ACTIVATE, CHEM_TAKE_ACT, CHEM_TAKE_SPENT, CHEM_ADD_ACT, CHEM_ADD_SPENT, CHEM_DEPOSIT_ACT, \
    CHEM_DEPOSIT_SPENT = range(8)
FORCE_FP = 256.0  # fixed-point scale of the force words (fluid.wgsl:280)
# FUMAROLE_PIXEL_STRENGTH (src/world/terrain.rs:247), FUMAROLE_GAIN_DEFAULT,
# SEED_DENSITY_DEFAULT and SEED_ACTIVE_DEFAULT (src/world/mod.rs:1980-1984)
FUMAROLE_PIXEL_STRENGTH = 5000.0
FUMAROLE_GAIN_DEFAULT = 0.3
SEED_DENSITY_DEFAULT = 0.83
SEED_ACTIVE_DEFAULT = 1.0


def force_words(f: np.ndarray) -> np.ndarray:
    """atomic_add_force's quantisation (fluid.wgsl:281-284) of forces f32[2F]
    into the words of FORCE_VECTORS (one agent's push per cell component)."""
    f32 = np.asarray(f, dtype=np.float32)
    return np.rint(np.clip(f32, np.float32(-1.0e5), np.float32(1.0e5)) * np.float32(FORCE_FP)).astype(np.int32)
MAX_STEPS_PER_FRAME = 64


class _CSolverResiduals(ctypes.Structure):
    _fields_ = [("last", ctypes.c_double), ("last_rounded", ctypes.c_double), ("max", ctypes.c_double),
                ("max_rounded", ctypes.c_double), ("restarts", ctypes.c_uint64)]


class RiboError(RuntimeError):
    pass


_lib = None


def lib() -> ctypes.CDLL:
    """Loads libribo_world.so (built in-tree by ``make -C ribossome-v2_amd``)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RiboError(f"{LIB_PATH} is missing: build it with `make -C ribossome-v2_amd`")
    L = ctypes.CDLL(LIB_PATH)
    P, U32, I, SZ = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int, ctypes.c_size_t
    sig = {
        "ribo_sim_params_default": (None, [P]),
        "ribo_world_settings_default": (None, [P]),
        "ribo_buffer_bytes": (SZ, [P, I]),
        "ribo_world_create": (I, [P, U32, ctypes.POINTER(P)]),
        "ribo_world_destroy": (I, [P]),
        "ribo_world_set_params": (I, [P, P]),
        "ribo_world_get_params": (I, [P, P]),
        "ribo_world_set_settings": (I, [P, P]),
        "ribo_world_get_settings": (I, [P, P]),
        "ribo_world_mg_levels": (I, [P, P, U32, ctypes.POINTER(U32)]),
        "ribo_world_upload": (I, [P, I, P, SZ]),
        "ribo_world_download": (I, [P, I, P, SZ]),
        "ribo_world_encode_steps": (I, [P, U32]),
        "ribo_world_run": (I, [P, U32]),
        "ribo_world_sync": (I, [P]),
        "ribo_world_idle": (I, [P]),
        "ribo_world_invalidate_light": (I, [P]),
        "ribo_world_ledger": (I, [P, P]),
        "ribo_world_chem_ops": (I, [P, P, U32, P]),
        "ribo_world_ledger_status": (I, [P, ctypes.POINTER(U32)]),
        "ribo_world_set_timing": (I, [P, I]),
        "ribo_world_set_pipeline": (I, [P, I]),
        "ribo_world_kernel_times": (I, [P, P, P, P, U32, ctypes.POINTER(U32)]),
        "ribo_world_kernel_launch_ms": (I, [P, ctypes.c_char_p, P, U32, ctypes.POINTER(U32)]),
        "ribo_world_reset_kernel_times": (I, [P]),
        "ribo_world_mg_stats": (I, [P, ctypes.POINTER(U32), ctypes.POINTER(ctypes.c_double)]),
        "ribo_world_solver_stats": (I, [P, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64),
                                        ctypes.POINTER(ctypes.c_uint64)]),
        "ribo_world_solver_residuals": (I, [P, P, I]),
        "ribo_ledger_noalloc": (I, [P, ctypes.c_uint64, P, P]),
        "ribo_rasterize_fumaroles": (I, [P, P, U32, P]),
        "ribo_vent_heat_src": (I, [P, P, P, ctypes.c_float, P]),
        "ribo_world_set_vents": (I, [P, P, U32, ctypes.c_float, P, P]),
        "ribo_seed_cells": (I, [P, ctypes.c_uint64, ctypes.c_float, ctypes.c_float, P]),
        "ribo_fit_matter_to_terrain": (I, [P, P, ctypes.c_uint64, ctypes.c_uint64]),
        "ribo_refit_matter_to_terrain": (I, [P, P, ctypes.c_uint64]),
        "ribo_world_restart_matter": (I, [P, ctypes.c_uint64, ctypes.c_float, ctypes.c_float, P]),
        "ribo_world_apply_terrain_live": (I, [P, P]),
        "ribo_last_error": (ctypes.c_char_p, []),
        "ribo_abi_version": (U32, []),
        "ribo_gen_gamma": (I, [P, ctypes.c_uint64, ctypes.c_float, I, P]),
        "ribo_gen_velocity": (I, [P, ctypes.c_uint64, U32, ctypes.c_float, P, P]),
        "ribo_gen_chem": (I, [P, ctypes.c_uint64, ctypes.c_float, P, P]),
        "ribo_gen_vent_heat": (I, [P, U32, ctypes.c_float, P]),
        "ribo_world_footprint": (SZ, [P, P]),
        "ribo_world_get_config": (I, [P, P]),
        "ribo_world_autotune": (I, [P, I]),
        "ribo_world_mg_tiles": (I, [P, P, U32, ctypes.POINTER(U32)]),
        "ribo_scene_write": (I, [ctypes.c_char_p, P, P, P, P]),
        "ribo_scene_read": (I, [ctypes.c_char_p, P, P, P, P]),
        "ribo_scene_write_agents": (I, [ctypes.c_char_p, P, P, P, P, U32, P, U32]),
        "ribo_scene_read_block": (I, [ctypes.c_char_p, ctypes.c_char_p, P, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)]),
        "ribo_scene_agents": (I, [ctypes.c_char_p, ctypes.POINTER(U32)]),
        "ribo_world_save_scene": (I, [P, ctypes.c_char_p]),
        "ribo_world_load_scene": (I, [P, ctypes.c_char_p]),
        "ribo_lz4_bound": (SZ, [SZ]),
        "ribo_lz4_compress": (ctypes.c_int64, [P, SZ, P, SZ]),
        "ribo_lz4_decompress": (ctypes.c_int64, [P, SZ, P, SZ]),
        "ribo_world_device_bytes": (I, [P, ctypes.POINTER(SZ)]),
        "ribo_batch_create": (I, [ctypes.POINTER(P), U32, ctypes.POINTER(P)]),
        "ribo_batch_destroy": (I, [P]),
        "ribo_batch_encode_steps": (I, [P, U32]),
        "ribo_batch_run": (I, [P, U32]),
        "ribo_batch_sync": (I, [P]),
        "ribo_batch_idle": (I, [P]),
        "ribo_batch_launches": (I, [P, ctypes.POINTER(ctypes.c_uint64)]),
        "ribo_batch_accepts": (I, [P]),
        "ribo_server_set_batching": (I, [P, I]),
        "ribo_server_create": (I, [ctypes.c_uint64, U32, U32, ctypes.POINTER(P)]),
        "ribo_server_destroy": (I, [P]),
        "ribo_server_submit": (I, [P, P, ctypes.POINTER(ctypes.c_uint64)]),
        "ribo_server_status": (I, [P, ctypes.c_uint64, P]),
        "ribo_server_download": (I, [P, ctypes.c_uint64, I, P, SZ]),
        "ribo_server_release": (I, [P, ctypes.c_uint64]),
        "ribo_server_error": (I, [P, ctypes.c_uint64, ctypes.c_char_p, SZ]),
        "ribo_server_stats": (I, [P, ctypes.POINTER(U32), ctypes.POINTER(U32), ctypes.POINTER(U32),
                                  ctypes.POINTER(ctypes.c_uint64)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _check(rc: int, what: str) -> int:
    if rc < 0:
        raise RiboError(f"{what} failed ({rc}): {lib().ribo_last_error().decode()}")
    return rc


def _ptr(a: np.ndarray) -> ctypes.c_void_p:
    return ctypes.c_void_p(a.ctypes.data)


def buffer_bytes(cfg: WorldConfig, which: int) -> int:
    c = cfg.c()
    return int(lib().ribo_buffer_bytes(ctypes.byref(c), which))


def buffer_shape(cfg: WorldConfig, which: int) -> tuple:
    n = buffer_bytes(cfg, which) // 4
    return (n,)


@dataclass
class Ledger:
    """Ledger (src/world/mod.rs:39-82): exact matter per channel and state."""

    act: tuple
    spent: tuple
    held: tuple = (0, 0, 0, 0)

    @classmethod
    def from_cells(cls, cells: np.ndarray) -> "Ledger":
        c = np.asarray(cells, dtype=np.uint32).reshape(-1, 4)
        act = tuple(int((c[:, k] & 0xFFFF).astype(np.uint64).sum() & 0xFFFFFFFF) for k in range(4))
        spent = tuple(int((c[:, k] >> 16).astype(np.uint64).sum() & 0xFFFFFFFF) for k in range(4))
        return cls(act, spent, (0, 0, 0, 0))

    @classmethod
    def from_words(cls, w) -> "Ledger":
        w = [int(x) for x in w]
        return cls(tuple(w[0:4]), tuple(w[4:8]), tuple(w[8:12]))

    def total(self) -> int:
        return sum(self.act) + sum(self.spent) + sum(self.held)

    def channel(self, ch: int) -> int:
        return self.act[ch] + self.spent[ch] + self.held[ch]

    def channels(self) -> tuple:
        return tuple(self.channel(c) for c in range(4))


class World:
    """World (src/world/mod.rs): one world resident on the current HIP device."""

    def __init__(self, cfg: WorldConfig, seed: int = 1):
        self.cfg = cfg
        self._h = ctypes.c_void_p()
        c = cfg.c()
        _check(lib().ribo_world_create(ctypes.byref(c), seed, ctypes.byref(self._h)), "ribo_world_create")

    def close(self):
        if self._h:
            lib().ribo_world_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- params / settings (pub fields of World) --
    @property
    def params(self) -> SimParams:
        p = SimParams()
        _check(lib().ribo_world_get_params(self._h, ctypes.byref(p)), "get_params")
        return p

    @params.setter
    def params(self, p: SimParams):
        _check(lib().ribo_world_set_params(self._h, ctypes.byref(p)), "set_params")

    def set_param(self, **kw):
        p = self.params
        for k, v in kw.items():
            setattr(p, k, v)
        self.params = p

    @property
    def settings(self) -> WorldSettings:
        s = WorldSettings()
        _check(lib().ribo_world_get_settings(self._h, ctypes.byref(s)), "get_settings")
        return s

    @settings.setter
    def settings(self, s: WorldSettings):
        _check(lib().ribo_world_set_settings(self._h, ctypes.byref(s)), "set_settings")

    def set_setting(self, **kw):
        s = self.settings
        for k, v in kw.items():
            setattr(s, k, v)
        self.settings = s

    def mg_levels(self) -> list:
        arr = (MgLevel * 32)()
        n = ctypes.c_uint32()
        _check(lib().ribo_world_mg_levels(self._h, arr, 32, ctypes.byref(n)), "mg_levels")
        return [(arr[i].n, arr[i].off, arr[i].n_c, arr[i].off_c, arr[i].eps) for i in range(n.value)]

    # -- data movement --
    def upload(self, which: int, data: np.ndarray):
        a = np.ascontiguousarray(data, dtype=_BUF_DTYPE[which])
        _check(lib().ribo_world_upload(self._h, which, _ptr(a), a.nbytes), f"upload {BUFFER_NAMES[which]}")

    def download(self, which: int) -> np.ndarray:
        a = np.empty(buffer_shape(self.cfg, which), dtype=_BUF_DTYPE[which])
        _check(lib().ribo_world_download(self._h, which, _ptr(a), a.nbytes), f"download {BUFFER_NAMES[which]}")
        return a

    # -- stepping --
    def encode_steps(self, steps: int) -> int:
        return _check(lib().ribo_world_encode_steps(self._h, steps), "encode_steps")

    def run(self, steps: int):
        _check(lib().ribo_world_run(self._h, steps), "run")

    def sync(self):
        _check(lib().ribo_world_sync(self._h), "sync")

    def idle(self) -> bool:
        """True when every recorded step has finished (does not wait)."""
        return _check(lib().ribo_world_idle(self._h), "idle") == 1

    def invalidate_light(self):
        """World::invalidate_light: the next step runs the full light sweep."""
        _check(lib().ribo_world_invalidate_light(self._h), "invalidate_light")

    def autotune(self, force: int = 0) -> list:
        """Measures (force 0) or sets (16, 32) the multigrid tile sides; returns them per level."""
        _check(lib().ribo_world_autotune(self._h, int(force)), "autotune")
        return self.mg_tiles()

    def mg_tiles(self) -> list:
        out = (ctypes.c_uint8 * 32)()
        n = ctypes.c_uint32()
        _check(lib().ribo_world_mg_tiles(self._h, out, 32, ctypes.byref(n)), "mg_tiles")
        return list(out[:n.value])

    def save_scene(self, path: str):
        """World::save_scene: the grids, parameters and settings as a .ribo file."""
        _check(lib().ribo_world_save_scene(self._h, os.fsencode(path)), "save_scene")

    def load_scene(self, path: str):
        """Scene::read + World::load_scene into this world (same configuration)."""
        _check(lib().ribo_world_load_scene(self._h, os.fsencode(path)), "load_scene")

    def device_bytes(self) -> int:
        """Device bytes this world holds now."""
        n = ctypes.c_size_t()
        _check(lib().ribo_world_device_bytes(self._h, ctypes.byref(n)), "device_bytes")
        return int(n.value)

    # -- readback helpers named like the reference's test API --
    def read_cells_blocking(self) -> np.ndarray:
        return self.download(CHEM)

    def read_f32_blocking(self, which: int) -> np.ndarray:
        return self.download(which)

    def chem_ops(self, ops) -> np.ndarray:
        """The agents' exchanges with the grid: rows (cell, channel, kind,
        count), applied at once; returns the monomers each moved."""
        a = np.ascontiguousarray(ops, dtype=np.uint32).reshape(-1, 4)
        done = np.zeros(len(a), dtype=np.uint32)
        _check(lib().ribo_world_chem_ops(self._h, _ptr(a), len(a), _ptr(done)), "chem_ops")
        return done

    def ledger_blocking(self) -> Ledger:
        out = (ctypes.c_uint32 * 12)()
        _check(lib().ribo_world_ledger(self._h, out), "ledger")
        return Ledger.from_words(out)

    def ledger_status(self) -> int | None:
        e = ctypes.c_uint32()
        rc = lib().ribo_world_ledger_status(self._h, ctypes.byref(e))
        return None if rc == 0 else int(e.value)

    # -- measurement --
    def set_timing(self, on: bool):
        _check(lib().ribo_world_set_timing(self._h, 1 if on else 0), "set_timing")

    def kernel_times(self) -> dict:
        cap = 64
        names = (ctypes.c_char_p * cap)()
        ms = (ctypes.c_double * cap)()
        cnt = (ctypes.c_uint64 * cap)()
        n = ctypes.c_uint32()
        _check(lib().ribo_world_kernel_times(self._h, names, ms, cnt, cap, ctypes.byref(n)), "kernel_times")
        return {names[i].decode(): (ms[i], int(cnt[i])) for i in range(min(n.value, cap))}

    def set_pipeline(self, enable: bool):
        """Overlap each step's transport with the next step's fluid where the
        step allows it (on by default): a second stream in the reference
        schedule, the scatter inside the next solve's V-cycles in the
        tolerance mode."""
        _check(lib().ribo_world_set_pipeline(self._h, 1 if enable else 0), "set_pipeline")

    def kernel_launch_ms(self, name: str) -> np.ndarray:
        """Every launch's milliseconds of kernel `name` since the last reset."""
        n = ctypes.c_uint32()
        _check(lib().ribo_world_kernel_launch_ms(self._h, name.encode(), None, 0, ctypes.byref(n)), "kernel_launch_ms")
        out = np.zeros(n.value, np.float32)
        if n.value:
            _check(lib().ribo_world_kernel_launch_ms(self._h, name.encode(), out.ctypes.data, n.value, ctypes.byref(n)),
                   "kernel_launch_ms")
        return out

    def reset_kernel_times(self):
        _check(lib().ribo_world_reset_kernel_times(self._h), "reset_kernel_times")

    def solver_stats(self) -> tuple:
        a, b, c = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        _check(lib().ribo_world_solver_stats(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)), "solver_stats")
        return int(a.value), int(b.value), int(c.value)

    def solver_residuals(self, reset: bool = False) -> dict:
        """Tolerance-mode residuals: the refined iterate's (binary64) and the
        stored binary32 pressure's, last solve and maxima since the last
        reset, and the restarts from the true residual."""
        r = _CSolverResiduals()
        _check(lib().ribo_world_solver_residuals(self._h, ctypes.byref(r), 1 if reset else 0), "solver_residuals")
        return {"last": r.last, "last_rounded": r.last_rounded, "max": r.max, "max_rounded": r.max_rounded,
                "restarts": int(r.restarts)}

    def set_vents(self, fumaroles, gain: float = FUMAROLE_GAIN_DEFAULT, heat_image=None, chem_image=None):
        """World::upload_heat: fumaroles and images -> RIBO_BUF_HEAT_SRC."""
        arr, n = _fumaroles(fumaroles)
        hi = None if heat_image is None else np.ascontiguousarray(heat_image, dtype=np.float32)
        ci = None if chem_image is None else np.ascontiguousarray(chem_image, dtype=np.float32)
        _check(lib().ribo_world_set_vents(self._h, arr, n, gain, None if hi is None else _ptr(hi),
                                          None if ci is None else _ptr(ci)), "set_vents")

    def restart_matter(self, seed: int, density: float = SEED_DENSITY_DEFAULT,
                       active: float = SEED_ACTIVE_DEFAULT) -> "Ledger":
        out = (ctypes.c_uint32 * 12)()
        _check(lib().ribo_world_restart_matter(self._h, seed, density, active, out), "restart_matter")
        return Ledger.from_words(out)

    def apply_terrain_live(self, gamma: np.ndarray):
        g = np.ascontiguousarray(gamma, dtype=np.uint32)
        _check(lib().ribo_world_apply_terrain_live(self._h, _ptr(g)), "apply_terrain_live")

    def mg_stats(self) -> tuple:
        c = ctypes.c_uint32()
        r = ctypes.c_double()
        _check(lib().ribo_world_mg_stats(self._h, ctypes.byref(c), ctypes.byref(r)), "mg_stats")
        return int(c.value), float(r.value)


def ledger_noalloc(chem_device_ptr: int, cells: int, stream: int = 0) -> Ledger:
    """Ledger of a grid in the caller's device memory (`cells` cells of 4
    u32 words at chem_device_ptr), computed without allocating on `stream`."""
    out = (ctypes.c_uint32 * 12)()
    _check(lib().ribo_ledger_noalloc(ctypes.c_void_p(chem_device_ptr), cells, ctypes.c_void_p(stream), out),
           "ledger_noalloc")
    return Ledger.from_words(out)


# ---- vents and matter seeding (the reference's host code, in the library) ----
class Fumarole(ctypes.Structure):
    """Fumarole (src/params.rs:795-810), 32 bytes."""

    _fields_ = [("x_frac", ctypes.c_float), ("y_frac", ctypes.c_float), ("strength", ctypes.c_float),
                ("spread", ctypes.c_float), ("enabled", ctypes.c_uint32), ("_pad0", ctypes.c_uint32),
                ("_pad1", ctypes.c_uint32), ("_pad2", ctypes.c_uint32)]




def _fumaroles(fs):
    arr = (Fumarole * max(len(fs), 1))()
    for i, f in enumerate(fs):
        arr[i] = f if isinstance(f, Fumarole) else Fumarole(*f[:4], f[4] if len(f) > 4 else 1, 0, 0, 0)
    return arr, len(fs)


def rasterize_fumaroles(cfg: WorldConfig, fumaroles) -> np.ndarray:
    """rasterize_fumaroles: heat f32[G]; fumaroles as Fumarole or (x_frac,
    y_frac, strength, spread[, enabled]) tuples."""
    arr, n = _fumaroles(fumaroles)
    out = np.empty(cfg.cells(), np.float32)
    c = cfg.c()
    _check(lib().ribo_rasterize_fumaroles(ctypes.byref(c), arr, n, _ptr(out)), "rasterize_fumaroles")
    return out


def vent_heat_src(cfg: WorldConfig, heat: np.ndarray, chem: np.ndarray, gain: float) -> np.ndarray:
    h = np.ascontiguousarray(heat, dtype=np.float32)
    ch = np.ascontiguousarray(chem, dtype=np.float32)
    out = np.empty(2 * cfg.fluid_cells(), np.float32)
    c = cfg.c()
    _check(lib().ribo_vent_heat_src(ctypes.byref(c), _ptr(h), _ptr(ch), gain, _ptr(out)), "vent_heat_src")
    return out


def seed_cells(cfg: WorldConfig, seed: int, density: float = SEED_DENSITY_DEFAULT,
               active: float = SEED_ACTIVE_DEFAULT) -> np.ndarray:
    out = np.empty(4 * cfg.cells(), np.uint32)
    c = cfg.c()
    _check(lib().ribo_seed_cells(ctypes.byref(c), seed, density, active, _ptr(out)), "seed_cells")
    return out


def fit_matter_to_terrain(cells: np.ndarray, gamma: np.ndarray, seed: int) -> np.ndarray:
    out = np.array(cells, dtype=np.uint32, copy=True)
    g = np.ascontiguousarray(gamma, dtype=np.uint32)
    _check(lib().ribo_fit_matter_to_terrain(_ptr(out), _ptr(g), g.size, seed), "fit_matter_to_terrain")
    return out


def refit_matter_to_terrain(cells: np.ndarray, gamma: np.ndarray) -> np.ndarray:
    out = np.array(cells, dtype=np.uint32, copy=True)
    g = np.ascontiguousarray(gamma, dtype=np.uint32)
    _check(lib().ribo_refit_matter_to_terrain(_ptr(out), _ptr(g), g.size), "refit_matter_to_terrain")
    return out


# ---- synthetic inputs (host generators in the library) ----
def gen_gamma(cfg: WorldConfig, seed: int, rock_frac: float, rubble: bool) -> np.ndarray:
    out = np.empty(cfg.cells(), dtype=np.uint32)
    c = cfg.c()
    _check(lib().ribo_gen_gamma(ctypes.byref(c), seed, rock_frac, 1 if rubble else 0, _ptr(out)), "gen_gamma")
    return out


def gen_velocity(cfg: WorldConfig, seed: int, vmax: float, gamma: np.ndarray | None, modes: int = 16) -> np.ndarray:
    out = np.empty(2 * cfg.fluid_cells(), dtype=np.float32)
    c = cfg.c()
    g = None if gamma is None else np.ascontiguousarray(gamma, dtype=np.uint32)
    _check(lib().ribo_gen_velocity(ctypes.byref(c), seed, modes, vmax, None if g is None else _ptr(g), _ptr(out)),
           "gen_velocity")
    return out


def gen_chem(cfg: WorldConfig, seed: int, lam: float, gamma: np.ndarray | None) -> np.ndarray:
    out = np.empty(4 * cfg.cells(), dtype=np.uint32)
    c = cfg.c()
    g = None if gamma is None else np.ascontiguousarray(gamma, dtype=np.uint32)
    _check(lib().ribo_gen_chem(ctypes.byref(c), seed, lam, None if g is None else _ptr(g), _ptr(out)), "gen_chem")
    return out


def gen_vent_heat(cfg: WorldConfig, rows: int, h: float) -> np.ndarray:
    out = np.empty(2 * cfg.fluid_cells(), dtype=np.float32)
    c = cfg.c()
    _check(lib().ribo_gen_vent_heat(ctypes.byref(c), rows, h, _ptr(out)), "gen_vent_heat")
    return out


def device_available() -> bool:
    """True when a gfx950 device is visible (without creating a world)."""
    try:
        import subprocess  # noqa: F401
        h = ctypes.c_void_p()
        c = WorldConfig(16, 8).c()
        rc = lib().ribo_world_create(ctypes.byref(c), 1, ctypes.byref(h))
        if rc == 0:
            lib().ribo_world_destroy(h)
            return True
        return False
    except Exception:
        return False


def world_footprint(cfg: WorldConfig, settings: "WorldSettings | None" = None) -> int:
    """Device bytes a world of this configuration holds (ribo_world_footprint)."""
    c = cfg.c()
    return int(lib().ribo_world_footprint(ctypes.byref(c), None if settings is None else ctypes.byref(settings)))


class _CScene(ctypes.Structure):
    _fields_ = [("cfg", _CWorldConfig), ("seed", ctypes.c_uint32), ("steps", ctypes.c_uint32),
                ("params", SimParams), ("settings", WorldSettings),
                ("input", ctypes.c_void_p * BUF_COUNT), ("input_bytes", ctypes.c_size_t * BUF_COUNT)]


class SceneStatus(ctypes.Structure):
    _fields_ = [("state", ctypes.c_int32), ("error", ctypes.c_int32), ("steps_done", ctypes.c_uint32),
                ("_pad", ctypes.c_uint32), ("t_submit", ctypes.c_double), ("t_admit", ctypes.c_double),
                ("t_done", ctypes.c_double), ("footprint", ctypes.c_uint64)]


class Batch:
    """Worlds stepped together with one launch per kernel (ribo_batch_*).
    The worlds stay the caller's: they are read with World.download after
    the batch's steps, and must not be closed while the batch exists."""

    def __init__(self, worlds):
        self._worlds = list(worlds)
        arr = (ctypes.c_void_p * len(self._worlds))(*[w._h.value for w in self._worlds])
        self._h = ctypes.c_void_p()
        _check(lib().ribo_batch_create(arr, len(self._worlds), ctypes.byref(self._h)), "batch_create")

    def close(self):
        if self._h:
            lib().ribo_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def encode_steps(self, steps: int) -> int:
        return _check(lib().ribo_batch_encode_steps(self._h, steps), "batch_encode_steps")

    def run(self, steps: int):
        _check(lib().ribo_batch_run(self._h, steps), "batch_run")

    def sync(self):
        _check(lib().ribo_batch_sync(self._h), "batch_sync")

    def idle(self) -> bool:
        return _check(lib().ribo_batch_idle(self._h), "batch_idle") == 1

    def launches(self) -> int:
        n = ctypes.c_uint64()
        _check(lib().ribo_batch_launches(self._h, ctypes.byref(n)), "batch_launches")
        return n.value

    @staticmethod
    def accepts(world: "World") -> bool:
        return _check(lib().ribo_batch_accepts(world._h), "batch_accepts") == 1


SCENE_QUEUED, SCENE_RUNNING, SCENE_DONE, SCENE_FAILED = range(4)


class Server:
    """The serving form (ribo_server_*): scenes admitted by an HBM budget,
    resident in HBM; with batching (the default) the scenes that fit a
    batch are stepped together, one launch per kernel, the others on their
    own streams by the library's scheduler threads."""

    def __init__(self, hbm_budget: int, max_running: int = 0, threads: int = 0, batching: bool = True):
        self._h = ctypes.c_void_p()
        _check(lib().ribo_server_create(int(hbm_budget), int(max_running), int(threads), ctypes.byref(self._h)),
               "server_create")
        _check(lib().ribo_server_set_batching(self._h, 1 if batching else 0), "server_set_batching")

    def close(self):
        if self._h:
            lib().ribo_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, cfg: WorldConfig, seed: int, steps: int, params: SimParams, settings: WorldSettings,
               inputs: dict) -> int:
        sc = _CScene()
        sc.cfg = cfg.c()
        sc.seed = seed
        sc.steps = steps
        sc.params = params
        sc.settings = settings
        keep = []
        for which, data in inputs.items():
            a = np.ascontiguousarray(data, dtype=_BUF_DTYPE[which])
            keep.append(a)
            sc.input[which] = a.ctypes.data
            sc.input_bytes[which] = a.nbytes
        t = ctypes.c_uint64()
        _check(lib().ribo_server_submit(self._h, ctypes.byref(sc), ctypes.byref(t)), "server_submit")
        return int(t.value)

    def status(self, ticket: int) -> SceneStatus:
        st = SceneStatus()
        _check(lib().ribo_server_status(self._h, ticket, ctypes.byref(st)), "server_status")
        return st

    def download(self, ticket: int, cfg: WorldConfig, which: int) -> np.ndarray:
        n = buffer_bytes(cfg, which)
        a = np.empty(n // 4, dtype=_BUF_DTYPE[which])
        _check(lib().ribo_server_download(self._h, ticket, which, _ptr(a), n), "server_download")
        return a

    def release(self, ticket: int):
        _check(lib().ribo_server_release(self._h, ticket), "server_release")

    def error(self, ticket: int) -> str:
        """The library's message for a failed scene."""
        buf = ctypes.create_string_buffer(512)
        _check(lib().ribo_server_error(self._h, ticket, buf, 512), "server_error")
        return buf.value.decode()

    def stats(self) -> tuple:
        q, r, f, u = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint64()
        _check(lib().ribo_server_stats(self._h, ctypes.byref(q), ctypes.byref(r), ctypes.byref(f), ctypes.byref(u)),
               "server_stats")
        return int(q.value), int(r.value), int(f.value), int(u.value)


class _CSceneGrids(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in ("chem", "gamma", "velocity", "pressure", "temperature", "redox",
                                                "heat_src")]


SCENE_GRIDS = (("chem", CHEM), ("gamma", GAMMA), ("velocity", VELOCITY), ("pressure", PRESSURE),
               ("temperature", TEMP), ("redox", REDOX), ("heat_src", HEAT_SRC))


class _CSceneBlock(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char_p), ("data", ctypes.c_void_p), ("bytes", ctypes.c_uint64)]


def scene_write(path: str, cfg: WorldConfig, params: SimParams, settings: WorldSettings, grids: dict,
                agents: int = 0, blocks: "dict | None" = None):
    """ribo_scene_write (ribo_scene_write_agents with agents or blocks): grids
    maps CHEM, GAMMA, VELOCITY, PRESSURE, TEMP, REDOX (and HEAT_SRC) to host
    arrays; blocks maps the reference's agent block names to their bytes."""
    g = _CSceneGrids()
    keep = []
    for name, which in SCENE_GRIDS:
        if which == HEAT_SRC and which not in grids:
            continue  # optional block
        a = np.ascontiguousarray(grids[which], dtype=_BUF_DTYPE[which])
        if a.nbytes != buffer_bytes(cfg, which):
            raise RiboError(f"scene_write: {name} has {a.nbytes} bytes")
        keep.append(a)
        setattr(g, name, a.ctypes.data)
    c = cfg.c()
    if not agents and not blocks:
        _check(lib().ribo_scene_write(os.fsencode(path), ctypes.byref(c), ctypes.byref(params), ctypes.byref(settings),
                                      ctypes.byref(g)), "scene_write")
        return
    items = list((blocks or {}).items())
    arr = (_CSceneBlock * max(len(items), 1))()
    for i, (name, data) in enumerate(items):
        b = np.frombuffer(bytes(data), dtype=np.uint8).copy()
        keep.append(b)
        arr[i].name = name.encode()
        arr[i].data = b.ctypes.data if b.size else None
        arr[i].bytes = b.size
    _check(lib().ribo_scene_write_agents(os.fsencode(path), ctypes.byref(c), ctypes.byref(params),
                                         ctypes.byref(settings), ctypes.byref(g), int(agents), arr, len(items)),
           "scene_write_agents")


def scene_read_block(path: str, name: str) -> bytes:
    """ribo_scene_read_block: one block of a scene file, decompressed."""
    n = ctypes.c_uint64()
    _check(lib().ribo_scene_read_block(os.fsencode(path), name.encode(), None, 0, ctypes.byref(n)),
           f"scene_read_block {name}")
    if n.value == 0:
        return b""
    buf = np.empty(int(n.value), dtype=np.uint8)
    _check(lib().ribo_scene_read_block(os.fsencode(path), name.encode(), buf.ctypes.data, int(n.value), ctypes.byref(n)),
           f"scene_read_block {name}")
    return buf[:n.value].tobytes()


def scene_agents(path: str) -> int:
    """ribo_scene_agents: the header's agent count."""
    n = ctypes.c_uint32()
    _check(lib().ribo_scene_agents(os.fsencode(path), ctypes.byref(n)), "scene_agents")
    return int(n.value)


def scene_read(path: str, cfg: WorldConfig, seed: int = 1, settings: "WorldSettings | None" = None):
    """ribo_scene_read: (params, settings, grids) of a .ribo file for a world of configuration cfg."""
    p = SimParams.default(seed)
    s = settings if settings is not None else WorldSettings.default()
    g = _CSceneGrids()
    out = {}
    for name, which in SCENE_GRIDS:
        a = np.zeros(buffer_bytes(cfg, which) // 4, dtype=_BUF_DTYPE[which])
        out[which] = a
        setattr(g, name, a.ctypes.data)
    c = cfg.c()
    _check(lib().ribo_scene_read(os.fsencode(path), ctypes.byref(c), ctypes.byref(p), ctypes.byref(s), ctypes.byref(g)),
           "scene_read")
    return p, s, out


def lz4_compress(data: bytes) -> bytes:
    src = np.frombuffer(bytes(data), dtype=np.uint8)
    cap = int(lib().ribo_lz4_bound(len(src)))
    dst = np.empty(max(cap, 1), dtype=np.uint8)
    n = _check(int(lib().ribo_lz4_compress(_ptr(src) if len(src) else None, len(src), _ptr(dst), cap)), "lz4_compress")
    return dst[:n].tobytes()


def lz4_decompress(block: bytes, size: int) -> bytes:
    src = np.frombuffer(bytes(block), dtype=np.uint8)
    dst = np.empty(max(size, 1), dtype=np.uint8)
    n = int(lib().ribo_lz4_decompress(_ptr(src) if len(src) else None, len(src), _ptr(dst), size))
    if n < 0:
        raise RiboError("lz4_decompress: malformed block")
    return dst[:n].tobytes()
