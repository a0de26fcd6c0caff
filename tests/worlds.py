"""Shared world builders for the tests: the same seeded inputs go to the HIP
world (ribossome_amd.World) and to the CPU oracle (oracle_ctypes.OracleWorld)."""
from __future__ import annotations

import numpy as np

import ribossome_amd as ra

FIELDS = (ra.VELOCITY, ra.PRESSURE, ra.TEMP, ra.REDOX, ra.DIVERGENCE, ra.NET_FLUX)


def parity_params(seed: int = 1, **over) -> ra.SimParams:
    """Reference defaults with the parity profile: substep 1, transport every
    step, no UV light, no solar heating (light is a frozen input)."""
    p = ra.SimParams.default(seed)
    p.transport_every = 1
    p.uv_strength = 0.0
    p.sun_heat = 0.0
    for k, v in over.items():
        setattr(p, k, v)
    return p


def parity_settings(**over) -> ra.WorldSettings:
    """The parity profile: reference defaults with the fluid every step, a
    frozen terrain and a frozen (uploaded) light field."""
    s = ra.WorldSettings.default()
    s.fluid_substep = 1
    s.terrain_enabled = 0
    s.light_mode = 0
    for k, v in over.items():
        setattr(s, k, v)
    return s


def make_inputs(cfg: ra.WorldConfig, seed: int, rock=0.10, rubble=False, lam=4.0, vent_rows=0, vent_h=1500.0,
                vcells_per_step=0.5):
    gamma = ra.gen_gamma(cfg, seed, rock, rubble)
    vel = ra.gen_velocity(cfg, seed, vcells_per_step / 0.017, gamma)
    chem = ra.gen_chem(cfg, seed, lam, gamma)
    heat = ra.gen_vent_heat(cfg, vent_rows, vent_h)
    return {ra.GAMMA: gamma, ra.VELOCITY: vel, ra.CHEM: chem, ra.HEAT_SRC: heat}


def load(world, inputs: dict, params: ra.SimParams, settings: ra.WorldSettings):
    """Order matters: gamma first (derived terrain), then the rest."""
    world.params = params
    if hasattr(world, "set_settings"):
        world.set_settings(settings)
    else:
        world.settings = settings
    world.upload(ra.GAMMA, inputs[ra.GAMMA])
    for k, v in inputs.items():
        if k != ra.GAMMA:
            world.upload(k, v)


def bit_diff(a: np.ndarray, b: np.ndarray) -> int:
    """Number of elements whose bit patterns differ (NaN-safe)."""
    return int(np.count_nonzero(a.view(np.uint32) != b.view(np.uint32)))
