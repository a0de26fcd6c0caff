/*
 * ribo_params_default.h - SimParams::default() (src/params.rs:553-690) as a
 * header-only initialiser, shared by the library and by host tools that do
 * not link it.  Values are the reference's, field by field.
 */
#ifndef RIBO_PARAMS_DEFAULT_H
#define RIBO_PARAMS_DEFAULT_H

#include <string.h>

#include "ribo_world.h"

static inline void ribo_sim_params_fill_default(RiboSimParams* p) {
    memset(p, 0, sizeof(*p));
    p->epoch = 0u;
    p->seed = 1u;
    p->dt = 0.017f;
    p->diffusion = 12.0f;
    p->settle = 0.0f;
    p->fluid_dt = 0.017f * 2.0f;
    p->fluid_decay = 0.99f;
    p->fluid_vorticity = 7.7f;
    p->fluid_viscosity = 3.7f;
    p->fumarole_count = 0u;
    p->uv_strength = 3.0f;
    p->uv_depth = 0.0f;
    p->fluid_enabled = 1u;
    p->fluid_obstacle_strength = 1000.0f;
    p->slope_steer_rate = 210.0f;
    p->cohesion = 0.25f;
    p->spawn_count = 0u;
    p->max_agents = 0u;
    p->death_probability = 0.0006f;
    p->spawn_energy = 5.0f;
    p->food_power = 6.0f;
    p->maintenance_cost = 0.002f;
    p->pairing_rate = 3.0f;
    p->mutation_rate = 0.005f;
    p->uv_damage = 1.0f;
    p->uptake_rate = 0.0003f;
    p->brownian = 0.0f;
    p->phoretic_gain = 0.0f;
    p->chain_stiffness = 100.0f;
    p->thermal_kt = 0.1f;
    p->motor_amplitude = 0.4f;
    p->rft_enabled = 1u;
    p->require_start = 1u;
    p->joint_coupling = 0.0f;
    p->hunger_regulation = 1u;
    p->pairing_cost = 0.1f;
    p->reactivation_rate = 0.00001f;
    p->swim_gain = 2.0f;
    p->bioturbation = 0.39f;
    p->bioturbation_cost = 0.01f;
    p->monomer_uv_absorb = 0.3f;
    p->sedimentation = 0.0f;
    p->direct_photoactivation = 0.02f;
    p->activation_decay = 0.0f;
    p->monomer_pressure = 0.0f;
    p->agent_fluid_push = 0.001f;
    p->swim_wobble = 1.0f;
    p->fluid_swim_only = 0u;
    p->motion_cost = 0.06f;
    p->bond_rate = 0.05f;
    p->bond_energy_share = 0.1f;
    p->bond_signal = 0.5f;
    p->aggregation = 0.03f;
    p->photo_yield = 0.5f;
    p->heat_kill = 0.27f;
    p->flow_coupling = 0.9f;
    p->inertia = 0.0f;
    p->death_energy_cap = 1.0f;
    p->chemo_yield = 0.9f;
    p->redox_decay = 0.02f;
    p->thermal_activation = 0.045f;
    p->denature_temp = 2.0f;
    p->day_period = 8000.0f;
    p->sun_slope = 0.0f;
    p->sun_now = 1.0f;
    p->sun_angle = 80.0f;
    p->day_fraction = 0.75f;
    p->metabolic_q10 = 2.0f;
    p->metabolic_span = 3.0f;
    p->metabolic_ref = 0.0f;
    p->sun_heat = 1.0f;
    p->sediment_transport = 1.9f;
    p->sediment_threshold = 0.5f;
    p->sediment_settle = 0.0f;
    p->signal_mode = 0.0f;
    p->clock_mute = 0.0f;
    p->paint_x = 0.0f;
    p->paint_y = 0.0f;
    p->paint_radius = 0.0f;
    p->paint_grains = 0.0f;
    p->rest_angle_mult = 1.0f;
    p->protease_power = 1.5f;
    p->lysis_yield = 0.7f;
    p->protease_direct = 0.5f;
    p->joint_load = 1.0f;
    p->swim_grip = 10.0f;
    p->swim_memory = 1.0f;
    p->mouth_leak = 0.3f;
    p->leak_base = 0.05f;
    p->skin_uptake = 0.1f;
    p->death_metab = 0.2f;
    p->salvage = 0.0f;
    p->bond_matter_share = 0.1f;
    p->brownian_rot = 0.0f;
    p->transport_every = 3u;
    p->signal_crosstalk = 0.1f;
    p->flow_mass = 0.0f;
    p->chemo_take = 0.008f;
    p->vent_cycle_period = 7000.0f;
    p->vent_cycle_amp = 0.91f;
    p->sediment_cohesion = 0.5f;
    p->sediment_compaction = 0.00001f;
    p->rock_crumble = 0.0001f;
    p->lysis_matter = 0.5f;
    p->scramble_conduction = 0u;
    p->scramble_angles = 0u;
    p->scramble_response = 0u;
    p->scramble_amount = 1.0f;
    p->pairing_energy_half = 0.5f;
    p->pairing_min_energy = 0.0f;
    p->uv_mutation = 0.0f;
    p->heat_mutation = 0.0f;
    p->birth_crowding = 80.0f;
    p->photo_damage = 0.005f;
    p->naked_decay = 0.005f;
    p->store_loss = 0.0001f;
}

#endif
