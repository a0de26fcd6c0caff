/*
 * ribo_world.h - C ABI of the MI355X world-step library (libribo_world.so).
 *
 * The library owns one world per handle: the monomer grid, the terrain, the
 * fluid fields and the multigrid hierarchy, all resident in HBM, and advances
 * it with hand-written HIP kernels for gfx950 on one HIP stream per world.
 *
 * Every entry point returns an int status (RIBO_OK = 0, negative on error)
 * and never aborts; ribo_last_error() returns a thread-local message for the
 * last failing call on the calling thread.  Host pointers are borrowed for
 * the duration of the call only (copy in / copy out).  A handle is not
 * thread-safe: use one world from one thread at a time.
 *
 * Reference interface this boundary replaces (manalokos/Ribossome-v2):
 *   World::new / seed_matter / encode_steps / encode_ledger / ledger_blocking /
 *   read_cells_blocking / read_f32_blocking      src/world/mod.rs:416,969,1406,1661,1834,1807,1802
 *   SimParams (repr(C), 116 x 4 B = 464 B)        src/params.rs:86-487, defaults :553-690
 *   MgLevel                                        src/params.rs:727-740
 *   WorldConfig                                    src/params.rs:1003-1034
 *   WorldSettings                                  src/world/mod.rs:91-134
 *   buffer layout of groups 1 / 2 / 4              shaders/world/bindings.wgsl:4-62,
 *                                                  shaders/world/multigrid.wgsl:19-23
 */
#ifndef RIBO_WORLD_H
#define RIBO_WORLD_H

#include <stddef.h>
#include <stdint.h>

#if defined(__GNUC__)
#define RIBO_API __attribute__((visibility("default")))
#else
#define RIBO_API
#endif

#ifdef __cplusplus
extern "C" {
#endif

#define RIBO_ABI_VERSION 5u

/* Opaque world handle. */
typedef struct RiboWorld RiboWorld;

enum {
    RIBO_OK = 0,
    RIBO_ERR_ARG = -1,          /* bad handle, size or enum value */
    RIBO_ERR_HIP = -2,          /* a HIP runtime call failed */
    RIBO_ERR_NO_DEVICE = -3,    /* no gfx950 device visible */
    RIBO_ERR_UNSUPPORTED = -4,  /* a setting outside the world-step path */
    RIBO_ERR_LEDGER = -5,       /* the exact per-channel ledger check failed */
    RIBO_ERR_OOM = -6           /* device allocation failed */
};

/* WorldConfig (src/params.rs:1003-1034).  grid_size must be a multiple of
 * fluid_size (the reference runs 2048/1024 and 256/128). */
typedef struct RiboWorldConfig {
    uint32_t grid_size;
    uint32_t fluid_size;
    uint32_t world_units_per_cell;
    uint32_t max_agents;        /* accepted for layout parity; agents are not simulated */
} RiboWorldConfig;

/* SimParams, byte-identical to the reference's repr(C) struct
 * (src/params.rs:86-487): same 116 fields in the same order, 464 bytes.
 * Only the fields the world step reads are used; the rest are carried. */
typedef struct RiboSimParams {
    uint32_t epoch;
    uint32_t seed;
    float dt;
    float diffusion;
    float settle;
    float fluid_dt;
    float fluid_decay;
    float fluid_vorticity;
    float fluid_viscosity;
    uint32_t fumarole_count;
    float uv_strength;
    float uv_depth;
    uint32_t fluid_enabled;
    float fluid_obstacle_strength;
    float slope_steer_rate;
    float cohesion;
    uint32_t spawn_count;
    uint32_t max_agents;
    float death_probability;
    float spawn_energy;
    float food_power;
    float maintenance_cost;
    float pairing_rate;
    float mutation_rate;
    float uv_damage;
    float uptake_rate;
    float brownian;
    float phoretic_gain;
    float chain_stiffness;
    float thermal_kt;
    float motor_amplitude;
    uint32_t rft_enabled;
    uint32_t require_start;
    float joint_coupling;
    uint32_t hunger_regulation;
    float pairing_cost;
    float reactivation_rate;
    float swim_gain;
    float bioturbation;
    float bioturbation_cost;
    float monomer_uv_absorb;
    float sedimentation;
    float direct_photoactivation;
    float activation_decay;
    float monomer_pressure;
    float agent_fluid_push;
    float swim_wobble;
    uint32_t fluid_swim_only;
    float motion_cost;
    float bond_rate;
    float bond_energy_share;
    float bond_signal;
    float aggregation;
    float photo_yield;
    float heat_kill;
    float flow_coupling;
    float inertia;
    float death_energy_cap;
    float chemo_yield;
    float redox_decay;
    float thermal_activation;
    float denature_temp;
    float day_period;
    float sun_slope;
    float sun_now;
    float sun_angle;
    float day_fraction;
    float metabolic_q10;
    float metabolic_span;
    float metabolic_ref;
    float sun_heat;
    float sediment_transport;
    float sediment_threshold;
    float sediment_settle;
    float signal_mode;
    float clock_mute;
    float paint_x;
    float paint_y;
    float paint_radius;
    float paint_grains;
    float rest_angle_mult;
    float protease_power;
    float lysis_yield;
    float protease_direct;
    float joint_load;
    float swim_grip;
    float swim_memory;
    float mouth_leak;
    float leak_base;
    float skin_uptake;
    float death_metab;
    float salvage;
    float bond_matter_share;
    float brownian_rot;
    uint32_t transport_every;
    float signal_crosstalk;
    float flow_mass;
    float chemo_take;
    float vent_cycle_period;
    float vent_cycle_amp;
    float sediment_cohesion;
    float sediment_compaction;
    float rock_crumble;
    float lysis_matter;
    uint32_t scramble_conduction;
    uint32_t scramble_angles;
    uint32_t scramble_response;
    float scramble_amount;
    float pairing_energy_half;
    float pairing_min_energy;
    float uv_mutation;
    float heat_mutation;
    float birth_crowding;
    float photo_damage;
    float naked_decay;
    float store_loss;
} RiboSimParams;

/* MgLevel (src/params.rs:727-740): one multigrid level of the pressure
 * hierarchy; sizes halve from fluid_size while > mg_min_size and even. */
typedef struct RiboMgLevel {
    uint32_t n, off, n_c, off_c;
    float eps;
    uint32_t _pad0, _pad1, _pad2;
} RiboMgLevel;

/* WorldSettings (src/world/mod.rs:91-134) plus the solver extensions.
 * terrain_enabled = 1 runs the grain dynamics each step (relax_gamma_a,
 * relax_gamma_b, grain_fall when sediment_settle > 0, then the slope;
 * src/world/mod.rs:1495-1503), in one fixed admissible order of the
 * reference's concurrent moves (each pass decides on the counts at its
 * start, then moves the grains, then evicts the monomers that no longer fit,
 * cell colour (x mod 3, y mod 3) after colour); with 0 the terrain is a
 * frozen input.  With the tolerance mode the coarsest level's exact inverse
 * follows the moving terrain (rebuilt on the device when the coarsest fluid
 * mask changes; that level must have at most 256 cells).  Light follows
 * light_mode (an extension): 0 keeps the light
 * field a frozen input (RIBO_BUF_LIGHT), 1 runs the reference's light
 * schedule with light_rows_per_step and light_interval.  contact_enabled is
 * accepted and ignored (agents are not part of the path). */
typedef struct RiboWorldSettings {
    uint32_t fluid_enabled;       /* reference default 1 */
    uint32_t fluid_substep;       /* reference default 2, clamped to 1..4 */
    uint32_t multigrid;           /* 1 = multigrid, 0 = Jacobi ping-pong */
    uint32_t mg_cycles;           /* V-cycles per fluid step (default 1) */
    uint32_t jacobi_iters;        /* rounded up to even, clamped 2..128 */
    uint32_t light_rows_per_step; /* light_mode 1: propagation passes per step
                                     (at most 16; reference default 1); 0 = the
                                     full sweep every light_interval steps */
    uint32_t light_interval;      /* reference default 10 */
    uint32_t terrain_enabled;     /* grain dynamics; reference default 1 */
    uint32_t contact_enabled;     /* ignored */
    /* --- extensions (not in the reference) --- */
    uint32_t mg_min_size;         /* coarsest level bound, reference MG_MIN_SIZE = 8 */
    float    mg_tol;              /* 0: run exactly mg_cycles V-cycles (reference
                                     behaviour); > 0: tolerance mode, conjugate
                                     gradients preconditioned by a symmetric
                                     V-cycle until ||r||2/||rhs||2 <= mg_tol over
                                     fluid cells of level 0 */
    uint32_t mg_max_cycles;       /* iteration cap of the tolerance mode.  Before
                                     every 8 steps it records, per solve, the
                                     most iterations any of the last 32 solves
                                     took plus max(4, a third of it), rounded up
                                     to a multiple of 4, at most the cap; a solve
                                     that runs out is counted (see
                                     ribo_world_solver_stats) and the next 8
                                     steps record twice as many. */
    uint32_t precision;           /* RIBO_PRECISION_FP32 (drop-in) or _MIXED */
    uint32_t ledger_check;        /* 1: exact per-channel totals checked on the
                                     device after every transport pass */
    uint32_t no_graphs;           /* 1: record every step eagerly instead of
                                     replaying captured HIP graphs */
    uint32_t light_mode;          /* 0: the light field is a frozen input; 1 (the
                                     default): the reference's light step
                                     before the fluid of every step
                                     (src/world/mod.rs:1506-1525): the full sweep
                                     when the light is stale (after an upload of
                                     GAMMA or CHEM, or ribo_world_invalidate_light)
                                     or, with light_rows_per_step = 0, every
                                     light_interval steps; else
                                     light_rows_per_step propagation passes.  An
                                     upload of LIGHT makes it current. */
    uint32_t mg_stationary;       /* with mg_tol > 0: 0 = conjugate gradients
                                     preconditioned by the symmetric V-cycle (the
                                     default); 1 = the reference V-cycle repeated
                                     until the relative residual reaches mg_tol
                                     (binary32 only; at most mg_max_cycles) */
} RiboWorldSettings;

enum { RIBO_PRECISION_FP32 = 0, RIBO_PRECISION_MIXED = 1 };

/* Buffers a caller can move in or out.  Sizes in bytes (G = grid_size^2,
 * F = fluid_size^2, L = (grid_size/2)^2):
 *   CHEM        u32[4G]  chem_grid, (y*grid+x)*4+ch, low 16 bits activated,
 *                        high 16 spent (bindings.wgsl:5-7, chem.wgsl:6-7)
 *   GAMMA       u32[G]   grains per grid cell, >= 3 is rock (bindings.wgsl:10-11)
 *   LIGHT       f32[2L]  UV light, 2 bands (bindings.wgsl:12-14); frozen input
 *   SLOPE       f32[2G]  terrain slope, derived from GAMMA (download only)
 *   VELOCITY    f32[2F]  velocity_a, fluid cells per second (bindings.wgsl:21)
 *   PRESSURE    f32[F]   pressure_a (warm start of the solve)
 *   TEMP        f32[F]   temperature (temp_in)
 *   REDOX       f32[F]   reductant (redox_in)
 *   HEAT_SRC    f32[2F]  vent (heat, reductant) source per fluid cell
 *   DIVERGENCE  f32[F]   last divergence (download only)
 *   SOLID_MASK  u32[F]   fluid wall mask, 1 = solid (download only)
 *   NET_FLUX    f32[2*fluid_size] row means of v_y then column means of v_x
 *   REDOX_EATEN u32[F]   fixed-point reductant eaten since the last fluid step
 *   FORCE_VECTORS i32[2F] agent forces, fixed point x 256 per fluid cell (x, y):
 *                        force_vectors (fluid.wgsl:280-288) as agents leave it
 *                        (fold.wgsl:495-509); added to buoyancy by the next
 *                        fluid step, then cleared (upload only)
 *   VELOCITY_SMOOTH f32[2F] smooth_velocity (fluid.wgsl:680-693) of the
 *                        current velocity, computed at download (download only)
 *   AGG         u32[G]   activated count in each 3x3 block (agg_nb + agg_act;
 *                        download only, valid after a transport pass) */
typedef enum RiboBuffer {
    RIBO_BUF_CHEM = 0,
    RIBO_BUF_GAMMA = 1,
    RIBO_BUF_LIGHT = 2,
    RIBO_BUF_SLOPE = 3,
    RIBO_BUF_VELOCITY = 4,
    RIBO_BUF_PRESSURE = 5,
    RIBO_BUF_TEMP = 6,
    RIBO_BUF_REDOX = 7,
    RIBO_BUF_HEAT_SRC = 8,
    RIBO_BUF_DIVERGENCE = 9,
    RIBO_BUF_SOLID_MASK = 10,
    RIBO_BUF_NET_FLUX = 11,
    RIBO_BUF_REDOX_EATEN = 12,
    RIBO_BUF_AGG = 13,
    RIBO_BUF_FORCE_VECTORS = 14,
    RIBO_BUF_VELOCITY_SMOOTH = 15,
    RIBO_BUF_COUNT = 16
} RiboBuffer;

/* Maximum steps one ribo_world_encode_steps call records
 * (MAX_STEPS_PER_FRAME, src/world/mod.rs:16). */
#define RIBO_MAX_STEPS_PER_CALL 64u

/* ---- defaults ---------------------------------------------------------- */

/* SimParams::default() (src/params.rs:553-690), with seed = 1. */
RIBO_API void ribo_sim_params_default(RiboSimParams* out);
/* WorldSettings::default() (src/world/mod.rs:121-134) except terrain_enabled
 * = 0 (the frozen terrain of the benchmark's profile), plus mg_min_size 8, mg_tol 0, mg_max_cycles 1, precision FP32. */
RIBO_API void ribo_world_settings_default(RiboWorldSettings* out);
/* Byte size of a buffer for a config (0 for an unknown buffer). */
RIBO_API size_t ribo_buffer_bytes(const RiboWorldConfig* cfg, RiboBuffer which);

/* ---- lifetime (World::new, src/world/mod.rs:416-932) --------------------- */

/* Creates a world on the current HIP device: every hot-path buffer, the
 * multigrid level table and the stream.  All state starts at zero; params =
 * SimParams::default() with the given seed; settings = defaults above. */
RIBO_API int ribo_world_create(const RiboWorldConfig* cfg, uint32_t seed, RiboWorld** out);
/* Device bytes a world of this configuration holds with these settings
 * (NULL: the defaults); the exact coarse solve of the tolerance mode is
 * counted at its largest.  The reference sizes its buffers in World::new
 * (src/world/mod.rs:416-674); the serving form admits scenes by this figure. */
RIBO_API size_t ribo_world_footprint(const RiboWorldConfig* cfg, const RiboWorldSettings* settings);
/* Device bytes this world holds now. */
RIBO_API int ribo_world_device_bytes(const RiboWorld* w, size_t* out);
RIBO_API int ribo_world_destroy(RiboWorld* w);
/* The world's configuration (World::cfg, src/world/mod.rs:205). */
RIBO_API int ribo_world_get_config(const RiboWorld* w, RiboWorldConfig* out);

/* World.params / World.settings (pub fields, src/world/mod.rs:205-207). */
RIBO_API int ribo_world_set_params(RiboWorld* w, const RiboSimParams* p);
RIBO_API int ribo_world_get_params(const RiboWorld* w, RiboSimParams* p);
RIBO_API int ribo_world_set_settings(RiboWorld* w, const RiboWorldSettings* s);
RIBO_API int ribo_world_get_settings(const RiboWorld* w, RiboWorldSettings* s);
/* Multigrid level table (src/world/mod.rs:635-671): writes up to cap levels,
 * returns the level count in *count. */
RIBO_API int ribo_world_mg_levels(const RiboWorld* w, RiboMgLevel* out, uint32_t cap, uint32_t* count);

/* ---- data movement (queue.write_buffer / read_*_blocking) ---------------- */

/* Copies host bytes into a world buffer (synchronous). bytes must equal
 * ribo_buffer_bytes().  Uploading GAMMA rebuilds the derived terrain data
 * (slope, wall mask, fill, drag, wall normals, multigrid masks). */
RIBO_API int ribo_world_upload(RiboWorld* w, RiboBuffer which, const void* src, size_t bytes);
/* Copies a world buffer out to host memory after the world's queued work. */
RIBO_API int ribo_world_download(RiboWorld* w, RiboBuffer which, void* dst, size_t bytes);

/* ---- stepping (World::encode_steps, src/world/mod.rs:1406-1632) --------- */

/* Records min(steps, RIBO_MAX_STEPS_PER_CALL) world steps on the world's
 * stream, asynchronously, with the reference's per-step epoch, fluid_substep
 * and transport_every semantics, and advances params.epoch by that count.
 * Returns the number of steps recorded (>= 0) or a negative status. */
RIBO_API int ribo_world_encode_steps(RiboWorld* w, uint32_t steps);
/* Convenience: records `steps` steps in chunks of RIBO_MAX_STEPS_PER_CALL
 * (the tests' run() loop, tests/conservation.rs:14-24). */
RIBO_API int ribo_world_run(RiboWorld* w, uint32_t steps);
/* Waits for the world's stream (gpu.wait_idle, src/gpu/mod.rs:50-52). */
RIBO_API int ribo_world_sync(RiboWorld* w);
/* Non-blocking form of ribo_world_sync, the wgpu device.poll without wait
 * (src/gpu/mod.rs:51 waits; a serving loop polls): 1 when every recorded
 * step has finished, 0 while work is still running, negative on error. */
RIBO_API int ribo_world_idle(RiboWorld* w);

/* The tile side (16 or 32 cells) of each multigrid level's fused down / up
 * kernels, measured on this world's levels for its current solver (no
 * counterpart in the reference, which dispatches fixed 16 x 16 workgroups):
 * force 0 times both sides per level and keeps the faster (the choice is
 * remembered per level side for every world of the process); force 16 or 32
 * sets that side on every level.  Results do not depend on the tiles.
 * ribo_world_mg_tiles reads the sides back (0: the size rule, 32 above 256
 * cells a side, else 16). */
RIBO_API int ribo_world_autotune(RiboWorld* w, int force);
RIBO_API int ribo_world_mg_tiles(const RiboWorld* w, uint8_t* out, uint32_t cap, uint32_t* count);

/* World::invalidate_light (src/world/mod.rs:1400-1402): the next step runs
 * the full light sweep (light_mode 1). */
RIBO_API int ribo_world_invalidate_light(RiboWorld* w);

/* ---- agent coupling: monomer exchanges (chem.wgsl:40-98) ----------------- */

/* One exchange between the monomer grid and the caller (the life system's
 * agents): `count` single-monomer atomic steps on channel `channel` of grid
 * cell `cell`, each of kind:
 *   RIBO_CHEM_SPEND       activated -> spent in place (chem_spend_one)
 *   RIBO_CHEM_ACTIVATE    spent -> activated in place (chem_activate_one)
 *   RIBO_CHEM_TAKE_ACT    remove an activated monomer (chem_take_state_one)
 *   RIBO_CHEM_TAKE_SPENT  remove a spent monomer
 *   RIBO_CHEM_ADD_ACT     deposit activated monomers (chem_add_state)
 *   RIBO_CHEM_ADD_SPENT   deposit spent monomers
 *   RIBO_CHEM_DEPOSIT_ACT, RIBO_CHEM_DEPOSIT_SPENT  deposit in
 *                         chem_open_cell(cell): rock sends the deposit to the
 *                         nearest water or rubble cell (chem.wgsl:76-98)
 * A step that finds nothing to take stops that exchange. */
typedef struct RiboChemOp {
    uint32_t cell;
    uint32_t channel;
    uint32_t kind;
    uint32_t count;
} RiboChemOp;
enum {
    RIBO_CHEM_SPEND = 0,
    RIBO_CHEM_ACTIVATE = 1,
    RIBO_CHEM_TAKE_ACT = 2,
    RIBO_CHEM_TAKE_SPENT = 3,
    RIBO_CHEM_ADD_ACT = 4,
    RIBO_CHEM_ADD_SPENT = 5,
    RIBO_CHEM_DEPOSIT_ACT = 6,
    RIBO_CHEM_DEPOSIT_SPENT = 7
};
/* Deposits (ADD_*, DEPOSIT_*) add their count to the cell's 16-bit field at
 * once, as chem_add_state does (chem.wgsl:70-74): a count above 0xFFFF is
 * refused (RIBO_ERR_ARG) before anything changes; like the reference, a
 * deposit into a field that is nearly full wraps into the next field. */
/* Applies n exchanges on the device, all at once, between steps (the agents
 * of the reference act between world kernels in the same frame); done[i]
 * (may be NULL) receives the monomers exchange i moved.  Exchanges on
 * different channel words are independent; exchanges that take from the same
 * word compete as the reference's atomics do.  The per-channel totals the
 * ledger check expects follow what was taken and deposited (the reference's
 * Ledger counts it as held by agents). */
RIBO_API int ribo_world_chem_ops(RiboWorld* w, const RiboChemOp* ops, uint32_t n, uint32_t* done);

/* ---- ledger (Ledger, src/world/mod.rs:39-82; ledger.wgsl:5-27) ---------- */

/* Exact free-matter ledger of the current grid: out[0..3] activated A U G C,
 * out[4..7] spent, out[8..11] held in agents (always 0 here).  u32 wrap. */
RIBO_API int ribo_world_ledger(RiboWorld* w, uint32_t out[12]);
/* The same totals as ribo_world_ledger for a grid the caller holds in its
 * own device memory (`cells` cells of 4 u32 words on the current device),
 * computed on `stream` (a hipStream_t; NULL = the null stream), which the
 * call then waits for.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (serving loops, harnesses that forbid
 * hipMalloc).  It allocates nothing: its partial sums live in a fixed
 * device scratch of ~33 KB per GPU, held until the process ends and shared
 * by every call on that GPU, so calls that overlap on one GPU race.  At
 * most RIBO_LEDGER_NOALLOC_MAX_CELLS cells (RIBO_ERR_ARG beyond). */
#define RIBO_LEDGER_NOALLOC_MAX_CELLS (1ull << 30)
RIBO_API int ribo_ledger_noalloc(const uint32_t* chem, uint64_t cells, void* stream, uint32_t out[12]);
/* With settings.ledger_check = 1: returns RIBO_OK if every checked step kept
 * the per-channel totals seen when checking was armed (when ledger_check was
 * switched on, and again at every upload of RIBO_BUF_CHEM), else
 * RIBO_ERR_LEDGER and the first failing epoch in *epoch (0xFFFFFFFF when
 * none). */
RIBO_API int ribo_world_ledger_status(RiboWorld* w, uint32_t* epoch);

/* ---- measurement (KernelTiming / profile_kernels.rs) -------------------- */

/* Steps whose fluid part reads nothing the transport writes (the light
 * frozen, no solar heating, the terrain still) overlap a step's transport
 * with the next step's fluid (on by default; the results are the same):
 * with the reference's fixed V-cycles the transport chain runs on a second
 * stream of the world, ordered by events; in the conjugate-gradient
 * tolerance mode (one fluid step and one transport per step) the scatter's
 * tiles run inside the next solve's V-cycles, beside the coarse levels'
 * one-workgroup kernel.  Many worlds stepping at once already fill the
 * device: a caller that steps them on their own streams turns it off. */
RIBO_API int ribo_world_set_pipeline(RiboWorld* w, int enable);
/* Enables per-kernel HIP-event timing on the world's stream (off by default;
 * it adds an event pair per launch). */
RIBO_API int ribo_world_set_timing(RiboWorld* w, int enable);
/* Per-kernel totals since the last reset: names[i] (static strings), total
 * milliseconds, launch counts.  *count returns the number of kernels. */
RIBO_API int ribo_world_kernel_times(RiboWorld* w, const char** names, double* total_ms,
                            uint64_t* launches, uint32_t cap, uint32_t* count);
/* Every launch's milliseconds of kernel `name` since the last reset, in
 * launch order (up to cap of them into ms; *count returns how many there
 * are, 0 for an unknown name). */
RIBO_API int ribo_world_kernel_launch_ms(RiboWorld* w, const char* name, float* ms, uint32_t cap, uint32_t* count);
RIBO_API int ribo_world_reset_kernel_times(RiboWorld* w);
/* Last multigrid statistics of the world: V-cycles run in the last fluid
 * step and the final relative residual (tolerance mode only, else -1). */
RIBO_API int ribo_world_mg_stats(RiboWorld* w, uint32_t* cycles, double* rel_residual);

/* Tolerance-mode totals since the world was created: solves, iterations,
 * and solves that stopped on their recorded iteration slots before reaching
 * mg_tol. */
RIBO_API int ribo_world_solver_stats(RiboWorld* w, uint64_t* solves, uint64_t* iterations, uint64_t* missed);

/* Residuals of the tolerance mode (no counterpart in the reference, which
 * runs a fixed number of V-cycles).  The solve iterates on a double-float
 * x = hi + lo (hi is the stored binary32 pressure), evaluates w = A u and
 * every residual in binary64 and stops when the true residual
 * ||div - A x||2 / ||div||2 of x reaches mg_tol, restarting from that true
 * residual when the recursive one got there first (iterative refinement
 * with a binary64 refinement residual).  last / max: that residual of the
 * last solve and its maximum over the solves since the last reset;
 * last_rounded / max_rounded: the same for the stored binary32 pressure
 * (hi alone), whose rounding sets a floor of about 2 ulp(|p|) / rms(div);
 * restarts: restarts since the last reset.  reset != 0 clears the maxima and
 * the restart count after reading them. */
typedef struct RiboSolverResiduals {
    double last, last_rounded, max, max_rounded;
    uint64_t restarts;
} RiboSolverResiduals;
RIBO_API int ribo_world_solver_residuals(RiboWorld* w, RiboSolverResiduals* out, int reset);

/* ---- errors -------------------------------------------------------------- */
RIBO_API const char* ribo_last_error(void);
RIBO_API uint32_t ribo_abi_version(void);

/* ---- synthetic inputs (host only; no device needed) ---------------------- */

/* Deterministic generators for the benchmark worlds (SplitMix64 seeded):
 *   velocity: sum of `modes` sinusoid modes scaled so max |v| = vmax (fluid
 *             cells per second), zero in wall cells.  out f32[2F]
 *   gamma:    Gaussian blobs thresholded to rock_frac rock cells (6 grains),
 *             plus a rubble shell (1-2 grains, 1-3 cells) when rubble != 0.
 *   chem:     8 Poisson(lambda) counts per cell (4 bases x act/spent),
 *             0 in rock, each clamped to 0xFFFF.  out u32[4G]
 *   heat:     vent band: (h, h) on fluid rows [0, rows).  out f32[2F] */
RIBO_API int ribo_gen_gamma(const RiboWorldConfig* cfg, uint64_t seed, float rock_frac, int rubble, uint32_t* out);
RIBO_API int ribo_gen_velocity(const RiboWorldConfig* cfg, uint64_t seed, uint32_t modes, float vmax,
                      const uint32_t* gamma, float* out);
RIBO_API int ribo_gen_chem(const RiboWorldConfig* cfg, uint64_t seed, float lambda, const uint32_t* gamma, uint32_t* out);
RIBO_API int ribo_gen_vent_heat(const RiboWorldConfig* cfg, uint32_t rows, float h, float* out);

/* ---- vents and matter seeding (host only; the reference's CPU code) ----- */

/* Fumarole (src/params.rs:795-810): a point vent, position as a fraction of
 * the world, radius in world units; 32 bytes as the reference's repr(C). */
typedef struct RiboFumarole {
    float x_frac, y_frac;
    float strength;
    float spread;
    uint32_t enabled, _pad0, _pad1, _pad2;
} RiboFumarole;
/* Heat of a pure-red image pixel, FUMAROLE_PIXEL_STRENGTH (src/world/terrain.rs:247),
 * the default vent gain FUMAROLE_GAIN_DEFAULT and the seeding defaults
 * SEED_DENSITY_DEFAULT / SEED_ACTIVE_DEFAULT (src/world/mod.rs:1980-1984). */
#define RIBO_FUMAROLE_PIXEL_STRENGTH 5000.0f
#define RIBO_FUMAROLE_GAIN_DEFAULT 0.3f
#define RIBO_SEED_DENSITY_DEFAULT 0.83f
#define RIBO_SEED_ACTIVE_DEFAULT 1.0f

/* rasterize_fumaroles (src/world/terrain.rs:316-336): heat f32[G] of the
 * enabled fumaroles, strength (1 - d/r)^2 inside the radius, as a fraction
 * of RIBO_FUMAROLE_PIXEL_STRENGTH. */
RIBO_API int ribo_rasterize_fumaroles(const RiboWorldConfig* cfg, const RiboFumarole* f, uint32_t count, float* heat);
/* upload_heat's grid -> fluid step (src/world/mod.rs:1366-1392): out f32[2F]
 * = (heat, chemistry) box sums x RIBO_FUMAROLE_PIXEL_STRENGTH x max(gain, 0)
 * / k^2, k = grid / fluid, from the grid-resolution heat and chemistry. */
RIBO_API int ribo_vent_heat_src(const RiboWorldConfig* cfg, const float* heat_grid, const float* chem_grid, float gain,
                                float* out);
/* World::upload_heat with heat_grid / chem_grid (src/world/mod.rs:1343-1392):
 * the fumaroles' raster plus the optional heat image f32[G] (heat_image,
 * NULL for none) give the heat; the raster plus chem_image (or, without it,
 * the heat image) the chemistry; the result becomes RIBO_BUF_HEAT_SRC. */
RIBO_API int ribo_world_set_vents(RiboWorld* w, const RiboFumarole* f, uint32_t count, float gain,
                                  const float* heat_image, const float* chem_image);
/* seed_cells (src/world/mod.rs:1990-2014): the initial matter u32[4G], per
 * channel an fBm density field of 6 octaves, Poisson-like rounding and the
 * activated fraction `active` (SplitMix64 from seed, bit-identical to the
 * reference). */
RIBO_API int ribo_seed_cells(const RiboWorldConfig* cfg, uint64_t seed, float density, float active, uint32_t* cells);
/* fit_matter_to_terrain (src/world/mod.rs:1970-1983): rock empty, rubble
 * full to capacity with half activated on average, n cells. */
RIBO_API int ribo_fit_matter_to_terrain(uint32_t* cells, const uint32_t* gamma, uint64_t n, uint64_t seed);
/* refit_matter_to_terrain (src/world/mod.rs:1870-1968): matter over a
 * cell's capacity leaves it in proportion to its kinds and is spread over
 * every cell with room in proportion to the room; exact, nothing lost. */
RIBO_API int ribo_refit_matter_to_terrain(uint32_t* cells, const uint32_t* gamma, uint64_t n);
/* World::restart_keeping_terrain (src/world/mod.rs:990-1004): new matter
 * from seed_cells + fit_matter_to_terrain on the world's current terrain,
 * epoch 0; out[12] the seeded ledger (may be NULL). */
RIBO_API int ribo_world_restart_matter(RiboWorld* w, uint64_t seed, float density, float active, uint32_t out[12]);
/* World::apply_terrain_live (src/world/mod.rs:1283-1297) for the terrain:
 * the live matter is refitted to gamma u32[G] (refit_matter_to_terrain),
 * then both are uploaded. */
RIBO_API int ribo_world_apply_terrain_live(RiboWorld* w, const uint32_t* gamma);

#ifdef __cplusplus
}
#endif
#endif
