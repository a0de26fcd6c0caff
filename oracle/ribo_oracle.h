/*
 * ribo_oracle.h - CPU restatement of the reference world step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity oracle and the CPU
 * baseline for the world-step hot path of manalokos/Ribossome-v2.  Only the
 * tests, the smoke check and the CPU-baseline leg of the benchmark load it;
 * the product library (libribo_world.so) never links or calls it.
 *
 * It restates, kernel by kernel and in the reference's dispatch order, the
 * WGSL in shaders/world/{fluid,multigrid,transport,ledger}.wgsl and
 * shaders/common/{rng,chem,terrain}.wgsl as scheduled by
 * World::encode_steps (src/world/mod.rs:1406-1632), and the light step of
 * shaders/world/light.wgsl (light_mode 1).  Every function in
 * ribo_oracle.cpp cites the reference lines it follows.
 *
 * The fluid and multigrid are templated on the field type: the float
 * instance is the FP32 parity oracle, the double instance is the FP64 error
 * oracle.  Transport decisions are always binary32.
 *
 * Pinning (what checks each part, see tests/): PCG4D against golden vectors
 * from an independent numpy restatement; exp/sin against libm; the
 * multigrid against scipy sparse direct solves of the same eps-anchored
 * 5-point operator and against manufactured solutions; transport against
 * exact invariants (per-channel conservation, v = 0 identity, integer
 * shifts, rock blocking, edge clamping, the 256-mover cap) and hop
 * statistics; the ledger against a plain numpy count; the light against
 * closed forms (water attenuation per row, transmission of known grain and
 * monomer loads, the sun's sideways shift) and the propagation's fixed
 * point, which is the sweep.
 */
#ifndef RIBO_ORACLE_H
#define RIBO_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#include "../include/ribo_world.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct OracleWorld OracleWorld;

/* fp64 = 0: binary32 fields (parity oracle); 1: binary64 fluid fields. */
OracleWorld* oracle_create(const RiboWorldConfig* cfg, uint32_t seed, int fp64);
void oracle_destroy(OracleWorld* w);
/* World::invalidate_light: the next step runs the full light sweep. */
void oracle_invalidate_light(OracleWorld* w);
void oracle_set_params(OracleWorld* w, const RiboSimParams* p);
void oracle_get_params(const OracleWorld* w, RiboSimParams* p);
int oracle_set_settings(OracleWorld* w, const RiboWorldSettings* s);
/* Same buffer ids and byte sizes as the product ABI; binary64 worlds
 * convert to and from binary32 at the boundary. */
int oracle_upload(OracleWorld* w, int which, const void* src, size_t bytes);
int oracle_download(const OracleWorld* w, int which, void* dst, size_t bytes);
/* Binary64 fields of a binary64 world (VELOCITY, PRESSURE, TEMP, REDOX,
 * DIVERGENCE); for a binary32 world they are widened. */
int oracle_download_f64(const OracleWorld* w, int which, double* dst, size_t count);
/* encode_steps: min(steps, 64) steps; returns the count. */
int oracle_encode_steps(OracleWorld* w, uint32_t steps);
// The agents' exchanges with the grid, applied one after another in order
// (ribo_world_chem_ops applies them at once; the results agree whenever no
// two exchanges take from the same word beyond what it holds).
int oracle_chem_ops(OracleWorld* w, const RiboChemOp* ops, uint32_t n, uint32_t* done);
void oracle_ledger(const OracleWorld* w, uint32_t out[12]);
/* One reference kernel on the current state (closed-form tests): names
 * build_solid_mask, temperature, forces, diffuse, advect, vorticity,
 * divergence, subtract_gradient, net_flux, boundaries, transport,
 * thermal_activation.  Velocity results land in velocity_a. */
int oracle_run_kernel(OracleWorld* w, const char* name);
/* Relative residual ||r||2 / ||rhs||2 over fluid cells of level 0 for the
 * current pressure and divergence (eps-anchored operator). */
double oracle_mg_residual(const OracleWorld* w);
/* Iterations and final relative residual of the last tolerance-mode solve. */
void oracle_mg_stats(const OracleWorld* w, uint32_t* cycles, double* rel);
/* Tolerance mode, last solve: the refined (double-float) iterate's relative
 * residual in binary64, the stored pressure's, and the restarts from the
 * true residual. */
void oracle_solver_residuals(const OracleWorld* w, double* refined, double* rounded, uint32_t* restarts);
/* Threads used by the parallel loops (OpenMP); 0 = library default. */
void oracle_set_threads(int n);
int oracle_get_threads(void);

/* ---- single kernels for unit tests ---- */
/* pcg4d (rng.wgsl:10-22) on n vectors of 4 u32. */
void oracle_pcg4d(const uint32_t* in, uint32_t* out, size_t n);
/* hash (rng.wgsl:46-50). */
uint32_t oracle_hash(uint32_t v);
float oracle_expf(float x);
float oracle_sinf(float x);
/* One multigrid V-cycle schedule on a standalone problem: n x n level 0,
 * fluid mask u32[n*n], rhs f32[n*n], p f32[n*n] in/out, `cycles` V-cycles
 * with the reference schedule (pre 2, post 2, coarse 30) and mg_min_size.
 * fp64 selects the field type.  Returns the level count. */
int oracle_mg_solve(uint32_t n, uint32_t min_size, const uint32_t* fluid, const float* rhs, float* p,
                    uint32_t cycles, int fp64);
/* Same with binary64 arrays (for the FP64 pinning tests). */
int oracle_mg_solve_f64(uint32_t n, uint32_t min_size, const uint32_t* fluid, const double* rhs, double* p,
                        uint32_t cycles);

/* One restriction (mg_restrict, multigrid.wgsl:82-105) of an n x n level
 * (n even): fluid u32[n*n], rhs and p f32[n*n] in; the coarse right-hand
 * side f32[(n/2)^2] and fluid mask u32[(n/2)^2] out. */
int oracle_mg_restrict(uint32_t n, const uint32_t* fluid, const float* rhs, const float* p, float* rhs_c,
                       uint32_t* fluid_c);

/* One prolongation (mg_prolong, multigrid.wgsl:108-114): p += the parent's
 * coarse value on fluid cells of an n x n level. */
int oracle_mg_prolong(uint32_t n, const uint32_t* fluid, float* p, const float* p_c);
/* The tolerance mode's bilinear transfer, binary64: p_out = P c_in on the
 * n x n level, rhs_c = R r_in on its coarse level. */
int oracle_bl_transfer(uint32_t n, const uint32_t* fluid, const uint32_t* fluid_c, const double* c_in, double* p_out,
                       const double* r_in, double* rhs_c);
/* max |A inv - I| of the tolerance mode's exact coarse solve over an n x n
 * mask (-1 when it is not built). */
double oracle_coarse_inverse_error(uint32_t n, const uint32_t* fluid, float eps);

#ifdef __cplusplus
}
#endif
#endif
