// kernels.h - device data layout and kernel launchers of the world step.
#ifndef RIBO_KERNELS_H
#define RIBO_KERNELS_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/ribo_world.h"

namespace ribo {

// Bits of the per-fluid-cell flag byte derived from the terrain.
enum : uint8_t {
    CF_SOLID = 1,   // fluid wall (solid_mask = 1)
    CF_NORMAL = 2,  // solid_normal_from_neighbors is non-zero
};

// The sun of each step over a window of epochs [base, base + SUN_TAB_LEN):
// (sun_now, sun_slope) as the host computes them in binary64 for every
// epoch (src/params.rs:505-537, src/world/mod.rs:1439-1441).  Kernels that
// read the sun take it from here when a day cycle makes it change from step
// to step, so that captured steps replay with the sun of their epoch.
constexpr uint32_t SUN_TAB_LEN = 8192;
struct SunTab {
    uint32_t base, pad;
    float2 v[SUN_TAB_LEN];
};

// Terrain-derived, per-world constant data (rebuilt when gamma changes) and
// the sizes every kernel needs.  Passed by value.
struct Geo {
    uint32_t G, F, Lsz, wupc;
    float sim_size;
    const uint32_t* gamma;    // u32[G]
    const float* light;       // f32[2 * Lsz^2]
    const float2* slope;      // f32x2[G], compute_gamma_slope
    const float2* slope_f;    // slope at env_cell_for_fluid, f32x2[F]
    const uint8_t* cflags;    // CF_* per fluid cell
    const float2* normal;     // solid_normal_from_neighbors per fluid cell
    const float* fill;        // gamma_fill_at_fluid_cell per fluid cell
    const float* dragg;       // rubble grain fraction g of rubble_drag per fluid cell
    const float2* heat;       // heat_src, f32x2[F]
    const SunTab* sun;        // per-epoch sun (null: the sun in the step's parameters)
};

// bfloat16 storage word (round to nearest even on the high 16 bits of the
// binary32 value; infinities and NaN pass through).
struct bf16 { uint16_t bits; };
__host__ __device__ inline uint16_t bf16_bits(float v) {
    union { float f; uint32_t u; } c;
    c.f = v;
    uint32_t u = c.u;
    if ((u & 0x7F800000u) == 0x7F800000u) return (uint16_t)(u >> 16);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// One multigrid level resident in HBM; S is the storage type of its arrays.
// One multigrid level.  p holds the level's unknowns at the start of a
// V-cycle (0 below level 0) and its result at the end; pp holds them after
// the downward pre-smoothing, for the upward pass.  The two arrays keep a
// tile kernel that reads a halo from other tiles from reading values those
// tiles have already rewritten.
template <class S> struct MgLevelT {
    uint32_t n;
    float eps;
    S* p;
    S* pp;
    S* rhs;
    const uint8_t* fluid;
};
using MgLevelDev = MgLevelT<float>;

// Level-0 vectors of the preconditioned conjugate gradients (tolerance
// mode, Chronopoulos-Gear recurrences).  r and s are double-buffered: the
// kernel that updates them reads their old values with a halo.  u is the
// preconditioner's output and lives in its storage type S; rs is the
// preconditioner's copy of r (null when S is float: it reads r itself).
template <class S> struct CgVec {
    uint32_t n;
    float eps;
    const uint8_t* fl;
    float* x;        // the pressure: the high part of the iterate
    float* xlo;      // the iterate's low part (x + xlo is the double-float iterate)
    float* xold;     // the previous solve's pressure (null: no extrapolated start)
    const float* b;  // the divergence
    float* r[2];
    float* s[2];
    float* pd;       // search direction
    float* w;        // A u
    S* u;
    S* zpre;         // u after the downward pre-smoothing of level 0
    S* rs;
};

constexpr uint32_t PCG_HOST_RING = 4, PCG_HOST_RING_LEN = 64;
// The scalars of the conjugate-gradient recurrences after an iteration.
// fresh: the next inner products are the first of a (re)start; near: the
// V-cycle after this update also forms the true residual's norm.
struct PcgScal {
    double gamma, pi, beta, alpha, rel;
    uint32_t it, done, fresh, near, restart, restarts;
};
// Once the recursive residual is within PCG_NEAR x tol, every V-cycle also
// forms the true residual ||b - A x||^2 in binary64 and the solve stops on it.
constexpr double PCG_NEAR = 4.0;
// Solver state.  The scalars are double-buffered: the kernel that applies an
// iteration's inner products reads one slot (every workgroup forms the same
// update) and its workgroup 0 writes the other.  `done` is the published
// flag the V-cycle's kernels test; it, rel and true_rel describe the last
// finished solve.
struct PcgState {
    double bb, rel, true_rel;
    double rounded_rel;             // residual of the stored (binary32) pressure of the last solve
    double max_true_rel, max_rounded_rel;  // maxima over the solves since the last reset
    uint32_t restarts;              // restarts from the true residual since the last reset
    uint32_t near;                  // the current V-cycle forms the true residual (see PcgScal)
    double theta;       // the start's step along the last two solves' change
    PcgScal sc[2];
    uint32_t it, done, steps, iters, missed;
    uint32_t hist;      // solves since the pressure history restarted, at most 2: at 2 the
                        // pressure and the previous one are two solves' results
    float tol_seen;
    double* part;       // per-block partial sums (up to 5 per block of 16 x 16 cells)
    uint32_t* host_out; // host-mapped words: last iterations, misses, solves, then the
                        // iterations of the last PCG_HOST_RING_LEN solves (a ring)
};

constexpr uint32_t LEDGER_MAX_BLOCKS = 1024;
struct LedgerDev {
    uint32_t part[8 * LEDGER_MAX_BLOCKS];  // per-block act[4], spent[4]
    uint32_t armed;         // 1 once expected[] is set
    uint32_t expected[4];   // per-channel act + spent
    uint32_t first_bad;     // first epoch whose totals differed (0xFFFFFFFF none)
    uint32_t last[8];       // last completed reduction
};

// Arguments of the one-workgroup multigrid kernel (k_mg_block): every level
// from the first that fits one workgroup's LDS down to the coarsest.
constexpr int MG_BLOCK_LEVELS = 8;
template <class S>
struct MgBlockArgs {
    int levels;
    int n[MG_BLOCK_LEVELS];
    float eps[MG_BLOCK_LEVELS];
    const uint8_t* fluid[MG_BLOCK_LEVELS];
    int off[MG_BLOCK_LEVELS + 1];  // cell offsets of each level in LDS
    S* p0;
    const S* rhs0;
    const double* cinv;            // m x m inverse of the coarsest operator (SYM only)
    const int32_t* ccell;          // unknown -> coarsest cell
    int cm;                        // unknowns of the coarsest level (0: use sweeps)
    const int* cm_dev;             // when set, cm is read from here (the inverse is rebuilt on the device)
};
// The arguments and dynamic LDS bytes of k_mg_block over `count` levels.
template <class S>
MgBlockArgs<S> mg_block_args(const MgLevelT<S>* lv, int count, const double* cinv, const int32_t* ccell, int cm,
                             size_t* lds, const int* cm_dev = nullptr);
// Rebuilds the coarsest level's exact inverse on the device when its fluid
// mask (n x n <= 1024 cells, at most `cap` of them fluid) differs from the
// snapshot the inverse was built for: the host's Gauss-Jordan elimination
// (build_coarse_inverse), operation for operation.  `work` holds cap x cap
// doubles; cinv, ccell and cm are written as the host writes them.
void launch_coarse_refresh(hipStream_t s, uint32_t n, float eps, const uint8_t* fluid, uint8_t* snap, uint32_t cap,
                           double* work, double* cinv, int32_t* ccell, int* cm);

// ---- many worlds in one launch ----
// One workgroup of a batched launch: the world (index into the argument
// table), the block index the kernel body sees, and that world's grid size
// (gx | gy << 16).
struct BatchBlock { uint32_t w, bx, by, gdim; };
// Per-world arguments of the batched kernels (the single-world kernels'
// parameters; the step's epoch offset comes with the launch).
struct BTemperature {
    Geo g; RiboSimParams P; const uint32_t* ep;
    const float2* vel; const float* T; const float* Q; float* t_out; float* r_out; uint32_t* eaten; const uint32_t* chem;
};
struct BForces { Geo g; RiboSimParams P; const uint32_t* ep; const float* T; float2* forces; };
// velocity kernels: add_forces (aux = forces), diffuse, advect (decay), vorticity
struct BVel { Geo g; RiboSimParams P; float decay; const float2* aux; const float2* vin; float2* vout; };
// divergence (f0 = div, f1 = p0), subtract_gradient (f0 = p), net_flux and
// enforce_boundaries (f1 = net flux)
struct BProj { Geo g; const float2* vin; float2* vout; float* f0; float* f1; float2* vcopy; };  // vcopy: enforce_boundaries' copy
// thermal_activation, agg, the clear of chem_next, the scatter, the ledger
struct BChem {
    Geo g; RiboSimParams P; const uint32_t* ep;
    uint32_t* chem; uint32_t* next; const float2* vel; const float* temp; uint32_t* agg_e;
    LedgerDev* led; uint32_t ledger_blocks;
};
struct BMgPair { MgLevelDev L, C; };
enum BatchKernel {
    BK_TEMPERATURE, BK_FORCES, BK_ADD_FORCES, BK_DIFFUSE, BK_ADVECT, BK_VORTICITY, BK_DIVERGENCE, BK_SUBGRAD,
    BK_NET_FLUX, BK_BOUNDARIES, BK_THERMAL, BK_AGG, BK_CLEAR_NEXT, BK_SCATTER, BK_LEDGER, BK_LEDGER_FIN,
    BK_MG_DOWN, BK_MG_UP, BK_MG_BLOCK
};
// One batched launch of kernel k over `blocks` workgroups (map: device
// BatchBlock[blocks], tab: device table of the kernel's argument type).
void launch_batch(hipStream_t s, BatchKernel k, const void* tab, const BatchBlock* map, uint32_t blocks, uint32_t step,
                  size_t lds = 0);

// The tolerance mode's solve over a batch (preconditioner storage S): per
// world its conjugate-gradient vectors, level 1 of its preconditioner, its
// solver state, tolerance and cap, and its iteration slots for this solve
// (nit: the world's V-cycles are j = 0 .. nit; a batch launch for a larger
// j leaves it alone, as the single world records no such V-cycle).
template <class S> struct BCg {
    CgVec<S> V;
    MgLevelT<S> C;
    PcgState* st;
    float tol;
    uint32_t maxit;
    const uint32_t* nit;
    uint32_t nt2;   // tiles of the level-0 V-cycle kernels (their partial sums)
    uint32_t nb16;  // 16 x 16 blocks of the init and final kernels
};
template <class S> struct BMgPairT { MgLevelT<S> L, C; const uint32_t* done; };
template <class S> struct BMgBlock { MgBlockArgs<S> a; const uint32_t* done; };
enum BatchTolKernel {
    BT_CG_INIT, BT_FIN_INIT, BT_SHIFT, BT_CG_DOWN0, BT_MG_DOWN, BT_MG_BLOCK, BT_MG_UP, BT_CG_UP0, BT_CG_FINAL, BT_FIN_FINAL
};
// j: the V-cycle (CG_DOWN0, CG_UP0), stage and rd as launch_cg_down0 takes them.
template <class S>
void launch_batch_tol(hipStream_t s, BatchTolKernel k, const void* tab, const BatchBlock* map, uint32_t blocks, uint32_t j,
                      int stage, int rd, size_t lds = 0);
// The grid of a tolerance-mode kernel for one world of fluid side n
// (level_n: the multigrid level's side for MG_DOWN / MG_UP).
void batch_tol_grid(BatchTolKernel k, uint32_t n, uint32_t level_n, uint32_t* gx, uint32_t* gy);
// The grid of kernel k for one world (level_n: the multigrid level's side).
void batch_grid(BatchKernel k, const Geo& g, uint32_t level_n, uint32_t* gx, uint32_t* gy);

// ---- launchers (kernels.hip) ----
// terrain-derived data
void launch_slope(hipStream_t s, const Geo& g, float2* slope);
void launch_fluid_static(hipStream_t s, const Geo& g, uint8_t* cflags, float* fill, float* dragg, float2* slope_f);
void launch_fluid_normals(hipStream_t s, const Geo& g, uint8_t* cflags, float2* normal);
void launch_mg_mask0(hipStream_t s, const Geo& g, uint8_t* fluid0);
void launch_mg_mask_coarsen(hipStream_t s, uint32_t n, const uint8_t* fine, uint8_t* coarse);
void launch_mask_u32(hipStream_t s, const Geo& g, uint32_t* out);

// fluid step
// Kernels that read the epoch take `ep`: null, or a device word added to
// P.epoch (graph replays keep the epoch base on the device).
void launch_temperature(hipStream_t s, const Geo& g, const RiboSimParams& P, const uint32_t* ep, const float2* vel,
                        const float* t_in, const float* r_in, float* t_out, float* r_out,
                        uint32_t* eaten, int eaten_active, const uint32_t* chem);
// agent: the agents' fixed-point force words (i32[2F]), added and cleared,
// or null.
void launch_forces(hipStream_t s, const Geo& g, const RiboSimParams& P, const uint32_t* ep, const float* temp,
                   float2* forces, uint32_t* agent);
// smooth_velocity of vin into vout.
void launch_smooth_velocity(hipStream_t s, const Geo& g, const float2* vin, float2* vout);
void launch_add_forces(hipStream_t s, const Geo& g, const RiboSimParams& P, const float2* forces,
                       const float2* vin, float2* vout);
void launch_diffuse(hipStream_t s, const Geo& g, const RiboSimParams& P, const float2* vin, float2* vout);
void launch_advect(hipStream_t s, const Geo& g, const RiboSimParams& P, float decay, const float2* vin, float2* vout);
void launch_vorticity(hipStream_t s, const Geo& g, const RiboSimParams& P, const float2* vin, float2* vout);
void launch_divergence(hipStream_t s, const Geo& g, const float2* vin, float* div, float* p0);
void launch_jacobi(hipStream_t s, const Geo& g, const float* pin, float* pout, const float* div);
void launch_mg_smooth(hipStream_t s, const MgLevelDev& L, uint32_t color, const uint32_t* done);
// fused multigrid: pre-smooth + restrict, prolong + post-smooth per level,
// and the coarse end of the cycle (at most 8 levels) in one workgroup.
// sym = the symmetric V-cycle of the tolerance mode.
// tile: cells a side per workgroup, 16 or 32 (0: 32 above MG_SMALL cells a
// side, else 16); the results do not depend on it.
template <class S> void launch_mg_down(hipStream_t s, const MgLevelT<S>& L, const MgLevelT<S>& C, bool sym,
                                       const uint32_t* done, int tile = 0);
template <class S> void launch_mg_up(hipStream_t s, const MgLevelT<S>& L, const MgLevelT<S>& C, bool sym,
                                     const uint32_t* done, int tile = 0);
// Zeroes a level's unknowns (the symmetric cycle's coarsest level when no
// suffix of levels fits one workgroup).
template <class S> void launch_zero_level(hipStream_t s, const MgLevelT<S>& L);
// The tolerance mode solves its coarsest level exactly (dense inverse) when
// it has at most MG_SYM_COARSE_MAX fluid cells; the workgroup kernel stages
// that right-hand side in LDS.
constexpr uint32_t MG_SYM_COARSE_MAX = 1024;
size_t mg_block_lds_bytes(const uint32_t* n, int count);
template <class S> void launch_mg_block(hipStream_t s, const MgLevelT<S>* lv, int count, bool sym, const double* cinv,
                                        const int32_t* ccell, int cm, const uint32_t* done, const int* cm_dev = nullptr);
// preconditioned conjugate gradients
// r = b - A x and the solve's initial state (zeroes the other vectors); when
// V.xold is set and the last two solves' pressures are on hand, the start
// moves along their change to the point of least residual.
template <class S> void launch_cg_init(hipStream_t s, const CgVec<S>& V, PcgState* st, float tol, uint32_t maxit);
// Pre-smoothing and restriction of level 0, fused with the vector updates of
// the iteration (p, s, x, r; none on the first application, first = 1);
// parity selects which r and s buffers are read.  Iteration j >= 1 first
// applies the inner products of iteration j - 1 to the scalar state
// (reading slot rd, writing 1 - rd; iteration 0 reads slot 1).
template <class S>
void launch_cg_down0(hipStream_t s, const CgVec<S>& V, const MgLevelT<S>& C, PcgState* st, int parity, int first,
                     int rd, float tol, uint32_t maxit);
// Prolongation and post-smoothing of level 0 fused with w = A u and the
// inner products.
template <class S>
void launch_cg_up0(hipStream_t s, const CgVec<S>& V, const MgLevelT<S>& C, PcgState* st, int parity);
// The last iteration's scalar update (slot rd to 1 - rd), the true residual
// ||b - A x|| / ||b|| of the solve, and the statistics.
template <class S> void launch_cg_final(hipStream_t s, const CgVec<S>& V, PcgState* st, int rd, float tol, uint32_t maxit);
// Stationary tolerance mode around the reference V-cycle: phase 0 starts the
// solve (residual of the warm start), 1 checks the residual after a V-cycle
// and publishes `done`, 2 records the solve's statistics.
void launch_vc_step(hipStream_t s, int phase, uint32_t n, float eps, const uint8_t* fl, const float* x, const float* b,
                    PcgState* st, float tol, uint32_t maxit);
void launch_subgrad(hipStream_t s, const Geo& g, const float2* vin, float2* vout, const float* p);
void launch_net_flux(hipStream_t s, const Geo& g, const float2* vin, float* nf);
void launch_boundaries(hipStream_t s, const Geo& g, const float2* vin, float2* vout, const float* nf, float2* vcopy = nullptr);
void launch_thermal(hipStream_t s, const Geo& g, const RiboSimParams& P, const uint32_t* ep, const float* temp,
                    uint32_t* chem);

// light (two bands of Lsz x Lsz cells, band after band)
// Transmission of every light cell into `light` (the sweep's pre-pass).
void launch_light_transmit(hipStream_t s, const Geo& g, const RiboSimParams& P, const uint32_t* chem, float* light);
// The full top-to-bottom sweep, in place over the transmissions.
void launch_light_sweep(hipStream_t s, const Geo& g, const RiboSimParams& P, const uint32_t* ep, float* light);
// One propagation pass: lout from the rows above in lin (distinct buffers).
void launch_light_propagate(hipStream_t s, const Geo& g, const RiboSimParams& P, const uint32_t* ep, const uint32_t* chem,
                            const float* lin,
                            float* lout);

// exchanges between the grid and the agents: n RiboChemOp records in device
// memory (validated), done[n] (may be null), the ledger's expected totals.
void launch_chem_ops(hipStream_t s, const Geo& g, uint32_t* chem, const uint32_t* ops, uint32_t n, uint32_t* done,
                     uint32_t* expected);

// terrain grain dynamics: one relax pass (phase 0 or 1: decisions on the
// pass's counts into code[G], the grain moves with the incoming-grain masks
// in inc[G], the colour-ordered evictions) and grain_fall (three column
// colours; a no-op off the clock)
void launch_relax_pass(hipStream_t s, const Geo& g, const RiboSimParams& P, const uint32_t* ep, uint32_t phase,
                       const float2* vel, uint32_t* gamma, uint32_t* chem, uint8_t* code, uint8_t* inc);
void launch_grain_fall(hipStream_t s, const Geo& g, const RiboSimParams& P, const uint32_t* ep, uint32_t* gamma,
                       uint32_t* chem);

// transport
void launch_agg(hipStream_t s, const Geo& g, const uint32_t* chem, uint32_t* agg_e);
void launch_fill_u32(hipStream_t s, uint32_t* p, size_t n, uint32_t v);
void launch_scatter(hipStream_t s, const Geo& g, const RiboSimParams& P, const uint32_t* ep, const uint32_t* chem,
                    uint32_t* next, const float2* vel, const float* temp, const uint32_t* agg_e);
// The scatter's operands, for the launch that runs it beside the symmetric
// V-cycle's one-workgroup levels: one workgroup is launch_mg_block's (sym;
// with two workgroups to a CU, the last of the first pass over the device's
// `cus` CUs), the others take FuseCfg<S>::tiles of the scatter's 16 x 16
// tiles each, tiles [t0, t1) of the scatter_tiles(G) in row-major order.
// Fits when the V-cycle's LDS does not exceed the tiles'.
struct ScatterArgs {
    Geo g; RiboSimParams P; const uint32_t* ep;
    const uint32_t* chem; uint32_t* next; const float2* vel; const float* temp; const uint32_t* agg_e;
};
// Tiles of a workgroup and workgroups a CU holds at once: binary32 levels
// take 3 tiles (768 threads, 80 KB of LDS, two workgroups per CU); the
// bfloat16 ones, whose V-cycle needs more registers, 4 (1024 threads, one).
template <class S> struct FuseCfg { static constexpr uint32_t tiles = 3, per_cu = 2; };
template <> struct FuseCfg<bf16> { static constexpr uint32_t tiles = 4, per_cu = 1; };
uint32_t scatter_tiles(uint32_t G);
template <class S> bool mg_block_scatter_fits(const MgLevelT<S>* lv, int count);
// Tiles [t0, t1) of the scatter alone, one 256-thread workgroup a tile.
void launch_scatter_range(hipStream_t s, const ScatterArgs& a, uint32_t t0, uint32_t t1);
template <class S>
void launch_mg_block_scatter(hipStream_t s, const MgLevelT<S>* lv, int count, const double* cinv, const int32_t* ccell, int cm,
                             const uint32_t* done, const int* cm_dev, const ScatterArgs& sc, uint32_t t0, uint32_t t1,
                             uint32_t cus);
void launch_ledger(hipStream_t s, const Geo& g, const uint32_t* chem, LedgerDev* led, uint32_t epoch, const uint32_t* ep,
                   int check);
void launch_epoch_add(hipStream_t s, uint32_t* ep, uint32_t n);
// *p = v on the stream (a value fixed at launch: no host buffer to outlive the call)
void launch_set_u32(hipStream_t s, uint32_t* p, uint32_t v);
// The free-matter totals of `cells` cells of a grid into led->last (no
// check, no epoch).
void launch_ledger_cells(hipStream_t s, const uint32_t* chem, uint32_t cells, LedgerDev* led);
// Device address of this device's ledger scratch (a file-scope device
// array of kernels.hip, looked up once per device): the only buffer
// ribo_ledger_noalloc uses beyond its input.
hipError_t ledger_scratch(LedgerDev** out);

}  // namespace ribo

#endif
