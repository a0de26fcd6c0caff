// world_state.h - the state of one world (the library's RiboWorld handle),
// shared by the world's own code (world.hip) and the batched step
// (batch.hip).  Internal: not part of the C ABI.
#ifndef RIBO_WORLD_STATE_H
#define RIBO_WORLD_STATE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>
#include <string>
#include <vector>

#include "../../include/ribo_world.h"
#include "kernels.h"

namespace ribo {
struct Timer {
    const char* name;
    double total_ms = 0;
    uint64_t launches = 0;
    std::vector<float> each;  // every launch's milliseconds since the last reset
};


// Sets the calling thread's last error (ribo_last_error) and returns code.
int set_error(int code, const std::string& msg);
// pow(clamp(fluid_decay, 0, 1), dt * 60) of advect_velocity (fluid.wgsl:473),
// evaluated on the host in binary64 once per step.
float step_decay_factor(float decay, float fluid_dt);
// Steps between two choices of the tolerance mode's iteration slots (a
// graph holds one such chunk).
constexpr uint32_t TOL_CHUNK = 8u;

// Serialises device allocation and release with graph capture across threads.
std::recursive_mutex& alloc_mutex();
}  // namespace ribo

using namespace ribo;  // this internal header is shared by world.hip and batch.hip only

struct RiboWorld {
    RiboWorldConfig cfg{};
    uint32_t G = 0, F = 0, Lsz = 0;
    float sim_size = 0;
    RiboSimParams params{};
    RiboWorldSettings st{};
    bool fluid_was_on = true;  // src/world/mod.rs:929
    int device = 0;
    hipStream_t stream = nullptr;
    // Pipelined steps (nothing of the fluid step reads the monomers: the
    // light frozen, no solar heating, the terrain still): the transport
    // chain (thermal activation, aggregation, scatter, ledger) runs on
    // tstream beside the next fluid step on `stream`.  It reads the step's
    // velocity from vel_s, which enforce_boundaries writes beside vel_a.
    // Events order the streams: pe_fluid (a fluid step done), pe_temp[b]
    // and pe_vel (the transport's last reads of temp[b] and vel_s, which the
    // fluid step after next rewrites), pe_fork and pe_join (a recording's
    // start and end).  A recording starts and ends joined to `stream`.
    hipStream_t tstream = nullptr;
    float2* vel_s = nullptr;
    hipEvent_t pe_fork = nullptr, pe_join = nullptr, pe_fluid = nullptr, pe_temp[2] = {nullptr, nullptr}, pe_vel = nullptr;
    bool pipe = false, pipe_allowed = true;
    bool pipe_temp_read[2] = {false, false}, pipe_vel_read = false;
    // Tolerance mode under the same conditions (and one fluid step and one
    // transport per step): a step's scatter runs inside the next step's
    // solve, its tiles spread over the V-cycles' k_mg_block launches, which
    // otherwise keep one CU busy (launch_mg_block_scatter).  pend holds the
    // deferred scatter from its transport step to the next solve; it reads
    // the step's velocity from vel_s and writes chem[cur] (already
    // committed), which nothing reads before the solve ends.  A recording
    // runs what is still pending at its end.
    bool fuse = false;
    struct PendingScatter {
        bool on = false;
        RiboSimParams P{};
        const uint32_t* ep = nullptr;
        const uint32_t* chem = nullptr;
        uint32_t* next = nullptr;
        const float* temp = nullptr;
    } pend;
    uint32_t cu_count = 0;

    // world-resolution buffers (bindings.wgsl group 1)
    uint32_t* chem[2] = {nullptr, nullptr};  // chem_grid / chem_next, ping-pong
    int cur = 0;
    uint32_t* gamma = nullptr;
    float* light[2] = {nullptr, nullptr};  // light_grid / light_next, ping-pong
    int lcur = 0;
    bool light_dirty = true;  // src/world/mod.rs:209
    float2* slope = nullptr;
    uint32_t* agg_e = nullptr;
    // fluid-resolution buffers (group 2)
    float2* vel_a = nullptr;
    float2* vel_b = nullptr;
    float* p_a = nullptr;
    float* p_b = nullptr;
    float* p_old = nullptr;  // tolerance mode: the previous solve's pressure
    float* div = nullptr;
    float* temp[2] = {nullptr, nullptr};
    int tcur = 0;
    float* redox[2] = {nullptr, nullptr};
    uint32_t* eaten = nullptr;
    bool eaten_active = false;
    float2* forces = nullptr;
    uint32_t* agent_force = nullptr;  // force_vectors from the agents, i32[2F] fixed point
    uint8_t* terrain_code = nullptr;  // per grid cell: the move a relax pass decided (0: none)
    uint8_t* terrain_inc = nullptr;   // per grid cell: the neighbours whose grain arrived (bit mask)
    bool agent_force_active = false;  // consumed (added, then cleared) by the next fluid step
    float2* heat = nullptr;
    float* net_flux = nullptr;
    // terrain-derived
    uint8_t* cflags = nullptr;
    float2* normal = nullptr;
    float* fill = nullptr;
    float* dragg = nullptr;
    float2* slope_f = nullptr;
    // multigrid (group 4); level 0 aliases p_a / div
    std::vector<RiboMgLevel> levels;
    std::vector<MgLevelDev> mg;
    float* mg_p = nullptr;
    float* mg_pp = nullptr;     // pre-smoothed unknowns of every level
    float* mg_rhs = nullptr;
    uint8_t* mg_fluid = nullptr;
    size_t mg_block_start = 0;  // first level solved inside one workgroup
    std::vector<uint8_t> mg_tile;  // tile side of each level's down / up kernels (0: the size rule)
    size_t sym_coarse = 0;      // level the tolerance mode solves exactly (the coarsest)
    size_t lds_limit = 65536;
    // tolerance mode (preconditioned conjugate gradients)
    PcgState* pcg = nullptr;
    double* pcg_part = nullptr;
    float* pcg_r = nullptr;   // residual, two buffers
    float* pcg_r1 = nullptr;
    float* pcg_s0 = nullptr;  // A p, two buffers
    float* pcg_s1 = nullptr;
    float* pcg_pd = nullptr;  // search direction
    float* pcg_xlo = nullptr; // low part of the double-float iterate
    float* pcg_w = nullptr;   // A u
    float* pcg_z = nullptr;
    bf16* pcg_r16 = nullptr;
    bf16* pcg_z16 = nullptr;
    bf16* mg_p16 = nullptr;     // coarse levels of the bfloat16 preconditioner
    bf16* mg_pp16 = nullptr;    // pre-smoothed unknowns, every level
    bf16* mg_rhs16 = nullptr;
    std::vector<MgLevelT<float>> mgz;   // preconditioner levels, binary32 (level 0 = z, r)
    std::vector<MgLevelT<bf16>> mgh;    // preconditioner levels, bfloat16 (level 0 = z16, r16)
    double* cinv = nullptr;             // inverse of the exactly solved level's operator (binary64, by columns)
    int32_t* ccell = nullptr;
    int cm = 0;
    bool cinv_valid = false;
    // moving terrain in the tolerance mode: the coarsest inverse is rebuilt
    // on the device whenever the coarsest fluid mask changes
    bool cmoving = false;
    uint32_t ccap = 0;       // capacity (coarsest cells) of cinv / ccell
    uint8_t* csnap = nullptr;  // the mask the inverse was built for
    double* cwork = nullptr;   // 2 ccap^2 doubles of elimination work
    int* cm_dev = nullptr;     // the inverse's unknowns
    uint32_t* pcg_host = nullptr;   // host-mapped: last iterations, misses, solves
    uint32_t pcg_slots = 0;         // iteration slots recorded per solve (0: not set)
    uint32_t missed_seen = 0;
    uint32_t solves_seen = 0;
    uint32_t solves_at_settings = 0;  // solves recorded before the last settings change
    LedgerDev* ledger = nullptr;
    uint32_t* scratch_u32 = nullptr;
    // device bytes: allocated at creation (multigrid arrays aside), by the
    // level table, and by the exact coarse solve's inverse
    size_t base_bytes = 0, mg_bytes = 0, cinv_bytes = 0;  // F words for downloads of derived masks
    // graph replay: the epoch base lives on the device; one executable graph
    // per (schedule phase, step count, ping-pong parities)
    uint32_t* epoch_dev = nullptr;
    uint32_t epoch_dev_value = 0;
    // the sun of each epoch over a window, when a day cycle moves it
    SunTab* sun_tab = nullptr;
    std::vector<SunTab> sun_host;  // one entry: the last window written
    bool sun_active = false, sun_valid = false;
    float sun_key[4] = {0, 0, 0, 0};
    bool epoch_dev_known = false;
    struct GraphEntry { uint32_t phase, n, slots; int cur, tcur, lcur, cur_end, tcur_end, lcur_end; hipGraphExec_t exec; };
    std::vector<GraphEntry> graphs;
    RiboSimParams graph_params{};
    RiboWorldSettings graph_settings{};

    // timing
    bool timing = false;
    std::vector<ribo::Timer> timers;
    struct Pending { size_t timer; hipEvent_t a, b; };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> event_pool;

    Geo geo() const {
        Geo g;
        g.G = G; g.F = F; g.Lsz = Lsz; g.wupc = cfg.world_units_per_cell; g.sim_size = sim_size;
        g.gamma = gamma; g.light = light[lcur]; g.slope = slope; g.slope_f = slope_f; g.cflags = cflags;
        g.normal = normal; g.fill = fill; g.dragg = dragg; g.heat = heat;
        g.sun = sun_active ? sun_tab : nullptr;
        return g;
    }
};

namespace ribo {
// The tolerance mode's per-chunk set-up of a world stepped in a batch (its
// coarsest inverse and iteration slots; world.hip).
int batch_prepare_solve(RiboWorld* w);
}  // namespace ribo

#endif
