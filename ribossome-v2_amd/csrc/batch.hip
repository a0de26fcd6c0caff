// batch.hip - many worlds stepped together with one launch per kernel.
//
// A step of the batch records the world step's kernels once each over the
// blocks of every world (launch_batch): a block map, built once, tells each
// workgroup its world and block, and an argument table per kernel (one per
// ping-pong state of the chem and temperature buffers) holds each world's
// pointers and parameters.  The kernel bodies are the single-world kernels'
// own, so a world in a batch computes what it computes alone.
//
// The schedule is World::encode_steps' (src/world/mod.rs:1406-1632) for the
// profile the batch supports: fluid every fluid_substep epochs (the
// reference V-cycle, mg_cycles times), transport every transport_every
// epochs with the ledger check, no light step, frozen terrain.

#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/ribo_batch.h"
#include "kernels.h"
#include "world_state.h"

using namespace ribo;

namespace {

#define BCHK(expr)                                                                                  \
    do {                                                                                            \
        hipError_t e_ = (expr);                                                                     \
        if (e_ != hipSuccess) return set_error(RIBO_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

constexpr uint32_t MAX_BATCH = 4096;

// A device block map of one launch shape.
struct Shape {
    const BatchBlock* map = nullptr;
    uint32_t blocks = 0;
};

}  // namespace

struct RiboBatch {
    std::vector<RiboWorld*> worlds;
    hipStream_t stream = nullptr;
    std::vector<hipEvent_t> world_done;  // one per world: its queued work
    hipEvent_t batch_done = nullptr;
    uint32_t* ep = nullptr;              // the device epoch base of every world (they step together)
    char* dev = nullptr;                 // maps and tables, one allocation
    size_t dev_bytes = 0;
    uint64_t launches = 0;
    uint32_t substep = 1, te = 1, cycles = 1, depth = 0;
    bool ledger = false, agg = false;
    size_t block_lds = 0;
    // shapes
    Shape fluid, grid, netflux, clear, ledger_blocks, one;
    std::vector<Shape> mg;               // per multigrid depth d: worlds with more than d tile levels
    // tables (device pointers): [tcur] or [cur][tcur] or per depth
    const BTemperature* temperature[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};  // [cur][tcur]
    const BForces* forces[2] = {nullptr, nullptr};
    const BVel *add_forces = nullptr, *diffuse = nullptr, *advect = nullptr, *vorticity = nullptr;
    const BProj *divergence = nullptr, *subgrad = nullptr, *netflux_t = nullptr, *boundaries = nullptr;
    const BChem* chem[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
    std::vector<const BMgPair*> mgpair;
    const MgBlockArgs<float>* mgblock = nullptr;
    // what each world looked like at creation (parameters and settings must not change)
    std::vector<RiboSimParams> params0;
    std::vector<RiboWorldSettings> settings0;
};

namespace {

int check_world(const RiboWorld* w, const RiboWorld* first) {
    const RiboWorldSettings& s = w->st;
    const RiboSimParams& p = w->params;
    if (!s.fluid_enabled || !s.multigrid || s.mg_tol > 0.0f)
        return set_error(RIBO_ERR_UNSUPPORTED, "a batch runs the reference V-cycle: fluid and multigrid on, mg_tol 0");
    if (s.terrain_enabled || s.light_mode)
        return set_error(RIBO_ERR_UNSUPPORTED, "a batch runs a frozen terrain and light field (terrain_enabled 0, light_mode 0)");
    if (p.sun_heat > 0.0f && p.day_period >= 1.0f)
        return set_error(RIBO_ERR_UNSUPPORTED, "a batch does not run solar heating under a moving sun");
    if (w->eaten_active || w->agent_force_active)
        return set_error(RIBO_ERR_UNSUPPORTED, "a world with pending agent words cannot join a batch");
    if (w->mg_block_start >= w->mg.size())
        return set_error(RIBO_ERR_UNSUPPORTED, "a world whose coarse levels do not fit one workgroup cannot join a batch");
    if (!first) return RIBO_OK;
    const RiboWorldSettings& f = first->st;
    auto clampss = [](uint32_t v) { return std::min(std::max(v, 1u), 4u); };
    if (clampss(s.fluid_substep) != clampss(f.fluid_substep) ||
        std::max(p.transport_every, 1u) != std::max(first->params.transport_every, 1u) ||
        std::max(s.mg_cycles, 1u) != std::max(f.mg_cycles, 1u) || (s.ledger_check != 0) != (f.ledger_check != 0) ||
        p.epoch != first->params.epoch || w->cur != first->cur || w->tcur != first->tcur || w->device != first->device)
        return set_error(RIBO_ERR_ARG, "the worlds of a batch must share fluid_substep, transport_every, mg_cycles, "
                                       "ledger_check, the epoch, the ping-pong state and the device");
    return RIBO_OK;
}

// The step's parameters as encode_steps forms them (src/world/mod.rs:1431-1445).
RiboSimParams step_params(const RiboWorld* w, uint32_t substep) {
    RiboSimParams P = w->params;
    P.fluid_dt = P.dt * (float)substep;
    P.fluid_enabled = 1u;
    P.epoch = 0u;  // each launch adds its step; the device base adds the chunk's first epoch
    P.sun_now = P.uv_strength > 0.0f ? (P.day_period < 1.0f ? 1.0f : 0.0f) : 0.0f;
    P.sun_slope = 0.0f;
    return P;
}

template <class T> size_t align_up(size_t n) { return (n + alignof(T) + 255) & ~(size_t)255; }

}  // namespace

extern "C" {

int ribo_batch_create(RiboWorld* const* worlds, uint32_t count, RiboBatch** out) {
    if (!worlds || !out || count == 0) return set_error(RIBO_ERR_ARG, "null argument or empty batch");
    if (count > MAX_BATCH) return set_error(RIBO_ERR_ARG, "at most 4096 worlds in a batch");
    *out = nullptr;
    for (uint32_t i = 0; i < count; ++i) {
        if (!worlds[i]) return set_error(RIBO_ERR_ARG, "null world");
        int rc = check_world(worlds[i], i ? worlds[0] : nullptr);
        if (rc) return rc;
        for (uint32_t j = 0; j < i; ++j)
            if (worlds[j] == worlds[i]) return set_error(RIBO_ERR_ARG, "a world appears twice in the batch");
    }
    RiboBatch* b = new RiboBatch();
    b->worlds.assign(worlds, worlds + count);
    const RiboWorld* w0 = worlds[0];
    b->substep = std::min(std::max(w0->st.fluid_substep, 1u), 4u);
    b->te = std::max(w0->params.transport_every, 1u);
    b->cycles = std::max(w0->st.mg_cycles, 1u);
    b->ledger = w0->st.ledger_check != 0;
    for (RiboWorld* w : b->worlds) {
        b->agg = b->agg || w->params.aggregation > 0.0f;
        b->depth = std::max<uint32_t>(b->depth, (uint32_t)w->mg_block_start);
        b->params0.push_back(w->params);
        b->settings0.push_back(w->st);
    }
    // ---- block maps (world-major) ----
    std::vector<BatchBlock> blocks;
    std::vector<std::pair<Shape*, size_t>> shape_at;  // shape -> first entry
    auto add_shape = [&](Shape* sh, BatchKernel k, int depth) {
        const size_t first = blocks.size();
        for (uint32_t wi = 0; wi < count; ++wi) {
            const RiboWorld* w = b->worlds[wi];
            uint32_t gx = 0, gy = 0, n = 0;
            if (depth >= 0) {
                if ((size_t)depth >= w->mg_block_start) continue;  // this world has no tile level at that depth
                n = w->mg[(size_t)depth].n;
            }
            batch_grid(k, w->geo(), n, &gx, &gy);
            for (uint32_t by = 0; by < gy; ++by)
                for (uint32_t bx = 0; bx < gx; ++bx) blocks.push_back(BatchBlock{wi, bx, by, gx | (gy << 16)});
        }
        sh->blocks = (uint32_t)(blocks.size() - first);
        shape_at.push_back({sh, first});
    };
    add_shape(&b->fluid, BK_DIFFUSE, -1);
    add_shape(&b->grid, BK_SCATTER, -1);
    add_shape(&b->netflux, BK_NET_FLUX, -1);
    add_shape(&b->clear, BK_CLEAR_NEXT, -1);
    add_shape(&b->ledger_blocks, BK_LEDGER, -1);
    add_shape(&b->one, BK_MG_BLOCK, -1);
    b->mg.resize(b->depth);
    for (uint32_t d = 0; d < b->depth; ++d) add_shape(&b->mg[d], BK_MG_DOWN, (int)d);
    // ---- argument tables ----
    std::vector<BTemperature> t_temp(4 * count);  // [cur][tcur]
    std::vector<BForces> t_forces(2 * count);
    std::vector<BVel> t_vel(4 * count);    // add_forces, diffuse, advect, vorticity
    std::vector<BProj> t_proj(4 * count);  // divergence, subgrad, net_flux, boundaries
    std::vector<BChem> t_chem(4 * count);  // [cur][tcur]
    std::vector<BMgPair> t_mg((size_t)b->depth * count);
    std::vector<MgBlockArgs<float>> t_block(count);
    for (uint32_t wi = 0; wi < count; ++wi) {
        RiboWorld* w = b->worlds[wi];
        Geo g = w->geo();
        g.sun = nullptr;  // the sun of every step is step_params' (no table: a moving sun is refused)
        const RiboSimParams P = step_params(w, b->substep);
        for (int t = 0; t < 2; ++t) {
            for (int c = 0; c < 2; ++c)
                t_temp[(size_t)(c * 2 + t) * count + wi] = BTemperature{
                    g, P, b->ep, w->vel_a, w->temp[t], w->redox[t], w->temp[1 - t], w->redox[1 - t], w->eaten, w->chem[c]};
            t_forces[t * count + wi] = BForces{g, P, b->ep, w->temp[t], w->forces};
        }
        const float decay = step_decay_factor(P.fluid_decay, P.fluid_dt);
        t_vel[0 * count + wi] = BVel{g, P, decay, w->forces, w->vel_a, w->vel_b};
        t_vel[1 * count + wi] = BVel{g, P, decay, nullptr, w->vel_b, w->vel_a};
        t_vel[2 * count + wi] = BVel{g, P, decay, nullptr, w->vel_a, w->vel_b};
        t_vel[3 * count + wi] = BVel{g, P, decay, nullptr, w->vel_b, w->vel_a};
        t_proj[0 * count + wi] = BProj{g, w->vel_a, nullptr, w->div, w->p_a};
        t_proj[1 * count + wi] = BProj{g, w->vel_a, w->vel_b, w->p_a, nullptr};
        t_proj[2 * count + wi] = BProj{g, w->vel_b, nullptr, nullptr, w->net_flux};
        t_proj[3 * count + wi] = BProj{g, w->vel_b, w->vel_a, nullptr, w->net_flux};
        const uint32_t cells = w->G * w->G;
        const uint32_t lb = std::min<uint32_t>((cells + 255u) / 256u, LEDGER_MAX_BLOCKS);
        for (int c = 0; c < 2; ++c)
            for (int t = 0; t < 2; ++t)
                t_chem[(size_t)(c * 2 + t) * count + wi] =
                    BChem{g, P, b->ep, w->chem[c], w->chem[1 - c], w->vel_a, w->temp[t], w->agg_e, w->ledger, lb};
        for (uint32_t d = 0; d < b->depth; ++d)
            if (d < w->mg_block_start) t_mg[(size_t)d * count + wi] = BMgPair{w->mg[d], w->mg[d + 1]};
        size_t lds = 0;
        const size_t bs = w->mg_block_start;
        t_block[wi] = mg_block_args<float>(w->mg.data() + bs, (int)(w->mg.size() - bs), nullptr, nullptr, 0, &lds);
        b->block_lds = std::max(b->block_lds, lds);
    }
    // ---- one device allocation: the epoch word, maps, tables ----
    size_t bytes = 256;
    auto reserve = [&](size_t n) {
        const size_t at = bytes;
        bytes += (n + 255) & ~(size_t)255;
        return at;
    };
    const size_t o_map = reserve(blocks.size() * sizeof(BatchBlock));
    const size_t o_temp = reserve(t_temp.size() * sizeof(BTemperature));
    const size_t o_forces = reserve(t_forces.size() * sizeof(BForces));
    const size_t o_vel = reserve(t_vel.size() * sizeof(BVel));
    const size_t o_proj = reserve(t_proj.size() * sizeof(BProj));
    const size_t o_chem = reserve(t_chem.size() * sizeof(BChem));
    const size_t o_mg = reserve(t_mg.size() * sizeof(BMgPair));
    const size_t o_block = reserve(t_block.size() * sizeof(MgBlockArgs<float>));
    {
        std::lock_guard<std::recursive_mutex> lk(alloc_mutex());
        hipError_t e = hipMalloc((void**)&b->dev, bytes);
        if (e != hipSuccess) {
            delete b;
            return set_error(RIBO_ERR_OOM, std::string("hipMalloc: ") + hipGetErrorString(e));
        }
    }
    b->dev_bytes = bytes;
    b->ep = (uint32_t*)b->dev;
    // the tables hold the epoch pointer: fill it now that it is known
    for (auto& a : t_temp) a.ep = b->ep;
    for (auto& a : t_forces) a.ep = b->ep;
    for (auto& a : t_chem) a.ep = b->ep;
    int rc = RIBO_OK;
    hipError_t he = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    auto up = [&](size_t off, const void* src, size_t n) {
        if (he == hipSuccess && n) he = hipMemcpyAsync(b->dev + off, src, n, hipMemcpyHostToDevice, b->stream);
    };
    up(o_map, blocks.data(), blocks.size() * sizeof(BatchBlock));
    up(o_temp, t_temp.data(), t_temp.size() * sizeof(BTemperature));
    up(o_forces, t_forces.data(), t_forces.size() * sizeof(BForces));
    up(o_vel, t_vel.data(), t_vel.size() * sizeof(BVel));
    up(o_proj, t_proj.data(), t_proj.size() * sizeof(BProj));
    up(o_chem, t_chem.data(), t_chem.size() * sizeof(BChem));
    up(o_mg, t_mg.data(), t_mg.size() * sizeof(BMgPair));
    up(o_block, t_block.data(), t_block.size() * sizeof(MgBlockArgs<float>));
    if (he == hipSuccess) he = hipStreamSynchronize(b->stream);
    if (he == hipSuccess) he = hipEventCreateWithFlags(&b->batch_done, hipEventDisableTiming);
    b->world_done.assign(count, nullptr);
    for (uint32_t i = 0; i < count && he == hipSuccess; ++i) he = hipEventCreateWithFlags(&b->world_done[i], hipEventDisableTiming);
    if (he != hipSuccess) rc = set_error(RIBO_ERR_HIP, std::string("batch set-up: ") + hipGetErrorString(he));
    if (rc) {
        ribo_batch_destroy(b);
        return rc;
    }
    const BatchBlock* map = (const BatchBlock*)(b->dev + o_map);
    for (auto& sa : shape_at) sa.first->map = map + sa.second;
    for (int t = 0; t < 2; ++t) {
        for (int c = 0; c < 2; ++c) b->temperature[c][t] = (const BTemperature*)(b->dev + o_temp) + (size_t)(c * 2 + t) * count;
        b->forces[t] = (const BForces*)(b->dev + o_forces) + t * count;
    }
    const BVel* tv = (const BVel*)(b->dev + o_vel);
    b->add_forces = tv;
    b->diffuse = tv + count;
    b->advect = tv + 2 * count;
    b->vorticity = tv + 3 * count;
    const BProj* tp = (const BProj*)(b->dev + o_proj);
    b->divergence = tp;
    b->subgrad = tp + count;
    b->netflux_t = tp + 2 * count;
    b->boundaries = tp + 3 * count;
    for (int c = 0; c < 2; ++c)
        for (int t = 0; t < 2; ++t) b->chem[c][t] = (const BChem*)(b->dev + o_chem) + (size_t)(c * 2 + t) * count;
    for (uint32_t d = 0; d < b->depth; ++d) b->mgpair.push_back((const BMgPair*)(b->dev + o_mg) + (size_t)d * count);
    b->mgblock = (const MgBlockArgs<float>*)(b->dev + o_block);
    *out = b;
    return RIBO_OK;
}

int ribo_batch_destroy(RiboBatch* b) {
    if (!b) return set_error(RIBO_ERR_ARG, "null batch");
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (hipEvent_t e : b->world_done)
        if (e) (void)hipEventDestroy(e);
    if (b->batch_done) (void)hipEventDestroy(b->batch_done);
    {
        std::lock_guard<std::recursive_mutex> lk(alloc_mutex());
        if (b->dev) (void)hipFree(b->dev);
    }
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return RIBO_OK;
}

int ribo_batch_encode_steps(RiboBatch* b, uint32_t steps) {
    if (!b) return set_error(RIBO_ERR_ARG, "null batch");
    steps = std::min(steps, RIBO_MAX_STEPS_PER_CALL);
    if (steps == 0) return 0;
    const size_t n = b->worlds.size();
    for (size_t i = 0; i < n; ++i) {
        const RiboWorld* w = b->worlds[i];
        RiboSimParams a = w->params, c = b->params0[i];
        a.epoch = c.epoch = 0;
        a.fluid_dt = c.fluid_dt = 0.0f;  // set from dt by every encode_steps call
        a.fluid_enabled = c.fluid_enabled = 0;
        if (memcmp(&a, &c, sizeof a) != 0 || memcmp(&w->st, &b->settings0[i], sizeof w->st) != 0)
            return set_error(RIBO_ERR_ARG, "a world's parameters or settings changed while it is in a batch");
        int rc = check_world(w, i ? b->worlds[0] : nullptr);
        if (rc) return rc;
    }
    hipStream_t s = b->stream;
    // after every world's queued work
    for (size_t i = 0; i < n; ++i) {
        BCHK(hipEventRecord(b->world_done[i], b->worlds[i]->stream));
        BCHK(hipStreamWaitEvent(s, b->world_done[i], 0));
    }
    RiboWorld* w0 = b->worlds[0];
    const uint32_t epoch0 = w0->params.epoch;
    launch_set_u32(s, b->ep, epoch0);
    int cur = w0->cur, tc = w0->tcur;
    uint64_t launched = 1;
    auto L = [&](BatchKernel k, const void* tab, const Shape& sh, uint32_t step, size_t lds = 0) {
        launch_batch(s, k, tab, sh.map, sh.blocks, step, lds);
        ++launched;
    };
    for (uint32_t i = 0; i < steps; ++i) {
        const uint32_t e = epoch0 + i;
        if (e % b->substep == 0) {  // fluid_step (src/world/mod.rs:1527-1588)
            L(BK_TEMPERATURE, b->temperature[cur][tc], b->fluid, i);
            tc = 1 - tc;  // copy_temperature
            L(BK_FORCES, b->forces[tc], b->fluid, i);
            L(BK_ADD_FORCES, b->add_forces, b->fluid, i);
            L(BK_DIFFUSE, b->diffuse, b->fluid, i);
            L(BK_ADVECT, b->advect, b->fluid, i);
            L(BK_VORTICITY, b->vorticity, b->fluid, i);
            L(BK_DIVERGENCE, b->divergence, b->fluid, i);
            for (uint32_t c = 0; c < b->cycles; ++c) {  // the V-cycle (:1555-1575)
                for (uint32_t d = 0; d < b->depth; ++d) L(BK_MG_DOWN, b->mgpair[d], b->mg[d], i);
                L(BK_MG_BLOCK, b->mgblock, b->one, i, b->block_lds);
                for (uint32_t d = b->depth; d-- > 0;) L(BK_MG_UP, b->mgpair[d], b->mg[d], i);
            }
            L(BK_SUBGRAD, b->subgrad, b->fluid, i);
            L(BK_NET_FLUX, b->netflux_t, b->netflux, i);
            L(BK_BOUNDARIES, b->boundaries, b->fluid, i);
            L(BK_THERMAL, b->chem[cur][tc], b->grid, i);
        }
        if (e % b->te == 0) {  // transport (:1594-1601)
            if (b->agg) L(BK_AGG, b->chem[cur][tc], b->grid, i);
            L(BK_CLEAR_NEXT, b->chem[cur][tc], b->clear, i);
            L(BK_SCATTER, b->chem[cur][tc], b->grid, i);
            cur = 1 - cur;  // transport_commit
            if (b->ledger) {
                L(BK_LEDGER, b->chem[cur][tc], b->ledger_blocks, i);
                L(BK_LEDGER_FIN, b->chem[cur][tc], b->one, i);
            }
        }
    }
    BCHK(hipGetLastError());
    b->launches += launched;
    // work queued on a world's stream from now on runs after the batch's steps
    BCHK(hipEventRecord(b->batch_done, s));
    for (RiboWorld* w : b->worlds) {
        BCHK(hipStreamWaitEvent(w->stream, b->batch_done, 0));
        w->params.epoch += steps;
        w->params.fluid_dt = w->params.dt * (float)b->substep;
        w->params.fluid_enabled = 1u;
        w->cur = cur;
        w->tcur = tc;
        w->epoch_dev_known = false;  // the world's own epoch word did not follow
        w->fluid_was_on = true;
    }
    return (int)steps;
}

int ribo_batch_run(RiboBatch* b, uint32_t steps) {
    if (!b) return set_error(RIBO_ERR_ARG, "null batch");
    uint32_t done = 0;
    while (done < steps) {
        const int k = ribo_batch_encode_steps(b, std::min(steps - done, RIBO_MAX_STEPS_PER_CALL));
        if (k < 0) return k;
        done += (uint32_t)k;
    }
    return RIBO_OK;
}

int ribo_batch_sync(RiboBatch* b) {
    if (!b) return set_error(RIBO_ERR_ARG, "null batch");
    BCHK(hipStreamSynchronize(b->stream));
    return RIBO_OK;
}

int ribo_batch_launches(const RiboBatch* b, uint64_t* launches) {
    if (!b || !launches) return set_error(RIBO_ERR_ARG, "null argument");
    *launches = b->launches;
    return RIBO_OK;
}

}  // extern "C"
