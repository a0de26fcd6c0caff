// batch.hip - many worlds stepped together with one launch per kernel.
//
// A step of the batch records the world step's kernels once each over the
// blocks of every world (launch_batch, launch_batch_tol): a block map, built
// once, tells each workgroup its world and block, and an argument table per
// kernel holds each world's pointers and parameters.  The kernel bodies are
// the single-world kernels' own, so a world in a batch computes what it
// computes alone.
//
// The schedule is World::encode_steps' (src/world/mod.rs:1406-1632) for the
// profile a batch supports: fluid every fluid_substep epochs, transport every
// transport_every epochs with the ledger check, no light step, frozen
// terrain.  The pressure solve is the reference V-cycle (mg_cycles of them)
// or the tolerance mode's conjugate gradients (one precision per batch).
//
// Worlds keep their own epochs (a device word each) and their own buffer
// parities: the tables are built relative to each world's parity at
// creation, and every world must have stepped the same way since (checked
// at every call).  Worlds at different epochs step together when they share
// the schedule's phase (always, with fluid and transport every step).

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

uint32_t substep_of(const RiboWorld* w) { return std::min(std::max(w->st.fluid_substep, 1u), 4u); }
uint32_t te_of(const RiboWorld* w) { return std::max(w->params.transport_every, 1u); }
bool tol_of(const RiboWorld* w) { return w->st.mg_tol > 0.0f; }
uint32_t gcd_u(uint32_t a, uint32_t b) { return b ? gcd_u(b, a % b) : a; }

}  // namespace

struct RiboBatch {
    std::vector<RiboWorld*> worlds;
    hipStream_t stream = nullptr;
    std::vector<hipEvent_t> world_done;  // one per world: its queued work
    hipEvent_t batch_done = nullptr;
    hipEvent_t staged = nullptr;         // the last upload from the pinned words has run
    uint32_t* words = nullptr;           // device: each world's epoch base, then its iteration slots
    uint32_t* host = nullptr;            // pinned staging of the same words
    char* dev = nullptr;                 // maps and tables, one allocation
    uint64_t launches = 0;
    uint32_t substep = 1, te = 1, period = 1, cycles = 1, depth = 0;
    bool ledger = false, agg = false, tol = false, mixed = false;
    // Pipelined batches (the reference V-cycle, no solar heating): the
    // transport chain runs on tstream beside the next fluid step, ordered by
    // events as a world stepped alone orders them (world_state.h).
    bool pipe = false;
    hipStream_t tstream = nullptr;
    hipEvent_t pe_fork = nullptr, pe_join = nullptr, pe_fluid = nullptr, pe_temp[2] = {nullptr, nullptr}, pe_vel = nullptr;
    size_t block_lds = 0, tol_lds = 0;
    std::vector<int> cur0, tcur0;        // each world's parity when the tables were built
    std::vector<const double*> cinv0;    // each world's coarsest inverse the tables point at
    // shapes (the reference V-cycle and the kernels every profile shares)
    Shape fluid, grid, netflux, clear, ledger_blocks, one;
    std::vector<Shape> mg;               // per depth d: worlds with more than d tile levels
    // shapes of the tolerance mode
    Shape cg16, cgshift, cgtile;
    std::vector<Shape> mgt;              // per depth d: worlds whose level d + 1 is a tile level
    // tables; [c][t] = chem and temperature parities relative to each world's own
    const BTemperature* temperature[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
    const BForces* forces[2] = {nullptr, nullptr};
    const BVel *add_forces = nullptr, *diffuse = nullptr, *advect = nullptr, *vorticity = nullptr;
    const BProj *divergence = nullptr, *subgrad = nullptr, *netflux_t = nullptr, *boundaries = nullptr;
    const BChem* chem[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
    std::vector<const BMgPair*> mgpair;
    const MgBlockArgs<float>* mgblock = nullptr;
    const void* cg = nullptr;            // BCg<S>[count]
    std::vector<const void*> mgpair_t;   // BMgPairT<S>[count] per depth
    const void* mgblock_t = nullptr;     // BMgBlock<S>[count]
    // captured chunks: replayed when the same chunk comes again
    struct GraphEntry {
        uint32_t m, J, phase;
        int cur, tc, cur_end, tc_end;
        uint64_t launches;
        hipGraphExec_t exec;
    };
    std::vector<GraphEntry> graphs;
    struct ChunkKey { uint32_t m, J, phase; int cur, tc; };
    std::vector<ChunkKey> seen;  // chunks recorded once, eagerly: the second time they are captured
    bool use_graphs = true;
    // what each world looked like at creation (parameters and settings must not change)
    std::vector<RiboSimParams> params0;
    std::vector<RiboWorldSettings> settings0;
    uint32_t* ep() const { return words; }
    uint32_t* nit() const { return words + worlds.size(); }
};

namespace {

int check_world(const RiboWorld* w, const RiboWorld* first) {
    const RiboWorldSettings& s = w->st;
    const RiboSimParams& p = w->params;
    if (!s.fluid_enabled || !s.multigrid || s.mg_stationary)
        return set_error(RIBO_ERR_UNSUPPORTED, "a batch runs the multigrid pressure solve: fluid and multigrid on, "
                                               "the reference V-cycle or the conjugate-gradient tolerance mode");
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
    const bool same = substep_of(w) == substep_of(first) && te_of(w) == te_of(first) && tol_of(w) == tol_of(first) &&
                      (s.ledger_check != 0) == (f.ledger_check != 0) && w->device == first->device &&
                      (tol_of(w) ? s.precision == f.precision : std::max(s.mg_cycles, 1u) == std::max(f.mg_cycles, 1u));
    if (!same)
        return set_error(RIBO_ERR_ARG, "the worlds of a batch must share fluid_substep, transport_every, the solver "
                                       "(mg_cycles, or the tolerance mode's precision), ledger_check and the device");
    return RIBO_OK;
}

// The step's parameters as encode_steps forms them (src/world/mod.rs:1431-1445).
RiboSimParams step_params(const RiboWorld* w, uint32_t substep) {
    RiboSimParams P = w->params;
    P.fluid_dt = P.dt * (float)substep;
    P.fluid_enabled = 1u;
    P.epoch = 0u;  // each launch adds its step; the world's device word adds the chunk's first epoch
    P.sun_now = P.uv_strength > 0.0f ? (P.day_period < 1.0f ? 1.0f : 0.0f) : 0.0f;
    P.sun_slope = 0.0f;
    return P;
}

// ---- the tolerance mode ----
template <class S> const std::vector<MgLevelT<S>>& levels_of(const RiboWorld* w);
template <> const std::vector<MgLevelT<float>>& levels_of<float>(const RiboWorld* w) { return w->mgz; }
template <> const std::vector<MgLevelT<bf16>>& levels_of<bf16>(const RiboWorld* w) { return w->mgh; }
// The conjugate-gradient vectors of one world, as pcg_solve forms them.
template <class S> CgVec<S> cg_vec(RiboWorld* w) {
    const std::vector<MgLevelT<S>>& lv = levels_of<S>(w);
    CgVec<S> V{};
    V.n = w->F;
    V.eps = w->levels[0].eps;
    V.fl = w->mg_fluid;
    V.x = w->p_a;
    V.xlo = w->pcg_xlo;
    V.xold = w->p_old;
    V.b = w->div;
    V.r[0] = w->pcg_r;
    V.r[1] = w->pcg_r1;
    V.s[0] = w->pcg_s0;
    V.s[1] = w->pcg_s1;
    V.pd = w->pcg_pd;
    V.w = w->pcg_w;
    V.u = lv[0].p;  // the preconditioner's level 0: z (binary32) or z16
    V.zpre = lv[0].pp;
    V.rs = sizeof(S) == 2 ? lv[0].rhs : nullptr;
    return V;
}
// First level the symmetric V-cycle solves inside one workgroup
// (vcycle_sym_inner); its tile levels are 1 .. sym_block - 1.
size_t sym_block(const RiboWorld* w) { return std::min(std::max<size_t>(w->mg_block_start, 1), w->sym_coarse); }

template <class S> struct TolTables {
    std::vector<BCg<S>> cg;
    std::vector<BMgPairT<S>> pair;  // depth-major
    std::vector<BMgBlock<S>> block;
};

template <class S>
void build_tol(RiboBatch* b, std::vector<BatchBlock>& blocks, std::vector<std::pair<Shape*, size_t>>& shape_at,
               TolTables<S>& T) {
    const uint32_t count = (uint32_t)b->worlds.size();
    auto add_shape = [&](Shape* sh, BatchTolKernel k, int depth) {
        const size_t first = blocks.size();
        for (uint32_t wi = 0; wi < count; ++wi) {
            const RiboWorld* w = b->worlds[wi];
            uint32_t n = 0, gx = 0, gy = 0;
            if (depth >= 0) {
                const size_t l = (size_t)depth + 1;
                if (l >= sym_block(w)) continue;
                n = levels_of<S>(w)[l].n;
            }
            batch_tol_grid(k, w->F, n, &gx, &gy);
            for (uint32_t by = 0; by < gy; ++by)
                for (uint32_t bx = 0; bx < gx; ++bx) blocks.push_back(BatchBlock{wi, bx, by, gx | (gy << 16)});
        }
        sh->blocks = (uint32_t)(blocks.size() - first);
        shape_at.push_back({sh, first});
    };
    uint32_t depth = 0;
    for (RiboWorld* w : b->worlds) depth = std::max<uint32_t>(depth, (uint32_t)sym_block(w) - 1u);
    add_shape(&b->cg16, BT_CG_INIT, -1);
    add_shape(&b->cgshift, BT_SHIFT, -1);
    add_shape(&b->cgtile, BT_CG_DOWN0, -1);
    b->mgt.assign(depth, Shape{});
    for (uint32_t d = 0; d < depth; ++d) add_shape(&b->mgt[d], BT_MG_DOWN, (int)d);
    T.cg.resize(count);
    T.pair.resize((size_t)depth * count);
    T.block.resize(count);
    for (uint32_t wi = 0; wi < count; ++wi) {
        RiboWorld* w = b->worlds[wi];
        const std::vector<MgLevelT<S>>& lv = levels_of<S>(w);
        uint32_t tx = 0, ty = 0;
        batch_tol_grid(BT_CG_DOWN0, w->F, 0, &tx, &ty);
        BCg<S>& c = T.cg[wi];
        c.V = cg_vec<S>(w);
        c.C = lv[1];
        c.st = w->pcg;
        c.tol = w->st.mg_tol;
        c.maxit = std::max(w->st.mg_max_cycles, 1u);
        c.nit = b->nit() + wi;
        c.nt2 = tx * ty;
        c.nb16 = ((w->F + 15u) / 16u) * ((w->F + 15u) / 16u);
        const size_t sb = sym_block(w);
        for (size_t d = 0; d + 1 < sb; ++d) T.pair[d * count + wi] = BMgPairT<S>{lv[d + 1], lv[d + 2], &w->pcg->done};
        size_t lds = 0;
        T.block[wi].a = mg_block_args<S>(lv.data() + sb, (int)(w->sym_coarse - sb + 1), w->cinv, w->ccell, w->cm, &lds,
                                         w->cmoving ? w->cm_dev : nullptr);
        T.block[wi].done = &w->pcg->done;
        b->tol_lds = std::max(b->tol_lds, lds);
    }
}

// One solve of the tolerance mode over the batch: pcg_solve's launches for
// V-cycles j = 0 .. J, J the most slots any world has for this chunk.
template <class S> uint64_t record_solve(RiboBatch* b, uint32_t J) {
    hipStream_t s = b->stream;
    uint64_t n = 0;
    auto T = [&](BatchTolKernel k, const void* tab, const Shape& sh, uint32_t j = 0, int stage = 0, int rd = 0,
                 size_t lds = 0) {
        launch_batch_tol<S>(s, k, tab, sh.map, sh.blocks, j, stage, rd, lds);
        ++n;
    };
    T(BT_CG_INIT, b->cg, b->cg16);
    T(BT_FIN_INIT, b->cg, b->one);
    T(BT_SHIFT, b->cg, b->cgshift);
    const uint32_t D = (uint32_t)b->mgt.size();
    for (uint32_t j = 0; j <= J; ++j) {
        const int stage = j == 0 ? 0 : (j == 1 ? 1 : 2);
        const int rd = j == 0 ? 1 : (int)(j & 1u);
        T(BT_CG_DOWN0, b->cg, b->cgtile, j, stage, rd);
        for (uint32_t d = 0; d < D; ++d) T(BT_MG_DOWN, b->mgpair_t[d], b->mgt[d]);
        T(BT_MG_BLOCK, b->mgblock_t, b->one, 0, 0, 0, b->tol_lds);
        for (uint32_t d = D; d-- > 0;) T(BT_MG_UP, b->mgpair_t[d], b->mgt[d]);
        T(BT_CG_UP0, b->cg, b->cgtile, j);
    }
    T(BT_CG_FINAL, b->cg, b->cg16);
    T(BT_FIN_FINAL, b->cg, b->one);
    return n;
}

// The relative parities (chem, temperature) the batch's worlds are at, or
// -1 when they disagree.
int relative_parity(const RiboBatch* b, int* c, int* t) {
    for (size_t i = 0; i < b->worlds.size(); ++i) {
        const int ci = b->worlds[i]->cur ^ b->cur0[i], ti = b->worlds[i]->tcur ^ b->tcur0[i];
        if (i == 0) {
            *c = ci;
            *t = ti;
        } else if (ci != *c || ti != *t) {
            return -1;
        }
    }
    return 0;
}

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
    b->substep = substep_of(w0);
    b->te = te_of(w0);
    b->period = b->substep / gcd_u(b->substep, b->te) * b->te;
    b->cycles = std::max(w0->st.mg_cycles, 1u);
    b->ledger = w0->st.ledger_check != 0;
    b->tol = tol_of(w0);
    b->mixed = b->tol && w0->st.precision == RIBO_PRECISION_MIXED;
    b->pipe = !b->tol;
    for (const RiboWorld* w : b->worlds) b->pipe = b->pipe && !(w->params.sun_heat > 0.0f);
    for (const RiboWorld* w : b->worlds) b->use_graphs = b->use_graphs && w->st.no_graphs == 0;
    for (RiboWorld* w : b->worlds) {
        b->agg = b->agg || w->params.aggregation > 0.0f;
        if (!b->tol) b->depth = std::max<uint32_t>(b->depth, (uint32_t)w->mg_block_start);
        b->params0.push_back(w->params);
        b->settings0.push_back(w->st);
        b->cur0.push_back(w->cur);
        b->tcur0.push_back(w->tcur);
    }
    int rc = RIBO_OK;
    // the tolerance mode's coarsest inverses, built as a world's own first solve builds them
    for (size_t i = 0; b->tol && i < b->worlds.size() && !rc; ++i) rc = batch_prepare_solve(b->worlds[i]);
    for (RiboWorld* w : b->worlds) b->cinv0.push_back(w->cinv);
    hipError_t he = hipSuccess;
    if (!rc) {
        std::lock_guard<std::recursive_mutex> lk(alloc_mutex());
        he = hipMalloc((void**)&b->words, 2 * (size_t)count * sizeof(uint32_t));
        if (he == hipSuccess) he = hipHostMalloc((void**)&b->host, 2 * (size_t)count * sizeof(uint32_t));
        if (he != hipSuccess) rc = set_error(RIBO_ERR_OOM, std::string("batch words: ") + hipGetErrorString(he));
    }
    if (rc) {
        ribo_batch_destroy(b);
        return rc;
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
    TolTables<float> tf;
    TolTables<bf16> th;
    if (b->tol && b->mixed) build_tol<bf16>(b, blocks, shape_at, th);
    else if (b->tol) build_tol<float>(b, blocks, shape_at, tf);
    // ---- argument tables ----
    std::vector<BTemperature> t_temp(4 * (size_t)count);
    std::vector<BForces> t_forces(2 * (size_t)count);
    std::vector<BVel> t_vel(4 * (size_t)count);    // add_forces, diffuse, advect, vorticity
    std::vector<BProj> t_proj(4 * (size_t)count);  // divergence, subgrad, net_flux, boundaries
    std::vector<BChem> t_chem(4 * (size_t)count);
    std::vector<BMgPair> t_mg((size_t)b->depth * count);
    std::vector<MgBlockArgs<float>> t_block(b->tol ? 0 : count);
    for (uint32_t wi = 0; wi < count; ++wi) {
        RiboWorld* w = b->worlds[wi];
        Geo g = w->geo();
        g.sun = nullptr;  // the sun of every step is step_params' (no table: a moving sun is refused)
        const RiboSimParams P = step_params(w, b->substep);
        const uint32_t* ep = b->ep() + wi;
        for (int t = 0; t < 2; ++t) {
            const int tw = b->tcur0[wi] ^ t;
            for (int c = 0; c < 2; ++c) {
                const int cw = b->cur0[wi] ^ c;
                t_temp[(size_t)(c * 2 + t) * count + wi] = BTemperature{
                    g, P, ep, w->vel_a, w->temp[tw], w->redox[tw], w->temp[1 - tw], w->redox[1 - tw], w->eaten, w->chem[cw]};
                t_chem[(size_t)(c * 2 + t) * count + wi] = BChem{
                    g, P, ep, w->chem[cw], w->chem[1 - cw], b->pipe ? w->vel_s : w->vel_a, w->temp[tw], w->agg_e, w->ledger,
                    std::min<uint32_t>((w->G * w->G + 255u) / 256u, LEDGER_MAX_BLOCKS)};
            }
            t_forces[t * (size_t)count + wi] = BForces{g, P, ep, w->temp[tw], w->forces};
        }
        const float decay = step_decay_factor(P.fluid_decay, P.fluid_dt);
        t_vel[0 * (size_t)count + wi] = BVel{g, P, decay, w->forces, w->vel_a, w->vel_b};
        t_vel[1 * (size_t)count + wi] = BVel{g, P, decay, nullptr, w->vel_b, w->vel_a};
        t_vel[2 * (size_t)count + wi] = BVel{g, P, decay, nullptr, w->vel_a, w->vel_b};
        t_vel[3 * (size_t)count + wi] = BVel{g, P, decay, nullptr, w->vel_b, w->vel_a};
        t_proj[0 * (size_t)count + wi] = BProj{g, w->vel_a, nullptr, w->div, w->p_a};
        t_proj[1 * (size_t)count + wi] = BProj{g, w->vel_a, w->vel_b, w->p_a, nullptr};
        t_proj[2 * (size_t)count + wi] = BProj{g, w->vel_b, nullptr, nullptr, w->net_flux};
        t_proj[3 * (size_t)count + wi] = BProj{g, w->vel_b, w->vel_a, nullptr, w->net_flux, b->pipe ? w->vel_s : nullptr};
        for (uint32_t d = 0; d < b->depth; ++d)
            if (d < w->mg_block_start) t_mg[(size_t)d * count + wi] = BMgPair{w->mg[d], w->mg[d + 1]};
        if (!b->tol) {
            size_t lds = 0;
            const size_t bs = w->mg_block_start;
            t_block[wi] = mg_block_args<float>(w->mg.data() + bs, (int)(w->mg.size() - bs), nullptr, nullptr, 0, &lds);
            b->block_lds = std::max(b->block_lds, lds);
        }
    }
    // ---- one device allocation: maps and tables ----
    size_t bytes = 0;
    struct Up { size_t at; const void* src; size_t n; };
    std::vector<Up> ups;
    auto put = [&](const void* src, size_t n) {
        const size_t at = bytes;
        bytes += (n + 255) & ~(size_t)255;
        ups.push_back(Up{at, src, n});
        return at;
    };
    const size_t o_map = put(blocks.data(), blocks.size() * sizeof(BatchBlock));
    const size_t o_temp = put(t_temp.data(), t_temp.size() * sizeof(BTemperature));
    const size_t o_forces = put(t_forces.data(), t_forces.size() * sizeof(BForces));
    const size_t o_vel = put(t_vel.data(), t_vel.size() * sizeof(BVel));
    const size_t o_proj = put(t_proj.data(), t_proj.size() * sizeof(BProj));
    const size_t o_chem = put(t_chem.data(), t_chem.size() * sizeof(BChem));
    const size_t o_mg = put(t_mg.data(), t_mg.size() * sizeof(BMgPair));
    const size_t o_block = put(t_block.data(), t_block.size() * sizeof(MgBlockArgs<float>));
    size_t o_cg = 0, o_pair = 0, o_tblock = 0, pair_size = 0;
    if (b->tol && b->mixed) {
        o_cg = put(th.cg.data(), th.cg.size() * sizeof(BCg<bf16>));
        o_pair = put(th.pair.data(), th.pair.size() * sizeof(BMgPairT<bf16>));
        o_tblock = put(th.block.data(), th.block.size() * sizeof(BMgBlock<bf16>));
        pair_size = sizeof(BMgPairT<bf16>);
    } else if (b->tol) {
        o_cg = put(tf.cg.data(), tf.cg.size() * sizeof(BCg<float>));
        o_pair = put(tf.pair.data(), tf.pair.size() * sizeof(BMgPairT<float>));
        o_tblock = put(tf.block.data(), tf.block.size() * sizeof(BMgBlock<float>));
        pair_size = sizeof(BMgPairT<float>);
    }
    {
        std::lock_guard<std::recursive_mutex> lk(alloc_mutex());
        he = hipMalloc((void**)&b->dev, std::max<size_t>(bytes, 256));
    }
    if (he == hipSuccess) he = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    for (const Up& u : ups)
        if (he == hipSuccess && u.n) he = hipMemcpyAsync(b->dev + u.at, u.src, u.n, hipMemcpyHostToDevice, b->stream);
    if (he == hipSuccess) he = hipStreamSynchronize(b->stream);
    if (he == hipSuccess) he = hipEventCreateWithFlags(&b->batch_done, hipEventDisableTiming);
    if (he == hipSuccess) he = hipEventCreateWithFlags(&b->staged, hipEventDisableTiming);
    if (he == hipSuccess) he = hipEventRecord(b->staged, b->stream);
    if (he == hipSuccess && b->pipe) he = hipStreamCreateWithFlags(&b->tstream, hipStreamNonBlocking);
    if (b->pipe)
        for (hipEvent_t* e : {&b->pe_fork, &b->pe_join, &b->pe_fluid, &b->pe_temp[0], &b->pe_temp[1], &b->pe_vel})
            if (he == hipSuccess) he = hipEventCreateWithFlags(e, hipEventDisableTiming);
    b->world_done.assign(count, nullptr);
    for (uint32_t i = 0; i < count && he == hipSuccess; ++i)
        he = hipEventCreateWithFlags(&b->world_done[i], hipEventDisableTiming);
    if (he != hipSuccess) {
        rc = set_error(RIBO_ERR_HIP, std::string("batch set-up: ") + hipGetErrorString(he));
        ribo_batch_destroy(b);
        return rc;
    }
    const BatchBlock* map = (const BatchBlock*)(b->dev + o_map);
    for (auto& sa : shape_at) sa.first->map = map + sa.second;
    for (int t = 0; t < 2; ++t) {
        for (int c = 0; c < 2; ++c) {
            b->temperature[c][t] = (const BTemperature*)(b->dev + o_temp) + (size_t)(c * 2 + t) * count;
            b->chem[c][t] = (const BChem*)(b->dev + o_chem) + (size_t)(c * 2 + t) * count;
        }
        b->forces[t] = (const BForces*)(b->dev + o_forces) + (size_t)t * count;
    }
    const BVel* tv = (const BVel*)(b->dev + o_vel);
    b->add_forces = tv;
    b->diffuse = tv + count;
    b->advect = tv + 2 * (size_t)count;
    b->vorticity = tv + 3 * (size_t)count;
    const BProj* tp = (const BProj*)(b->dev + o_proj);
    b->divergence = tp;
    b->subgrad = tp + count;
    b->netflux_t = tp + 2 * (size_t)count;
    b->boundaries = tp + 3 * (size_t)count;
    for (uint32_t d = 0; d < b->depth; ++d) b->mgpair.push_back((const BMgPair*)(b->dev + o_mg) + (size_t)d * count);
    b->mgblock = (const MgBlockArgs<float>*)(b->dev + o_block);
    if (b->tol) {
        b->cg = b->dev + o_cg;
        for (size_t d = 0; d < b->mgt.size(); ++d) b->mgpair_t.push_back(b->dev + o_pair + d * count * pair_size);
        b->mgblock_t = b->dev + o_tblock;
    }
    *out = b;
    return RIBO_OK;
}

int ribo_batch_destroy(RiboBatch* b) {
    if (!b) return set_error(RIBO_ERR_ARG, "null batch");
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->tstream) (void)hipStreamSynchronize(b->tstream);
    for (auto& g : b->graphs) (void)hipGraphExecDestroy(g.exec);
    for (hipEvent_t e : {b->pe_fork, b->pe_join, b->pe_fluid, b->pe_temp[0], b->pe_temp[1], b->pe_vel})
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : b->world_done)
        if (e) (void)hipEventDestroy(e);
    if (b->batch_done) (void)hipEventDestroy(b->batch_done);
    if (b->staged) (void)hipEventDestroy(b->staged);
    {
        std::lock_guard<std::recursive_mutex> lk(alloc_mutex());
        if (b->dev) (void)hipFree(b->dev);
        if (b->words) (void)hipFree(b->words);
        if (b->host) (void)hipHostFree(b->host);
    }
    if (b->stream) (void)hipStreamDestroy(b->stream);
    if (b->tstream) (void)hipStreamDestroy(b->tstream);
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
        if (w->params.epoch % b->period != b->worlds[0]->params.epoch % b->period)
            return set_error(RIBO_ERR_ARG, "the worlds of a batch must be at the same phase of the fluid and transport "
                                           "schedule (epoch modulo its period)");
    }
    int rc_c = 0, rc_t = 0;
    if (relative_parity(b, &rc_c, &rc_t) != 0)
        return set_error(RIBO_ERR_ARG, "a world was stepped on its own by a different count of fluid or transport "
                                       "steps than the others since the batch was built; build a new batch");
    int cur = rc_c, tc = rc_t;
    hipStream_t s = b->stream;
    // after every world's queued work
    for (size_t i = 0; i < n; ++i) {
        BCHK(hipEventRecord(b->world_done[i], b->worlds[i]->stream));
        BCHK(hipStreamWaitEvent(s, b->world_done[i], 0));
    }
    const uint32_t e0 = b->worlds[0]->params.epoch;
    uint64_t launched = 0;
    hipStream_t ls = s;  // the stream the next launches go to
    auto L = [&](BatchKernel k, const void* tab, const Shape& sh, uint32_t step, size_t lds = 0) {
        launch_batch(ls, k, tab, sh.map, sh.blocks, step, lds);
        ++launched;
    };
    if (b->pipe)  // the transport's copy of each world's velocity, until a fluid step writes it
        for (size_t i = 0; i < n; ++i) {
            const RiboWorld* w = b->worlds[i];
            BCHK(hipMemcpyAsync(w->vel_s, w->vel_a, sizeof(float2) * (size_t)w->F * w->F, hipMemcpyDeviceToDevice, s));
        }
    bool temp_read[2] = {false, false}, vel_read = false;
    const uint32_t chunk = b->tol ? TOL_CHUNK : RIBO_MAX_STEPS_PER_CALL;
    for (uint32_t first = 0; first < steps; first += chunk) {
        const uint32_t m = std::min(chunk, steps - first);
        // the words of this chunk: each world's epoch base and, in the
        // tolerance mode, its iteration slots (chosen after the previous
        // chunk's solves finished, as a world stepped alone chooses them)
        uint32_t J = 0;
        if (b->tol) {
            BCHK(hipStreamSynchronize(s));
            for (size_t i = 0; i < n; ++i) {
                RiboWorld* w = b->worlds[i];
                const int rc = batch_prepare_solve(w);
                if (rc) return rc;
                if (w->cinv != b->cinv0[i])
                    return set_error(RIBO_ERR_ARG, "a world's terrain was replaced while it is in a batch; build a new batch");
                const uint32_t slots = std::max(std::min(w->pcg_slots, std::max(w->st.mg_max_cycles, 1u)), 1u);
                b->host[n + i] = slots;
                J = std::max(J, slots);
            }
        }
        BCHK(hipEventSynchronize(b->staged));  // the pinned words are free again
        for (size_t i = 0; i < n; ++i) b->host[i] = b->worlds[i]->params.epoch + first;
        BCHK(hipMemcpyAsync(b->words, b->host, (b->tol ? 2 : 1) * n * sizeof(uint32_t), hipMemcpyHostToDevice, s));
        BCHK(hipEventRecord(b->staged, s));
        ++launched;
        // the chunk's launches: replayed from a graph captured the first time
        // the same chunk (length, slots, parities, schedule phase) comes
        const uint32_t phase = (e0 + first) % b->period;
        RiboBatch::GraphEntry* ge = nullptr;
        if (b->use_graphs)
            for (auto& g : b->graphs)
                if (g.m == m && g.J == J && g.phase == phase && g.cur == cur && g.tc == tc) ge = &g;
        if (ge) {
            BCHK(hipGraphLaunch(ge->exec, s));
            launched += ge->launches;
            cur = ge->cur_end;
            tc = ge->tc_end;
            continue;
        }
        const int cur_start = cur, tc_start = tc;
        const uint64_t before = launched;
        // a chunk is captured the second time it comes (capture and
        // instantiation cost more than one eager recording)
        bool capture = false;
        if (b->use_graphs) {
            for (const auto& k : b->seen)
                if (k.m == m && k.J == J && k.phase == phase && k.cur == cur && k.tc == tc) capture = true;
            if (!capture) {
                if (b->seen.size() >= 256) b->seen.clear();
                b->seen.push_back(RiboBatch::ChunkKey{m, J, phase, cur, tc});
            }
        }
        std::unique_lock<std::recursive_mutex> cap_lock(alloc_mutex(), std::defer_lock);
        if (capture) {
            cap_lock.lock();
            if (b->graphs.size() >= 64) {
                for (auto& g : b->graphs) (void)hipGraphExecDestroy(g.exec);
                b->graphs.clear();
            }
            BCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        }
        if (b->pipe) {  // the transport stream starts after everything before
            BCHK(hipEventRecord(b->pe_fork, s));
            BCHK(hipStreamWaitEvent(b->tstream, b->pe_fork, 0));
            temp_read[0] = temp_read[1] = vel_read = false;
        }
        for (uint32_t i = 0; i < m; ++i) {
            const uint32_t e = e0 + first + i;
            if (e % b->substep == 0) {  // fluid_step (src/world/mod.rs:1527-1588)
                ls = s;
                // the temperature buffer rewritten now was read by the transport before last
                if (b->pipe && temp_read[1 - tc]) BCHK(hipStreamWaitEvent(s, b->pe_temp[1 - tc], 0));
                L(BK_TEMPERATURE, b->temperature[cur][tc], b->fluid, i);
                tc = 1 - tc;  // copy_temperature
                L(BK_FORCES, b->forces[tc], b->fluid, i);
                L(BK_ADD_FORCES, b->add_forces, b->fluid, i);
                L(BK_DIFFUSE, b->diffuse, b->fluid, i);
                L(BK_ADVECT, b->advect, b->fluid, i);
                L(BK_VORTICITY, b->vorticity, b->fluid, i);
                L(BK_DIVERGENCE, b->divergence, b->fluid, i);
                if (b->tol) {  // the tolerance mode (pcg_solve)
                    launched += b->mixed ? record_solve<bf16>(b, J) : record_solve<float>(b, J);
                } else {
                    for (uint32_t c = 0; c < b->cycles; ++c) {  // the V-cycle (:1555-1575)
                        for (uint32_t d = 0; d < b->depth; ++d) L(BK_MG_DOWN, b->mgpair[d], b->mg[d], i);
                        L(BK_MG_BLOCK, b->mgblock, b->one, i, b->block_lds);
                        for (uint32_t d = b->depth; d-- > 0;) L(BK_MG_UP, b->mgpair[d], b->mg[d], i);
                    }
                }
                L(BK_SUBGRAD, b->subgrad, b->fluid, i);
                L(BK_NET_FLUX, b->netflux_t, b->netflux, i);
                if (b->pipe && vel_read) BCHK(hipStreamWaitEvent(s, b->pe_vel, 0));  // the velocity copy's last reader
                L(BK_BOUNDARIES, b->boundaries, b->fluid, i);
                if (b->pipe) {  // the thermal activation follows on the transport stream
                    BCHK(hipEventRecord(b->pe_fluid, s));
                    BCHK(hipStreamWaitEvent(b->tstream, b->pe_fluid, 0));
                    ls = b->tstream;
                }
                L(BK_THERMAL, b->chem[cur][tc], b->grid, i);
                if (b->pipe) {
                    BCHK(hipEventRecord(b->pe_temp[tc], b->tstream));
                    temp_read[tc] = true;
                }
            }
            if (e % b->te == 0) {  // transport (:1594-1601)
                ls = b->pipe ? b->tstream : s;
                if (b->agg) L(BK_AGG, b->chem[cur][tc], b->grid, i);
                L(BK_CLEAR_NEXT, b->chem[cur][tc], b->clear, i);
                L(BK_SCATTER, b->chem[cur][tc], b->grid, i);
                if (b->pipe) {
                    BCHK(hipEventRecord(b->pe_temp[tc], b->tstream));
                    BCHK(hipEventRecord(b->pe_vel, b->tstream));
                    temp_read[tc] = vel_read = true;
                }
                cur = 1 - cur;  // transport_commit
                if (b->ledger) {
                    L(BK_LEDGER, b->chem[cur][tc], b->ledger_blocks, i);
                    L(BK_LEDGER_FIN, b->chem[cur][tc], b->one, i);
                }
            }
        }
        ls = s;
        if (b->pipe) {  // ... and everything after waits for it
            BCHK(hipEventRecord(b->pe_join, b->tstream));
            BCHK(hipStreamWaitEvent(s, b->pe_join, 0));
        }
        if (capture) {
            hipGraph_t graph = nullptr;
            const hipError_t ce = hipStreamEndCapture(s, &graph);
            if (ce != hipSuccess) return set_error(RIBO_ERR_HIP, std::string("batch capture: ") + hipGetErrorString(ce));
            RiboBatch::GraphEntry g{m, J, phase, cur_start, tc_start, cur, tc, launched - before, nullptr};
            const hipError_t ie = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (ie != hipSuccess) return set_error(RIBO_ERR_HIP, std::string("batch graph: ") + hipGetErrorString(ie));
            cap_lock.unlock();
            b->graphs.push_back(g);
            BCHK(hipGraphLaunch(g.exec, s));
        }
    }
    BCHK(hipGetLastError());
    b->launches += launched;
    // work queued on a world's stream from now on runs after the batch's steps
    BCHK(hipEventRecord(b->batch_done, s));
    for (size_t i = 0; i < n; ++i) {
        RiboWorld* w = b->worlds[i];
        BCHK(hipStreamWaitEvent(w->stream, b->batch_done, 0));
        w->params.epoch += steps;
        w->params.fluid_dt = w->params.dt * (float)b->substep;
        w->params.fluid_enabled = 1u;
        w->cur = b->cur0[i] ^ cur;
        w->tcur = b->tcur0[i] ^ tc;
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

int ribo_batch_accepts(const RiboWorld* w) {
    if (!w) return set_error(RIBO_ERR_ARG, "null world");
    return check_world(w, nullptr) == RIBO_OK ? 1 : 0;
}

int ribo_batch_idle(RiboBatch* b) {
    if (!b) return set_error(RIBO_ERR_ARG, "null batch");
    const hipError_t e = hipStreamQuery(b->stream);
    if (e == hipSuccess) return 1;
    if (e == hipErrorNotReady) return 0;
    return set_error(RIBO_ERR_HIP, std::string("batch stream: ") + hipGetErrorString(e));
}

int ribo_batch_launches(const RiboBatch* b, uint64_t* launches) {
    if (!b || !launches) return set_error(RIBO_ERR_ARG, "null argument");
    *launches = b->launches;
    return RIBO_OK;
}

}  // extern "C"
