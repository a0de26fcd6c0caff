// world.hip - the world object, the per-step schedule and the C ABI of
// libribo_world.so.
//
// The host side mirrors World (src/world/mod.rs): it owns every hot-path
// buffer in HBM, keeps params/settings, builds the multigrid level table
// (:635-671) and records the steps of encode_steps (:1406-1632) on one HIP
// stream per world, with the reference's epoch, fluid_substep and
// transport_every semantics.  Buffers that the reference copies between
// steps ping-pong instead (chem/chem_next, temperature, reductant).

#include <hip/hip_runtime.h>
#include <math.h>
#include <stddef.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <string>
#include <vector>

#include "../../include/ribo_params_default.h"
#include "../../include/ribo_world.h"
#include "kernels.h"
#include "world_state.h"

using namespace ribo;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define HIPCHK(expr)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess) return fail(RIBO_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

constexpr uint32_t LIGHT_DIV = 2u;  // src/shaders.rs:132
constexpr float MG_EPS = 1e-4f;     // src/world/mod.rs:202
constexpr uint32_t MG_COARSE = 30;  // src/world/mod.rs:198
constexpr float MAX_DT = 0.07f;     // fluid.wgsl:15

// Host-side per-step values (src/params.rs:505-537), binary64 as the reference.
void day_phase(const RiboSimParams& p, uint32_t epoch, bool* day, double* phase) {
    double ph = (double)epoch / (double)p.day_period;
    ph = ph - floor(ph);
    double frac = std::min(std::max((double)p.day_fraction, 0.05), 1.0);
    *day = ph < frac;
    *phase = ph / frac;
}
float day_light(const RiboSimParams& p, uint32_t epoch) {
    if (p.day_period < 1.0f) return 1.0f;
    bool day;
    double ph;
    day_phase(p, epoch, &day, &ph);
    if (!day) return 0.0f;
    return (float)std::max(sin(M_PI * ph), 0.0);
}
float sun_slope_at(const RiboSimParams& p, uint32_t epoch) {
    if (p.day_period < 1.0f) return 0.0f;
    bool day;
    double ph;
    day_phase(p, epoch, &day, &ph);
    double max_angle = std::min(std::max((double)p.sun_angle, 0.0), 85.0) * (M_PI / 180.0);
    if (!day) ph = 1.0;
    double angle = (ph - 0.5) * 2.0 * max_angle;
    return (float)tan(angle);
}
// pow(clamp(fluid_decay, 0, 1), dt * 60) of advect_velocity (fluid.wgsl:473),
// uniform over a step, evaluated once on the host in binary64.
float decay_factor(float decay, float fluid_dt_param) {
    float dt = std::min(std::max(fluid_dt_param, 0.0f), MAX_DT);
    float d = std::min(std::max(decay, 0.0f), 1.0f);
    return (float)pow((double)d, (double)(dt * 60.0f));
}

}  // namespace



namespace {

// Bytes of device memory allocated while an AllocCount is alive on this
// thread (the world's footprint, by part).
thread_local size_t* g_alloc_acc = nullptr;
struct AllocCount {
    size_t* saved;
    explicit AllocCount(size_t* acc) : saved(g_alloc_acc) { *acc = 0; g_alloc_acc = acc; }
    ~AllocCount() { g_alloc_acc = saved; }
};

// Allocation and release of device memory, and graph capture, are serialised
// across the threads of the process: a world created or destroyed on one
// thread while another thread captures a graph would otherwise touch the
// capturing stream through the runtime's implicit synchronisation.
std::recursive_mutex& device_mutex() {
    static std::recursive_mutex m;
    return m;
}

// Zeroed device allocation; the clear is ordered on the world's stream (no
// legacy-stream operation).
template <class T> int dalloc(hipStream_t s, T** p, size_t count) {
    std::lock_guard<std::recursive_mutex> lk(device_mutex());
    if (count == 0) count = 1;
    if (g_alloc_acc) *g_alloc_acc += count * sizeof(T);
    hipError_t e = hipMalloc((void**)p, count * sizeof(T));
    if (e != hipSuccess) return fail(RIBO_ERR_OOM, std::string("hipMalloc: ") + hipGetErrorString(e));
    e = hipMemsetAsync(*p, 0, count * sizeof(T), s);
    if (e != hipSuccess) return fail(RIBO_ERR_HIP, std::string("hipMemset: ") + hipGetErrorString(e));
    return RIBO_OK;
}
template <class T> void dfree(T*& p) {
    std::lock_guard<std::recursive_mutex> lk(device_mutex());
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}

hipEvent_t get_event(RiboWorld* w) {
    if (!w->event_pool.empty()) {
        hipEvent_t e = w->event_pool.back();
        w->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}

void resolve_timers(RiboWorld* w) {
    for (auto& p : w->pending) {
        float ms = 0;
        if (hipEventSynchronize(p.b) == hipSuccess && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            w->timers[p.timer].total_ms += ms;
            w->timers[p.timer].launches += 1;
            w->timers[p.timer].each.push_back(ms);
        }
        w->event_pool.push_back(p.a);
        w->event_pool.push_back(p.b);
    }
    w->pending.clear();
}

// Records `fn` (one kernel launch) and, when timing is on, an event pair
// around it on the world's stream.
template <class Fn> void run(RiboWorld* w, const char* name, Fn&& fn) {
    if (!w->timing) {
        fn();
        return;
    }
    size_t t = 0;
    for (; t < w->timers.size(); ++t)
        if (w->timers[t].name == name) break;
    if (t == w->timers.size()) w->timers.push_back(Timer{name});
    hipEvent_t a = get_event(w), b = get_event(w);
    (void)hipEventRecord(a, w->stream);
    fn();
    (void)hipEventRecord(b, w->stream);
    w->pending.push_back({t, a, b});
    if (w->pending.size() > 4096) resolve_timers(w);
}

// Multigrid level table (src/world/mod.rs:637-671): sizes halve while > min
// and even; eps_l = MG_EPS * 2^l.
std::vector<RiboMgLevel> make_levels(uint32_t n0, uint32_t min_size) {
    std::vector<uint32_t> sizes{n0};
    while (sizes.back() > min_size && sizes.back() % 2 == 0) sizes.push_back(sizes.back() / 2);
    std::vector<uint32_t> offs{0};
    for (uint32_t n : sizes) offs.push_back(offs.back() + n * n);
    std::vector<RiboMgLevel> lv;
    for (size_t l = 0; l < sizes.size(); ++l) {
        RiboMgLevel L{};
        L.n = sizes[l];
        L.off = offs[l];
        L.n_c = l + 1 < sizes.size() ? sizes[l + 1] : 0;
        L.off_c = offs[l + 1];
        L.eps = MG_EPS * (float)(1u << l);
        lv.push_back(L);
    }
    return lv;
}

void clear_graphs(RiboWorld* w) {
    for (auto& e : w->graphs) (void)hipGraphExecDestroy(e.exec);
    w->graphs.clear();
}

int build_mg(RiboWorld* w) {
    AllocCount count(&w->mg_bytes);
    clear_graphs(w);
    dfree(w->mg_p);
    dfree(w->mg_pp);
    dfree(w->mg_rhs);
    dfree(w->mg_fluid);
    w->levels = make_levels(w->F, std::max(w->st.mg_min_size, 1u));
    w->mg_tile.assign(w->levels.size(), 0);
    size_t total = w->levels.back().off_c;
    size_t fine = (size_t)w->F * w->F;
    size_t coarse = total - fine;
    int rc;
    if ((rc = dalloc(w->stream, &w->mg_p, coarse)) || (rc = dalloc(w->stream, &w->mg_pp, total)) || (rc = dalloc(w->stream, &w->mg_rhs, coarse)) ||
        (rc = dalloc(w->stream, &w->mg_fluid, total)))
        return rc;
    dfree(w->mg_p16);
    dfree(w->mg_pp16);
    dfree(w->mg_rhs16);
    if ((rc = dalloc(w->stream, &w->mg_p16, coarse)) || (rc = dalloc(w->stream, &w->mg_pp16, total)) || (rc = dalloc(w->stream, &w->mg_rhs16, coarse))) return rc;
    w->mg.clear();
    w->mgz.clear();
    w->mgh.clear();
    for (size_t l = 0; l < w->levels.size(); ++l) {
        MgLevelDev d;
        d.n = w->levels[l].n;
        d.eps = w->levels[l].eps;
        MgLevelT<bf16> h;
        h.n = d.n;
        h.eps = d.eps;
        if (l == 0) {
            d.p = w->p_a;
            d.rhs = w->div;
            h.p = w->pcg_z16;
            h.rhs = w->pcg_r16;
        } else {
            d.p = w->mg_p + (w->levels[l].off - fine);
            d.rhs = w->mg_rhs + (w->levels[l].off - fine);
            h.p = w->mg_p16 + (w->levels[l].off - fine);
            h.rhs = w->mg_rhs16 + (w->levels[l].off - fine);
        }
        d.pp = w->mg_pp + w->levels[l].off;
        h.pp = w->mg_pp16 + w->levels[l].off;
        d.fluid = w->mg_fluid + w->levels[l].off;
        h.fluid = d.fluid;
        w->mg.push_back(d);
        MgLevelDev z = d;
        if (l == 0) {
            z.p = w->pcg_z;
            z.rhs = w->pcg_r;
        }
        w->mgz.push_back(z);
        w->mgh.push_back(h);
    }
    w->cinv_valid = false;
    // the longest suffix of levels (at most 8) whose unknowns, right-hand
    // sides and masks fit in one workgroup's LDS together
    std::vector<uint32_t> sizes;
    for (auto& L : w->levels) sizes.push_back(L.n);
    size_t b = w->mg.size();
    while (b > 0 && w->mg.size() - (b - 1) <= 8 &&
           mg_block_lds_bytes(sizes.data() + (b - 1), (int)(w->mg.size() - (b - 1))) <= w->lds_limit)
        --b;
    w->mg_block_start = b;
    w->sym_coarse = w->levels.size() - 1;
    return RIBO_OK;
}

// Terrain-derived data, rebuilt whenever gamma or the level table changes.
int rebuild_terrain(RiboWorld* w) {
    Geo g = w->geo();
    launch_slope(w->stream, g, w->slope);
    launch_fluid_static(w->stream, g, w->cflags, w->fill, w->dragg, w->slope_f);
    launch_fluid_normals(w->stream, g, w->cflags, w->normal);
    launch_mg_mask0(w->stream, g, w->mg_fluid);
    for (size_t l = 1; l < w->levels.size(); ++l)
        launch_mg_mask_coarsen(w->stream, w->levels[l].n, w->mg_fluid + w->levels[l - 1].off, w->mg_fluid + w->levels[l].off);
    HIPCHK(hipGetLastError());
    w->cinv_valid = false;
    return RIBO_OK;
}

// Whether the coarsest inverse follows a moving terrain: the conjugate-
// gradient tolerance mode with terrain_enabled.
bool coarse_moving(const RiboWorld* w) {
    return w->st.terrain_enabled && w->st.multigrid && w->st.mg_tol > 0.0f && !w->st.mg_stationary;
}

// Inverse of the coarsest level's eps-anchored operator over its fluid
// cells, in binary64 (Gauss-Jordan with partial pivoting on the host), for
// the exact coarse solve of the tolerance mode.  Skipped above 1024 unknowns.
// Under a moving terrain the buffers take every cell of the coarsest level
// (at most MOVING_COARSE_MAX), with the mask the inverse was built for and
// its unknown count on the device: k_coarse_refresh rebuilds the inverse
// there, in the same operations, whenever a fluid step changes that mask.
constexpr uint32_t MOVING_COARSE_MAX = 256;
int build_coarse_inverse(RiboWorld* w) {
    const RiboMgLevel& C = w->levels[w->sym_coarse];
    const uint32_t n = C.n;
    AllocCount count(&w->cinv_bytes);
    dfree(w->cinv);
    dfree(w->ccell);
    dfree(w->csnap);
    dfree(w->cwork);
    dfree(w->cm_dev);
    w->cm = 0;
    w->cmoving = coarse_moving(w);
    w->ccap = 0;
    if (w->cmoving) {
        // buffers for every cell of the coarsest level and an invalid mask
        // snapshot: the first fluid step's k_coarse_refresh builds the inverse
        const uint32_t cap = n * n;
        if (cap > MOVING_COARSE_MAX)
            return fail(RIBO_ERR_UNSUPPORTED, "moving terrain with the tolerance mode needs a coarsest multigrid level of "
                                              "at most 256 cells (lower mg_min_size)");
        int rc;
        if ((rc = dalloc(w->stream, &w->cinv, (size_t)cap * cap)) || (rc = dalloc(w->stream, &w->ccell, cap)) ||
            (rc = dalloc(w->stream, &w->csnap, cap)) || (rc = dalloc(w->stream, &w->cwork, 2 * (size_t)cap * cap)) ||
            (rc = dalloc(w->stream, &w->cm_dev, 1)))
            return rc;
        w->ccap = cap;
        HIPCHK(hipMemsetAsync(w->csnap, 0xFF, cap, w->stream));
        HIPCHK(hipMemsetAsync(w->cm_dev, 0, sizeof(int), w->stream));
        HIPCHK(hipStreamSynchronize(w->stream));
        w->cinv_valid = true;
        return RIBO_OK;
    }
    std::vector<uint8_t> fl((size_t)n * n);
    HIPCHK(hipMemcpyAsync(fl.data(), w->mg_fluid + C.off, fl.size(), hipMemcpyDeviceToHost, w->stream));
    HIPCHK(hipStreamSynchronize(w->stream));
    std::vector<int32_t> idx((size_t)n * n, -1), cell;
    for (uint32_t i = 0; i < n * n; ++i)
        if (fl[i]) { idx[i] = (int32_t)cell.size(); cell.push_back((int32_t)i); }
    const size_t m = cell.size();
    if (m == 0 || m > 1024) {
        w->cinv_valid = true;
        return RIBO_OK;
    }
    std::vector<double> a(m * m, 0.0), inv(m * m, 0.0);
    for (size_t r = 0; r < m; ++r) {
        const uint32_t c = (uint32_t)cell[r], x = c % n, y = c / n;
        int k = 0;
        const int32_t nb[4] = {x > 0 ? idx[c - 1] : -1, x + 1 < n ? idx[c + 1] : -1, y > 0 ? idx[c - n] : -1,
                               y + 1 < n ? idx[c + n] : -1};
        for (int d = 0; d < 4; ++d)
            if (nb[d] >= 0) { a[r * m + (size_t)nb[d]] = 1.0; ++k; }
        a[r * m + r] = -((double)k + (double)C.eps);
        inv[r * m + r] = 1.0;
    }
    for (size_t c = 0; c < m; ++c) {
        size_t piv = c;
        for (size_t r = c + 1; r < m; ++r)
            if (fabs(a[r * m + c]) > fabs(a[piv * m + c])) piv = r;
        if (piv != c)
            for (size_t j = 0; j < m; ++j) { std::swap(a[c * m + j], a[piv * m + j]); std::swap(inv[c * m + j], inv[piv * m + j]); }
        const double dg = a[c * m + c];
        for (size_t j = 0; j < m; ++j) { a[c * m + j] /= dg; inv[c * m + j] /= dg; }
        for (size_t r = 0; r < m; ++r) {
            if (r == c) continue;
            const double f = a[r * m + c];
            if (f == 0.0) continue;
            for (size_t j = 0; j < m; ++j) { a[r * m + j] -= f * a[c * m + j]; inv[r * m + j] -= f * inv[c * m + j]; }
        }
    }
    int rc;
    if ((rc = dalloc(w->stream, &w->cinv, m * m)) || (rc = dalloc(w->stream, &w->ccell, m))) return rc;
    // by columns, so that the rows a workgroup computes read consecutive words
    std::vector<double> inv_t(m * m);
    for (size_t r = 0; r < m; ++r)
        for (size_t j = 0; j < m; ++j) inv_t[j * m + r] = inv[r * m + j];
    HIPCHK(hipMemcpyAsync(w->cinv, inv_t.data(), m * m * sizeof(double), hipMemcpyHostToDevice, w->stream));
    HIPCHK(hipMemcpyAsync(w->ccell, cell.data(), m * sizeof(int32_t), hipMemcpyHostToDevice, w->stream));
    HIPCHK(hipStreamSynchronize(w->stream));
    w->cm = (int)m;
    w->cinv_valid = true;
    return RIBO_OK;
}

// One V-cycle (src/world/mod.rs:1555-1575).  Levels [0, block) run as one
// fused pre-smooth + restrict launch and one prolong + post-smooth launch
// each; levels [block, L) run inside a single workgroup.
void vcycle(RiboWorld* w, const uint32_t* done) {
    const size_t L = w->mg.size();
    hipStream_t s = w->stream;
    const size_t b = w->mg_block_start;
    if (b < L) {
        for (size_t l = 0; l < b; ++l)
            run(w, "mg_down", [&] { launch_mg_down<float>(s, w->mg[l], w->mg[l + 1], false, done, w->mg_tile[l]); });
        run(w, "mg_block", [&] { launch_mg_block<float>(s, w->mg.data() + b, (int)(L - b), false, nullptr, nullptr, 0, done); });
        for (size_t l = b; l-- > 0;)
            run(w, "mg_up", [&] { launch_mg_up<float>(s, w->mg[l], w->mg[l + 1], false, done, w->mg_tile[l]); });
        return;
    }
    // coarsest level too large for one workgroup: per-operation dispatches
    for (size_t l = 0; l + 1 < L; ++l)
        run(w, "mg_down", [&] { launch_mg_down<float>(s, w->mg[l], w->mg[l + 1], false, done, w->mg_tile[l]); });
    for (uint32_t k = 0; k < MG_COARSE; ++k) {
        run(w, "mg_smooth", [&] { launch_mg_smooth(s, w->mg[L - 1], 0, done); });
        run(w, "mg_smooth", [&] { launch_mg_smooth(s, w->mg[L - 1], 1, done); });
    }
    for (size_t l = L - 1; l-- > 0;)
        run(w, "mg_up", [&] { launch_mg_up<float>(s, w->mg[l], w->mg[l + 1], false, done, w->mg_tile[l]); });
}

// Tiles [t0, t1) of the pending scatter that V-cycle j of n carries: the
// tiles spread evenly over as few slices as fit the CUs beside k_mg_block
// (FuseCfg: tiles a workgroup, workgroups a CU; every workgroup of a launch
// runs at once, and the CU of k_mg_block's holds no scatter workgroup beside
// it).  With fewer V-cycles than that, each
// takes a full slice and the tiles from fused_rest(n) on run on their own
// after the solve.
uint32_t fused_per(const RiboWorld* w, uint32_t n, uint32_t* slices) {
    const uint32_t tiles = scatter_tiles(w->G);
    const bool half = w->st.precision == RIBO_PRECISION_MIXED;
    const uint32_t per_cu = half ? FuseCfg<bf16>::per_cu : FuseCfg<float>::per_cu;
    const uint32_t cap = (per_cu * w->cu_count - per_cu) * (half ? FuseCfg<bf16>::tiles : FuseCfg<float>::tiles);
    const uint32_t need = (tiles + cap - 1u) / cap;
    *slices = std::max(std::min(need, n), 1u);
    return *slices == need ? (tiles + need - 1u) / need : cap;
}
void fused_slice(const RiboWorld* w, uint32_t j, uint32_t n, uint32_t* t0, uint32_t* t1) {
    uint32_t slices = 0;
    const uint32_t per = fused_per(w, n, &slices), tiles = scatter_tiles(w->G);
    *t0 = j < slices ? std::min(j * per, tiles) : tiles;
    *t1 = j < slices ? std::min(*t0 + per, tiles) : tiles;
}
uint32_t fused_rest(const RiboWorld* w, uint32_t n) {
    uint32_t slices = 0;
    const uint32_t per = fused_per(w, n, &slices);
    return std::min(slices * per, scatter_tiles(w->G));
}

// Levels 1 .. sym_coarse (the coarsest, solved exactly with its dense
// inverse) of the symmetric V-cycle; level 0 is fused into the
// conjugate-gradient kernels.  (Stopping at 16 x 16 with a 256-unknown
// inverse was measured: the iterations stayed the same at 1024^2 and the
// one-workgroup inverse product was slower than the levels it replaced.)
template <class S>
void vcycle_sym_inner(RiboWorld* w, const std::vector<MgLevelT<S>>& lv, const uint32_t* done, uint32_t j = 0, uint32_t n = 0) {
    const size_t L = lv.size(), lc = w->sym_coarse;
    hipStream_t s = w->stream;
    if (w->mg_block_start < L) {
        const size_t b = std::min(std::max<size_t>(w->mg_block_start, 1), lc);
        for (size_t l = 1; l < b; ++l)
            run(w, "mg_down", [&] { launch_mg_down<S>(s, lv[l], lv[l + 1], true, done, w->mg_tile[l]); });
        uint32_t t0 = 0, t1 = 0;
        if (w->pend.on) fused_slice(w, j, n, &t0, &t1);
        run(w, "mg_block", [&] {
            const int* cmd = w->cmoving ? w->cm_dev : nullptr;
            if (t1 > t0) {
                const RiboWorld::PendingScatter& q = w->pend;
                const ScatterArgs sa{w->geo(), q.P, q.ep, q.chem, q.next, w->vel_s, q.temp, w->agg_e};
                launch_mg_block_scatter<S>(s, lv.data() + b, (int)(lc - b + 1), w->cinv, w->ccell, w->cm, done, cmd, sa, t0, t1,
                                           w->cu_count);
            } else {
                launch_mg_block<S>(s, lv.data() + b, (int)(lc - b + 1), true, w->cinv, w->ccell, w->cm, done, cmd);
            }
        });
        for (size_t l = b; l-- > 1;) run(w, "mg_up", [&] { launch_mg_up<S>(s, lv[l], lv[l + 1], true, done, w->mg_tile[l]); });
        return;
    }
    // levels too large for one workgroup: sweeps only (the coarsest level
    // keeps the zero correction its restriction implies)
    for (size_t l = 1; l + 1 < L; ++l)
        run(w, "mg_down", [&] { launch_mg_down<S>(s, lv[l], lv[l + 1], true, done, w->mg_tile[l]); });
    run(w, "mg_zero", [&] { launch_zero_level<S>(s, lv[L - 1]); });
    for (size_t l = L - 1; l-- > 1;) run(w, "mg_up", [&] { launch_mg_up<S>(s, lv[l], lv[l + 1], true, done, w->mg_tile[l]); });
}

// The end of a deferred scatter: its tiles below `rest` ran in the
// V-cycles, the others run now on their own (all of them when rest is 0);
// then the ledger of its output.
void finish_scatter(RiboWorld* w, uint32_t rest) {
    RiboWorld::PendingScatter& q = w->pend;
    if (!q.on) return;
    hipStream_t s = w->stream;
    const Geo g = w->geo();
    const uint32_t tiles = scatter_tiles(w->G);
    if (rest == 0u) run(w, "transport_scatter", [&] { launch_scatter(s, g, q.P, q.ep, q.chem, q.next, w->vel_s, q.temp, w->agg_e); });
    else if (rest < tiles)
        run(w, "transport_scatter", [&] {
            launch_scatter_range(s, ScatterArgs{g, q.P, q.ep, q.chem, q.next, w->vel_s, q.temp, w->agg_e}, rest, tiles);
        });
    q.on = false;
    if (w->st.ledger_check) run(w, "ledger", [&] { launch_ledger(s, g, q.next, w->ledger, q.P.epoch, q.ep, 1); });
}

// Tolerance mode: preconditioned conjugate gradients (Chronopoulos-Gear
// recurrences) on A p = div over the fluid cells of level 0, warm-started
// from the pressure, preconditioned by the symmetric V-cycle stored in S
// (binary32, or bfloat16 for the mixed-precision form); stops on
// ||div - A p|| / ||div|| <= mg_tol or after mg_max_cycles iterations.  The
// first V-cycle forms u = M r for the initial residual; each of the `slots`
// iterations is one more V-cycle whose first and last kernels carry the
// vector updates and inner products.  Every kernel returns at once after
// convergence, so the recorded work does not depend on host round trips.
template <class S> void pcg_solve(RiboWorld* w, const std::vector<MgLevelT<S>>& lv, S* rs, S* z, uint32_t slots) {
    hipStream_t s = w->stream;
    CgVec<S> V;
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
    V.u = z;
    V.zpre = lv[0].pp;
    V.rs = rs;
    const float tol = w->st.mg_tol;
    const uint32_t maxit = std::max(w->st.mg_max_cycles, 1u);
    const uint32_t* done = &w->pcg->done;
    run(w, "pcg_init", [&] { launch_cg_init<S>(s, V, w->pcg, tol, maxit); });
    const uint32_t n_it = std::max(std::min(slots, maxit), 1u);
    // V-cycle j applies the inner products of V-cycle j - 1 to the scalars
    // kept in slot rd (the initial state is in slot 1) and writes slot 1 - rd
    for (uint32_t j = 0; j <= n_it; ++j) {
        const int stage = j == 0 ? 0 : (j == 1 ? 1 : 2);
        const int rd = j == 0 ? 1 : (int)(j & 1u);
        run(w, "cg_down0", [&] { launch_cg_down0<S>(s, V, lv[1], w->pcg, (int)(j & 1u), stage, rd, tol, maxit); });
        vcycle_sym_inner<S>(w, lv, done, j, n_it + 1u);
        run(w, "cg_up0", [&] { launch_cg_up0<S>(s, V, lv[1], w->pcg, (int)(j & 1u)); });
    }
    run(w, "pcg_final", [&] { launch_cg_final<S>(s, V, w->pcg, (int)((n_it - 1u) & 1u), tol, maxit); });
    finish_scatter(w, w->mg_block_start < lv.size() ? fused_rest(w, n_it + 1u) : 0u);
}

// Stationary tolerance mode: the reference V-cycle (as vcycle() records it)
// repeated until ||div - A p|| / ||div|| over level-0 fluid cells reaches
// mg_tol or mg_max_cycles V-cycles ran; the residual is checked after every
// V-cycle and the kernels of the remaining slots return at once.
void vcycle_solve(RiboWorld* w, uint32_t slots) {
    hipStream_t s = w->stream;
    const float tol = w->st.mg_tol;
    const uint32_t maxit = std::max(w->st.mg_max_cycles, 1u);
    const uint32_t* done = &w->pcg->done;
    auto check = [&](int phase) {
        run(w, "vc_check",
            [&] { launch_vc_step(s, phase, w->F, w->levels[0].eps, w->mg_fluid, w->p_a, w->div, w->pcg, tol, maxit); });
    };
    check(0);
    const uint32_t n_it = std::max(std::min(slots, maxit), 1u);
    for (uint32_t j = 0; j < n_it; ++j) {
        vcycle(w, done);
        check(1);
    }
    check(2);
}

// The terrain stage of a step (src/world/mod.rs:1495-1503): two relax
// passes, grain_fall when grains settle, then the slope.
void terrain_step(RiboWorld* w, const RiboSimParams& P, const uint32_t* ep) {
    hipStream_t s = w->stream;
    const Geo g = w->geo();
    for (uint32_t phase = 0; phase < 2u; ++phase)
        run(w, phase ? "relax_gamma_b" : "relax_gamma_a",
            [&] { launch_relax_pass(s, g, P, ep, phase, w->vel_a, w->gamma, w->chem[w->cur], w->terrain_code, w->terrain_inc); });
    if (P.sediment_settle > 0.0f) run(w, "grain_fall", [&] { launch_grain_fall(s, g, P, ep, w->gamma, w->chem[w->cur]); });
    run(w, "compute_gamma_slope", [&] { launch_slope(s, g, w->slope); });
}

// The fluid's terrain-derived data (wall flags, fill, drag, slope at the
// fluid cells, wall normals, multigrid masks), rebuilt at each fluid step
// while the terrain moves, as build_solid_mask is (fluid.wgsl:185-189).
void fluid_terrain(RiboWorld* w) {
    hipStream_t s = w->stream;
    const Geo g = w->geo();
    run(w, "build_solid_mask", [&] {
        launch_fluid_static(s, g, w->cflags, w->fill, w->dragg, w->slope_f);
        launch_fluid_normals(s, g, w->cflags, w->normal);
        launch_mg_mask0(s, g, w->mg_fluid);
        for (size_t l = 1; l < w->levels.size(); ++l)
            launch_mg_mask_coarsen(s, w->levels[l].n, w->mg_fluid + w->levels[l - 1].off, w->mg_fluid + w->levels[l].off);
    });
}

void fluid_step(RiboWorld* w, const RiboSimParams& P, const uint32_t* ep) {  // src/world/mod.rs:1527-1588
    hipStream_t s = w->stream;
    if (w->st.terrain_enabled) {
        fluid_terrain(w);
        if (w->cmoving) {  // the coarsest inverse follows the coarsest fluid mask
            const RiboMgLevel& C = w->levels[w->sym_coarse];
            run(w, "coarse_refresh", [&] {
                launch_coarse_refresh(s, C.n, C.eps, w->mg_fluid + C.off, w->csnap, w->ccap, w->cwork, w->cinv, w->ccell,
                                      w->cm_dev);
            });
        }
    }
    Geo g = w->geo();
    const int tc = w->tcur;
    // pipelined: update_temperature rewrites temp[1 - tc], which the
    // transport before last read
    if (w->pipe && w->pipe_temp_read[1 - tc]) (void)hipStreamWaitEvent(s, w->pe_temp[1 - tc], 0);
    run(w, "update_temperature", [&] {
        launch_temperature(s, g, P, ep, w->vel_a, w->temp[tc], w->redox[tc], w->temp[1 - tc], w->redox[1 - tc], w->eaten,
                           w->eaten_active ? 1 : 0, w->chem[w->cur]);
    });
    w->eaten_active = false;
    w->tcur = 1 - tc;  // copy_temperature
    const float* T = w->temp[w->tcur];
    run(w, "buoyancy_gather", [&] { launch_forces(s, g, P, ep, T, w->forces, w->agent_force_active ? w->agent_force : nullptr); });
    w->agent_force_active = false;
    run(w, "add_forces", [&] { launch_add_forces(s, g, P, w->forces, w->vel_a, w->vel_b); });
    run(w, "diffuse_velocity", [&] { launch_diffuse(s, g, P, w->vel_b, w->vel_a); });
    const float decay = decay_factor(P.fluid_decay, P.fluid_dt);
    run(w, "advect_velocity", [&] { launch_advect(s, g, P, decay, w->vel_a, w->vel_b); });
    run(w, "vorticity", [&] { launch_vorticity(s, g, P, w->vel_b, w->vel_a); });
    run(w, "divergence", [&] { launch_divergence(s, g, w->vel_a, w->div, w->p_a); });
    if (w->st.multigrid) {
        if (w->st.mg_tol > 0.0f && w->st.mg_stationary) {
            vcycle_solve(w, w->pcg_slots);
        } else if (w->st.mg_tol > 0.0f) {
            if (w->st.precision == RIBO_PRECISION_MIXED) pcg_solve<bf16>(w, w->mgh, w->pcg_r16, w->pcg_z16, w->pcg_slots);
            else pcg_solve<float>(w, w->mgz, nullptr, w->pcg_z, w->pcg_slots);
        } else {
            const uint32_t cycles = std::max(w->st.mg_cycles, 1u);
            for (uint32_t c = 0; c < cycles; ++c) vcycle(w, nullptr);
        }
    } else {
        const uint32_t iters = (std::min(std::max(w->st.jacobi_iters, 2u), 128u) + 1u) & ~1u;  // :1442
        for (uint32_t k = 0; k < iters; ++k) {
            if (k % 2 == 0) run(w, "jacobi", [&] { launch_jacobi(s, g, w->p_a, w->p_b, w->div); });
            else run(w, "jacobi", [&] { launch_jacobi(s, g, w->p_b, w->p_a, w->div); });
        }
    }
    run(w, "subtract_gradient", [&] { launch_subgrad(s, g, w->vel_a, w->vel_b, w->p_a); });
    run(w, "net_flux", [&] { launch_net_flux(s, g, w->vel_b, w->net_flux); });
    if (!w->pipe) {
        run(w, "enforce_boundaries", [&] { launch_boundaries(s, g, w->vel_b, w->vel_a, w->net_flux, w->fuse ? w->vel_s : nullptr); });
        run(w, "thermal_activation", [&] { launch_thermal(s, g, P, ep, T, w->chem[w->cur]); });
        return;
    }
    // pipelined: vel_s waits for the last transport's reads; the thermal
    // activation follows on the transport stream
    if (w->pipe_vel_read) (void)hipStreamWaitEvent(s, w->pe_vel, 0);
    run(w, "enforce_boundaries", [&] { launch_boundaries(s, g, w->vel_b, w->vel_a, w->net_flux, w->vel_s); });
    (void)hipEventRecord(w->pe_fluid, s);
    hipStream_t ts = w->tstream;
    (void)hipStreamWaitEvent(ts, w->pe_fluid, 0);
    run(w, "thermal_activation", [&] { launch_thermal(ts, g, P, ep, T, w->chem[w->cur]); });
    (void)hipEventRecord(w->pe_temp[w->tcur], ts);
    w->pipe_temp_read[w->tcur] = true;
}

void transport_step(RiboWorld* w, const RiboSimParams& P, const uint32_t* ep, bool defer) {  // src/world/mod.rs:1594-1601
    hipStream_t s = w->pipe ? w->tstream : w->stream;
    Geo g = w->geo();
    uint32_t* cur = w->chem[w->cur];
    uint32_t* nxt = w->chem[1 - w->cur];
    if (P.aggregation > 0.0f) run(w, "agg", [&] { launch_agg(s, g, cur, w->agg_e); });
    run(w, "clear_next", [&] { launch_fill_u32(s, nxt, (size_t)4 * w->G * w->G, 0u); });
    if (defer) {  // the scatter rides on the next solve (finish_scatter adds the ledger)
        RiboWorld::PendingScatter& q = w->pend;
        q.on = true;
        q.P = P;
        q.ep = ep;
        q.chem = cur;
        q.next = nxt;
        q.temp = w->temp[w->tcur];
        w->cur = 1 - w->cur;  // transport_commit
        return;
    }
    const float2* vel = w->pipe ? w->vel_s : w->vel_a;
    run(w, "transport_scatter", [&] { launch_scatter(s, g, P, ep, cur, nxt, vel, w->temp[w->tcur], w->agg_e); });
    if (w->pipe) {
        (void)hipEventRecord(w->pe_temp[w->tcur], s);
        (void)hipEventRecord(w->pe_vel, s);
        w->pipe_temp_read[w->tcur] = true;
        w->pipe_vel_read = true;
    }
    w->cur = 1 - w->cur;  // transport_commit
    if (w->st.ledger_check) run(w, "ledger", [&] { launch_ledger(s, g, w->chem[w->cur], w->ledger, P.epoch, ep, 1); });
}

// Light (src/world/mod.rs:1506-1525) with light_mode 1, before the fluid:
// the full sweep (light_transmit_pass, then compute_uv_light) when the light
// is stale or, with light_rows_per_step = 0, every light_interval steps;
// otherwise light_rows_per_step (at most 16) propagation passes, each a
// ping-pong of the two fields instead of light_commit's copy.  Agent shade
// (clear_shade, agents_shade) is zero: there are no agents.
void light_step(RiboWorld* w, const RiboSimParams& P, uint32_t epoch, const uint32_t* ep) {
    if (!w->st.light_mode) return;
    hipStream_t s = w->stream;
    const bool sweep = w->light_dirty || (w->st.light_rows_per_step == 0 && epoch % std::max(w->st.light_interval, 1u) == 0);
    if (sweep) {
        w->light_dirty = false;
        Geo g = w->geo();
        run(w, "light_transmit", [&] { launch_light_transmit(s, g, P, w->chem[w->cur], w->light[w->lcur]); });
        run(w, "light_sweep", [&] { launch_light_sweep(s, g, P, ep, w->light[w->lcur]); });
        return;
    }
    for (uint32_t k = 0; k < std::min(w->st.light_rows_per_step, 16u); ++k) {
        Geo g = w->geo();
        run(w, "light_propagate",
            [&] { launch_light_propagate(s, g, P, ep, w->chem[w->cur], w->light[w->lcur], w->light[1 - w->lcur]); });
        w->lcur = 1 - w->lcur;
    }
}

// Steps [first, first + count) of one encode_steps call: epochs
// params.epoch + first + i, passed as is (eager) or as i on top of the
// device epoch base `ep` (graph capture).
void record_steps(RiboWorld* w, uint32_t first, uint32_t count, uint32_t substep, const uint32_t* ep) {
    w->pend.on = false;  // a recording cut short by an error leaves nothing pending for the next
    if (w->pipe) {  // the transport stream starts after everything before
        (void)hipEventRecord(w->pe_fork, w->stream);
        (void)hipStreamWaitEvent(w->tstream, w->pe_fork, 0);
        w->pipe_temp_read[0] = w->pipe_temp_read[1] = w->pipe_vel_read = false;
    }
    for (uint32_t i = 0; i < count; ++i) {
        const uint32_t epoch = w->params.epoch + first + i;
        RiboSimParams P = w->params;
        P.epoch = ep ? i : epoch;
        P.sun_now = w->params.uv_strength > 0.0f ? day_light(w->params, epoch) : 0.0f;
        P.sun_slope = sun_slope_at(w->params, epoch);
        if (w->st.terrain_enabled) terrain_step(w, P, ep);
        light_step(w, P, epoch, ep);
        if (w->st.fluid_enabled && epoch % substep == 0) fluid_step(w, P, ep);
        if (epoch % std::max(P.transport_every, 1u) == 0) transport_step(w, P, ep, w->fuse && i + 1 < count);
    }
    finish_scatter(w, 0u);
    if (w->pipe) {  // ... and everything after waits for it
        (void)hipEventRecord(w->pe_join, w->tstream);
        (void)hipStreamWaitEvent(w->stream, w->pe_join, 0);
    }
}

// Steps pipeline when nothing of a fluid step reads what the transport
// writes: the light frozen (the light step reads the monomers), no solar
// heating (update_temperature's absorbed light reads them), the terrain
// still (grain moves rewrite monomers and walls); and not while kernels are
// timed one by one.  The tolerance modes do not pipeline: their solver is a
// chain of short, latency-bound launches, and the scatter beside it takes
// every CU's LDS and waves and delays the chain by more than it hides (a
// binary32 step measured 2.72 ms pipelined against 2.45 ms in one stream;
// the reference schedule 0.68 against 0.75 ms).
bool pipeline_ok(const RiboWorld* w) {
    return w->pipe_allowed && !w->timing && !w->st.light_mode && !w->st.terrain_enabled && !(w->params.sun_heat > 0.0f) &&
           !(w->st.multigrid && w->st.mg_tol > 0.0f);
}

// The scatter rides on the next step's solve (RiboWorld::fuse) under the
// conditions of pipeline_ok, in the conjugate-gradient tolerance mode with a
// one-workgroup V-cycle kernel whose LDS fits the scatter tiles', when every
// step has one fluid step and one transport.
bool fuse_ok(const RiboWorld* w, uint32_t substep) {
    if (!w->pipe_allowed || w->timing || w->st.light_mode || w->st.terrain_enabled || w->params.sun_heat > 0.0f) return false;
    if (!w->st.fluid_enabled || !w->st.multigrid || !(w->st.mg_tol > 0.0f) || w->st.mg_stationary) return false;
    if (substep != 1u || std::max(w->params.transport_every, 1u) != 1u) return false;
    const size_t L = w->mgz.size();
    if (w->mg_block_start >= L || L < 2) return false;
    const size_t b = std::min(std::max<size_t>(w->mg_block_start, 1), w->sym_coarse);
    const int count = (int)(w->sym_coarse - b + 1);
    return w->st.precision == RIBO_PRECISION_MIXED ? mg_block_scatter_fits<bf16>(w->mgh.data() + b, count)
                                                   : mg_block_scatter_fits<float>(w->mgz.data() + b, count);
}

// The transport stream and the events, made when a world first pipelines
// (worlds that never do, such as a server's, keep one stream).
int ensure_pipeline(RiboWorld* w) {
    if (w->tstream) return RIBO_OK;
    hipError_t he = hipStreamCreateWithFlags(&w->tstream, hipStreamNonBlocking);
    for (hipEvent_t* e : {&w->pe_fork, &w->pe_join, &w->pe_fluid, &w->pe_temp[0], &w->pe_temp[1], &w->pe_vel})
        if (he == hipSuccess) he = hipEventCreateWithFlags(e, hipEventDisableTiming);
    if (he != hipSuccess) return fail(RIBO_ERR_HIP, std::string("pipeline stream: ") + hipGetErrorString(he));
    return RIBO_OK;
}

// Graph replay is used when nothing in the recorded work depends on host
// values that change from step to step: the sun of a day cycle comes from
// the per-epoch table on the device, but a stale light's sweep is a
// one-off, and a pending redox_eaten or force_vectors upload is consumed
// once.
bool graph_eligible(const RiboWorld* w) {
    return w->st.no_graphs == 0 && !w->timing && !w->eaten_active && !w->agent_force_active &&
           !(w->st.light_mode && w->light_dirty);
}

// The sun table (SunTab) holds record_steps' (sun_now, sun_slope) of every
// epoch of a window; it is active while a day cycle moves the sun
// (day_period >= 1), and rewritten (after draining the stream) when the
// steps about to be recorded leave the window or the sun's parameters change.
int ensure_sun(RiboWorld* w, uint32_t steps) {
    const RiboSimParams& p = w->params;
    w->sun_active = p.day_period >= 1.0f;
    if (!w->sun_active) return RIBO_OK;
    const float key[4] = {p.day_period, p.day_fraction, p.sun_angle, p.uv_strength > 0.0f ? 1.0f : 0.0f};
    const uint32_t base = w->sun_valid ? w->sun_host[0].base : 0u;
    const bool inside = w->sun_valid && p.epoch >= base && (uint64_t)p.epoch + steps <= (uint64_t)base + SUN_TAB_LEN;
    if (inside && memcmp(key, w->sun_key, sizeof key) == 0) return RIBO_OK;
    HIPCHK(hipStreamSynchronize(w->stream));  // no recorded step still reads the old window
    SunTab& t = w->sun_host[0];
    t.base = p.epoch;
    t.pad = 0u;
    for (uint32_t i = 0; i < SUN_TAB_LEN; ++i) {
        const uint32_t e = p.epoch + i;
        t.v[i] = make_float2(p.uv_strength > 0.0f ? day_light(p, e) : 0.0f, sun_slope_at(p, e));
    }
    HIPCHK(hipMemcpyAsync(w->sun_tab, &t, sizeof(SunTab), hipMemcpyHostToDevice, w->stream));
    HIPCHK(hipStreamSynchronize(w->stream));
    memcpy(w->sun_key, key, sizeof key);
    w->sun_valid = true;
    return RIBO_OK;
}

// Arms the per-step ledger check with the totals of the current grid (when
// checking is switched on, and again after every upload of the grid: a
// host-side change of matter is not a transport error).
int arm_ledger(RiboWorld* w) {
    LedgerDev init{};
    init.first_bad = 0xFFFFFFFFu;
    HIPCHK(hipMemcpyAsync(w->ledger, &init, sizeof(init), hipMemcpyHostToDevice, w->stream));
    launch_ledger(w->stream, w->geo(), w->chem[w->cur], w->ledger, w->params.epoch, nullptr, 1);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(w->stream));
    return RIBO_OK;
}

// The tolerance mode starts each solve from the extrapolation of the two
// previous solves' pressures; a new pressure, terrain, velocity or settings
// starts the history again (the next solve starts from the pressure itself).
int reset_pressure_history(RiboWorld* w) {
    HIPCHK(hipMemsetAsync((char*)w->pcg + offsetof(PcgState, hist), 0, sizeof(uint32_t), w->stream));
    return RIBO_OK;
}

int validate_settings(const RiboWorldSettings* s) {
    if (s->terrain_enabled > 1u) return fail(RIBO_ERR_ARG, "terrain_enabled must be 0 or 1");
    if (s->light_mode > 1u) return fail(RIBO_ERR_ARG, "light_mode must be 0 (frozen) or 1 (the reference's light schedule)");
    if (s->precision != RIBO_PRECISION_FP32 && s->precision != RIBO_PRECISION_MIXED)
        return fail(RIBO_ERR_ARG, "unknown precision");
    if (s->precision == RIBO_PRECISION_MIXED && !(s->mg_tol > 0.0f))
        return fail(RIBO_ERR_ARG, "the mixed-precision form is a tolerance-mode solver: set mg_tol > 0");
    if (!(s->mg_tol >= 0.0f)) return fail(RIBO_ERR_ARG, "mg_tol must be >= 0");
    if (s->mg_stationary > 1u) return fail(RIBO_ERR_ARG, "mg_stationary must be 0 or 1");
    if (s->mg_stationary && s->precision != RIBO_PRECISION_FP32)
        return fail(RIBO_ERR_UNSUPPORTED, "the stationary tolerance mode is binary32 (precision 0)");
    return RIBO_OK;
}

bool buffer_size_ok(RiboWorld* w, RiboBuffer which, size_t bytes) { return bytes == ribo_buffer_bytes(&w->cfg, which) && bytes > 0; }

}  // namespace

namespace ribo {
int set_error(int code, const std::string& msg) { return fail(code, msg); }
float step_decay_factor(float decay, float fluid_dt) { return decay_factor(decay, fluid_dt); }
std::recursive_mutex& alloc_mutex() { return device_mutex(); }
}  // namespace ribo

// ============================================================================
extern "C" {

const char* ribo_last_error(void) { return g_err.c_str(); }
uint32_t ribo_abi_version(void) { return RIBO_ABI_VERSION; }

void ribo_sim_params_default(RiboSimParams* out) {
    if (out) ribo_sim_params_fill_default(out);
}

void ribo_world_settings_default(RiboWorldSettings* s) {  // src/world/mod.rs:121-134
    if (!s) return;
    memset(s, 0, sizeof(*s));
    s->fluid_enabled = 1;
    s->fluid_substep = 2;
    s->multigrid = 1;
    s->mg_cycles = 1;
    s->jacobi_iters = 128;
    s->light_interval = 10;
    s->light_rows_per_step = 1;
    s->terrain_enabled = 1;
    s->contact_enabled = 1;
    s->light_mode = 1;  // the reference runs its light step every step
    s->mg_min_size = 8;
    s->mg_tol = 0.0f;
    s->mg_max_cycles = 1;
    s->precision = RIBO_PRECISION_FP32;
    s->ledger_check = 0;
}

size_t ribo_buffer_bytes(const RiboWorldConfig* cfg, RiboBuffer which) {
    if (!cfg) return 0;
    size_t G = (size_t)cfg->grid_size * cfg->grid_size, F = (size_t)cfg->fluid_size * cfg->fluid_size;
    size_t L = (size_t)(cfg->grid_size / LIGHT_DIV) * (cfg->grid_size / LIGHT_DIV);
    switch (which) {
        case RIBO_BUF_CHEM: return 16 * G;
        case RIBO_BUF_GAMMA: return 4 * G;
        case RIBO_BUF_LIGHT: return 8 * L;
        case RIBO_BUF_SLOPE: return 8 * G;
        case RIBO_BUF_VELOCITY: return 8 * F;
        case RIBO_BUF_PRESSURE: return 4 * F;
        case RIBO_BUF_TEMP: return 4 * F;
        case RIBO_BUF_REDOX: return 4 * F;
        case RIBO_BUF_HEAT_SRC: return 8 * F;
        case RIBO_BUF_DIVERGENCE: return 4 * F;
        case RIBO_BUF_SOLID_MASK: return 4 * F;
        case RIBO_BUF_NET_FLUX: return 8 * (size_t)cfg->fluid_size;
        case RIBO_BUF_REDOX_EATEN: return 4 * F;
        case RIBO_BUF_AGG: return 4 * G;
        case RIBO_BUF_FORCE_VECTORS: return 8 * F;
        case RIBO_BUF_VELOCITY_SMOOTH: return 8 * F;
        default: return 0;
    }
}

size_t ribo_world_footprint(const RiboWorldConfig* cfg, const RiboWorldSettings* st) {
    if (!cfg || cfg->grid_size == 0 || cfg->fluid_size == 0) return 0;
    auto a = [](size_t count, size_t elem) { return std::max<size_t>(count, 1) * elem; };  // dalloc's sizes
    const size_t G = (size_t)cfg->grid_size * cfg->grid_size, F = (size_t)cfg->fluid_size * cfg->fluid_size;
    const size_t Ls = cfg->grid_size / LIGHT_DIV, nb = (cfg->fluid_size + 15) / 16;
    size_t b = 2 * a(4 * G, 4) + a(G, 4) + 2 * a(2 * Ls * Ls, 4) + a(G, 8) + a(G, 4);  // chem x2, gamma, light x2, slope, agg
    b += 3 * a(F, 8) + 4 * a(F, 4) + 4 * a(F, 4) + a(F, 4) + 2 * a(F, 8);  // velocity x3, p x3, div, temp/redox x2, eaten, forces, heat
    b += a(2 * F, 4) + 2 * a(G, 1);  // the agents' force words, the grain move codes and masks
    b += a(2 * (size_t)cfg->fluid_size, 4) + a(F, 1) + a(F, 8) + 2 * a(F, 4) + a(F, 8);  // net flux, terrain-derived
    b += a(1, sizeof(PcgState)) + a(5 * nb * nb, 8) + 8 * a(F, 4) + 2 * a(F, 2);  // tolerance-mode vectors
    b += a(1, sizeof(LedgerDev)) + a(F, 4) + a(1, 4) + a(1, sizeof(SunTab));  // ledger, scratch, device epoch, sun table
    const uint32_t min_size = st ? std::max(st->mg_min_size, 1u) : 8u;
    const std::vector<RiboMgLevel> lv = make_levels(cfg->fluid_size, min_size);
    const size_t total = lv.back().off_c, coarse = total - F;
    b += 2 * a(coarse, 4) + a(total, 4) + a(total, 1) + 2 * a(coarse, 2) + a(total, 2);
    if (st && st->multigrid && st->mg_tol > 0.0f) {  // the exact coarse solve: at most every cell of the coarsest level
        const size_t m = (size_t)lv.back().n * lv.back().n;
        if (m <= 1024) b += a(m * m, 8) + a(m, 4);
        // under a moving terrain: the mask snapshot, the elimination work and the device count
        if (st->terrain_enabled && !st->mg_stationary && m <= MOVING_COARSE_MAX) b += a(m, 1) + a(2 * m * m, 8) + a(1, 4);
    }
    return b;
}

int ribo_world_device_bytes(const RiboWorld* w, size_t* out) {
    if (!w || !out) return fail(RIBO_ERR_ARG, "null argument");
    *out = w->base_bytes + w->mg_bytes + w->cinv_bytes;
    return RIBO_OK;
}

int ribo_world_create(const RiboWorldConfig* cfg, uint32_t seed, RiboWorld** out) {
    if (!cfg || !out) return fail(RIBO_ERR_ARG, "null argument");
    std::lock_guard<std::recursive_mutex> lk(device_mutex());
    *out = nullptr;
    if (cfg->grid_size == 0 || cfg->fluid_size == 0 || cfg->grid_size % cfg->fluid_size != 0 || cfg->grid_size < 2 ||
        cfg->world_units_per_cell == 0 || cfg->grid_size > 32768)
        return fail(RIBO_ERR_ARG, "grid_size must be a non-zero multiple of fluid_size (<= 32768) and world_units_per_cell > 0");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(RIBO_ERR_NO_DEVICE, "no HIP device visible");
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, dev));
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(RIBO_ERR_NO_DEVICE, std::string("this library is built for gfx950; device is ") + prop.gcnArchName);

    RiboWorld* w = new RiboWorld();
    w->sun_host.resize(1);
    w->cfg = *cfg;
    w->G = cfg->grid_size;
    w->F = cfg->fluid_size;
    w->Lsz = cfg->grid_size / LIGHT_DIV;
    w->sim_size = (float)(cfg->grid_size * cfg->world_units_per_cell);
    w->device = dev;
    // k_mg_block takes its LDS dynamically, up to the workgroup maximum
    // (160 KB on gfx950)
    w->lds_limit = (size_t)prop.sharedMemPerBlock;
    w->cu_count = (uint32_t)std::max(prop.multiProcessorCount, 2);
    ribo_sim_params_fill_default(&w->params);
    w->params.seed = seed;
    ribo_world_settings_default(&w->st);
    size_t G = (size_t)w->G * w->G, F = (size_t)w->F * w->F;
    int rc;
    hipError_t he = hipStreamCreateWithFlags(&w->stream, hipStreamNonBlocking);
    if (he != hipSuccess) {
        delete w;
        return fail(RIBO_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(he));
    }
    AllocCount count(&w->base_bytes);
    if ((rc = dalloc(w->stream, &w->chem[0], 4 * G)) || (rc = dalloc(w->stream, &w->chem[1], 4 * G)) || (rc = dalloc(w->stream, &w->gamma, G)) ||
        (rc = dalloc(w->stream, &w->light[0], 2 * (size_t)w->Lsz * w->Lsz)) ||
        (rc = dalloc(w->stream, &w->light[1], 2 * (size_t)w->Lsz * w->Lsz)) || (rc = dalloc(w->stream, &w->slope, G)) || (rc = dalloc(w->stream, &w->agg_e, G)) ||
        (rc = dalloc(w->stream, &w->vel_a, F)) || (rc = dalloc(w->stream, &w->vel_b, F)) || (rc = dalloc(w->stream, &w->vel_s, F)) ||
        (rc = dalloc(w->stream, &w->p_a, F)) || (rc = dalloc(w->stream, &w->p_b, F)) ||
        (rc = dalloc(w->stream, &w->p_old, F)) ||
        (rc = dalloc(w->stream, &w->div, F)) || (rc = dalloc(w->stream, &w->temp[0], F)) || (rc = dalloc(w->stream, &w->temp[1], F)) ||
        (rc = dalloc(w->stream, &w->redox[0], F)) || (rc = dalloc(w->stream, &w->redox[1], F)) || (rc = dalloc(w->stream, &w->eaten, F)) ||
        (rc = dalloc(w->stream, &w->forces, F)) || (rc = dalloc(w->stream, &w->agent_force, 2 * F)) || (rc = dalloc(w->stream, &w->terrain_code, G)) || (rc = dalloc(w->stream, &w->terrain_inc, G)) || (rc = dalloc(w->stream, &w->heat, F)) || (rc = dalloc(w->stream, &w->net_flux, 2 * (size_t)w->F)) ||
        (rc = dalloc(w->stream, &w->cflags, F)) || (rc = dalloc(w->stream, &w->normal, F)) || (rc = dalloc(w->stream, &w->fill, F)) ||
        (rc = dalloc(w->stream, &w->dragg, F)) || (rc = dalloc(w->stream, &w->slope_f, F)) || (rc = dalloc(w->stream, &w->pcg, 1)) ||
        (rc = dalloc(w->stream, &w->pcg_part, 5 * (size_t)((w->F + 15) / 16) * ((w->F + 15) / 16))) || (rc = dalloc(w->stream, &w->pcg_r, F)) ||
        (rc = dalloc(w->stream, &w->pcg_r1, F)) || (rc = dalloc(w->stream, &w->pcg_s0, F)) || (rc = dalloc(w->stream, &w->pcg_s1, F)) ||
        (rc = dalloc(w->stream, &w->pcg_pd, F)) || (rc = dalloc(w->stream, &w->pcg_w, F)) || (rc = dalloc(w->stream, &w->pcg_z, F)) ||
        (rc = dalloc(w->stream, &w->pcg_xlo, F)) ||
        (rc = dalloc(w->stream, &w->pcg_r16, F)) || (rc = dalloc(w->stream, &w->pcg_z16, F)) ||
        (rc = dalloc(w->stream, &w->ledger, 1)) || (rc = dalloc(w->stream, &w->scratch_u32, F)) || (rc = dalloc(w->stream, &w->epoch_dev, 1)) ||
        (rc = dalloc(w->stream, &w->sun_tab, 1)) ||
        (rc = build_mg(w))) {
        ribo_world_destroy(w);
        return rc;
    }
    LedgerDev init{};
    init.first_bad = 0xFFFFFFFFu;
    he = hipMemcpyAsync(w->ledger, &init, sizeof(init), hipMemcpyHostToDevice, w->stream);
    PcgState pst{};
    pst.part = w->pcg_part;
    const size_t host_words = PCG_HOST_RING + PCG_HOST_RING_LEN;
    if (he == hipSuccess && hipHostMalloc((void**)&w->pcg_host, host_words * sizeof(uint32_t), hipHostMallocMapped) == hipSuccess) {
        memset(w->pcg_host, 0, host_words * sizeof(uint32_t));
        if (hipHostGetDevicePointer((void**)&pst.host_out, w->pcg_host, 0) != hipSuccess) pst.host_out = nullptr;
    }
    if (he == hipSuccess) he = hipMemcpyAsync(w->pcg, &pst, sizeof(pst), hipMemcpyHostToDevice, w->stream);
    if (he != hipSuccess || (rc = rebuild_terrain(w)) != RIBO_OK || hipStreamSynchronize(w->stream) != hipSuccess) {
        ribo_world_destroy(w);
        return fail(RIBO_ERR_HIP, "world initialisation failed");
    }
    *out = w;
    return RIBO_OK;
}

int ribo_world_destroy(RiboWorld* w) {
    if (!w) return fail(RIBO_ERR_ARG, "null world");
    std::lock_guard<std::recursive_mutex> lk(device_mutex());
    if (w->stream) (void)hipStreamSynchronize(w->stream);
    if (w->tstream) (void)hipStreamSynchronize(w->tstream);
    resolve_timers(w);
    for (auto e : w->event_pool) (void)hipEventDestroy(e);
    for (hipEvent_t e : {w->pe_fork, w->pe_join, w->pe_fluid, w->pe_temp[0], w->pe_temp[1], w->pe_vel})
        if (e) (void)hipEventDestroy(e);
    dfree(w->chem[0]); dfree(w->chem[1]); dfree(w->gamma); dfree(w->light[0]); dfree(w->light[1]); dfree(w->slope); dfree(w->agg_e);
    dfree(w->vel_a); dfree(w->vel_b); dfree(w->vel_s); dfree(w->p_a); dfree(w->p_b); dfree(w->p_old); dfree(w->div); dfree(w->temp[0]); dfree(w->temp[1]);
    dfree(w->redox[0]); dfree(w->redox[1]); dfree(w->eaten); dfree(w->forces); dfree(w->agent_force); dfree(w->terrain_code); dfree(w->terrain_inc); dfree(w->heat); dfree(w->net_flux);
    dfree(w->cflags); dfree(w->normal); dfree(w->fill); dfree(w->dragg); dfree(w->slope_f);
    dfree(w->mg_p); dfree(w->mg_pp); dfree(w->mg_pp16); dfree(w->mg_rhs); dfree(w->mg_fluid); dfree(w->pcg); dfree(w->pcg_part); dfree(w->pcg_r);
    dfree(w->pcg_r1); dfree(w->pcg_s0); dfree(w->pcg_s1); dfree(w->pcg_pd); dfree(w->pcg_w); dfree(w->pcg_xlo);
    dfree(w->pcg_z); dfree(w->pcg_r16); dfree(w->pcg_z16); dfree(w->mg_p16);
    dfree(w->mg_rhs16); dfree(w->cinv); dfree(w->ccell); dfree(w->csnap); dfree(w->cwork); dfree(w->cm_dev); dfree(w->ledger); dfree(w->scratch_u32); dfree(w->epoch_dev);
    dfree(w->sun_tab);
    clear_graphs(w);
    if (w->pcg_host) (void)hipHostFree(w->pcg_host);
    if (w->stream) (void)hipStreamDestroy(w->stream);
    if (w->tstream) (void)hipStreamDestroy(w->tstream);
    delete w;
    return RIBO_OK;
}

int ribo_world_get_config(const RiboWorld* w, RiboWorldConfig* out) {
    if (!w || !out) return fail(RIBO_ERR_ARG, "null argument");
    *out = w->cfg;
    return RIBO_OK;
}

int ribo_world_set_params(RiboWorld* w, const RiboSimParams* p) {
    if (!w || !p) return fail(RIBO_ERR_ARG, "null argument");
    w->params = *p;
    return RIBO_OK;
}
int ribo_world_get_params(const RiboWorld* w, RiboSimParams* p) {
    if (!w || !p) return fail(RIBO_ERR_ARG, "null argument");
    *p = w->params;
    return RIBO_OK;
}
int ribo_world_set_settings(RiboWorld* w, const RiboWorldSettings* s) {
    if (!w || !s) return fail(RIBO_ERR_ARG, "null argument");
    int rc = validate_settings(s);
    if (rc) return rc;
    if (s->multigrid && s->mg_tol > 0.0f && make_levels(w->F, std::max(s->mg_min_size, 1u)).size() < 2)
        return fail(RIBO_ERR_UNSUPPORTED, "the tolerance mode needs at least two multigrid levels (an even fluid size above mg_min_size)");
    const bool relevel = std::max(s->mg_min_size, 1u) != std::max(w->st.mg_min_size, 1u);
    const bool arm = s->ledger_check && !w->st.ledger_check;
    w->st = *s;
    if ((rc = reset_pressure_history(w))) return rc;
    // the slot history starts again with the new settings (read after draining)
    HIPCHK(hipStreamSynchronize(w->stream));
    if (w->pcg_host) w->solves_at_settings = __atomic_load_n(&w->pcg_host[2], __ATOMIC_RELAXED);
    if (relevel) {
        HIPCHK(hipStreamSynchronize(w->stream));
        if ((rc = build_mg(w)) || (rc = rebuild_terrain(w))) return rc;
    }
    if (arm) {
        rc = arm_ledger(w);
        if (rc) return rc;
    }
    return RIBO_OK;
}
int ribo_world_get_settings(const RiboWorld* w, RiboWorldSettings* s) {
    if (!w || !s) return fail(RIBO_ERR_ARG, "null argument");
    *s = w->st;
    return RIBO_OK;
}
int ribo_world_mg_levels(const RiboWorld* w, RiboMgLevel* out, uint32_t cap, uint32_t* count) {
    if (!w || !count) return fail(RIBO_ERR_ARG, "null argument");
    *count = (uint32_t)w->levels.size();
    for (uint32_t i = 0; i < cap && i < *count; ++i) out[i] = w->levels[i];
    return RIBO_OK;
}

int ribo_world_upload(RiboWorld* w, RiboBuffer which, const void* src, size_t bytes) {
    if (!w || !src) return fail(RIBO_ERR_ARG, "null argument");
    if (!buffer_size_ok(w, which, bytes)) return fail(RIBO_ERR_ARG, "buffer size does not match the world configuration");
    void* dst = nullptr;
    switch (which) {
        case RIBO_BUF_CHEM: dst = w->chem[w->cur]; break;
        case RIBO_BUF_GAMMA: dst = w->gamma; break;
        case RIBO_BUF_LIGHT: dst = w->light[w->lcur]; w->light_dirty = false; break;
        case RIBO_BUF_VELOCITY: dst = w->vel_a; break;
        case RIBO_BUF_PRESSURE: dst = w->p_a; break;
        case RIBO_BUF_TEMP: dst = w->temp[w->tcur]; break;
        case RIBO_BUF_REDOX: dst = w->redox[w->tcur]; break;
        case RIBO_BUF_HEAT_SRC: dst = w->heat; break;
        case RIBO_BUF_REDOX_EATEN: dst = w->eaten; w->eaten_active = true; break;
        case RIBO_BUF_FORCE_VECTORS: dst = w->agent_force; w->agent_force_active = true; break;
        default: return fail(RIBO_ERR_ARG, "buffer is derived and cannot be uploaded");
    }
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, w->stream));
    // new terrain or matter leaves the light stale (src/world/mod.rs:986,1026)
    if (which == RIBO_BUF_GAMMA || which == RIBO_BUF_CHEM) w->light_dirty = true;
    if (which == RIBO_BUF_GAMMA || which == RIBO_BUF_VELOCITY || which == RIBO_BUF_PRESSURE) {
        int rc = reset_pressure_history(w);
        if (rc) return rc;
    }
    if (which == RIBO_BUF_GAMMA) {
        int rc = rebuild_terrain(w);
        if (rc) return rc;
    }
    if (which == RIBO_BUF_CHEM && w->st.ledger_check) {
        int rc = arm_ledger(w);
        if (rc) return rc;
    }
    HIPCHK(hipStreamSynchronize(w->stream));
    return RIBO_OK;
}

int ribo_world_download(RiboWorld* w, RiboBuffer which, void* dst, size_t bytes) {
    if (!w || !dst) return fail(RIBO_ERR_ARG, "null argument");
    if (!buffer_size_ok(w, which, bytes)) return fail(RIBO_ERR_ARG, "buffer size does not match the world configuration");
    const void* src = nullptr;
    switch (which) {
        case RIBO_BUF_CHEM: src = w->chem[w->cur]; break;
        case RIBO_BUF_GAMMA: src = w->gamma; break;
        case RIBO_BUF_LIGHT: src = w->light[w->lcur]; break;
        case RIBO_BUF_SLOPE: src = w->slope; break;
        case RIBO_BUF_VELOCITY: src = w->vel_a; break;
        case RIBO_BUF_PRESSURE: src = w->p_a; break;
        case RIBO_BUF_TEMP: src = w->temp[w->tcur]; break;
        case RIBO_BUF_REDOX: src = w->redox[w->tcur]; break;
        case RIBO_BUF_HEAT_SRC: src = w->heat; break;
        case RIBO_BUF_DIVERGENCE: src = w->div; break;
        case RIBO_BUF_NET_FLUX: src = w->net_flux; break;
        case RIBO_BUF_REDOX_EATEN: src = w->eaten; break;
        case RIBO_BUF_AGG: src = w->agg_e; break;
        case RIBO_BUF_FORCE_VECTORS: src = w->agent_force; break;
        case RIBO_BUF_SOLID_MASK:
            launch_mask_u32(w->stream, w->geo(), w->scratch_u32);
            HIPCHK(hipGetLastError());
            src = w->scratch_u32;
            break;
        case RIBO_BUF_VELOCITY_SMOOTH:  // the force buffer is dead between steps
            run(w, "smooth_velocity", [&] { launch_smooth_velocity(w->stream, w->geo(), w->vel_a, w->forces); });
            HIPCHK(hipGetLastError());
            src = w->forces;
            break;
        default: return fail(RIBO_ERR_ARG, "unknown buffer");
    }
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, w->stream));
    HIPCHK(hipStreamSynchronize(w->stream));
    return RIBO_OK;
}

// Iteration slots per solve for the next graph chunk of the tolerance mode
// (read after draining the stream, so the choice depends only on the
// sequence of calls): the most iterations any of the last SLOT_WINDOW
// solves took, plus max(3, 1/4 of it) (max(8, 1/2 of it) for the stationary
// V-cycle, whose counts move more), rounded up to an even count (fewer
// distinct graphs).  The choice is made again before every chunk of at most
// TOL_CHUNK steps, so it is never older than that.  The window spans two
// periods of the buoyancy noise (a new bucket every 16 epochs, fluid.wgsl:384,
// each adding ~8 iterations for a few steps on the 2048^2 benchmark world)
// and is short enough that the first solves of a cold start (31-53
// iterations there, 10-20 in steady flow) leave it soon; an unused slot is
// ten kernels that return at once (~3 us each).  A solve that ran out of
// slots (counted on the device) gives the next chunk twice the maximum + 4;
// until 8 solves have run with the current settings every solve gets the
// full cap.
constexpr uint32_t SLOT_WINDOW = 32u;
static_assert(SLOT_WINDOW <= PCG_HOST_RING_LEN, "the window lies in the host ring");
static int choose_slots(RiboWorld* w) {
    const uint32_t maxit = std::max(w->st.mg_max_cycles, 1u);
    HIPCHK(hipStreamSynchronize(w->stream));
    uint32_t slots = maxit;
    uint32_t missed = 0u;
    if (w->pcg_host) {
        missed = __atomic_load_n(&w->pcg_host[1], __ATOMIC_RELAXED);
        const uint32_t solves = __atomic_load_n(&w->pcg_host[2], __ATOMIC_RELAXED);
        const uint32_t known = std::min(solves - w->solves_at_settings, SLOT_WINDOW);
        uint32_t most = 0u;
        for (uint32_t k = 0; k < known; ++k)
            most = std::max(most, __atomic_load_n(&w->pcg_host[PCG_HOST_RING + (solves - 1u - k) % PCG_HOST_RING_LEN],
                                                  __ATOMIC_RELAXED));
        if (known >= 8u) {
            const uint32_t margin = w->st.mg_stationary ? std::max(8u, most / 2u) : std::max(3u, most / 4u);
            const uint32_t want = missed != w->missed_seen ? 2u * most + 4u : most + margin;
            slots = std::min(maxit, (want + 1u) & ~1u);
        }
        w->solves_seen = solves;
    }
    w->missed_seen = missed;
    w->pcg_slots = std::max(slots, 1u);
    return RIBO_OK;
}

int ribo_world_encode_steps(RiboWorld* w, uint32_t steps) {  // src/world/mod.rs:1406-1446,1622,1631
    if (!w) return fail(RIBO_ERR_ARG, "null world");
    int rc = validate_settings(&w->st);
    if (rc) return rc;
    steps = std::min(steps, RIBO_MAX_STEPS_PER_CALL);
    if (steps == 0) return 0;
    if (w->st.multigrid && w->st.mg_tol > 0.0f) {
        if (w->cinv_valid && w->cmoving != coarse_moving(w)) w->cinv_valid = false;
        if (!w->cinv_valid) {
            clear_graphs(w);
            rc = build_coarse_inverse(w);
            if (rc) return rc;
        }
        if ((rc = choose_slots(w))) return rc;
    }
    if ((rc = ensure_sun(w, steps))) return rc;
    const RiboWorldSettings st = w->st;
    const uint32_t substep = std::min(std::max(st.fluid_substep, 1u), 4u);
    if (!st.fluid_enabled && w->fluid_was_on) {
        const size_t F = (size_t)w->F * w->F;
        launch_fill_u32(w->stream, (uint32_t*)w->vel_a, 2 * F, 0u);
        launch_fill_u32(w->stream, (uint32_t*)w->temp[w->tcur], F, 0u);
        launch_fill_u32(w->stream, (uint32_t*)w->redox[w->tcur], F, 0u);
    }
    w->fluid_was_on = st.fluid_enabled != 0;
    w->params.fluid_dt = w->params.dt * (float)substep;
    w->params.fluid_enabled = st.fluid_enabled ? 1u : 0u;
    w->pipe = pipeline_ok(w);
    w->fuse = !w->pipe && fuse_ok(w, substep);
    if (w->pipe && (rc = ensure_pipeline(w))) return rc;
    if (w->pipe)  // the transport's copy of the velocity, until a fluid step writes it
        HIPCHK(hipMemcpyAsync(w->vel_s, w->vel_a, sizeof(float2) * (size_t)w->F * w->F, hipMemcpyDeviceToDevice, w->stream));
    if (graph_eligible(w)) {
        // any change of the by-value kernel arguments invalidates the graphs
        RiboSimParams a = w->params, b = w->graph_params;
        a.epoch = b.epoch = 0;
        if (memcmp(&a, &b, sizeof(a)) != 0 || memcmp(&w->st, &w->graph_settings, sizeof(w->st)) != 0) {
            clear_graphs(w);
            w->graph_params = w->params;
            w->graph_settings = w->st;
        }
        if (!w->epoch_dev_known || w->epoch_dev_value != w->params.epoch) {
            HIPCHK(hipMemcpyAsync(w->epoch_dev, &w->params.epoch, sizeof(uint32_t), hipMemcpyHostToDevice, w->stream));
            w->epoch_dev_value = w->params.epoch;
            w->epoch_dev_known = true;
        }
        const uint32_t te = std::max(w->params.transport_every, 1u);
        uint32_t period = substep;
        while (period % te != 0) period += substep;
        if (w->st.light_mode && w->st.light_rows_per_step == 0) {  // periodic sweeps
            const uint32_t li = std::max(w->st.light_interval, 1u);
            const uint32_t base = period;
            while (period % li != 0) period += base;
        }
        // tolerance-mode steps carry their solver iterations: shorter graphs
        // (the stationary mode records up to mg_max_cycles V-cycles per step:
        // one step per graph)
        const bool tol_mode = w->st.multigrid && w->st.mg_tol > 0.0f;
        const uint32_t max_graph = !tol_mode ? RIBO_MAX_STEPS_PER_CALL : (w->st.mg_stationary ? 1u : TOL_CHUNK);
        for (uint32_t first = 0; first < steps;) {
            const uint32_t n = std::min(max_graph, steps - first);
            if (tol_mode && first > 0 && (rc = choose_slots(w))) return rc;
            const uint32_t phase = (w->params.epoch + first) % period;
            RiboWorld::GraphEntry* ge = nullptr;
            for (auto& e : w->graphs)
                if (e.phase == phase && e.n == n && e.slots == w->pcg_slots && e.cur == w->cur && e.tcur == w->tcur &&
                    e.lcur == w->lcur)
                    ge = &e;
            if (!ge) {
                if (w->graphs.size() >= 32) clear_graphs(w);
                RiboWorld::GraphEntry e{phase, n, w->pcg_slots, w->cur, w->tcur, w->lcur, 0, 0, 0, nullptr};
                hipGraph_t graph = nullptr;
                std::lock_guard<std::recursive_mutex> lk(device_mutex());
                HIPCHK(hipStreamBeginCapture(w->stream, hipStreamCaptureModeThreadLocal));
                record_steps(w, first, n, substep, w->epoch_dev);
                launch_epoch_add(w->stream, w->epoch_dev, n);
                hipError_t ce = hipStreamEndCapture(w->stream, &graph);
                e.cur_end = w->cur;
                e.tcur_end = w->tcur;
                e.lcur_end = w->lcur;
                w->cur = e.cur;
                w->tcur = e.tcur;
                w->lcur = e.lcur;
                if (ce != hipSuccess) return fail(RIBO_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(ce));
                hipError_t ie = hipGraphInstantiate(&e.exec, graph, nullptr, nullptr, 0);
                (void)hipGraphDestroy(graph);
                if (ie != hipSuccess) return fail(RIBO_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(ie));
                w->graphs.push_back(e);
                ge = &w->graphs.back();
            }
            HIPCHK(hipGraphLaunch(ge->exec, w->stream));
            w->cur = ge->cur_end;
            w->tcur = ge->tcur_end;
            w->lcur = ge->lcur_end;
            w->epoch_dev_value += n;
            first += n;
        }
    } else if (w->st.multigrid && w->st.mg_tol > 0.0f) {
        for (uint32_t first = 0; first < steps; first += TOL_CHUNK) {
            if (first > 0 && (rc = choose_slots(w))) return rc;
            record_steps(w, first, std::min(TOL_CHUNK, steps - first), substep, nullptr);
        }
    } else {
        record_steps(w, 0, steps, substep, nullptr);
    }
    HIPCHK(hipGetLastError());
    w->params.epoch += steps;
    return (int)steps;
}

int ribo_world_run(RiboWorld* w, uint32_t steps) {
    if (!w) return fail(RIBO_ERR_ARG, "null world");
    uint32_t done = 0;
    while (done < steps) {
        int n = ribo_world_encode_steps(w, std::min(steps - done, RIBO_MAX_STEPS_PER_CALL));
        if (n < 0) return n;
        done += (uint32_t)n;
    }
    return RIBO_OK;
}

// Tile sides measured per (level side, V-cycle kind, storage) by any world
// of the process, so that many worlds of one size measure once.
std::mutex g_tune_mu;
std::map<std::tuple<uint32_t, int, int>, uint8_t> g_tune_cache;

int ribo_world_autotune(RiboWorld* w, int force) {
    if (!w) return fail(RIBO_ERR_ARG, "null world");
    if (force != 0 && force != 16 && force != 32) return fail(RIBO_ERR_ARG, "force must be 0, 16 or 32");
    const size_t L = w->levels.size();
    clear_graphs(w);
    const bool sym = w->st.multigrid && w->st.mg_tol > 0.0f && !w->st.mg_stationary;
    const int bf = sym && w->st.precision == RIBO_PRECISION_MIXED ? 1 : 0;
    // the levels that run as tile kernels: below the one-workgroup block
    // (level 0 of the symmetric cycle belongs to the conjugate-gradient
    // kernels), all but the coarsest when no suffix fits one workgroup
    const size_t lo = sym ? 1 : 0;
    const size_t hi = w->mg_block_start < L ? w->mg_block_start : L - 1;
    for (size_t l = 0; l < L; ++l) w->mg_tile[l] = force ? (uint8_t)force : 0;
    if (force || hi <= lo) return RIBO_OK;
    HIPCHK(hipStreamSynchronize(w->stream));
    // scratch levels (the world's masks, zeroed values): timing only, the
    // world's unknowns are not touched
    const size_t n0 = w->levels[lo].n;
    float* buf = nullptr;
    int rc = dalloc(w->stream, &buf, 6 * n0 * n0);
    if (rc) return rc;
    hipEvent_t e0 = get_event(w), e1 = get_event(w);
    auto measure = [&]() -> int {
        for (size_t l = lo; l < hi; ++l) {
            const uint32_t n = w->levels[l].n, nc = w->levels[l + 1].n;
            const auto key = std::make_tuple(n, sym ? 1 : 0, bf);
            {
                std::lock_guard<std::mutex> lk(g_tune_mu);
                auto it = g_tune_cache.find(key);
                if (it != g_tune_cache.end()) {
                    w->mg_tile[l] = it->second;
                    continue;
                }
            }
            MgLevelT<float> A{n, w->levels[l].eps, buf, buf + n0 * n0, buf + 2 * n0 * n0, w->mg[l].fluid};
            MgLevelT<float> C{nc, w->levels[l + 1].eps, buf + 3 * n0 * n0, buf + 4 * n0 * n0, buf + 5 * n0 * n0,
                              w->mg[l + 1].fluid};
            MgLevelT<bf16> Ah{n, A.eps, (bf16*)A.p, (bf16*)A.pp, (bf16*)A.rhs, A.fluid};
            MgLevelT<bf16> Ch{nc, C.eps, (bf16*)C.p, (bf16*)C.pp, (bf16*)C.rhs, C.fluid};
            float best = 0.0f;
            uint8_t pick = 0;
            for (int tile : {16, 32}) {
                auto cycle = [&]() {
                    if (bf) {
                        launch_mg_down<bf16>(w->stream, Ah, Ch, sym, nullptr, tile);
                        launch_mg_up<bf16>(w->stream, Ah, Ch, sym, nullptr, tile);
                    } else {
                        launch_mg_down<float>(w->stream, A, C, sym, nullptr, tile);
                        launch_mg_up<float>(w->stream, A, C, sym, nullptr, tile);
                    }
                };
                for (int k = 0; k < 3; ++k) cycle();
                HIPCHK(hipEventRecord(e0, w->stream));
                for (int k = 0; k < 20; ++k) cycle();
                HIPCHK(hipEventRecord(e1, w->stream));
                HIPCHK(hipEventSynchronize(e1));
                float ms = 0.0f;
                HIPCHK(hipEventElapsedTime(&ms, e0, e1));
                if (pick == 0 || ms < best) {
                    best = ms;
                    pick = (uint8_t)tile;
                }
            }
            w->mg_tile[l] = pick;
            std::lock_guard<std::mutex> lk(g_tune_mu);
            g_tune_cache[key] = pick;
        }
        return RIBO_OK;
    };
    rc = measure();
    (void)hipStreamSynchronize(w->stream);
    w->event_pool.push_back(e0);
    w->event_pool.push_back(e1);
    dfree(buf);
    return rc;
}

int ribo_world_mg_tiles(const RiboWorld* w, uint8_t* out, uint32_t cap, uint32_t* count) {
    if (!w || !count) return fail(RIBO_ERR_ARG, "null argument");
    *count = (uint32_t)w->mg_tile.size();
    for (uint32_t i = 0; i < cap && i < *count; ++i) out[i] = w->mg_tile[i];
    return RIBO_OK;
}

int ribo_world_invalidate_light(RiboWorld* w) {  // src/world/mod.rs:1400-1402
    if (!w) return fail(RIBO_ERR_ARG, "null world");
    w->light_dirty = true;
    return RIBO_OK;
}

int ribo_world_sync(RiboWorld* w) {
    if (!w) return fail(RIBO_ERR_ARG, "null world");
    HIPCHK(hipStreamSynchronize(w->stream));
    return RIBO_OK;
}
int ribo_world_idle(RiboWorld* w) {
    if (!w) return fail(RIBO_ERR_ARG, "null world");
    const hipError_t e = hipStreamQuery(w->stream);
    if (e == hipSuccess) return 1;
    if (e == hipErrorNotReady) return 0;
    return fail(RIBO_ERR_HIP, std::string("hipStreamQuery: ") + hipGetErrorString(e));
}

int ribo_world_chem_ops(RiboWorld* w, const RiboChemOp* ops, uint32_t n, uint32_t* done) {  // chem.wgsl:40-98
    if (!w || (n > 0 && !ops)) return fail(RIBO_ERR_ARG, "null argument");
    const uint64_t cells = (uint64_t)w->cfg.grid_size * w->cfg.grid_size;
    for (uint32_t i = 0; i < n; ++i) {
        if (ops[i].cell >= cells || ops[i].channel >= 4u || ops[i].kind > RIBO_CHEM_DEPOSIT_SPENT)
            return fail(RIBO_ERR_ARG, "exchange " + std::to_string(i) + ": cell, channel or kind out of range");
        // a deposit adds its count to a 16-bit field at once (chem_add_state,
        // chem.wgsl:70-74): more than 0xFFFF cannot be held by any cell
        if (ops[i].kind >= RIBO_CHEM_ADD_ACT && ops[i].count > 0xFFFFu)
            return fail(RIBO_ERR_ARG, "exchange " + std::to_string(i) + ": a deposit of more than 65535 monomers");
    }
    static_assert(sizeof(RiboChemOp) == 16, "an exchange is four words");
    // Between steps the force buffer (8 bytes a fluid cell) and the scratch
    // words are free: the exchanges go through them in batches of F / 2.
    const uint32_t batch = std::max<uint32_t>((uint32_t)std::min<size_t>((size_t)w->F * w->F / 2, 0xFFFFFFFFu), 1u);
    const Geo g = w->geo();
    for (uint32_t b = 0; b < n; b += batch) {
        const uint32_t m = std::min(batch, n - b);
        HIPCHK(hipMemcpyAsync(w->forces, ops + b, (size_t)m * sizeof(RiboChemOp), hipMemcpyHostToDevice, w->stream));
        run(w, "chem_ops",
            [&] { launch_chem_ops(w->stream, g, w->chem[w->cur], (const uint32_t*)w->forces, m, w->scratch_u32, w->ledger->expected); });
        HIPCHK(hipGetLastError());
        if (done) HIPCHK(hipMemcpyAsync(done + b, w->scratch_u32, (size_t)m * 4, hipMemcpyDeviceToHost, w->stream));
    }
    HIPCHK(hipStreamSynchronize(w->stream));
    return RIBO_OK;
}

int ribo_world_ledger(RiboWorld* w, uint32_t out[12]) {  // encode_ledger + ledger_blocking (src/world/mod.rs:1661-1674,1834-1839)
    if (!w || !out) return fail(RIBO_ERR_ARG, "null argument");
    launch_ledger(w->stream, w->geo(), w->chem[w->cur], w->ledger, w->params.epoch, nullptr, 0);
    HIPCHK(hipGetLastError());
    uint32_t last[8];
    HIPCHK(hipMemcpyAsync(last, w->ledger->last, sizeof(last), hipMemcpyDeviceToHost, w->stream));
    HIPCHK(hipStreamSynchronize(w->stream));
    for (int i = 0; i < 8; ++i) out[i] = last[i];
    for (int i = 8; i < 12; ++i) out[i] = 0u;
    return RIBO_OK;
}

int ribo_ledger_noalloc(const uint32_t* chem, uint64_t cells, void* stream, uint32_t out[12]) {
    if (!out || (cells > 0 && !chem)) return fail(RIBO_ERR_ARG, "null argument");
    if (cells > RIBO_LEDGER_NOALLOC_MAX_CELLS)
        return fail(RIBO_ERR_ARG, "ribo_ledger_noalloc: " + std::to_string(cells) + " cells exceed the maximum of " +
                                      std::to_string(RIBO_LEDGER_NOALLOC_MAX_CELLS) + " (a 32768^2 grid)");
    for (int i = 0; i < 12; ++i) out[i] = 0u;
    if (cells == 0) return RIBO_OK;
    LedgerDev* led = nullptr;
    const hipError_t e = ledger_scratch(&led);
    if (e != hipSuccess) return fail(RIBO_ERR_HIP, std::string("ribo_ledger_noalloc: scratch address: ") + hipGetErrorString(e));
    hipStream_t s = static_cast<hipStream_t>(stream);
    // k_ledger rewrites every partial k_ledger_fin reads, and k_ledger_fin
    // every total: nothing of the last call's scratch needs clearing
    launch_ledger_cells(s, chem, (uint32_t)cells, led);
    HIPCHK(hipGetLastError());
    uint32_t last[8];
    HIPCHK(hipMemcpyAsync(last, led->last, sizeof(last), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    for (int i = 0; i < 8; ++i) out[i] = last[i];
    return RIBO_OK;
}

int ribo_world_ledger_status(RiboWorld* w, uint32_t* epoch) {
    if (!w || !epoch) return fail(RIBO_ERR_ARG, "null argument");
    uint32_t fb = 0;
    HIPCHK(hipMemcpyAsync(&fb, &w->ledger->first_bad, sizeof(fb), hipMemcpyDeviceToHost, w->stream));
    HIPCHK(hipStreamSynchronize(w->stream));
    *epoch = fb;
    if (fb != 0xFFFFFFFFu) return fail(RIBO_ERR_LEDGER, "per-channel monomer totals changed at epoch " + std::to_string(fb));
    return RIBO_OK;
}

int ribo_world_set_pipeline(RiboWorld* w, int enable) {
    if (!w) return fail(RIBO_ERR_ARG, "null world");
    if (w->pipe_allowed != (enable != 0)) clear_graphs(w);  // recordings differ
    w->pipe_allowed = enable != 0;
    return RIBO_OK;
}
int ribo_world_set_timing(RiboWorld* w, int enable) {
    if (!w) return fail(RIBO_ERR_ARG, "null world");
    w->timing = enable != 0;
    return RIBO_OK;
}
int ribo_world_kernel_times(RiboWorld* w, const char** names, double* total_ms, uint64_t* launches, uint32_t cap,
                            uint32_t* count) {
    if (!w || !count) return fail(RIBO_ERR_ARG, "null argument");
    HIPCHK(hipStreamSynchronize(w->stream));
    resolve_timers(w);
    *count = (uint32_t)w->timers.size();
    for (uint32_t i = 0; i < cap && i < *count; ++i) {
        if (names) names[i] = w->timers[i].name;
        if (total_ms) total_ms[i] = w->timers[i].total_ms;
        if (launches) launches[i] = w->timers[i].launches;
    }
    return RIBO_OK;
}
int ribo_world_kernel_launch_ms(RiboWorld* w, const char* name, float* ms, uint32_t cap, uint32_t* count) {
    if (!w || !name || !count) return fail(RIBO_ERR_ARG, "null argument");
    HIPCHK(hipStreamSynchronize(w->stream));
    resolve_timers(w);
    *count = 0;
    for (const Timer& t : w->timers) {
        if (strcmp(t.name, name) != 0) continue;
        *count = (uint32_t)t.each.size();
        for (uint32_t i = 0; ms && i < cap && i < *count; ++i) ms[i] = t.each[i];
    }
    return RIBO_OK;
}
int ribo_world_reset_kernel_times(RiboWorld* w) {
    if (!w) return fail(RIBO_ERR_ARG, "null world");
    HIPCHK(hipStreamSynchronize(w->stream));
    resolve_timers(w);
    w->timers.clear();
    return RIBO_OK;
}
int ribo_world_mg_stats(RiboWorld* w, uint32_t* cycles, double* rel_residual) {
    if (!w || !cycles || !rel_residual) return fail(RIBO_ERR_ARG, "null argument");
    PcgState c;
    HIPCHK(hipMemcpyAsync(&c, w->pcg, sizeof(c), hipMemcpyDeviceToHost, w->stream));
    HIPCHK(hipStreamSynchronize(w->stream));
    if (w->st.mg_tol > 0.0f) {
        *cycles = c.it;
        *rel_residual = c.true_rel;
    } else {
        *cycles = std::max(w->st.mg_cycles, 1u);
        *rel_residual = -1.0;
    }
    return RIBO_OK;
}

int ribo_world_solver_stats(RiboWorld* w, uint64_t* solves, uint64_t* iterations, uint64_t* missed) {
    if (!w || !solves || !iterations || !missed) return fail(RIBO_ERR_ARG, "null argument");
    PcgState c;
    HIPCHK(hipMemcpyAsync(&c, w->pcg, sizeof(c), hipMemcpyDeviceToHost, w->stream));
    HIPCHK(hipStreamSynchronize(w->stream));
    *solves = c.steps;
    *iterations = c.iters;
    *missed = c.missed;
    return RIBO_OK;
}

int ribo_world_set_vents(RiboWorld* w, const RiboFumarole* f, uint32_t count, float gain, const float* heat_image,
                         const float* chem_image) {  // src/world/mod.rs:1343-1392
    if (!w || (count > 0 && !f)) return fail(RIBO_ERR_ARG, "null argument");
    const size_t G = (size_t)w->G * w->G, F = (size_t)w->F * w->F;
    std::vector<float> heat(G), chem(G), src(2 * F);
    int rc = ribo_rasterize_fumaroles(&w->cfg, f, count, heat.data());
    if (rc) return fail(rc, "ribo_world_set_vents: bad fumarole list");
    chem = heat;
    const float* cimg = chem_image ? chem_image : heat_image;
    if (heat_image)
        for (size_t i = 0; i < G; ++i) heat[i] += heat_image[i];
    if (cimg)
        for (size_t i = 0; i < G; ++i) chem[i] += cimg[i];
    if ((rc = ribo_vent_heat_src(&w->cfg, heat.data(), chem.data(), gain, src.data()))) return fail(rc, "ribo_vent_heat_src");
    return ribo_world_upload(w, RIBO_BUF_HEAT_SRC, src.data(), src.size() * sizeof(float));
}

int ribo_world_restart_matter(RiboWorld* w, uint64_t seed, float density, float active, uint32_t out[12]) {
    if (!w) return fail(RIBO_ERR_ARG, "null world");
    const size_t G = (size_t)w->G * w->G;
    std::vector<uint32_t> gamma(G), cells(4 * G);
    int rc;
    if ((rc = ribo_world_download(w, RIBO_BUF_GAMMA, gamma.data(), 4 * G))) return rc;
    if ((rc = ribo_seed_cells(&w->cfg, seed, density, active, cells.data()))) return fail(rc, "ribo_seed_cells");
    if ((rc = ribo_fit_matter_to_terrain(cells.data(), gamma.data(), G, seed))) return fail(rc, "ribo_fit_matter_to_terrain");
    if ((rc = ribo_world_upload(w, RIBO_BUF_CHEM, cells.data(), 16 * G))) return rc;
    w->params.epoch = 0;
    if (out) {
        uint64_t t[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (size_t i = 0; i < G; ++i)
            for (int c = 0; c < 4; ++c) {
                t[c] += cells[4 * i + c] & 0xFFFFu;
                t[4 + c] += cells[4 * i + c] >> 16;
            }
        for (int k = 0; k < 8; ++k) out[k] = (uint32_t)t[k];
        for (int k = 8; k < 12; ++k) out[k] = 0u;
    }
    return RIBO_OK;
}

int ribo_world_apply_terrain_live(RiboWorld* w, const uint32_t* gamma) {  // src/world/mod.rs:1283-1297
    if (!w || !gamma) return fail(RIBO_ERR_ARG, "null argument");
    const size_t G = (size_t)w->G * w->G;
    std::vector<uint32_t> cells(4 * G);
    int rc;
    if ((rc = ribo_world_download(w, RIBO_BUF_CHEM, cells.data(), 16 * G))) return rc;
    if ((rc = ribo_refit_matter_to_terrain(cells.data(), gamma, G))) return fail(rc, "ribo_refit_matter_to_terrain");
    if ((rc = ribo_world_upload(w, RIBO_BUF_GAMMA, gamma, 4 * G))) return rc;
    return ribo_world_upload(w, RIBO_BUF_CHEM, cells.data(), 16 * G);
}

int ribo_world_solver_residuals(RiboWorld* w, RiboSolverResiduals* out, int reset) {
    if (!w || !out) return fail(RIBO_ERR_ARG, "null argument");
    PcgState c;
    HIPCHK(hipMemcpyAsync(&c, w->pcg, sizeof(c), hipMemcpyDeviceToHost, w->stream));
    HIPCHK(hipStreamSynchronize(w->stream));
    out->last = c.true_rel;
    out->last_rounded = c.rounded_rel;
    out->max = c.max_true_rel;
    out->max_rounded = c.max_rounded_rel;
    out->restarts = c.restarts;
    if (reset) {
        struct { double a, b; uint32_t r; } z{0.0, 0.0, 0u};
        static_assert(offsetof(PcgState, max_rounded_rel) == offsetof(PcgState, max_true_rel) + 8 &&
                          offsetof(PcgState, restarts) == offsetof(PcgState, max_true_rel) + 16,
                      "the reset fields are consecutive");
        HIPCHK(hipMemcpyAsync((char*)w->pcg + offsetof(PcgState, max_true_rel), &z, 20, hipMemcpyHostToDevice, w->stream));
        HIPCHK(hipStreamSynchronize(w->stream));
    }
    return RIBO_OK;
}

}  // extern "C"

namespace ribo {
// The tolerance mode's set-up of a world before a chunk of batched steps:
// its coarsest inverse (built when missing, as ribo_world_encode_steps
// builds it) and its iteration slots for the chunk's solves.
int batch_prepare_solve(RiboWorld* w) {
    if (w->cinv_valid && w->cmoving != coarse_moving(w)) w->cinv_valid = false;
    if (!w->cinv_valid) {
        clear_graphs(w);
        const int rc = build_coarse_inverse(w);
        if (rc) return rc;
    }
    return choose_slots(w);
}
}  // namespace ribo
