// ribo_oracle.cpp - CPU restatement of the reference world step.
//
// TEST INFRASTRUCTURE ONLY: the parity oracle and CPU baseline for the
// world-step hot path.  It follows the WGSL kernels one by one, in the order
// World::encode_steps dispatches them (src/world/mod.rs:1406-1632), with
// the reference's ping-pong buffers ("ab" reads a / writes b, "ba" the
// reverse, src/world/mod.rs:622-633).  Nothing here is fused or reordered:
// clarity and faithfulness over speed.  Parallel loops (OpenMP) only run
// over cells whose updates are independent; integer scatter uses atomic adds
// exactly like the reference, so results do not depend on thread count.
//
// Paths in citations are relative to the reference repository root.

#include "ribo_oracle.h"

#include "../include/ribo_params_default.h"

#include <math.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <string>
#include <vector>

#include "../ribossome-v2_amd/csrc/wgsl_sem.h"

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// ---- constants --------------------------------------------------------------
// fluid.wgsl:15-56
constexpr float MAX_DT = 0.07f;
constexpr float MAX_FORCE = 200000.0f;
constexpr float MAX_VEL = 2000.0f;
constexpr float FORCE_SMOOTH_MIX = 0.25f;
constexpr float SOLID_PERM_THRESHOLD = 0.02f;
constexpr float TEMP_HEAT_RATE = 0.01f;
constexpr float TEMP_COOL_RATE = 0.12f;
constexpr float TEMP_MAX = 12.0f;
constexpr float SUN_HEAT_RATE = 0.15f;
constexpr float SUN_IR = 1.2f;
constexpr float SUN_IR_DEPTH = 0.08f;
constexpr float TEMP_BUOYANCY = 12.0f;
constexpr float TEMP_AMBIENT_SURFACE = 0.0f;
constexpr float TEMP_AMBIENT_ATTEN = 4.0f;
constexpr float TEMP_DIFFUSE = 0.08f;
constexpr float TEMP_ROCK_CONDUCT = 4.0f;
constexpr float REDOX_RATE = 0.01f;
constexpr float VENT_WAVES = 1.5f;
constexpr float REDOX_FP = 10000.0f;
constexpr float REDOX_MAX = 50.0f;
constexpr float RUBBLE_DRAG = 300.0f;
constexpr float FORCE_FP = 256.0f;
// transport.wgsl:20-51
constexpr float DIFF_HOP_P = 0.01f;
constexpr float DIFF_HOP_FLOOR = 0.15f;
constexpr float DIFF_AGITATION_SPEED = 0.5f;
constexpr float BURIED_DIFF_FACTOR = 1.0f;
constexpr float GAMMA_POROSITY_K = 0.3f;
constexpr float PRESSURE_HOP_P = 0.01f;
constexpr float RUBBLE_DISPERSION = 1.0f;
constexpr float CHEM_SQUEEZE_P = 0.25f;
constexpr float MONOMER_SETTLE_P = 0.002f;
constexpr uint32_t MAX_MOVERS_PER_CH = 256u;
constexpr float LIGHT_ACT_P = 0.006f;
constexpr float CHEM_SENSITIZE = 0.5f;
constexpr float CHEM_SHIELD = 0.5f;
constexpr float TEMP_ACT_THRESHOLD = 2.0f;
constexpr float FUMAROLE_ACT_P = 0.15f;
constexpr uint32_t SMOOTH_CELL = 16u;
// rng.wgsl:25-33
constexpr uint32_t S_PHOTO = 1u;
constexpr uint32_t S_DECAY = 2u;
constexpr uint32_t S_THERMAL = 3u;
constexpr uint32_t S_BUOYANCY = 4u;
constexpr uint32_t S_MOVE = 1u << 16;
constexpr uint32_t S_DISPERSE = 6u << 16;
// chem.wgsl:6-7, common/terrain.wgsl:4, src/shaders.rs:130
constexpr uint32_t CHEM_STATE_MASK = 0xFFFFu;
constexpr uint32_t CHEM_SPENT_ONE = 0x10000u;
constexpr uint32_t CHEM_CELL_CAP = 48u;
constexpr uint32_t GAMMA_SOLID_THRESHOLD = 3u;
// light.wgsl:18-40, 181
constexpr float UV_SHADOW_ABSORB = 0.6f;
constexpr int LIGHT_TAPS = 7;
constexpr float SUN_WATER_ABSORB = 0.05f;
constexpr uint32_t LIGHT_DIV = 2u;  // src/shaders.rs:132
// terrain.wgsl:6
constexpr float GAMMA_QUANTUM = 0.25f;
// src/world/mod.rs:195-202
constexpr uint32_t MG_PRE = 2, MG_POST = 2, MG_COARSE = 30;
constexpr float MG_EPS = 1e-4f;

int g_threads = 0;

// ---- RNG (rng.wgsl:10-50), restated independently of the product -----------
inline void pcg4d(uint32_t v[4]) {
    for (int i = 0; i < 4; ++i) v[i] = v[i] * 1664525u + 1013904223u;
    v[0] += v[1] * v[3];
    v[1] += v[2] * v[0];
    v[2] += v[0] * v[1];
    v[3] += v[1] * v[2];
    for (int i = 0; i < 4; ++i) v[i] ^= v[i] >> 16;
    v[0] += v[1] * v[3];
    v[1] += v[2] * v[0];
    v[2] += v[0] * v[1];
    v[3] += v[1] * v[2];
}
struct F4 { float x, y, z, w; };
inline F4 rng_f4(uint32_t key, uint32_t epoch, uint32_t stream, uint32_t seed) {
    uint32_t v[4] = {key, epoch, stream, seed};
    pcg4d(v);
    const float s = 1.0f / 16777216.0f;
    return {(float)(v[0] >> 8) * s, (float)(v[1] >> 8) * s, (float)(v[2] >> 8) * s, (float)(v[3] >> 8) * s};
}
inline uint32_t hash_u32(uint32_t v) {
    uint32_t x = v;
    x = x ^ (x >> 16); x = x * 0x7feb352du; x = x ^ (x >> 15); x = x * 0x846ca68bu; x = x ^ (x >> 16);
    return x;
}

// ---- field-type dependent transcendentals ------------------------------------
inline float ex(float x) { return rb_expf(x); }
inline double ex(double x) { return exp(x); }
inline float sn(float x) { return rb_sinf(x); }
inline double sn(double x) { return sin(x); }
inline float rint_(float x) { return rb_rint(x); }
inline double rint_(double x) { return nearbyint(x); }
inline float sq(float x) { return rb_sqrt(x); }
inline double sq(double x) { return sqrt(x); }
template <class R> inline R mn(R a, R b) { return b < a ? b : a; }
template <class R> inline R mx(R a, R b) { return a < b ? b : a; }
template <class R> inline R cl(R e, R lo, R hi) { return mn(mx(e, lo), hi); }
template <class R> inline R mixr(R a, R b, R t) { return a * (R(1) - t) + b * t; }
template <class R> inline R fractr(R x) { return x - floor(x); }
inline int32_t to_i32(float x) { return rb_f2i(x); }
inline int32_t to_i32(double x) {
    if (x != x) return 0;
    if (x <= -2147483648.0) return (int32_t)0x80000000u;
    if (x >= 2147483647.0) return 2147483647;
    return (int32_t)x;
}

template <class R> struct V2 { R x, y; };
template <class R> inline V2<R> vadd(V2<R> a, V2<R> b) { return {a.x + b.x, a.y + b.y}; }
template <class R> inline V2<R> vsub(V2<R> a, V2<R> b) { return {a.x - b.x, a.y - b.y}; }
template <class R> inline V2<R> vmul(V2<R> a, R s) { return {a.x * s, a.y * s}; }
template <class R> inline V2<R> vmix(V2<R> a, V2<R> b, R t) { return {mixr(a.x, b.x, t), mixr(a.y, b.y, t)}; }
template <class R> inline R vlen(V2<R> v) { return sq(v.x * v.x + v.y * v.y); }
template <class R> inline R vdot(V2<R> a, V2<R> b) { return a.x * b.x + a.y * b.y; }

template <class R> struct Fluid {
    std::vector<R> vel_a, vel_b, p_a, p_b, div, temp_in, temp_out, redox_in, redox_out, forces, net_flux;
    std::vector<R> mg_p, mg_rhs;
    std::vector<R> p_old;  // tolerance mode: the previous solve's pressure
    int hist = 0;          // solves since the history restarted (at most 2): at 2, p_a and p_old are two solves' results
    void alloc(size_t F, size_t N, size_t mg_total) {
        vel_a.assign(2 * F, 0); vel_b.assign(2 * F, 0);
        p_a.assign(F, 0); p_b.assign(F, 0); div.assign(F, 0); p_old.assign(F, 0);
        temp_in.assign(F, 0); temp_out.assign(F, 0); redox_in.assign(F, 0); redox_out.assign(F, 0);
        forces.assign(2 * F, 0); net_flux.assign(2 * N, 0);
        mg_p.assign(mg_total, 0); mg_rhs.assign(mg_total, 0);
    }
};

// Per-step uniforms: SimParams with epoch, sun_now and sun_slope of the step
// (src/world/mod.rs:1431-1445).
struct StepParams {
    RiboSimParams p;
};

}  // namespace

struct OracleWorld {
    uint32_t G = 0, F = 0, wupc = 30, Lsz = 0;
    float sim_size = 0;
    int fp64 = 0;
    RiboSimParams params;
    RiboWorldSettings st;
    bool fluid_was_on = true;  // src/world/mod.rs:929
    std::vector<uint32_t> chem, chem_next, gamma, agg_act, agg_nb, solid_mask, redox_eaten, mg_fluid;
    std::vector<uint32_t> force_vectors;
    std::vector<float> light, light_next, slope, heat_src;
    bool light_dirty = true;  // src/world/mod.rs:209
    std::vector<uint32_t> mg_n, mg_off;
    uint32_t mg_total = 0;
    Fluid<float> f32;
    Fluid<double> f64;
    double last_rel_res = -1.0;      // tolerance mode: the refined iterate's residual (binary64)
    double last_rounded_rel = -1.0;  // and that of the pressure as stored (binary32 in this world)
    uint32_t last_cycles = 0;
    uint32_t last_restarts = 0;
};

namespace {

// ---- host-side per-step values (src/params.rs:505-537) -----------------------
void day_phase(const RiboSimParams& p, uint32_t epoch, bool* day, double* phase) {
    double ph = (double)epoch / (double)p.day_period;
    ph = ph - floor(ph);
    double frac = std::min(std::max((double)p.day_fraction, 0.05), 1.0);
    *day = ph < frac;
    *phase = ph / frac;
}
float day_light(const RiboSimParams& p, uint32_t epoch) {
    if (p.day_period < 1.0f) return 1.0f;
    bool day; double ph;
    day_phase(p, epoch, &day, &ph);
    if (!day) return 0.0f;
    return (float)std::max(sin(M_PI * ph), 0.0);
}
float sun_slope_at(const RiboSimParams& p, uint32_t epoch) {
    if (p.day_period < 1.0f) return 0.0f;
    bool day; double ph;
    day_phase(p, epoch, &day, &ph);
    double max_angle = std::min(std::max((double)p.sun_angle, 0.0), 85.0) * (M_PI / 180.0);
    if (!day) ph = 1.0;
    double angle = (ph - 0.5) * 2.0 * max_angle;
    return (float)tan(angle);
}

// pow(clamp(fluid_decay, 0, 1), dt * 60) of advect_velocity (fluid.wgsl:473),
// uniform over the step: evaluated once on the host in binary64.
float decay_factor_f32(float decay, float dt) {
    return (float)pow((double)rb_clamp(decay, 0.0f, 1.0f), (double)(dt * 60.0f));
}
double decay_factor_f64(float decay, double dt) {
    return pow((double)rb_clamp(decay, 0.0f, 1.0f), dt * 60.0);
}

// ---- chem / terrain accessors (chem.wgsl:17-33, common/terrain.wgsl:6-8) ----
inline uint32_t cell_total(const uint32_t* chem, uint32_t cell) {
    uint32_t t = 0;
    for (uint32_t c = 0; c < 4; ++c) {
        uint32_t v = chem[cell * 4 + c];
        t += (v & CHEM_STATE_MASK) + (v >> 16);
    }
    return t;
}
inline uint32_t capacity(const uint32_t* gamma, uint32_t cell) {
    uint32_t g = gamma[cell];
    if (g >= GAMMA_SOLID_THRESHOLD) return 0;
    return CHEM_CELL_CAP * (GAMMA_SOLID_THRESHOLD - g) / GAMMA_SOLID_THRESHOLD;
}

// ---- fluid helpers (fluid.wgsl:58-273) ---------------------------------------
struct Ctx {
    const OracleWorld* w;
    uint32_t G, F;
    const RiboSimParams* p;
    uint32_t fg(uint32_t x, uint32_t y) const { return y * F + x; }
    bool solid(uint32_t x, uint32_t y) const { return w->solid_mask[fg(x, y)] != 0; }
    uint32_t gcount(uint32_t c) const { return w->gamma[c]; }
};

template <class R> R temp_ambient_at(const Ctx& c, uint32_t y) {  // fluid.wgsl:62-64
    return R(TEMP_AMBIENT_SURFACE) * ex(R(-TEMP_AMBIENT_ATTEN) * (R(1) - (R(y) + R(0.5)) / R(c.F)));
}
template <class R> R fluid_dt(const Ctx& c) { return cl<R>(R(c.p->fluid_dt), R(0), R(MAX_DT)); }  // :66-68
template <class R> bool is_bad(R x) { return (x != x) || (fabs((double)x) > (double)1e20f); }  // :70-72
template <class R> V2<R> sanitize(V2<R> v) {  // :74-77
    if (is_bad(v.x) || is_bad(v.y)) return {R(0), R(0)};
    return v;
}
template <class R> V2<R> clamp_len(V2<R> v, R m) {  // :79-83
    R len = vlen(v);
    if (len > m) return vmul(v, m / mx<R>(len, R(1e-12f)));
    return v;
}
inline uint32_t env_cell_for_fluid(const Ctx& c, uint32_t x, uint32_t y) {  // :87-92
    float s = (float)c.G / (float)c.F;
    uint32_t gx = rb_f2u(rb_clamp(((float)x + 0.5f) * s, 0.0f, (float)(c.G - 1u)));
    uint32_t gy = rb_f2u(rb_clamp(((float)y + 0.5f) * s, 0.0f, (float)(c.G - 1u)));
    return gy * c.G + gx;
}
inline float gamma_solidity_at_fluid_cell(const Ctx& c, uint32_t x, uint32_t y) {  // :95-109
    float scale = (float)c.G / (float)c.F;
    float gx = ((float)x + 0.5f) * scale;
    float gy = ((float)y + 0.5f) * scale;
    float h = scale * 0.25f;
    uint32_t n = 0;
    for (uint32_t s = 0; s < 4; ++s) {
        float ox = (s & 1u) == 1u ? h : -h;
        float oy = (s & 2u) == 2u ? h : -h;
        uint32_t sx = rb_f2u(rb_clamp(gx + ox, 0.0f, (float)(c.G - 1u)));
        uint32_t sy = rb_f2u(rb_clamp(gy + oy, 0.0f, (float)(c.G - 1u)));
        if (c.gcount(sy * c.G + sx) >= GAMMA_SOLID_THRESHOLD) n += 1;
    }
    return (float)n * 0.25f;
}
template <class R> R gamma_fill_at_fluid_cell(const Ctx& c, uint32_t x, uint32_t y) {  // :114-128
    float scale = (float)c.G / (float)c.F;
    float gx = ((float)x + 0.5f) * scale;
    float gy = ((float)y + 0.5f) * scale;
    float h = scale * 0.25f;
    R fill = 0;
    for (uint32_t s = 0; s < 4; ++s) {
        float ox = (s & 1u) == 1u ? h : -h;
        float oy = (s & 2u) == 2u ? h : -h;
        uint32_t sx = rb_f2u(rb_clamp(gx + ox, 0.0f, (float)(c.G - 1u)));
        uint32_t sy = rb_f2u(rb_clamp(gy + oy, 0.0f, (float)(c.G - 1u)));
        fill += R(std::min(c.gcount(sy * c.G + sx), GAMMA_SOLID_THRESHOLD)) / R(GAMMA_SOLID_THRESHOLD);
    }
    return fill * R(0.25);
}
inline float permeability(const Ctx& c, uint32_t x, uint32_t y) {  // :138-142
    float solidity = gamma_solidity_at_fluid_cell(c, x, y);
    if (solidity >= 0.5f) return 0.0f;
    return rb_clamp(1.0f - solidity * 2.0f, 0.0f, 1.0f);
}
template <class R> R rubble_drag(const Ctx& c, uint32_t x, uint32_t y, R dt) {  // :150-161
    uint32_t scale = c.G / c.F;
    R g = 0;
    for (uint32_t dy = 0; dy < scale; ++dy)
        for (uint32_t dx = 0; dx < scale; ++dx) {
            uint32_t cnt = c.gcount((y * scale + dy) * c.G + x * scale + dx);
            if (cnt < GAMMA_SOLID_THRESHOLD) g += R(cnt);
        }
    g /= R(scale * scale) * R(GAMMA_SOLID_THRESHOLD);
    return R(1) / (R(1) + R(RUBBLE_DRAG) * g * dt);
}
template <class R> V2<R> slope_steer_velocity(const Ctx& c, uint32_t x, uint32_t y, V2<R> v_in, R dt) {  // :165-178
    uint32_t e = env_cell_for_fluid(c, x, y);
    V2<R> sr = sanitize(V2<R>{R(c.w->slope[2 * e]), R(c.w->slope[2 * e + 1])});
    V2<R> s = {-sr.x, -sr.y};
    R v_len = vlen(v_in);
    R s_len = vlen(s);
    if (v_len < R(1e-5f) || s_len < R(1e-5f)) return v_in;
    V2<R> v_dir = {v_in.x / v_len, v_in.y / v_len};
    V2<R> s_dir = {s.x / s_len, s.y / s_len};
    R misalign = cl<R>(R(1) - vdot(v_dir, s_dir), R(0), R(2));
    R t = cl<R>(R(1) - ex(-mx<R>(R(c.p->slope_steer_rate), R(0)) * misalign * dt), R(0), R(1));
    V2<R> dir_raw = vadd(v_dir, vmul(vsub(s_dir, v_dir), t));
    R dir_len = vlen(dir_raw);
    if (dir_len < R(1e-5f)) return v_in;
    return vmul(V2<R>{dir_raw.x / dir_len, dir_raw.y / dir_len}, v_len);
}
template <class R> V2<R> raw_velocity_cell(const Ctx& c, const R* vin, uint32_t x, uint32_t y) {  // :191-194
    if (c.solid(x, y)) return {R(0), R(0)};
    uint32_t i = c.fg(x, y);
    return sanitize(V2<R>{vin[2 * i], vin[2 * i + 1]});
}
template <class R> V2<R> solid_normal_from_neighbors(const Ctx& c, uint32_t x, uint32_t y) {  // :197-216
    const uint32_t F = c.F;
    uint32_t xm = x == 0u ? x : x - 1u;
    uint32_t xp = x + 1u >= F ? x : x + 1u;
    uint32_t ym = y == 0u ? y : y - 1u;
    uint32_t yp = y + 1u >= F ? y : y + 1u;
    R s_l = (x > 0u && c.solid(xm, y)) ? R(1) : R(0);
    R s_r = (x + 1u < F && c.solid(xp, y)) ? R(1) : R(0);
    R s_b = (y > 0u && c.solid(x, ym)) ? R(1) : R(0);
    R s_t = (y + 1u < F && c.solid(x, yp)) ? R(1) : R(0);
    R s_bl = (x > 0u && y > 0u && c.solid(xm, ym)) ? R(1) : R(0);
    R s_tl = (x > 0u && y + 1u < F && c.solid(xm, yp)) ? R(1) : R(0);
    R s_br = (x + 1u < F && y > 0u && c.solid(xp, ym)) ? R(1) : R(0);
    R s_tr = (x + 1u < F && y + 1u < F && c.solid(xp, yp)) ? R(1) : R(0);
    R diag = R(0.70710678f);
    R nx = (s_l - s_r) + diag * ((s_bl + s_tl) - (s_br + s_tr));
    R ny = (s_b - s_t) + diag * ((s_bl + s_br) - (s_tl + s_tr));
    V2<R> n = {nx, ny};
    R n_len = vlen(n);
    return n_len > R(1e-6f) ? V2<R>{n.x / n_len, n.y / n_len} : V2<R>{R(0), R(0)};
}
template <class R> V2<R> reflect_if_into_solid(const Ctx& c, uint32_t x, uint32_t y, V2<R> v) {  // :219-224
    V2<R> n = solid_normal_from_neighbors<R>(c, x, y);
    R d = vdot(v, n);
    if (vlen(n) > R(1e-6f) && d < R(0)) return vsub(v, vmul(n, d));
    return v;
}
template <class R> bool is_solid_at_pos(const Ctx& c, V2<R> pos) {  // :226-231
    int32_t ix = to_i32(floor(pos.x - R(0.5)));
    int32_t iy = to_i32(floor(pos.y - R(0.5)));
    if (ix < 0 || ix >= (int32_t)c.F || iy < 0 || iy >= (int32_t)c.F) return true;
    return c.solid((uint32_t)ix, (uint32_t)iy);
}
template <class R> V2<R> sample_velocity(const Ctx& c, const R* vin, V2<R> pos) {  // :238-255
    R min_pos = R(0.5), max_pos = R(c.F) - R(0.5);
    if (pos.x < min_pos || pos.x > max_pos || pos.y < min_pos || pos.y > max_pos) return {R(0), R(0)};
    R x = pos.x - R(0.5), y = pos.y - R(0.5);
    int32_t x0 = to_i32(floor(x)), y0 = to_i32(floor(y));
    R fx = fractr(x), fy = fractr(y);
    auto cc = [&](int32_t a) { return (uint32_t)rb_clampi(a, 0, (int32_t)c.F - 1); };
    V2<R> v0 = vmix(raw_velocity_cell(c, vin, cc(x0), cc(y0)), raw_velocity_cell(c, vin, cc(x0 + 1), cc(y0)), fx);
    V2<R> v1 = vmix(raw_velocity_cell(c, vin, cc(x0), cc(y0 + 1)), raw_velocity_cell(c, vin, cc(x0 + 1), cc(y0 + 1)), fx);
    return vmix(v0, v1, fy);
}
// Transport sampler, binary32 (fluid.wgsl:260-273).
V2<float> fluid_velocity_at_world(const Ctx& c, const float* vin, float sim_size, V2<float> pw) {
    float g = (float)c.F;
    float gx = rb_clamp(pw.x / sim_size * g, 0.5f, g - 0.5f) - 0.5f;
    float gy = rb_clamp(pw.y / sim_size * g, 0.5f, g - 0.5f) - 0.5f;
    uint32_t x0 = rb_f2u(rb_floor(gx)), y0 = rb_f2u(rb_floor(gy));
    uint32_t x1 = std::min(x0 + 1u, c.F - 1u), y1 = std::min(y0 + 1u, c.F - 1u);
    float tx = rb_fract(gx), ty = rb_fract(gy);
    auto at = [&](uint32_t x, uint32_t y) {
        uint32_t i = c.fg(x, y);
        return sanitize(V2<float>{vin[2 * i], vin[2 * i + 1]});
    };
    V2<float> v0 = vmix(at(x0, y0), at(x1, y0), tx);
    V2<float> v1 = vmix(at(x0, y1), at(x1, y1), tx);
    return vmix(v0, v1, ty);
}

// ---- light helpers on the frozen light field (light.wgsl:46-102,182-189) ----
float light_above(const OracleWorld& w, const RiboSimParams& p, uint32_t x, uint32_t row, uint32_t band) {
    const uint32_t LS = w.Lsz, LC = LS * LS;
    float sh = -p.sun_slope;
    int32_t base = rb_f2i(rb_floor(sh));
    float f = sh - (float)base;
    float s = 0.0f;
    for (int d = 0; d < LIGHT_TAPS; ++d) {
        int32_t x0 = rb_clampi((int32_t)x + d - LIGHT_TAPS / 2 + base, 0, (int32_t)LS - 1);
        int32_t x1 = rb_clampi(x0 + 1, 0, (int32_t)LS - 1);
        s += rb_mix(w.light[band * LC + row * LS + (uint32_t)x0], w.light[band * LC + row * LS + (uint32_t)x1], f);
    }
    float cpos = (float)x + sh;
    float inside = rb_clamp(cpos + 1.0f, 0.0f, 1.0f) * rb_clamp((float)LS - cpos, 0.0f, 1.0f);
    return s / (float)LIGHT_TAPS * inside;
}
float light_transmit(const OracleWorld& w, const RiboSimParams& p, uint32_t lx, uint32_t ly, uint32_t band) {
    // Agent shade is zero: agents are not part of the world-step path.
    float g = 0.0f;
    uint32_t m = 0;
    for (uint32_t dy = 0; dy < LIGHT_DIV; ++dy)
        for (uint32_t dx = 0; dx < LIGHT_DIV; ++dx) {
            uint32_t cc = (ly * LIGHT_DIV + dy) * w.G + lx * LIGHT_DIV + dx;
            g += (float)w.gamma[cc];
            m += cell_total(w.chem.data(), cc);
        }
    g /= (float)LIGHT_DIV;
    float dens = (float)m / (float)(LIGHT_DIV * LIGHT_DIV);
    float tau_m = rb_max(p.monomer_uv_absorb, 0.0f) * dens * (float)LIGHT_DIV / (float)w.G * (band == 1u ? 0.35f : 1.0f);
    float shade = 0.0f;
    return rb_expf(-UV_SHADOW_ABSORB * g - 0.15f * shade - tau_m);
}
float light_absorbed_at(const OracleWorld& w, const RiboSimParams& p, uint32_t env) {
    uint32_t lx = (env % w.G) / LIGHT_DIV, ly = (env / w.G) / LIGHT_DIV;
    float incoming = p.sun_now;
    if (ly + 1u < w.Lsz) incoming = light_above(w, p, lx, ly + 1u, 0u);
    incoming = rb_max(incoming, 0.0f);
    return incoming * ((1.0f - light_transmit(w, p, lx, ly, 0u)) + SUN_WATER_ABSORB);
}
inline float uv_light_at_cell(const OracleWorld& w, uint32_t x, uint32_t y) {  // light.wgsl:46-50
    uint32_t lx = std::min(x, w.G - 1u) / LIGHT_DIV, ly = std::min(y, w.G - 1u) / LIGHT_DIV;
    return w.light[ly * w.Lsz + lx];
}

// ---- light propagation (light.wgsl:44-46,104-176) ------------------------------
float row_water_of(const OracleWorld& w, const RiboSimParams& p, uint32_t band) {  // light.wgsl:44-46
    return rb_expf(-rb_max(p.uv_depth, 0.0f) * (band == 1u ? 0.03f : 1.0f) * (float)LIGHT_DIV / (float)w.G);
}
// light_transmit_pass (light.wgsl:121-127): every light cell's transmission,
// both bands, written into the light field.
void light_transmit_pass(OracleWorld& w, const RiboSimParams& p) {
    const uint32_t LS = w.Lsz, LC = LS * LS;
    for (uint32_t i = 0; i < LC; ++i) {
        w.light[i] = light_transmit(w, p, i % LS, i / LS, 0u);
        w.light[LC + i] = light_transmit(w, p, i % LS, i / LS, 1u);
    }
}
// compute_uv_light (light.wgsl:129-160): top row times the sun and the water,
// then row by row downwards, each cell times the light above and the water.
void compute_uv_light(OracleWorld& w, const RiboSimParams& p) {
    const uint32_t LS = w.Lsz, LC = LS * LS, top = LS - 1u;
    for (uint32_t x = 0; x < LS; ++x)
        for (uint32_t band = 0; band < 2u; ++band) w.light[band * LC + top * LS + x] *= row_water_of(w, p, band) * p.sun_now;
    for (uint32_t t = 1; t < LS; ++t) {
        const uint32_t y = LS - 1u - t;
        for (uint32_t x = 0; x < LS; ++x)
            for (uint32_t band = 0; band < 2u; ++band) {
                float above = light_above(w, p, x, y + 1u, band);
                w.light[band * LC + y * LS + x] *= above * row_water_of(w, p, band);
            }
    }
}
// light_propagate + light_commit (light.wgsl:162-183): every cell takes the
// row above from the previous field.
void light_propagate(OracleWorld& w, const RiboSimParams& p) {
    const uint32_t LS = w.Lsz, LC = LS * LS;
    for (uint32_t i = 0; i < LC; ++i) {
        const uint32_t x = i % LS, y = i / LS;
        for (uint32_t band = 0; band < 2u; ++band) {
            float above = p.sun_now;
            if (y + 1u < LS) above = light_above(w, p, x, y + 1u, band);
            w.light_next[band * LC + i] = above * row_water_of(w, p, band) * light_transmit(w, p, x, y, band);
        }
    }
    w.light = w.light_next;
}
// The light step of encode_steps (src/world/mod.rs:1506-1525).
void light_step(OracleWorld& w, const RiboSimParams& p) {
    if (!w.st.light_mode) return;
    const bool sweep = w.light_dirty || (w.st.light_rows_per_step == 0 && p.epoch % std::max(w.st.light_interval, 1u) == 0);
    if (sweep) {
        w.light_dirty = false;
        light_transmit_pass(w, p);
        compute_uv_light(w, p);
    } else {
        for (uint32_t k = 0; k < std::min(w.st.light_rows_per_step, 16u); ++k) light_propagate(w, p);
    }
}

template <class R> R vent_cycle(const Ctx& c, uint32_t x) {  // fluid.wgsl:46-54
    R m = R(1);
    if (c.p->vent_cycle_amp > 0.0f && c.p->vent_cycle_period >= 1.0f) {
        uint32_t period = rb_f2u(c.p->vent_cycle_period);
        R t = R(c.p->epoch % period) / R(period);
        m = mx<R>(R(1) + R(c.p->vent_cycle_amp) * sn(R(6.2831853f) * (t + R(VENT_WAVES) * R(x) / R(c.F))), R(0));
    }
    return m;
}

// ---- fluid kernels ------------------------------------------------------------
template <class R> void build_solid_mask(OracleWorld& w, const Ctx& c) {  // fluid.wgsl:185-189
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)c.F * c.F; ++i) {
        uint32_t x = (uint32_t)(i % c.F), y = (uint32_t)(i / c.F);
        w.solid_mask[i] = permeability(c, x, y) < SOLID_PERM_THRESHOLD ? 1u : 0u;
    }
}

template <class R> void update_temperature(OracleWorld& w, Fluid<R>& fl, const Ctx& c) {  // fluid.wgsl:300-363
    const uint32_t F = c.F;
    const R* vin = fl.vel_a.data();
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)F * F; ++ii) {
        uint32_t x = (uint32_t)(ii % F), y = (uint32_t)(ii / F);
        uint32_t idx = c.fg(x, y);
        R dt = fluid_dt<R>(c);
        V2<R> vel = raw_velocity_cell(c, vin, x, y);
        V2<R> src = vsub(V2<R>{R(x) + R(0.5), R(y) + R(0.5)}, vmul(vel, dt));
        R sxf = cl<R>(src.x - R(0.5), R(0), R(F - 1u));
        R syf = cl<R>(src.y - R(0.5), R(0), R(F - 1u));
        uint32_t x0 = (uint32_t)floor(sxf), y0 = (uint32_t)floor(syf);
        uint32_t x1 = std::min(x0 + 1u, F - 1u), y1 = std::min(y0 + 1u, F - 1u);
        R fx = fractr(sxf), fy = fractr(syf);
        const R* T = fl.temp_in.data();
        R t = mixr(mixr(T[c.fg(x0, y0)], T[c.fg(x1, y0)], fx), mixr(T[c.fg(x0, y1)], T[c.fg(x1, y1)], fx), fy);
        R tl = T[c.fg(x == 0 ? 0 : x - 1, y)];
        R tr = T[c.fg(std::min(x + 1u, F - 1u), y)];
        R tb = T[c.fg(x, y == 0 ? 0 : y - 1)];
        R tt = T[c.fg(x, std::min(y + 1u, F - 1u))];
        R fill = gamma_fill_at_fluid_cell<R>(c, x, y);
        t = mixr(t, (tl + tr + tb + tt) * R(0.25), R(TEMP_DIFFUSE) * mixr(R(1), R(TEMP_ROCK_CONDUCT), fill));
        t += R(w.heat_src[2 * idx]) * vent_cycle<R>(c, x) * R(TEMP_HEAT_RATE) * dt;
        if (c.p->sun_heat > 0.0f) {
            uint32_t env = env_cell_for_fluid(c, x, y);
            R depth = R(1) - (R(y) + R(0.5)) / R(F);
            R ir = R(c.p->sun_now) * R(SUN_IR) * ex(-depth / R(SUN_IR_DEPTH));
            t += R(mx(c.p->sun_heat, 0.0f)) * R(SUN_HEAT_RATE) * (R(light_absorbed_at(w, *c.p, env)) + ir) * dt;
        }
        R amb = temp_ambient_at<R>(c, y);
        t = amb + (t - amb) * ex(R(-TEMP_COOL_RATE) * (R(1) - fill) * dt);
        fl.temp_out[idx] = cl<R>(t, R(0), R(TEMP_MAX));

        const R* Q = fl.redox_in.data();
        R r = mixr(mixr(Q[c.fg(x0, y0)], Q[c.fg(x1, y0)], fx), mixr(Q[c.fg(x0, y1)], Q[c.fg(x1, y1)], fx), fy);
        R rl = Q[c.fg(x == 0 ? 0 : x - 1, y)];
        R rr = Q[c.fg(std::min(x + 1u, F - 1u), y)];
        R rb = Q[c.fg(x, y == 0 ? 0 : y - 1)];
        R rt = Q[c.fg(x, std::min(y + 1u, F - 1u))];
        r = mixr(r, (rl + rr + rb + rt) * R(0.25), R(TEMP_DIFFUSE));
        r += R(w.heat_src[2 * idx + 1]) * vent_cycle<R>(c, x) * R(REDOX_RATE) * dt;
        r = mx<R>(r - R((float)w.redox_eaten[idx]) / R(REDOX_FP), R(0));
        r *= ex(-mx<R>(R(c.p->redox_decay), R(0)) * dt);
        fl.redox_out[idx] = cl<R>(r, R(0), R(REDOX_MAX));
    }
}

template <class R> void copy_temperature(OracleWorld& w, Fluid<R>& fl) {  // fluid.wgsl:365-372
    fl.temp_in = fl.temp_out;
    fl.redox_in = fl.redox_out;
    std::fill(w.redox_eaten.begin(), w.redox_eaten.end(), 0u);
}

template <class R> void atomic_add_force(OracleWorld& w, uint32_t i, R v) {  // fluid.wgsl:281-284
    int32_t q = to_i32(rint_(cl<R>(v, R(-1.0e5f), R(1.0e5f)) * R(FORCE_FP)));
    if (q != 0) w.force_vectors[i] += (uint32_t)q;
}
template <class R> R force_at_fp(const OracleWorld& w, uint32_t i) {  // :286-288
    return R((int32_t)w.force_vectors[i]) / R(FORCE_FP);
}

template <class R> void buoyancy(OracleWorld& w, Fluid<R>& fl, const Ctx& c) {  // fluid.wgsl:376-389
    const uint32_t F = c.F;
    // Each cell adds only to its own two words: no contention.
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)F * F; ++ii) {
        uint32_t x = (uint32_t)(ii % F), y = (uint32_t)(ii / F);
        if (x == 0u || y == 0u || x >= F - 1u || y >= F - 1u) continue;
        uint32_t idx = c.fg(x, y);
        R dev = fl.temp_in[idx] - temp_ambient_at<R>(c, y);
        if (fabs((double)dev) <= 1e-3f) continue;
        uint32_t time_bucket = c.p->epoch / 16u;
        R ad = R(fabs((double)dev));
        R side = (R(rng_f4(idx, time_bucket, S_BUOYANCY, c.p->seed).x) * R(2) - R(1)) * R(0.25) * ad / mx<R>(ad, R(1));
        R fx = side * R(TEMP_BUOYANCY) * ad;
        R fy = R(TEMP_BUOYANCY) * dev;
        atomic_add_force<R>(w, idx * 2u, fx);
        atomic_add_force<R>(w, idx * 2u + 1u, fy);
    }
}

template <class R> void gather_forces(OracleWorld& w, Fluid<R>& fl, const Ctx& c) {  // fluid.wgsl:392-398
    for (uint32_t i = 0; i < c.F * c.F; ++i) {
        V2<R> f = {force_at_fp<R>(w, 2 * i), force_at_fp<R>(w, 2 * i + 1)};
        V2<R> g = clamp_len(sanitize(f), R(MAX_FORCE));
        fl.forces[2 * i] = g.x;
        fl.forces[2 * i + 1] = g.y;
    }
}

template <class R> void add_forces(OracleWorld& w, Fluid<R>& fl, const Ctx& c, const R* vin, R* vout) {  // :410-432
    const uint32_t F = c.F;
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)F * F; ++ii) {
        uint32_t x = (uint32_t)(ii % F), y = (uint32_t)(ii / F);
        uint32_t idx = c.fg(x, y);
        if (c.solid(x, y)) { vout[2 * idx] = 0; vout[2 * idx + 1] = 0; continue; }
        R dt = fluid_dt<R>(c);
        uint32_t xm = x > 0u ? x - 1u : x, xp = x + 1u < F ? x + 1u : x;
        uint32_t ym = y > 0u ? y - 1u : y, yp = y + 1u < F ? y + 1u : y;
        auto fa = [&](uint32_t a, uint32_t b) {
            uint32_t i = c.fg(a, b);
            return sanitize(V2<R>{fl.forces[2 * i], fl.forces[2 * i + 1]});
        };
        V2<R> f_c = fa(x, y);
        V2<R> axis = vadd(vadd(vadd(fa(xm, y), fa(xp, y)), fa(x, ym)), fa(x, yp));
        V2<R> diag = vadd(vadd(vadd(fa(xm, ym), fa(xp, ym)), fa(xm, yp)), fa(xp, yp));
        V2<R> f_avg = vmul(vadd(vadd(vmul(axis, R(4)), diag), vmul(f_c, R(20))), R(1.0f / 36.0f));
        V2<R> f_user = clamp_len(vmix(f_c, f_avg, R(FORCE_SMOOTH_MIX)), R(MAX_FORCE));
        V2<R> v = sanitize(vadd(sanitize(V2<R>{vin[2 * idx], vin[2 * idx + 1]}), vmul(f_user, dt)));
        v = slope_steer_velocity<R>(c, x, y, v, dt);
        v = vmul(v, rubble_drag<R>(c, x, y, dt));
        v = reflect_if_into_solid<R>(c, x, y, v);
        v = clamp_len(v, R(MAX_VEL));
        vout[2 * idx] = v.x;
        vout[2 * idx + 1] = v.y;
    }
}

template <class R> void diffuse_velocity(const Ctx& c, const R* vin, R* vout) {  // fluid.wgsl:435-456
    const uint32_t F = c.F;
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)F * F; ++ii) {
        uint32_t x = (uint32_t)(ii % F), y = (uint32_t)(ii / F);
        uint32_t idx = c.fg(x, y);
        if (c.solid(x, y)) { vout[2 * idx] = 0; vout[2 * idx + 1] = 0; continue; }
        if (x == 0u || x == F - 1u || y == 0u || y == F - 1u) {
            vout[2 * idx] = vin[2 * idx];
            vout[2 * idx + 1] = vin[2 * idx + 1];
            continue;
        }
        auto rv = [&](uint32_t a, uint32_t b) { return raw_velocity_cell(c, vin, a, b); };
        V2<R> v_c = rv(x, y);
        V2<R> axis_sum = vadd(vadd(vadd(rv(x - 1, y), rv(x + 1, y)), rv(x, y - 1)), rv(x, y + 1));
        V2<R> diag_sum = vadd(vadd(vadd(rv(x - 1, y - 1), rv(x + 1, y - 1)), rv(x - 1, y + 1)), rv(x + 1, y + 1));
        V2<R> lap = vmul(vsub(vadd(vmul(axis_sum, R(4)), diag_sum), vmul(v_c, R(20))), R(1.0f / 6.0f));
        R a = mx<R>(R(c.p->fluid_viscosity), R(0)) * fluid_dt<R>(c);
        V2<R> v = sanitize(vadd(v_c, vmul(lap, a)));
        v = reflect_if_into_solid<R>(c, x, y, v);
        v = clamp_len(v, R(MAX_VEL));
        vout[2 * idx] = v.x;
        vout[2 * idx + 1] = v.y;
    }
}

template <class R> void advect_velocity(const Ctx& c, const R* vin, R* vout, R decay_factor) {  // :458-477
    const uint32_t F = c.F;
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)F * F; ++ii) {
        uint32_t x = (uint32_t)(ii % F), y = (uint32_t)(ii / F);
        uint32_t idx = c.fg(x, y);
        if (c.solid(x, y)) { vout[2 * idx] = 0; vout[2 * idx + 1] = 0; continue; }
        V2<R> pos = {R(x) + R(0.5), R(y) + R(0.5)};
        R dt = fluid_dt<R>(c);
        V2<R> vel = reflect_if_into_solid<R>(c, x, y, clamp_len(raw_velocity_cell(c, vin, x, y), R(MAX_VEL)));
        V2<R> trace = vsub(pos, vmul(vel, dt));
        if (is_solid_at_pos<R>(c, trace)) trace = vsub(pos, vmul(reflect_if_into_solid<R>(c, x, y, vel), dt));
        V2<R> out = sanitize(vmul(sample_velocity<R>(c, vin, trace), decay_factor));
        out = reflect_if_into_solid<R>(c, x, y, out);
        out = clamp_len(out, R(MAX_VEL));
        vout[2 * idx] = out.x;
        vout[2 * idx + 1] = out.y;
    }
}

template <class R> R curl_at(const Ctx& c, const R* vin, uint32_t x, uint32_t y) {  // fluid.wgsl:479-489
    int32_t Fm = (int32_t)c.F - 1;
    uint32_t xm = (uint32_t)rb_clampi((int32_t)x - 1, 0, Fm), xp = (uint32_t)rb_clampi((int32_t)x + 1, 0, Fm);
    uint32_t ym = (uint32_t)rb_clampi((int32_t)y - 1, 0, Fm), yp = (uint32_t)rb_clampi((int32_t)y + 1, 0, Fm);
    V2<R> v_l = reflect_if_into_solid<R>(c, xm, y, raw_velocity_cell(c, vin, xm, y));
    V2<R> v_r = reflect_if_into_solid<R>(c, xp, y, raw_velocity_cell(c, vin, xp, y));
    V2<R> v_b = reflect_if_into_solid<R>(c, x, ym, raw_velocity_cell(c, vin, x, ym));
    V2<R> v_t = reflect_if_into_solid<R>(c, x, yp, raw_velocity_cell(c, vin, x, yp));
    return R(0.5) * (v_r.y - v_l.y) - R(0.5) * (v_t.x - v_b.x);
}

template <class R> void vorticity_confinement(const Ctx& c, const R* vin, R* vout) {  // fluid.wgsl:491-518
    const uint32_t F = c.F;
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)F * F; ++ii) {
        uint32_t x = (uint32_t)(ii % F), y = (uint32_t)(ii / F);
        uint32_t idx = c.fg(x, y);
        if (c.solid(x, y)) { vout[2 * idx] = 0; vout[2 * idx + 1] = 0; continue; }
        const uint32_t border = 2u;
        R eps = cl<R>(R(c.p->fluid_vorticity), R(0), R(10));
        if (x < border || x >= F - border || y < border || y >= F - border || eps < R(1e-6f)) {
            vout[2 * idx] = vin[2 * idx]; vout[2 * idx + 1] = vin[2 * idx + 1];
            continue;
        }
        R wv = curl_at<R>(c, vin, x, y);
        if (fabs((double)wv) < 5e-4f) {
            vout[2 * idx] = vin[2 * idx]; vout[2 * idx + 1] = vin[2 * idx + 1];
            continue;
        }
        V2<R> grad = vmul(V2<R>{R(fabs((double)curl_at<R>(c, vin, x + 1, y))) - R(fabs((double)curl_at<R>(c, vin, x - 1, y))),
                                R(fabs((double)curl_at<R>(c, vin, x, y + 1))) - R(fabs((double)curl_at<R>(c, vin, x, y - 1)))},
                          R(0.5));
        R gl = mx<R>(vlen(grad), R(1e-5f));
        V2<R> n = {grad.x / gl, grad.y / gl};
        V2<R> f = vmul(V2<R>{n.y, -n.x}, wv * eps);
        V2<R> dv = clamp_len(vmul(f, fluid_dt<R>(c)), R(2));
        V2<R> v = sanitize(vadd(V2<R>{vin[2 * idx], vin[2 * idx + 1]}, dv));
        v = reflect_if_into_solid<R>(c, x, y, v);
        v = clamp_len(v, R(MAX_VEL));
        vout[2 * idx] = v.x;
        vout[2 * idx + 1] = v.y;
    }
}

template <class R> V2<R> div_neighbor(const Ctx& c, const R* vin, V2<R> v_c, int32_t nx, int32_t ny, bool axis_x) {  // :525-531
    bool outside = nx < 0 || ny < 0 || nx >= (int32_t)c.F || ny >= (int32_t)c.F;
    if (outside || c.solid((uint32_t)nx, (uint32_t)ny)) return axis_x ? V2<R>{-v_c.x, v_c.y} : V2<R>{v_c.x, -v_c.y};
    return raw_velocity_cell(c, vin, (uint32_t)nx, (uint32_t)ny);
}
template <class R> void compute_divergence(const Ctx& c, const R* vin, R* div) {  // fluid.wgsl:533-551
    const uint32_t F = c.F;
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)F * F; ++ii) {
        uint32_t x = (uint32_t)(ii % F), y = (uint32_t)(ii / F);
        uint32_t idx = c.fg(x, y);
        if (c.solid(x, y)) { div[idx] = 0; continue; }
        V2<R> v_c = raw_velocity_cell(c, vin, x, y);
        int32_t xi = (int32_t)x, yi = (int32_t)y;
        V2<R> v_l = div_neighbor<R>(c, vin, v_c, xi - 1, yi, true);
        V2<R> v_r = div_neighbor<R>(c, vin, v_c, xi + 1, yi, true);
        V2<R> v_b = div_neighbor<R>(c, vin, v_c, xi, yi - 1, false);
        V2<R> v_t = div_neighbor<R>(c, vin, v_c, xi, yi + 1, false);
        div[idx] = R(0.5) * ((v_r.x - v_l.x) + (v_t.y - v_b.y));
    }
}

template <class R> void jacobi_pressure(const Ctx& c, const R* pin, R* pout, const R* div) {  // fluid.wgsl:553-608
    const uint32_t F = c.F;
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)F * F; ++ii) {
        int32_t x = (int32_t)(ii % F), y = (int32_t)(ii / F);
        uint32_t idx = c.fg(x, y);
        auto sol = [&](int32_t a, int32_t b) {
            if (a < 0 || b < 0 || a >= (int32_t)F || b >= (int32_t)F) return true;
            return c.solid((uint32_t)a, (uint32_t)b);
        };
        auto pv = [&](int32_t a, int32_t b) -> R {
            if (a < 0 || b < 0 || a >= (int32_t)F || b >= (int32_t)F) return R(0);
            return pin[c.fg((uint32_t)a, (uint32_t)b)];
        };
        if (sol(x, y)) { pout[idx] = 0; continue; }
        R p_c = pv(x, y);
        R p_l = sol(x - 1, y) ? p_c : pv(x - 1, y);
        R p_r = sol(x + 1, y) ? p_c : pv(x + 1, y);
        R p_b = sol(x, y - 1) ? p_c : pv(x, y - 1);
        R p_t = sol(x, y + 1) ? p_c : pv(x, y + 1);
        R d = div[idx];
        if (is_bad(d) || is_bad(p_c)) { pout[idx] = 0; continue; }
        R p_new = (p_l + p_r + p_b + p_t - d) * R(0.25);
        pout[idx] = is_bad(p_new) ? R(0) : p_new;
    }
}

template <class R> void subtract_gradient(const Ctx& c, const R* vin, R* vout, const R* P) {  // fluid.wgsl:610-630
    const uint32_t F = c.F;
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)F * F; ++ii) {
        uint32_t x = (uint32_t)(ii % F), y = (uint32_t)(ii / F);
        uint32_t idx = c.fg(x, y);
        if (c.solid(x, y)) { vout[2 * idx] = 0; vout[2 * idx + 1] = 0; continue; }
        R p_c = P[idx];
        R p_l = p_c, p_r = p_c, p_b = p_c, p_t = p_c;
        if (x > 0u && !c.solid(x - 1u, y)) p_l = P[c.fg(x - 1u, y)];
        if (x + 1u < F && !c.solid(x + 1u, y)) p_r = P[c.fg(x + 1u, y)];
        if (y > 0u && !c.solid(x, y - 1u)) p_b = P[c.fg(x, y - 1u)];
        if (y + 1u < F && !c.solid(x, y + 1u)) p_t = P[c.fg(x, y + 1u)];
        V2<R> grad = vmul(V2<R>{p_r - p_l, p_t - p_b}, R(0.5));
        V2<R> v = vsub(sanitize(V2<R>{vin[2 * idx], vin[2 * idx + 1]}), grad);
        v = reflect_if_into_solid<R>(c, x, y, v);
        v = clamp_len(sanitize(v), R(MAX_VEL));
        vout[2 * idx] = v.x;
        vout[2 * idx + 1] = v.y;
    }
}

template <class R> void net_flux(const Ctx& c, const R* vin, R* nf) {  // fluid.wgsl:640-657
    const uint32_t F = c.F;
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)2 * F; ++ii) {
        uint32_t i = (uint32_t)ii;
        bool row = i < F;
        uint32_t j = row ? i : i - F;
        R sum = 0, n = 0;
        for (uint32_t k = 0; k < F; ++k) {  // sequential order, as the reference
            uint32_t x = row ? k : j, y = row ? j : k;
            if (c.solid(x, y)) continue;
            uint32_t q = c.fg(x, y);
            V2<R> v = sanitize(V2<R>{vin[2 * q], vin[2 * q + 1]});
            sum += row ? v.y : v.x;
            n += R(1);
        }
        nf[i] = n > R(0) ? sum / n : R(0);
    }
}

template <class R> void enforce_boundaries(const Ctx& c, const R* vin, R* vout, const R* nf) {  // fluid.wgsl:663-677
    const uint32_t F = c.F;
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)F * F; ++ii) {
        uint32_t x = (uint32_t)(ii % F), y = (uint32_t)(ii / F);
        uint32_t idx = c.fg(x, y);
        if (c.solid(x, y)) { vout[2 * idx] = 0; vout[2 * idx + 1] = 0; continue; }
        V2<R> v = sanitize(V2<R>{vin[2 * idx], vin[2 * idx + 1]});
        v = vsub(v, V2<R>{nf[F + x], nf[y]});
        if (x == 0u || x == F - 1u) v.x = 0;
        if (y == 0u || y == F - 1u) v.y = 0;
        v = reflect_if_into_solid<R>(c, x, y, v);
        v = clamp_len(sanitize(v), R(MAX_VEL));
        vout[2 * idx] = v.x;
        vout[2 * idx + 1] = v.y;
    }
}

// ---- multigrid (multigrid.wgsl:25-121; schedule src/world/mod.rs:1545-1577) ----
struct Level { uint32_t n, off, n_c, off_c; float eps; };

// bfloat16 storage of the mixed-precision preconditioner: round to nearest
// even on the 16 high bits.
inline float round_bf16(float x) {
    uint32_t u = rb_as_u32(x);
    if ((u & 0x7F800000u) == 0x7F800000u) return x;
    u += 0x7FFFu + ((u >> 16) & 1u);
    return rb_as_f32(u & 0xFFFF0000u);
}

// Exact solve of the coarsest level of the tolerance mode: the dense inverse
// of its eps-anchored operator over fluid cells, in binary64 (Gauss-Jordan
// with partial pivoting; -A is symmetric positive definite).
struct CoarseInv {
    uint32_t n = 0;
    std::vector<int32_t> idx;   // cell -> fluid unknown, -1 for walls
    std::vector<uint32_t> cell; // unknown -> cell
    std::vector<double> inv;    // m x m, row major
    bool build(uint32_t n_, const uint32_t* fluid, float eps) {
        n = n_;
        idx.assign((size_t)n * n, -1);
        cell.clear();
        for (uint32_t i = 0; i < n * n; ++i)
            if (fluid[i]) { idx[i] = (int32_t)cell.size(); cell.push_back(i); }
        const size_t m = cell.size();
        if (m == 0 || m > 1024) return false;
        std::vector<double> a(m * m, 0.0);
        inv.assign(m * m, 0.0);
        for (size_t r = 0; r < m; ++r) {
            const uint32_t c = cell[r], x = c % n, y = c / n;
            int k = 0;
            const int32_t nbr[4] = {x > 0 ? idx[c - 1] : -1, x + 1 < n ? idx[c + 1] : -1,
                                    y > 0 ? idx[c - n] : -1, y + 1 < n ? idx[c + n] : -1};
            for (int d = 0; d < 4; ++d)
                if (nbr[d] >= 0) { a[r * m + (size_t)nbr[d]] = 1.0; ++k; }
            a[r * m + r] = -((double)k + (double)eps);
            inv[r * m + r] = 1.0;
        }
        for (size_t c = 0; c < m; ++c) {
            size_t piv = c;
            for (size_t r = c + 1; r < m; ++r)
                if (fabs(a[r * m + c]) > fabs(a[piv * m + c])) piv = r;
            if (piv != c)
                for (size_t j = 0; j < m; ++j) { std::swap(a[c * m + j], a[piv * m + j]); std::swap(inv[c * m + j], inv[piv * m + j]); }
            const double d = a[c * m + c];
            for (size_t j = 0; j < m; ++j) { a[c * m + j] /= d; inv[c * m + j] /= d; }
            for (size_t r = 0; r < m; ++r) {
                if (r == c) continue;
                const double f = a[r * m + c];
                if (f == 0.0) continue;
                for (size_t j = 0; j < m; ++j) { a[r * m + j] -= f * a[c * m + j]; inv[r * m + j] -= f * inv[c * m + j]; }
            }
        }
        return true;
    }
};

template <class R> struct MG {
    const std::vector<Level>* lv;
    R* p;
    R* rhs;
    uint32_t* fluid;
    bool bf16 = false;              // round stores of the preconditioner to bfloat16
    const CoarseInv* cinv = nullptr;  // exact solve of the last level of vcycle_sym (tolerance mode)
    size_t sym_levels = 0;            // levels vcycle_sym descends through (0: all)
    std::vector<R> rn_scratch;        // res / W of a fine level during restriction
    bool is_fluid(const Level& L, int32_t x, int32_t y) const {  // :29-32
        if (x < 0 || y < 0 || x >= (int32_t)L.n || y >= (int32_t)L.n) return false;
        return fluid[L.off + (uint32_t)y * L.n + (uint32_t)x] != 0u;
    }
    void neighbors(const Level& L, uint32_t x, uint32_t y, R* s, R* k) const {  // :35-45
        R ss = 0, kk = 0;
        int32_t xi = (int32_t)x, yi = (int32_t)y;
        if (is_fluid(L, xi - 1, yi)) { ss += p[L.off + y * L.n + x - 1u]; kk += R(1); }
        if (is_fluid(L, xi + 1, yi)) { ss += p[L.off + y * L.n + x + 1u]; kk += R(1); }
        if (is_fluid(L, xi, yi - 1)) { ss += p[L.off + (y - 1u) * L.n + x]; kk += R(1); }
        if (is_fluid(L, xi, yi + 1)) { ss += p[L.off + (y + 1u) * L.n + x]; kk += R(1); }
        *s = ss; *k = kk;
    }
    void smooth(const Level& L, uint32_t color) {  // :60-77
#pragma omp parallel for schedule(static)
        for (int64_t ii = 0; ii < (int64_t)L.n * L.n; ++ii) {
            uint32_t x = (uint32_t)(ii % L.n), y = (uint32_t)(ii / L.n);
            if (((x + y) & 1u) != color) continue;
            uint32_t i = L.off + y * L.n + x;
            if (fluid[i] == 0u) continue;
            R s, k;
            neighbors(L, x, y, &s, &k);
            p[i] = (s - rhs[i]) / (k + R(L.eps));
        }
    }
    void restrict_(const Level& L) {  // :82-105
#pragma omp parallel for schedule(static)
        for (int64_t ii = 0; ii < (int64_t)L.n_c * L.n_c; ++ii) {
            uint32_t gx = (uint32_t)(ii % L.n_c), gy = (uint32_t)(ii / L.n_c);
            R sum = 0;
            uint32_t nf = 0;
            for (uint32_t c = 0; c < 4; ++c) {
                uint32_t fx = gx * 2u + (c & 1u), fy = gy * 2u + (c >> 1);
                uint32_t i = L.off + fy * L.n + fx;
                if (fluid[i] == 0u) continue;
                R s, k;
                neighbors(L, fx, fy, &s, &k);
                sum += rhs[i] - (s - (k + R(L.eps)) * p[i]);
                nf += 1;
            }
            uint32_t ci = L.off_c + gy * L.n_c + gx;
            fluid[ci] = nf > 0 ? 1u : 0u;
            rhs[ci] = nf > 0 ? R(2) * sum / R(std::max(nf, 1u)) : R(0);
            p[ci] = 0;
        }
    }
    void prolong(const Level& L) {  // :108-114
#pragma omp parallel for schedule(static)
        for (int64_t ii = 0; ii < (int64_t)L.n * L.n; ++ii) {
            uint32_t x = (uint32_t)(ii % L.n), y = (uint32_t)(ii / L.n);
            uint32_t i = L.off + y * L.n + x;
            if (fluid[i] == 0u) continue;
            p[i] += p[L.off_c + (y / 2u) * L.n_c + x / 2u];
        }
    }
    void vcycle() {  // src/world/mod.rs:1555-1575
        const std::vector<Level>& L = *lv;
        size_t levels = L.size();
        for (size_t l = 0; l + 1 < levels; ++l) {
            for (uint32_t s = 0; s < MG_PRE; ++s) { smooth(L[l], 0); smooth(L[l], 1); }
            restrict_(L[l]);
        }
        for (uint32_t s = 0; s < MG_COARSE; ++s) { smooth(L[levels - 1], 0); smooth(L[levels - 1], 1); }
        for (size_t l = levels - 1; l-- > 0;) {
            prolong(L[l]);
            for (uint32_t s = 0; s < MG_POST; ++s) { smooth(L[l], 0); smooth(L[l], 1); }
        }
    }
    void round_level(const Level& L, R* a) {
        if (!bf16) return;
        for (uint32_t i = 0; i < L.n * L.n; ++i) a[L.off + i] = R(round_bf16((float)a[L.off + i]));
    }
    // Bilinear (cell-centred) transfer of the tolerance mode between a level
    // of even side n and its coarse level (n / 2).  In each direction fine
    // index f takes its parent a0 = f / 2 with weight 3/4 and the coarse
    // cell on its side, a1 (clamped at the edge, where it repeats a0), with
    // weight 1/4.  The four products, in the order (a0y, a0x), (a0y, a1x),
    // (a1y, a0x), (a1y, a1x), count only fluid coarse cells and are divided
    // by the sum W of their weights (W >= 9/16: the parent of a fluid cell
    // is fluid).  Restriction is the transpose: a fluid coarse cell gathers
    // wt(f, c) res_f / W_f over its 4 x 4 fine cells 2c - 1 .. 2c + 2, rows
    // then columns, where wt(f, c) = 3/4 [a0(f) = c] + 1/4 [a1(f) = c].
    static void bl_nf(uint32_t f, uint32_t nc, uint32_t* a0, uint32_t* a1) {
        *a0 = f >> 1;
        *a1 = (f & 1u) ? std::min(*a0 + 1u, nc - 1u) : (*a0 > 0u ? *a0 - 1u : 0u);
    }
    static R bl_wt(uint32_t f, uint32_t c, uint32_t nc) {
        uint32_t a0, a1;
        bl_nf(f, nc, &a0, &a1);
        return (a0 == c ? R(0.75) : R(0)) + (a1 == c ? R(0.25) : R(0));
    }
    // W_f and the interpolated coarse value at fine cell (fx, fy) of L.
    void bl_interp(const Level& L, uint32_t fx, uint32_t fy, R* wsum, R* val) const {
        const uint32_t nc = L.n_c;
        uint32_t x0, x1, y0, y1;
        bl_nf(fx, nc, &x0, &x1);
        bl_nf(fy, nc, &y0, &y1);
        const uint32_t cx[4] = {x0, x1, x0, x1}, cy[4] = {y0, y0, y1, y1};
        const R wk[4] = {R(0.5625), R(0.1875), R(0.1875), R(0.0625)};
        R acc = 0, ws = 0;
        for (int k = 0; k < 4; ++k) {
            const uint32_t ci = L.off_c + cy[k] * nc + cx[k];
            const bool f = fluid[ci] != 0u;
            acc = acc + (f ? wk[k] * p[ci] : R(0));
            ws = ws + (f ? wk[k] : R(0));
        }
        *wsum = ws;
        *val = acc;
    }
    void restrict_bl(const Level& Lf, std::vector<R>& rn) {
        const uint32_t n = Lf.n, nc = Lf.n_c;
        rn.assign((size_t)n * n, R(0));
        for (uint32_t fy = 0; fy < n; ++fy)
            for (uint32_t fx = 0; fx < n; ++fx) {
                const uint32_t i = Lf.off + fy * n + fx;
                if (fluid[i] == 0u) continue;
                R s, k, ws, v;
                neighbors(Lf, fx, fy, &s, &k);
                bl_interp(Lf, fx, fy, &ws, &v);
                rn[(size_t)fy * n + fx] = (rhs[i] - (s - (k + R(Lf.eps)) * p[i])) / ws;
            }
        for (uint32_t gy = 0; gy < nc; ++gy)
            for (uint32_t gx = 0; gx < nc; ++gx) {
                R sum = 0;
                for (int dy = -1; dy <= 2; ++dy) {
                    const int64_t fy = 2 * (int64_t)gy + dy;
                    for (int dx = -1; dx <= 2; ++dx) {
                        const int64_t fx = 2 * (int64_t)gx + dx;
                        if (fy < 0 || fx < 0 || fy >= n || fx >= n) continue;
                        const R w = bl_wt((uint32_t)fy, gy, nc) * bl_wt((uint32_t)fx, gx, nc);
                        sum = sum + w * rn[(size_t)fy * n + (size_t)fx];
                    }
                }
                const uint32_t ci = Lf.off_c + gy * nc + gx;
                rhs[ci] = fluid[ci] ? sum : R(0);
                p[ci] = 0;
            }
    }
    void prolong_bl(const Level& L) {
        for (uint32_t fy = 0; fy < L.n; ++fy)
            for (uint32_t fx = 0; fx < L.n; ++fx) {
                const uint32_t i = L.off + fy * L.n + fx;
                if (fluid[i] == 0u) continue;
                R ws, v;
                bl_interp(L, fx, fy, &ws, &v);
                p[i] += v / ws;
            }
    }
    // Restriction proportional to the transpose of the injection: half the
    // sum over fluid children (the reference's twice their mean, on
    // all-fluid cells).
    void restrict_half_sum(const Level& Lf) {
        for (uint32_t gy = 0; gy < Lf.n_c; ++gy)
            for (uint32_t gx = 0; gx < Lf.n_c; ++gx) {
                R sum = 0;
                uint32_t nf = 0;
                for (uint32_t c = 0; c < 4; ++c) {
                    const uint32_t fx = gx * 2u + (c & 1u), fy = gy * 2u + (c >> 1);
                    const uint32_t i = Lf.off + fy * Lf.n + fx;
                    if (fluid[i] == 0u) continue;
                    R s, k;
                    neighbors(Lf, fx, fy, &s, &k);
                    sum += rhs[i] - (s - (k + R(Lf.eps)) * p[i]);
                    nf += 1;
                }
                const uint32_t ci = Lf.off_c + gy * Lf.n_c + gx;
                rhs[ci] = nf > 0 ? R(0.5) * sum : R(0);
                p[ci] = 0;
            }
    }
    // The symmetric V-cycle of the tolerance mode: the reference V-cycle with
    // the post-smoother in reverse colour order (black, red), a symmetric
    // transfer -- the bilinear one above in binary32 and binary64; with
    // bfloat16 storage, injection and half the sum over fluid children
    // (rounding the bilinear cycle's output to bfloat16 stalls conjugate
    // gradients) -- and an exact (or colour-symmetric) coarsest solve.  It
    // is a symmetric preconditioner, as conjugate gradients requires.
    void vcycle_sym() {
        const std::vector<Level>& L = *lv;
        const size_t levels = sym_levels ? std::min(sym_levels, L.size()) : L.size();
        for (size_t l = 0; l + 1 < levels; ++l) {
            for (uint32_t s2 = 0; s2 < MG_PRE; ++s2) { smooth(L[l], 0); smooth(L[l], 1); }
            round_level(L[l], p);
            if (bf16) restrict_half_sum(L[l]);
            else restrict_bl(L[l], rn_scratch);
            round_level(L[l + 1], rhs);
        }
        const Level& C = L[levels - 1];
        if (cinv) {
            const size_t m = cinv->cell.size();
            std::vector<double> b(m);
            for (size_t r = 0; r < m; ++r) b[r] = (double)rhs[C.off + cinv->cell[r]];
            for (size_t r = 0; r < m; ++r) {
                double acc = 0.0;
                for (size_t j = 0; j < m; ++j) acc += cinv->inv[r * m + j] * b[j];
                p[C.off + cinv->cell[r]] = R(acc);
            }
        } else {
            for (uint32_t s2 = 0; s2 < MG_COARSE / 2; ++s2) { smooth(C, 0); smooth(C, 1); }
            for (uint32_t s2 = 0; s2 < MG_COARSE / 2; ++s2) { smooth(C, 1); smooth(C, 0); }
        }
        for (size_t l = levels - 1; l-- > 0;) {
            if (bf16) prolong(L[l]);
            else prolong_bl(L[l]);
            for (uint32_t s2 = 0; s2 < MG_POST; ++s2) { smooth(L[l], 1); smooth(L[l], 0); }
            round_level(L[l], p);
        }
    }
    // A v on fluid cells of level 0: sum of fluid neighbours - (k + eps) v.
    void apply(const R* v, R* out) const {
        const Level& L = (*lv)[0];
        for (uint32_t y = 0; y < L.n; ++y)
            for (uint32_t x = 0; x < L.n; ++x) {
                const uint32_t i = y * L.n + x;
                if (fluid[L.off + i] == 0u) { out[i] = 0; continue; }
                R s = 0, k = 0;
                if (x > 0 && fluid[L.off + i - 1]) { s += v[i - 1]; k += R(1); }
                if (x + 1 < L.n && fluid[L.off + i + 1]) { s += v[i + 1]; k += R(1); }
                if (y > 0 && fluid[L.off + i - L.n]) { s += v[i - L.n]; k += R(1); }
                if (y + 1 < L.n && fluid[L.off + i + L.n]) { s += v[i + L.n]; k += R(1); }
                out[i] = s - (k + R(L.eps)) * v[i];
            }
    }
    // A v in binary64 on fluid cells of level 0, for v = hi + lo formed in
    // binary64 per cell (lo may be null): the fluid neighbours summed in the
    // order left, right, down, up, minus (k + eps) v; 0 in walls.
    void apply64(const R* hi, const R* lo, double* out) const {
        const Level& L = (*lv)[0];
        auto val = [&](uint32_t j) { return (double)hi[j] + (lo ? (double)lo[j] : 0.0); };
        for (uint32_t y = 0; y < L.n; ++y)
            for (uint32_t x = 0; x < L.n; ++x) {
                const uint32_t i = y * L.n + x;
                if (fluid[L.off + i] == 0u) { out[i] = 0.0; continue; }
                double s = 0.0, k = 0.0;
                if (x > 0 && fluid[L.off + i - 1]) { s += val(i - 1); k += 1.0; }
                if (x + 1 < L.n && fluid[L.off + i + 1]) { s += val(i + 1); k += 1.0; }
                if (y > 0 && fluid[L.off + i - L.n]) { s += val(i - L.n); k += 1.0; }
                if (y + 1 < L.n && fluid[L.off + i + L.n]) { s += val(i + L.n); k += 1.0; }
                out[i] = s - (k + (double)L.eps) * val(i);
            }
    }
    // ||b - A p||2 / ||b||2 over fluid cells of level 0 for level-0 vectors.
    double rel_residual_of(const R* p0, const R* b) const {
        const Level& L = (*lv)[0];
        double rr = 0, bb = 0;
        for (uint32_t y = 0; y < L.n; ++y)
            for (uint32_t x = 0; x < L.n; ++x) {
                const uint32_t i = y * L.n + x;
                if (fluid[L.off + i] == 0u) continue;
                double s = 0, k = 0;
                if (x > 0 && fluid[L.off + i - 1]) { s += (double)p0[i - 1]; k += 1; }
                if (x + 1 < L.n && fluid[L.off + i + 1]) { s += (double)p0[i + 1]; k += 1; }
                if (y > 0 && fluid[L.off + i - L.n]) { s += (double)p0[i - L.n]; k += 1; }
                if (y + 1 < L.n && fluid[L.off + i + L.n]) { s += (double)p0[i + L.n]; k += 1; }
                const double r = (double)b[i] - (s - (k + (double)L.eps) * (double)p0[i]);
                rr += r * r;
                bb += (double)b[i] * (double)b[i];
            }
        return bb > 0 ? sqrt(rr / bb) : 0.0;
    }
    // ||rhs - A p||2 / ||rhs||2 over fluid cells of level 0 (binary64 sums).
    double rel_residual() const {
        const Level& L = (*lv)[0];
        double rr = 0, bb = 0;
        for (uint32_t y = 0; y < L.n; ++y)
            for (uint32_t x = 0; x < L.n; ++x) {
                uint32_t i = L.off + y * L.n + x;
                if (fluid[i] == 0u) continue;
                R s, k;
                neighbors(L, x, y, &s, &k);
                double r = (double)rhs[i] - ((double)s - ((double)k + (double)L.eps) * (double)p[i]);
                rr += r * r;
                bb += (double)rhs[i] * (double)rhs[i];
            }
        return bb > 0 ? sqrt(rr / bb) : 0.0;
    }
};

std::vector<Level> make_levels(uint32_t n0, uint32_t min_size, uint32_t* total) {  // src/world/mod.rs:637-671
    std::vector<uint32_t> sizes{n0};
    while (sizes.back() > min_size && sizes.back() % 2 == 0) sizes.push_back(sizes.back() / 2);
    std::vector<uint32_t> offs{0};
    for (uint32_t n : sizes) offs.push_back(offs.back() + n * n);
    std::vector<Level> lv;
    for (size_t l = 0; l < sizes.size(); ++l) {
        Level L;
        L.n = sizes[l];
        L.off = offs[l];
        L.n_c = l + 1 < sizes.size() ? sizes[l + 1] : 0;
        L.off_c = offs[l + 1];
        L.eps = MG_EPS * (float)(1u << l);
        lv.push_back(L);
    }
    *total = offs.back();
    return lv;
}

// ---- transport (transport.wgsl:61-438) -----------------------------------------
struct TCtx {
    const OracleWorld* w;
    const RiboSimParams* p;
    const uint32_t* chem;
    const uint32_t* gamma;
    uint32_t G, F;
};

V2<float> lattice_rand(const TCtx& t, uint32_t ix, uint32_t iy) {  // transport.wgsl:97-99
    F4 r = rng_f4((ix * 73856093u) ^ (iy * 19349663u), t.p->epoch, S_MOVE + MAX_MOVERS_PER_CH + 1u, t.p->seed);
    return {r.x, r.y};
}
V2<float> smooth_start(const TCtx& t, uint32_t x0, uint32_t y0) {  // transport.wgsl:101-117
    F4 jit = rng_f4(0x85EBCA6Bu, t.p->epoch, S_MOVE + MAX_MOVERS_PER_CH + 2u, t.p->seed);
    uint32_t x = x0 + rb_f2u(jit.x * (float)SMOOTH_CELL);
    uint32_t y = y0 + rb_f2u(jit.y * (float)SMOOTH_CELL);
    uint32_t gx = x / SMOOTH_CELL, gy = y / SMOOTH_CELL;
    float tx = rb_smoothstep01(((float)(x % SMOOTH_CELL) + 0.5f) / (float)SMOOTH_CELL);
    float ty = rb_smoothstep01(((float)(y % SMOOTH_CELL) + 0.5f) / (float)SMOOTH_CELL);
    V2<float> n = vmix(vmix(lattice_rand(t, gx, gy), lattice_rand(t, gx + 1u, gy), tx),
                       vmix(lattice_rand(t, gx, gy + 1u), lattice_rand(t, gx + 1u, gy + 1u), tx), ty);
    F4 g = rng_f4(0x9E3779B9u, t.p->epoch, S_MOVE + MAX_MOVERS_PER_CH, t.p->seed);
    float vx = rb_fract(n.x + g.x), vy = rb_fract(n.y + g.y);
    return {1.0f - rb_abs(2.0f * vx - 1.0f), 1.0f - rb_abs(2.0f * vy - 1.0f)};
}

void transport_scatter_cell(const TCtx& t, const float* vel, const float* temp, const uint32_t* agg_act,
                            const uint32_t* agg_nb, std::atomic<uint32_t>* next, uint32_t x, uint32_t y) {
    // transport.wgsl:119-406
    const uint32_t G = t.G, F = t.F;
    const RiboSimParams& P = *t.p;
    const OracleWorld& W = *t.w;
    const uint32_t idx = y * G + x;
    const uint32_t src_total = cell_total(t.chem, idx);
    if (src_total == 0u) return;
    const float cell_w = (float)W.wupc;
    const uint32_t g_src = t.gamma[idx];
    Ctx c{&W, G, F, &P};
    V2<float> v = {0.0f, 0.0f};
    if (P.fluid_enabled != 0u && g_src < GAMMA_SOLID_THRESHOLD) {
        V2<float> cw = {((float)x + 0.5f) * cell_w, ((float)y + 0.5f) * cell_w};
        V2<float> v0 = fluid_velocity_at_world(c, vel, W.sim_size, cw);
        float world_per_fluid = W.sim_size / (float)F;
        float k = rb_max(P.dt, 0.0f);
        V2<float> q = {cw.x + v0.x * world_per_fluid * k * 0.5f, cw.y + v0.y * world_per_fluid * k * 0.5f};
        v = fluid_velocity_at_world(c, vel, W.sim_size, q);
    }
    if (v.x != 0.0f || v.y != 0.0f) {  // :153-162
        bool rock_r = x + 1u < G && t.gamma[idx + 1u] >= GAMMA_SOLID_THRESHOLD;
        bool rock_l = x > 0u && t.gamma[idx - 1u] >= GAMMA_SOLID_THRESHOLD;
        bool rock_u = y + 1u < G && t.gamma[idx + G] >= GAMMA_SOLID_THRESHOLD;
        bool rock_d = y > 0u && t.gamma[idx - G] >= GAMMA_SOLID_THRESHOLD;
        if ((v.x > 0.0f && rock_r) || (v.x < 0.0f && rock_l)) v.x = 0.0f;
        else if ((v.x > 0.0f && rock_l) || (v.x < 0.0f && rock_r)) v.x *= 0.5f;
        if ((v.y > 0.0f && rock_u) || (v.y < 0.0f && rock_d)) v.y = 0.0f;
        else if ((v.y > 0.0f && rock_d) || (v.y < 0.0f && rock_u)) v.y *= 0.5f;
    }
    const float tn = (float)std::max(P.transport_every, 1u);
    const float ratio = (float)G / (float)F;
    const float dtp = rb_max(P.dt, 0.0f);
    const V2<float> disp = {v.x * ratio * dtp * tn, v.y * ratio * dtp * tn};
    const float agitation = rb_clamp(rb_len2(v.x, v.y) / DIFF_AGITATION_SPEED, 0.0f, 1.0f);
    float p_diff = rb_clamp(DIFF_HOP_P * rb_mix(DIFF_HOP_FLOOR, 1.0f, agitation) * rb_max(P.diffusion, 0.0f) * tn, 0.0f, 0.5f);
    if (g_src > 0u) p_diff = rb_min(DIFF_HOP_P * BURIED_DIFF_FACTOR / (1.0f + GAMMA_POROSITY_K * (float)g_src) * tn, 0.5f);
    const uint32_t cap_self = capacity(t.gamma, idx);
    if (src_total > cap_self) p_diff = rb_min(CHEM_SQUEEZE_P * tn, 0.5f);
    float p_settle = 0.0f;
    if (g_src == 0u) p_settle = rb_min(MONOMER_SETTLE_P * rb_max(P.settle, 0.0f) * tn, 0.5f);
    const int32_t base_x = rb_addi(rb_f2i(rb_floor(disp.x)), -1);
    const int32_t base_y = rb_addi(rb_f2i(rb_floor(disp.y)), -1);

    float press[4] = {1.0f, 1.0f, 1.0f, 1.0f};
    float max_drop = 0.0f;
    if (P.monomer_pressure > 0.0f) {  // :191-206
        float f_self = (float)src_total / (float)std::max(cap_self, 1u);
        int32_t nb[4] = {1, -1, (int32_t)G, -(int32_t)G};
        bool inside[4] = {x + 1u < G, x > 0u, y + 1u < G, y > 0u};
        for (int d = 0; d < 4; ++d) {
            if (!inside[d]) continue;
            uint32_t n = (uint32_t)((int32_t)idx + nb[d]);
            uint32_t cap_n = capacity(t.gamma, n);
            if (cap_n == 0u) continue;
            float f_n = (float)cell_total(t.chem, n) / (float)cap_n;
            press[d] = rb_max(0.05f, 1.0f + P.monomer_pressure * (f_self - f_n));
            max_drop = rb_max(max_drop, f_self - f_n);
        }
    }
    p_diff = rb_clamp(p_diff + PRESSURE_HOP_P * P.monomer_pressure * max_drop * tn, 0.0f, 0.5f);
    const float light_t = uv_light_at_cell(W, x, y);

    float p_bound = 0.0f;
    float accept[4] = {1.0f, 1.0f, 1.0f, 1.0f};
    float accept_spent[4] = {1.0f, 1.0f, 1.0f, 1.0f};
    if (P.aggregation > 0.0f && src_total <= cap_self) {  // :228-260
        float e[5] = {(float)(agg_nb[idx] + agg_act[idx]), 0.0f, 0.0f, 0.0f, 0.0f};
        if (x + 1u < G) e[1] = (float)(agg_nb[idx + 1u] + agg_act[idx + 1u]);
        if (x > 0u) e[2] = (float)(agg_nb[idx - 1u] + agg_act[idx - 1u]);
        if (y + 1u < G) e[3] = (float)(agg_nb[idx + G] + agg_act[idx + G]);
        if (y > 0u) e[4] = (float)(agg_nb[idx - G] + agg_act[idx - G]);
        uint32_t fx = std::min((x * F) / G, F - 1u), fy = std::min((y * F) / G, F - 1u);
        float t_rel = 1.0f + rb_max(temp[fy * F + fx], 0.0f) / TEMP_ACT_THRESHOLD;
        float room = rb_clamp(5.0f * (1.0f - (float)src_total / (float)std::max(cap_self, 1u)), 0.0f, 1.0f);
        float k_e = P.aggregation * room / t_rel;
        for (int d = 0; d < 4; ++d) {
            float drop = e[0] - e[d + 1];
            if (drop > 0.0f) accept[d] = rb_expf(-k_e * drop);
            float rise = e[d + 1] - e[0];
            if (rise > 0.0f) accept_spent[d] = rb_expf(-k_e * rise);
        }
        p_bound = 1.0f - rb_expf(-k_e * rb_min(rb_max(e[0] - 1.0f, 0.0f), 48.0f));
    }

    for (uint32_t ch = 0; ch < 4; ++ch) {  // :262-405
        const uint32_t slot = idx * 4u + ch;
        const uint32_t v_ch = t.chem[slot];
        uint32_t act_n = v_ch & CHEM_STATE_MASK;
        uint32_t spent_n = v_ch >> 16;
        if (act_n + spent_n == 0u) continue;
        if (spent_n > 0u) {
            float sens = 1.0f + CHEM_SENSITIZE * (float)std::min(act_n, 8u);
            float exp_act = (float)std::min(spent_n, 8u) * tn *
                            (LIGHT_ACT_P * rb_max(P.uv_strength, 0.0f) * rb_max(P.direct_photoactivation, 0.0f) * light_t * sens +
                             rb_max(P.reactivation_rate, 0.0f));
            uint32_t na = rb_f2u(rb_floor(exp_act));
            float frac = exp_act - (float)na;
            if (frac > 0.0f && rng_f4(slot, P.epoch, S_PHOTO, P.seed).x < frac) na += 1u;
            na = std::min(na, std::min(spent_n, 8u));
            act_n += na;
            spent_n -= na;
        }
        if (act_n > 0u) {
            float shield = 1.0f / (1.0f + CHEM_SHIELD * (float)(std::min(act_n, 8u) - 1u));
            float exp_dec = (float)std::min(act_n, 8u) * rb_max(P.activation_decay, 0.0f) * shield * tn;
            if (exp_dec > 0.0f && rng_f4(slot, P.epoch, S_DECAY, P.seed).x < exp_dec) {
                act_n -= 1u;
                spent_n += 1u;
            }
        }
        float coh[4] = {1.0f, 1.0f, 1.0f, 1.0f};
        if (act_n > 0u && P.cohesion > 0.0f) {
            auto ac = [&](uint32_t cell) { return std::min(t.chem[cell * 4u + ch] & CHEM_STATE_MASK, 8u); };
            if (x + 1u < G) coh[0] = 1.0f + P.cohesion * (float)ac(idx + 1u);
            if (x > 0u) coh[1] = 1.0f + P.cohesion * (float)ac(idx - 1u);
            if (y + 1u < G) coh[2] = 1.0f + P.cohesion * (float)ac(idx + G);
            if (y > 0u) coh[3] = 1.0f + P.cohesion * (float)ac(idx - G);
        }
        uint32_t bins[16] = {0};
        uint32_t stay = 0;
        const uint32_t n = act_n + spent_n;
        const uint32_t movers = std::min(n, MAX_MOVERS_PER_CH);
        const uint32_t act_stay = act_n - std::min(act_n, movers);
        stay += act_stay + (n - movers - act_stay) * CHEM_SPENT_ONE;
        V2<float> bound_start = {0.0f, 0.0f};
        bool have_bound_start = false;
        const float dlen = rb_len2(disp.x, disp.y);
        for (uint32_t k = 0; k < movers; ++k) {
            const uint32_t unit = k < act_n ? 1u : CHEM_SPENT_ONE;
            F4 rf = rng_f4(slot, P.epoch, S_MOVE + k, P.seed);
            const bool bound = k < act_n && p_bound > 0.0f && rb_fract(rf.z * 7.31f + rf.w * 3.17f) < p_bound;
            if (bound && !have_bound_start) { bound_start = smooth_start(t, x, y); have_bound_start = true; }
            V2<float> p = bound ? bound_start : V2<float>{rf.x, rf.y};
            p.x += disp.x;
            p.y += disp.y;
            if (g_src > 0u && (disp.x != 0.0f || disp.y != 0.0f)) {
                F4 rj = rng_f4(slot, P.epoch, S_DISPERSE + k, P.seed);
                float jx = (rj.x * 2.0f - 1.0f) * dlen * RUBBLE_DISPERSION;
                float jy = (rj.y * 2.0f - 1.0f) * dlen * RUBBLE_DISPERSION;
                p.x += rb_clamp(jx, -0.9f, 0.9f);
                p.y += rb_clamp(jy, -0.9f, 0.9f);
            }
            const float r = rf.z;
            if (r < p_diff) {
                float wgt[4] = {press[0], press[1], press[2], press[3]};
                if (k < act_n) for (int cd = 0; cd < 4; ++cd) wgt[cd] *= coh[cd];
                float u = rf.w * (wgt[0] + wgt[1] + wgt[2] + wgt[3]);
                uint32_t d = 3u;
                for (uint32_t cd = 0; cd < 4; ++cd) {
                    if (u < wgt[cd]) { d = cd; break; }
                    u -= wgt[cd];
                }
                float acc_d = k < act_n ? accept[d] : accept_spent[d];
                bool stay_put = rb_fract(rf.w * 13.73f + rf.x * 5.31f) >= acc_d;
                if (!stay_put) {
                    if (d == 0u) p.x += 1.0f;
                    else if (d == 1u) p.x -= 1.0f;
                    else if (d == 2u) p.y += 1.0f;
                    else p.y -= 1.0f;
                }
            } else if (r < p_diff + p_settle) {
                p.y -= 1.0f;
            }
            int32_t ox = rb_addi(rb_f2i(rb_floor(p.x)), -base_x);
            int32_t oy = rb_addi(rb_f2i(rb_floor(p.y)), -base_y);
            uint32_t bi = (uint32_t)rb_clampi(oy, 0, 3) * 4u + (uint32_t)rb_clampi(ox, 0, 3);
            bins[bi] += unit;
        }
        for (uint32_t b = 0; b < 16; ++b) {
            uint32_t cnt = bins[b];
            if (cnt == 0u) continue;
            int32_t tx = rb_clampi(rb_addi(rb_addi((int32_t)x, base_x), (int32_t)(b & 3u)), 0, (int32_t)G - 1);
            int32_t ty = rb_clampi(rb_addi(rb_addi((int32_t)y, base_y), (int32_t)(b >> 2)), 0, (int32_t)G - 1);
            uint32_t t_idx = (uint32_t)ty * G + (uint32_t)tx;
            bool blocked = false;
            if (t_idx != idx) {
                uint32_t g_tgt = t.gamma[t_idx];
                if (g_tgt >= GAMMA_SOLID_THRESHOLD) blocked = g_src < GAMMA_SOLID_THRESHOLD;
                else if (cell_total(t.chem, t_idx) > capacity(t.gamma, t_idx)) blocked = true;
            }
            if (blocked) stay += cnt;
            else next[t_idx * 4u + ch].fetch_add(cnt, std::memory_order_relaxed);
        }
        if (stay > 0u) next[slot].fetch_add(stay, std::memory_order_relaxed);
    }
}

template <class R> const std::vector<R>& fluid_temp_in(const Fluid<R>& fl) { return fl.temp_in; }

template <class R> void thermal_activation(OracleWorld& w, const Fluid<R>& fl, const RiboSimParams& P) {  // transport.wgsl:419-438
    const uint32_t G = w.G, F = w.F;
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)G * G; ++ii) {
        uint32_t x = (uint32_t)(ii % G), y = (uint32_t)(ii / G);
        uint32_t fx = std::min((x * F) / G, F - 1u), fy = std::min((y * F) / G, F - 1u);
        float t_loc = (float)fl.temp_in[fy * F + fx];
        if (t_loc <= TEMP_ACT_THRESHOLD) continue;
        float p_act = rb_clamp(FUMAROLE_ACT_P * rb_max(P.thermal_activation, 0.0f) * (t_loc - TEMP_ACT_THRESHOLD) / TEMP_ACT_THRESHOLD,
                               0.0f, 0.9f);
        uint32_t idx = y * G + x;
        for (uint32_t ch = 0; ch < 4; ++ch) {
            uint32_t slot = idx * 4u + ch;
            if ((w.chem[slot] >> 16) > 0u && rng_f4(slot, P.epoch, S_THERMAL, P.seed).x < p_act) {
                // chem_activate_one (chem.wgsl:90-110): spent -> activated in place.
                if ((w.chem[slot] >> 16) != 0u) w.chem[slot] += 0xFFFF0001u;
            }
        }
    }
}

void agg_count(OracleWorld& w) {  // transport.wgsl:61-68
    const uint32_t G = w.G;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)G * G; ++i) {
        uint32_t cnt = 0;
        for (uint32_t ch = 0; ch < 4; ++ch) cnt += w.chem[i * 4 + ch] & CHEM_STATE_MASK;
        w.agg_act[i] = cnt;
    }
}
void agg_neighbours(OracleWorld& w) {  // transport.wgsl:71-87
    const int32_t G = (int32_t)w.G;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)G * G; ++i) {
        int32_t x = (int32_t)(i % G), y = (int32_t)(i / G);
        uint32_t sum = 0;
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                if (dx == 0 && dy == 0) continue;
                int32_t nx = x + dx, ny = y + dy;
                if (nx < 0 || ny < 0 || nx >= G || ny >= G) continue;
                sum += w.agg_act[(uint32_t)ny * (uint32_t)G + (uint32_t)nx];
            }
        w.agg_nb[i] = sum;
    }
}

template <class R> void transport(OracleWorld& w, const Fluid<R>& fl, const RiboSimParams& P) {
    const uint32_t G = w.G, F = w.F;
    // binary32 views of the velocity and temperature the transport reads.
    std::vector<float> vel(fl.vel_a.size()), temp(fl.temp_in.size());
    for (size_t i = 0; i < vel.size(); ++i) vel[i] = (float)fl.vel_a[i];
    for (size_t i = 0; i < temp.size(); ++i) temp[i] = (float)fl.temp_in[i];
    if (P.aggregation > 0.0f) {  // src/world/mod.rs:1595-1598
        agg_count(w);
        agg_neighbours(w);
    }
    static_assert(sizeof(std::atomic<uint32_t>) == sizeof(uint32_t), "atomic layout");
    std::atomic<uint32_t>* next = reinterpret_cast<std::atomic<uint32_t>*>(w.chem_next.data());
    TCtx t{&w, &P, w.chem.data(), w.gamma.data(), G, F};
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < (int64_t)G * G; ++i)
        transport_scatter_cell(t, vel.data(), temp.data(), w.agg_act.data(), w.agg_nb.data(), next,
                               (uint32_t)(i % G), (uint32_t)(i / G));
    // transport_commit (transport.wgsl:409-415)
    w.chem.swap(w.chem_next);
    std::fill(w.chem_next.begin(), w.chem_next.end(), 0u);
}

// Tolerance mode (an extension; the reference runs a fixed number of
// V-cycles): preconditioned conjugate gradients on the eps-anchored operator
// A p = div over fluid cells of level 0, started from the previous pressure
// moved along the change between the last two solves to the point of least
// residual (once two solves are done; the history restarts on new settings
// and on uploads of pressure, velocity or terrain), preconditioned by the
// symmetric V-cycle (MG::vcycle_sym), until
// ||div - A p||2 / ||div||2 <= mg_tol or mg_max_cycles iterations.
//
// The recurrences put all inner products of an iteration at one point,
// after the preconditioner (Chronopoulos and Gear, 1989): with u = M r and
// w = A u,
//   gamma = (r, u), delta = (w, u), e = (u, s), f = (p, w), rr = (r, r)
// for the current p and s, then
//   beta = gamma / gamma_prev (0 on the first application),
//   pi = delta + beta (e + f) + beta^2 pi_prev   (= (p', s') below),
//   alpha = gamma / pi,
//   p' = u + beta p,  s' = w + beta s,  x += alpha p',  r -= alpha s'.
// pi is the expansion of (p', s') itself rather than Chronopoulos and Gear's
// shorter form delta - beta gamma / alpha_prev, which holds only for an
// exactly linear preconditioner: with the bfloat16 preconditioner the short
// form stalls and this one converges like binary32.
//
// Precision.  Scalars and inner products are binary64; vectors binary32
// (binary64 in the FP64 oracle), with three exceptions that keep the
// recursively updated residual on the true one:
//   - the iterate is a double-float pair x = hi + lo (hi is the pressure as
//     stored; every update adds alpha p' with an exact two-sum and
//     renormalises, so hi stays the nearest stored value of x);
//   - w = A u and every residual b - A x are evaluated in binary64 (A x on
//     hi + lo) and rounded once; in binary32 the stencil's cancellation on
//     smooth vectors (|A v| ~ lambda |v| with lambda down to ~1e-5) would
//     otherwise put an error of ~3 eps32 / lambda on w;
//   - the stop is the true residual: once the recursive residual is within
//     NEAR x tol, each V-cycle also forms ||b - A x||^2 in binary64, and the
//     solve ends when that reaches tol.  If the recursive residual reaches
//     tol while the true one has not, the iteration restarts from the true
//     residual (p = s = 0, the next products applied as a first step): one
//     step of iterative refinement, with the refinement residual in binary64.
// precision = MIXED stores the preconditioner's unknowns and right-hand
// sides (level 0 included) in bfloat16.
constexpr double PCG_NEAR = 4.0;

// x = hi + lo += t: the exact two-sum of hi and t, the error added to lo,
// then renormalised so that |lo| <= ulp(hi) / 2.
template <class R> inline void df_add(R& hi, R& lo, R t) {
    const R s = hi + t;
    const R bv = s - hi;
    const R e = (hi - (s - bv)) + (t - bv);
    const R l = lo + e;
    const R h = s + l;
    lo = l - (h - s);
    hi = h;
}

template <class R> void pcg_solve(OracleWorld& w, Fluid<R>& fl, MG<R>& mg, const std::vector<Level>& lv) {
    const RiboWorldSettings& st = w.st;
    const size_t n2 = (size_t)lv[0].n * lv[0].n;
    mg.bf16 = st.precision == RIBO_PRECISION_MIXED;
    for (size_t l = 0; l + 1 < lv.size(); ++l)  // coarse masks: any fluid child (multigrid.wgsl:98)
        for (uint32_t gy = 0; gy < lv[l].n_c; ++gy)
            for (uint32_t gx = 0; gx < lv[l].n_c; ++gx) {
                uint32_t nf = 0;
                for (uint32_t c = 0; c < 4; ++c)
                    nf += w.mg_fluid[lv[l].off + (gy * 2u + (c >> 1)) * lv[l].n + gx * 2u + (c & 1u)] ? 1u : 0u;
                w.mg_fluid[lv[l].off_c + gy * lv[l].n_c + gx] = nf > 0 ? 1u : 0u;
            }
    mg.sym_levels = lv.size();
    CoarseInv ci;
    const Level& C = lv.back();
    mg.cinv = ci.build(C.n, w.mg_fluid.data() + C.off, C.eps) ? &ci : nullptr;
    std::vector<R> r(n2), u(n2), wv(n2), pd(n2, R(0)), sd(n2, R(0)), xlo(n2, R(0));
    std::vector<double> a64(n2);
    R* x = fl.p_a.data();
    const R* b = fl.div.data();
    const double tol = (double)st.mg_tol;
    auto fluid0 = [&](size_t i) { return w.mg_fluid[i] != 0u; };
    // r = b - A x in binary64, rounded once
    auto true_residual = [&]() {
        mg.apply64(x, xlo.data(), a64.data());
        for (size_t i = 0; i < n2; ++i) r[i] = fluid0(i) ? R((double)b[i] - a64[i]) : R(0);
    };
    auto true_rr = [&]() {
        mg.apply64(x, xlo.data(), a64.data());
        double t = 0.0;
        for (size_t i = 0; i < n2; ++i) {
            if (!fluid0(i)) continue;
            const double v = (double)b[i] - a64[i];
            t += v * v;
        }
        return t;
    };
    true_residual();
    double bb = 0.0, rr = 0.0, rq = 0.0, qq = 0.0, theta = 0.0;
    // The start: once the pressure and p_old are the last two solves'
    // results, move along their change d = p - p_old to the point of least
    // residual, x + theta d with theta = (r, A d) / (A d, A d) (theta = 1 is
    // the linear extrapolation in time); p_old takes the old pressure.  The
    // residual of the start follows from the sums:
    // |r - theta q|^2 = |r|^2 - (r, q)^2 / (q, q).
    std::vector<R> d, q;
    if (fl.hist >= 2) {
        d.resize(n2);
        q.resize(n2);
        for (size_t i = 0; i < n2; ++i) d[i] = x[i] - fl.p_old[i];
        mg.apply64(d.data(), nullptr, a64.data());
        for (size_t i = 0; i < n2; ++i) q[i] = fluid0(i) ? R(a64[i]) : R(0);
    }
    for (size_t i = 0; i < n2; ++i) {
        if (!fluid0(i)) continue;
        bb += (double)b[i] * (double)b[i];
        rr += (double)r[i] * (double)r[i];
        if (!q.empty()) {
            rq += (double)r[i] * (double)q[i];
            qq += (double)q[i] * (double)q[i];
        }
    }
    if (!q.empty()) {
        theta = qq > 0.0 ? rq / qq : 0.0;
        const R th = R(theta);
        for (size_t i = 0; i < n2; ++i) {
            const R a = x[i];
            if (th != R(0) && fluid0(i)) {
                df_add(x[i], xlo[i], th * (a - fl.p_old[i]));
                r[i] = r[i] - th * q[i];
            }
            fl.p_old[i] = a;
        }
    } else {
        for (size_t i = 0; i < n2; ++i) fl.p_old[i] = x[i];
    }
    double rel = bb > 0.0 ? sqrt(std::max(rr - theta * rq, 0.0) / bb) : 0.0;
    bool near = rel <= PCG_NEAR * tol;
    // u = M r, w = A u and the V-cycle's inner products (with the true
    // residual of the current x when near)
    double gamma = 0.0, delta = 0.0, e = 0.0, f = 0.0, tt = -1.0;
    auto precond = [&]() {
        for (size_t i = 0; i < n2; ++i) {
            fl.mg_rhs[i] = mg.bf16 ? R(round_bf16((float)r[i])) : r[i];
            fl.mg_p[i] = 0;
        }
        mg.vcycle_sym();
        for (size_t i = 0; i < n2; ++i) u[i] = fl.mg_p[i];
        mg.apply64(u.data(), nullptr, a64.data());
        for (size_t i = 0; i < n2; ++i) wv[i] = fluid0(i) ? R(a64[i]) : R(0);
        gamma = delta = e = f = rr = 0.0;
        for (size_t i = 0; i < n2; ++i) {
            if (!fluid0(i)) continue;
            gamma += (double)r[i] * (double)u[i];
            delta += (double)wv[i] * (double)u[i];
            e += (double)u[i] * (double)sd[i];
            f += (double)pd[i] * (double)wv[i];
            rr += (double)r[i] * (double)r[i];
        }
        tt = near ? true_rr() : -1.0;
    };
    uint32_t it = 0, restarts = 0;
    const uint32_t maxit = std::max(st.mg_max_cycles, 1u);
    if (!(rel <= tol)) {
        precond();
        bool fresh = true;
        double beta = 0.0, pi = 0.0, alpha = 0.0, g_prev = 0.0;
        for (;;) {
            // the scalars from the last V-cycle's inner products
            bool restart = false;
            if (fresh) {
                g_prev = gamma;
                pi = delta;
                alpha = delta != 0.0 ? gamma / delta : 0.0;
                beta = 0.0;
                fresh = false;
            } else {
                ++it;
                rel = bb > 0.0 ? sqrt(rr / bb) : 0.0;
                const bool have_true = tt >= 0.0;
                const double trel = have_true && bb > 0.0 ? sqrt(tt / bb) : (have_true ? 0.0 : -1.0);
                if (have_true && trel <= tol) break;
                if (it >= maxit) break;
                if (have_true && rel <= tol) {
                    restart = true;
                } else {
                    beta = g_prev != 0.0 ? gamma / g_prev : 0.0;
                    pi = delta + beta * (e + f) + beta * beta * pi;
                    alpha = pi != 0.0 ? gamma / pi : 0.0;
                    g_prev = gamma;
                }
                near = rel <= PCG_NEAR * tol;
            }
            if (restart) {
                // one refinement step: the iteration starts again from the
                // true residual of the current iterate
                true_residual();
                std::fill(pd.begin(), pd.end(), R(0));
                std::fill(sd.begin(), sd.end(), R(0));
                fresh = true;
                ++restarts;
            } else {
                const R a = R(alpha), be = R(beta);
                for (size_t i = 0; i < n2; ++i) {
                    if (!fluid0(i)) continue;
                    pd[i] = u[i] + be * pd[i];
                    sd[i] = wv[i] + be * sd[i];
                    df_add(x[i], xlo[i], a * pd[i]);
                    r[i] -= a * sd[i];
                }
            }
            precond();
        }
    }
    w.last_cycles = it;
    w.last_restarts = restarts;
    w.last_rel_res = bb > 0.0 ? sqrt(true_rr() / bb) : 0.0;
    w.last_rounded_rel = mg.rel_residual_of(x, b);
    fl.hist = std::min(fl.hist + 1, 2);
}

// ---- one fluid step (src/world/mod.rs:1527-1588) ---------------------------------
template <class R> void fluid_step(OracleWorld& w, Fluid<R>& fl, const RiboSimParams& P) {
    Ctx c{&w, w.G, w.F, &P};
    const RiboWorldSettings& st = w.st;
    build_solid_mask<R>(w, c);
    update_temperature<R>(w, fl, c);
    copy_temperature<R>(w, fl);
    buoyancy<R>(w, fl, c);
    gather_forces<R>(w, fl, c);
    std::fill(w.force_vectors.begin(), w.force_vectors.end(), 0u);  // clear_force_vectors
    add_forces<R>(w, fl, c, fl.vel_a.data(), fl.vel_b.data());       // a -> b
    std::fill(fl.forces.begin(), fl.forces.end(), R(0));            // clear_forces
    diffuse_velocity<R>(c, fl.vel_b.data(), fl.vel_a.data());       // b -> a
    R dt = fluid_dt<R>(c);
    R decay = sizeof(R) == 4 ? R(decay_factor_f32(P.fluid_decay, (float)dt)) : R(decay_factor_f64(P.fluid_decay, (double)dt));
    advect_velocity<R>(c, fl.vel_a.data(), fl.vel_b.data(), decay); // a -> b
    vorticity_confinement<R>(c, fl.vel_b.data(), fl.vel_a.data());  // b -> a
    compute_divergence<R>(c, fl.vel_a.data(), fl.div.data());       // reads a
    if (st.multigrid) {
        uint32_t total;
        std::vector<Level> lv = make_levels(w.F, st.mg_min_size, &total);
        MG<R> mg{&lv, fl.mg_p.data(), fl.mg_rhs.data(), w.mg_fluid.data()};
        const Level& L0 = lv[0];
        for (uint32_t i = 0; i < L0.n * L0.n; ++i) {  // mg_init (multigrid.wgsl:49-58)
            bool fluid = w.solid_mask[i] == 0u;
            w.mg_fluid[i] = fluid ? 1u : 0u;
            fl.mg_rhs[i] = fluid ? fl.div[i] : R(0);
            fl.mg_p[i] = fluid ? fl.p_a[i] : R(0);
        }
        uint32_t cycles = std::max(st.mg_cycles, 1u);
        if (st.mg_tol > 0.0f && st.mg_stationary) {
            // the reference V-cycle repeated until ||rhs - A p|| / ||rhs||
            // over level-0 fluid cells reaches mg_tol (the residual and the
            // norms in binary64)
            const uint32_t maxit = std::max(st.mg_max_cycles, 1u);
            const size_t n0 = (size_t)L0.n * L0.n;
            std::vector<double> ap(n0);
            auto rel_of = [&]() {
                mg.apply64(fl.mg_p.data(), nullptr, ap.data());
                double bb = 0.0, rr = 0.0;
                for (size_t i = 0; i < n0; ++i) {
                    if (!w.mg_fluid[i]) continue;
                    const double r = (double)fl.mg_rhs[i] - ap[i];
                    bb += (double)fl.mg_rhs[i] * (double)fl.mg_rhs[i];
                    rr += r * r;
                }
                return bb > 0.0 ? sqrt(rr / bb) : 0.0;
            };
            uint32_t it = 0;
            double rel = rel_of();
            while (rel > (double)st.mg_tol && it < maxit) {
                mg.vcycle();
                ++it;
                rel = rel_of();
            }
            w.last_cycles = it;
            w.last_rel_res = rel;
            w.last_rounded_rel = rel;
        } else if (st.mg_tol > 0.0f) {
            // the iterate starts from mg_init's pressure: zero at the walls
            // (a cell the moving terrain has just made solid drops its value)
            for (uint32_t i = 0; i < L0.n * L0.n; ++i) fl.p_a[i] = fl.mg_p[i];
            pcg_solve<R>(w, fl, mg, lv);
            for (uint32_t i = 0; i < L0.n * L0.n; ++i) fl.mg_p[i] = fl.p_a[i];
        } else {
            for (uint32_t k = 0; k < cycles; ++k) mg.vcycle();
        }
        for (uint32_t i = 0; i < L0.n * L0.n; ++i) fl.p_a[i] = fl.mg_p[i];  // mg_finish
    } else {
        uint32_t iters = (std::min(std::max(st.jacobi_iters, 2u), 128u) + 1u) & ~1u;  // src/world/mod.rs:1442
        for (uint32_t k = 0; k < iters; ++k) {
            if (k % 2 == 0) jacobi_pressure<R>(c, fl.p_a.data(), fl.p_b.data(), fl.div.data());
            else jacobi_pressure<R>(c, fl.p_b.data(), fl.p_a.data(), fl.div.data());
        }
    }
    subtract_gradient<R>(c, fl.vel_a.data(), fl.vel_b.data(), fl.p_a.data());  // a -> b
    net_flux<R>(c, fl.vel_b.data(), fl.net_flux.data());                      // reads b
    enforce_boundaries<R>(c, fl.vel_b.data(), fl.vel_a.data(), fl.net_flux.data());  // b -> a
    thermal_activation<R>(w, fl, P);
}

}  // namespace
extern "C" {
static void compute_slope(OracleWorld* w);
}
namespace {

// ---- terrain grain dynamics (terrain.wgsl:38-102,159-430) -------------------
// The reference runs every cell's decision at once, each reading its
// neighbours' grain counts while other cells move grains; no order of those
// reads is fixed.  This restatement fixes one admissible order, the same as
// the HIP kernels': every decision of a pass reads the counts at the start
// of the pass; then every grain moves (a cell only ever gives its own grain
// away, so each take succeeds); then the monomers that no longer fit leave
// the cells that received grains, cell colour (x mod 3, y mod 3) after cell
// colour, and within a cell the grains from its eight neighbours in row
// order.  Cells of one colour are 3 apart, so their evictions touch
// disjoint cells and their order does not matter.
constexpr uint32_t SANDPILE_DIFF = 3u, GAMMA_ROCK_FULL = 6u;
constexpr uint32_t S_RELAX = 16u, S_SAND = 32u;
constexpr float GAMMA_RELAX_P = 0.12f, GAMMA_STRAY_WALK_P = 0.05f, GAMMA_SEDIMENT_FACTOR = 0.5f;
constexpr float GAMMA_SETTLE_SPEED = 0.5f, VENT_BLOW = 8.0f, GAMMA_SHED_P = 0.03f, GRAIN_ADV_CAP_T = 0.9f;

float hop_p(const OracleWorld& w, const RiboSimParams& P, float vx, float vy) {  // parcel_hop_p, transport.wgsl:54-58
    const float env_per_fluid = (float)w.G / (float)w.F;
    const float l1 = fabsf(vx) + fabsf(vy);
    return rb_clamp(l1 * env_per_fluid * rb_max(P.dt, 1e-3f), 0.0f, GRAIN_ADV_CAP_T);
}

// evict_target (terrain.wgsl:49-66)
uint32_t evict_target(const OracleWorld& w, uint32_t src, uint32_t dst) {
    const uint32_t* ch = w.chem.data();
    const uint32_t* g = w.gamma.data();
    if (cell_total(ch, src) < capacity(g, src)) return src;
    const int x = (int)(dst % w.G), y = (int)(dst / w.G);
    const int nb[4][2] = {{x + 1, y}, {x - 1, y}, {x, y + 1}, {x, y - 1}};
    uint32_t not_rock = g[src] < GAMMA_SOLID_THRESHOLD ? src : 0xFFFFFFFFu;
    for (int i = 0; i < 4; ++i) {
        if (nb[i][0] < 0 || nb[i][1] < 0 || nb[i][0] >= (int)w.G || nb[i][1] >= (int)w.G) continue;
        const uint32_t n = (uint32_t)nb[i][1] * w.G + (uint32_t)nb[i][0];
        if (cell_total(ch, n) < capacity(g, n)) return n;
        if (not_rock == 0xFFFFFFFFu && g[n] < GAMMA_SOLID_THRESHOLD) not_rock = n;
    }
    return not_rock == 0xFFFFFFFFu ? src : not_rock;
}

// The eviction loop of gamma_move_one (terrain.wgsl:76-96), after the grain
// has arrived in dst.
void evict(OracleWorld& w, uint32_t src, uint32_t dst) {
    const uint32_t cap = capacity(w.gamma.data(), dst);
    for (uint32_t guard = 0; guard < 64u; ++guard) {
        if (cell_total(w.chem.data(), dst) <= cap) break;
        const uint32_t to = evict_target(w, src, dst);
        bool moved = false;
        for (uint32_t c = 0; c < 4u && !moved; ++c) {
            uint32_t& slot = w.chem[(size_t)dst * 4 + c];
            if ((slot >> 16) != 0u) {  // chem_take_state_one(slot, spent)
                slot -= CHEM_SPENT_ONE;
                w.chem[(size_t)to * 4 + c] += CHEM_SPENT_ONE;
                moved = true;
            } else if ((slot & CHEM_STATE_MASK) != 0u) {
                slot -= 1u;
                w.chem[(size_t)to * 4 + c] += 1u;
                moved = true;
            }
        }
        if (!moved) break;
    }
}

// compaction_stable (terrain.wgsl:139-157) on the pass's counts g.
bool compaction_stable(const OracleWorld& w, const uint32_t* g, uint32_t src, uint32_t dst, uint32_t n, uint32_t c) {
    const int d4[4][2] = {{1, 0}, {-1, 0}, {0, 1}, {0, -1}};
    bool ok = true;
    const int G = (int)w.G;
    for (int k = 0; k < 4; ++k) {
        const int qdx = (int)(dst % w.G) + d4[k][0], qdy = (int)(dst / w.G) + d4[k][1];
        if (qdx >= 0 && qdy >= 0 && qdx < G && qdy < G) {
            const uint32_t qi = (uint32_t)qdy * w.G + (uint32_t)qdx;
            const uint32_t low = qi == src ? n - 1u : g[qi];
            if (c + 1u >= low + SANDPILE_DIFF) ok = false;
        }
        const int qsx = (int)(src % w.G) + d4[k][0], qsy = (int)(src / w.G) + d4[k][1];
        if (qsx >= 0 && qsy >= 0 && qsx < G && qsy < G) {
            const uint32_t qi = (uint32_t)qsy * w.G + (uint32_t)qsx;
            const uint32_t high = qi == dst ? c + 1u : g[qi];
            if (high >= n - 1u + SANDPILE_DIFF) ok = false;
        }
    }
    return ok;
}

// The decision of relax_gamma_pass (terrain.wgsl:159-406) for one cell:
// where its grain goes (idx itself for none).
uint32_t relax_decide(const OracleWorld& w, const RiboSimParams& P, const uint32_t* g, const float* vel, uint32_t idx,
                      uint32_t phase) {
    const uint32_t G = w.G, F = w.F, x = idx % G, y = idx / G;
    const uint32_t n = g[idx];
    if (n >= 1u && n < SANDPILE_DIFF) {
        uint32_t bonds = 0, best_idx = idx, best_c = 0xFFFFu, em_idx = idx;
        int best_score = -1, em_score = -1;
        bool em_found = false;
        uint32_t rr[4] = {idx, P.epoch, S_RELAX + phase, P.seed};
        pcg4d(rr);
        uint32_t tie_h = rr[0];
        uint32_t comp_idx = idx, comp_c = 0;
        const float comp_roll = (float)(hash_u32(rr[0] ^ 0x5bd1e995u) >> 8) * (1.0f / 16777216.0f);
        const bool try_comp = comp_roll < P.sediment_compaction;
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                if (dx == 0 && dy == 0) continue;
                const int nx = (int)x + dx, ny = (int)y + dy;
                if (nx < 0 || ny < 0 || nx >= (int)G || ny >= (int)G) continue;
                const uint32_t ni = (uint32_t)ny * G + (uint32_t)nx;
                const uint32_t c = g[ni];
                if (c >= 1u) bonds += 1u;
                if (try_comp && c >= 1u && c < GAMMA_ROCK_FULL &&
                    (c > comp_c || (c == comp_c && (hash_u32(tie_h ^ (ni * 31u)) & 1u) == 1u)) &&
                    compaction_stable(w, g, idx, ni, n, c)) {
                    comp_c = c;
                    comp_idx = ni;
                }
                if (c <= n) {
                    int score = 0;
                    for (int sy = -1; sy <= 1; ++sy)
                        for (int sx = -1; sx <= 1; ++sx) {
                            if (sx == 0 && sy == 0) continue;
                            const int qx = nx + sx, qy = ny + sy;
                            if (qx < 0 || qy < 0 || qx >= (int)G || qy >= (int)G) continue;
                            const uint32_t qi = (uint32_t)qy * G + (uint32_t)qx;
                            if (qi == idx) continue;
                            if (g[qi] >= 1u) score += 1;
                        }
                    tie_h = hash_u32(tie_h ^ ni);
                    if (score > best_score || (score == best_score && c < best_c) ||
                        (score == best_score && c == best_c && (tie_h & 1u) == 1u)) {
                        best_score = score;
                        best_c = c;
                        best_idx = ni;
                    }
                    if (c == 0u && (score > em_score || (score == em_score && (tie_h & 2u) == 2u))) {
                        em_score = score;
                        em_idx = ni;
                        em_found = true;
                    }
                }
            }
        const float mob = ldexpf(1.0f, -2 * (int)bonds);  // pow(0.25, bonds), exact
        if (try_comp && bonds >= 3u && comp_idx != idx &&
            comp_roll < P.sediment_compaction * (float)(bonds - 2u) / 6.0f)
            return comp_idx;
        float vsx = 0.0f, vsy = 0.0f;
        if (P.fluid_enabled != 0u) {
            const int nb4[4][2] = {{1, 0}, {-1, 0}, {0, 1}, {0, -1}};
            for (int k = 0; k < 4; ++k) {
                const int qx = (int)x + nb4[k][0], qy = (int)y + nb4[k][1];
                if (qx < 0 || qy < 0 || qx >= (int)G || qy >= (int)G) continue;
                if (g[(uint32_t)qy * G + (uint32_t)qx] != 0u) continue;
                const uint32_t fxi = std::min(((uint32_t)qx * F) / G, F - 1u), fyi = std::min(((uint32_t)qy * F) / G, F - 1u);
                float vx = vel[2 * (fyi * F + fxi)], vy = vel[2 * (fyi * F + fxi) + 1];
                if (is_bad(vx) || is_bad(vy)) vx = vy = 0.0f;  // sanitize_vec2
                if (vx * vx + vy * vy > vsx * vsx + vsy * vsy) { vsx = vx; vsy = vy; }
            }
        }
        const bool unsupported = y > 0u && g[idx - G] == 0u;
        float mvx = 0.0f, mvy = 0.0f, p_sed = 0.0f;
        if (unsupported && bonds == 0u) {
            mvx = vsx; mvy = vsy;
            p_sed = hop_p(w, P, mvx, mvy);
        } else {
            const float speed = rb_len2(vsx, vsy);
            const float crit = rb_max(P.sediment_threshold, 0.0f) * (1.0f + rb_max(P.sediment_cohesion, 0.0f) * (float)bonds);
            const float excess = rb_max(speed - crit, 0.0f);
            if (speed > 1e-6f) {
                const float k = excess / rb_max(speed, 1e-6f);
                mvx = vsx * k; mvy = vsy * k;
            }
            p_sed = hop_p(w, P, mvx, mvy) * GAMMA_SEDIMENT_FACTOR * rb_max(P.sediment_transport, 0.0f);
        }
        const uint32_t hf = std::min((y * F) / G, F - 1u) * F + std::min((x * F) / G, F - 1u);
        const float blow = VENT_BLOW * rb_max(w.heat_src[2 * hf], w.heat_src[2 * hf + 1]);
        if (blow > 0.0f) {
            mvx = ((rr[1] & 1u) == 1u ? 0.6f : -0.6f) * blow;
            mvy = 1.0f * blow;
            p_sed = rb_max(p_sed, hop_p(w, P, mvx, mvy));
        }
        if (bonds >= 4u && p_sed <= 0.0f) return idx;
        const uint32_t hw = rr[1];
        const float rw = (float)(rr[2] >> 8) * (1.0f / 16777216.0f);
        const float rw2 = (float)(rr[3] >> 8) * (1.0f / 16777216.0f);
        uint32_t dest = idx;
        if (n >= 2u && em_found && rw2 < GAMMA_SHED_P * rb_clamp(P.sediment_settle, 0.0f, 1.0f)) {
            dest = em_idx;
        } else if (rw < p_sed) {
            int ddx = 0, ddy = 0;
            const float ax = fabsf(mvx), ay = fabsf(mvy);
            if ((float)(hw >> 8) * (1.0f / 16777216.0f) < ax / rb_max(ax + ay, 1e-6f)) ddx = mvx > 0.0f ? 1 : -1;
            else ddy = mvy > 0.0f ? 1 : -1;
            const int wx = (int)x + ddx, wy = (int)y + ddy;
            if (wx >= 0 && wy >= 0 && wx < (int)G && wy < (int)G) {
                const uint32_t wi = (uint32_t)wy * G + (uint32_t)wx;
                if (g[wi] <= n) dest = wi;
            }
        } else if (rw < p_sed + GAMMA_STRAY_WALK_P * mob) {
            if (best_score > (int)bonds) {
                dest = best_idx;
            } else if (bonds == 0u) {
                static const int dir[8][2] = {{1, 0}, {-1, 0}, {0, 1}, {0, -1}, {1, 1}, {-1, 1}, {1, -1}, {-1, -1}};
                const uint32_t k = (hw >> 8) % 8u;
                const int wx = (int)x + dir[k][0], wy = (int)y + dir[k][1];
                if (wx >= 0 && wy >= 0 && wx < (int)G && wy < (int)G) {
                    const uint32_t wi = (uint32_t)wy * G + (uint32_t)wx;
                    if (g[wi] <= n) dest = wi;
                }
            }
        }
        return dest;
    }
    if (n < SANDPILE_DIFF) return idx;
    uint32_t best_idx = idx, best_count = n;
    if (x > 0u && g[idx - 1u] < best_count) { best_count = g[idx - 1u]; best_idx = idx - 1u; }
    if (x + 1u < G && g[idx + 1u] < best_count) { best_count = g[idx + 1u]; best_idx = idx + 1u; }
    if (y > 0u && g[idx - G] < best_count) { best_count = g[idx - G]; best_idx = idx - G; }
    if (y + 1u < G && g[idx + G] < best_count) { best_count = g[idx + G]; best_idx = idx + G; }
    if (best_idx == idx || n - best_count < SANDPILE_DIFF) return idx;
    const float p = GAMMA_RELAX_P * rb_max(rb_clamp(P.sediment_settle, 0.0f, 1.0f), rb_clamp(P.rock_crumble, 0.0f, 1.0f));
    if (rng_f4(idx, P.epoch, S_SAND + phase, P.seed).x < p) return best_idx;
    return idx;
}

// One relax pass (relax_gamma_a / relax_gamma_b) in the order described above.
void relax_pass(OracleWorld& w, const RiboSimParams& P, const float* vel, uint32_t phase) {
    const uint32_t G = w.G;
    const std::vector<uint32_t> g0 = w.gamma;
    std::vector<uint32_t> dest((size_t)G * G);
    for (uint32_t i = 0; i < G * G; ++i) dest[i] = relax_decide(w, P, g0.data(), vel, i, phase);
    for (uint32_t i = 0; i < G * G; ++i)
        if (dest[i] != i) { w.gamma[i] -= 1u; w.gamma[dest[i]] += 1u; }
    for (uint32_t color = 0; color < 9u; ++color)
        for (uint32_t y = color / 3u; y < G; y += 3u)
            for (uint32_t x = color % 3u; x < G; x += 3u) {
                const uint32_t d = y * G + x;
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        if (dx == 0 && dy == 0) continue;
                        const int sx = (int)x + dx, sy = (int)y + dy;
                        if (sx < 0 || sy < 0 || sx >= (int)G || sy >= (int)G) continue;
                        const uint32_t src = (uint32_t)sy * G + (uint32_t)sx;
                        if (dest[src] == d) evict(w, src, d);
                    }
            }
}

// grain_fall (terrain.wgsl:405-437): one column at a time, columns of
// colour x mod 3 = 0, 1, 2 in turn (their evictions touch columns x - 1 .. x + 1).
void grain_fall(OracleWorld& w, const RiboSimParams& P) {
    const uint32_t G = w.G;
    const float w_s = GAMMA_SETTLE_SPEED * rb_max(P.sediment_settle, 0.0f);
    const float rate = hop_p(w, P, 0.0f, w_s);
    if (rate <= 0.0f) return;
    const uint32_t rf = rb_f2u(rate * 65536.0f);
    if (((P.epoch * rf) >> 16) == (((P.epoch + 1u) * rf) >> 16)) return;
    for (uint32_t color = 0; color < 3u; ++color)
        for (uint32_t x = color; x < G; x += 3u) {
            bool supported = true;
            for (uint32_t y = 0; y < G; ++y) {
                const uint32_t idx = y * G + x;
                const uint32_t n = w.gamma[idx];
                if (n == 0u) { supported = false; continue; }
                if (n >= GAMMA_SOLID_THRESHOLD) { supported = true; continue; }
                if (supported) continue;
                for (uint32_t k = 0; k < n; ++k) {
                    w.gamma[idx] -= 1u;
                    w.gamma[idx - G] += 1u;
                    evict(w, idx, idx - G);
                }
            }
        }
}

// The terrain stage of a step (src/world/mod.rs:1495-1503).
void terrain_step(OracleWorld& w, const float* vel, const RiboSimParams& P) {
    relax_pass(w, P, vel, 0u);
    relax_pass(w, P, vel, 1u);
    if (P.sediment_settle > 0.0f) grain_fall(w, P);
    compute_slope(&w);
}

template <class R> void world_step(OracleWorld& w, Fluid<R>& fl, const RiboSimParams& P, uint32_t substep) {
    if (w.st.terrain_enabled) {
        if (sizeof(R) == 4) {
            terrain_step(w, (const float*)fl.vel_a.data(), P);
        } else {
            std::vector<float> v(fl.vel_a.begin(), fl.vel_a.end());
            terrain_step(w, v.data(), P);
        }
    }
    light_step(w, P);
    if (w.st.fluid_enabled && P.epoch % substep == 0) fluid_step<R>(w, fl, P);
    if (P.epoch % std::max(P.transport_every, 1u) == 0) transport<R>(w, fl, P);
}

}  // namespace

// Byte sizes of the buffers (same table as the product ABI; the oracle does
// not link the product library).
static size_t oracle_buffer_bytes(const RiboWorldConfig* cfg, RiboBuffer which) {
    if (!cfg) return 0;
    size_t G = (size_t)cfg->grid_size * cfg->grid_size, F = (size_t)cfg->fluid_size * cfg->fluid_size;
    size_t L = (size_t)(cfg->grid_size / 2) * (cfg->grid_size / 2);
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

namespace {

// smooth_velocity (fluid.wgsl:680-693): the sanitized velocity of the cell
// and its four neighbours (clamped at the edges), summed in that order, x 0.2.
template <class R> void smooth_velocity(uint32_t F, const R* vin, R* vout) {
    auto at = [&](uint32_t x, uint32_t y) { return sanitize(V2<R>{vin[2 * (y * F + x)], vin[2 * (y * F + x) + 1]}); };
    for (uint32_t y = 0; y < F; ++y)
        for (uint32_t x = 0; x < F; ++x) {
            const uint32_t xm = x > 0 ? x - 1 : x, xp = std::min(x + 1, F - 1);
            const uint32_t ym = y > 0 ? y - 1 : y, yp = std::min(y + 1, F - 1);
            V2<R> a = at(x, y), b = at(xm, y), c = at(xp, y), d = at(x, ym), e = at(x, yp);
            R vx = a.x + b.x, vy = a.y + b.y;
            vx = vx + c.x; vy = vy + c.y;
            vx = vx + d.x; vy = vy + d.y;
            vx = vx + e.x; vy = vy + e.y;
            const R fifth = sizeof(R) == 4 ? R(0.2f) : R(0.2);
            vout[2 * (y * F + x)] = vx * fifth;
            vout[2 * (y * F + x) + 1] = vy * fifth;
        }
}

size_t buf_bytes(const OracleWorld& w, int which) {
    RiboWorldConfig cfg{w.G, w.F, w.wupc, 0};
    return oracle_buffer_bytes(&cfg, (RiboBuffer)which);
}

template <class R> R* fluid_ptr(OracleWorld& w, Fluid<R>& fl, int which, size_t* count) {
    size_t F = (size_t)w.F * w.F;
    switch (which) {
        case RIBO_BUF_VELOCITY: *count = 2 * F; return fl.vel_a.data();
        case RIBO_BUF_PRESSURE: *count = F; return fl.p_a.data();
        case RIBO_BUF_TEMP: *count = F; return fl.temp_in.data();
        case RIBO_BUF_REDOX: *count = F; return fl.redox_in.data();
        case RIBO_BUF_DIVERGENCE: *count = F; return fl.div.data();
        case RIBO_BUF_NET_FLUX: *count = 2 * (size_t)w.F; return fl.net_flux.data();
        default: *count = 0; return nullptr;
    }
}

}  // namespace

// ---- C API -------------------------------------------------------------------------
extern "C" {

OracleWorld* oracle_create(const RiboWorldConfig* cfg, uint32_t seed, int fp64) {
    if (!cfg || cfg->fluid_size == 0 || cfg->grid_size % cfg->fluid_size != 0) return nullptr;
    OracleWorld* w = new OracleWorld();
    w->G = cfg->grid_size;
    w->F = cfg->fluid_size;
    w->wupc = cfg->world_units_per_cell;
    w->Lsz = cfg->grid_size / LIGHT_DIV;
    w->sim_size = (float)(cfg->grid_size * cfg->world_units_per_cell);
    w->fp64 = fp64;
    ribo_sim_params_fill_default(&w->params);  // SimParams::default() (src/params.rs:553-690)
    w->params.seed = seed;
    RiboWorldSettings s;
    memset(&s, 0, sizeof(s));
    s.fluid_enabled = 1; s.fluid_substep = 2; s.multigrid = 1; s.mg_cycles = 1; s.jacobi_iters = 128;
    s.light_rows_per_step = 1; s.light_interval = 10; s.terrain_enabled = 1; s.contact_enabled = 1; s.light_mode = 1;
    s.mg_min_size = 8; s.mg_tol = 0.0f; s.mg_max_cycles = 1; s.precision = 0; s.ledger_check = 0;
    w->st = s;
    size_t G = (size_t)w->G * w->G, F = (size_t)w->F * w->F;
    w->chem.assign(4 * G, 0); w->chem_next.assign(4 * G, 0); w->gamma.assign(G, 0);
    w->agg_act.assign(G, 0); w->agg_nb.assign(G, 0);
    w->solid_mask.assign(F, 0); w->redox_eaten.assign(F, 0); w->force_vectors.assign(2 * F, 0);
    w->light.assign(2 * (size_t)w->Lsz * w->Lsz, 0.0f); w->light_next = w->light; w->slope.assign(2 * G, 0.0f); w->heat_src.assign(2 * F, 0.0f);
    make_levels(w->F, 8, &w->mg_total);
    w->mg_fluid.assign(w->mg_total, 0);
    if (fp64) w->f64.alloc(F, w->F, w->mg_total);
    else w->f32.alloc(F, w->F, w->mg_total);
    return w;
}

void oracle_destroy(OracleWorld* w) { delete w; }
void oracle_invalidate_light(OracleWorld* w) { w->light_dirty = true; }  // src/world/mod.rs:1400-1402
void oracle_set_params(OracleWorld* w, const RiboSimParams* p) { w->params = *p; }
void oracle_get_params(const OracleWorld* w, RiboSimParams* p) { *p = w->params; }

int oracle_set_settings(OracleWorld* w, const RiboWorldSettings* s) {
    if (s->mg_min_size < 1 || s->light_mode > 1u) return RIBO_ERR_ARG;
    w->st = *s;
    w->f32.hist = w->f64.hist = 0;  // new settings start the pressure history again
    uint32_t total;
    make_levels(w->F, s->mg_min_size, &total);
    if (total != w->mg_total) {
        w->mg_total = total;
        w->mg_fluid.assign(total, 0);
        if (w->fp64) { w->f64.mg_p.assign(total, 0); w->f64.mg_rhs.assign(total, 0); }
        else { w->f32.mg_p.assign(total, 0); w->f32.mg_rhs.assign(total, 0); }
    }
    return RIBO_OK;
}

static void compute_slope(OracleWorld* w) {  // terrain.wgsl:104-133
    const int32_t G = (int32_t)w->G;
    auto h = [&](int32_t ix, int32_t iy) {
        uint32_t x = (uint32_t)rb_clampi(ix, 0, G - 1), y = (uint32_t)rb_clampi(iy, 0, G - 1);
        return (float)w->gamma[y * (uint32_t)G + x] * GAMMA_QUANTUM;
    };
    for (int32_t iy = 0; iy < G; ++iy)
        for (int32_t ix = 0; ix < G; ++ix) {
            float l = h(ix - 1, iy), r = h(ix + 1, iy), b = h(ix, iy - 1), t = h(ix, iy + 1);
            float bl = h(ix - 1, iy - 1), br = h(ix + 1, iy - 1), tl = h(ix - 1, iy + 1), tr = h(ix + 1, iy + 1);
            const float k = (float)(1.0 / (4.0 * 1.41421356237));
            float dx = (r - l) * 0.5f + (tr + br - tl - bl) * k;
            float dy = (t - b) * 0.5f + (tl + tr - bl - br) * k;
            size_t i = (size_t)iy * G + ix;
            w->slope[2 * i] = dx / (float)w->wupc;
            w->slope[2 * i + 1] = dy / (float)w->wupc;
        }
}

int oracle_upload(OracleWorld* w, int which, const void* src, size_t bytes) {
    if (bytes != buf_bytes(*w, which)) return RIBO_ERR_ARG;
    if (which == RIBO_BUF_GAMMA || which == RIBO_BUF_VELOCITY || which == RIBO_BUF_PRESSURE)
        w->f32.hist = w->f64.hist = 0;  // the pressure history starts again
    switch (which) {
        case RIBO_BUF_CHEM: memcpy(w->chem.data(), src, bytes); w->light_dirty = true; return 0;
        case RIBO_BUF_GAMMA: memcpy(w->gamma.data(), src, bytes); compute_slope(w); w->light_dirty = true;
            {
                Ctx c{w, w->G, w->F, &w->params};
                build_solid_mask<float>(*w, c);
            }
            return 0;
        case RIBO_BUF_LIGHT: memcpy(w->light.data(), src, bytes); w->light_dirty = false; return 0;
        case RIBO_BUF_HEAT_SRC: memcpy(w->heat_src.data(), src, bytes); return 0;
        case RIBO_BUF_REDOX_EATEN: memcpy(w->redox_eaten.data(), src, bytes); return 0;
        case RIBO_BUF_FORCE_VECTORS: memcpy(w->force_vectors.data(), src, bytes); return 0;  // fold.wgsl:495-509
        case RIBO_BUF_VELOCITY_SMOOTH: return RIBO_ERR_ARG;
        default: break;
    }
    size_t count;
    const float* s = (const float*)src;
    if (w->fp64) {
        double* d = fluid_ptr(*w, w->f64, which, &count);
        if (!d || which == RIBO_BUF_DIVERGENCE || which == RIBO_BUF_NET_FLUX) return RIBO_ERR_ARG;
        for (size_t i = 0; i < count; ++i) d[i] = s[i];
    } else {
        float* d = fluid_ptr(*w, w->f32, which, &count);
        if (!d || which == RIBO_BUF_DIVERGENCE || which == RIBO_BUF_NET_FLUX) return RIBO_ERR_ARG;
        memcpy(d, s, bytes);
    }
    return 0;
}

int oracle_download(const OracleWorld* cw, int which, void* dst, size_t bytes) {
    OracleWorld* w = const_cast<OracleWorld*>(cw);
    if (bytes != buf_bytes(*w, which)) return RIBO_ERR_ARG;
    switch (which) {
        case RIBO_BUF_CHEM: memcpy(dst, w->chem.data(), bytes); return 0;
        case RIBO_BUF_GAMMA: memcpy(dst, w->gamma.data(), bytes); return 0;
        case RIBO_BUF_LIGHT: memcpy(dst, w->light.data(), bytes); return 0;
        case RIBO_BUF_SLOPE: memcpy(dst, w->slope.data(), bytes); return 0;
        case RIBO_BUF_HEAT_SRC: memcpy(dst, w->heat_src.data(), bytes); return 0;
        case RIBO_BUF_SOLID_MASK: memcpy(dst, w->solid_mask.data(), bytes); return 0;
        case RIBO_BUF_REDOX_EATEN: memcpy(dst, w->redox_eaten.data(), bytes); return 0;
        case RIBO_BUF_AGG:
            for (size_t i = 0; i < w->agg_act.size(); ++i) ((uint32_t*)dst)[i] = w->agg_act[i] + w->agg_nb[i];
            return 0;
        case RIBO_BUF_FORCE_VECTORS: memcpy(dst, w->force_vectors.data(), bytes); return 0;
        case RIBO_BUF_VELOCITY_SMOOTH:
            if (w->fp64) {
                std::vector<double> v(2 * (size_t)w->F * w->F);
                smooth_velocity<double>(w->F, w->f64.vel_a.data(), v.data());
                for (size_t i = 0; i < v.size(); ++i) ((float*)dst)[i] = (float)v[i];
            } else {
                smooth_velocity<float>(w->F, w->f32.vel_a.data(), (float*)dst);
            }
            return 0;
        default: break;
    }
    size_t count;
    float* d = (float*)dst;
    if (w->fp64) {
        double* s = fluid_ptr(*w, w->f64, which, &count);
        if (!s) return RIBO_ERR_ARG;
        for (size_t i = 0; i < count; ++i) d[i] = (float)s[i];
    } else {
        float* s = fluid_ptr(*w, w->f32, which, &count);
        if (!s) return RIBO_ERR_ARG;
        memcpy(d, s, bytes);
    }
    return 0;
}

int oracle_download_f64(const OracleWorld* cw, int which, double* dst, size_t count) {
    OracleWorld* w = const_cast<OracleWorld*>(cw);
    size_t n;
    if (w->fp64) {
        double* s = fluid_ptr(*w, w->f64, which, &n);
        if (!s || n != count) return RIBO_ERR_ARG;
        memcpy(dst, s, n * sizeof(double));
    } else {
        float* s = fluid_ptr(*w, w->f32, which, &n);
        if (!s || n != count) return RIBO_ERR_ARG;
        for (size_t i = 0; i < n; ++i) dst[i] = s[i];
    }
    return 0;
}

int oracle_encode_steps(OracleWorld* w, uint32_t steps) {  // src/world/mod.rs:1406-1446,1622,1631
    steps = std::min(steps, RIBO_MAX_STEPS_PER_CALL);
    if (steps == 0) return 0;
    const RiboWorldSettings st = w->st;
    const uint32_t substep = std::min(std::max(st.fluid_substep, 1u), 4u);
    if (!st.fluid_enabled && w->fluid_was_on) {
        if (w->fp64) {
            std::fill(w->f64.vel_a.begin(), w->f64.vel_a.end(), 0.0);
            std::fill(w->f64.temp_in.begin(), w->f64.temp_in.end(), 0.0);
            std::fill(w->f64.redox_in.begin(), w->f64.redox_in.end(), 0.0);
        } else {
            std::fill(w->f32.vel_a.begin(), w->f32.vel_a.end(), 0.0f);
            std::fill(w->f32.temp_in.begin(), w->f32.temp_in.end(), 0.0f);
            std::fill(w->f32.redox_in.begin(), w->f32.redox_in.end(), 0.0f);
        }
    }
    w->fluid_was_on = st.fluid_enabled != 0;
    w->params.fluid_dt = w->params.dt * (float)substep;
    w->params.fluid_enabled = st.fluid_enabled ? 1u : 0u;
    for (uint32_t i = 0; i < steps; ++i) {
        RiboSimParams p = w->params;
        p.epoch = w->params.epoch + i;
        p.sun_now = w->params.uv_strength > 0.0f ? day_light(w->params, p.epoch) : 0.0f;
        p.sun_slope = sun_slope_at(w->params, p.epoch);
        if (w->fp64) world_step<double>(*w, w->f64, p, substep);
        else world_step<float>(*w, w->f32, p, substep);
    }
    w->params.epoch += steps;
    return (int)steps;
}

// Runs one reference kernel on the current state with the bindings of
// encode_steps, for closed-form tests of the restatement.  Velocity results
// are written back to velocity_a.  Returns 0, or -1 for an unknown name.
int oracle_run_kernel(OracleWorld* w, const char* name) {
    if (w->fp64) return -1;
    Fluid<float>& fl = w->f32;
    const RiboSimParams& P = w->params;
    Ctx c{w, w->G, w->F, &P};
    const std::string n(name);
    if (n == "relax_a" || n == "relax_b") relax_pass(*w, P, fl.vel_a.data(), n == "relax_b" ? 1u : 0u);
    else if (n == "grain_fall") grain_fall(*w, P);
    else if (n == "build_solid_mask") build_solid_mask<float>(*w, c);
    else if (n == "temperature") { update_temperature<float>(*w, fl, c); copy_temperature<float>(*w, fl); }
    else if (n == "forces") {
        buoyancy<float>(*w, fl, c);
        gather_forces<float>(*w, fl, c);
        std::fill(w->force_vectors.begin(), w->force_vectors.end(), 0u);
        add_forces<float>(*w, fl, c, fl.vel_a.data(), fl.vel_b.data());
        std::fill(fl.forces.begin(), fl.forces.end(), 0.0f);
        fl.vel_a = fl.vel_b;
    } else if (n == "diffuse") { diffuse_velocity<float>(c, fl.vel_a.data(), fl.vel_b.data()); fl.vel_a = fl.vel_b; }
    else if (n == "advect") {
        float decay = decay_factor_f32(P.fluid_decay, fluid_dt<float>(c));
        advect_velocity<float>(c, fl.vel_a.data(), fl.vel_b.data(), decay);
        fl.vel_a = fl.vel_b;
    } else if (n == "vorticity") { vorticity_confinement<float>(c, fl.vel_a.data(), fl.vel_b.data()); fl.vel_a = fl.vel_b; }
    else if (n == "divergence") compute_divergence<float>(c, fl.vel_a.data(), fl.div.data());
    else if (n == "subtract_gradient") { subtract_gradient<float>(c, fl.vel_a.data(), fl.vel_b.data(), fl.p_a.data()); fl.vel_a = fl.vel_b; }
    else if (n == "net_flux") net_flux<float>(c, fl.vel_a.data(), fl.net_flux.data());
    else if (n == "boundaries") { enforce_boundaries<float>(c, fl.vel_a.data(), fl.vel_b.data(), fl.net_flux.data()); fl.vel_a = fl.vel_b; }
    else if (n == "jacobi") { jacobi_pressure<float>(c, fl.p_a.data(), fl.p_b.data(), fl.div.data()); fl.p_a = fl.p_b; }
    else if (n == "transport") transport<float>(*w, fl, P);
    else if (n == "thermal_activation") thermal_activation<float>(*w, fl, P);
    else return -1;
    return 0;
}

// The exchanges of chem.wgsl:40-98, one after another in the order given
// (chem_cas, chem_add_state, chem_open_cell).
int oracle_chem_ops(OracleWorld* w, const RiboChemOp* ops, uint32_t n, uint32_t* done) {
    const uint64_t cells = (uint64_t)w->G * w->G;
    for (uint32_t i = 0; i < n; ++i)
        if (ops[i].cell >= cells || ops[i].channel >= 4u || ops[i].kind > RIBO_CHEM_DEPOSIT_SPENT) return RIBO_ERR_ARG;
    for (uint32_t i = 0; i < n; ++i) {
        const RiboChemOp& op = ops[i];
        uint32_t moved = 0;
        if (op.kind <= RIBO_CHEM_TAKE_SPENT) {
            static const uint32_t delta[4] = {CHEM_SPENT_ONE - 1u, 0xFFFF0001u, 0xFFFFFFFFu, 0xFFFF0000u};
            const bool need_act = op.kind == RIBO_CHEM_SPEND || op.kind == RIBO_CHEM_TAKE_ACT;
            uint32_t& word = w->chem[(size_t)op.cell * 4 + op.channel];
            while (moved < op.count) {
                if (need_act ? (word & CHEM_STATE_MASK) == 0u : (word >> 16) == 0u) break;
                word += delta[op.kind];
                ++moved;
            }
        } else {
            uint32_t cell = op.cell;
            if (op.kind >= RIBO_CHEM_DEPOSIT_ACT && w->gamma[cell] >= GAMMA_SOLID_THRESHOLD) {  // chem_open_cell
                const int gx = (int)(cell % w->G), gy = (int)(cell / w->G);
                bool found = false;
                for (int r = 1; r <= 12 && !found; ++r)
                    for (int dy = -r; dy <= r && !found; ++dy)
                        for (int dx = -r; dx <= r && !found; ++dx) {
                            if (abs(dx) != r && abs(dy) != r) continue;
                            const int qx = gx + dx, qy = gy + dy;
                            if (qx < 0 || qy < 0 || qx >= (int)w->G || qy >= (int)w->G) continue;
                            const uint32_t q = (uint32_t)qy * w->G + (uint32_t)qx;
                            if (w->gamma[q] < GAMMA_SOLID_THRESHOLD) { cell = q; found = true; }
                        }
            }
            const bool spent = op.kind == RIBO_CHEM_ADD_SPENT || op.kind == RIBO_CHEM_DEPOSIT_SPENT;
            w->chem[(size_t)cell * 4 + op.channel] += spent ? op.count * CHEM_SPENT_ONE : op.count;
            moved = op.count;
        }
        if (done) done[i] = moved;
    }
    return 0;
}

void oracle_ledger(const OracleWorld* w, uint32_t out[12]) {  // Ledger::from_cells (src/world/mod.rs:65-73)
    for (int i = 0; i < 12; ++i) out[i] = 0;
    for (size_t i = 0; i < w->chem.size(); ++i) {
        out[i % 4] += w->chem[i] & 0xFFFFu;
        out[4 + i % 4] += w->chem[i] >> 16;
    }
}

void oracle_mg_stats(const OracleWorld* w, uint32_t* cycles, double* rel) {
    *cycles = w->last_cycles;
    *rel = w->last_rel_res;
}
void oracle_solver_residuals(const OracleWorld* w, double* refined, double* rounded, uint32_t* restarts) {
    *refined = w->last_rel_res;
    *rounded = w->last_rounded_rel;
    *restarts = w->last_restarts;
}

double oracle_mg_residual(const OracleWorld* cw) {
    OracleWorld* w = const_cast<OracleWorld*>(cw);
    uint32_t total;
    std::vector<Level> lv = make_levels(w->F, w->st.mg_min_size, &total);
    std::vector<uint32_t> fl((size_t)w->F * w->F);
    for (size_t i = 0; i < fl.size(); ++i) fl[i] = w->solid_mask[i] == 0u ? 1u : 0u;
    if (w->fp64) {
        std::vector<double> rhs(w->f64.div), p(w->f64.p_a);
        for (size_t i = 0; i < fl.size(); ++i) if (!fl[i]) { rhs[i] = 0; p[i] = 0; }
        MG<double> mg{&lv, p.data(), rhs.data(), fl.data()};
        return mg.rel_residual();
    }
    std::vector<float> rhs(w->f32.div), p(w->f32.p_a);
    for (size_t i = 0; i < fl.size(); ++i) if (!fl[i]) { rhs[i] = 0; p[i] = 0; }
    MG<float> mg{&lv, p.data(), rhs.data(), fl.data()};
    return mg.rel_residual();
}

void oracle_set_threads(int n) {
    g_threads = n;
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#endif
}
int oracle_get_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void oracle_pcg4d(const uint32_t* in, uint32_t* out, size_t n) {
    for (size_t i = 0; i < n; ++i) {
        uint32_t v[4] = {in[4 * i], in[4 * i + 1], in[4 * i + 2], in[4 * i + 3]};
        pcg4d(v);
        memcpy(out + 4 * i, v, sizeof(v));
    }
}
uint32_t oracle_hash(uint32_t v) { return hash_u32(v); }
float oracle_expf(float x) { return rb_expf(x); }
float oracle_sinf(float x) { return rb_sinf(x); }

}  // extern "C"

template <class R>
static int mg_solve_t(uint32_t n, uint32_t min_size, const uint32_t* fluid, const R* rhs, R* p, uint32_t cycles) {
    uint32_t total;
    std::vector<Level> lv = make_levels(n, min_size, &total);
    std::vector<R> P(total, 0), RHS(total, 0);
    std::vector<uint32_t> FL(total, 0);
    for (uint32_t i = 0; i < n * n; ++i) {
        FL[i] = fluid[i] ? 1u : 0u;
        RHS[i] = fluid[i] ? rhs[i] : R(0);
        P[i] = fluid[i] ? p[i] : R(0);
    }
    MG<R> mg{&lv, P.data(), RHS.data(), FL.data()};
    for (uint32_t k = 0; k < cycles; ++k) mg.vcycle();
    for (uint32_t i = 0; i < n * n; ++i) p[i] = P[i];
    return (int)lv.size();
}

extern "C" {

int oracle_mg_solve(uint32_t n, uint32_t min_size, const uint32_t* fluid, const float* rhs, float* p, uint32_t cycles,
                    int fp64) {
    if (!fp64) return mg_solve_t<float>(n, min_size, fluid, rhs, p, cycles);
    std::vector<double> r(rhs, rhs + (size_t)n * n), q(p, p + (size_t)n * n);
    int l = mg_solve_t<double>(n, min_size, fluid, r.data(), q.data(), cycles);
    for (size_t i = 0; i < q.size(); ++i) p[i] = (float)q[i];
    return l;
}
int oracle_mg_solve_f64(uint32_t n, uint32_t min_size, const uint32_t* fluid, const double* rhs, double* p,
                        uint32_t cycles) {
    return mg_solve_t<double>(n, min_size, fluid, rhs, p, cycles);
}

// One mg_prolong (multigrid.wgsl:108-114): p += coarse p of the parent on
// fluid cells of an n x n level.
int oracle_mg_prolong(uint32_t n, const uint32_t* fluid, float* p, const float* p_c) {
    if (n < 2 || n % 2) return -1;
    uint32_t total;
    std::vector<Level> lv = make_levels(n, n / 2, &total);
    std::vector<float> P(total, 0.0f), RHS(total, 0.0f);
    std::vector<uint32_t> FL(total, 0);
    for (uint32_t i = 0; i < n * n; ++i) { FL[i] = fluid[i] ? 1u : 0u; P[i] = p[i]; }
    for (uint32_t i = 0; i < (n / 2) * (n / 2); ++i) P[lv[0].off_c + i] = p_c[i];
    MG<float> mg{&lv, P.data(), RHS.data(), FL.data()};
    mg.prolong(lv[0]);
    for (uint32_t i = 0; i < n * n; ++i) p[i] = P[i];
    return 0;
}

// The bilinear transfer of the tolerance mode between an n x n level and
// its coarse level, binary64: prolong adds the interpolated coarse values
// (P c) to p, restrict writes R r (p = 0, so the residual is rhs) into rhs_c.
// Fluid masks: fine u32[n*n], coarse u32[(n/2)^2].
int oracle_bl_transfer(uint32_t n, const uint32_t* fluid, const uint32_t* fluid_c, const double* c_in, double* p_out,
                       const double* r_in, double* rhs_c) {
    if (n < 2 || n % 2) return -1;
    uint32_t total;
    std::vector<Level> lv = make_levels(n, n / 2, &total);
    const uint32_t m = (n / 2) * (n / 2);
    std::vector<double> P(total, 0.0), RHS(total, 0.0);
    std::vector<uint32_t> FL(total, 0);
    for (uint32_t i = 0; i < n * n; ++i) FL[i] = fluid[i] ? 1u : 0u;
    for (uint32_t i = 0; i < m; ++i) FL[lv[0].off_c + i] = fluid_c[i] ? 1u : 0u;
    MG<double> mg{&lv, P.data(), RHS.data(), FL.data()};
    for (uint32_t i = 0; i < m; ++i) P[lv[0].off_c + i] = c_in[i];
    mg.prolong_bl(lv[0]);
    for (uint32_t i = 0; i < n * n; ++i) p_out[i] = P[i];
    std::fill(P.begin(), P.end(), 0.0);
    for (uint32_t i = 0; i < n * n; ++i) RHS[i] = r_in[i];
    std::vector<double> rn;
    mg.restrict_bl(lv[0], rn);
    for (uint32_t i = 0; i < m; ++i) rhs_c[i] = RHS[lv[0].off_c + i];
    return 0;
}

// The exact coarse solve of the tolerance mode: max |A inv - I| over the
// fluid unknowns of an n x n mask (eps as given), or -1 when not built
// (no fluid, or more than 1024 unknowns).
double oracle_coarse_inverse_error(uint32_t n, const uint32_t* fluid, float eps) {
    CoarseInv ci;
    if (!ci.build(n, fluid, eps)) return -1.0;
    const size_t m = ci.cell.size();
    double worst = 0.0;
    for (size_t r = 0; r < m; ++r) {
        const uint32_t c = ci.cell[r], x = c % n, y = c / n;
        for (size_t j = 0; j < m; ++j) {
            // row r of A (fluid neighbours 1, diagonal -(k + eps)) times column j of inv
            double acc = 0.0, k = 0.0;
            const int64_t nb[4] = {x > 0 ? (int64_t)c - 1 : -1, x + 1 < n ? (int64_t)c + 1 : -1,
                                   y > 0 ? (int64_t)c - n : -1, y + 1 < n ? (int64_t)c + n : -1};
            for (int d = 0; d < 4; ++d)
                if (nb[d] >= 0 && ci.idx[(size_t)nb[d]] >= 0) {
                    acc += ci.inv[(size_t)ci.idx[(size_t)nb[d]] * m + j];
                    k += 1.0;
                }
            acc -= (k + (double)eps) * ci.inv[r * m + j];
            worst = std::max(worst, fabs(acc - (r == j ? 1.0 : 0.0)));
        }
    }
    return worst;
}

// One mg_restrict (multigrid.wgsl:82-105) of an n x n level (n even).
int oracle_mg_restrict(uint32_t n, const uint32_t* fluid, const float* rhs, const float* p, float* rhs_c,
                       uint32_t* fluid_c) {
    if (n < 2 || n % 2) return -1;
    uint32_t total;
    std::vector<Level> lv = make_levels(n, n / 2, &total);
    std::vector<float> P(total, 0.0f), RHS(total, 0.0f);
    std::vector<uint32_t> FL(total, 0);
    for (uint32_t i = 0; i < n * n; ++i) {
        FL[i] = fluid[i] ? 1u : 0u;
        RHS[i] = rhs[i];
        P[i] = p[i];
    }
    MG<float> mg{&lv, P.data(), RHS.data(), FL.data()};
    mg.restrict_(lv[0]);
    const uint32_t m = (n / 2) * (n / 2);
    for (uint32_t i = 0; i < m; ++i) {
        rhs_c[i] = RHS[lv[0].off_c + i];
        fluid_c[i] = FL[lv[0].off_c + i];
    }
    return 0;
}

}  // extern "C"
