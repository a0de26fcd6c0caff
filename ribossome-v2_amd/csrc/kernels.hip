// kernels.hip - HIP kernels of the world step for gfx950 (MI355X).
//
// Every kernel computes each output value with the same sequence of
// binary32 operations as the reference WGSL (the file:line next to each
// kernel), compiled without contraction and with correctly rounded divide
// and sqrt, so the results are bit-identical to a faithful restatement.
// What differs from the reference is the data path: terrain-derived values
// (wall mask, wall normals, slope at the fluid sample, grain fill, rubble
// fraction, multigrid masks) are computed once per terrain instead of once
// per use; the temperature ping-pongs instead of being copied; the level-0
// multigrid arrays alias the pressure and divergence fields; force
// quantisation and gathering fuse into one pass.  All buffers are
// HBM-resident.  The work is stencils, gathers, integer scatter and
// reductions: nothing here is a dense contraction, so no MFMA.

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>
#include <type_traits>

#include "kernels.h"
#include "wgsl_sem.h"

namespace ribo {
namespace {

// ---- constants (fluid.wgsl:15-56, transport.wgsl:20-51, rng.wgsl:25-33,
// chem.wgsl:6-7, common/terrain.wgsl:4, light.wgsl:18-40,181) ----------------
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
constexpr uint32_t S_PHOTO = 1u;
constexpr uint32_t S_DECAY = 2u;
constexpr uint32_t S_THERMAL = 3u;
constexpr uint32_t S_BUOYANCY = 4u;
constexpr uint32_t S_MOVE = 1u << 16;
constexpr uint32_t S_DISPERSE = 6u << 16;
constexpr uint32_t CHEM_STATE_MASK = 0xFFFFu;
constexpr uint32_t CHEM_SPENT_ONE = 0x10000u;
constexpr uint32_t CHEM_CELL_CAP = 48u;
constexpr uint32_t GAMMA_SOLID = 3u;
constexpr float UV_SHADOW_ABSORB = 0.6f;
constexpr int LIGHT_TAPS = 7;
constexpr float SUN_WATER_ABSORB = 0.05f;
constexpr uint32_t LIGHT_DIV = 2u;
constexpr float GAMMA_QUANTUM = 0.25f;

// ---- PCG4D counter RNG (rng.wgsl:10-43; Jarzynski & Olano, JCGT 2020) --------
struct U4 { uint32_t x, y, z, w; };
__device__ __forceinline__ U4 pcg4d(U4 v) {
    v.x = v.x * 1664525u + 1013904223u;
    v.y = v.y * 1664525u + 1013904223u;
    v.z = v.z * 1664525u + 1013904223u;
    v.w = v.w * 1664525u + 1013904223u;
    v.x += v.y * v.w;
    v.y += v.z * v.x;
    v.z += v.x * v.y;
    v.w += v.y * v.z;
    v.x ^= v.x >> 16; v.y ^= v.y >> 16; v.z ^= v.z >> 16; v.w ^= v.w >> 16;
    v.x += v.y * v.w;
    v.y += v.z * v.x;
    v.z += v.x * v.y;
    v.w += v.y * v.z;
    return v;
}
struct F4 { float x, y, z, w; };
__device__ __forceinline__ F4 u4_to_f4(U4 v) {
    const float s = 1.0f / 16777216.0f;
    return {(float)(v.x >> 8) * s, (float)(v.y >> 8) * s, (float)(v.z >> 8) * s, (float)(v.w >> 8) * s};
}
__device__ __forceinline__ F4 rng_f4(uint32_t key, uint32_t epoch, uint32_t stream, uint32_t seed) {
    return u4_to_f4(pcg4d({key, epoch, stream, seed}));
}

// ---- small vector helpers -------------------------------------------------------
__device__ __forceinline__ float2 f2(float x, float y) { return make_float2(x, y); }
__device__ __forceinline__ float2 add2(float2 a, float2 b) { return f2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 sub2(float2 a, float2 b) { return f2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 mul2(float2 a, float s) { return f2(a.x * s, a.y * s); }
__device__ __forceinline__ float2 mix2(float2 a, float2 b, float t) { return f2(rb_mix(a.x, b.x, t), rb_mix(a.y, b.y, t)); }
__device__ __forceinline__ float len2(float2 v) { return rb_len2(v.x, v.y); }
__device__ __forceinline__ float dot2(float2 a, float2 b) { return a.x * b.x + a.y * b.y; }

// fluid.wgsl:70-83
__device__ __forceinline__ bool is_bad(float x) { return (x != x) || (rb_abs(x) > 1e20f); }
__device__ __forceinline__ float2 sanitize(float2 v) {
    if (is_bad(v.x) || is_bad(v.y)) return f2(0.0f, 0.0f);
    return v;
}
__device__ __forceinline__ float2 clamp_len(float2 v, float m) {
    float len = len2(v);
    if (len > m) return mul2(v, m / rb_max(len, 1e-12f));
    return v;
}
__device__ __forceinline__ float fluid_dt(const RiboSimParams& P) { return rb_clamp(P.fluid_dt, 0.0f, MAX_DT); }
__device__ __forceinline__ float temp_ambient_at(uint32_t y, uint32_t F) {  // fluid.wgsl:62-64
    return TEMP_AMBIENT_SURFACE * rb_expf(-TEMP_AMBIENT_ATTEN * (1.0f - ((float)y + 0.5f) / (float)F));
}
__device__ __forceinline__ bool solid(const Geo& g, uint32_t i) { return (g.cflags[i] & CF_SOLID) != 0; }
// reflect_if_into_solid (fluid.wgsl:219-224) with the precomputed normal.
__device__ __forceinline__ float2 reflect(const Geo& g, uint32_t i, float2 v) {
    if (g.cflags[i] & CF_NORMAL) {
        float2 n = g.normal[i];
        float d = dot2(v, n);
        if (len2(n) > 1e-6f && d < 0.0f) return sub2(v, mul2(n, d));
    }
    return v;
}
// (the flag and the velocity are both loaded and the result selected, so a
// stencil's loads go out together instead of flag-then-velocity per cell)
__device__ __forceinline__ float2 raw_vel(const Geo& g, const float2* vin, uint32_t i) {  // fluid.wgsl:191-194
    const bool wall = solid(g, i);
    const float2 v = sanitize(vin[i]);
    return wall ? f2(0.0f, 0.0f) : v;
}
__device__ __forceinline__ uint32_t env_cell_for_fluid(const Geo& g, uint32_t x, uint32_t y) {  // fluid.wgsl:87-92
    float s = (float)g.G / (float)g.F;
    uint32_t gx = rb_f2u(rb_clamp(((float)x + 0.5f) * s, 0.0f, (float)(g.G - 1u)));
    uint32_t gy = rb_f2u(rb_clamp(((float)y + 0.5f) * s, 0.0f, (float)(g.G - 1u)));
    return gy * g.G + gx;
}
__device__ __forceinline__ uint32_t cell_total(const uint32_t* chem, uint32_t cell) {  // chem.wgsl:17-24
    uint4 v = *reinterpret_cast<const uint4*>(chem + 4u * cell);
    return (v.x & CHEM_STATE_MASK) + (v.x >> 16) + (v.y & CHEM_STATE_MASK) + (v.y >> 16) + (v.z & CHEM_STATE_MASK) +
           (v.z >> 16) + (v.w & CHEM_STATE_MASK) + (v.w >> 16);
}
__device__ __forceinline__ uint32_t capacity(uint32_t gam) {  // chem.wgsl:29-33
    if (gam >= GAMMA_SOLID) return 0u;
    return CHEM_CELL_CAP * (GAMMA_SOLID - gam) / GAMMA_SOLID;
}

// ---- terrain-derived data (on gamma upload) ---------------------------------------
// compute_gamma_slope (terrain.wgsl:104-133)
__global__ __launch_bounds__(256) void k_slope(Geo g, float2* slope) {
    const int32_t G = (int32_t)g.G;
    const int32_t ix = (int32_t)(blockIdx.x * 16 + threadIdx.x), iy = (int32_t)(blockIdx.y * 16 + threadIdx.y);
    if (ix >= G || iy >= G) return;
    auto h = [&](int32_t a, int32_t b) {
        uint32_t x = (uint32_t)rb_clampi(a, 0, G - 1), y = (uint32_t)rb_clampi(b, 0, G - 1);
        return (float)g.gamma[y * (uint32_t)G + x] * GAMMA_QUANTUM;
    };
    float l = h(ix - 1, iy), r = h(ix + 1, iy), b = h(ix, iy - 1), t = h(ix, iy + 1);
    float bl = h(ix - 1, iy - 1), br = h(ix + 1, iy - 1), tl = h(ix - 1, iy + 1), tr = h(ix + 1, iy + 1);
    const float k = (float)(1.0 / (4.0 * 1.41421356237));
    float dx = (r - l) * 0.5f + (tr + br - tl - bl) * k;
    float dy = (t - b) * 0.5f + (tl + tr - bl - br) * k;
    slope[(uint32_t)iy * (uint32_t)G + (uint32_t)ix] = f2(dx / (float)g.wupc, dy / (float)g.wupc);
}

// build_solid_mask (fluid.wgsl:185-189) with permeability and
// gamma_solidity_at_fluid_cell (:95-109,138-142); gamma_fill_at_fluid_cell
// (:114-128); the grain fraction of rubble_drag (:150-161); the slope sample
// of slope_steer_velocity (:166).
__global__ __launch_bounds__(256) void k_fluid_static(Geo g, uint8_t* cflags, float* fill, float* dragg, float2* slope_f) {
    const uint32_t x = blockIdx.x * 16 + threadIdx.x, y = blockIdx.y * 16 + threadIdx.y;
    if (x >= g.F || y >= g.F) return;
    const uint32_t i = y * g.F + x;
    const float scale = (float)g.G / (float)g.F;
    const float gx = ((float)x + 0.5f) * scale, gy = ((float)y + 0.5f) * scale;
    const float h = scale * 0.25f;
    uint32_t n_solid = 0;
    float fl = 0.0f;
    for (uint32_t s = 0; s < 4; ++s) {
        float ox = (s & 1u) == 1u ? h : -h;
        float oy = (s & 2u) == 2u ? h : -h;
        uint32_t sx = rb_f2u(rb_clamp(gx + ox, 0.0f, (float)(g.G - 1u)));
        uint32_t sy = rb_f2u(rb_clamp(gy + oy, 0.0f, (float)(g.G - 1u)));
        uint32_t c = g.gamma[sy * g.G + sx];
        if (c >= GAMMA_SOLID) n_solid += 1u;
        fl += (float)rb_minu(c, GAMMA_SOLID) / (float)GAMMA_SOLID;
    }
    const float solidity = (float)n_solid * 0.25f;
    const float perm = solidity >= 0.5f ? 0.0f : rb_clamp(1.0f - solidity * 2.0f, 0.0f, 1.0f);
    cflags[i] = perm < SOLID_PERM_THRESHOLD ? CF_SOLID : 0;
    fill[i] = fl * 0.25f;
    const uint32_t sc = g.G / g.F;
    float gr = 0.0f;
    for (uint32_t dy = 0; dy < sc; ++dy)
        for (uint32_t dx = 0; dx < sc; ++dx) {
            uint32_t c = g.gamma[(y * sc + dy) * g.G + x * sc + dx];
            if (c < GAMMA_SOLID) gr += (float)c;
        }
    gr /= (float)(sc * sc) * (float)GAMMA_SOLID;
    dragg[i] = gr;
    slope_f[i] = g.slope[env_cell_for_fluid(g, x, y)];
}

// solid_normal_from_neighbors (fluid.wgsl:197-216)
__global__ __launch_bounds__(256) void k_fluid_normals(Geo g, uint8_t* cflags, float2* normal) {
    const uint32_t x = blockIdx.x * 16 + threadIdx.x, y = blockIdx.y * 16 + threadIdx.y;
    const uint32_t F = g.F;
    if (x >= F || y >= F) return;
    auto so = [&](uint32_t a, uint32_t b) { return (cflags[b * F + a] & CF_SOLID) != 0; };
    const uint32_t xm = x == 0u ? x : x - 1u, xp = x + 1u >= F ? x : x + 1u;
    const uint32_t ym = y == 0u ? y : y - 1u, yp = y + 1u >= F ? y : y + 1u;
    float s_l = (x > 0u && so(xm, y)) ? 1.0f : 0.0f;
    float s_r = (x + 1u < F && so(xp, y)) ? 1.0f : 0.0f;
    float s_b = (y > 0u && so(x, ym)) ? 1.0f : 0.0f;
    float s_t = (y + 1u < F && so(x, yp)) ? 1.0f : 0.0f;
    float s_bl = (x > 0u && y > 0u && so(xm, ym)) ? 1.0f : 0.0f;
    float s_tl = (x > 0u && y + 1u < F && so(xm, yp)) ? 1.0f : 0.0f;
    float s_br = (x + 1u < F && y > 0u && so(xp, ym)) ? 1.0f : 0.0f;
    float s_tr = (x + 1u < F && y + 1u < F && so(xp, yp)) ? 1.0f : 0.0f;
    const float diag = 0.70710678f;
    float nx = (s_l - s_r) + diag * ((s_bl + s_tl) - (s_br + s_tr));
    float ny = (s_b - s_t) + diag * ((s_bl + s_br) - (s_tl + s_tr));
    float n_len = rb_len2(nx, ny);
    float2 n = n_len > 1e-6f ? f2(nx / n_len, ny / n_len) : f2(0.0f, 0.0f);
    normal[y * F + x] = n;
    // Only this thread writes its own flag byte, and the reads of the other
    // threads test bit 0 (CF_SOLID), which no thread changes here.
    if (n_len > 1e-6f) cflags[y * F + x] = (uint8_t)(cflags[y * F + x] | CF_NORMAL);
}

__global__ __launch_bounds__(256) void k_mg_mask0(Geo g, uint8_t* fluid0) {  // mg_init (multigrid.wgsl:54-55)
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= g.F * g.F) return;
    fluid0[i] = (g.cflags[i] & CF_SOLID) ? 0 : 1;
}
__global__ __launch_bounds__(256) void k_mg_mask_coarsen(uint32_t nc, const uint8_t* fine, uint8_t* coarse) {  // multigrid.wgsl:98
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nc * nc) return;
    const uint32_t gx = i % nc, gy = i / nc, n = 2 * nc;
    uint32_t nf = 0;
    for (uint32_t c = 0; c < 4; ++c) nf += fine[(gy * 2u + (c >> 1)) * n + gx * 2u + (c & 1u)] ? 1u : 0u;
    coarse[i] = nf > 0u ? 1 : 0;
}
__global__ __launch_bounds__(256) void k_mask_u32(Geo g, uint32_t* out) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= g.F * g.F) return;
    out[i] = (g.cflags[i] & CF_SOLID) ? 1u : 0u;
}

// The step's sun from the per-epoch table, when the world keeps one.
__device__ __forceinline__ void sun_of_step(const Geo& g, RiboSimParams& P, const uint32_t* ep) {
    if (!g.sun) return;
    const uint32_t e = P.epoch + (ep ? *ep : 0u);
    const float2 v = g.sun->v[e - g.sun->base];
    P.sun_now = v.x;
    P.sun_slope = v.y;
}

// ---- light on the frozen light field (light.wgsl:46-102,182-189) -------------------
__device__ float light_above(const Geo& g, const RiboSimParams& P, uint32_t x, uint32_t row, uint32_t band) {
    const uint32_t LS = g.Lsz, LC = LS * LS;
    float sh = -P.sun_slope;
    int32_t base = rb_f2i(rb_floor(sh));
    float f = sh - (float)base;
    float s = 0.0f;
    for (int d = 0; d < LIGHT_TAPS; ++d) {
        int32_t x0 = rb_clampi((int32_t)x + d - LIGHT_TAPS / 2 + base, 0, (int32_t)LS - 1);
        int32_t x1 = rb_clampi(x0 + 1, 0, (int32_t)LS - 1);
        s += rb_mix(g.light[band * LC + row * LS + (uint32_t)x0], g.light[band * LC + row * LS + (uint32_t)x1], f);
    }
    float cpos = (float)x + sh;
    float inside = rb_clamp(cpos + 1.0f, 0.0f, 1.0f) * rb_clamp((float)LS - cpos, 0.0f, 1.0f);
    return s / (float)LIGHT_TAPS * inside;
}
__device__ float light_transmit0(const Geo& g, const RiboSimParams& P, const uint32_t* chem, uint32_t lx, uint32_t ly) {
    float gg = 0.0f;
    uint32_t m = 0;
    for (uint32_t dy = 0; dy < LIGHT_DIV; ++dy)
        for (uint32_t dx = 0; dx < LIGHT_DIV; ++dx) {
            uint32_t c = (ly * LIGHT_DIV + dy) * g.G + lx * LIGHT_DIV + dx;
            gg += (float)g.gamma[c];
            m += cell_total(chem, c);
        }
    gg /= (float)LIGHT_DIV;
    float dens = (float)m / (float)(LIGHT_DIV * LIGHT_DIV);
    float tau_m = rb_max(P.monomer_uv_absorb, 0.0f) * dens * (float)LIGHT_DIV / (float)g.G * 1.0f;
    float shade = 0.0f;  // agents are not part of the world step
    return rb_expf(-UV_SHADOW_ABSORB * gg - 0.15f * shade - tau_m);
}
__device__ float light_absorbed_at(const Geo& g, const RiboSimParams& P, const uint32_t* chem, uint32_t env) {
    uint32_t lx = (env % g.G) / LIGHT_DIV, ly = (env / g.G) / LIGHT_DIV;
    float incoming = P.sun_now;
    if (ly + 1u < g.Lsz) incoming = light_above(g, P, lx, ly + 1u, 0u);
    incoming = rb_max(incoming, 0.0f);
    return incoming * ((1.0f - light_transmit0(g, P, chem, lx, ly)) + SUN_WATER_ABSORB);
}
// ---- light propagation (light.wgsl:44-46,56-82,84-102,104-176) ----------------------
// The light field is two bands of LS x LS cells (LS = grid / LIGHT_DIV) stored
// band after band; row LS - 1 is the top, where the sun enters.
constexpr float DEEP_LIGHT_DEPTH = 0.35f;
constexpr float DEEP_WATER_DEPTH = 0.03f;

// row_water_of (light.wgsl:44-46): water attenuation across one light row.
__device__ __forceinline__ float row_water(const RiboSimParams& P, uint32_t G, uint32_t band) {
    return rb_expf(-rb_max(P.uv_depth, 0.0f) * (band == 1u ? DEEP_WATER_DEPTH : 1.0f) * (float)LIGHT_DIV / (float)G);
}
// The sun's sideways shift of light_above (light.wgsl:91-93): light at (x, y)
// comes from (x - sun_slope, y + 1), integer part `base`, fraction f.
struct SunShift { float sh, f; int32_t base; };
__device__ __forceinline__ SunShift sun_shift(const RiboSimParams& P) {
    SunShift s;
    s.sh = -P.sun_slope;
    s.base = rb_f2i(rb_floor(s.sh));
    s.f = s.sh - (float)s.base;
    return s;
}
// light_above (light.wgsl:90-102) over one row of one band, from HBM or LDS.
template <class Row> __device__ __forceinline__ float light_above_row(Row row, uint32_t LS, uint32_t x, const SunShift& s) {
    float acc = 0.0f;
#pragma unroll
    for (int d = 0; d < LIGHT_TAPS; ++d) {
        const int32_t x0 = rb_clampi((int32_t)x + d - LIGHT_TAPS / 2 + s.base, 0, (int32_t)LS - 1);
        const int32_t x1 = rb_clampi(x0 + 1, 0, (int32_t)LS - 1);
        acc += rb_mix(row[x0], row[x1], s.f);
    }
    const float cpos = (float)x + s.sh;
    const float inside = rb_clamp(cpos + 1.0f, 0.0f, 1.0f) * rb_clamp((float)LS - cpos, 0.0f, 1.0f);
    return acc / (float)LIGHT_TAPS * inside;
}
// light_transmit (light.wgsl:56-82) of light cell (lx, ly) for both bands:
// the grain and monomer sums over its LIGHT_DIV^2 grid cells are formed once
// (the bands differ only in the monomer depth factor).  Agent shade is zero.
__device__ __forceinline__ void light_transmit2(const Geo& g, const RiboSimParams& P, const uint32_t* chem, uint32_t lx,
                                                uint32_t ly, float* t0, float* t1) {
    const uint4* c4 = reinterpret_cast<const uint4*>(chem);
    float gg = 0.0f;
    uint32_t m = 0;
#pragma unroll
    for (uint32_t dy = 0; dy < LIGHT_DIV; ++dy)
#pragma unroll
        for (uint32_t dx = 0; dx < LIGHT_DIV; ++dx) {
            const uint32_t c = (ly * LIGHT_DIV + dy) * g.G + lx * LIGHT_DIV + dx;
            gg += (float)g.gamma[c];
            const uint4 v = c4[c];
            m += (v.x & CHEM_STATE_MASK) + (v.x >> 16) + (v.y & CHEM_STATE_MASK) + (v.y >> 16) +
                 (v.z & CHEM_STATE_MASK) + (v.z >> 16) + (v.w & CHEM_STATE_MASK) + (v.w >> 16);
        }
    gg /= (float)LIGHT_DIV;
    const float dens = (float)m / (float)(LIGHT_DIV * LIGHT_DIV);
    const float tau = rb_max(P.monomer_uv_absorb, 0.0f) * dens * (float)LIGHT_DIV / (float)g.G;
    const float shade = 0.0f;
    *t0 = rb_expf(-UV_SHADOW_ABSORB * gg - 0.15f * shade - tau * 1.0f);
    *t1 = rb_expf(-UV_SHADOW_ABSORB * gg - 0.15f * shade - tau * DEEP_LIGHT_DEPTH);
}
// light_transmit_pass (light.wgsl:121-127): the sweep's pre-pass writes each
// light cell's transmission into the light field.
__global__ __launch_bounds__(256) void k_light_transmit(Geo g, RiboSimParams P, const uint32_t* chem, float* light) {
    const uint32_t LS = g.Lsz, LC = LS * LS;
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= LC) return;
    float t0, t1;
    light_transmit2(g, P, chem, i % LS, i / LS, &t0, &t1);
    light[i] = t0;
    light[LC + i] = t1;
}
// compute_uv_light (light.wgsl:129-160): the sequential top-to-bottom sweep.
// The bands never mix, so each band is one workgroup; the finished row above
// lives in LDS (two row buffers, one barrier per row: a row's readers finish
// before the barrier that precedes its buffer's reuse two rows later).  Each
// thread owns columns tid + 1024 c (c < CPT) and loads the next row's
// transmissions while it works on the current row, so a row costs LDS reads
// and a barrier rather than an HBM round trip.  Rows below the current one
// still hold the pre-pass values when they are prefetched.
template <int CPT> __global__ __launch_bounds__(1024) void k_light_sweep(Geo g, RiboSimParams P, const uint32_t* ep, float* light) {
    extern __shared__ float rows[];
    sun_of_step(g, P, ep);
    const uint32_t LS = g.Lsz, LC = LS * LS, band = blockIdx.x, tid = threadIdx.x;
    float* L = light + band * LC;
    const float rw = row_water(P, g.G, band);
    const SunShift s = sun_shift(P);
    float* prev = rows;
    float* cur = rows + LS;
    const uint32_t top = LS - 1u;
    float tr[CPT];
#pragma unroll
    for (int c = 0; c < CPT; ++c) {
        const uint32_t x = tid + 1024u * (uint32_t)c;
        if (x < LS) {
            const float v = L[top * LS + x] * (rw * P.sun_now);
            L[top * LS + x] = v;
            prev[x] = v;
        }
        tr[c] = (x < LS && top >= 1u) ? L[(top - 1u) * LS + x] : 0.0f;
    }
    __syncthreads();
    for (uint32_t t = 1; t < LS; ++t) {
        const uint32_t y = LS - 1u - t;
        float tn[CPT];
#pragma unroll
        for (int c = 0; c < CPT; ++c) {
            const uint32_t x = tid + 1024u * (uint32_t)c;
            tn[c] = (x < LS && y >= 1u) ? L[(y - 1u) * LS + x] : 0.0f;
        }
#pragma unroll
        for (int c = 0; c < CPT; ++c) {
            const uint32_t x = tid + 1024u * (uint32_t)c;
            if (x < LS) {
                const float above = light_above_row(prev, LS, x, s);
                const float v = tr[c] * (above * rw);
                L[y * LS + x] = v;
                cur[x] = v;
            }
        }
        __syncthreads();
        float* tmp = prev;
        prev = cur;
        cur = tmp;
#pragma unroll
        for (int c = 0; c < CPT; ++c) tr[c] = tn[c];
    }
}
// light_propagate (light.wgsl:162-176): every light cell at once takes the
// row above from the previous field; light_commit's copy is a ping-pong of
// the two fields.
__global__ __launch_bounds__(256) void k_light_propagate(Geo g, RiboSimParams P, const uint32_t* ep, const uint32_t* chem,
                                                         const float* lin, float* lout) {
    const uint32_t LS = g.Lsz, LC = LS * LS;
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= LC) return;
    sun_of_step(g, P, ep);
    const uint32_t x = i % LS, y = i / LS;
    float t0, t1;
    light_transmit2(g, P, chem, x, y, &t0, &t1);
    float a0 = P.sun_now, a1 = P.sun_now;
    if (y + 1u < LS) {
        const SunShift s = sun_shift(P);
        a0 = light_above_row(lin + (y + 1u) * LS, LS, x, s);
        a1 = light_above_row(lin + LC + (y + 1u) * LS, LS, x, s);
    }
    lout[i] = a0 * row_water(P, g.G, 0u) * t0;
    lout[LC + i] = a1 * row_water(P, g.G, 1u) * t1;
}

__device__ __forceinline__ float vent_cycle(const RiboSimParams& P, uint32_t x, uint32_t F) {  // fluid.wgsl:46-54
    float m = 1.0f;
    if (P.vent_cycle_amp > 0.0f && P.vent_cycle_period >= 1.0f) {
        uint32_t period = rb_f2u(P.vent_cycle_period);
        float t = (float)(P.epoch % period) / (float)period;
        m = rb_max(1.0f + P.vent_cycle_amp * rb_sinf(6.2831853f * (t + VENT_WAVES * (float)x / (float)F)), 0.0f);
    }
    return m;
}

// ---- fluid step kernels -------------------------------------------------------------
#define FLUID_XY                                                              \
    const uint32_t x = blockIdx.x * 16 + threadIdx.x, y = blockIdx.y * 16 + threadIdx.y; \
    const uint32_t F = g.F;                                                   \
    if (x >= F || y >= F) return;                                             \
    const uint32_t idx = y * F + x;

// update_temperature (fluid.wgsl:300-363); the copy back (copy_temperature,
// :365-372) is a pointer swap on the host, and redox_eaten is cleared here.
__device__ __forceinline__ void b_temperature(const uint3 blockIdx, const uint3 gridDim, Geo g, RiboSimParams P, const uint32_t* ep, const float2* __restrict__ vel,
                                                     const float* __restrict__ T, const float* __restrict__ Q,
                                                     float* __restrict__ t_out, float* __restrict__ r_out,
                                                     uint32_t* eaten, int eaten_active, const uint32_t* chem) {
    FLUID_XY
    sun_of_step(g, P, ep);
    if (ep) P.epoch += *ep;
    const float dt = fluid_dt(P);
    const float2 v = raw_vel(g, vel, idx);
    const float2 src = sub2(f2((float)x + 0.5f, (float)y + 0.5f), mul2(v, dt));
    const float sxf = rb_clamp(src.x - 0.5f, 0.0f, (float)(F - 1u));
    const float syf = rb_clamp(src.y - 0.5f, 0.0f, (float)(F - 1u));
    const uint32_t x0 = rb_f2u(rb_floor(sxf)), y0 = rb_f2u(rb_floor(syf));
    const uint32_t x1 = rb_minu(x0 + 1u, F - 1u), y1 = rb_minu(y0 + 1u, F - 1u);
    const float fx = rb_fract(sxf), fy = rb_fract(syf);
    const uint32_t xl = x == 0u ? 0u : x - 1u, xr = rb_minu(x + 1u, F - 1u);
    const uint32_t yb = y == 0u ? 0u : y - 1u, yt = rb_minu(y + 1u, F - 1u);
    const float fill = g.fill[idx];
    const float2 hs = g.heat[idx];
    const float vc = vent_cycle(P, x, F);
    float t = rb_mix(rb_mix(T[y0 * F + x0], T[y0 * F + x1], fx), rb_mix(T[y1 * F + x0], T[y1 * F + x1], fx), fy);
    const float tl = T[y * F + xl], tr = T[y * F + xr], tb = T[yb * F + x], tt = T[yt * F + x];
    t = rb_mix(t, (tl + tr + tb + tt) * 0.25f, TEMP_DIFFUSE * rb_mix(1.0f, TEMP_ROCK_CONDUCT, fill));
    t += hs.x * vc * TEMP_HEAT_RATE * dt;
    if (P.sun_heat > 0.0f) {
        const uint32_t env = env_cell_for_fluid(g, x, y);
        const float depth = 1.0f - ((float)y + 0.5f) / (float)F;
        const float ir = P.sun_now * SUN_IR * rb_expf(-depth / SUN_IR_DEPTH);
        t += rb_max(P.sun_heat, 0.0f) * SUN_HEAT_RATE * (light_absorbed_at(g, P, chem, env) + ir) * dt;
    }
    const float amb = temp_ambient_at(y, F);
    t = amb + (t - amb) * rb_expf(-TEMP_COOL_RATE * (1.0f - fill) * dt);
    t_out[idx] = rb_clamp(t, 0.0f, TEMP_MAX);

    float r = rb_mix(rb_mix(Q[y0 * F + x0], Q[y0 * F + x1], fx), rb_mix(Q[y1 * F + x0], Q[y1 * F + x1], fx), fy);
    const float rl = Q[y * F + xl], rr = Q[y * F + xr], rbm = Q[yb * F + x], rt = Q[yt * F + x];
    r = rb_mix(r, (rl + rr + rbm + rt) * 0.25f, TEMP_DIFFUSE);
    r += hs.y * vc * REDOX_RATE * dt;
    uint32_t e = 0u;
    if (eaten_active) {
        e = eaten[idx];
        eaten[idx] = 0u;
    }
    r = rb_max(r - (float)e / REDOX_FP, 0.0f);
    r *= rb_expf(-rb_max(P.redox_decay, 0.0f) * dt);
    r_out[idx] = rb_clamp(r, 0.0f, REDOX_MAX);
}
__global__ __launch_bounds__(256) void k_temperature(Geo g, RiboSimParams P, const uint32_t* ep, const float2* __restrict__ vel,
                                                     const float* __restrict__ T, const float* __restrict__ Q,
                                                     float* __restrict__ t_out, float* __restrict__ r_out,
                                                     uint32_t* eaten, int eaten_active, const uint32_t* chem) {
    b_temperature(make_uint3(::blockIdx.x, ::blockIdx.y, ::blockIdx.z), make_uint3(::gridDim.x, ::gridDim.y, ::gridDim.z), g, P, ep, vel, T, Q, t_out, r_out, eaten, eaten_active, chem);
}

// buoyancy (fluid.wgsl:376-389) into the fixed-point force words, then
// gather_forces (:392-398) and clear_force_vectors (:290-296).  A word holds
// the agents' fixed-point forces (agent, when non-null) plus the cell's own
// buoyancy; the sum is an integer sum modulo 2^32, as the reference's atomic
// adds, so the order of the adds does not matter.
__device__ __forceinline__ void b_forces(const uint3 blockIdx, const uint3 gridDim, Geo g, RiboSimParams P, const uint32_t* ep, const float* __restrict__ T,
                                                float2* forces, uint32_t* agent) {
    FLUID_XY
    if (ep) P.epoch += *ep;
    int32_t qx = 0, qy = 0;
    if (!(x == 0u || y == 0u || x >= F - 1u || y >= F - 1u)) {
        const float dev = T[idx] - temp_ambient_at(y, F);
        if (!(rb_abs(dev) <= 1e-3f)) {
            const uint32_t time_bucket = P.epoch / 16u;
            const float ad = rb_abs(dev);
            const float side = (rng_f4(idx, time_bucket, S_BUOYANCY, P.seed).x * 2.0f - 1.0f) * 0.25f * ad / rb_max(ad, 1.0f);
            const float fx = side * TEMP_BUOYANCY * ad;
            const float fy = TEMP_BUOYANCY * dev;
            qx = rb_f2i(rb_rint(rb_clamp(fx, -1.0e5f, 1.0e5f) * FORCE_FP));
            qy = rb_f2i(rb_rint(rb_clamp(fy, -1.0e5f, 1.0e5f) * FORCE_FP));
        }
    }
    if (agent) {
        const uint2 a = reinterpret_cast<uint2*>(agent)[idx];
        qx = (int32_t)((uint32_t)qx + a.x);
        qy = (int32_t)((uint32_t)qy + a.y);
        reinterpret_cast<uint2*>(agent)[idx] = make_uint2(0u, 0u);
    }
    const float2 f = f2((float)qx / FORCE_FP, (float)qy / FORCE_FP);
    forces[idx] = clamp_len(sanitize(f), MAX_FORCE);
}
__global__ __launch_bounds__(256) void k_forces(Geo g, RiboSimParams P, const uint32_t* ep, const float* __restrict__ T,
                                                float2* forces, uint32_t* agent) {
    b_forces(make_uint3(::blockIdx.x, ::blockIdx.y, ::blockIdx.z), make_uint3(::gridDim.x, ::gridDim.y, ::gridDim.z), g, P, ep, T, forces, agent);
}

// smooth_velocity (fluid.wgsl:680-693): the mean of the cell and its four
// neighbours (clamped at the edges), summed in the reference's order.
__global__ __launch_bounds__(256) void k_smooth_velocity(Geo g, const float2* __restrict__ vin, float2* __restrict__ vout) {
    FLUID_XY
    const uint32_t xm = x > 0u ? x - 1u : x, xp = x + 1u < F ? x + 1u : F - 1u;
    const uint32_t ym = y > 0u ? y - 1u : y, yp = y + 1u < F ? y + 1u : F - 1u;
    float2 v = add2(sanitize(vin[idx]), sanitize(vin[y * F + xm]));
    v = add2(v, sanitize(vin[y * F + xp]));
    v = add2(v, sanitize(vin[ym * F + x]));
    v = add2(v, sanitize(vin[yp * F + x]));
    vout[idx] = mul2(v, 0.2f);
}

// slope_steer_velocity (fluid.wgsl:165-178)
__device__ __forceinline__ float2 slope_steer(const Geo& g, const RiboSimParams& P, uint32_t idx, float2 v_in, float dt) {
    const float2 sr = sanitize(g.slope_f[idx]);
    const float2 s = f2(-sr.x, -sr.y);
    const float v_len = len2(v_in);
    const float s_len = len2(s);
    if (v_len < 1e-5f || s_len < 1e-5f) return v_in;
    const float2 v_dir = f2(v_in.x / v_len, v_in.y / v_len);
    const float2 s_dir = f2(s.x / s_len, s.y / s_len);
    const float misalign = rb_clamp(1.0f - dot2(v_dir, s_dir), 0.0f, 2.0f);
    const float t = rb_clamp(1.0f - rb_expf(-rb_max(P.slope_steer_rate, 0.0f) * misalign * dt), 0.0f, 1.0f);
    const float2 dir_raw = add2(v_dir, mul2(sub2(s_dir, v_dir), t));
    const float dir_len = len2(dir_raw);
    if (dir_len < 1e-5f) return v_in;
    return mul2(f2(dir_raw.x / dir_len, dir_raw.y / dir_len), v_len);
}

// add_forces (fluid.wgsl:410-432) with rubble_drag (:150-161)
__device__ __forceinline__ void b_add_forces(const uint3 blockIdx, const uint3 gridDim, Geo g, RiboSimParams P, const float2* __restrict__ forces,
                                                    const float2* __restrict__ vin, float2* __restrict__ vout) {
    FLUID_XY
    if (solid(g, idx)) { vout[idx] = f2(0.0f, 0.0f); return; }
    const float dt = fluid_dt(P);
    const uint32_t xm = x > 0u ? x - 1u : x, xp = x + 1u < F ? x + 1u : x;
    const uint32_t ym = y > 0u ? y - 1u : y, yp = y + 1u < F ? y + 1u : y;
    auto fa = [&](uint32_t a, uint32_t b) { return sanitize(forces[b * F + a]); };
    const float2 f_c = fa(x, y);
    const float2 axis = add2(add2(add2(fa(xm, y), fa(xp, y)), fa(x, ym)), fa(x, yp));
    const float2 diag = add2(add2(add2(fa(xm, ym), fa(xp, ym)), fa(xm, yp)), fa(xp, yp));
    const float2 f_avg = mul2(add2(add2(mul2(axis, 4.0f), diag), mul2(f_c, 20.0f)), (float)(1.0 / 36.0));
    const float2 f_user = clamp_len(mix2(f_c, f_avg, FORCE_SMOOTH_MIX), MAX_FORCE);
    float2 v = sanitize(add2(sanitize(vin[idx]), mul2(f_user, dt)));
    v = slope_steer(g, P, idx, v, dt);
    v = mul2(v, 1.0f / (1.0f + RUBBLE_DRAG * g.dragg[idx] * dt));
    v = reflect(g, idx, v);
    vout[idx] = clamp_len(v, MAX_VEL);
}
__global__ __launch_bounds__(256) void k_add_forces(Geo g, RiboSimParams P, const float2* __restrict__ forces,
                                                    const float2* __restrict__ vin, float2* __restrict__ vout) {
    b_add_forces(make_uint3(::blockIdx.x, ::blockIdx.y, ::blockIdx.z), make_uint3(::gridDim.x, ::gridDim.y, ::gridDim.z), g, P, forces, vin, vout);
}

// diffuse_velocity (fluid.wgsl:435-456)
__device__ __forceinline__ void b_diffuse(const uint3 blockIdx, const uint3 gridDim, Geo g, RiboSimParams P, const float2* __restrict__ vin, float2* __restrict__ vout) {
    FLUID_XY
    if (solid(g, idx)) { vout[idx] = f2(0.0f, 0.0f); return; }
    if (x == 0u || x == F - 1u || y == 0u || y == F - 1u) { vout[idx] = vin[idx]; return; }
    auto rv = [&](uint32_t a, uint32_t b) { return raw_vel(g, vin, b * F + a); };
    const float2 v_c = rv(x, y);
    const float2 axis_sum = add2(add2(add2(rv(x - 1, y), rv(x + 1, y)), rv(x, y - 1)), rv(x, y + 1));
    const float2 diag_sum = add2(add2(add2(rv(x - 1, y - 1), rv(x + 1, y - 1)), rv(x - 1, y + 1)), rv(x + 1, y + 1));
    const float2 lap = mul2(sub2(add2(mul2(axis_sum, 4.0f), diag_sum), mul2(v_c, 20.0f)), (float)(1.0 / 6.0));
    const float a = rb_max(P.fluid_viscosity, 0.0f) * fluid_dt(P);
    float2 v = sanitize(add2(v_c, mul2(lap, a)));
    v = reflect(g, idx, v);
    vout[idx] = clamp_len(v, MAX_VEL);
}
__global__ __launch_bounds__(256) void k_diffuse(Geo g, RiboSimParams P, const float2* __restrict__ vin, float2* __restrict__ vout) {
    b_diffuse(make_uint3(::blockIdx.x, ::blockIdx.y, ::blockIdx.z), make_uint3(::gridDim.x, ::gridDim.y, ::gridDim.z), g, P, vin, vout);
}

// is_solid_at_pos and sample_velocity (fluid.wgsl:226-255)
__device__ __forceinline__ bool is_solid_at_pos(const Geo& g, float2 pos) {
    const int32_t ix = rb_f2i(rb_floor(pos.x - 0.5f)), iy = rb_f2i(rb_floor(pos.y - 0.5f));
    if (ix < 0 || ix >= (int32_t)g.F || iy < 0 || iy >= (int32_t)g.F) return true;
    return solid(g, (uint32_t)iy * g.F + (uint32_t)ix);
}
__device__ __forceinline__ float2 sample_velocity(const Geo& g, const float2* vin, float2 pos) {
    const float min_pos = 0.5f, max_pos = (float)g.F - 0.5f;
    if (pos.x < min_pos || pos.x > max_pos || pos.y < min_pos || pos.y > max_pos) return f2(0.0f, 0.0f);
    const float xx = pos.x - 0.5f, yy = pos.y - 0.5f;
    const int32_t x0 = rb_f2i(rb_floor(xx)), y0 = rb_f2i(rb_floor(yy));
    const float fx = rb_fract(xx), fy = rb_fract(yy);
    const int32_t Fm = (int32_t)g.F - 1;
    const uint32_t cx0 = (uint32_t)rb_clampi(x0, 0, Fm), cx1 = (uint32_t)rb_clampi(x0 + 1, 0, Fm);
    const uint32_t cy0 = (uint32_t)rb_clampi(y0, 0, Fm), cy1 = (uint32_t)rb_clampi(y0 + 1, 0, Fm);
    const float2 v0 = mix2(raw_vel(g, vin, cy0 * g.F + cx0), raw_vel(g, vin, cy0 * g.F + cx1), fx);
    const float2 v1 = mix2(raw_vel(g, vin, cy1 * g.F + cx0), raw_vel(g, vin, cy1 * g.F + cx1), fx);
    return mix2(v0, v1, fy);
}

// advect_velocity (fluid.wgsl:458-477); decay = pow(clamp(fluid_decay,0,1),
// dt*60), uniform over the step, comes from the host.
__device__ __forceinline__ void b_advect(const uint3 blockIdx, const uint3 gridDim, Geo g, RiboSimParams P, float decay, const float2* __restrict__ vin,
                                                float2* __restrict__ vout) {
    FLUID_XY
    if (solid(g, idx)) { vout[idx] = f2(0.0f, 0.0f); return; }
    const float2 pos = f2((float)x + 0.5f, (float)y + 0.5f);
    const float dt = fluid_dt(P);
    const float2 vel = reflect(g, idx, clamp_len(raw_vel(g, vin, idx), MAX_VEL));
    float2 trace = sub2(pos, mul2(vel, dt));
    if (is_solid_at_pos(g, trace)) trace = sub2(pos, mul2(reflect(g, idx, vel), dt));
    float2 out = sanitize(mul2(sample_velocity(g, vin, trace), decay));
    out = reflect(g, idx, out);
    vout[idx] = clamp_len(out, MAX_VEL);
}
__global__ __launch_bounds__(256) void k_advect(Geo g, RiboSimParams P, float decay, const float2* __restrict__ vin,
                                                float2* __restrict__ vout) {
    b_advect(make_uint3(::blockIdx.x, ::blockIdx.y, ::blockIdx.z), make_uint3(::gridDim.x, ::gridDim.y, ::gridDim.z), g, P, decay, vin, vout);
}

// curl_at (fluid.wgsl:479-489)
__device__ __forceinline__ float curl_at(const Geo& g, const float2* vin, uint32_t x, uint32_t y) {
    const int32_t Fm = (int32_t)g.F - 1;
    const uint32_t xm = (uint32_t)rb_clampi((int32_t)x - 1, 0, Fm), xp = (uint32_t)rb_clampi((int32_t)x + 1, 0, Fm);
    const uint32_t ym = (uint32_t)rb_clampi((int32_t)y - 1, 0, Fm), yp = (uint32_t)rb_clampi((int32_t)y + 1, 0, Fm);
    const uint32_t F = g.F;
    const float2 v_l = reflect(g, y * F + xm, raw_vel(g, vin, y * F + xm));
    const float2 v_r = reflect(g, y * F + xp, raw_vel(g, vin, y * F + xp));
    const float2 v_b = reflect(g, ym * F + x, raw_vel(g, vin, ym * F + x));
    const float2 v_t = reflect(g, yp * F + x, raw_vel(g, vin, yp * F + x));
    return 0.5f * (v_r.y - v_l.y) - 0.5f * (v_t.x - v_b.x);
}

// vorticity_confinement (fluid.wgsl:491-518).  Each cell needs the curl at
// itself and its 4 neighbours, and each curl the reflected raw velocity of 4
// cells, so the 16 x 16 tile stages those velocities with a 2-cell halo
// (coordinates clamped into the grid, as curl_at clamps them), then the
// curls with a 1-cell halo, in LDS; the arithmetic is curl_at's.
__device__ __forceinline__ void b_vorticity(const uint3 blockIdx, const uint3 gridDim, Geo g, RiboSimParams P, const float2* __restrict__ vin, float2* __restrict__ vout) {
    __shared__ float2 rv[20][21];
    __shared__ float cl[18][19];
    const uint32_t F = g.F;
    const int32_t Fm = (int32_t)F - 1;
    const int32_t bx = (int32_t)(blockIdx.x * 16), by = (int32_t)(blockIdx.y * 16);
    const uint32_t tid = threadIdx.y * 16 + threadIdx.x;
    for (uint32_t k = tid; k < 20u * 20u; k += 256u) {
        const int32_t lx = (int32_t)(k % 20u), ly = (int32_t)(k / 20u);
        const uint32_t gx = (uint32_t)rb_clampi(bx + lx - 2, 0, Fm), gy = (uint32_t)rb_clampi(by + ly - 2, 0, Fm);
        const uint32_t q = gy * F + gx;
        rv[ly][lx] = reflect(g, q, raw_vel(g, vin, q));
    }
    __syncthreads();
    for (uint32_t k = tid; k < 18u * 18u; k += 256u) {
        const uint32_t lx = k % 18u, ly = k / 18u;  // curl at tile cell (lx - 1, ly - 1)
        const float2 v_l = rv[ly + 1][lx], v_r = rv[ly + 1][lx + 2], v_b = rv[ly][lx + 1], v_t = rv[ly + 2][lx + 1];
        cl[ly][lx] = 0.5f * (v_r.y - v_l.y) - 0.5f * (v_t.x - v_b.x);
    }
    __syncthreads();
    const uint32_t x = blockIdx.x * 16 + threadIdx.x, y = blockIdx.y * 16 + threadIdx.y;
    if (x >= F || y >= F) return;
    const uint32_t idx = y * F + x;
    if (solid(g, idx)) { vout[idx] = f2(0.0f, 0.0f); return; }
    const uint32_t border = 2u;
    const float eps = rb_clamp(P.fluid_vorticity, 0.0f, 10.0f);
    if (x < border || x >= F - border || y < border || y >= F - border || eps < 1e-6f) { vout[idx] = vin[idx]; return; }
    const uint32_t cx = threadIdx.x + 1u, cy = threadIdx.y + 1u;
    const float w = cl[cy][cx];
    if (rb_abs(w) < 5e-4f) { vout[idx] = vin[idx]; return; }
    const float2 grad = mul2(f2(rb_abs(cl[cy][cx + 1]) - rb_abs(cl[cy][cx - 1]), rb_abs(cl[cy + 1][cx]) - rb_abs(cl[cy - 1][cx])),
                             0.5f);
    const float gl = rb_max(len2(grad), 1e-5f);
    const float2 n = f2(grad.x / gl, grad.y / gl);
    const float2 f = mul2(f2(n.y, -n.x), w * eps);
    const float2 dv = clamp_len(mul2(f, fluid_dt(P)), 2.0f);
    float2 v = sanitize(add2(vin[idx], dv));
    v = reflect(g, idx, v);
    vout[idx] = clamp_len(v, MAX_VEL);
}
__global__ __launch_bounds__(256) void k_vorticity(Geo g, RiboSimParams P, const float2* __restrict__ vin, float2* __restrict__ vout) {
    b_vorticity(make_uint3(::blockIdx.x, ::blockIdx.y, ::blockIdx.z), make_uint3(::gridDim.x, ::gridDim.y, ::gridDim.z), g, P, vin, vout);
}

// compute_divergence (fluid.wgsl:525-551).  Also zeroes the pressure of wall
// cells, which is what mg_init (multigrid.wgsl:57) does to the level-0
// multigrid unknowns that alias the pressure field.
__device__ __forceinline__ void b_divergence(const uint3 blockIdx, const uint3 gridDim, Geo g, const float2* __restrict__ vin, float* __restrict__ div, float* p0) {
    FLUID_XY
    if (solid(g, idx)) {
        div[idx] = 0.0f;
        p0[idx] = 0.0f;
        return;
    }
    const float2 v_c = raw_vel(g, vin, idx);
    auto nbv = [&](int32_t nx, int32_t ny, bool axis_x) -> float2 {
        const bool outside = nx < 0 || ny < 0 || nx >= (int32_t)F || ny >= (int32_t)F;
        const uint32_t q = outside ? idx : (uint32_t)ny * F + (uint32_t)nx;
        const bool wall = outside || solid(g, q);
        const float2 v = raw_vel(g, vin, q);
        return wall ? (axis_x ? f2(-v_c.x, v_c.y) : f2(v_c.x, -v_c.y)) : v;
    };
    const int32_t xi = (int32_t)x, yi = (int32_t)y;
    const float2 v_l = nbv(xi - 1, yi, true), v_r = nbv(xi + 1, yi, true);
    const float2 v_b = nbv(xi, yi - 1, false), v_t = nbv(xi, yi + 1, false);
    div[idx] = 0.5f * ((v_r.x - v_l.x) + (v_t.y - v_b.y));
}
__global__ __launch_bounds__(256) void k_divergence(Geo g, const float2* __restrict__ vin, float* __restrict__ div, float* p0) {
    b_divergence(make_uint3(::blockIdx.x, ::blockIdx.y, ::blockIdx.z), make_uint3(::gridDim.x, ::gridDim.y, ::gridDim.z), g, vin, div, p0);
}

// jacobi_pressure (fluid.wgsl:553-608): Neumann by mirroring p_c.
__global__ __launch_bounds__(256) void k_jacobi(Geo g, const float* __restrict__ pin, float* __restrict__ pout,
                                                const float* __restrict__ div) {
    FLUID_XY
    if (solid(g, idx)) { pout[idx] = 0.0f; return; }
    const float p_c = pin[idx];
    auto nb = [&](int32_t a, int32_t b) -> float {
        if (a < 0 || b < 0 || a >= (int32_t)F || b >= (int32_t)F) return p_c;
        const uint32_t j = (uint32_t)b * F + (uint32_t)a;
        return solid(g, j) ? p_c : pin[j];
    };
    const float p_l = nb((int32_t)x - 1, (int32_t)y), p_r = nb((int32_t)x + 1, (int32_t)y);
    const float p_b = nb((int32_t)x, (int32_t)y - 1), p_t = nb((int32_t)x, (int32_t)y + 1);
    const float d = div[idx];
    if (is_bad(d) || is_bad(p_c)) { pout[idx] = 0.0f; return; }
    const float p_new = (p_l + p_r + p_b + p_t - d) * 0.25f;
    pout[idx] = is_bad(p_new) ? 0.0f : p_new;
}

// ---- multigrid, one dispatch per operation (multigrid.wgsl:25-114) ----------------
__device__ __forceinline__ void mg_neighbors(const MgLevelDev& L, uint32_t x, uint32_t y, float* s, float* k) {
    const uint32_t n = L.n;
    float ss = 0.0f, kk = 0.0f;
    if (x > 0u && L.fluid[y * n + x - 1u]) { ss += L.p[y * n + x - 1u]; kk += 1.0f; }
    if (x + 1u < n && L.fluid[y * n + x + 1u]) { ss += L.p[y * n + x + 1u]; kk += 1.0f; }
    if (y > 0u && L.fluid[(y - 1u) * n + x]) { ss += L.p[(y - 1u) * n + x]; kk += 1.0f; }
    if (y + 1u < n && L.fluid[(y + 1u) * n + x]) { ss += L.p[(y + 1u) * n + x]; kk += 1.0f; }
    *s = ss;
    *k = kk;
}
__global__ __launch_bounds__(256) void k_mg_smooth(MgLevelDev L, uint32_t color, const uint32_t* done) {  // :60-77
    if (done && *done) return;
    const uint32_t x = blockIdx.x * 16 + threadIdx.x, y = blockIdx.y * 16 + threadIdx.y;
    if (x >= L.n || y >= L.n) return;
    if (((x + y) & 1u) != color) return;
    const uint32_t i = y * L.n + x;
    if (!L.fluid[i]) return;
    float s, k;
    mg_neighbors(L, x, y, &s, &k);
    L.p[i] = (s - L.rhs[i]) / (k + L.eps);
}
__global__ __launch_bounds__(256) void k_mg_restrict(MgLevelDev L, MgLevelDev C, const uint32_t* done) {  // :82-105
    if (done && *done) return;
    const uint32_t gx = blockIdx.x * 16 + threadIdx.x, gy = blockIdx.y * 16 + threadIdx.y;
    if (gx >= C.n || gy >= C.n) return;
    float sum = 0.0f;
    uint32_t nf = 0;
    for (uint32_t c = 0; c < 4; ++c) {
        const uint32_t fx = gx * 2u + (c & 1u), fy = gy * 2u + (c >> 1);
        const uint32_t i = fy * L.n + fx;
        if (!L.fluid[i]) continue;
        float s, k;
        mg_neighbors(L, fx, fy, &s, &k);
        sum += L.rhs[i] - (s - (k + L.eps) * L.p[i]);
        nf += 1u;
    }
    const uint32_t ci = gy * C.n + gx;
    C.rhs[ci] = nf > 0u ? 2.0f * sum / (float)rb_maxu(nf, 1u) : 0.0f;
    C.p[ci] = 0.0f;
}
__global__ __launch_bounds__(256) void k_mg_prolong(MgLevelDev L, MgLevelDev C, const uint32_t* done) {  // :108-114
    if (done && *done) return;
    const uint32_t x = blockIdx.x * 16 + threadIdx.x, y = blockIdx.y * 16 + threadIdx.y;
    if (x >= L.n || y >= L.n) return;
    const uint32_t i = y * L.n + x;
    if (!L.fluid[i]) return;
    L.p[i] += C.p[(y / 2u) * C.n + x / 2u];
}

// ---- multigrid, fused: one launch per level per direction -------------------
// A red-black half-sweep updates every cell of one colour from the current
// values of its 4 neighbours (the other colour), so the order inside a
// half-sweep is irrelevant and a tile with an h-cell halo can run h
// half-sweeps from its own copy and still hold exact values in its core:
// after half-sweep k every cell at distance >= k from the loaded edge is
// exact.  The tiles below therefore reproduce the per-dispatch schedule of
// the reference bit for bit while reading each level once per direction.
//
// Template parameters: S is the storage type of the level arrays in HBM
// (float, or bf16 for the mixed-precision preconditioner; arithmetic is
// always binary32 in LDS); SYM selects the symmetric V-cycle of the
// tolerance mode (post-smoothing black then red, restriction = half the sum
// over fluid children, exact coarsest solve) instead of the reference's.
// Output tiles are T x T with T even, so tile origins keep (x + y) parity:
// 32 on the large levels, 16 on levels of at most MG_SMALL cells a side,
// where there are few tiles and each block's latency is the whole cost.
constexpr int MG_SMALL = 256;

__device__ __forceinline__ float ldv(const float* a, int i) { return a[i]; }
__device__ __forceinline__ float ldv(const bf16* a, int i) { return rb_as_f32((uint32_t)a[i].bits << 16); }
__device__ __forceinline__ void stv(float* a, int i, float v) { a[i] = v; }
__device__ __forceinline__ void stv(bf16* a, int i, float v) { a[i].bits = bf16_bits(v); }

template <int H, int T>
struct MgTile {
    static constexpr int R = T + 2 * H;
    float p[R][R + 1];
    float rhs[R][R + 1];
    uint8_t fl[R][R + 1];
};

// Sum of the fluid 4-neighbours (left, right, down, up: the reference's
// order) and their count.  All flags and values are loaded before any is
// tested, and a missing neighbour is skipped by a select (not by adding 0,
// which would turn a -0 sum into +0), so the result equals the branchy
// form bit for bit while the loads go out together instead of one dependent
// LDS round trip per neighbour.
struct NbSum { float s, k; };
__device__ __forceinline__ NbSum nb_sum(bool fl, bool fr, bool fd, bool fu, float vl, float vr, float vd, float vu) {
    float s = 0.0f, k = 0.0f;
    s = fl ? s + vl : s;
    k = fl ? k + 1.0f : k;
    s = fr ? s + vr : s;
    k = fr ? k + 1.0f : k;
    s = fd ? s + vd : s;
    k = fd ? k + 1.0f : k;
    s = fu ? s + vu : s;
    k = fu ? k + 1.0f : k;
    return {s, k};
}

// Neighbour sum of a tile cell (the tile's halo covers every neighbour read).
template <int H, int T>
__device__ __forceinline__ NbSum tile_nb(const MgTile<H, T>& t, int ly, int lx) {
    return nb_sum(t.fl[ly][lx - 1] != 0, t.fl[ly][lx + 1] != 0, t.fl[ly - 1][lx] != 0, t.fl[ly + 1][lx] != 0,
                  t.p[ly][lx - 1], t.p[ly][lx + 1], t.p[ly - 1][lx], t.p[ly + 1][lx]);
}

// p_c <- (sum of fluid neighbours - rhs) / (k + eps), on the cells of one
// colour inside [lo, hi) of the tile (multigrid.wgsl:35-45,60-67); solid
// cells keep their value.
template <int H, int T>
__device__ __forceinline__ void tile_half_sweep(MgTile<H, T>& t, int color, float eps, int lo, int hi) {
    const int w = hi - lo;
    const int half = (w + 1) / 2;
    for (int k = threadIdx.x; k < w * half; k += blockDim.x) {
        const int ly = lo + k / half;
        int lx = lo + 2 * (k % half);
        if (((lx + ly) & 1) != color) ++lx;
        if (lx >= hi) continue;
        const bool f = t.fl[ly][lx] != 0;
        const float pc = t.p[ly][lx], rh = t.rhs[ly][lx];
        const NbSum nb = tile_nb(t, ly, lx);
        t.p[ly][lx] = f ? (nb.s - rh) / (nb.k + eps) : pc;
    }
}

// Loads the tile with its halo (unknowns from psrc, or 0 when psrc is
// null).  Every thread issues all of its loads
// before the first LDS store (one memory round trip per tile, not one per
// loop trip); values of solid or outside cells are replaced by 0 after the
// load, so the fluid mask is not on the load's critical path.
template <int H, int T, class S, int NT = 256>
__device__ __forceinline__ void tile_load(MgTile<H, T>& t, const MgLevelT<S>& L, const S* psrc, int x0, int y0,
                                          const MgLevelT<S>* C) {
    constexpr int R = MgTile<H, T>::R;
    constexpr int PER = (R * R + NT - 1) / NT;
    const int n = (int)L.n;
    float pv[PER], rv[PER], cv[PER];
    uint32_t fv[PER];
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        const int k = (int)threadIdx.x + NT * j;
        const int ly = k / R, lx = k % R;
        const int gx = x0 + lx, gy = y0 + ly;
        pv[j] = 0.0f;
        rv[j] = 0.0f;
        cv[j] = 0.0f;
        fv[j] = 0u;
        if (k < R * R && gx >= 0 && gy >= 0 && gx < n && gy < n) {
            const int i = gy * n + gx;
            fv[j] = L.fluid[i];
            pv[j] = psrc ? ldv(psrc, i) : 0.0f;
            rv[j] = ldv(L.rhs, i);
            if (C) cv[j] = ldv(C->p, (gy / 2) * (int)C->n + gx / 2);
        }
    }
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        const int k = (int)threadIdx.x + NT * j;
        if (k < R * R) {
            const int ly = k / R, lx = k % R;
            const bool f = fv[j] != 0u;
            // mg_prolong (multigrid.wgsl:108-114), pointwise, applied on load
            t.p[ly][lx] = f ? (C ? pv[j] + cv[j] : pv[j]) : 0.0f;
            t.rhs[ly][lx] = f ? rv[j] : 0.0f;
            t.fl[ly][lx] = (uint8_t)fv[j];
        }
    }
}

// Bilinear (cell-centred) transfer of the tolerance mode between a level
// of even side n and its coarse level (n / 2).  In each direction fine
// index f takes its parent a0 = f / 2 with weight 3/4 and the coarse cell
// on its side, a1 (clamped at the edge, where it repeats a0), with weight
// 1/4.  Prolongation adds the four products (a0y a0x, a0y a1x, a1y a0x,
// a1y a1x), over fluid coarse cells, divided by the sum W of their weights
// (W >= 9/16: the parent of a fluid cell is fluid).  Restriction is its
// transpose: a fluid coarse cell gathers wt(f, c) res_f / W_f over its 4 x 4
// fine cells 2c - 1 .. 2c + 2, rows then columns, with
// wt(f, c) = 3/4 [a0(f) = c] + 1/4 [a1(f) = c].
// The binary32 preconditioner uses it; the bfloat16 one keeps the
// piecewise-constant transfer (prolongation by injection, restriction half
// the sum over fluid children): rounding the bilinear V-cycle's output to
// bfloat16 stalls conjugate gradients.
template <class S> constexpr bool BL_TRANSFER = std::is_same<S, float>::value;
__device__ __forceinline__ void bl_nf(int f, int nc, int& a0, int& a1) {
    a0 = f >> 1;
    a1 = (f & 1) ? (a0 + 1 < nc ? a0 + 1 : a0) : (a0 > 0 ? a0 - 1 : 0);
}
// wt(2c - 1 + d, c) for d = 0..3 (wt(f, c) = 3/4 [a0(f) = c] +
// 1/4 [a1(f) = c]), 0 outside the level: 1/4, 3/4 (1 at the
// low edge, where the clamped far neighbour is c too), 3/4 (1 at the high
// edge), 1/4.
__device__ __forceinline__ void bl_w4(int c, int nc, float (&w)[4]) {
    w[0] = c > 0 ? 0.25f : 0.0f;
    w[1] = c == 0 ? 1.0f : 0.75f;
    w[2] = c == nc - 1 ? 1.0f : 0.75f;
    w[3] = c < nc - 1 ? 0.25f : 0.0f;
}
// W and the weighted sum of the coarse values at fine cell (fx, fy):
// cf(cx, cy) and cv(cx, cy) read a coarse flag and value.
template <class CF, class CV>
__device__ __forceinline__ void bl_interp(int fx, int fy, int nc, CF cf, CV cv, float& ws, float& acc) {
    int x0, x1, y0, y1;
    bl_nf(fx, nc, x0, x1);
    bl_nf(fy, nc, y0, y1);
    const bool f0 = cf(x0, y0), f1 = cf(x1, y0), f2 = cf(x0, y1), f3 = cf(x1, y1);
    const float v0 = cv(x0, y0), v1 = cv(x1, y0), v2 = cv(x0, y1), v3 = cv(x1, y1);
    acc = 0.0f;
    ws = 0.0f;
    acc = acc + (f0 ? 0.5625f * v0 : 0.0f);
    ws = ws + (f0 ? 0.5625f : 0.0f);
    acc = acc + (f1 ? 0.1875f * v1 : 0.0f);
    ws = ws + (f1 ? 0.1875f : 0.0f);
    acc = acc + (f2 ? 0.1875f * v2 : 0.0f);
    ws = ws + (f2 ? 0.1875f : 0.0f);
    acc = acc + (f3 ? 0.0625f * v3 : 0.0f);
    ws = ws + (f3 ? 0.0625f : 0.0f);
}
// The restriction half of a tile kernel: rn (TT + 2)^2 holds res / W of
// the fine cells one ring around the T x T core at fine origin fx0 - 1 (0
// outside the grid or off the fluid); cfl the coarse flags (T/2 + 2)^2 at
// coarse origin cx0 - 1.  rn is filled from the tile (whose values must be
// exact two cells beyond the core), then each coarse cell of the core
// gathers its 16 fine cells.
template <int H, int T, class S>
__device__ __forceinline__ void bl_restrict_tile(const MgTile<H, T>& t, float* rn, const uint8_t* cfl, int n, float eps,
                                                 const MgLevelT<S>& C, int x0, int y0) {
    constexpr int RN = T + 2, CF = T / 2 + 2;
    const int nc = (int)C.n;
    const int fx0 = x0 + H - 1, fy0 = y0 + H - 1;
    const int cx0 = (x0 + H) / 2 - 1, cy0 = (y0 + H) / 2 - 1;
    for (int k = threadIdx.x; k < RN * RN; k += blockDim.x) {
        const int ly = H - 1 + k / RN, lx = H - 1 + k % RN;
        const int gx = fx0 + k % RN, gy = fy0 + k / RN;
        float v = 0.0f;
        if (gx >= 0 && gy >= 0 && gx < n && gy < n && t.fl[ly][lx]) {
            const NbSum nb = tile_nb(t, ly, lx);
            const float res = t.rhs[ly][lx] - (nb.s - (nb.k + eps) * t.p[ly][lx]);
            float ws, acc;
            bl_interp(
                gx, gy, nc, [&](int cx, int cy) { return cfl[(cy - cy0) * CF + (cx - cx0)] != 0; },
                [](int, int) { return 0.0f; }, ws, acc);
            v = res / ws;
        }
        rn[k] = v;
    }
    __syncthreads();
    const int ct = threadIdx.x;  // one thread per coarse cell of the tile
    if (ct < (T / 2) * (T / 2)) {
        const int cx = (x0 + H) / 2 + ct % (T / 2), cy = (y0 + H) / 2 + ct / (T / 2);
        if (cx < nc && cy < nc) {
            float wy[4], wx[4];
            bl_w4(cy, nc, wy);
            bl_w4(cx, nc, wx);
            const float* r0 = rn + (2 * cy - 1 - fy0) * RN + (2 * cx - 1 - fx0);
            float sum = 0.0f;
#pragma unroll
            for (int dy = 0; dy < 4; ++dy)
#pragma unroll
                for (int dx = 0; dx < 4; ++dx) sum = sum + (wy[dy] * wx[dx]) * r0[dy * RN + dx];
            const int ci = cy * nc + cx;
            stv(C.rhs, ci, cfl[(cy - cy0) * CF + (cx - cx0)] ? sum : 0.0f);
        }
    }
}
// Coarse flags of the restriction: (T/2 + 2)^2 from coarse origin
// (x0 + H) / 2 - 1.
template <int H, int T, class S>
__device__ __forceinline__ void bl_load_cfl(uint8_t* cfl, const MgLevelT<S>& C, int x0, int y0) {
    constexpr int CF = T / 2 + 2;
    const int nc = (int)C.n;
    const int cx0 = (x0 + H) / 2 - 1, cy0 = (y0 + H) / 2 - 1;
    for (int k = threadIdx.x; k < CF * CF; k += blockDim.x) {
        const int cx = cx0 + k % CF, cy = cy0 + k / CF;
        cfl[k] = (cx >= 0 && cy >= 0 && cx < nc && cy < nc) ? C.fluid[cy * nc + cx] : (uint8_t)0;
    }
}
// Coarse values and flags of the prolongation onto a tile with halo H <= 5:
// (T/2 + 6)^2 from coarse origin (x0 + H) / 2 - 3.
template <int H, int T, class S>
__device__ __forceinline__ void bl_load_coarse(float* cpv, uint8_t* cfl, const MgLevelT<S>& C, int x0, int y0) {
    static_assert(H <= 5, "the coarse tile covers a fine halo of at most 5");
    constexpr int CU = T / 2 + 6;
    const int nc = (int)C.n;
    const int cx0 = (x0 + H) / 2 - 3, cy0 = (y0 + H) / 2 - 3;
    for (int k = threadIdx.x; k < CU * CU; k += blockDim.x) {
        const int cx = cx0 + k % CU, cy = cy0 + k / CU;
        const bool in = cx >= 0 && cy >= 0 && cx < nc && cy < nc;
        cfl[k] = in ? C.fluid[cy * nc + cx] : (uint8_t)0;
        cpv[k] = in ? ldv(C.p, cy * nc + cx) : 0.0f;
    }
}
// Adds the bilinear prolongation to every fluid cell of the tile.
template <int H, int T>
__device__ __forceinline__ void bl_prolong_tile(MgTile<H, T>& t, const float* cpv, const uint8_t* cfl, int n, int nc, int x0,
                                                int y0) {
    constexpr int R = MgTile<H, T>::R, CU = T / 2 + 6;
    const int cx0 = (x0 + H) / 2 - 3, cy0 = (y0 + H) / 2 - 3;
    for (int k = threadIdx.x; k < R * R; k += blockDim.x) {
        const int ly = k / R, lx = k % R;
        const int gx = x0 + lx, gy = y0 + ly;
        if (gx < 0 || gy < 0 || gx >= n || gy >= n || !t.fl[ly][lx]) continue;
        float ws, acc;
        bl_interp(
            gx, gy, nc, [&](int cx, int cy) { return cfl[(cy - cy0) * CU + (cx - cx0)] != 0; },
            [&](int cx, int cy) { return cpv[(cy - cy0) * CU + (cx - cx0)]; }, ws, acc);
        t.p[ly][lx] = t.p[ly][lx] + acc / ws;
    }
}

// Pre-smoothing (2 x red, black) and restriction of one level
// (src/world/mod.rs:1556-1563; multigrid.wgsl:60-105).
template <class S, bool SYM, int MG_T>
__device__ __forceinline__ void b_mg_down(const uint3 blockIdx, const uint3 gridDim, MgLevelT<S> L, MgLevelT<S> C, const uint32_t* done) {
    if (done && *done) return;
    // the bilinear restriction reads residuals one ring beyond the core,
    // so its tiles keep exact values two cells out
    constexpr bool BL = SYM && BL_TRANSFER<S>;
    constexpr int H = BL ? 6 : 5;
    __shared__ MgTile<H, MG_T> t;
    __shared__ float rn[BL ? (MG_T + 2) * (MG_T + 2) : 1];
    __shared__ uint8_t cfl[BL ? (MG_T / 2 + 2) * (MG_T / 2 + 2) : 1];
    const int x0 = (int)blockIdx.x * MG_T - H, y0 = (int)blockIdx.y * MG_T - H;
    if (BL) bl_load_cfl<H, MG_T, S>(cfl, C, x0, y0);
    // the symmetric cycle's coarse levels always start from 0 (nothing of
    // the previous cycle survives), so their unknowns are not read
    tile_load<H, MG_T, S>(t, L, SYM ? nullptr : L.p, x0, y0, nullptr);
    __syncthreads();
    constexpr int R = MgTile<H, MG_T>::R;
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        tile_half_sweep<H, MG_T>(t, s & 1, L.eps, 1 + s, R - 1 - s);
        __syncthreads();
    }
    const int n = (int)L.n;
    for (int k = threadIdx.x; k < MG_T * MG_T; k += blockDim.x) {
        const int gx = x0 + H + k % MG_T, gy = y0 + H + k / MG_T;
        if (gx < n && gy < n && t.fl[H + k / MG_T][H + k % MG_T]) stv(L.pp, gy * n + gx, t.p[H + k / MG_T][H + k % MG_T]);
    }
    if (BL) {
        bl_restrict_tile<H, MG_T, S>(t, rn, cfl, n, L.eps, C, x0, y0);
        return;
    }
    const int nc = (int)C.n;
    const int ct = threadIdx.x;  // one thread per coarse cell of the tile
    if (ct < (MG_T / 2) * (MG_T / 2)) {
        const int cx = (int)blockIdx.x * (MG_T / 2) + ct % (MG_T / 2), cy = (int)blockIdx.y * (MG_T / 2) + ct / (MG_T / 2);
        if (cx < nc && cy < nc) {
            float sum = 0.0f;
            uint32_t nf = 0;
            for (int c = 0; c < 4; ++c) {
                const int lx = 2 * (ct % (MG_T / 2)) + (c & 1) + H, ly = 2 * (ct / (MG_T / 2)) + (c >> 1) + H;
                const bool f = t.fl[ly][lx] != 0;
                const float pc = t.p[ly][lx], rh = t.rhs[ly][lx];
                const NbSum nb = tile_nb(t, ly, lx);
                sum = f ? sum + (rh - (nb.s - (nb.k + L.eps) * pc)) : sum;
                nf += f ? 1u : 0u;
            }
            const int ci = cy * nc + cx;
            const float rc = nf > 0u ? (SYM ? 0.5f * sum : 2.0f * sum / (float)rb_maxu(nf, 1u)) : 0.0f;
            stv(C.rhs, ci, rc);
            if (!SYM) stv(C.p, ci, 0.0f);
        }
    }
}
template <class S, bool SYM, int MG_T>
__global__ __launch_bounds__(256) void k_mg_down(MgLevelT<S> L, MgLevelT<S> C, const uint32_t* done) {
    b_mg_down<S, SYM, MG_T>(make_uint3(::blockIdx.x, ::blockIdx.y, ::blockIdx.z), make_uint3(::gridDim.x, ::gridDim.y, ::gridDim.z), L, C, done);
}

// Prolongation and post-smoothing (2 x red, black; black, red when SYM) of
// one level (src/world/mod.rs:1569-1575).
template <class S, bool SYM, int MG_T>
__device__ __forceinline__ void b_mg_up(const uint3 blockIdx, const uint3 gridDim, MgLevelT<S> L, MgLevelT<S> C, const uint32_t* done) {
    if (done && *done) return;
    constexpr int H = 4;
    __shared__ MgTile<H, MG_T> t;
    constexpr bool BL = SYM && BL_TRANSFER<S>;
    __shared__ float cpv[BL ? (MG_T / 2 + 6) * (MG_T / 2 + 6) : 1];
    __shared__ uint8_t cfl[BL ? (MG_T / 2 + 6) * (MG_T / 2 + 6) : 1];
    const int x0 = (int)blockIdx.x * MG_T - H, y0 = (int)blockIdx.y * MG_T - H;
    if (BL) bl_load_coarse<H, MG_T, S>(cpv, cfl, C, x0, y0);
    tile_load<H, MG_T, S>(t, L, L.pp, x0, y0, BL ? nullptr : &C);
    __syncthreads();
    if (BL) {
        bl_prolong_tile<H, MG_T>(t, cpv, cfl, (int)L.n, (int)C.n, x0, y0);
        __syncthreads();
    }
    constexpr int R = MgTile<H, MG_T>::R;
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        tile_half_sweep<H, MG_T>(t, SYM ? 1 - (s & 1) : (s & 1), L.eps, 1 + s, R - 1 - s);
        __syncthreads();
    }
    const int n = (int)L.n;
    for (int k = threadIdx.x; k < MG_T * MG_T; k += blockDim.x) {
        const int gx = x0 + H + k % MG_T, gy = y0 + H + k / MG_T;
        if (gx < n && gy < n && t.fl[H + k / MG_T][H + k % MG_T]) stv(L.p, gy * n + gx, t.p[H + k / MG_T][H + k % MG_T]);
    }
}
template <class S, bool SYM, int MG_T>
__global__ __launch_bounds__(256) void k_mg_up(MgLevelT<S> L, MgLevelT<S> C, const uint32_t* done) {
    b_mg_up<S, SYM, MG_T>(make_uint3(::blockIdx.x, ::blockIdx.y, ::blockIdx.z), make_uint3(::gridDim.x, ::gridDim.y, ::gridDim.z), L, C, done);
}

// The coarse end of the V-cycle in one workgroup: every level of `lv` (at
// most MG_BLOCK_LEVELS, small enough that all of them fit in LDS together)
// goes through pre-smoothing and restriction, the coarsest gets 30 x (red,
// black) -- or, in the tolerance mode, an exact solve with the binary64
// inverse of its operator -- then prolongation and post-smoothing back up;
// the top level's unknowns are read from (the symmetric cycle starts them
// from 0) and written back to HBM.

// The workgroup is 32 x 32 threads; level loops stride rows by 32 and, in a
// half-sweep, cells of one colour by 64 along the row, so no index needs a
// division and a 64 x 64 level is 2 cells per thread per half-sweep.
// Neighbour sum of cell (x, y) of an n x n level held whole in LDS; indices
// of neighbours outside the level are clamped to the cell (their flag is
// false, so the value is not used).
__device__ __forceinline__ NbSum blk_nb(const float* p, const uint8_t* fl, int n, int x, int y) {
    const int i = y * n + x;
    const int il = x > 0 ? i - 1 : i, ir = x + 1 < n ? i + 1 : i;
    const int id = y > 0 ? i - n : i, iu = y + 1 < n ? i + n : i;
    const bool gl = fl[il] != 0, gr = fl[ir] != 0, gd = fl[id] != 0, gu = fl[iu] != 0;
    const float vl = p[il], vr = p[ir], vd = p[id], vu = p[iu];
    return nb_sum((x > 0) & gl, (x + 1 < n) & gr, (y > 0) & gd, (y + 1 < n) & gu, vl, vr, vd, vu);
}

// The threads of k_mg_block as a 32 x 32 team: level loops stride rows by
// 32 and, in a half-sweep, cells of one colour by 64 along the row, so no
// index needs a division.
template <int TW, int TH>
struct Team {
    int tx, ty;
    __device__ explicit Team(int t) : tx(t % TW), ty(t / TW) {}
    __device__ void sync() const { __syncthreads(); }
};

// One half-sweep of an n x n level held whole in LDS: rows stride by TH,
// cells of the colour by 2 TW along the row.
template <int TW, int TH>
__device__ __forceinline__ void blk_sweep(const Team<TW, TH>& tm, float* p, const float* rhs, const uint8_t* fl, int n,
                                          float eps, int color) {
    for (int y = tm.ty; y < n; y += TH) {
        for (int x = ((color ^ y) & 1) + 2 * tm.tx; x < n; x += 2 * TW) {
            const int i = y * n + x;
            const bool f = fl[i] != 0;
            const float pc = p[i], rh = rhs[i];
            const NbSum nb = blk_nb(p, fl, n, x, y);
            p[i] = f ? (nb.s - rh) / (nb.k + eps) : pc;
        }
    }
}

// 2 x (red, black) and the restriction of level l into level l + 1
// (bilinear when SYM, through res / W of every fine cell in RN).
template <bool SYM, int TW, int TH, class S>
__device__ void blk_down(const Team<TW, TH>& tm, const MgBlockArgs<S>& a, float* P, float* RHS, const uint8_t* FL, float* RN,
                         int l) {
    float* p = P + a.off[l];
    const float* rhs = RHS + a.off[l];
    const uint8_t* fl = FL + a.off[l];
    const int n = a.n[l];
    for (int s = 0; s < 4; ++s) {
        blk_sweep(tm, p, rhs, fl, n, a.eps[l], s & 1);
        tm.sync();
    }
    const int nc = a.n[l + 1];
    float* pc = P + a.off[l + 1];
    float* rc = RHS + a.off[l + 1];
    if (SYM && BL_TRANSFER<S>) {
        const uint8_t* flc = FL + a.off[l + 1];
        for (int y = tm.ty; y < n; y += TH)
            for (int x = tm.tx; x < n; x += TW) {
                const int i = y * n + x;
                float v = 0.0f;
                if (fl[i]) {
                    const NbSum nb = blk_nb(p, fl, n, x, y);
                    const float res = rhs[i] - (nb.s - (nb.k + a.eps[l]) * p[i]);
                    float ws, acc;
                    bl_interp(
                        x, y, nc, [&](int cx, int cy) { return flc[cy * nc + cx] != 0; }, [](int, int) { return 0.0f; }, ws,
                        acc);
                    v = res / ws;
                }
                RN[i] = v;
            }
        tm.sync();
        for (int gy = tm.ty; gy < nc; gy += TH)
            for (int gx = tm.tx; gx < nc; gx += TW) {
                float wy[4], wx[4];
                bl_w4(gy, nc, wy);
                bl_w4(gx, nc, wx);
                float sum = 0.0f;
#pragma unroll
                for (int dy = 0; dy < 4; ++dy) {
                    const int fy = 2 * gy - 1 + dy;
                    const int ry = fy < 0 ? 0 : (fy >= n ? n - 1 : fy);  // clamped rows carry weight 0
#pragma unroll
                    for (int dx = 0; dx < 4; ++dx) {
                        const int fx = 2 * gx - 1 + dx;
                        const int rx = fx < 0 ? 0 : (fx >= n ? n - 1 : fx);
                        sum = sum + (wy[dy] * wx[dx]) * RN[ry * n + rx];
                    }
                }
                const int ci = gy * nc + gx;
                rc[ci] = flc[ci] ? sum : 0.0f;
                pc[ci] = 0.0f;
            }
        tm.sync();
        return;
    }
    for (int gy = tm.ty; gy < nc; gy += TH)
        for (int gx = tm.tx; gx < nc; gx += TW) {
            float sum = 0.0f;
            uint32_t nf = 0;
            for (int c = 0; c < 4; ++c) {
                const int fx = gx * 2 + (c & 1), fy = gy * 2 + (c >> 1);
                const int i = fy * n + fx;
                const bool f = fl[i] != 0;
                const float pcv = p[i], rh = rhs[i];
                const NbSum nb = blk_nb(p, fl, n, fx, fy);
                sum = f ? sum + (rh - (nb.s - (nb.k + a.eps[l]) * pcv)) : sum;
                nf += f ? 1u : 0u;
            }
            const int ci = gy * nc + gx;
            rc[ci] = nf > 0u ? (SYM ? 0.5f * sum : 2.0f * sum / (float)rb_maxu(nf, 1u)) : 0.0f;
            pc[ci] = 0.0f;
        }
    tm.sync();
}

// Prolongation from level l + 1 and 2 x (red, black) -- (black, red) when
// SYM -- on level l.
template <bool SYM, int TW, int TH, class S>
__device__ void blk_up(const Team<TW, TH>& tm, const MgBlockArgs<S>& a, float* P, const float* RHS, const uint8_t* FL, int l) {
    float* p = P + a.off[l];
    const float* rhs = RHS + a.off[l];
    const uint8_t* fl = FL + a.off[l];
    const int n = a.n[l], nc = a.n[l + 1];
    const float* pc = P + a.off[l + 1];
    const uint8_t* flc = FL + a.off[l + 1];
    for (int y = tm.ty; y < n; y += TH)
        for (int x = tm.tx; x < n; x += TW) {
            const int i = y * n + x;
            if (!fl[i]) continue;
            if (SYM && BL_TRANSFER<S>) {
                float ws, acc;
                bl_interp(
                    x, y, nc, [&](int cx, int cy) { return flc[cy * nc + cx] != 0; },
                    [&](int cx, int cy) { return pc[cy * nc + cx]; }, ws, acc);
                p[i] = p[i] + acc / ws;
            } else {
                p[i] += pc[(y / 2) * nc + x / 2];
            }
        }
    tm.sync();
    for (int s = 0; s < 4; ++s) {
        blk_sweep(tm, p, rhs, fl, n, a.eps[l], SYM ? 1 - (s & 1) : (s & 1));
        tm.sync();
    }
}

// The coarsest level: 30 x (red, black) (then (black, red) when SYM), or in
// the tolerance mode the exact solve with the binary64 inverse.  Lanes of
// the team share the rows of the inverse.  With at most MG_CINV_REG
// unknowns, each of the first 4 MG_CINV_REG lanes holds its quarter row in
// registers, loaded when the kernel starts (ci).
constexpr int MG_CINV_REG = 64;
template <bool SYM, int TW, int TH, class S>
__device__ void blk_coarse(const Team<TW, TH>& tm, const MgBlockArgs<S>& a, float* P, const float* RHS, const uint8_t* FL,
                           float* bv, const double (&ci)[MG_CINV_REG / 4], const int cm) {
    const int L = a.levels;
    float* p = P + a.off[L - 1];
    const float* rhs = RHS + a.off[L - 1];
    const uint8_t* fl = FL + a.off[L - 1];
    const int t = tm.ty * TW + tm.tx;
    if (SYM && cm > 0) {
        // exact solve: p = A^-1 rhs over the fluid cells, binary64.  The
        // right-hand side is staged compactly; four lanes share a row (the
        // inverse is stored by columns, so a row's quarter reads words that
        // neighbouring rows read too) and add their partial sums in a fixed
        // order.
        const int m = cm;
        for (int j = t; j < m; j += TW * TH) bv[j] = rhs[a.ccell[j]];
        tm.sync();
        const int q = t & 3;
        for (int r0 = t >> 2; r0 < ((m + 15) & ~15); r0 += (TW * TH) / 4) {
            double acc = 0.0;
            if (m <= MG_CINV_REG) {
#pragma unroll
                for (int k = 0; k < MG_CINV_REG / 4; ++k)
                    if (r0 < m && q + 4 * k < m) acc += ci[k] * (double)bv[q + 4 * k];
            } else if (r0 < m) {
                for (int j = q; j < m; j += 4) acc += a.cinv[(size_t)j * m + r0] * (double)bv[j];
            }
            acc += __shfl_down(acc, 2, 4);
            acc += __shfl_down(acc, 1, 4);
            if (q == 0 && r0 < m) p[a.ccell[r0]] = (float)acc;
        }
        tm.sync();
    } else if (a.n[L - 1] * a.n[L - 1] <= 64) {
        // at most 64 cells: one wave, one cell per lane, neighbours exchanged
        // through cross-lane permutes
        if (t < 64) {
            const int nn = a.n[L - 1], lane = t;
            const bool live = lane < nn * nn;
            const int cx = live ? lane % nn : 0, cy = live ? lane / nn : 0;
            const bool me = live && fl[lane];
            const bool fl_l = live && cx > 0 && fl[lane - 1];
            const bool fl_r = live && cx + 1 < nn && fl[lane + 1];
            const bool fl_b = live && cy > 0 && fl[lane - nn];
            const bool fl_t = live && cy + 1 < nn && fl[lane + nn];
            const float rv = live ? rhs[lane] : 0.0f;
            const float eps = a.eps[L - 1];
            const int color = (cx + cy) & 1;
            float pv = live ? p[lane] : 0.0f;
            for (int s = 0; s < 2 * 30; ++s) {
                const int want = (SYM && s >= 30) ? 1 - (s & 1) : (s & 1);
                const float pl = __shfl(pv, lane - 1, 64), pr = __shfl(pv, lane + 1, 64);
                const float pb = __shfl(pv, lane - nn, 64), pt = __shfl(pv, lane + nn, 64);
                if (me && color == want) {
                    const NbSum nb = nb_sum(fl_l, fl_r, fl_b, fl_t, pl, pr, pb, pt);
                    pv = (nb.s - rv) / (nb.k + eps);
                }
            }
            if (live) p[lane] = pv;
        }
        tm.sync();
    } else {
        for (int s = 0; s < 2 * 30; ++s) {
            const int color = (SYM && s >= 30) ? 1 - (s & 1) : (s & 1);
            blk_sweep(tm, p, rhs, fl, a.n[L - 1], a.eps[L - 1], color);
            tm.sync();
        }
    }
}

// TH: rows of the 32-wide team (the workgroup has 32 TH threads).
template <class S, bool SYM, int TH = 32>
__device__ __forceinline__ void b_mg_block(const uint3 blockIdx, const uint3 gridDim, const MgBlockArgs<S>& a, const uint32_t* done,
                                           float* lds = nullptr) {
    if (done && *done) return;
    extern __shared__ float dyn_sm[];
    float* const sm = lds ? lds : dyn_sm;  // the launch's dynamic LDS, or the caller's
    const Team<32, TH> blk((int)threadIdx.x);
    const int total = a.off[a.levels];
    float* P = sm;
    float* RHS = sm + total;
    uint8_t* FL = reinterpret_cast<uint8_t*>(sm + 2 * total);
    float* BV = reinterpret_cast<float*>(reinterpret_cast<uint8_t*>(sm) + ((9 * total + 3) & ~3));
    float* RN = BV + MG_SYM_COARSE_MAX;  // res / W of a fine level (SYM), n0^2 floats
    const int n0 = a.n[0];
    for (int y = blk.ty; y < n0; y += TH)
        for (int x = blk.tx; x < n0; x += 32) {
            const int i = y * n0 + x;
            const uint8_t f = a.fluid[0][i];
            const float pv = SYM ? 0.0f : ldv(a.p0, i), rv = ldv(a.rhs0, i);
            FL[i] = f;
            P[i] = f ? pv : 0.0f;
            RHS[i] = f ? rv : 0.0f;
        }
    for (int l = 1; l < a.levels; ++l) {
        const int n = a.n[l];
        for (int y = blk.ty; y < n; y += TH)
            for (int x = blk.tx; x < n; x += 32) FL[a.off[l] + y * n + x] = a.fluid[l][y * n + x];
    }
    // the coarsest level's unknowns: from the device when the inverse is
    // rebuilt there; this lane's quarter row of its inverse (see blk_coarse)
    const int cm = (SYM && a.cm_dev) ? (int)__builtin_amdgcn_readfirstlane((uint32_t)*a.cm_dev) : a.cm;
    double ci[MG_CINV_REG / 4];
    {
        const int m = cm, t = (int)threadIdx.x, q = t & 3, r0 = t >> 2;
#pragma unroll
        for (int k = 0; k < MG_CINV_REG / 4; ++k)
            ci[k] = (SYM && m > 0 && m <= MG_CINV_REG && r0 < m && q + 4 * k < m) ? a.cinv[(size_t)(q + 4 * k) * m + r0] : 0.0;
    }
    __syncthreads();
    const int L = a.levels;
    for (int l = 0; l + 1 < L; ++l) blk_down<SYM>(blk, a, P, RHS, FL, RN, l);
    blk_coarse<SYM>(blk, a, P, RHS, FL, BV, ci, cm);
    for (int l = L - 2; l >= 0; --l) blk_up<SYM>(blk, a, P, RHS, FL, l);
    for (int y = blk.ty; y < n0; y += TH)
        for (int x = blk.tx; x < n0; x += 32) {
            const int i = y * n0 + x;
            if (FL[i]) stv(a.p0, i, P[i]);
        }
}
template <class S, bool SYM>
__global__ __launch_bounds__(1024) void k_mg_block(MgBlockArgs<S> a, const uint32_t* done) {
    b_mg_block<S, SYM>(make_uint3(::blockIdx.x, ::blockIdx.y, ::blockIdx.z), make_uint3(::gridDim.x, ::gridDim.y, ::gridDim.z), a, done);
}

// ---- tolerance mode: preconditioned conjugate gradients ------------------------
// All inner products of an iteration are formed at one point, after the
// preconditioner (Chronopoulos and Gear, 1989): with u = M r (the symmetric
// V-cycle) and w = A u,
//   gamma = (r, u), delta = (w, u), e = (u, s), f = (p, w), rr = (r, r)
// for the current p and s; then beta = gamma / gamma_prev,
// pi = delta + beta (e + f) + beta^2 pi_prev (the expansion of (p', s')),
// alpha = gamma / pi, and
//   p' = u + beta p,  s' = w + beta s,  x += alpha p',  r -= alpha s'.
// The vector updates ride on the V-cycle's first kernel, w = A u and the
// inner products on its last, and the scalar step on the next V-cycle's
// first kernel (every workgroup sums the partials in the same order and
// forms the same scalars): an iteration is exactly the V-cycle's launches.
// Level-0 vectors cover fluid cells; inner products are binary64, summed
// per block and then over the blocks in a fixed order, so every run gives
// the same bits.  A v = sum of fluid neighbours - (k + eps) v.
__device__ __forceinline__ float apply_A(const uint8_t* fl, const float* v, int n, int x, int y, float eps) {
    const int i = y * n + x;
    const int il = x > 0 ? i - 1 : i, ir = x + 1 < n ? i + 1 : i;
    const int id = y > 0 ? i - n : i, iu = y + 1 < n ? i + n : i;
    const bool gl = fl[il] != 0, gr = fl[ir] != 0, gd = fl[id] != 0, gu = fl[iu] != 0;
    const float vl = v[il], vr = v[ir], vd = v[id], vu = v[iu], vc = v[i];
    const NbSum nb = nb_sum((x > 0) & gl, (x + 1 < n) & gr, (y > 0) & gd, (y + 1 < n) & gu, vl, vr, vd, vu);
    return nb.s - (nb.k + eps) * vc;
}

// Binary64 forms of the operator.  Values are widened and summed in the
// order of nb_sum; v = hi + lo per cell, formed in binary64 (lo may be null:
// v = hi + 0).  The tolerance mode evaluates w = A u and every residual
// b - A x this way and rounds once: in binary32 the stencil's cancellation on
// smooth vectors (|A v| ~ lambda |v|, lambda down to ~1e-5 at 1024^2) leaves
// an error of ~3 eps32 / lambda on the result, which the recursively updated
// residual then carries.
struct NbSum64 { double s, k; };
__device__ __forceinline__ NbSum64 nb_sum64(bool fl, bool fr, bool fd, bool fu, double vl, double vr, double vd, double vu) {
    double s = 0.0, k = 0.0;
    s = fl ? s + vl : s;
    k = fl ? k + 1.0 : k;
    s = fr ? s + vr : s;
    k = fr ? k + 1.0 : k;
    s = fd ? s + vd : s;
    k = fd ? k + 1.0 : k;
    s = fu ? s + vu : s;
    k = fu ? k + 1.0 : k;
    return {s, k};
}
__device__ __forceinline__ double dval(const float* hi, const float* lo, int j) {
    return (double)hi[j] + (lo ? (double)lo[j] : 0.0);
}
__device__ __forceinline__ double apply_A64(const uint8_t* fl, const float* hi, const float* lo, int n, int x, int y,
                                            float eps) {
    const int i = y * n + x;
    const int il = x > 0 ? i - 1 : i, ir = x + 1 < n ? i + 1 : i;
    const int id = y > 0 ? i - n : i, iu = y + 1 < n ? i + n : i;
    const bool gl = fl[il] != 0, gr = fl[ir] != 0, gd = fl[id] != 0, gu = fl[iu] != 0;
    const double vl = dval(hi, lo, il), vr = dval(hi, lo, ir), vd = dval(hi, lo, id), vu = dval(hi, lo, iu);
    const double vc = dval(hi, lo, i);
    const NbSum64 nb = nb_sum64((x > 0) & gl, (x + 1 < n) & gr, (y > 0) & gd, (y + 1 < n) & gu, vl, vr, vd, vu);
    return nb.s - (nb.k + (double)eps) * vc;
}

// x = hi + lo += t: the exact two-sum of hi and t, its error added to lo,
// then renormalised (|lo| <= ulp(hi) / 2), so that hi stays the nearest
// binary32 value of the double-float iterate.
__device__ __forceinline__ void df_add(float& hi, float& lo, float t) {
    const float s = hi + t;
    const float bv = s - hi;
    const float e = (hi - (s - bv)) + (t - bv);
    const float l = lo + e;
    const float h = s + l;
    lo = l - (h - s);
    hi = h;
}

// Value after a store to S and a load back.
__device__ __forceinline__ float round_s(float v, const float*) { return v; }
__device__ __forceinline__ float round_s(float v, const bf16*) { return rb_as_f32((uint32_t)bf16_bits(v) << 16); }

// Per-block sums of N values into part[N * block + j]: wave shuffles, then
// the block's waves in order.  The totals are formed in a later launch
// (k_pcg_fin), so no block signals or waits on another and the sums do not
// depend on block scheduling.
template <int N>
__device__ __forceinline__ void block_partial(const double (&v)[N], double* part, uint32_t blk) {
    __shared__ double sh[N][16];
    const int tid = threadIdx.y * blockDim.x + threadIdx.x;
    const int lane = tid & 63, wid = tid >> 6, nw = (blockDim.x * blockDim.y) >> 6;
    double a[N];
#pragma unroll
    for (int j = 0; j < N; ++j) {
        a[j] = v[j];
        for (int o = 32; o > 0; o >>= 1) a[j] += __shfl_down(a[j], o, 64);
    }
    if (lane == 0)
#pragma unroll
        for (int j = 0; j < N; ++j) sh[j][wid] = a[j];
    __syncthreads();
    if (tid < N) {
        double t = 0.0;
        for (int w = 0; w < nw; ++w) t += sh[tid][w];
        part[(size_t)N * blk + tid] = t;
    }
}

// Totals of nb partial N-tuples in a fixed order (thread t sums blocks t,
// t + 256, ..., then the wave tree, then the waves in order); valid in
// thread 0.
template <int N>
__device__ void sum_partials(const double* __restrict__ part, uint32_t nb, double (&t)[N]) {
    __shared__ double sh[N][4];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    double a[N];
#pragma unroll
    for (int j = 0; j < N; ++j) a[j] = 0.0;
    for (uint32_t k = (uint32_t)tid; k < nb; k += 256u)
#pragma unroll
        for (int j = 0; j < N; ++j) a[j] += part[(size_t)N * k + j];
#pragma unroll
    for (int j = 0; j < N; ++j) {
        for (int o = 32; o > 0; o >>= 1) a[j] += __shfl_down(a[j], o, 64);
        if (lane == 0) sh[j][wid] = a[j];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < N; ++j) t[j] = ((sh[j][0] + sh[j][1]) + sh[j][2]) + sh[j][3];
}

// Totals of nb partial N-tuples by a whole workgroup of NT threads, in a
// fixed order (thread t sums blocks t, t + NT, ..., then the wave tree, then
// the waves in order); every workgroup that calls it gets the same bits.
// Valid in thread 0.
template <int N, int NT>
__device__ void sum_partials_wg(const double* __restrict__ part, uint32_t nb, double (&t)[N]) {
    __shared__ double sh[N][NT / 64];
    const int tid = threadIdx.y * blockDim.x + threadIdx.x, lane = tid & 63, wid = tid >> 6;
    double a[N];
#pragma unroll
    for (int j = 0; j < N; ++j) a[j] = 0.0;
    for (uint32_t k = (uint32_t)tid; k < nb; k += NT)
#pragma unroll
        for (int j = 0; j < N; ++j) a[j] += part[(size_t)N * k + j];
#pragma unroll
    for (int j = 0; j < N; ++j) {
        for (int o = 32; o > 0; o >>= 1) a[j] += __shfl_down(a[j], o, 64);
        if (lane == 0) sh[j][wid] = a[j];
    }
    __syncthreads();
    if (tid == 0)
#pragma unroll
        for (int j = 0; j < N; ++j) {
            double v = 0.0;
            for (int w = 0; w < NT / 64; ++w) v += sh[j][w];
            t[j] = v;
        }
}

// The scalar recurrences after an iteration's inner products t = (gamma,
// delta, e, f, rr, tt) (see above; tt = ||b - A x||^2 in binary64, formed
// when a.near): state a to b.  The first application after a (re)start only
// sets gamma, pi = delta and alpha.  Later ones count the iteration and
// stop when the true residual reaches tol (or at maxit); when the recursive
// residual has reached tol but the true one has not, the next V-cycle
// restarts from the true residual (one step of iterative refinement).  A
// finished state passes unchanged.
__device__ inline PcgScal pcg_advance(const PcgState* st, const PcgScal& a, const double (&t)[6], float tol,
                                      uint32_t maxit) {
    PcgScal b = a;
    b.restart = 0u;
    if (a.done) return b;
    if (a.fresh) {
        b.gamma = t[0];
        b.pi = t[1];
        b.alpha = t[1] != 0.0 ? t[0] / t[1] : 0.0;
        b.beta = 0.0;
        b.fresh = 0u;
        return b;
    }
    b.it = a.it + 1u;
    b.rel = st->bb > 0.0 ? sqrt(t[4] / st->bb) : 0.0;
    const bool have_true = a.near != 0u;
    const double trel = st->bb > 0.0 ? sqrt(t[5] / st->bb) : 0.0;
    if ((have_true && trel <= (double)tol) || b.it >= maxit) {
        b.done = 1u;
        return b;
    }
    if (have_true && b.rel <= (double)tol) {
        b.restart = 1u;
        b.fresh = 1u;
        b.restarts = a.restarts + 1u;
    } else {
        const double g = t[0], beta = a.gamma != 0.0 ? g / a.gamma : 0.0;
        const double pi = t[1] + beta * (t[2] + t[3]) + beta * beta * a.pi;
        b.beta = beta;
        b.pi = pi;
        b.alpha = pi != 0.0 ? g / pi : 0.0;
        b.gamma = g;
    }
    b.near = b.rel <= PCG_NEAR * (double)tol ? 1u : 0u;
    return b;
}

// A workgroup of NT threads applies the partials of the last V-cycle to
// slot rd, shares the result through LDS and (workgroup 0) writes it to
// slot 1 - rd and publishes done and near.
template <int NT>
__device__ PcgScal pcg_step_wg(PcgState* st, int rd, uint32_t nb, float tol, uint32_t maxit, bool writer) {
    __shared__ PcgScal sh_sc;
    double t[6];
    sum_partials_wg<6, NT>(st->part, nb, t);
    const int tid = threadIdx.y * blockDim.x + threadIdx.x;
    if (tid == 0) {
        const PcgScal b = pcg_advance(st, st->sc[rd], t, tol, maxit);
        sh_sc = b;
        if (writer) {
            st->sc[1 - rd] = b;
            st->done = b.done;
            st->near = b.near;
        }
    }
    __syncthreads();
    return sh_sc;
}

#define PCG_XY                                                                         \
    const int x = (int)(blockIdx.x * 16 + threadIdx.x), y = (int)(blockIdx.y * 16 + threadIdx.y); \
    const int n = (int)V.n;                                                            \
    const bool in = x < n && y < n;                                                    \
    const int i = in ? y * n + x : 0;                                                  \
    const bool f = in && V.fl[i];

// A (x - xold) at fluid cell (x, y) in binary64: the operator on the change
// between the last two solves' pressures, each difference formed in
// binary32.
__device__ __forceinline__ double apply_A64_diff(const uint8_t* fl, const float* v, const float* o, int n, int x, int y,
                                                 float eps) {
    const int i = y * n + x;
    const int il = x > 0 ? i - 1 : i, ir = x + 1 < n ? i + 1 : i;
    const int id = y > 0 ? i - n : i, iu = y + 1 < n ? i + n : i;
    const bool gl = fl[il] != 0, gr = fl[ir] != 0, gd = fl[id] != 0, gu = fl[iu] != 0;
    const double dl = (double)(v[il] - o[il]) + 0.0, dr = (double)(v[ir] - o[ir]) + 0.0;
    const double dd = (double)(v[id] - o[id]) + 0.0, du = (double)(v[iu] - o[iu]) + 0.0;
    const double dc = (double)(v[i] - o[i]) + 0.0;
    const NbSum64 nb = nb_sum64((x > 0) & gl, (x + 1 < n) & gr, (y > 0) & gd, (y + 1 < n) & gu, dl, dr, dd, du);
    return nb.s - (nb.k + (double)eps) * dc;
}

// r = b - A x (r[0]); r[1], s, w start at 0; ||b||^2 and ||r||^2.  Once the
// pressure and xold are the last two solves' results, also q = A (x - xold)
// (into the search direction p, which k_pcg_shift clears) and (r, q), (q, q):
// the start x + theta (x - xold) with theta = (r, q) / (q, q) has the least
// residual on that line (theta = 1 is the linear extrapolation in time).
template <class S>
__device__ __forceinline__ void b_cg_init(const uint3 blockIdx, const uint3 gridDim, const CgVec<S>& V, PcgState* st) {
    PCG_XY
    const bool extrap = V.xold != nullptr && st->hist >= 2u;
    double bb = 0.0, rr = 0.0, rq = 0.0, qq = 0.0;
    if (in) {
        float r = 0.0f, q = 0.0f;
        if (f) {
            r = (float)((double)V.b[i] - apply_A64(V.fl, V.x, nullptr, n, x, y, V.eps));
            if (extrap) q = (float)apply_A64_diff(V.fl, V.x, V.xold, n, x, y, V.eps);
            bb = (double)V.b[i] * (double)V.b[i];
            rr = (double)r * (double)r;
            rq = (double)r * (double)q;
            qq = (double)q * (double)q;
        }
        V.r[0][i] = r;
        V.r[1][i] = 0.0f;
        V.s[0][i] = 0.0f;
        V.s[1][i] = 0.0f;
        V.pd[i] = q;
        V.w[i] = 0.0f;
        V.xlo[i] = 0.0f;
    }
    block_partial<4>({bb, rr, rq, qq}, st->part, blockIdx.y * gridDim.x + blockIdx.x);
}
template <class S>
__global__ __launch_bounds__(256) void k_cg_init(CgVec<S> V, PcgState* st) {
    b_cg_init<S>(make_uint3(::blockIdx.x, ::blockIdx.y, ::blockIdx.z), make_uint3(::gridDim.x, ::gridDim.y, ::gridDim.z), V, st);
}

// The start on the line: x += theta (x - xold), r -= theta q in fluid cells;
// xold takes the old x; the search direction (q until now) starts at 0.
__device__ __forceinline__ void b_pcg_shift(const uint3 blockIdx, uint32_t cells, const uint8_t* __restrict__ fl, float* x,
                                            float* xlo, float* xold, float* r, float* pd, const PcgState* st) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= cells) return;
    const float th = (float)st->theta;
    const float a = x[i];
    if (th != 0.0f && fl[i]) {
        float h = a, l = 0.0f;
        df_add(h, l, th * (a - xold[i]));
        x[i] = h;
        xlo[i] = l;
        r[i] = r[i] - th * pd[i];
    }
    xold[i] = a;
    pd[i] = 0.0f;
}
__global__ __launch_bounds__(256) void k_pcg_shift(uint32_t cells, const uint8_t* __restrict__ fl, float* x, float* xlo,
                                                   float* xold, float* r, float* pd, const PcgState* st) {
    b_pcg_shift(make_uint3(::blockIdx.x, ::blockIdx.y, ::blockIdx.z), cells, fl, x, xlo, xold, r, pd, st);
}

// Level-0 tiles of the tolerance mode: 64 x 64 cells per 1024-thread
// workgroup (one workgroup per CU at 1024^2), so that the 5- and 6-cell
// halos add 20-40 % to the core's work instead of 50-90 % at 32 x 32.
constexpr int CG_T = 64, CG_NT = 1024;
static_assert((CG_T / 2) * (CG_T / 2) <= CG_NT, "one thread per coarse cell of a tile");

// Level 0 of the V-cycle, downward, with the iteration's vector updates:
// every tile cell (halo included) forms s' = w + beta s and r' = r - alpha s'
// from the old r and s, so the tile's right-hand side is the updated
// residual; core cells also store r', s' and the updated p and x.  Then the
// reference's 2 x (red, black) pre-smoothing from u = 0 and the bilinear
// restriction (or, for bfloat16 storage, half the sum over fluid children),
// as k_mg_down.
template <class S>
__device__ __forceinline__ void b_cg_down0(const uint3 blockIdx, const uint3 gridDim, const CgVec<S>& V,
                                           const MgLevelT<S>& C, PcgState* st, int par, int stage, int rd, uint32_t nb,
                                           float tol, uint32_t maxit) {
    // stage 0: the first V-cycle (the initial state in slot rd); later
    // stages first apply the previous V-cycle's inner products to the
    // scalars.  A restart (the recursive residual reached tol, the true one
    // did not) takes the true residual b - A x in binary64 as r, with p and s
    // at 0 and x unchanged, like the first V-cycle.
    const int first = stage == 0 ? 1 : 0;
    if (st->sc[rd].done) {
        if (!first && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) st->sc[1 - rd] = st->sc[rd];
        return;
    }
    const PcgScal sc = first ? st->sc[rd] : pcg_step_wg<CG_NT>(st, rd, nb, tol, maxit, blockIdx.x == 0 && blockIdx.y == 0);
    if (sc.done) return;
    const bool rst = !first && sc.restart != 0u;
    const bool upd = !first && !rst;
    constexpr bool BL = BL_TRANSFER<S>;
    constexpr int H = BL ? 6 : 5, T = CG_T, NT = CG_NT;
    constexpr int R = T + 2 * H;
    constexpr int PER = (R * R + NT - 1) / NT;
    __shared__ MgTile<H, T> t;
    __shared__ float rn[BL ? (T + 2) * (T + 2) : 1];
    __shared__ uint8_t cfl[BL ? (T / 2 + 2) * (T / 2 + 2) : 1];
    const float alpha = (float)sc.alpha, beta = (float)sc.beta;
    const float* r_in = V.r[par];
    const float* s_in = V.s[par];
    float* r_out = V.r[1 - par];
    float* s_out = V.s[1 - par];
    const int n = (int)V.n;
    const int x0 = (int)blockIdx.x * T - H, y0 = (int)blockIdx.y * T - H;
    if (BL) bl_load_cfl<H, T, S>(cfl, C, x0, y0);
    float rv[PER], sv[PER], wv[PER], uv[PER], pv[PER], xv[PER], lv[PER];
    uint32_t fv[PER];
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        const int k = (int)threadIdx.x + NT * j;
        const int ly = k / R, lx = k % R;
        const int gx = x0 + lx, gy = y0 + ly;
        rv[j] = sv[j] = wv[j] = uv[j] = pv[j] = xv[j] = lv[j] = 0.0f;
        fv[j] = 0u;
        if (k < R * R && gx >= 0 && gy >= 0 && gx < n && gy < n) {
            const int i = gy * n + gx;
            fv[j] = V.fl[i];
            if (rst) rv[j] = fv[j] ? (float)((double)V.b[i] - apply_A64(V.fl, V.x, V.xlo, n, gx, gy, V.eps)) : 0.0f;
            else rv[j] = r_in[i];
            if (upd) {
                sv[j] = s_in[i];
                wv[j] = V.w[i];
                if (lx >= H && lx < H + T && ly >= H && ly < H + T) {
                    uv[j] = ldv(V.u, i);
                    pv[j] = V.pd[i];
                    xv[j] = V.x[i];
                    lv[j] = V.xlo[i];
                }
            }
        }
    }
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        const int k = (int)threadIdx.x + NT * j;
        if (k >= R * R) continue;
        const int ly = k / R, lx = k % R;
        const bool f = fv[j] != 0u;
        float s_new = 0.0f, r_new = rv[j];
        if (upd) {
            s_new = wv[j] + beta * sv[j];
            r_new = rv[j] - alpha * s_new;
        }
        t.p[ly][lx] = 0.0f;
        t.rhs[ly][lx] = f ? round_s(r_new, V.u) : 0.0f;
        t.fl[ly][lx] = (uint8_t)fv[j];
        const int gx = x0 + lx, gy = y0 + ly;
        if (f && lx >= H && lx < H + T && ly >= H && ly < H + T && gx < n && gy < n) {
            const int i = gy * n + gx;
            r_out[i] = r_new;
            s_out[i] = s_new;
            if (V.rs) stv(V.rs, i, r_new);
            if (upd) {
                const float p_new = uv[j] + beta * pv[j];
                V.pd[i] = p_new;
                float h = xv[j], l = lv[j];
                df_add(h, l, alpha * p_new);
                V.x[i] = h;
                V.xlo[i] = l;
            } else if (rst) {
                V.pd[i] = 0.0f;
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        tile_half_sweep<H, T>(t, s & 1, V.eps, 1 + s, R - 1 - s);
        __syncthreads();
    }
    for (int k = threadIdx.x; k < T * T; k += blockDim.x) {
        const int gx = x0 + H + k % T, gy = y0 + H + k / T;
        if (gx < n && gy < n && t.fl[H + k / T][H + k % T]) stv(V.zpre, gy * n + gx, t.p[H + k / T][H + k % T]);
    }
    if (BL) {
        bl_restrict_tile<H, T, S>(t, rn, cfl, n, V.eps, C, x0, y0);
        return;
    }
    const int nc = (int)C.n;
    const int ct = threadIdx.x;  // one thread per coarse cell of the tile
    if (ct < (T / 2) * (T / 2)) {
        const int cx = (int)blockIdx.x * (T / 2) + ct % (T / 2), cy = (int)blockIdx.y * (T / 2) + ct / (T / 2);
        if (cx < nc && cy < nc) {
            float sum = 0.0f;
            uint32_t nf = 0;
            for (int c = 0; c < 4; ++c) {
                const int lx = 2 * (ct % (T / 2)) + (c & 1) + H, ly = 2 * (ct / (T / 2)) + (c >> 1) + H;
                const bool f = t.fl[ly][lx] != 0;
                const float pc = t.p[ly][lx], rh = t.rhs[ly][lx];
                const NbSum nb = tile_nb(t, ly, lx);
                sum = f ? sum + (rh - (nb.s - (nb.k + V.eps) * pc)) : sum;
                nf += f ? 1u : 0u;
            }
            const int ci = cy * nc + cx;
            stv(C.rhs, ci, nf > 0u ? 0.5f * sum : 0.0f);
        }
    }
}
template <class S>
__global__ __launch_bounds__(CG_NT) void k_cg_down0(CgVec<S> V, MgLevelT<S> C, PcgState* st, int par, int stage, int rd,
                                                   uint32_t nb, float tol, uint32_t maxit) {
    b_cg_down0<S>(make_uint3(::blockIdx.x, ::blockIdx.y, ::blockIdx.z), make_uint3(::gridDim.x, ::gridDim.y, ::gridDim.z), V, C,
                  st, par, stage, rd, nb, tol, maxit);
}

// Level 0 of the V-cycle, upward: the bilinear prolongation added to the
// pre-smoothed u written by k_cg_down0, and the symmetric
// post-smoothing (black, red) x 2, as k_mg_up but with a 5-cell halo, so the
// smoothed u is exact one ring beyond the core; then w = A u on the core
// (from u as stored in S) and the block's partial sums of (r, u), (w, u),
// (u, s), (p, w) and (r, r).
template <class S>
__device__ __forceinline__ void b_cg_up0(const uint3 blockIdx, const uint3 gridDim, const CgVec<S>& V, const MgLevelT<S>& C,
                                         PcgState* st, int par) {
    if (st->done) return;
    const bool tr = st->near != 0u;  // also form ||b - A x||^2 (binary64)
    constexpr int H = 5, T = CG_T, NT = CG_NT;
    constexpr int R = T + 2 * H;
    __shared__ MgTile<H, T> t;
    constexpr bool BL = BL_TRANSFER<S>;
    __shared__ float cpv[BL ? (T / 2 + 6) * (T / 2 + 6) : 1];
    __shared__ uint8_t cfl[BL ? (T / 2 + 6) * (T / 2 + 6) : 1];
    const float* r_cur = V.r[1 - par];
    MgLevelT<S> L;
    L.n = V.n;
    L.eps = V.eps;
    L.p = V.u;
    L.pp = V.zpre;
    L.rhs = V.rs ? V.rs : reinterpret_cast<S*>(const_cast<float*>(r_cur));
    L.fluid = V.fl;
    const int n = (int)V.n;
    const int x0 = (int)blockIdx.x * T - H, y0 = (int)blockIdx.y * T - H;
    if (BL) bl_load_coarse<H, T, S>(cpv, cfl, C, x0, y0);
    tile_load<H, T, S, NT>(t, L, V.zpre, x0, y0, BL ? nullptr : &C);
    // the inner products' operands at this thread's core cells, loaded now so
    // that they arrive during the sweeps
    const float* s_cur = V.s[1 - par];
    constexpr int PC = (T * T) / NT;
    float sc[PC], pc[PC], rc[PC];
#pragma unroll
    for (int j = 0; j < PC; ++j) {
        const int k = (int)threadIdx.x + NT * j;
        const int gx = x0 + H + k % T, gy = y0 + H + k / T;
        sc[j] = pc[j] = rc[j] = 0.0f;
        if (gx < n && gy < n) {
            const int i = gy * n + gx;
            sc[j] = s_cur[i];
            pc[j] = V.pd[i];
            if (V.rs) rc[j] = r_cur[i];
        }
    }
    __syncthreads();
    if (BL) {
        bl_prolong_tile<H, T>(t, cpv, cfl, n, (int)C.n, x0, y0);
        __syncthreads();
    }
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        tile_half_sweep<H, T>(t, 1 - (s & 1), V.eps, 1 + s, R - 1 - s);
        __syncthreads();
    }
    for (int k = threadIdx.x; k < (T + 2) * (T + 2); k += blockDim.x) {
        const int ly = H - 1 + k / (T + 2), lx = H - 1 + k % (T + 2);
        t.p[ly][lx] = round_s(t.p[ly][lx], V.u);
    }
    __syncthreads();
    double gam = 0.0, del = 0.0, e = 0.0, f = 0.0, rr = 0.0, tt = 0.0;
#pragma unroll
    for (int j = 0; j < PC; ++j) {
        const int k = (int)threadIdx.x + NT * j;
        const int ly = H + k / T, lx = H + k % T;
        const int gx = x0 + lx, gy = y0 + ly;
        if (gx >= n || gy >= n || !t.fl[ly][lx]) continue;
        const int i = gy * n + gx;
        const float u = t.p[ly][lx];
        const NbSum64 nb = nb_sum64(t.fl[ly][lx - 1] != 0, t.fl[ly][lx + 1] != 0, t.fl[ly - 1][lx] != 0,
                                    t.fl[ly + 1][lx] != 0, (double)t.p[ly][lx - 1] + 0.0, (double)t.p[ly][lx + 1] + 0.0,
                                    (double)t.p[ly - 1][lx] + 0.0, (double)t.p[ly + 1][lx] + 0.0);
        const float wv = (float)(nb.s - (nb.k + (double)V.eps) * ((double)u + 0.0));
        stv(V.u, i, u);
        V.w[i] = wv;
        const float r = V.rs ? rc[j] : t.rhs[ly][lx];
        gam += (double)r * (double)u;
        del += (double)wv * (double)u;
        e += (double)u * (double)sc[j];
        f += (double)pc[j] * (double)wv;
        rr += (double)r * (double)r;
        if (tr) {
            const double rt = (double)V.b[i] - apply_A64(V.fl, V.x, V.xlo, n, gx, gy, V.eps);
            tt += rt * rt;
        }
    }
    block_partial<6>({gam, del, e, f, rr, tt}, st->part, blockIdx.y * gridDim.x + blockIdx.x);
}
template <class S>
__global__ __launch_bounds__(CG_NT) void k_cg_up0(CgVec<S> V, MgLevelT<S> C, PcgState* st, int par) {
    b_cg_up0<S>(make_uint3(::blockIdx.x, ::blockIdx.y, ::blockIdx.z), make_uint3(::gridDim.x, ::gridDim.y, ::gridDim.z), V, C, st,
                par);
}

// ||b - A x||^2 over fluid cells, for the reported residuals.
template <class S>
__device__ __forceinline__ void b_cg_final(const uint3 blockIdx, const uint3 gridDim, const CgVec<S>& V, PcgState* st, int rd,
                                           uint32_t nb, float tol, uint32_t maxit, uint32_t off) {
    // workgroup 0 applies the last V-cycle's inner products (slot rd to 1 - rd)
    if (blockIdx.x == 0 && blockIdx.y == 0) (void)pcg_step_wg<256>(st, rd, nb, tol, maxit, true);
    PCG_XY
    // the refined iterate's residual and the stored pressure's, binary64
    double rr = 0.0, rh = 0.0;
    if (f) {
        const double r = (double)V.b[i] - apply_A64(V.fl, V.x, V.xlo, n, x, y, V.eps);
        const double q = (double)V.b[i] - apply_A64(V.fl, V.x, nullptr, n, x, y, V.eps);
        rr = r * r;
        rh = q * q;
    }
    block_partial<2>({rr, rh}, st->part + off, blockIdx.y * gridDim.x + blockIdx.x);  // after the partials read above
}
template <class S>
__global__ __launch_bounds__(256) void k_cg_final(CgVec<S> V, PcgState* st, int rd, uint32_t nb, float tol, uint32_t maxit,
                                                  uint32_t off) {
    b_cg_final<S>(make_uint3(::blockIdx.x, ::blockIdx.y, ::blockIdx.z), make_uint3(::gridDim.x, ::gridDim.y, ::gridDim.z), V, st,
                  rd, nb, tol, maxit, off);
}

// The solve's scalar state from the partials of the kernel before: one
// block, after k_cg_init (the initial state, into slot 1) and after
// k_cg_final (the statistics of the solve whose last state is in slot rd).
enum PcgFin { FIN_INIT, FIN_FINAL };
template <int W>
__device__ __forceinline__ void b_pcg_fin(PcgState* st, uint32_t nb, float tol, uint32_t maxit, int rd, uint32_t off) {
    constexpr int N = W == FIN_INIT ? 4 : 2;
    double t[N];
    sum_partials<N>(st->part + off, nb, t);
    if (threadIdx.x != 0) return;
    if (W == FIN_INIT) {
        st->bb = t[0];
        // the start of least residual on the line through the last two
        // pressures (theta = 0 without that history); its residual norm
        // follows from the sums: |r - theta q|^2 = |r|^2 - (r, q)^2 / (q, q)
        const double theta = t[3] > 0.0 ? t[2] / t[3] : 0.0;
        st->theta = theta;
        const double rr0 = fmax(t[1] - theta * t[2], 0.0);
        const double rel = t[0] > 0.0 ? sqrt(rr0 / t[0]) : 0.0;
        PcgScal a{};
        a.rel = rel;
        a.done = (rel <= (double)tol || maxit == 0u) ? 1u : 0u;
        a.fresh = 1u;
        a.near = rel <= PCG_NEAR * (double)tol ? 1u : 0u;
        st->sc[1] = a;
        st->rel = rel;
        st->it = 0u;
        st->done = a.done;
        st->near = a.near;
        st->tol_seen = tol;
    } else {
        const PcgScal a = st->sc[rd];
        st->it = a.it;
        st->rel = a.rel;
        st->done = a.done;
        st->true_rel = st->bb > 0.0 ? sqrt(t[0] / st->bb) : 0.0;
        st->rounded_rel = st->bb > 0.0 ? sqrt(t[1] / st->bb) : 0.0;
        st->max_true_rel = fmax(st->max_true_rel, st->true_rel);
        st->max_rounded_rel = fmax(st->max_rounded_rel, st->rounded_rel);
        st->restarts += a.restarts;
        st->hist = st->hist < 2u ? st->hist + 1u : 2u;  // solves done since the history restarted (at most 2)
        st->steps += 1u;
        st->iters += a.it;
        // a solve that stopped on its iteration slots before converging
        if (!a.done) st->missed += 1u;
        if (st->host_out) {
            st->host_out[0] = a.it;
            st->host_out[1] = st->missed;
            st->host_out[2] = st->steps;
            st->host_out[PCG_HOST_RING + (st->steps - 1u) % PCG_HOST_RING_LEN] = a.it;
        }
    }
}
template <int W>
__global__ __launch_bounds__(256) void k_pcg_fin(PcgState* st, uint32_t nb, float tol, uint32_t maxit, int rd, uint32_t off) {
    b_pcg_fin<W>(st, nb, tol, maxit, rd, off);
}

// Stationary tolerance mode: the reference V-cycle repeated until the
// relative residual reaches the tolerance.  k_vc_resid leaves per-block
// partials of ||b||^2 and ||b - A x||^2 over level-0 fluid cells (at the
// start of a solve always, after a V-cycle only while the solve is not
// done); k_vc_fin (one block) sums them in a fixed order:
// VC_BEGIN starts the solve, VC_CHECK counts a V-cycle and publishes `done`
// (residual reached or iteration cap), VC_END records the solve's statistics
// as the conjugate-gradient mode does.
__global__ __launch_bounds__(256) void k_vc_resid(uint32_t n, float eps, const uint8_t* __restrict__ fl,
                                                  const float* __restrict__ xp, const float* __restrict__ b,
                                                  PcgState* st, int begin) {
    if (!begin && st->done) return;
    const int x = (int)(blockIdx.x * 16 + threadIdx.x), y = (int)(blockIdx.y * 16 + threadIdx.y);
    const int nn = (int)n;
    const bool in = x < nn && y < nn;
    const int i = in ? y * nn + x : 0;
    double bb = 0.0, rr = 0.0;
    if (in && fl[i]) {
        const double r = (double)b[i] - apply_A64(fl, xp, nullptr, nn, x, y, eps);
        bb = (double)b[i] * (double)b[i];
        rr = r * r;
    }
    block_partial<2>({bb, rr}, st->part, blockIdx.y * gridDim.x + blockIdx.x);
}
enum VcFin { VC_BEGIN, VC_CHECK, VC_END };
template <int W>
__global__ __launch_bounds__(256) void k_vc_fin(PcgState* st, uint32_t nb, float tol, uint32_t maxit) {
    if (W == VC_CHECK && st->done) return;
    double t[2] = {0.0, 0.0};
    if (W != VC_END) sum_partials<2>(st->part, nb, t);
    if (threadIdx.x != 0) return;
    if (W == VC_BEGIN) {
        st->bb = t[0];
        st->rel = t[0] > 0.0 ? sqrt(t[1] / t[0]) : 0.0;
        st->it = 0u;
        st->done = (st->rel <= (double)tol || maxit == 0u) ? 1u : 0u;
        st->tol_seen = tol;
    } else if (W == VC_CHECK) {
        st->rel = st->bb > 0.0 ? sqrt(t[1] / st->bb) : 0.0;
        st->it += 1u;
        st->done = (st->rel <= (double)tol || st->it >= maxit) ? 1u : 0u;
    } else {
        const bool reached = st->rel <= (double)tol;
        st->true_rel = st->rel;
        st->rounded_rel = st->rel;
        st->max_true_rel = fmax(st->max_true_rel, st->rel);
        st->max_rounded_rel = fmax(st->max_rounded_rel, st->rel);
        st->steps += 1u;
        st->iters += st->it;
        if (!reached) st->missed += 1u;
        st->done = 1u;
        if (st->host_out) {
            st->host_out[0] = st->it;
            st->host_out[1] = st->missed;
            st->host_out[2] = st->steps;
            st->host_out[PCG_HOST_RING + (st->steps - 1u) % PCG_HOST_RING_LEN] = st->it;
        }
    }
}

// subtract_gradient (fluid.wgsl:610-630)
__device__ __forceinline__ void b_subgrad(const uint3 blockIdx, const uint3 gridDim, Geo g, const float2* __restrict__ vin, float2* __restrict__ vout,
                                                 const float* __restrict__ P) {
    FLUID_XY
    if (solid(g, idx)) { vout[idx] = f2(0.0f, 0.0f); return; }
    // neighbours outside the box or in walls use p_c; every load is issued
    // (indices clamped to the cell) and the value selected
    const uint32_t il = x > 0u ? idx - 1u : idx, ir = x + 1u < F ? idx + 1u : idx;
    const uint32_t ib = y > 0u ? idx - F : idx, it = y + 1u < F ? idx + F : idx;
    const float p_c = P[idx];
    const float ql = P[il], qr = P[ir], qb = P[ib], qt = P[it];
    const bool wl = solid(g, il), wr = solid(g, ir), wb = solid(g, ib), wt = solid(g, it);
    const float p_l = (x > 0u && !wl) ? ql : p_c, p_r = (x + 1u < F && !wr) ? qr : p_c;
    const float p_b = (y > 0u && !wb) ? qb : p_c, p_t = (y + 1u < F && !wt) ? qt : p_c;
    const float2 grad = mul2(f2(p_r - p_l, p_t - p_b), 0.5f);
    float2 v = sub2(sanitize(vin[idx]), grad);
    v = reflect(g, idx, v);
    vout[idx] = clamp_len(sanitize(v), MAX_VEL);
}
__global__ __launch_bounds__(256) void k_subgrad(Geo g, const float2* __restrict__ vin, float2* __restrict__ vout,
                                                 const float* __restrict__ P) {
    b_subgrad(make_uint3(::blockIdx.x, ::blockIdx.y, ::blockIdx.z), make_uint3(::gridDim.x, ::gridDim.y, ::gridDim.z), g, vin, vout, P);
}

// net_flux (fluid.wgsl:640-657): the mean v_y of each row and the mean v_x
// of each column over fluid cells, summed in the reference's order (k = 0,
// 1, ...) so the fluid stays bit-exact.  A workgroup owns NF_C rows or
// columns; its threads stage up to NF_K values of each chain in LDS in one
// pass, k-major so that the walkers read consecutive words, with walls
// stored as -0 (x + -0 == x for every x, so adding it is the reference's
// skip) and count the fluid cells of their stretch; then NF_C lanes walk
// their chains with one dependent add per cell.
constexpr uint32_t NF_C = 8, NF_K = 1024, NF_T = 1024, NF_PER = NF_C * NF_K / NF_T;
__device__ __forceinline__ void b_net_flux(const uint3 blockIdx, const uint3 gridDim, Geo g, const float2* __restrict__ vin, float* __restrict__ nf) {
    __shared__ float buf[NF_K][NF_C];
    __shared__ uint32_t cnt[NF_T / NF_C][NF_C];
    const uint32_t F = g.F;
    const uint32_t nb = (F + NF_C - 1) / NF_C;
    const bool cols = blockIdx.x < nb;
    const uint32_t c0 = (cols ? blockIdx.x : blockIdx.x - nb) * NF_C;
    const uint32_t t = threadIdx.x;
    // thread t stages NF_PER consecutive cells of chain c
    const uint32_t c = cols ? t % NF_C : t / (NF_T / NF_C);
    const uint32_t h = cols ? t / NF_C : t % (NF_T / NF_C);
    const uint32_t chain = c0 + c;
    float sum = 0.0f;
    uint32_t n = 0u;
    for (uint32_t k0 = 0; k0 < F; k0 += NF_K) {
        const uint32_t len = rb_minu(NF_K, F - k0);
        if (k0 > 0) __syncthreads();
        uint32_t fl = 0u;
#pragma unroll
        for (uint32_t j = 0; j < NF_PER; ++j) {
            const uint32_t k = h * NF_PER + j, pos = k0 + k;
            float val = -0.0f;
            if (chain < F && k < len) {
                const uint32_t q = cols ? pos * F + chain : chain * F + pos;
                const float2 v = sanitize(vin[q]);
                const bool wall = solid(g, q);
                val = wall ? -0.0f : (cols ? v.x : v.y);
                fl += wall ? 0u : 1u;
            }
            buf[k][c] = val;
        }
        cnt[h][c] = fl;
        __syncthreads();
        if (t < NF_C) {
            uint32_t k = 0;
            for (; k + 16 <= len; k += 16) {
#pragma unroll
                for (uint32_t u = 0; u < 16; ++u) sum += buf[k + u][t];
            }
            for (; k < len; ++k) sum += buf[k][t];
            for (uint32_t i = 0; i < NF_T / NF_C; ++i) n += cnt[i][t];
        }
    }
    // the reference counts with n += 1.0 in binary32: exact below 2^24
    if (t < NF_C && c0 + t < F) nf[(cols ? F : 0u) + c0 + t] = n > 0u ? sum / (float)n : 0.0f;
}
__global__ __launch_bounds__(NF_T) void k_net_flux(Geo g, const float2* __restrict__ vin, float* __restrict__ nf) {
    b_net_flux(make_uint3(::blockIdx.x, ::blockIdx.y, ::blockIdx.z), make_uint3(::gridDim.x, ::gridDim.y, ::gridDim.z), g, vin, nf);
}

// enforce_boundaries (fluid.wgsl:663-677)
// vcopy (optional) receives the same velocity: the copy that the transport
// of a pipelined step reads while the next fluid step rewrites vout.
__device__ __forceinline__ void b_boundaries(const uint3 blockIdx, const uint3 gridDim, Geo g, const float2* __restrict__ vin, float2* __restrict__ vout,
                                                    const float* __restrict__ nf, float2* __restrict__ vcopy) {
    FLUID_XY
    float2 v = f2(0.0f, 0.0f);
    if (!solid(g, idx)) {
        v = sanitize(vin[idx]);
        v = sub2(v, f2(nf[F + x], nf[y]));
        if (x == 0u || x == F - 1u) v.x = 0.0f;
        if (y == 0u || y == F - 1u) v.y = 0.0f;
        v = reflect(g, idx, v);
        v = clamp_len(sanitize(v), MAX_VEL);
    }
    vout[idx] = v;
    if (vcopy) vcopy[idx] = v;
}
__global__ __launch_bounds__(256) void k_boundaries(Geo g, const float2* __restrict__ vin, float2* __restrict__ vout,
                                                    const float* __restrict__ nf, float2* __restrict__ vcopy) {
    b_boundaries(make_uint3(::blockIdx.x, ::blockIdx.y, ::blockIdx.z), make_uint3(::gridDim.x, ::gridDim.y, ::gridDim.z), g, vin, vout, nf,
                 vcopy);
}

// thermal_activation (transport.wgsl:419-438) with chem_activate_one
// (chem.wgsl:90-110).  Each thread owns its cell's four words, so the
// compare-and-swap of the reference reduces to a plain read-modify-write.
__device__ __forceinline__ void b_thermal(const uint3 blockIdx, const uint3 gridDim, Geo g, RiboSimParams P, const uint32_t* ep, const float* __restrict__ T, uint32_t* chem) {
    if (ep) P.epoch += *ep;
    const uint32_t x = blockIdx.x * 16 + threadIdx.x, y = blockIdx.y * 16 + threadIdx.y;
    const uint32_t G = g.G, F = g.F;
    if (x >= G || y >= G) return;
    const uint32_t fx = rb_minu((x * F) / G, F - 1u), fy = rb_minu((y * F) / G, F - 1u);
    const float t_loc = T[fy * F + fx];
    if (t_loc <= TEMP_ACT_THRESHOLD) return;
    const float p_act = rb_clamp(FUMAROLE_ACT_P * rb_max(P.thermal_activation, 0.0f) * (t_loc - TEMP_ACT_THRESHOLD) / TEMP_ACT_THRESHOLD, 0.0f, 0.9f);
    const uint32_t idx = y * G + x;
    uint4 v = *reinterpret_cast<uint4*>(chem + 4u * idx);
    uint32_t w[4] = {v.x, v.y, v.z, v.w};
    bool changed = false;
    for (uint32_t ch = 0; ch < 4; ++ch) {
        const uint32_t slot = idx * 4u + ch;
        if ((w[ch] >> 16) > 0u && rng_f4(slot, P.epoch, S_THERMAL, P.seed).x < p_act) {
            w[ch] += 0xFFFF0001u;
            changed = true;
        }
    }
    if (changed) *reinterpret_cast<uint4*>(chem + 4u * idx) = make_uint4(w[0], w[1], w[2], w[3]);
}
__global__ __launch_bounds__(256) void k_thermal(Geo g, RiboSimParams P, const uint32_t* ep, const float* __restrict__ T, uint32_t* chem) {
    b_thermal(make_uint3(::blockIdx.x, ::blockIdx.y, ::blockIdx.z), make_uint3(::gridDim.x, ::gridDim.y, ::gridDim.z), g, P, ep, T, chem);
}

// ---- exchanges with the agents (chem.wgsl:40-98) ---------------------------------------
// chem_cas: replaces the word by old + delta (modulo 2^32) when it holds an
// activated (need_act) or spent (need_spent) monomer.
__device__ __forceinline__ bool chem_cas(uint32_t* slot, uint32_t delta, bool need_act, bool need_spent) {
    uint32_t old = __hip_atomic_load(slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (;;) {
        if (need_act && (old & 0xFFFFu) == 0u) return false;
        if (need_spent && (old >> 16) == 0u) return false;
        const uint32_t prev = atomicCAS(slot, old, old + delta);
        if (prev == old) return true;
        old = prev;
    }
}

// chem_open_cell (chem.wgsl:76-98): the cell itself unless it is rock, else
// the first water or rubble cell on the square rings of radius 1..12 (rows
// from the top of each ring, cells from the left), else the cell itself.
__device__ uint32_t chem_open_cell(const uint32_t* __restrict__ gamma, uint32_t G, uint32_t idx) {
    if (gamma[idx] < GAMMA_SOLID) return idx;
    const int gx = (int)(idx % G), gy = (int)(idx / G);
    for (int r = 1; r <= 12; ++r)
        for (int dy = -r; dy <= r; ++dy)
            for (int dx = -r; dx <= r; ++dx) {
                if (abs(dx) != r && abs(dy) != r) continue;
                const int qx = gx + dx, qy = gy + dy;
                if (qx < 0 || qy < 0 || qx >= (int)G || qy >= (int)G) continue;
                const uint32_t q = (uint32_t)qy * G + (uint32_t)qx;
                if (gamma[q] < GAMMA_SOLID) return q;
            }
    return idx;
}

// One exchange per thread: ops holds (cell, channel, kind, count) words,
// validated on the host.  done[i] receives the monomers moved; the ledger's
// expected channel totals follow what left or entered the grid.
__global__ __launch_bounds__(256) void k_chem_ops(uint32_t* chem, const uint32_t* __restrict__ gamma, uint32_t G,
                                                  const uint4* __restrict__ ops, uint32_t n, uint32_t* __restrict__ done,
                                                  uint32_t* expected) {
    // the net change per channel is summed in the workgroup first: one
    // device-wide atomic per channel and workgroup instead of one per exchange
    __shared__ uint32_t acc[4];
    if (threadIdx.x < 4) acc[threadIdx.x] = 0u;
    __syncthreads();
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < n) {
        const uint4 op = ops[i];
        const uint32_t cell = op.x, ch = op.y, kind = op.z, count = op.w;
        uint32_t moved = 0, net = 0;
        if (kind <= RIBO_CHEM_TAKE_SPENT) {
            // spend: -1 activated +1 spent; activate: the reverse; take: -1 in that state
            const uint32_t delta = kind == RIBO_CHEM_SPEND ? CHEM_SPENT_ONE - 1u
                                   : kind == RIBO_CHEM_ACTIVATE ? 0xFFFF0001u
                                   : kind == RIBO_CHEM_TAKE_ACT ? 0xFFFFFFFFu : 0xFFFF0000u;
            const bool need_act = kind == RIBO_CHEM_SPEND || kind == RIBO_CHEM_TAKE_ACT;
            uint32_t* slot = chem + (size_t)cell * 4u + ch;
            while (moved < count && chem_cas(slot, delta, need_act, !need_act)) ++moved;
            if (kind >= RIBO_CHEM_TAKE_ACT) net = 0u - moved;
        } else {
            const bool spent = kind == RIBO_CHEM_ADD_SPENT || kind == RIBO_CHEM_DEPOSIT_SPENT;
            const uint32_t c = kind >= RIBO_CHEM_DEPOSIT_ACT ? chem_open_cell(gamma, G, cell) : cell;
            if (count > 0u) atomicAdd(chem + (size_t)c * 4u + ch, spent ? count * CHEM_SPENT_ONE : count);  // chem_add_state
            moved = net = count;
        }
        if (done) done[i] = moved;
        if (net != 0u) atomicAdd(&acc[ch], net);
    }
    __syncthreads();
    if (threadIdx.x < 4 && acc[threadIdx.x] != 0u) atomicAdd(expected + threadIdx.x, acc[threadIdx.x]);
}

// ---- terrain grain dynamics (terrain.wgsl:38-102,159-437) -------------------------------
// The reference decides every cell at once, each reading its neighbours'
// grain counts while other cells move grains; the order of those reads is
// not fixed.  These kernels fix one admissible order: every decision of a
// pass reads the counts at the start of the pass (k_relax_decide writes a
// move code per cell), every grain then moves (k_relax_apply: a cell only
// gives its own grain away, so every take succeeds), and the monomers that
// no longer fit a cell that received grains leave it in k_relax_evict, one
// launch per cell colour (x mod 3, y mod 3): cells of one colour are 3
// apart, so their evictions touch disjoint cells and need no atomics.
constexpr uint32_t SANDPILE_DIFF = 3u, GAMMA_ROCK_FULL = 6u, S_RELAX = 16u, S_SAND = 32u;
constexpr float GAMMA_RELAX_P = 0.12f, GAMMA_STRAY_WALK_P = 0.05f, GAMMA_SEDIMENT_FACTOR = 0.5f;
constexpr float GAMMA_SETTLE_SPEED = 0.5f, VENT_BLOW = 8.0f, GAMMA_SHED_P = 0.03f, GRAIN_ADV_CAP = 0.9f;

// The eight neighbours in row order; move code k + 1 means neighbour k.
__constant__ int8_t NB8_DX[8] = {-1, 0, 1, -1, 1, -1, 0, 1};
__constant__ int8_t NB8_DY[8] = {-1, -1, -1, 0, 0, 1, 1, 1};
__device__ __forceinline__ uint32_t nb8_code(int dx, int dy) {  // dx, dy in -1..1, not both 0
    const int k = (dy + 1) * 3 + (dx + 1);
    return (uint32_t)(k < 4 ? k + 1 : k);
}

__device__ __forceinline__ uint32_t mix_hash(uint32_t v) {  // hash (rng.wgsl:46-50)
    uint32_t x = v;
    x = x ^ (x >> 16); x = x * 0x7feb352du; x = x ^ (x >> 15); x = x * 0x846ca68bu; x = x ^ (x >> 16);
    return x;
}
// parcel_hop_p (transport.wgsl:54-58)
__device__ __forceinline__ float hop_p(const Geo& g, const RiboSimParams& P, float vx, float vy) {
    const float env_per_fluid = (float)g.G / (float)g.F;
    const float l1 = rb_abs(vx) + rb_abs(vy);
    return rb_clamp(l1 * env_per_fluid * rb_max(P.dt, 1e-3f), 0.0f, GRAIN_ADV_CAP);
}

// The pass's grain counts of a 16 x 16 tile and a 2-cell halo (every cell a
// decision reads lies within 2 of the deciding cell), and the occupied-cell
// count of the 8-neighbourhood of the tile and a 1-cell halo; cells outside
// the world read 0 and are never counted.
constexpr int RT = 16, RH = 2, RW = RT + 2 * RH, OW = RT + 2;
struct RelaxTile {
    uint32_t g[RW * RW];
    uint8_t occ[OW * OW];
};

// compaction_stable (terrain.wgsl:139-157): src = (sx, sy), dst = (tx, ty).
__device__ bool compaction_stable(const RelaxTile& T, int x0, int y0, int G, int sx, int sy, int tx, int ty, uint32_t n,
                                  uint32_t c) {
    bool ok = true;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int ddx = k == 0 ? 1 : k == 1 ? -1 : 0, ddy = k == 2 ? 1 : k == 3 ? -1 : 0;
        const int qdx = tx + ddx, qdy = ty + ddy;
        if (qdx >= 0 && qdy >= 0 && qdx < G && qdy < G) {
            const uint32_t low = (qdx == sx && qdy == sy) ? n - 1u : T.g[(qdy - y0 + RH) * RW + (qdx - x0 + RH)];
            if (c + 1u >= low + SANDPILE_DIFF) ok = false;
        }
        const int qsx = sx + ddx, qsy = sy + ddy;
        if (qsx >= 0 && qsy >= 0 && qsx < G && qsy < G) {
            const uint32_t high = (qsx == tx && qsy == ty) ? c + 1u : T.g[(qsy - y0 + RH) * RW + (qsx - x0 + RH)];
            if (high >= n - 1u + SANDPILE_DIFF) ok = false;
        }
    }
    return ok;
}

// relax_gamma_pass's decision (terrain.wgsl:159-406): the cell's move code.
__global__ __launch_bounds__(256) void k_relax_decide(Geo g, RiboSimParams P, const uint32_t* ep, uint32_t phase,
                                                      const float2* __restrict__ vel, uint8_t* __restrict__ code) {
    __shared__ RelaxTile T;
    const int G = (int)g.G;
    const uint32_t F = g.F;
    const int x0 = (int)blockIdx.x * RT, y0 = (int)blockIdx.y * RT;
    const int tid = (int)(threadIdx.y * 16 + threadIdx.x);
    for (int i = tid; i < RW * RW; i += 256) {
        const int gx = x0 - RH + i % RW, gy = y0 - RH + i / RW;
        T.g[i] = (gx >= 0 && gy >= 0 && gx < G && gy < G) ? g.gamma[(uint32_t)gy * (uint32_t)G + (uint32_t)gx] : 0u;
    }
    __syncthreads();
    for (int i = tid; i < OW * OW; i += 256) {
        const int lx = i % OW + 1, ly = i / OW + 1;  // in the staged tile
        uint32_t c = 0;
#pragma unroll
        for (int k = 0; k < 8; ++k) c += T.g[(ly + NB8_DY[k]) * RW + lx + NB8_DX[k]] >= 1u ? 1u : 0u;
        T.occ[i] = (uint8_t)c;
    }
    __syncthreads();
    const int x = x0 + (int)threadIdx.x, y = y0 + (int)threadIdx.y;
    if (x >= G || y >= G) return;
    if (ep) P.epoch += *ep;
    auto cnt = [&](int gx, int gy) { return T.g[(gy - y0 + RH) * RW + (gx - x0 + RH)]; };
    const uint32_t idx = (uint32_t)y * (uint32_t)G + (uint32_t)x;
    const uint32_t n = cnt(x, y);
    int dxm = 0, dym = 0;  // the move
    if (n >= 1u && n < SANDPILE_DIFF) {
        uint32_t bonds = 0, best_c = 0xFFFFu;
        int best_score = -1, em_score = -1, bdx = 0, bdy = 0, edx = 0, edy = 0, cdx = 0, cdy = 0;
        bool em_found = false;
        const U4 rr = pcg4d({idx, P.epoch, S_RELAX + phase, P.seed});
        uint32_t tie_h = rr.x;
        uint32_t comp_c = 0;
        const float comp_roll = (float)(mix_hash(rr.x ^ 0x5bd1e995u) >> 8) * (1.0f / 16777216.0f);
        const bool try_comp = comp_roll < P.sediment_compaction;
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                if (dx == 0 && dy == 0) continue;
                const int nx = x + dx, ny = y + dy;
                if (nx < 0 || ny < 0 || nx >= G || ny >= G) continue;
                const uint32_t ni = (uint32_t)ny * (uint32_t)G + (uint32_t)nx;
                const uint32_t c = cnt(nx, ny);
                if (c >= 1u) bonds += 1u;
                if (try_comp && c >= 1u && c < GAMMA_ROCK_FULL &&
                    (c > comp_c || (c == comp_c && (mix_hash(tie_h ^ (ni * 31u)) & 1u) == 1u)) &&
                    compaction_stable(T, x0, y0, G, x, y, nx, ny, n, c)) {
                    comp_c = c;
                    cdx = dx;
                    cdy = dy;
                }
                if (c <= n) {
                    // occupied cells around the neighbour, less this (occupied) cell
                    const int score = (int)T.occ[(ny - y0 + 1) * OW + (nx - x0 + 1)] - 1;
                    tie_h = mix_hash(tie_h ^ ni);
                    if (score > best_score || (score == best_score && c < best_c) ||
                        (score == best_score && c == best_c && (tie_h & 1u) == 1u)) {
                        best_score = score;
                        best_c = c;
                        bdx = dx;
                        bdy = dy;
                    }
                    if (c == 0u && (score > em_score || (score == em_score && (tie_h & 2u) == 2u))) {
                        em_score = score;
                        edx = dx;
                        edy = dy;
                        em_found = true;
                    }
                }
            }
        const float mob = __builtin_ldexpf(1.0f, -2 * (int)bonds);  // pow(0.25, bonds), exact
        if (try_comp && bonds >= 3u && (cdx != 0 || cdy != 0) &&
            comp_roll < P.sediment_compaction * (float)(bonds - 2u) / 6.0f) {
            dxm = cdx;
            dym = cdy;
        } else {
            float vsx = 0.0f, vsy = 0.0f;
            if (P.fluid_enabled != 0u) {
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const int qx = x + (k == 0 ? 1 : k == 1 ? -1 : 0), qy = y + (k == 2 ? 1 : k == 3 ? -1 : 0);
                    if (qx < 0 || qy < 0 || qx >= G || qy >= G) continue;
                    if (cnt(qx, qy) != 0u) continue;
                    const uint32_t fxi = rb_minu(((uint32_t)qx * F) / (uint32_t)G, F - 1u);
                    const uint32_t fyi = rb_minu(((uint32_t)qy * F) / (uint32_t)G, F - 1u);
                    const float2 v = sanitize(vel[fyi * F + fxi]);
                    if (v.x * v.x + v.y * v.y > vsx * vsx + vsy * vsy) { vsx = v.x; vsy = v.y; }
                }
            }
            const bool unsupported = y > 0 && cnt(x, y - 1) == 0u;
            float mvx = 0.0f, mvy = 0.0f, p_sed;
            if (unsupported && bonds == 0u) {
                mvx = vsx; mvy = vsy;
                p_sed = hop_p(g, P, mvx, mvy);
            } else {
                const float speed = rb_len2(vsx, vsy);
                const float crit = rb_max(P.sediment_threshold, 0.0f) * (1.0f + rb_max(P.sediment_cohesion, 0.0f) * (float)bonds);
                const float excess = rb_max(speed - crit, 0.0f);
                if (speed > 1e-6f) {
                    const float k = excess / rb_max(speed, 1e-6f);
                    mvx = vsx * k; mvy = vsy * k;
                }
                p_sed = hop_p(g, P, mvx, mvy) * GAMMA_SEDIMENT_FACTOR * rb_max(P.sediment_transport, 0.0f);
            }
            const float2 vent = g.heat[rb_minu(((uint32_t)y * F) / (uint32_t)G, F - 1u) * F +
                                       rb_minu(((uint32_t)x * F) / (uint32_t)G, F - 1u)];
            const float blow = VENT_BLOW * rb_max(vent.x, vent.y);
            if (blow > 0.0f) {
                mvx = ((rr.y & 1u) == 1u ? 0.6f : -0.6f) * blow;
                mvy = 1.0f * blow;
                p_sed = rb_max(p_sed, hop_p(g, P, mvx, mvy));
            }
            if (!(bonds >= 4u && p_sed <= 0.0f)) {
                const uint32_t hw = rr.y;
                const float rw = (float)(rr.z >> 8) * (1.0f / 16777216.0f);
                const float rw2 = (float)(rr.w >> 8) * (1.0f / 16777216.0f);
                int wdx = 0, wdy = 0;
                bool walk = false;
                if (n >= 2u && em_found && rw2 < GAMMA_SHED_P * rb_clamp(P.sediment_settle, 0.0f, 1.0f)) {
                    dxm = edx;
                    dym = edy;
                } else if (rw < p_sed) {
                    const float ax = rb_abs(mvx), ay = rb_abs(mvy);
                    if ((float)(hw >> 8) * (1.0f / 16777216.0f) < ax / rb_max(ax + ay, 1e-6f)) wdx = mvx > 0.0f ? 1 : -1;
                    else wdy = mvy > 0.0f ? 1 : -1;
                    walk = true;
                } else if (rw < p_sed + GAMMA_STRAY_WALK_P * mob) {
                    if (best_score > (int)bonds) {
                        dxm = bdx;
                        dym = bdy;
                    } else if (bonds == 0u) {
                        const uint32_t k = (hw >> 8) % 8u;  // +x, -x, +y, -y, then the diagonals
                        wdx = k == 0 || k == 4 || k == 6 ? 1 : k == 2 || k == 3 ? 0 : -1;
                        wdy = k == 2 || k == 4 || k == 5 ? 1 : k <= 1 ? 0 : -1;
                        walk = true;
                    }
                }
                if (walk) {  // a step onto a cell with no more grains than this one
                    const int wx = x + wdx, wy = y + wdy;
                    if (wx >= 0 && wy >= 0 && wx < G && wy < G && cnt(wx, wy) <= n) {
                        dxm = wdx;
                        dym = wdy;
                    }
                }
            }
        }
    } else if (n >= SANDPILE_DIFF) {
        uint32_t best_count = n;
        int sdx = 0, sdy = 0;
        if (x > 0 && cnt(x - 1, y) < best_count) { best_count = cnt(x - 1, y); sdx = -1; sdy = 0; }
        if (x + 1 < G && cnt(x + 1, y) < best_count) { best_count = cnt(x + 1, y); sdx = 1; sdy = 0; }
        if (y > 0 && cnt(x, y - 1) < best_count) { best_count = cnt(x, y - 1); sdx = 0; sdy = -1; }
        if (y + 1 < G && cnt(x, y + 1) < best_count) { best_count = cnt(x, y + 1); sdx = 0; sdy = 1; }
        if ((sdx != 0 || sdy != 0) && n - best_count >= SANDPILE_DIFF) {
            const float p = GAMMA_RELAX_P * rb_max(rb_clamp(P.sediment_settle, 0.0f, 1.0f), rb_clamp(P.rock_crumble, 0.0f, 1.0f));
            if (rng_f4(idx, P.epoch, S_SAND + phase, P.seed).x < p) { dxm = sdx; dym = sdy; }
        }
    }
    code[idx] = (uint8_t)((dxm != 0 || dym != 0) ? nb8_code(dxm, dym) : 0u);
}

// The grain moves of a pass: each cell loses its own grain (if it moves one)
// and gains one from every neighbour whose code points at it; inc[idx] gets
// bit j for a grain from neighbour j (row order), for the evictions.
__global__ __launch_bounds__(256) void k_relax_apply(uint32_t G, uint32_t* __restrict__ gamma, const uint8_t* __restrict__ code,
                                                     uint8_t* __restrict__ inc) {
    // the tile's move codes and a 1-cell halo (0 outside the world)
    __shared__ uint8_t C[OW * OW];
    const int x0 = (int)blockIdx.x * RT, y0 = (int)blockIdx.y * RT;
    for (int i = (int)(threadIdx.y * 16 + threadIdx.x); i < OW * OW; i += 256) {
        const int gx = x0 - 1 + i % OW, gy = y0 - 1 + i / OW;
        C[i] = (gx >= 0 && gy >= 0 && gx < (int)G && gy < (int)G) ? code[(uint32_t)gy * G + (uint32_t)gx] : (uint8_t)0;
    }
    __syncthreads();
    const uint32_t x = (uint32_t)x0 + threadIdx.x, y = (uint32_t)y0 + threadIdx.y;
    if (x >= G || y >= G) return;
    const uint32_t idx = y * G + x;
    const int lx = (int)threadIdx.x + 1, ly = (int)threadIdx.y + 1;
    const uint32_t own = C[ly * OW + lx] != 0u ? 1u : 0u;
    uint32_t in = 0, bits = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        // the neighbour whose move k lands here; outside the world its code reads 0
        if (C[(ly - NB8_DY[k]) * OW + (lx - NB8_DX[k])] == (uint8_t)(k + 1)) {
            in += 1u;
            bits |= 1u << (7 - k);  // that neighbour sits at offset -(move k), row-order index 7 - k
        }
    }
    if (own | in) gamma[idx] = gamma[idx] - own + in;
    inc[idx] = (uint8_t)bits;
}

// The eviction loop of gamma_move_one (terrain.wgsl:76-96) once the grain
// from src sits in dst; the caller owns every cell it touches.  The loop
// moves one monomer an iteration while dst is over capacity (at most 64),
// taking them in the order spent then activated of channel 0, 1, 2, 3, each
// to evict_target: src while it has room, then dst's neighbours in order
// while they have room, then the fallback.  Since only this loop changes
// those cells, each target takes a run of consecutive monomers until its
// room is used up, and the runs are moved here in a few steps.
__device__ void evict_one_move(uint32_t* chem, const uint32_t* gm, uint32_t G, uint32_t src, uint32_t dst) {
    const uint32_t cap = capacity(gm[dst]);
    const uint32_t tot = cell_total(chem, dst);
    if (tot <= cap) return;
    uint32_t left = rb_minu(tot - cap, 64u);
    uint4 dv = *reinterpret_cast<const uint4*>(chem + 4u * dst);
    uint32_t seg[8] = {dv.x >> 16, dv.x & CHEM_STATE_MASK, dv.y >> 16, dv.y & CHEM_STATE_MASK,
                       dv.z >> 16, dv.z & CHEM_STATE_MASK, dv.w >> 16, dv.w & CHEM_STATE_MASK};
    uint32_t si = 0;
    // moves `take` monomers of the removal order into cell `to`
    auto give = [&](uint32_t to, uint32_t take) {
        uint4 tv = *reinterpret_cast<const uint4*>(chem + 4u * to);
        while (take > 0u) {
            while (seg[si] == 0u) ++si;
            const uint32_t m = rb_minu(seg[si], take);
            seg[si] -= m;
            take -= m;
            const uint32_t add = (si & 1u) ? m : m * CHEM_SPENT_ONE;
            const uint32_t c = si >> 1;
            if (c == 0) tv.x += add; else if (c == 1) tv.y += add; else if (c == 2) tv.z += add; else tv.w += add;
        }
        *reinterpret_cast<uint4*>(chem + 4u * to) = tv;
    };
    auto room = [&](uint32_t cell) {
        const uint32_t t = cell_total(chem, cell), k = capacity(gm[cell]);
        return t < k ? k - t : 0u;
    };
    {
        const uint32_t r = rb_minu(room(src), left);
        if (r) { give(src, r); left -= r; }
    }
    const int x = (int)(dst % G), y = (int)(dst / G);
    uint32_t not_rock = gm[src] < GAMMA_SOLID ? src : 0xFFFFFFFFu;
    for (int i = 0; i < 4; ++i) {
        const int cx = x + (i == 0 ? 1 : i == 1 ? -1 : 0), cy = y + (i == 2 ? 1 : i == 3 ? -1 : 0);
        if (cx < 0 || cy < 0 || cx >= (int)G || cy >= (int)G) continue;
        const uint32_t n = (uint32_t)cy * G + (uint32_t)cx;
        if (left) {
            const uint32_t r = rb_minu(room(n), left);
            if (r) { give(n, r); left -= r; }
        }
        if (not_rock == 0xFFFFFFFFu && gm[n] < GAMMA_SOLID) not_rock = n;
    }
    if (left) give(not_rock == 0xFFFFFFFFu ? src : not_rock, left);
    *reinterpret_cast<uint4*>(chem + 4u * dst) =
        make_uint4((seg[0] << 16) | seg[1], (seg[2] << 16) | seg[3], (seg[4] << 16) | seg[5], (seg[6] << 16) | seg[7]);
}

// Evictions of the cells of colour (cx, cy) that received grains, their
// incoming grains in row order of the source.
__global__ __launch_bounds__(256) void k_relax_evict(uint32_t G, uint32_t cx, uint32_t cy, uint32_t* chem,
                                                     const uint32_t* __restrict__ gm, const uint8_t* __restrict__ inc) {
    const uint32_t nx = (G - cx + 2u) / 3u, ny = (G - cy + 2u) / 3u;
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (t >= nx * ny) return;
    const uint32_t x = cx + 3u * (t % nx), y = cy + 3u * (t / nx);
    const uint32_t d = y * G + x;
    uint32_t bits = inc[d];
    while (bits) {
        const int j = __builtin_ctz(bits);
        bits &= bits - 1u;
        const uint32_t src = (uint32_t)((int)y + NB8_DY[j]) * G + (uint32_t)((int)x + NB8_DX[j]);
        evict_one_move(chem, gm, G, src, d);
    }
}

// grain_fall (terrain.wgsl:405-437) for the columns x = colour + 3k, one
// thread a column from the bottom up (evictions touch columns x - 1 .. x + 1).
__global__ __launch_bounds__(64) void k_grain_fall(uint32_t G, uint32_t F, RiboSimParams P, const uint32_t* ep,
                                                   uint32_t colour, uint32_t* gamma, uint32_t* chem) {
    const uint32_t x = colour + 3u * (blockIdx.x * 64u + threadIdx.x);
    if (x >= G) return;
    if (ep) P.epoch += *ep;
    Geo gg;
    gg.G = G;
    gg.F = F;
    const float w_s = GAMMA_SETTLE_SPEED * rb_max(P.sediment_settle, 0.0f);
    const float rate = hop_p(gg, P, 0.0f, w_s);
    if (rate <= 0.0f) return;
    const uint32_t rf = rb_f2u(rate * 65536.0f);
    if (((P.epoch * rf) >> 16) == (((P.epoch + 1u) * rf) >> 16)) return;
    // Row y's count changes only when rows y and y + 1 are walked, so the
    // counts are loaded ahead in batches and the walk waits on none of them.
    constexpr uint32_t B = 32;
    bool supported = true;
    for (uint32_t y0 = 0; y0 < G; y0 += B) {
        uint32_t cnt[B];
#pragma unroll
        for (uint32_t j = 0; j < B; ++j) cnt[j] = y0 + j < G ? gamma[(y0 + j) * G + x] : 0u;
#pragma unroll
        for (uint32_t j = 0; j < B; ++j) {
            const uint32_t y = y0 + j;
            if (y >= G) break;
            const uint32_t n = cnt[j];
            if (n == 0u) { supported = false; continue; }
            if (n >= GAMMA_SOLID) { supported = true; continue; }
            if (supported) continue;
            const uint32_t idx = y * G + x;
            for (uint32_t k = 0; k < n; ++k) {  // gamma_move_one(idx, idx - G), grain by grain
                gamma[idx] -= 1u;
                gamma[idx - G] += 1u;
                evict_one_move(chem, gamma, G, idx, idx - G);
            }
        }
    }
}

// ---- transport ------------------------------------------------------------------------
// agg_count + agg_neighbours (transport.wgsl:61-87) in one pass: activated
// counts of the tile and its 1-cell ring go to LDS, then each cell sums its
// 3x3 block; the scatter only ever reads agg_nb + agg_act, stored as one word.
__device__ __forceinline__ void b_agg(const uint3 blockIdx, const uint3 gridDim, Geo g, const uint32_t* __restrict__ chem, uint32_t* __restrict__ agg_e) {
    __shared__ uint32_t act[18][18];
    const uint32_t G = g.G;
    const int32_t bx = (int32_t)(blockIdx.x * 16), by = (int32_t)(blockIdx.y * 16);
    const uint32_t tid = threadIdx.y * 16 + threadIdx.x;
    for (uint32_t k = tid; k < 18u * 18u; k += 256u) {
        const int32_t lx = (int32_t)(k % 18u), ly = (int32_t)(k / 18u);
        const int32_t gx = bx + lx - 1, gy = by + ly - 1;
        uint32_t a = 0u;
        if (gx >= 0 && gy >= 0 && gx < (int32_t)G && gy < (int32_t)G) {
            uint4 v = *reinterpret_cast<const uint4*>(chem + 4u * ((uint32_t)gy * G + (uint32_t)gx));
            a = (v.x & CHEM_STATE_MASK) + (v.y & CHEM_STATE_MASK) + (v.z & CHEM_STATE_MASK) + (v.w & CHEM_STATE_MASK);
        }
        act[ly][lx] = a;
    }
    __syncthreads();
    const uint32_t x = blockIdx.x * 16 + threadIdx.x, y = blockIdx.y * 16 + threadIdx.y;
    if (x >= G || y >= G) return;
    uint32_t s = 0u;
    for (uint32_t dy = 0; dy < 3; ++dy)
        for (uint32_t dx = 0; dx < 3; ++dx) s += act[threadIdx.y + dy][threadIdx.x + dx];
    agg_e[y * G + x] = s;
}
__global__ __launch_bounds__(256) void k_agg(Geo g, const uint32_t* __restrict__ chem, uint32_t* __restrict__ agg_e) {
    b_agg(make_uint3(::blockIdx.x, ::blockIdx.y, ::blockIdx.z), make_uint3(::gridDim.x, ::gridDim.y, ::gridDim.z), g, chem, agg_e);
}

__global__ void k_epoch_add(uint32_t* ep, uint32_t n) { *ep += n; }
__global__ void k_set_u32(uint32_t* p, uint32_t v) { *p = v; }

template <class S> __global__ __launch_bounds__(256) void k_zero(S* p, uint32_t n) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < n) stv(p, (int)i, 0.0f);
}

__device__ __forceinline__ void b_fill_u32(const uint3 blockIdx, const uint3 gridDim, uint32_t* p, size_t n, uint32_t v) {
    size_t i = (size_t)blockIdx.x * 1024u + threadIdx.x * 4u;
    if (i + 3 < n) {
        *reinterpret_cast<uint4*>(p + i) = make_uint4(v, v, v, v);
    } else {
        for (; i < n; ++i) p[i] = v;
    }
}
__global__ __launch_bounds__(256) void k_fill_u32(uint32_t* p, size_t n, uint32_t v) {
    b_fill_u32(make_uint3(::blockIdx.x, ::blockIdx.y, ::blockIdx.z), make_uint3(::gridDim.x, ::gridDim.y, ::gridDim.z), p, n, v);
}

// fluid_velocity_at_world (fluid.wgsl:260-273)
__device__ __forceinline__ float2 fluid_velocity_at_world(const Geo& g, const float2* vin, float2 pw) {
    const float gf = (float)g.F;
    const float gx = rb_clamp(pw.x / g.sim_size * gf, 0.5f, gf - 0.5f) - 0.5f;
    const float gy = rb_clamp(pw.y / g.sim_size * gf, 0.5f, gf - 0.5f) - 0.5f;
    const uint32_t x0 = rb_f2u(rb_floor(gx)), y0 = rb_f2u(rb_floor(gy));
    const uint32_t x1 = rb_minu(x0 + 1u, g.F - 1u), y1 = rb_minu(y0 + 1u, g.F - 1u);
    const float tx = rb_fract(gx), ty = rb_fract(gy);
    const float2 v0 = mix2(sanitize(vin[y0 * g.F + x0]), sanitize(vin[y0 * g.F + x1]), tx);
    const float2 v1 = mix2(sanitize(vin[y1 * g.F + x0]), sanitize(vin[y1 * g.F + x1]), tx);
    return mix2(v0, v1, ty);
}

// lattice_rand / smooth_start (transport.wgsl:95-117)
__device__ __forceinline__ float2 lattice_rand(const RiboSimParams& P, uint32_t ix, uint32_t iy) {
    F4 r = rng_f4((ix * 73856093u) ^ (iy * 19349663u), P.epoch, S_MOVE + MAX_MOVERS_PER_CH + 1u, P.seed);
    return f2(r.x, r.y);
}
// The lattice points of smooth_start a 16 x 16 tile at (bx, by) * 16 reads:
// x + jitter lies in [16 bx, 16 bx + 30], so gx and gx + 1 in bx .. bx + 2.
// lat holds lattice_rand(bx + i % 3, by + i / 3) at i.
__device__ __forceinline__ void smooth_jitter(const RiboSimParams& P, uint32_t& jx, uint32_t& jy) {
    const F4 jit = rng_f4(0x85EBCA6Bu, P.epoch, S_MOVE + MAX_MOVERS_PER_CH + 2u, P.seed);
    jx = rb_f2u(jit.x * (float)SMOOTH_CELL);
    jy = rb_f2u(jit.y * (float)SMOOTH_CELL);
}
__device__ float2 smooth_start(const RiboSimParams& P, uint32_t x0, uint32_t y0, const float2* lat, uint32_t bx, uint32_t by) {
    uint32_t jx, jy;
    smooth_jitter(P, jx, jy);
    const uint32_t x = x0 + jx;
    const uint32_t y = y0 + jy;
    const uint32_t ix = x / SMOOTH_CELL - bx, iy = y / SMOOTH_CELL - by;
    const float tx = rb_smoothstep01(((float)(x % SMOOTH_CELL) + 0.5f) / (float)SMOOTH_CELL);
    const float ty = rb_smoothstep01(((float)(y % SMOOTH_CELL) + 0.5f) / (float)SMOOTH_CELL);
    const float2* l0 = lat + iy * 3u + ix;
    const float2 n = mix2(mix2(l0[0], l0[1], tx), mix2(l0[3], l0[4], tx), ty);
    const F4 gg = rng_f4(0x9E3779B9u, P.epoch, S_MOVE + MAX_MOVERS_PER_CH, P.seed);
    const float vx = rb_fract(n.x + gg.x), vy = rb_fract(n.y + gg.y);
    return f2(1.0f - rb_abs(2.0f * vx - 1.0f), 1.0f - rb_abs(2.0f * vy - 1.0f));
}

// transport_scatter (transport.wgsl:119-406) over a 16 x 16 tile of grid
// cells per workgroup, in three phases.
//  1. One thread per cell forms what the cell's movers share: the advection
//     displacement, the hop, settle and binding probabilities, the pressure
//     weights and Kawasaki acceptances of the four hop directions, and the
//     blocked and in-window masks of the 16 destination bins.  They go to an
//     LDS record per cell.
//  2. The tile's (cell, channel) runs that have movers are sorted by mover
//     count (a counting sort in LDS, with the runs of cells under the rubble
//     jitter in a second class) and dealt to the threads in that order, 256
//     at a time, so the 64 runs of a wave have about the same length and
//     seldom mix jitter and plain runs.
//  3. A thread takes each run dealt to it: the photoactivation and decay of
//     its channel, the cohesion weights of its active movers, then its
//     movers.  Mover k of channel c draws PCG4D at key idx * 4 + c and stream
//     S_MOVE + k (S_DISPERSE + k for the jitter).  With key, epoch and seed
//     fixed, the first half of the hash is a polynomial in k (its y and z
//     words affine, its w word quadratic), so a run steps it by additions and
//     a draw costs the four multiplies of the second half.
// Each mover adds its unit to its destination at once: within SC_R cells of
// the tile into an LDS window that reaches HBM as one coalesced atomic add
// per non-empty (cell, channel) at the end; farther destinations add to HBM
// directly.  Integer sums do not depend on order, so the result is the
// reference's exactly.
constexpr int SC_T = 16, SC_R = 3, SC_W = SC_T + 2 * SC_R;
constexpr int SC_TS = 24;                   // LDS row stride of the window flags (bytes, a multiple of 4)
constexpr uint32_t SC_MB = 64;              // mover-count buckets per class
constexpr uint32_t SC_BUCKETS = 2 * SC_MB;  // plain runs, then runs under the rubble jitter
static_assert(SMOOTH_CELL == (uint32_t)SC_T, "the smooth_start lattice is staged per tile");
constexpr uint32_t LCG_A = 1664525u, LCG_C = 1013904223u;  // the PCG4D input LCG
constexpr float SC_U24 = 1.0f / 16777216.0f;
// A draw's word u gives r = (u >> 8) 2^-24 exactly, so r * c rounds as
// (float)(u >> 8) * (c 2^-24) with c 2^-24 exact, and r + d as one fma.
constexpr float SC_C731 = 7.31f * SC_U24, SC_C317 = 3.17f * SC_U24, SC_C1373 = 13.73f * SC_U24, SC_C531 = 5.31f * SC_U24;
__device__ __forceinline__ uint32_t movers_of(uint32_t w) {
    return rb_minu((w & CHEM_STATE_MASK) + (w >> 16), MAX_MOVERS_PER_CH);
}
// For r = (u >> 8) * 2^-24 and p in [0, 1]: r < p exactly when
// (u >> 8) < ceil(p * 2^24); p * 2^24 is exact.
__device__ __forceinline__ uint32_t u24_threshold(float p) { return (uint32_t)__builtin_ceilf(p * 16777216.0f); }

// rng_f4(key, epoch, stream, seed).x from the key's x1 = LCG(key) +
// LCG(epoch) LCG(seed) and x1s = x1 ^ (x1 >> 16), with y0 = LCG(epoch),
// w0 = LCG(seed), z0 = LCG(stream): the x word needs y, z and w of the
// first round and y, w of the second.
__device__ __forceinline__ float pcg_x_of(uint32_t x1, uint32_t x1s, uint32_t y0, uint32_t w0, uint32_t z0) {
    const uint32_t y1 = y0 + z0 * x1;
    const uint32_t z1 = z0 + x1 * y1;
    const uint32_t w1 = w0 + y1 * z1;
    const uint32_t rx = x1s + (y1 ^ (y1 >> 16)) * (w1 ^ (w1 >> 16));
    return (float)(rx >> 8) * SC_U24;
}

// What the movers of one cell share (written in phase 1, read in phase 3).
// 64 bytes, so that six workgroups fit in a CU's LDS (a run forms the bin
// base, |displacement| and the light itself).  A hop's Kawasaki acceptance
// is exp(-k_e |drop|) for active movers when the drop is positive and for
// spent ones when it is negative, else 1: one value per hop and two sign
// bits.
struct ScCell {
    float4 a;   // displacement x, y; smooth_start x, y
    uint4 b;    // p_bound (bits), hop-or-settle threshold | jitter << 31, direct | blocked << 16,
                // hop threshold | positive drops << 24 | negative drops << 28
    float4 pr;  // pressure weights of the hops +x, -x, +y, -y
    float4 ac;  // exp(-k_e |drop|) of each hop (1 without aggregation)
};
constexpr uint32_t SC_JITTER_BIT = 1u << 31;  // the thresholds are at most 2^24
constexpr uint32_t SC_TF_WORDS = SC_W * SC_TS / 4;

// atomicAdd on a word of LDS, kept an LDS atomic whatever the compiler
// infers of the pointer (a generic atomic would take a flat one).
__device__ __forceinline__ uint32_t lds_add(uint32_t* p, uint32_t v) {
    return __atomic_fetch_add((__attribute__((address_space(3))) uint32_t*)p, v, __ATOMIC_RELAXED);
}

// The LDS of one tile (26.7 KB).
struct ScLds {
    ScCell cell[256];
    uint4 acc4[SC_W * SC_W];  // [window cell][channel]
    // phase 1 reads the window flags (rows of SC_TS bytes: 1 rock, 2 over
    // capacity) and the smooth_start lattice; phases 2 and 3 reuse the words
    // for the sorted runs (cell * 4 + channel)
    uint32_t u[512];
    uint32_t hist[SC_BUCKETS];
    uint32_t nruns;
    uint32_t org;  // the tile: blockIdx.x | blockIdx.y << 16
};
static_assert(SC_TF_WORDS + 18 <= 512 && SC_TF_WORDS % 2 == 0, "flags and lattice fit the run list's words");

// The tile at blockIdx (x, y) by 256 threads (thx, thy in 0..15) that share
// Ls[sub] and every barrier of the workgroup.  NS tiles of one workgroup
// (sub = 0 .. NS - 1) sort their runs together and share them out over all
// NS x 256 threads, so that no tile's threads wait at the barrier after the
// runs for another tile's longer ones.
template <int NS>
__device__ __forceinline__ void b_scatter(ScLds* Ls, const uint32_t sub, const uint32_t thx, const uint32_t thy, const uint3 blockIdx,
                                          Geo g, RiboSimParams P, const uint32_t* ep, const uint32_t* __restrict__ chem,
                                          uint32_t* __restrict__ next, const float2* __restrict__ vel,
                                          const float* __restrict__ temp, const uint32_t* __restrict__ agg_e) {
    ScLds& L = Ls[sub];
    ScCell* const s_cell = L.cell;
    uint4* const s_acc4 = L.acc4;
    uint32_t* const s_hist = L.hist;
    uint32_t* const s_u = L.u;
    uint32_t* const s_tf4 = s_u;
    float2* const s_lat = reinterpret_cast<float2*>(s_u + SC_TF_WORDS);
    uint16_t* const s_run = reinterpret_cast<uint16_t*>(s_u);
    uint32_t* const s_acc = reinterpret_cast<uint32_t*>(s_acc4);
    uint8_t* const s_tf = reinterpret_cast<uint8_t*>(s_tf4);
    if (ep) P.epoch += *ep;
    const uint32_t tid = thy * 16 + thx;
    const uint32_t G = g.G, F = g.F;
    const int wx0 = (int)(blockIdx.x * SC_T) - SC_R, wy0 = (int)(blockIdx.y * SC_T) - SC_R;
    for (uint32_t k = tid; k < SC_W * SC_W; k += 256) s_acc4[k] = make_uint4(0u, 0u, 0u, 0u);
    for (uint32_t k = tid; k < SC_W * SC_W; k += 256) {
        const uint32_t kx = k % SC_W, ky = k / SC_W;
        const int gx = wx0 + (int)kx, gy = wy0 + (int)ky;
        uint8_t f = 0;
        if (gx >= 0 && gy >= 0 && gx < (int)G && gy < (int)G) {
            const uint32_t c = (uint32_t)gy * G + (uint32_t)gx;
            const uint32_t gm = g.gamma[c];
            if (gm >= GAMMA_SOLID) f = 1;
            else if (cell_total(chem, c) > capacity(gm)) f = 2;
        }
        s_tf[ky * SC_TS + kx] = f;
    }
    if (tid < SC_BUCKETS) s_hist[tid] = 0u;
    if (NS > 1 && tid == 0) L.org = blockIdx.x | (blockIdx.y << 16);
    if (tid < 9u) s_lat[tid] = lattice_rand(P, blockIdx.x + tid % 3u, blockIdx.y + tid / 3u);
    __syncthreads();

    // ---- phase 1: the cell's shared quantities ----
    const uint32_t x = blockIdx.x * 16 + thx, y = blockIdx.y * 16 + thy;
    const uint32_t idx = y * G + x;
    uint32_t src_total = 0u;
    uint4 cw = make_uint4(0u, 0u, 0u, 0u);
    if (x < G && y < G) {
        cw = *reinterpret_cast<const uint4*>(chem + 4u * idx);
        src_total = (cw.x & CHEM_STATE_MASK) + (cw.x >> 16) + (cw.y & CHEM_STATE_MASK) + (cw.y >> 16) +
                    (cw.z & CHEM_STATE_MASK) + (cw.z >> 16) + (cw.w & CHEM_STATE_MASK) + (cw.w >> 16);
    }
    bool jitter_on = false;
    if (src_total != 0u) {
        const float cell_w = (float)g.wupc;
        const uint32_t g_src = g.gamma[idx];
        float2 v = f2(0.0f, 0.0f);
        if (P.fluid_enabled != 0u && g_src < GAMMA_SOLID) {
            const float2 c = f2(((float)x + 0.5f) * cell_w, ((float)y + 0.5f) * cell_w);
            const float2 v0 = fluid_velocity_at_world(g, vel, c);
            const float world_per_fluid = g.sim_size / (float)F;
            const float k = rb_max(P.dt, 0.0f);
            v = fluid_velocity_at_world(g, vel, f2(c.x + v0.x * world_per_fluid * k * 0.5f, c.y + v0.y * world_per_fluid * k * 0.5f));
        }
        if (v.x != 0.0f || v.y != 0.0f) {
            const bool rock_r = x + 1u < G && g.gamma[idx + 1u] >= GAMMA_SOLID;
            const bool rock_l = x > 0u && g.gamma[idx - 1u] >= GAMMA_SOLID;
            const bool rock_u = y + 1u < G && g.gamma[idx + G] >= GAMMA_SOLID;
            const bool rock_d = y > 0u && g.gamma[idx - G] >= GAMMA_SOLID;
            if ((v.x > 0.0f && rock_r) || (v.x < 0.0f && rock_l)) v.x = 0.0f;
            else if ((v.x > 0.0f && rock_l) || (v.x < 0.0f && rock_r)) v.x *= 0.5f;
            if ((v.y > 0.0f && rock_u) || (v.y < 0.0f && rock_d)) v.y = 0.0f;
            else if ((v.y > 0.0f && rock_d) || (v.y < 0.0f && rock_u)) v.y *= 0.5f;
        }
        const float tn = (float)rb_maxu(P.transport_every, 1u);
        const float ratio = (float)G / (float)F;
        const float dtp = rb_max(P.dt, 0.0f);
        const float2 disp = f2(v.x * ratio * dtp * tn, v.y * ratio * dtp * tn);
        const float agitation = rb_clamp(len2(v) / DIFF_AGITATION_SPEED, 0.0f, 1.0f);
        float p_diff = rb_clamp(DIFF_HOP_P * rb_mix(DIFF_HOP_FLOOR, 1.0f, agitation) * rb_max(P.diffusion, 0.0f) * tn, 0.0f, 0.5f);
        if (g_src > 0u) p_diff = rb_min(DIFF_HOP_P * BURIED_DIFF_FACTOR / (1.0f + GAMMA_POROSITY_K * (float)g_src) * tn, 0.5f);
        const uint32_t cap_self = capacity(g_src);
        if (src_total > cap_self) p_diff = rb_min(CHEM_SQUEEZE_P * tn, 0.5f);
        float p_settle = 0.0f;
        if (g_src == 0u) p_settle = rb_min(MONOMER_SETTLE_P * rb_max(P.settle, 0.0f) * tn, 0.5f);
        const int32_t base_x = rb_addi(rb_f2i(rb_floor(disp.x)), -1);
        const int32_t base_y = rb_addi(rb_f2i(rb_floor(disp.y)), -1);

        float pr0 = 1.0f, pr1 = 1.0f, pr2 = 1.0f, pr3 = 1.0f;
        float max_drop = 0.0f;
        if (P.monomer_pressure > 0.0f) {  // transport.wgsl:191-206
            const float f_self = (float)src_total / (float)rb_maxu(cap_self, 1u);
            auto press_of = [&](bool inside, uint32_t n, float cur) {
                if (!inside) return cur;
                const uint32_t cap_n = capacity(g.gamma[n]);
                if (cap_n == 0u) return cur;
                const float f_n = (float)cell_total(chem, n) / (float)cap_n;
                max_drop = rb_max(max_drop, f_self - f_n);
                return rb_max(0.05f, 1.0f + P.monomer_pressure * (f_self - f_n));
            };
            pr0 = press_of(x + 1u < G, idx + 1u, pr0);
            pr1 = press_of(x > 0u, idx - 1u, pr1);
            pr2 = press_of(y + 1u < G, idx + G, pr2);
            pr3 = press_of(y > 0u, idx - G, pr3);
        }
        p_diff = rb_clamp(p_diff + PRESSURE_HOP_P * P.monomer_pressure * max_drop * tn, 0.0f, 0.5f);

        float p_bound = 0.0f;
        float ac0 = 1.0f, ac1 = 1.0f, ac2 = 1.0f, ac3 = 1.0f;
        uint32_t signs = 0u;
        if (P.aggregation > 0.0f && src_total <= cap_self) {  // transport.wgsl:228-260
            const float e0 = (float)agg_e[idx];
            const float e1 = x + 1u < G ? (float)agg_e[idx + 1u] : 0.0f;
            const float e2 = x > 0u ? (float)agg_e[idx - 1u] : 0.0f;
            const float e3 = y + 1u < G ? (float)agg_e[idx + G] : 0.0f;
            const float e4 = y > 0u ? (float)agg_e[idx - G] : 0.0f;
            const uint32_t fx = rb_minu((x * F) / G, F - 1u), fy = rb_minu((y * F) / G, F - 1u);
            const float t_rel = 1.0f + rb_max(temp[fy * F + fx], 0.0f) / TEMP_ACT_THRESHOLD;
            const float room = rb_clamp(5.0f * (1.0f - (float)src_total / (float)rb_maxu(cap_self, 1u)), 0.0f, 1.0f);
            const float k_e = P.aggregation * room / t_rel;
            // en - e0 is exactly -(e0 - en), so one exp of -k_e |e0 - en|
            // serves the side that is positive
            auto acc_of = [&](float en, float& acc, uint32_t bit) {
                const float drop = e0 - en;
                acc = rb_expf(-k_e * rb_abs(drop));
                signs |= (drop > 0.0f ? 1u : 0u) << bit;
                signs |= (drop < 0.0f ? 1u : 0u) << (bit + 4u);
            };
            acc_of(e1, ac0, 0u);
            acc_of(e2, ac1, 1u);
            acc_of(e3, ac2, 2u);
            acc_of(e4, ac3, 3u);
            p_bound = 1.0f - rb_expf(-k_e * rb_min(rb_max(e0 - 1.0f, 0.0f), 48.0f));
        }
        jitter_on = g_src > 0u && (disp.x != 0.0f || disp.y != 0.0f);
        // smooth_start is a function of the cell alone; every bound mover of
        // the cell starts there (transport.wgsl:95-117,290-296)
        float2 bound_start = f2(0.0f, 0.0f);
        if (p_bound > 0.0f) bound_start = smooth_start(P, x, y, s_lat, blockIdx.x, blockIdx.y);
        const int lself = ((int)y - wy0) * SC_W + ((int)x - wx0);

        // The 4 x 4 destinations of the bins (transport.wgsl:330-380):
        // a bin in `direct` adds to the window at lbase + oy * SC_W + ox, or
        // to the cell itself when it is in `blocked`; the other bins (clamped
        // at the grid edge, or outside the window) take the general path.
        const int32_t lbase = rb_addi(rb_addi(lself, (int32_t)((uint32_t)base_y * (uint32_t)SC_W)), base_x);
        uint32_t direct, blocked = 0u;
        const int32_t lx0 = rb_addi((int)x - wx0, base_x), ly0 = rb_addi((int)y - wy0, base_y);
        const bool block4 = base_x >= -SC_W && base_x <= SC_W && base_y >= -SC_W && base_y <= SC_W &&
                            lx0 >= 0 && lx0 + 3 < SC_W && ly0 >= 0 && ly0 + 3 < SC_W &&
                            (int)x + base_x >= 0 && (int)x + base_x + 3 < (int)G &&
                            (int)y + base_y >= 0 && (int)y + base_y + 3 < (int)G;
        if (block4) {
            // all 16 destinations inside the grid and the window: one
            // 4-byte row of flags per bin row; a byte is blocked when it
            // is 2, or 1 while this cell is not rock
            const uint32_t rockm = g_src < GAMMA_SOLID ? 0x01010101u : 0u;
            for (int i = 0; i < 4; ++i) {
                const uint32_t off = (uint32_t)((ly0 + i) * SC_TS + lx0);
                const uint32_t v = __builtin_amdgcn_alignbyte(s_tf4[(off >> 2) + 1u], s_tf4[off >> 2], off & 3u);
                const uint32_t b = ((v >> 1) & 0x01010101u) | (v & rockm);
                blocked |= ((b * 0x01020408u) >> 24) << (4 * i);  // byte k's low bit to bit k
            }
            if (base_x <= 0 && base_x >= -3 && base_y <= 0 && base_y >= -3)  // the cell itself
                blocked &= ~(1u << (uint32_t)(-base_y * 4 - base_x));
            direct = 0xFFFFu;
        } else {
            direct = 0u;
            for (uint32_t b = 0; b < 16; ++b) {
                const int32_t tx = rb_clampi(rb_addi(rb_addi((int32_t)x, base_x), (int32_t)(b & 3u)), 0, (int32_t)G - 1);
                const int32_t ty = rb_clampi(rb_addi(rb_addi((int32_t)y, base_y), (int32_t)(b >> 2)), 0, (int32_t)G - 1);
                const uint32_t t_idx = (uint32_t)ty * G + (uint32_t)tx;
                const int lx = tx - wx0, ly = ty - wy0;
                const bool in_win = lx >= 0 && ly >= 0 && lx < SC_W && ly < SC_W;
                bool blk = false;
                if (t_idx != idx) {
                    uint32_t tf;
                    if (in_win) {
                        tf = s_tf[ly * SC_TS + lx];
                    } else {
                        const uint32_t g_tgt = g.gamma[t_idx];
                        tf = g_tgt >= GAMMA_SOLID ? 1u : (cell_total(chem, t_idx) > capacity(g_tgt) ? 2u : 0u);
                    }
                    blk = (tf == 1u) ? (g_src < GAMMA_SOLID) : (tf == 2u);
                }
                const bool fast = in_win && ly * SC_W + lx == rb_addi(lbase, (int32_t)((b >> 2) * SC_W + (b & 3u)));
                blocked |= (uint32_t)blk << b;
                direct |= (uint32_t)(blk || fast) << b;
            }
        }
        // hop: r < p_diff; settle: r < p_diff + p_settle (transport.wgsl:337-357)
        ScCell& R = s_cell[tid];
        R.a = make_float4(disp.x, disp.y, bound_start.x, bound_start.y);
        R.b = make_uint4(rb_as_u32(p_bound), u24_threshold(p_diff + p_settle) | (jitter_on ? SC_JITTER_BIT : 0u),
                         direct | (blocked << 16), u24_threshold(p_diff) | (signs << 24));
        R.pr = make_float4(pr0, pr1, pr2, pr3);
        R.ac = make_float4(ac0, ac1, ac2, ac3);
    }

    // ---- phase 2: sort the runs by mover count ----
    uint32_t rk0 = 0u, rk1 = 0u, rk2 = 0u, rk3 = 0u;
    const uint32_t cls = jitter_on ? SC_MB : 0u;
    const uint32_t m0 = movers_of(cw.x), m1 = movers_of(cw.y), m2 = movers_of(cw.z), m3 = movers_of(cw.w);
    const uint32_t b0 = cls + rb_minu(m0, SC_MB - 1u), b1 = cls + rb_minu(m1, SC_MB - 1u);
    const uint32_t b2 = cls + rb_minu(m2, SC_MB - 1u), b3 = cls + rb_minu(m3, SC_MB - 1u);
    if (m0) rk0 = lds_add(&s_hist[b0], 1u);
    if (m1) rk1 = lds_add(&s_hist[b1], 1u);
    if (m2) rk2 = lds_add(&s_hist[b2], 1u);
    if (m3) rk3 = lds_add(&s_hist[b3], 1u);
    __syncthreads();
    if (sub == 0 && tid < 64) {  // exclusive scan of the bucket counts; lane l holds buckets 2l and 2l + 1
        uint32_t h0[NS], h1[NS], t0 = 0u, t1 = 0u;
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            h0[k] = Ls[k].hist[2 * tid];
            h1[k] = Ls[k].hist[2 * tid + 1];
            t0 += h0[k];
            t1 += h1[k];
        }
        uint32_t v = t0 + t1;
        for (uint32_t off = 1; off < 64; off <<= 1) {
            const uint32_t u = __shfl_up(v, off, 64);
            if (tid >= off) v += u;
        }
        // a bucket's runs: tile 0's, then tile 1's, ...
        uint32_t o0 = v - (t0 + t1), o1 = o0 + t0;
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            Ls[k].hist[2 * tid] = o0;
            Ls[k].hist[2 * tid + 1] = o1;
            o0 += h0[k];
            o1 += h1[k];
        }
        if (tid == 63) Ls[0].nruns = v;
    }
    __syncthreads();
    // run list entries: tile << 10 | cell * 4 + channel; entry p lies in
    // tile p >> 10's words
    auto put_run = [&](uint32_t p, uint32_t c) {
        const uint16_t e = (uint16_t)((NS > 1 ? sub << 10 : 0u) | (tid * 4u + c));
        if (NS == 1) s_run[p] = e;
        else reinterpret_cast<uint16_t*>(Ls[p >> 10].u)[p & 1023u] = e;
    };
    if (m0) put_run(s_hist[b0] + rk0, 0u);
    if (m1) put_run(s_hist[b1] + rk1, 1u);
    if (m2) put_run(s_hist[b2] + rk2, 2u);
    if (m3) put_run(s_hist[b3] + rk3, 3u);
    __syncthreads();

    // ---- phase 3: the runs ----
    const uint32_t nruns = Ls[0].nruns;
    const float tn = (float)rb_maxu(P.transport_every, 1u);
    // PCG4D words that do not depend on key or stream
    const uint32_t y0 = P.epoch * LCG_A + LCG_C, w0 = P.seed * LCG_A + LCG_C, y0w0 = y0 * w0;
    constexpr uint32_t Z0 = S_MOVE * LCG_A + LCG_C, ZJ0 = S_DISPERSE * LCG_A + LCG_C;
    for (uint32_t pos = sub * 256u + tid; pos < nruns; pos += 256u * NS) {
        const uint32_t e = NS == 1 ? s_run[pos] : reinterpret_cast<const uint16_t*>(Ls[pos >> 10].u)[pos & 1023u];
        const uint32_t run = e & 1023u, cell = run >> 2, ch = run & 3u;
        ScLds& Lr = Ls[NS == 1 ? 0u : e >> 10];  // the run's tile
        uint32_t* const r_acc = reinterpret_cast<uint32_t*>(Lr.acc4);
        const uint32_t org = NS == 1 ? 0u : Lr.org;
        const uint32_t rbx = NS == 1 ? blockIdx.x : (org & 0xFFFFu), rby = NS == 1 ? blockIdx.y : (org >> 16);
        const int rwx0 = (int)(rbx * SC_T) - SC_R, rwy0 = (int)(rby * SC_T) - SC_R;
        const uint32_t cx = rbx * 16 + (cell & 15u), cy = rby * 16 + (cell >> 4);
        const uint32_t cidx = cy * G + cx;
        const ScCell& R = Lr.cell[cell];
        const float4 ra = R.a, rp = R.pr, rv = R.ac;
        const uint4 rd = R.b;
        const float p_bound = rb_as_f32(rd.x);
        const float pr0 = rp.x, pr1 = rp.y, pr2 = rp.z, pr3 = rp.w;
        const int32_t base_x = rb_addi(rb_f2i(rb_floor(ra.x)), -1);
        const int32_t base_y = rb_addi(rb_f2i(rb_floor(ra.y)), -1);
        const int32_t lself = (int32_t)((cell >> 4) + SC_R) * SC_W + (int32_t)((cell & 15u) + SC_R);
        const int32_t lbase = rb_addi(rb_addi(lself, (int32_t)((uint32_t)base_y * (uint32_t)SC_W)), base_x);
        const uint32_t direct = rd.z & 0xFFFFu, blocked = rd.z >> 16;
        const uint32_t t_hop = rd.w & 0xFFFFFFu, t_set = rd.y & ~SC_JITTER_BIT;
        const bool jit = (rd.y & SC_JITTER_BIT) != 0u;
        const uint32_t sg = rd.w >> 24;
        const float4 rac = make_float4((sg & 1u) ? rv.x : 1.0f, (sg & 2u) ? rv.y : 1.0f, (sg & 4u) ? rv.z : 1.0f, (sg & 8u) ? rv.w : 1.0f);
        const float4 ras = make_float4((sg & 16u) ? rv.x : 1.0f, (sg & 32u) ? rv.y : 1.0f, (sg & 64u) ? rv.z : 1.0f, (sg & 128u) ? rv.w : 1.0f);

        // photoactivation and decay of the channel (transport.wgsl:269-294);
        // every draw of the run shares the key's first PCG4D words
        const uint32_t slot = cidx * 4u + ch;
        const uint32_t x1 = slot * LCG_A + LCG_C + y0w0;
        const uint32_t x1s = x1 ^ (x1 >> 16);
        const uint32_t v_ch = chem[slot];
        uint32_t act_n = v_ch & CHEM_STATE_MASK;
        uint32_t spent_n = v_ch >> 16;
        if (spent_n > 0u) {
            const uint32_t lxi = rb_minu(cx, G - 1u) / LIGHT_DIV, lyi = rb_minu(cy, G - 1u) / LIGHT_DIV;
            const float light_t = g.light[lyi * g.Lsz + lxi];
            const float sens = 1.0f + CHEM_SENSITIZE * (float)rb_minu(act_n, 8u);
            const float exp_act = (float)rb_minu(spent_n, 8u) * tn *
                                  (LIGHT_ACT_P * rb_max(P.uv_strength, 0.0f) * rb_max(P.direct_photoactivation, 0.0f) * light_t * sens +
                                   rb_max(P.reactivation_rate, 0.0f));
            uint32_t na = rb_f2u(rb_floor(exp_act));
            const float frac = exp_act - (float)na;
            if (frac > 0.0f && pcg_x_of(x1, x1s, y0, w0, S_PHOTO * LCG_A + LCG_C) < frac) na += 1u;
            na = rb_minu(na, rb_minu(spent_n, 8u));
            act_n += na;
            spent_n -= na;
        }
        // (a decay rate of 0 or below makes exp_dec 0: no draw, no change)
        if (act_n > 0u && P.activation_decay > 0.0f) {
            const float shield = 1.0f / (1.0f + CHEM_SHIELD * (float)(rb_minu(act_n, 8u) - 1u));
            const float exp_dec = (float)rb_minu(act_n, 8u) * rb_max(P.activation_decay, 0.0f) * shield * tn;
            if (exp_dec > 0.0f && pcg_x_of(x1, x1s, y0, w0, S_DECAY * LCG_A + LCG_C) < exp_dec) {
                act_n -= 1u;
                spent_n += 1u;
            }
        }
        // cohesion of active movers (transport.wgsl:296-304)
        float ch0 = 1.0f, ch1 = 1.0f, ch2 = 1.0f, ch3 = 1.0f;
        if (act_n > 0u && P.cohesion > 0.0f) {
            auto ac = [&](uint32_t c) { return (float)rb_minu(chem[c * 4u + ch] & CHEM_STATE_MASK, 8u); };
            if (cx + 1u < G) ch0 = 1.0f + P.cohesion * ac(cidx + 1u);
            if (cx > 0u) ch1 = 1.0f + P.cohesion * ac(cidx - 1u);
            if (cy + 1u < G) ch2 = 1.0f + P.cohesion * ac(cidx + G);
            if (cy > 0u) ch3 = 1.0f + P.cohesion * ac(cidx - G);
        }
        const float wa0 = pr0 * ch0, wa1 = pr1 * ch1, wa2 = pr2 * ch2, wa3 = pr3 * ch3;
        const float wsum_a = wa0 + wa1 + wa2 + wa3, wsum_s = pr0 + pr1 + pr2 + pr3;
        // the sums are at least 0.2 (or inf, or NaN), so scaling them by
        // 2^-24 is exact
        const float wsum_a24 = wsum_a * SC_U24, wsum_s24 = wsum_s * SC_U24;
        const float bsdx = ra.z + ra.x, bsdy = ra.w + ra.y;  // a bound mover's start plus the displacement
        // With |displacement| < 2^22 a position stays below 2^22 + 3 in
        // magnitude, so floor(p) - base is an exact float and the bin needs
        // no saturating conversion; the wave takes that form when all its
        // runs allow it.
        const bool fast_bins = __all(rb_abs(ra.x) < 4194304.0f && rb_abs(ra.y) < 4194304.0f);
        // every bin of every run of the wave lies in the LDS window (the
        // common case away from the grid's edge and from fast flow)
        const bool all_direct = __all(direct == 0xFFFFu);
        const float bxf = (float)base_x, byf = (float)base_y;
        const uint32_t n = act_n + spent_n;
        const uint32_t movers = rb_minu(n, MAX_MOVERS_PER_CH);
        const uint32_t act_mv = rb_minu(act_n, movers);
        const uint32_t act_stay = act_n - act_mv;
        const uint32_t stay = act_stay + (n - movers - act_stay) * CHEM_SPENT_ONE;
        if (stay > 0u) lds_add(r_acc + lself * 4 + (int)ch, stay);

        // the run's PCG4D polynomials in k: y = Y + k dY, z = Z + k dZ,
        // w = W + k D0 + k (k - 1) / 2 DD (mod 2^32)
        const uint32_t dY = LCG_A * x1;
        const uint32_t dZ = LCG_A + x1 * dY;
        const uint32_t dYZ = dY * dZ, DD = 2u * dYZ;
        uint32_t Y = y0 + Z0 * x1;
        uint32_t Z = Z0 + x1 * Y;
        uint32_t W = w0 + Y * Z;
        uint32_t D = Y * dZ + Z * dY + dYZ;
        uint32_t Yj = 0u, Zj = 0u, Wj = 0u, Dj = 0u;
        float dlen = 0.0f;
        if (jit) {
            dlen = len2(f2(ra.x, ra.y));
            Yj = y0 + ZJ0 * x1;
            Zj = ZJ0 + x1 * Yj;
            Wj = w0 + Yj * Zj;
            Dj = Yj * dZ + Zj * dY + dYZ;
        }
        for (uint32_t k = 0; k < movers; ++k) {
            uint32_t y1 = Y, z1 = Z, w1 = W;
            Y += dY;
            Z += dZ;
            W += D;
            D += DD;
            y1 ^= y1 >> 16;
            z1 ^= z1 >> 16;
            w1 ^= w1 >> 16;
            const uint32_t rx = x1s + y1 * w1;
            const uint32_t ry = y1 + z1 * rx;
            const uint32_t rz = z1 + rx * ry;
            const uint32_t rw = w1 + ry * rz;
            const float kx = (float)(rx >> 8), ky = (float)(ry >> 8), kw = (float)(rw >> 8);
            const uint32_t kz = rz >> 8;
            const bool is_act = k < act_mv;
            const bool bound = is_act && p_bound > 0.0f && rb_fract((float)kz * SC_C731 + kw * SC_C317) < p_bound;
            float px = bound ? bsdx : rb_fma(kx, SC_U24, ra.x);
            float py = bound ? bsdy : rb_fma(ky, SC_U24, ra.y);
            if (jit) {
                uint32_t yj = Yj, zj = Zj, wj = Wj;
                Yj += dY;
                Zj += dZ;
                Wj += Dj;
                Dj += DD;
                yj ^= yj >> 16;
                zj ^= zj >> 16;
                wj ^= wj >> 16;
                const uint32_t jxu = x1s + yj * wj;
                const uint32_t jyu = yj + zj * jxu;
                const float jx = ((float)(jxu >> 8) * SC_U24 * 2.0f - 1.0f) * dlen * RUBBLE_DISPERSION;
                const float jy = ((float)(jyu >> 8) * SC_U24 * 2.0f - 1.0f) * dlen * RUBBLE_DISPERSION;
                px += rb_clamp(jx, -0.9f, 0.9f);
                py += rb_clamp(jy, -0.9f, 0.9f);
            }
            // the hop (transport.wgsl:337-357) as selects: direction d is
            // the first of u < w0, u - w0 < w1, u - w0 - w1 < w2 to hold
            // (else 3); a position steps by +-1 or by 0, and x + 0 differs
            // from x only for x = -0, which floors to the same bin
            const bool hop = kz < t_hop;
            const float w0s = is_act ? wa0 : pr0, w1s = is_act ? wa1 : pr1, w2s = is_act ? wa2 : pr2;
            const float u0 = kw * (is_act ? wsum_a24 : wsum_s24);
            const float u1 = u0 - w0s, u2 = u1 - w1s;
            const bool d0 = u0 < w0s, d1 = u1 < w1s, d2 = u2 < w2s;
            const float a0 = is_act ? rac.x : ras.x, a1 = is_act ? rac.y : ras.y;
            const float a2 = is_act ? rac.z : ras.z, a3 = is_act ? rac.w : ras.w;
            const float acc_d = d0 ? a0 : (d1 ? a1 : (d2 ? a2 : a3));
            const bool moves = hop && !(rb_fract(kw * SC_C1373 + kx * SC_C531) >= acc_d);
            const bool vert = !d0 && !d1;
            px += (moves && !vert) ? (d0 ? 1.0f : -1.0f) : 0.0f;
            py += (moves && vert) ? (d2 ? 1.0f : -1.0f) : ((!hop && kz < t_set) ? -1.0f : 0.0f);
            int32_t ox, oy;
            if (fast_bins) {
                ox = (int32_t)__builtin_amdgcn_fmed3f(rb_floor(px) - bxf, 0.0f, 3.0f);
                oy = (int32_t)__builtin_amdgcn_fmed3f(rb_floor(py) - byf, 0.0f, 3.0f);
            } else {
                ox = rb_clampi(rb_addi(rb_f2i(rb_floor(px)), -base_x), 0, 3);
                oy = rb_clampi(rb_addi(rb_f2i(rb_floor(py)), -base_y), 0, 3);
            }
            const uint32_t bi = (uint32_t)oy * 4u + (uint32_t)ox;
            const uint32_t unit = is_act ? 1u : CHEM_SPENT_ONE;
            if (all_direct) {
                const int32_t tgt = ((blocked >> bi) & 1u) ? lself : rb_addi(lbase, (int32_t)__umul24((uint32_t)oy, SC_W) + ox);
                lds_add(r_acc + tgt * 4 + (int)ch, unit);
            } else if ((direct >> bi) & 1u) {
                const int32_t tgt = ((blocked >> bi) & 1u) ? lself : rb_addi(lbase, (int32_t)__umul24((uint32_t)oy, SC_W) + ox);
                lds_add(r_acc + tgt * 4 + (int)ch, unit);
            } else {
                const int32_t tx = rb_clampi(rb_addi(rb_addi((int32_t)cx, base_x), ox), 0, (int32_t)G - 1);
                const int32_t ty = rb_clampi(rb_addi(rb_addi((int32_t)cy, base_y), oy), 0, (int32_t)G - 1);
                const int lx = tx - rwx0, ly = ty - rwy0;
                if (lx >= 0 && ly >= 0 && lx < SC_W && ly < SC_W) lds_add(r_acc + (ly * SC_W + lx) * 4 + (int)ch, unit);
                else atomicAdd(next + ((uint32_t)ty * G + (uint32_t)tx) * 4u + ch, unit);
            }
        }
    }
    __syncthreads();
    // consecutive threads take consecutive words, so a wave's atomics cover
    // whole cache lines of chem_next (only cells inside the grid receive adds)
    for (uint32_t k = tid; k < SC_W * SC_W * 4; k += 256) {
        const uint32_t v = s_acc[k];
        if (v == 0u) continue;
        const int gx = wx0 + (int)((k / 4) % SC_W), gy = wy0 + (int)((k / 4) / SC_W);
        atomicAdd(next + ((uint32_t)gy * G + (uint32_t)gx) * 4u + (k & 3u), v);
    }
}
__global__ __launch_bounds__(256, 6) void k_scatter(Geo g, RiboSimParams P, const uint32_t* ep, const uint32_t* __restrict__ chem,
                                                 uint32_t* __restrict__ next, const float2* __restrict__ vel,
                                                 const float* __restrict__ temp, const uint32_t* __restrict__ agg_e) {
    __shared__ ScLds lds;
    b_scatter<1>(&lds, 0u, threadIdx.x, threadIdx.y, make_uint3(::blockIdx.x, ::blockIdx.y, ::blockIdx.z), g, P, ep, chem, next, vel,
                 temp, agg_e);
}

// Tiles t0 + blockIdx.x of the scatter (row-major over the grid's tiles).
__global__ __launch_bounds__(256, 6) void k_scatter_range(ScatterArgs a, uint32_t t0) {
    __shared__ ScLds lds;
    const uint32_t tiles_x = (a.g.G + 15u) / 16u, t = t0 + blockIdx.x;
    b_scatter<1>(&lds, 0u, threadIdx.x, threadIdx.y, make_uint3(t % tiles_x, t / tiles_x, 0u), a.g, a.P, a.ep, a.chem, a.next, a.vel,
                 a.temp, a.agg_e);
}

// The transport scatter beside the solver's one-workgroup coarse levels, in
// one launch: workgroup mgw runs the symmetric V-cycle's k_mg_block, which
// leaves every other CU idle; every other workgroup (w-th in order) runs
// NT = FuseCfg<S>::tiles scatter tiles of 256 threads each, tiles t0 + w NT
// + (threadIdx.x >> 8) up to t1 (a sub-tile past t1 takes a tile past the
// grid's last row: no cells, only the barriers).  Neither part reads what
// the other writes; the scatter runs whether or not the solve is done.  The
// V-cycle's LDS overlays the tiles' (it is at most NT sizeof(ScLds) bytes).
template <class S>
__global__ __launch_bounds__(FuseCfg<S>::tiles * 256, FuseCfg<S>::tiles * FuseCfg<S>::per_cu) void k_mg_block_scatter(
    MgBlockArgs<S> a, const uint32_t* done, ScatterArgs sc, uint32_t t0, uint32_t t1, uint32_t mgw) {
    constexpr uint32_t NT = FuseCfg<S>::tiles;
    __shared__ ScLds lds[NT];
    if (blockIdx.x == mgw) {
        b_mg_block<S, true, NT * 8>(make_uint3(0u, 0u, 0u), make_uint3(1u, 1u, 1u), a, done, reinterpret_cast<float*>(lds));
        return;
    }
    const uint32_t sub = threadIdx.x >> 8, t = threadIdx.x & 255u;
    const uint32_t tiles_x = (sc.g.G + 15u) / 16u;
    const uint32_t tile = t0 + (blockIdx.x < mgw ? blockIdx.x : blockIdx.x - 1u) * NT + sub;
    const uint3 tb = tile < t1 ? make_uint3(tile % tiles_x, tile / tiles_x, 0u) : make_uint3(0u, tiles_x, 0u);
    b_scatter<NT>(lds, sub, t & 15u, t >> 4, tb, sc.g, sc.P, sc.ep, sc.chem, sc.next, sc.vel, sc.temp, sc.agg_e);
}

// ledger_reduce (ledger.wgsl:5-27): exact u32 sums per channel and state
// (wrapping, as the reference's atomics).  k_ledger leaves one partial per
// block; k_ledger_fin (one block, the next launch) sums them in a fixed
// order, publishes the totals and, when asked, checks the per-channel
// act + spent totals against the armed values.
__device__ __forceinline__ void b_ledger(const uint3 blockIdx, const uint3 gridDim, const uint32_t* __restrict__ chem, uint32_t cells, LedgerDev* led) {
    __shared__ uint32_t sh[4][8];
    uint32_t a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < cells; i += gridDim.x * 256) {
        const uint4 v = *reinterpret_cast<const uint4*>(chem + 4u * i);
        a[0] += v.x & CHEM_STATE_MASK; a[4] += v.x >> 16;
        a[1] += v.y & CHEM_STATE_MASK; a[5] += v.y >> 16;
        a[2] += v.z & CHEM_STATE_MASK; a[6] += v.z >> 16;
        a[3] += v.w & CHEM_STATE_MASK; a[7] += v.w >> 16;
    }
    for (int k = 0; k < 8; ++k)
        for (int o = 32; o > 0; o >>= 1) a[k] += __shfl_down(a[k], o, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0)
        for (int k = 0; k < 8; ++k) sh[wid][k] = a[k];
    __syncthreads();
    if (threadIdx.x < 8) led->part[blockIdx.x * 8 + threadIdx.x] = sh[0][threadIdx.x] + sh[1][threadIdx.x] + sh[2][threadIdx.x] + sh[3][threadIdx.x];
}
__global__ __launch_bounds__(256) void k_ledger(const uint32_t* __restrict__ chem, uint32_t cells, LedgerDev* led) {
    b_ledger(make_uint3(::blockIdx.x, ::blockIdx.y, ::blockIdx.z), make_uint3(::gridDim.x, ::gridDim.y, ::gridDim.z), chem, cells, led);
}

__device__ __forceinline__ void b_ledger_fin(const uint3 blockIdx, const uint3 gridDim, LedgerDev* led, uint32_t nb, uint32_t epoch, const uint32_t* ep, int check) {
    if (ep) epoch += *ep;
    __shared__ uint32_t sh[4][8];
    uint32_t t8[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (uint32_t b = threadIdx.x; b < nb; b += 256u)
        for (int k = 0; k < 8; ++k) t8[k] += led->part[b * 8 + k];
    for (int k = 0; k < 8; ++k)
        for (int o = 32; o > 0; o >>= 1) t8[k] += __shfl_down(t8[k], o, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0)
        for (int k = 0; k < 8; ++k) sh[wid][k] = t8[k];
    __syncthreads();
    if (threadIdx.x != 0) return;
    uint32_t tot[8];
    for (int k = 0; k < 8; ++k) {
        tot[k] = sh[0][k] + sh[1][k] + sh[2][k] + sh[3][k];
        led->last[k] = tot[k];
    }
    if (check) {
        if (!led->armed) {
            for (int c = 0; c < 4; ++c) led->expected[c] = tot[c] + tot[4 + c];
            led->armed = 1u;
        } else {
            bool ok = true;
            for (int c = 0; c < 4; ++c) ok = ok && (led->expected[c] == tot[c] + tot[4 + c]);
            if (!ok && epoch < led->first_bad) led->first_bad = epoch;
        }
    }
}
__global__ __launch_bounds__(256) void k_ledger_fin(LedgerDev* led, uint32_t nb, uint32_t epoch, const uint32_t* ep, int check) {
    b_ledger_fin(make_uint3(::blockIdx.x, ::blockIdx.y, ::blockIdx.z), make_uint3(::gridDim.x, ::gridDim.y, ::gridDim.z), led, nb, epoch, ep, check);
}

// ---- many worlds in one launch ---------------------------------------------------
// A batched launch runs one kernel over the blocks of several worlds: its
// block map gives, per workgroup, the world (an index into the kernel's
// argument table) and the block index and grid size the kernel body sees
// for that world.  The bodies are the single-world kernels' own, so every
// world computes exactly what it computes alone.  The map entry and the
// world index are uniform over the workgroup (scalar loads).
#define BATCH_BLOCK(tab)                                                                                   \
    const BatchBlock bb = map[::blockIdx.x];                                                               \
    const uint32_t bw = __builtin_amdgcn_readfirstlane(bb.w);                                              \
    const uint32_t bgd = __builtin_amdgcn_readfirstlane(bb.gdim);                                          \
    const uint3 bid = make_uint3(__builtin_amdgcn_readfirstlane(bb.bx), __builtin_amdgcn_readfirstlane(bb.by), 0u); \
    const uint3 gdim = make_uint3(bgd & 0xFFFFu, bgd >> 16, 1u);                                          \
    const auto& a = tab[bw];                                                                               \
    (void)bid;                                                                                             \
    (void)gdim;

__global__ __launch_bounds__(256) void kb_temperature(const BTemperature* __restrict__ tab, const BatchBlock* __restrict__ map,
                                                      uint32_t step) {
    BATCH_BLOCK(tab)
    RiboSimParams P = a.P;
    P.epoch = step;
    b_temperature(bid, gdim, a.g, P, a.ep, a.vel, a.T, a.Q, a.t_out, a.r_out, a.eaten, 0, a.chem);
}
__global__ __launch_bounds__(256) void kb_forces(const BForces* __restrict__ tab, const BatchBlock* __restrict__ map,
                                                 uint32_t step) {
    BATCH_BLOCK(tab)
    RiboSimParams P = a.P;
    P.epoch = step;
    b_forces(bid, gdim, a.g, P, a.ep, a.T, a.forces, nullptr);
}
__global__ __launch_bounds__(256) void kb_add_forces(const BVel* __restrict__ tab, const BatchBlock* __restrict__ map) {
    BATCH_BLOCK(tab)
    b_add_forces(bid, gdim, a.g, a.P, a.aux, a.vin, a.vout);
}
__global__ __launch_bounds__(256) void kb_diffuse(const BVel* __restrict__ tab, const BatchBlock* __restrict__ map) {
    BATCH_BLOCK(tab)
    b_diffuse(bid, gdim, a.g, a.P, a.vin, a.vout);
}
__global__ __launch_bounds__(256) void kb_advect(const BVel* __restrict__ tab, const BatchBlock* __restrict__ map) {
    BATCH_BLOCK(tab)
    b_advect(bid, gdim, a.g, a.P, a.decay, a.vin, a.vout);
}
__global__ __launch_bounds__(256) void kb_vorticity(const BVel* __restrict__ tab, const BatchBlock* __restrict__ map) {
    BATCH_BLOCK(tab)
    b_vorticity(bid, gdim, a.g, a.P, a.vin, a.vout);
}
__global__ __launch_bounds__(256) void kb_divergence(const BProj* __restrict__ tab, const BatchBlock* __restrict__ map) {
    BATCH_BLOCK(tab)
    b_divergence(bid, gdim, a.g, a.vin, a.f0, a.f1);
}
__global__ __launch_bounds__(256) void kb_subgrad(const BProj* __restrict__ tab, const BatchBlock* __restrict__ map) {
    BATCH_BLOCK(tab)
    b_subgrad(bid, gdim, a.g, a.vin, a.vout, a.f0);
}
__global__ __launch_bounds__(NF_T) void kb_net_flux(const BProj* __restrict__ tab, const BatchBlock* __restrict__ map) {
    BATCH_BLOCK(tab)
    b_net_flux(bid, gdim, a.g, a.vin, a.f1);
}
__global__ __launch_bounds__(256) void kb_boundaries(const BProj* __restrict__ tab, const BatchBlock* __restrict__ map) {
    BATCH_BLOCK(tab)
    b_boundaries(bid, gdim, a.g, a.vin, a.vout, a.f1, a.vcopy);
}
__global__ __launch_bounds__(256) void kb_thermal(const BChem* __restrict__ tab, const BatchBlock* __restrict__ map,
                                                  uint32_t step) {
    BATCH_BLOCK(tab)
    RiboSimParams P = a.P;
    P.epoch = step;
    b_thermal(bid, gdim, a.g, P, a.ep, a.temp, a.chem);
}
__global__ __launch_bounds__(256) void kb_agg(const BChem* __restrict__ tab, const BatchBlock* __restrict__ map) {
    BATCH_BLOCK(tab)
    b_agg(bid, gdim, a.g, a.chem, a.agg_e);
}
__global__ __launch_bounds__(256) void kb_clear_next(const BChem* __restrict__ tab, const BatchBlock* __restrict__ map) {
    BATCH_BLOCK(tab)
    b_fill_u32(bid, gdim, a.next, (size_t)4 * a.g.G * a.g.G, 0u);
}
__global__ __launch_bounds__(256, 6) void kb_scatter(const BChem* __restrict__ tab, const BatchBlock* __restrict__ map,
                                                  uint32_t step) {
    BATCH_BLOCK(tab)
    RiboSimParams P = a.P;
    P.epoch = step;
    __shared__ ScLds lds;
    b_scatter<1>(&lds, 0u, threadIdx.x, threadIdx.y, bid, a.g, P, a.ep, a.chem, a.next, a.vel, a.temp, a.agg_e);
}
__global__ __launch_bounds__(256) void kb_ledger(const BChem* __restrict__ tab, const BatchBlock* __restrict__ map) {
    BATCH_BLOCK(tab)
    b_ledger(bid, gdim, a.chem, a.g.G * a.g.G, a.led);
}
__global__ __launch_bounds__(256) void kb_ledger_fin(const BChem* __restrict__ tab, const BatchBlock* __restrict__ map,
                                                     uint32_t step) {
    BATCH_BLOCK(tab)
    b_ledger_fin(bid, gdim, a.led, a.ledger_blocks, step, a.ep, 1);
}
__global__ __launch_bounds__(256) void kb_mg_down(const BMgPair* __restrict__ tab, const BatchBlock* __restrict__ map) {
    BATCH_BLOCK(tab)
    b_mg_down<float, false, 16>(bid, gdim, a.L, a.C, nullptr);
}
__global__ __launch_bounds__(256) void kb_mg_up(const BMgPair* __restrict__ tab, const BatchBlock* __restrict__ map) {
    BATCH_BLOCK(tab)
    b_mg_up<float, false, 16>(bid, gdim, a.L, a.C, nullptr);
}
__global__ __launch_bounds__(1024) void kb_mg_block(const MgBlockArgs<float>* __restrict__ tab,
                                                    const BatchBlock* __restrict__ map) {
    BATCH_BLOCK(tab)
    b_mg_block<float, false>(bid, gdim, a, nullptr);
}
// tolerance mode
template <class S>
__global__ __launch_bounds__(256) void kb_cg_init(const BCg<S>* __restrict__ tab, const BatchBlock* __restrict__ map) {
    BATCH_BLOCK(tab)
    b_cg_init<S>(bid, gdim, a.V, a.st);
}
template <class S>
__global__ __launch_bounds__(256) void kb_fin_init(const BCg<S>* __restrict__ tab, const BatchBlock* __restrict__ map) {
    BATCH_BLOCK(tab)
    b_pcg_fin<FIN_INIT>(a.st, a.nb16, a.tol, a.maxit, 0, 0u);
}
template <class S>
__global__ __launch_bounds__(256) void kb_shift(const BCg<S>* __restrict__ tab, const BatchBlock* __restrict__ map) {
    BATCH_BLOCK(tab)
    if (a.V.xold) b_pcg_shift(bid, a.V.n * a.V.n, a.V.fl, a.V.x, a.V.xlo, a.V.xold, a.V.r[0], a.V.pd, a.st);
}
template <class S>
__global__ __launch_bounds__(CG_NT) void kb_cg_down0(const BCg<S>* __restrict__ tab, const BatchBlock* __restrict__ map,
                                                    uint32_t j, int stage, int rd) {
    BATCH_BLOCK(tab)
    if (j > *a.nit) return;
    b_cg_down0<S>(bid, gdim, a.V, a.C, a.st, (int)(j & 1u), stage, rd, a.nt2, a.tol, a.maxit);
}
template <class S>
__global__ __launch_bounds__(256) void kb_mg_down_sym(const BMgPairT<S>* __restrict__ tab, const BatchBlock* __restrict__ map) {
    BATCH_BLOCK(tab)
    b_mg_down<S, true, 16>(bid, gdim, a.L, a.C, a.done);
}
template <class S>
__global__ __launch_bounds__(256) void kb_mg_up_sym(const BMgPairT<S>* __restrict__ tab, const BatchBlock* __restrict__ map) {
    BATCH_BLOCK(tab)
    b_mg_up<S, true, 16>(bid, gdim, a.L, a.C, a.done);
}
template <class S>
__global__ __launch_bounds__(1024) void kb_mg_block_sym(const BMgBlock<S>* __restrict__ tab,
                                                        const BatchBlock* __restrict__ map) {
    BATCH_BLOCK(tab)
    b_mg_block<S, true>(bid, gdim, a.a, a.done);
}
template <class S>
__global__ __launch_bounds__(CG_NT) void kb_cg_up0(const BCg<S>* __restrict__ tab, const BatchBlock* __restrict__ map,
                                                  uint32_t j) {
    BATCH_BLOCK(tab)
    if (j > *a.nit) return;
    b_cg_up0<S>(bid, gdim, a.V, a.C, a.st, (int)(j & 1u));
}
template <class S>
__global__ __launch_bounds__(256) void kb_cg_final(const BCg<S>* __restrict__ tab, const BatchBlock* __restrict__ map) {
    BATCH_BLOCK(tab)
    const int rd = (int)((*a.nit - 1u) & 1u);
    b_cg_final<S>(bid, gdim, a.V, a.st, rd, a.nt2, a.tol, a.maxit, 6u * a.nt2);
}
template <class S>
__global__ __launch_bounds__(256) void kb_fin_final(const BCg<S>* __restrict__ tab, const BatchBlock* __restrict__ map) {
    BATCH_BLOCK(tab)
    const int rd = (int)((*a.nit - 1u) & 1u);
    b_pcg_fin<FIN_FINAL>(a.st, a.nb16, 0.0f, 0u, 1 - rd, 6u * a.nt2);
}
#undef BATCH_BLOCK

// ---- the coarsest inverse under moving terrain (tolerance mode) --------------------------
// One workgroup of 1024 threads, one coarsest cell per thread (n^2 <= 1024).
// Returns at once while the coarsest fluid mask equals `snap` (the mask the
// inverse was built for); otherwise rebuilds the inverse exactly as the host
// build does: the unknowns are the fluid cells in index order; row r of A has
// 1 at each fluid 4-neighbour (left, right, down, up) and -(k + eps) on the
// diagonal; Gauss-Jordan with partial pivoting (the first row of largest
// |a[r][c]|), the pivot row divided by its pivot, every other row with a
// non-zero factor f updated as a[r][j] - f * a[c][j]; the inverse stored by
// columns (cinv[j * m + r]).  Each element sees the host's operations in the
// host's order, so the inverse is bit-identical to the host's.  More than
// `cap` fluid cells, or none, leaves cm = 0 (the coarsest level is smoothed).
__global__ __launch_bounds__(1024) void k_coarse_refresh(uint32_t n, float eps, const uint8_t* __restrict__ fl,
                                                         uint8_t* snap, uint32_t cap, double* a, double* inv,
                                                         double* cinv, int32_t* ccell, int* cm) {
    __shared__ int s_changed, s_piv;
    __shared__ int32_t s_idx[1024];
    __shared__ uint32_t s_wsum[16];
    __shared__ double s_bv[16];
    __shared__ int s_bi[16];
    __shared__ double s_d;
    __shared__ double s_f[1024];
    const int t = (int)threadIdx.x;
    const int nn = (int)(n * n);
    const uint8_t cur = (t < nn && fl[t]) ? 1u : 0u;
    if (t == 0) s_changed = 0;
    __syncthreads();
    if (t < nn && snap[t] != cur) atomicOr(&s_changed, 1);
    __syncthreads();
    if (!s_changed) return;
    // unknown index of each cell: exclusive scan of the fluid flags
    const uint64_t bal = __ballot(cur != 0);
    const int lane = t & 63, wv = t >> 6;
    const uint32_t below = (uint32_t)__popcll(bal & ((1ull << lane) - 1ull));
    if (lane == 0) s_wsum[wv] = (uint32_t)__popcll(bal);
    __syncthreads();
    uint32_t base = 0, m = 0;
    for (int k = 0; k < 16; ++k) {
        base += k < wv ? s_wsum[k] : 0u;
        m += s_wsum[k];
    }
    s_idx[t] = cur ? (int32_t)(base + below) : -1;
    __syncthreads();
    if (t < nn) snap[t] = cur;
    if (m == 0u || m > cap) {
        if (t == 0) *cm = 0;
        return;
    }
    const int M = (int)m;
    if (cur) ccell[s_idx[t]] = t;
    for (int k = t; k < M * M; k += 1024) {
        a[k] = 0.0;
        inv[k] = 0.0;
    }
    __syncthreads();
    if (cur) {
        const int r = s_idx[t], x = t % (int)n, y = t / (int)n;
        const int nb[4] = {x > 0 ? s_idx[t - 1] : -1, x + 1 < (int)n ? s_idx[t + 1] : -1,
                           y > 0 ? s_idx[t - (int)n] : -1, y + 1 < (int)n ? s_idx[t + (int)n] : -1};
        int k = 0;
        for (int d = 0; d < 4; ++d)
            if (nb[d] >= 0) {
                a[(size_t)r * M + nb[d]] = 1.0;
                ++k;
            }
        a[(size_t)r * M + r] = -((double)k + (double)eps);
        inv[(size_t)r * M + r] = 1.0;
    }
    __syncthreads();
    for (int c = 0; c < M; ++c) {
        // pivot: the first row r >= c of largest |a[r][c]|
        double v = -1.0;
        int vi = M;
        if (t >= c && t < M) {
            v = fabs(a[(size_t)t * M + c]);
            vi = t;
        }
        for (int off = 32; off > 0; off >>= 1) {
            const double ov = __shfl_down(v, off, 64);
            const int oi = __shfl_down(vi, off, 64);
            if (ov > v || (ov == v && oi < vi)) {
                v = ov;
                vi = oi;
            }
        }
        if (lane == 0) {
            s_bv[wv] = v;
            s_bi[wv] = vi;
        }
        __syncthreads();
        if (t == 0) {
            double bv = s_bv[0];
            int bi = s_bi[0];
            for (int k = 1; k < 16; ++k)
                if (s_bv[k] > bv || (s_bv[k] == bv && s_bi[k] < bi)) {
                    bv = s_bv[k];
                    bi = s_bi[k];
                }
            s_piv = bi;
        }
        __syncthreads();
        const int piv = s_piv;
        if (piv != c)
            for (int j = t; j < M; j += 1024) {
                const double x0 = a[(size_t)c * M + j], x1 = a[(size_t)piv * M + j];
                a[(size_t)c * M + j] = x1;
                a[(size_t)piv * M + j] = x0;
                const double y0 = inv[(size_t)c * M + j], y1 = inv[(size_t)piv * M + j];
                inv[(size_t)c * M + j] = y1;
                inv[(size_t)piv * M + j] = y0;
            }
        __syncthreads();
        if (t == 0) s_d = a[(size_t)c * M + c];
        __syncthreads();
        const double dg = s_d;
        for (int j = t; j < M; j += 1024) {
            a[(size_t)c * M + j] = a[(size_t)c * M + j] / dg;
            inv[(size_t)c * M + j] = inv[(size_t)c * M + j] / dg;
        }
        __syncthreads();
        // every row's factor before any row changes
        for (int r = t; r < M; r += 1024) s_f[r] = a[(size_t)r * M + c];
        __syncthreads();
        for (int k = t; k < M * M; k += 1024) {
            const int r = k / M, jj = k - r * M;
            const double fr = s_f[r];
            if (r == c || fr == 0.0) continue;
            a[k] = a[k] - fr * a[(size_t)c * M + jj];
            inv[k] = inv[k] - fr * inv[(size_t)c * M + jj];
        }
        __syncthreads();
    }
    for (int k = t; k < M * M; k += 1024) {
        const int r = k / M, jj = k - r * M;
        cinv[(size_t)jj * M + r] = inv[k];
    }
    if (t == 0) *cm = M;
}

inline dim3 grid2(uint32_t n) { return dim3((n + 15) / 16, (n + 15) / 16); }
const dim3 B16(16, 16);
inline uint32_t blocks1(size_t n, uint32_t per) { return (uint32_t)((n + per - 1) / per); }

}  // namespace

// ---- launchers -------------------------------------------------------------------------------
void launch_slope(hipStream_t s, const Geo& g, float2* slope) { k_slope<<<grid2(g.G), B16, 0, s>>>(g, slope); }
void launch_fluid_static(hipStream_t s, const Geo& g, uint8_t* cflags, float* fill, float* dragg, float2* slope_f) {
    k_fluid_static<<<grid2(g.F), B16, 0, s>>>(g, cflags, fill, dragg, slope_f);
}
void launch_fluid_normals(hipStream_t s, const Geo& g, uint8_t* cflags, float2* normal) {
    k_fluid_normals<<<grid2(g.F), B16, 0, s>>>(g, cflags, normal);
}
void launch_mg_mask0(hipStream_t s, const Geo& g, uint8_t* fluid0) {
    k_mg_mask0<<<blocks1((size_t)g.F * g.F, 256), 256, 0, s>>>(g, fluid0);
}
void launch_mg_mask_coarsen(hipStream_t s, uint32_t nc, const uint8_t* fine, uint8_t* coarse) {
    k_mg_mask_coarsen<<<blocks1((size_t)nc * nc, 256), 256, 0, s>>>(nc, fine, coarse);
}
void launch_mask_u32(hipStream_t s, const Geo& g, uint32_t* out) {
    k_mask_u32<<<blocks1((size_t)g.F * g.F, 256), 256, 0, s>>>(g, out);
}
void launch_temperature(hipStream_t s, const Geo& g, const RiboSimParams& P, const uint32_t* ep, const float2* vel,
                        const float* t_in, const float* r_in, float* t_out, float* r_out, uint32_t* eaten,
                        int eaten_active, const uint32_t* chem) {
    k_temperature<<<grid2(g.F), B16, 0, s>>>(g, P, ep, vel, t_in, r_in, t_out, r_out, eaten, eaten_active, chem);
}
void launch_forces(hipStream_t s, const Geo& g, const RiboSimParams& P, const uint32_t* ep, const float* temp,
                   float2* forces, uint32_t* agent) {
    k_forces<<<grid2(g.F), B16, 0, s>>>(g, P, ep, temp, forces, agent);
}
void launch_smooth_velocity(hipStream_t s, const Geo& g, const float2* vin, float2* vout) {
    k_smooth_velocity<<<grid2(g.F), B16, 0, s>>>(g, vin, vout);
}
void launch_add_forces(hipStream_t s, const Geo& g, const RiboSimParams& P, const float2* forces, const float2* vin,
                       float2* vout) {
    k_add_forces<<<grid2(g.F), B16, 0, s>>>(g, P, forces, vin, vout);
}
void launch_diffuse(hipStream_t s, const Geo& g, const RiboSimParams& P, const float2* vin, float2* vout) {
    k_diffuse<<<grid2(g.F), B16, 0, s>>>(g, P, vin, vout);
}
void launch_advect(hipStream_t s, const Geo& g, const RiboSimParams& P, float decay, const float2* vin, float2* vout) {
    k_advect<<<grid2(g.F), B16, 0, s>>>(g, P, decay, vin, vout);
}
void launch_vorticity(hipStream_t s, const Geo& g, const RiboSimParams& P, const float2* vin, float2* vout) {
    k_vorticity<<<grid2(g.F), B16, 0, s>>>(g, P, vin, vout);
}
void launch_divergence(hipStream_t s, const Geo& g, const float2* vin, float* div, float* p0) {
    k_divergence<<<grid2(g.F), B16, 0, s>>>(g, vin, div, p0);
}
void launch_jacobi(hipStream_t s, const Geo& g, const float* pin, float* pout, const float* div) {
    k_jacobi<<<grid2(g.F), B16, 0, s>>>(g, pin, pout, div);
}
void launch_mg_smooth(hipStream_t s, const MgLevelDev& L, uint32_t color, const uint32_t* done) {
    k_mg_smooth<<<grid2(L.n), B16, 0, s>>>(L, color, done);
}
void launch_mg_restrict(hipStream_t s, const MgLevelDev& L, const MgLevelDev& C, const uint32_t* done) {
    k_mg_restrict<<<grid2(C.n), B16, 0, s>>>(L, C, done);
}
void launch_mg_prolong(hipStream_t s, const MgLevelDev& L, const MgLevelDev& C, const uint32_t* done) {
    k_mg_prolong<<<grid2(L.n), B16, 0, s>>>(L, C, done);
}
template <class S>
void launch_mg_down(hipStream_t s, const MgLevelT<S>& L, const MgLevelT<S>& C, bool sym, const uint32_t* done, int tile) {
    if (tile ? tile == 16 : L.n <= (uint32_t)MG_SMALL) {
        const uint32_t nt = (L.n + 15) / 16;
        if (sym) k_mg_down<S, true, 16><<<dim3(nt, nt), 256, 0, s>>>(L, C, done);
        else k_mg_down<S, false, 16><<<dim3(nt, nt), 256, 0, s>>>(L, C, done);
    } else {
        const uint32_t nt = (L.n + 31) / 32;
        if (sym) k_mg_down<S, true, 32><<<dim3(nt, nt), 256, 0, s>>>(L, C, done);
        else k_mg_down<S, false, 32><<<dim3(nt, nt), 256, 0, s>>>(L, C, done);
    }
}
template <class S>
void launch_mg_up(hipStream_t s, const MgLevelT<S>& L, const MgLevelT<S>& C, bool sym, const uint32_t* done, int tile) {
    if (tile ? tile == 16 : L.n <= (uint32_t)MG_SMALL) {
        const uint32_t nt = (L.n + 15) / 16;
        if (sym) k_mg_up<S, true, 16><<<dim3(nt, nt), 256, 0, s>>>(L, C, done);
        else k_mg_up<S, false, 16><<<dim3(nt, nt), 256, 0, s>>>(L, C, done);
    } else {
        const uint32_t nt = (L.n + 31) / 32;
        if (sym) k_mg_up<S, true, 32><<<dim3(nt, nt), 256, 0, s>>>(L, C, done);
        else k_mg_up<S, false, 32><<<dim3(nt, nt), 256, 0, s>>>(L, C, done);
    }
}
// LDS of k_mg_block: unknowns and right-hand sides (binary32) and masks of
// every level, the staged coarsest right-hand side, and res / W of the top
// level for the bilinear restriction.
size_t mg_block_lds_bytes_total(size_t cells, size_t top_cells) {
    return ((cells * 9 + 3) & ~(size_t)3) + 4 * MG_SYM_COARSE_MAX + 4 * top_cells;
}
size_t mg_block_lds_bytes(const uint32_t* n, int count) {
    size_t cells = 0;
    for (int i = 0; i < count; ++i) cells += (size_t)n[i] * n[i];
    return mg_block_lds_bytes_total(cells, count > 0 ? (size_t)n[0] * n[0] : 0);
}
template <class S> void launch_zero_level(hipStream_t s, const MgLevelT<S>& L) {
    const uint32_t n = L.n * L.n;
    k_zero<S><<<(n + 255) / 256, 256, 0, s>>>(L.p, n);
}
template <class S>
MgBlockArgs<S> mg_block_args(const MgLevelT<S>* lv, int count, const double* cinv, const int32_t* ccell, int cm,
                             size_t* lds, const int* cm_dev) {
    MgBlockArgs<S> a{};
    a.levels = count;
    int off = 0;
    for (int i = 0; i < count; ++i) {
        a.n[i] = (int)lv[i].n;
        a.eps[i] = lv[i].eps;
        a.fluid[i] = lv[i].fluid;
        a.off[i] = off;
        off += (int)(lv[i].n * lv[i].n);
    }
    a.off[count] = off;
    a.p0 = lv[0].p;
    a.rhs0 = lv[0].rhs;
    a.cinv = cinv;
    a.ccell = ccell;
    a.cm = cm;
    a.cm_dev = cm_dev;
    *lds = mg_block_lds_bytes_total((size_t)off, (size_t)lv[0].n * lv[0].n);
    return a;
}
template <class S>
void launch_mg_block(hipStream_t s, const MgLevelT<S>* lv, int count, bool sym, const double* cinv, const int32_t* ccell,
                     int cm, const uint32_t* done, const int* cm_dev) {
    size_t lds = 0;
    const MgBlockArgs<S> a = mg_block_args<S>(lv, count, cinv, ccell, cm, &lds, cm_dev);
    if (sym) k_mg_block<S, true><<<1, 1024, lds, s>>>(a, done);
    else k_mg_block<S, false><<<1, 1024, lds, s>>>(a, done);
}
template MgBlockArgs<float> mg_block_args<float>(const MgLevelT<float>*, int, const double*, const int32_t*, int, size_t*,
                                                  const int*);
template MgBlockArgs<bf16> mg_block_args<bf16>(const MgLevelT<bf16>*, int, const double*, const int32_t*, int, size_t*,
                                                const int*);
inline uint32_t pcg_blocks(uint32_t n) { return ((n + 15) / 16) * ((n + 15) / 16); }
inline uint32_t cg_tiles(uint32_t n) { return (n + CG_T - 1) / CG_T; }
template <class S> void launch_cg_init(hipStream_t s, const CgVec<S>& V, PcgState* st, float tol, uint32_t maxit) {
    k_cg_init<S><<<grid2(V.n), B16, 0, s>>>(V, st);
    k_pcg_fin<FIN_INIT><<<1, 256, 0, s>>>(st, pcg_blocks(V.n), tol, maxit, 0, 0u);
    if (V.xold) k_pcg_shift<<<blocks1((size_t)V.n * V.n, 256), 256, 0, s>>>(V.n * V.n, V.fl, V.x, V.xlo, V.xold, V.r[0], V.pd, st);
}
template <class S>
void launch_cg_down0(hipStream_t s, const CgVec<S>& V, const MgLevelT<S>& C, PcgState* st, int parity, int stage, int rd,
                     float tol, uint32_t maxit) {
    const uint32_t nt = cg_tiles(V.n);
    k_cg_down0<S><<<dim3(nt, nt), CG_NT, 0, s>>>(V, C, st, parity & 1, stage, rd & 1, nt * nt, tol, maxit);
}
template <class S> void launch_cg_up0(hipStream_t s, const CgVec<S>& V, const MgLevelT<S>& C, PcgState* st, int parity) {
    const uint32_t nt = cg_tiles(V.n);
    k_cg_up0<S><<<dim3(nt, nt), CG_NT, 0, s>>>(V, C, st, parity & 1);
}
void launch_vc_step(hipStream_t s, int phase, uint32_t n, float eps, const uint8_t* fl, const float* x, const float* b,
                    PcgState* st, float tol, uint32_t maxit) {
    if (phase != 2) k_vc_resid<<<grid2(n), B16, 0, s>>>(n, eps, fl, x, b, st, phase == 0 ? 1 : 0);
    if (phase == 0) k_vc_fin<VC_BEGIN><<<1, 256, 0, s>>>(st, pcg_blocks(n), tol, maxit);
    else if (phase == 1) k_vc_fin<VC_CHECK><<<1, 256, 0, s>>>(st, pcg_blocks(n), tol, maxit);
    else k_vc_fin<VC_END><<<1, 256, 0, s>>>(st, 0u, tol, maxit);
}
template <class S> void launch_cg_final(hipStream_t s, const CgVec<S>& V, PcgState* st, int rd, float tol, uint32_t maxit) {
    const uint32_t nt = cg_tiles(V.n), off = 6u * nt * nt;
    k_cg_final<S><<<grid2(V.n), B16, 0, s>>>(V, st, rd & 1, nt * nt, tol, maxit, off);
    k_pcg_fin<FIN_FINAL><<<1, 256, 0, s>>>(st, pcg_blocks(V.n), 0.0f, 0u, 1 - (rd & 1), off);
}
#define RIBO_INSTANTIATE(S)                                                                                         \
    template void launch_mg_down<S>(hipStream_t, const MgLevelT<S>&, const MgLevelT<S>&, bool, const uint32_t*, int); \
    template void launch_mg_up<S>(hipStream_t, const MgLevelT<S>&, const MgLevelT<S>&, bool, const uint32_t*, int); \
    template void launch_zero_level<S>(hipStream_t, const MgLevelT<S>&);                                            \
    template void launch_mg_block<S>(hipStream_t, const MgLevelT<S>*, int, bool, const double*, const int32_t*, int, \
                                     const uint32_t*, const int*);                                                  \
    template void launch_cg_init<S>(hipStream_t, const CgVec<S>&, PcgState*, float, uint32_t);                     \
    template void launch_cg_down0<S>(hipStream_t, const CgVec<S>&, const MgLevelT<S>&, PcgState*, int, int, int,    \
                                     float, uint32_t);                                                             \
    template void launch_cg_up0<S>(hipStream_t, const CgVec<S>&, const MgLevelT<S>&, PcgState*, int);             \
    template void launch_cg_final<S>(hipStream_t, const CgVec<S>&, PcgState*, int, float, uint32_t);
RIBO_INSTANTIATE(float)
RIBO_INSTANTIATE(bf16)
#undef RIBO_INSTANTIATE
void launch_subgrad(hipStream_t s, const Geo& g, const float2* vin, float2* vout, const float* p) {
    k_subgrad<<<grid2(g.F), B16, 0, s>>>(g, vin, vout, p);
}
void launch_net_flux(hipStream_t s, const Geo& g, const float2* vin, float* nf) {
    uint32_t nb = (g.F + NF_C - 1) / NF_C;
    k_net_flux<<<2 * nb, NF_T, 0, s>>>(g, vin, nf);
}
void launch_boundaries(hipStream_t s, const Geo& g, const float2* vin, float2* vout, const float* nf, float2* vcopy) {
    k_boundaries<<<grid2(g.F), B16, 0, s>>>(g, vin, vout, nf, vcopy);
}
void launch_thermal(hipStream_t s, const Geo& g, const RiboSimParams& P, const uint32_t* ep, const float* temp,
                    uint32_t* chem) {
    k_thermal<<<grid2(g.G), B16, 0, s>>>(g, P, ep, temp, chem);
}
void launch_light_transmit(hipStream_t s, const Geo& g, const RiboSimParams& P, const uint32_t* chem, float* light) {
    k_light_transmit<<<blocks1((size_t)g.Lsz * g.Lsz, 256), 256, 0, s>>>(g, P, chem, light);
}
void launch_light_sweep(hipStream_t s, const Geo& g, const RiboSimParams& P, const uint32_t* ep, float* light) {
    const size_t lds = 2 * (size_t)g.Lsz * sizeof(float);  // grid <= 32768: Lsz <= 16384, 128 KB
    const uint32_t cpt = (g.Lsz + 1023u) / 1024u;
    if (cpt <= 1) k_light_sweep<1><<<2, 1024, lds, s>>>(g, P, ep, light);
    else if (cpt <= 2) k_light_sweep<2><<<2, 1024, lds, s>>>(g, P, ep, light);
    else if (cpt <= 4) k_light_sweep<4><<<2, 1024, lds, s>>>(g, P, ep, light);
    else if (cpt <= 8) k_light_sweep<8><<<2, 1024, lds, s>>>(g, P, ep, light);
    else k_light_sweep<16><<<2, 1024, lds, s>>>(g, P, ep, light);
}
void launch_light_propagate(hipStream_t s, const Geo& g, const RiboSimParams& P, const uint32_t* ep, const uint32_t* chem,
                            const float* lin, float* lout) {
    k_light_propagate<<<blocks1((size_t)g.Lsz * g.Lsz, 256), 256, 0, s>>>(g, P, ep, chem, lin, lout);
}
void launch_chem_ops(hipStream_t s, const Geo& g, uint32_t* chem, const uint32_t* ops, uint32_t n, uint32_t* done,
                     uint32_t* expected) {
    if (n == 0) return;
    k_chem_ops<<<(n + 255u) / 256u, 256, 0, s>>>(chem, g.gamma, g.G, (const uint4*)ops, n, done, expected);
}
void launch_relax_pass(hipStream_t s, const Geo& g, const RiboSimParams& P, const uint32_t* ep, uint32_t phase,
                       const float2* vel, uint32_t* gamma, uint32_t* chem, uint8_t* code, uint8_t* inc) {
    k_relax_decide<<<grid2(g.G), B16, 0, s>>>(g, P, ep, phase, vel, code);
    k_relax_apply<<<grid2(g.G), B16, 0, s>>>(g.G, gamma, code, inc);
    for (uint32_t c = 0; c < 9u; ++c) {
        const uint32_t cx = c % 3u, cy = c / 3u;
        if (cx >= g.G || cy >= g.G) continue;
        const uint32_t n = ((g.G - cx + 2u) / 3u) * ((g.G - cy + 2u) / 3u);
        k_relax_evict<<<(n + 255u) / 256u, 256, 0, s>>>(g.G, cx, cy, chem, gamma, inc);
    }
}
void launch_grain_fall(hipStream_t s, const Geo& g, const RiboSimParams& P, const uint32_t* ep, uint32_t* gamma,
                       uint32_t* chem) {
    for (uint32_t c = 0; c < 3u && c < g.G; ++c) {
        const uint32_t cols = (g.G - c + 2u) / 3u;
        k_grain_fall<<<(cols + 63u) / 64u, 64, 0, s>>>(g.G, g.F, P, ep, c, gamma, chem);
    }
}
void launch_agg(hipStream_t s, const Geo& g, const uint32_t* chem, uint32_t* agg_e) {
    k_agg<<<grid2(g.G), B16, 0, s>>>(g, chem, agg_e);
}
void launch_fill_u32(hipStream_t s, uint32_t* p, size_t n, uint32_t v) {
    k_fill_u32<<<blocks1(n, 1024), 256, 0, s>>>(p, n, v);
}
void launch_scatter(hipStream_t s, const Geo& g, const RiboSimParams& P, const uint32_t* ep, const uint32_t* chem,
                    uint32_t* next, const float2* vel, const float* temp, const uint32_t* agg_e) {
    k_scatter<<<grid2(g.G), B16, 0, s>>>(g, P, ep, chem, next, vel, temp, agg_e);
}
uint32_t scatter_tiles(uint32_t G) { return ((G + 15u) / 16u) * ((G + 15u) / 16u); }
template <class S> bool mg_block_scatter_fits(const MgLevelT<S>* lv, int count) {
    size_t lds = 0;
    (void)mg_block_args<S>(lv, count, nullptr, nullptr, 0, &lds, nullptr);
    return lds <= FuseCfg<S>::tiles * sizeof(ScLds);
}
template <class S>
void launch_mg_block_scatter(hipStream_t s, const MgLevelT<S>* lv, int count, const double* cinv, const int32_t* ccell, int cm,
                             const uint32_t* done, const int* cm_dev, const ScatterArgs& sc, uint32_t t0, uint32_t t1,
                             uint32_t cus) {
    size_t lds = 0;
    const MgBlockArgs<S> a = mg_block_args<S>(lv, count, cinv, ccell, cm, &lds, cm_dev);
    constexpr uint32_t NT = FuseCfg<S>::tiles;
    const uint32_t wgs = 1u + (t1 - t0 + NT - 1u) / NT;
    // with two workgroups to a CU, the V-cycle's is the last of the first
    // pass over the CUs, so that (with no more than 2 cus - 1 workgroups)
    // its CU takes no second one
    const uint32_t mgw = FuseCfg<S>::per_cu > 1 ? std::min(wgs, cus) - 1u : 0u;
    k_mg_block_scatter<S><<<wgs, NT * 256, 0, s>>>(a, done, sc, t0, t1, mgw);
}
void launch_scatter_range(hipStream_t s, const ScatterArgs& a, uint32_t t0, uint32_t t1) {
    if (t1 > t0) k_scatter_range<<<t1 - t0, B16, 0, s>>>(a, t0);
}
template bool mg_block_scatter_fits<float>(const MgLevelT<float>*, int);
template bool mg_block_scatter_fits<bf16>(const MgLevelT<bf16>*, int);
template void launch_mg_block_scatter<float>(hipStream_t, const MgLevelT<float>*, int, const double*, const int32_t*, int,
                                             const uint32_t*, const int*, const ScatterArgs&, uint32_t, uint32_t, uint32_t);
template void launch_mg_block_scatter<bf16>(hipStream_t, const MgLevelT<bf16>*, int, const double*, const int32_t*, int,
                                            const uint32_t*, const int*, const ScatterArgs&, uint32_t, uint32_t, uint32_t);
void launch_ledger(hipStream_t s, const Geo& g, const uint32_t* chem, LedgerDev* led, uint32_t epoch, const uint32_t* ep,
                   int check) {
    uint32_t cells = g.G * g.G;
    uint32_t nb = rb_minu(blocks1(cells, 256), LEDGER_MAX_BLOCKS);
    k_ledger<<<nb, 256, 0, s>>>(chem, cells, led);
    k_ledger_fin<<<1, 256, 0, s>>>(led, nb, epoch, ep, check);
}
void launch_ledger_cells(hipStream_t s, const uint32_t* chem, uint32_t cells, LedgerDev* led) {
    const uint32_t nb = rb_minu(blocks1(cells, 256), LEDGER_MAX_BLOCKS);
    k_ledger<<<nb, 256, 0, s>>>(chem, cells, led);
    k_ledger_fin<<<1, 256, 0, s>>>(led, nb, 0u, nullptr, 0);
}
void launch_coarse_refresh(hipStream_t s, uint32_t n, float eps, const uint8_t* fluid, uint8_t* snap, uint32_t cap,
                           double* work, double* cinv, int32_t* ccell, int* cm) {
    k_coarse_refresh<<<1, 1024, 0, s>>>(n, eps, fluid, snap, cap, work, work + (size_t)cap * cap, cinv, ccell, cm);
}
void launch_epoch_add(hipStream_t s, uint32_t* ep, uint32_t n) {
    k_epoch_add<<<1, 1, 0, s>>>(ep, n);
}
void launch_set_u32(hipStream_t s, uint32_t* p, uint32_t v) { k_set_u32<<<1, 1, 0, s>>>(p, v); }


// ---- batched launches (see BatchBlock) ----
void launch_batch(hipStream_t s, BatchKernel k, const void* tab, const BatchBlock* map, uint32_t blocks, uint32_t step,
                  size_t lds) {
    if (blocks == 0) return;
    switch (k) {
        case BK_TEMPERATURE: kb_temperature<<<blocks, B16, 0, s>>>((const BTemperature*)tab, map, step); break;
        case BK_FORCES: kb_forces<<<blocks, B16, 0, s>>>((const BForces*)tab, map, step); break;
        case BK_ADD_FORCES: kb_add_forces<<<blocks, B16, 0, s>>>((const BVel*)tab, map); break;
        case BK_DIFFUSE: kb_diffuse<<<blocks, B16, 0, s>>>((const BVel*)tab, map); break;
        case BK_ADVECT: kb_advect<<<blocks, B16, 0, s>>>((const BVel*)tab, map); break;
        case BK_VORTICITY: kb_vorticity<<<blocks, B16, 0, s>>>((const BVel*)tab, map); break;
        case BK_DIVERGENCE: kb_divergence<<<blocks, B16, 0, s>>>((const BProj*)tab, map); break;
        case BK_SUBGRAD: kb_subgrad<<<blocks, B16, 0, s>>>((const BProj*)tab, map); break;
        case BK_NET_FLUX: kb_net_flux<<<blocks, NF_T, 0, s>>>((const BProj*)tab, map); break;
        case BK_BOUNDARIES: kb_boundaries<<<blocks, B16, 0, s>>>((const BProj*)tab, map); break;
        case BK_THERMAL: kb_thermal<<<blocks, B16, 0, s>>>((const BChem*)tab, map, step); break;
        case BK_AGG: kb_agg<<<blocks, B16, 0, s>>>((const BChem*)tab, map); break;
        case BK_CLEAR_NEXT: kb_clear_next<<<blocks, 256, 0, s>>>((const BChem*)tab, map); break;
        case BK_SCATTER: kb_scatter<<<blocks, B16, 0, s>>>((const BChem*)tab, map, step); break;
        case BK_LEDGER: kb_ledger<<<blocks, 256, 0, s>>>((const BChem*)tab, map); break;
        case BK_LEDGER_FIN: kb_ledger_fin<<<blocks, 256, 0, s>>>((const BChem*)tab, map, step); break;
        case BK_MG_DOWN: kb_mg_down<<<blocks, 256, 0, s>>>((const BMgPair*)tab, map); break;
        case BK_MG_UP: kb_mg_up<<<blocks, 256, 0, s>>>((const BMgPair*)tab, map); break;
        case BK_MG_BLOCK: kb_mg_block<<<blocks, 1024, lds, s>>>((const MgBlockArgs<float>*)tab, map); break;
    }
}
template <class S>
void launch_batch_tol(hipStream_t s, BatchTolKernel k, const void* tab, const BatchBlock* map, uint32_t blocks, uint32_t j,
                      int stage, int rd, size_t lds) {
    if (blocks == 0) return;
    const BCg<S>* cg = (const BCg<S>*)tab;
    switch (k) {
        case BT_CG_INIT: kb_cg_init<S><<<blocks, B16, 0, s>>>(cg, map); break;
        case BT_FIN_INIT: kb_fin_init<S><<<blocks, 256, 0, s>>>(cg, map); break;
        case BT_SHIFT: kb_shift<S><<<blocks, 256, 0, s>>>(cg, map); break;
        case BT_CG_DOWN0: kb_cg_down0<S><<<blocks, CG_NT, 0, s>>>(cg, map, j, stage, rd & 1); break;
        case BT_MG_DOWN: kb_mg_down_sym<S><<<blocks, 256, 0, s>>>((const BMgPairT<S>*)tab, map); break;
        case BT_MG_BLOCK: kb_mg_block_sym<S><<<blocks, 1024, lds, s>>>((const BMgBlock<S>*)tab, map); break;
        case BT_MG_UP: kb_mg_up_sym<S><<<blocks, 256, 0, s>>>((const BMgPairT<S>*)tab, map); break;
        case BT_CG_UP0: kb_cg_up0<S><<<blocks, CG_NT, 0, s>>>(cg, map, j); break;
        case BT_CG_FINAL: kb_cg_final<S><<<blocks, B16, 0, s>>>(cg, map); break;
        case BT_FIN_FINAL: kb_fin_final<S><<<blocks, 256, 0, s>>>(cg, map); break;
    }
}
template void launch_batch_tol<float>(hipStream_t, BatchTolKernel, const void*, const BatchBlock*, uint32_t, uint32_t, int,
                                      int, size_t);
template void launch_batch_tol<bf16>(hipStream_t, BatchTolKernel, const void*, const BatchBlock*, uint32_t, uint32_t, int,
                                     int, size_t);
void batch_tol_grid(BatchTolKernel k, uint32_t n, uint32_t level_n, uint32_t* gx, uint32_t* gy) {
    *gy = 1;
    switch (k) {
        case BT_CG_INIT: case BT_CG_FINAL: *gx = *gy = (n + 15) / 16; break;
        case BT_FIN_INIT: case BT_FIN_FINAL: case BT_MG_BLOCK: *gx = 1; break;
        case BT_SHIFT: *gx = blocks1((size_t)n * n, 256); break;
        case BT_CG_DOWN0: case BT_CG_UP0: *gx = *gy = cg_tiles(n); break;
        case BT_MG_DOWN: case BT_MG_UP: *gx = *gy = (level_n + 15) / 16; break;
    }
}
void batch_grid(BatchKernel k, const Geo& g, uint32_t level_n, uint32_t* gx, uint32_t* gy) {
    *gy = 1;
    switch (k) {
        case BK_THERMAL: case BK_AGG: case BK_SCATTER: *gx = *gy = (g.G + 15) / 16; break;
        case BK_NET_FLUX: *gx = 2 * ((g.F + NF_C - 1) / NF_C); break;
        case BK_CLEAR_NEXT: *gx = blocks1((size_t)4 * g.G * g.G, 1024); break;
        case BK_LEDGER: *gx = rb_minu(blocks1((size_t)g.G * g.G, 256), LEDGER_MAX_BLOCKS); break;
        case BK_LEDGER_FIN: case BK_MG_BLOCK: *gx = 1; break;
        case BK_MG_DOWN: case BK_MG_UP: *gx = *gy = (level_n + 15) / 16; break;
        default: *gx = *gy = (g.F + 15) / 16; break;  // fluid kernels: 16 x 16 tiles
    }
}

// ---- scratch of the ledger that allocates nothing (ribo_ledger_noalloc) ----
// One ledger's per-block partials and totals (sizeof(LedgerDev), ~33 KB),
// resident on every device the library runs on until the process ends.  In
// host code `scratch` is only a stand-in for the symbol: its device address
// comes from hipGetSymbolAddress, once per device, and only that address is
// passed to kernels or copies.
static __device__ __align__(256) uint32_t scratch[sizeof(LedgerDev) / sizeof(uint32_t)];
static_assert(sizeof(LedgerDev) % sizeof(uint32_t) == 0, "the ledger is whole words");

hipError_t ledger_scratch(LedgerDev** out) {
    constexpr int MAX_DEVICES = 64;
    static std::mutex mu;
    static LedgerDev* ptr[MAX_DEVICES] = {};
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(mu);
    if (!ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        ptr[dev] = static_cast<LedgerDev*>(p);
    }
    *out = ptr[dev];
    return hipSuccess;
}

}  // namespace ribo
