// seed.cpp - the reference's host-side preparation of world inputs, on the
// host as in the reference (no device needed):
//   - the vents: rasterize_fumaroles (src/world/terrain.rs:316-336) and the
//     grid -> fluid box average of upload_heat (src/world/mod.rs:1366-1392),
//     which turn point fumaroles and heat / chemistry images into the
//     per-fluid-cell (heat, reductant) source the fluid step reads;
//   - the matter seeding: seed_cells with its fBm density fields
//     (src/world/mod.rs:1985-2070), fit_matter_to_terrain (:1970-1983) and
//     refit_matter_to_terrain (:1870-1968), the exact redistribution of
//     matter that no longer fits a changed terrain.
// Every float operation follows the reference's order in binary32, and the
// RNG is its SplitMix64, so the results are bit-identical to the reference's
// for the same inputs.  Compiled without contraction (-ffp-contract=off).

#include <math.h>
#include <stdint.h>

#include <algorithm>
#include <vector>

#include "../../include/ribo_world.h"

namespace {

constexpr uint32_t CHEM_CELL_CAP = 48u;  // src/shaders.rs:130
constexpr uint32_t GAMMA_SOLID = 3u;     // GAMMA_SOLID_THRESHOLD

// splitmix64 of the reference's CPU seeding (src/world/mod.rs:1842-1855).
struct SplitMix {
    uint64_t s;
    uint64_t next() {
        s += 0x9E3779B97F4A7C15ull;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    float f32() { return (float)(next() >> 40) / (float)(1u << 24); }
};

// Rust's `f as u32` / `as usize`: truncation, saturating at the ends, NaN -> 0.
uint64_t sat_u(float f, uint64_t hi) {
    if (!(f > 0.0f)) return 0;
    if (f >= (float)hi) return hi;
    return (uint64_t)f;
}

uint32_t cell_capacity(uint32_t g) {  // chem_capacity (chem.wgsl:29-33)
    return g >= GAMMA_SOLID ? 0u : CHEM_CELL_CAP * (GAMMA_SOLID - g) / GAMMA_SOLID;
}

// fBm of value noise on an m x m grid (src/world/mod.rs:2016-2049).
std::vector<float> fbm_field(size_t m, uint64_t seed) {
    auto lattice = [&](uint64_t o, int64_t ix, int64_t iy) {
        SplitMix h{seed ^ (o * 0x9E3779B97F4A7C15ull) ^ ((uint64_t)ix * 0xD6E8FEB86659FD93ull) ^
                   ((uint64_t)iy * 0xCA5A826395121157ull)};
        return h.f32();
    };
    std::vector<float> out(m * m, 0.0f);
    float amp = 0.5f, norm = 0.0f, freq = 4.0f;
    for (uint64_t o = 0; o < 6; ++o) {
        for (size_t y = 0; y < m; ++y)
            for (size_t x = 0; x < m; ++x) {
                const float px = (float)x / (float)m * freq;
                const float py = (float)y / (float)m * freq;
                const float fx = floorf(px), fy = floorf(py);
                const float tx = px - fx, ty = py - fy;
                const float sx = tx * tx * (3.0f - 2.0f * tx), sy = ty * ty * (3.0f - 2.0f * ty);
                const int64_t ix = (int64_t)fx, iy = (int64_t)fy;
                const float a = lattice(o, ix, iy) + (lattice(o, ix + 1, iy) - lattice(o, ix, iy)) * sx;
                const float b = lattice(o, ix, iy + 1) + (lattice(o, ix + 1, iy + 1) - lattice(o, ix, iy + 1)) * sx;
                out[y * m + x] += amp * (a + (b - a) * sy);
            }
        norm += amp;
        amp *= 0.65f;
        freq *= 2.0f;
    }
    for (float& v : out) v /= norm;
    return out;
}

float bilinear(const std::vector<float>& f, size_t m, float x, float y) {  // :2051-2062
    x = std::min(std::max(x, 0.0f), (float)(m - 1));
    y = std::min(std::max(y, 0.0f), (float)(m - 1));
    const size_t x0 = (size_t)floorf(x), y0 = (size_t)floorf(y);
    const size_t x1 = std::min(x0 + 1, m - 1), y1 = std::min(y0 + 1, m - 1);
    const float tx = x - (float)x0, ty = y - (float)y0;
    const float a = f[y0 * m + x0] + (f[y0 * m + x1] - f[y0 * m + x0]) * tx;
    const float b = f[y1 * m + x0] + (f[y1 * m + x1] - f[y1 * m + x0]) * tx;
    return a + (b - a) * ty;
}

bool cfg_ok(const RiboWorldConfig* c) {
    return c && c->fluid_size > 0 && c->grid_size >= c->fluid_size && c->grid_size % c->fluid_size == 0 &&
           c->world_units_per_cell > 0;
}

}  // namespace

extern "C" {

int ribo_rasterize_fumaroles(const RiboWorldConfig* cfg, const RiboFumarole* fum, uint32_t count, float* heat) {
    if (!cfg_ok(cfg) || !heat || (count > 0 && !fum)) return RIBO_ERR_ARG;
    const size_t n = cfg->grid_size;
    std::fill(heat, heat + n * n, 0.0f);
    const float cell_world = (float)cfg->world_units_per_cell;
    for (uint32_t i = 0; i < count; ++i) {
        const RiboFumarole& f = fum[i];
        if (f.enabled == 0 || !(f.strength > 0.0f)) continue;
        const float cx = f.x_frac * (float)n, cy = f.y_frac * (float)n;
        const float r = std::max(f.spread / cell_world, 1.0f);
        const size_t x0 = (size_t)sat_u(std::max(floorf(cx - r), 0.0f), SIZE_MAX);
        const size_t x1 = std::min((size_t)sat_u(ceilf(cx + r), SIZE_MAX), n);
        const size_t y0 = (size_t)sat_u(std::max(floorf(cy - r), 0.0f), SIZE_MAX);
        const size_t y1 = std::min((size_t)sat_u(ceilf(cy + r), SIZE_MAX), n);
        for (size_t y = y0; y < y1; ++y)
            for (size_t x = x0; x < x1; ++x) {
                const float dx = (float)x + 0.5f - cx, dy = (float)y + 0.5f - cy;
                const float d = sqrtf(dx * dx + dy * dy);
                if (d < r) {
                    const float w0 = 1.0f - d / r;
                    heat[y * n + x] += f.strength * w0 * w0 / RIBO_FUMAROLE_PIXEL_STRENGTH;
                }
            }
    }
    return RIBO_OK;
}

int ribo_vent_heat_src(const RiboWorldConfig* cfg, const float* heat_grid, const float* chem_grid, float gain,
                       float* out) {
    if (!cfg_ok(cfg) || !heat_grid || !chem_grid || !out) return RIBO_ERR_ARG;
    const size_t n = cfg->grid_size, f = cfg->fluid_size;
    const size_t k = std::max<size_t>(n / f, 1);
    const float scale = RIBO_FUMAROLE_PIXEL_STRENGTH * std::max(gain, 0.0f) / (float)(k * k);
    std::fill(out, out + f * f * 2, 0.0f);
    const size_t lim = std::min(n, f * k);
    for (size_t y = 0; y < lim; ++y)
        for (size_t x = 0; x < lim; ++x) {
            const size_t o = ((y / k) * f + x / k) * 2;
            out[o] += heat_grid[y * n + x] * scale;
            out[o + 1] += chem_grid[y * n + x] * scale;
        }
    return RIBO_OK;
}

int ribo_seed_cells(const RiboWorldConfig* cfg, uint64_t seed, float density, float active, uint32_t* cells) {
    if (!cfg_ok(cfg) || !cells) return RIBO_ERR_ARG;
    const size_t n = cfg->grid_size;
    SplitMix rng{seed};
    const uint32_t max_per_channel = CHEM_CELL_CAP / 4u;
    const size_t m = std::max<size_t>(n / 4, 2);
    std::vector<std::vector<float>> fields;
    for (uint64_t ch = 0; ch < 4; ++ch) fields.push_back(fbm_field(m, seed ^ ((ch + 1) * 0xA24BAED4963EE407ull)));
    const float dens = std::min(std::max(density, 0.0f), 1.0f), act_p = std::min(std::max(active, 0.0f), 1.0f);
    for (size_t y = 0; y < n; ++y)
        for (size_t x = 0; x < n; ++x) {
            const float fx = ((float)x + 0.5f) / (float)n * (float)m - 0.5f;
            const float fy = ((float)y + 0.5f) / (float)n * (float)m - 0.5f;
            for (size_t ch = 0; ch < 4; ++ch) {
                const float d = std::min(std::max(0.5f + 1.6f * (bilinear(fields[ch], m, fx, fy) - 0.5f), 0.0f), 1.0f);
                const float expect = d * ((float)max_per_channel - 2.0f) * dens;
                const uint32_t count = std::min((uint32_t)sat_u(expect + rng.f32(), UINT32_MAX), max_per_channel);
                uint32_t act = 0;
                for (uint32_t i = 0; i < count; ++i) act += rng.f32() < act_p ? 1u : 0u;
                cells[(y * n + x) * 4 + ch] = act | ((count - act) << 16);
            }
        }
    return RIBO_OK;
}

int ribo_fit_matter_to_terrain(uint32_t* cells, const uint32_t* gamma, uint64_t n, uint64_t seed) {
    if ((n > 0 && (!cells || !gamma))) return RIBO_ERR_ARG;
    SplitMix rng{seed ^ 0x0E170B0DE5EED5EDull};
    for (uint64_t i = 0; i < n; ++i) {
        const uint32_t g = gamma[i];
        if (g == 0) continue;
        const uint32_t per_ch = g >= GAMMA_SOLID ? 0u : CHEM_CELL_CAP * (GAMMA_SOLID - g) / GAMMA_SOLID / 4u;
        for (int c = 0; c < 4; ++c) {
            uint32_t act = 0;
            for (uint32_t k = 0; k < per_ch; ++k) act += rng.f32() < 0.5f ? 1u : 0u;
            cells[i * 4 + c] = act | ((per_ch - act) << 16);
        }
    }
    return RIBO_OK;
}

int ribo_refit_matter_to_terrain(uint32_t* cells, const uint32_t* gamma, uint64_t n) {
    if (n > 0 && (!cells || !gamma)) return RIBO_ERR_ARG;
    // kind k = channel * 2 + state (0 activated, 1 spent)
    auto get = [&](uint64_t i, int k) -> uint32_t {
        const uint32_t v = cells[i * 4 + k / 2];
        return k % 2 == 0 ? (v & 0xFFFFu) : (v >> 16);
    };
    auto add = [&](uint64_t i, int k, uint32_t d) { cells[i * 4 + k / 2] += k % 2 == 0 ? d : d << 16; };
    auto sub = [&](uint64_t i, int k, uint32_t d) { cells[i * 4 + k / 2] -= k % 2 == 0 ? d : d << 16; };
    auto total = [&](uint64_t i) {
        uint32_t t = 0;
        for (int k = 0; k < 8; ++k) t += get(i, k);
        return t;
    };
    // 1. the excess of each cell leaves in proportion to what it holds
    uint64_t carry[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    uint64_t room_total = 0;
    for (uint64_t i = 0; i < n; ++i) {
        const uint32_t cap = cell_capacity(gamma[i]);
        const uint32_t tot = total(i);
        if (tot <= cap) {
            room_total += cap - tot;
            continue;
        }
        const uint32_t over = tot - cap;
        uint32_t left = over;
        for (int k = 0; k < 8; ++k) {
            const uint32_t t = (uint32_t)((uint64_t)get(i, k) * over / tot);
            sub(i, k, t);
            carry[k] += t;
            left -= t;
        }
        int k = (int)(i % 8);
        while (left > 0) {
            if (get(i, k) > 0) {
                sub(i, k, 1);
                carry[k] += 1;
                left -= 1;
            }
            k = (k + 1) % 8;
        }
    }
    uint64_t pending = 0;
    for (int k = 0; k < 8; ++k) pending += carry[k];
    if (pending == 0) return RIBO_OK;
    // 2. every cell with room receives carry * room_i / room_total of each
    // kind (cumulative rounding, a phase per kind); what rounding leaves over
    // is `rest`
    uint64_t rest[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (room_total > 0) {
        uint64_t give[8], done[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int k = 0; k < 8; ++k) give[k] = std::min(carry[k], carry[k] * std::min(room_total, pending) / pending);
        uint64_t cum_room = 0;
        for (uint64_t i = 0; i < n; ++i) {
            const uint32_t cap = cell_capacity(gamma[i]);
            const uint32_t tot = total(i);
            if (tot >= cap) continue;
            uint32_t room = cap - tot;
            cum_room += room;
            for (int k = 0; k < 8; ++k) {
                const unsigned __int128 phase = (unsigned __int128)k * room_total / 8;
                const uint64_t target = (uint64_t)(((unsigned __int128)give[k] * cum_room + phase) / room_total);
                const uint32_t want = (uint32_t)(target - done[k]);
                done[k] = target;
                const uint32_t d = std::min(std::min(want, room), 0xFFFFu - get(i, k));
                add(i, k, d);
                room -= d;
                rest[k] += want - d;
            }
        }
        for (int k = 0; k < 8; ++k) rest[k] += carry[k] - give[k];
    } else {
        for (int k = 0; k < 8; ++k) rest[k] = carry[k];
    }
    // 3. the few left over: through the cells with room, one of each kind at a time
    uint64_t left = 0;
    for (int k = 0; k < 8; ++k) left += rest[k];
    uint64_t i = 0, idle = 0;
    int k = 0;
    while (left > 0 && idle < n) {
        const uint32_t cap = cell_capacity(gamma[i]);
        const uint32_t tot = total(i);
        if (tot < cap) {
            while (rest[k] == 0) k = (k + 1) % 8;
            if (get(i, k) < 0xFFFFu) {
                add(i, k, 1);
                rest[k] -= 1;
                left -= 1;
                idle = 0;
            }
            k = (k + 1) % 8;
        } else {
            idle += 1;
        }
        i = (i + 1) % n;
    }
    // no room anywhere: the rest stays in the first cell that is not rock,
    // over its capacity (transport spreads it later); nothing is lost
    if (left > 0) {
        uint64_t j = 0;
        while (j < n && gamma[j] >= GAMMA_SOLID) ++j;
        if (j == n) j = 0;
        for (int kk = 0; kk < 8; ++kk) add(j, kk, (uint32_t)rest[kk]);
    }
    return RIBO_OK;
}

}  // extern "C"
