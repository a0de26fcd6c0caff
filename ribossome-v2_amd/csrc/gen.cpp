// gen.cpp - deterministic synthetic worlds for the benchmark configurations.
//
// Host-only code (no device needed).  Every generator is seeded SplitMix64
// (the reference's CPU seeding RNG, src/world/mod.rs:1842-1855) and writes
// the reference buffer layouts:
//   gamma     u32[G]   Gaussian blobs thresholded to an exact rock fraction
//                      (6 grains, the reference's ROCK, src/world/terrain.rs:11),
//                      with an optional 1-3 cell rubble shell of 1-2 grains;
//   velocity  f32x2[F] sum of sinusoid modes scaled to max |v| = vmax,
//                      zero in fluid wall cells;
//   chem      u32[4G]  8 Poisson(lambda) counts per cell (4 bases x
//                      activated/spent), 0 in rock, clamped to 0xFFFF;
//   heat      f32x2[F] a vent band (h, h) on the bottom fluid rows.

#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "../../include/ribo_world.h"

namespace {

struct SplitMix {
    uint64_t s;
    uint64_t next() {
        s += 0x9E3779B97F4A7C15ull;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double unit() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

bool cfg_ok(const RiboWorldConfig* c) {
    return c && c->fluid_size > 0 && c->grid_size >= c->fluid_size && c->grid_size % c->fluid_size == 0;
}

// Fluid wall test of build_solid_mask (fluid.wgsl:95-109,138-142,185-189):
// a fluid cell is a wall when at least 2 of its 4 grid samples are rock.
bool fluid_wall(const RiboWorldConfig* c, const uint32_t* gamma, uint32_t x, uint32_t y) {
    const uint32_t G = c->grid_size;
    const float scale = (float)G / (float)c->fluid_size;
    const float gx = ((float)x + 0.5f) * scale, gy = ((float)y + 0.5f) * scale, h = scale * 0.25f;
    int n = 0;
    for (int s = 0; s < 4; ++s) {
        float sx = gx + ((s & 1) ? h : -h), sy = gy + ((s & 2) ? h : -h);
        sx = std::min(std::max(sx, 0.0f), (float)(G - 1));
        sy = std::min(std::max(sy, 0.0f), (float)(G - 1));
        if (gamma[(uint32_t)sy * G + (uint32_t)sx] >= 3u) ++n;
    }
    return n >= 2;
}

}  // namespace

extern "C" {

int ribo_gen_gamma(const RiboWorldConfig* cfg, uint64_t seed, float rock_frac, int rubble, uint32_t* out) {
    if (!cfg_ok(cfg) || !out || !(rock_frac >= 0.0f && rock_frac < 1.0f)) return RIBO_ERR_ARG;
    const uint32_t G = cfg->grid_size;
    const size_t N = (size_t)G * G;
    SplitMix rng{seed ^ 0x6A09E667F3BCC908ull};
    std::vector<float> field(N, 0.0f);
    const uint32_t blobs = std::max<uint32_t>(8u, (uint32_t)(N / 4096u));
    for (uint32_t b = 0; b < blobs; ++b) {
        const double cx = rng.unit() * G, cy = rng.unit() * G;
        const double sigma = G / 128.0 + rng.unit() * (G / 32.0 - G / 128.0) + 1.0;
        const double amp = 0.5 + rng.unit();
        const int r = (int)ceil(3.0 * sigma);
        const int x0 = std::max(0, (int)cx - r), x1 = std::min((int)G - 1, (int)cx + r);
        const int y0 = std::max(0, (int)cy - r), y1 = std::min((int)G - 1, (int)cy + r);
        const double inv = 1.0 / (2.0 * sigma * sigma);
        for (int y = y0; y <= y1; ++y)
            for (int x = x0; x <= x1; ++x) {
                const double dx = x + 0.5 - cx, dy = y + 0.5 - cy;
                field[(size_t)y * G + x] += (float)(amp * exp(-(dx * dx + dy * dy) * inv));
            }
    }
    const size_t want = (size_t)llround((double)rock_frac * (double)N);
    memset(out, 0, N * sizeof(uint32_t));
    if (want > 0) {
        std::vector<uint32_t> order(N);
        for (size_t i = 0; i < N; ++i) order[i] = (uint32_t)i;
        // the `want` largest field values become rock (ties by index)
        std::nth_element(order.begin(), order.begin() + (want - 1), order.end(), [&](uint32_t a, uint32_t b) {
            return field[a] > field[b] || (field[a] == field[b] && a < b);
        });
        for (size_t i = 0; i < want; ++i) out[order[i]] = 6u;
    }
    if (rubble) {
        // Chebyshev distance to rock up to 3; each cell's shell width is 1-3.
        std::vector<uint8_t> dist(N, 255);
        for (size_t i = 0; i < N; ++i)
            if (out[i] >= 3u) dist[i] = 0;
        for (int pass = 1; pass <= 3; ++pass) {
            std::vector<uint8_t> nd = dist;
            for (uint32_t y = 0; y < G; ++y)
                for (uint32_t x = 0; x < G; ++x) {
                    size_t i = (size_t)y * G + x;
                    if (dist[i] != 255) continue;
                    bool near = false;
                    for (int dy = -1; dy <= 1 && !near; ++dy)
                        for (int dx = -1; dx <= 1; ++dx) {
                            int nx = (int)x + dx, ny = (int)y + dy;
                            if (nx < 0 || ny < 0 || nx >= (int)G || ny >= (int)G) continue;
                            if (dist[(size_t)ny * G + nx] == pass - 1) { near = true; break; }
                        }
                    if (near) nd[i] = (uint8_t)pass;
                }
            dist.swap(nd);
        }
        SplitMix rr{seed ^ 0xBB67AE8584CAA73Bull};
        for (size_t i = 0; i < N; ++i) {
            const uint64_t h = rr.next();
            const uint32_t width = 1u + (uint32_t)(h % 3u);
            if (dist[i] >= 1 && dist[i] != 255 && dist[i] <= width) out[i] = dist[i] == 1 ? 2u : 1u;
        }
    }
    return RIBO_OK;
}

int ribo_gen_velocity(const RiboWorldConfig* cfg, uint64_t seed, uint32_t modes, float vmax, const uint32_t* gamma,
                      float* out) {
    if (!cfg_ok(cfg) || !out) return RIBO_ERR_ARG;
    const uint32_t F = cfg->fluid_size;
    SplitMix rng{seed ^ 0x3C6EF372FE94F82Bull};
    struct Mode { double kx, ky, pu, pv, au, av; };
    std::vector<Mode> m(modes);
    for (auto& md : m) {
        md.kx = 1 + (double)(rng.next() % 8);
        md.ky = 1 + (double)(rng.next() % 8);
        md.pu = rng.unit() * 2 * M_PI;
        md.pv = rng.unit() * 2 * M_PI;
        md.au = 0.5 + 0.5 * rng.unit();
        md.av = 0.5 + 0.5 * rng.unit();
    }
    std::vector<double> u((size_t)F * F * 2, 0.0);
    double vm = 0.0;
    for (uint32_t y = 0; y < F; ++y)
        for (uint32_t x = 0; x < F; ++x) {
            double a = 0, b = 0;
            for (const auto& md : m) {
                const double ph = 2 * M_PI * (md.kx * (x + 0.5) + md.ky * (y + 0.5)) / F;
                a += md.au * sin(ph + md.pu);
                b += md.av * cos(ph + md.pv);
            }
            const bool wall = gamma && fluid_wall(cfg, gamma, x, y);
            if (wall) a = b = 0;
            u[2 * ((size_t)y * F + x)] = a;
            u[2 * ((size_t)y * F + x) + 1] = b;
            vm = std::max(vm, sqrt(a * a + b * b));
        }
    const double s = vm > 0 ? (double)vmax / vm : 0.0;
    for (size_t i = 0; i < u.size(); ++i) out[i] = (float)(u[i] * s);
    return RIBO_OK;
}

int ribo_gen_chem(const RiboWorldConfig* cfg, uint64_t seed, float lambda, const uint32_t* gamma, uint32_t* out) {
    if (!cfg_ok(cfg) || !out || !(lambda >= 0.0f)) return RIBO_ERR_ARG;
    const size_t N = (size_t)cfg->grid_size * cfg->grid_size;
    SplitMix rng{seed ^ 0xA54FF53A5F1D36F1ull};
    const double e = exp(-(double)lambda);
    for (size_t i = 0; i < N; ++i) {
        uint32_t c[8];
        for (int k = 0; k < 8; ++k) {
            // inverse-CDF Poisson draw
            const double uu = rng.unit();
            uint32_t n = 0;
            double p = e, cdf = e;
            while (uu > cdf && n < 0xFFFFu) {
                ++n;
                p *= (double)lambda / n;
                cdf += p;
                if (p < 1e-300) break;
            }
            c[k] = n;
        }
        const bool rock = gamma && gamma[i] >= 3u;
        for (int ch = 0; ch < 4; ++ch) {
            const uint32_t act = rock ? 0u : std::min(c[2 * ch], 0xFFFFu);
            const uint32_t spent = rock ? 0u : std::min(c[2 * ch + 1], 0xFFFFu);
            out[4 * i + ch] = act | (spent << 16);
        }
    }
    return RIBO_OK;
}

int ribo_gen_vent_heat(const RiboWorldConfig* cfg, uint32_t rows, float h, float* out) {
    if (!cfg_ok(cfg) || !out) return RIBO_ERR_ARG;
    const uint32_t F = cfg->fluid_size;
    for (uint32_t y = 0; y < F; ++y)
        for (uint32_t x = 0; x < F; ++x) {
            const float v = y < rows ? h : 0.0f;
            out[2 * ((size_t)y * F + x)] = v;
            out[2 * ((size_t)y * F + x) + 1] = v;
        }
    return RIBO_OK;
}

}  // extern "C"
