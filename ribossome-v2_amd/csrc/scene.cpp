// scene.cpp - .ribo scene files (src/world/snapshot.rs): the LZ4 block codec,
// the JSON header and the world blocks.  Host code only.
//
// The LZ4 block format (public "LZ4 Block Format Description"): a block is
// a sequence of (token, literal-length bytes, literals, 2-byte offset,
// match-length bytes); the token holds the literal length (high 4 bits) and
// the match length - 4 (low 4 bits), 15 meaning "more bytes follow, each
// adding up to 255".  The last sequence has literals only; the last 5 bytes
// of a block are literals and no match starts in its last 12 bytes.

#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <map>
#include <new>
#include <memory>
#include <string>
#include <vector>

#include "../../include/ribo_params_default.h"
#include "../../include/ribo_scene.h"

namespace {

// ---- LZ4 block codec ------------------------------------------------------------
constexpr size_t LZ4_MIN_MATCH = 4, LZ4_LAST_LITERALS = 5, LZ4_MFLIMIT = 12, LZ4_MAX_OFFSET = 65535;

inline uint32_t read32(const uint8_t* p) {
    uint32_t v;
    memcpy(&v, p, 4);
    return v;
}
inline uint32_t hash4(uint32_t v) { return (v * 2654435761u) >> 16; }  // 16-bit Fibonacci hash

void put_length(std::vector<uint8_t>& out, size_t extra) {  // the bytes after a 15 in the token
    while (extra >= 255) {
        out.push_back(255);
        extra -= 255;
    }
    out.push_back((uint8_t)extra);
}

void emit_sequence(std::vector<uint8_t>& out, const uint8_t* lit, size_t nlit, size_t offset, size_t mlen) {
    const size_t ml = mlen ? mlen - LZ4_MIN_MATCH : 0;
    out.push_back((uint8_t)(((nlit < 15 ? nlit : 15) << 4) | (mlen ? (ml < 15 ? ml : 15) : 0)));
    if (nlit >= 15) put_length(out, nlit - 15);
    out.insert(out.end(), lit, lit + nlit);
    if (!mlen) return;
    out.push_back((uint8_t)(offset & 0xFF));
    out.push_back((uint8_t)(offset >> 8));
    if (ml >= 15) put_length(out, ml - 15);
}

// Greedy compression: a 64 Ki-entry table of the last position of each
// 4-byte hash, matches extended forward as far as the end rules allow.
std::vector<uint8_t> lz4_compress(const uint8_t* src, size_t n) {
    std::vector<uint8_t> out;
    out.reserve(n + n / 255 + 16);
    size_t anchor = 0;
    if (n >= LZ4_MFLIMIT + 1) {
        std::vector<int64_t> table(1u << 16, -1);
        size_t i = 0;
        while (i + LZ4_MFLIMIT <= n) {
            const uint32_t v = read32(src + i);
            const uint32_t h = hash4(v);
            const int64_t cand = table[h];
            table[h] = (int64_t)i;
            if (cand >= 0 && i - (size_t)cand <= LZ4_MAX_OFFSET && read32(src + cand) == v) {
                size_t mlen = LZ4_MIN_MATCH;
                while (i + mlen < n - LZ4_LAST_LITERALS && src[(size_t)cand + mlen] == src[i + mlen]) ++mlen;
                emit_sequence(out, src + anchor, i - anchor, i - (size_t)cand, mlen);
                i += mlen;
                anchor = i;
            } else {
                ++i;
            }
        }
    }
    emit_sequence(out, src + anchor, n - anchor, 0, 0);
    return out;
}

// Decompression with every read and write checked; -1 on a malformed block.
int64_t lz4_decompress(const uint8_t* src, size_t n, uint8_t* dst, size_t cap) {
    size_t ip = 0, op = 0;
    while (ip < n) {
        const uint8_t token = src[ip++];
        size_t lit = token >> 4;
        if (lit == 15) {
            uint8_t b;
            do {
                if (ip >= n) return -1;
                b = src[ip++];
                lit += b;
            } while (b == 255);
        }
        if (lit > n - ip || lit > cap - op) return -1;
        memcpy(dst + op, src + ip, lit);
        ip += lit;
        op += lit;
        if (ip == n) break;  // the last sequence: literals only
        if (n - ip < 2) return -1;
        const size_t off = (size_t)src[ip] | ((size_t)src[ip + 1] << 8);
        ip += 2;
        if (off == 0 || off > op) return -1;
        size_t ml = (size_t)(token & 15) + LZ4_MIN_MATCH;
        if ((token & 15) == 15) {
            uint8_t b;
            do {
                if (ip >= n) return -1;
                b = src[ip++];
                ml += b;
            } while (b == 255);
        }
        if (ml > cap - op) return -1;
        for (size_t k = 0; k < ml; ++k, ++op) dst[op] = dst[op - off];  // copies may overlap
    }
    return (int64_t)op;
}

// ---- JSON --------------------------------------------------------------------------
struct Json {
    enum Kind { NUL, BOOL, NUM, STR, ARR, OBJ } kind = NUL;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<Json> arr;
    std::vector<std::pair<std::string, Json>> obj;
    const Json* get(const std::string& k) const {
        if (kind != OBJ) return nullptr;
        for (const auto& kv : obj)
            if (kv.first == k) return &kv.second;
        return nullptr;
    }
};

struct JsonParser {
    const char* p;
    const char* e;
    bool ok = true;
    void ws() {
        while (p < e && (*p == ' ' || *p == '\n' || *p == '\r' || *p == '\t')) ++p;
    }
    bool lit(const char* s) {
        const size_t k = strlen(s);
        if ((size_t)(e - p) < k || strncmp(p, s, k) != 0) return false;
        p += k;
        return true;
    }
    std::string string() {
        std::string s;
        ++p;  // opening quote
        while (p < e && *p != '"') {
            if (*p == '\\') {
                if (++p >= e) break;
                const char c = *p++;
                if (c == 'n') s += '\n';
                else if (c == 't') s += '\t';
                else if (c == 'r') s += '\r';
                else if (c == 'b') s += '\b';
                else if (c == 'f') s += '\f';
                else if (c == 'u') {
                    if (e - p < 4) { ok = false; return s; }
                    const unsigned cp = (unsigned)strtoul(std::string(p, 4).c_str(), nullptr, 16);
                    p += 4;
                    if (cp < 0x80) s += (char)cp;
                    else if (cp < 0x800) { s += (char)(0xC0 | (cp >> 6)); s += (char)(0x80 | (cp & 0x3F)); }
                    else { s += (char)(0xE0 | (cp >> 12)); s += (char)(0x80 | ((cp >> 6) & 0x3F)); s += (char)(0x80 | (cp & 0x3F)); }
                } else s += c;  // \" \\ \/
            } else {
                s += *p++;
            }
        }
        if (p >= e) ok = false;
        else ++p;  // closing quote
        return s;
    }
    Json value(int depth = 0) {
        Json v;
        ws();
        if (p >= e || depth > 64) { ok = false; return v; }
        if (*p == '{') {
            v.kind = Json::OBJ;
            ++p;
            ws();
            if (p < e && *p == '}') { ++p; return v; }
            while (ok) {
                ws();
                if (p >= e || *p != '"') { ok = false; break; }
                std::string k = string();
                ws();
                if (p >= e || *p != ':') { ok = false; break; }
                ++p;
                v.obj.emplace_back(std::move(k), value(depth + 1));
                ws();
                if (p < e && *p == ',') { ++p; continue; }
                if (p < e && *p == '}') { ++p; break; }
                ok = false;
            }
        } else if (*p == '[') {
            v.kind = Json::ARR;
            ++p;
            ws();
            if (p < e && *p == ']') { ++p; return v; }
            while (ok) {
                v.arr.push_back(value(depth + 1));
                ws();
                if (p < e && *p == ',') { ++p; continue; }
                if (p < e && *p == ']') { ++p; break; }
                ok = false;
            }
        } else if (*p == '"') {
            v.kind = Json::STR;
            v.str = string();
        } else if (lit("true")) {
            v.kind = Json::BOOL;
            v.b = true;
        } else if (lit("false")) {
            v.kind = Json::BOOL;
        } else if (lit("null")) {
        } else {
            char* end = nullptr;
            const std::string tok(p, (size_t)std::min<ptrdiff_t>(e - p, 64));
            v.num = strtod(tok.c_str(), &end);
            if (end == tok.c_str()) { ok = false; return v; }
            v.kind = Json::NUM;
            p += end - tok.c_str();
        }
        return v;
    }
};

std::string num(double v) {  // shortest form that reads back to the same binary64
    char buf[40];
    for (int prec = 15; prec <= 17; ++prec) {
        snprintf(buf, sizeof buf, "%.*g", prec, v);
        if (strtod(buf, nullptr) == v) break;
    }
    return buf;
}

// ---- SimParams by name (src/params.rs:86-487, to_named / set_named) -------------
enum ParamType { PT_U32, PT_F32 };
struct ParamField { const char* name; size_t off; ParamType type; };
const ParamField PARAM_FIELDS[] = {
    {"epoch", offsetof(RiboSimParams, epoch), PT_U32},
    {"seed", offsetof(RiboSimParams, seed), PT_U32},
    {"dt", offsetof(RiboSimParams, dt), PT_F32},
    {"diffusion", offsetof(RiboSimParams, diffusion), PT_F32},
    {"settle", offsetof(RiboSimParams, settle), PT_F32},
    {"fluid_dt", offsetof(RiboSimParams, fluid_dt), PT_F32},
    {"fluid_decay", offsetof(RiboSimParams, fluid_decay), PT_F32},
    {"fluid_vorticity", offsetof(RiboSimParams, fluid_vorticity), PT_F32},
    {"fluid_viscosity", offsetof(RiboSimParams, fluid_viscosity), PT_F32},
    {"fumarole_count", offsetof(RiboSimParams, fumarole_count), PT_U32},
    {"uv_strength", offsetof(RiboSimParams, uv_strength), PT_F32},
    {"uv_depth", offsetof(RiboSimParams, uv_depth), PT_F32},
    {"fluid_enabled", offsetof(RiboSimParams, fluid_enabled), PT_U32},
    {"fluid_obstacle_strength", offsetof(RiboSimParams, fluid_obstacle_strength), PT_F32},
    {"slope_steer_rate", offsetof(RiboSimParams, slope_steer_rate), PT_F32},
    {"cohesion", offsetof(RiboSimParams, cohesion), PT_F32},
    {"spawn_count", offsetof(RiboSimParams, spawn_count), PT_U32},
    {"max_agents", offsetof(RiboSimParams, max_agents), PT_U32},
    {"death_probability", offsetof(RiboSimParams, death_probability), PT_F32},
    {"spawn_energy", offsetof(RiboSimParams, spawn_energy), PT_F32},
    {"food_power", offsetof(RiboSimParams, food_power), PT_F32},
    {"maintenance_cost", offsetof(RiboSimParams, maintenance_cost), PT_F32},
    {"pairing_rate", offsetof(RiboSimParams, pairing_rate), PT_F32},
    {"mutation_rate", offsetof(RiboSimParams, mutation_rate), PT_F32},
    {"uv_damage", offsetof(RiboSimParams, uv_damage), PT_F32},
    {"uptake_rate", offsetof(RiboSimParams, uptake_rate), PT_F32},
    {"brownian", offsetof(RiboSimParams, brownian), PT_F32},
    {"phoretic_gain", offsetof(RiboSimParams, phoretic_gain), PT_F32},
    {"chain_stiffness", offsetof(RiboSimParams, chain_stiffness), PT_F32},
    {"thermal_kt", offsetof(RiboSimParams, thermal_kt), PT_F32},
    {"motor_amplitude", offsetof(RiboSimParams, motor_amplitude), PT_F32},
    {"rft_enabled", offsetof(RiboSimParams, rft_enabled), PT_U32},
    {"require_start", offsetof(RiboSimParams, require_start), PT_U32},
    {"joint_coupling", offsetof(RiboSimParams, joint_coupling), PT_F32},
    {"hunger_regulation", offsetof(RiboSimParams, hunger_regulation), PT_U32},
    {"pairing_cost", offsetof(RiboSimParams, pairing_cost), PT_F32},
    {"reactivation_rate", offsetof(RiboSimParams, reactivation_rate), PT_F32},
    {"swim_gain", offsetof(RiboSimParams, swim_gain), PT_F32},
    {"bioturbation", offsetof(RiboSimParams, bioturbation), PT_F32},
    {"bioturbation_cost", offsetof(RiboSimParams, bioturbation_cost), PT_F32},
    {"monomer_uv_absorb", offsetof(RiboSimParams, monomer_uv_absorb), PT_F32},
    {"sedimentation", offsetof(RiboSimParams, sedimentation), PT_F32},
    {"direct_photoactivation", offsetof(RiboSimParams, direct_photoactivation), PT_F32},
    {"activation_decay", offsetof(RiboSimParams, activation_decay), PT_F32},
    {"monomer_pressure", offsetof(RiboSimParams, monomer_pressure), PT_F32},
    {"agent_fluid_push", offsetof(RiboSimParams, agent_fluid_push), PT_F32},
    {"swim_wobble", offsetof(RiboSimParams, swim_wobble), PT_F32},
    {"fluid_swim_only", offsetof(RiboSimParams, fluid_swim_only), PT_U32},
    {"motion_cost", offsetof(RiboSimParams, motion_cost), PT_F32},
    {"bond_rate", offsetof(RiboSimParams, bond_rate), PT_F32},
    {"bond_energy_share", offsetof(RiboSimParams, bond_energy_share), PT_F32},
    {"bond_signal", offsetof(RiboSimParams, bond_signal), PT_F32},
    {"aggregation", offsetof(RiboSimParams, aggregation), PT_F32},
    {"photo_yield", offsetof(RiboSimParams, photo_yield), PT_F32},
    {"heat_kill", offsetof(RiboSimParams, heat_kill), PT_F32},
    {"flow_coupling", offsetof(RiboSimParams, flow_coupling), PT_F32},
    {"inertia", offsetof(RiboSimParams, inertia), PT_F32},
    {"death_energy_cap", offsetof(RiboSimParams, death_energy_cap), PT_F32},
    {"chemo_yield", offsetof(RiboSimParams, chemo_yield), PT_F32},
    {"redox_decay", offsetof(RiboSimParams, redox_decay), PT_F32},
    {"thermal_activation", offsetof(RiboSimParams, thermal_activation), PT_F32},
    {"denature_temp", offsetof(RiboSimParams, denature_temp), PT_F32},
    {"day_period", offsetof(RiboSimParams, day_period), PT_F32},
    {"sun_slope", offsetof(RiboSimParams, sun_slope), PT_F32},
    {"sun_now", offsetof(RiboSimParams, sun_now), PT_F32},
    {"sun_angle", offsetof(RiboSimParams, sun_angle), PT_F32},
    {"day_fraction", offsetof(RiboSimParams, day_fraction), PT_F32},
    {"metabolic_q10", offsetof(RiboSimParams, metabolic_q10), PT_F32},
    {"metabolic_span", offsetof(RiboSimParams, metabolic_span), PT_F32},
    {"metabolic_ref", offsetof(RiboSimParams, metabolic_ref), PT_F32},
    {"sun_heat", offsetof(RiboSimParams, sun_heat), PT_F32},
    {"sediment_transport", offsetof(RiboSimParams, sediment_transport), PT_F32},
    {"sediment_threshold", offsetof(RiboSimParams, sediment_threshold), PT_F32},
    {"sediment_settle", offsetof(RiboSimParams, sediment_settle), PT_F32},
    {"signal_mode", offsetof(RiboSimParams, signal_mode), PT_F32},
    {"clock_mute", offsetof(RiboSimParams, clock_mute), PT_F32},
    {"paint_x", offsetof(RiboSimParams, paint_x), PT_F32},
    {"paint_y", offsetof(RiboSimParams, paint_y), PT_F32},
    {"paint_radius", offsetof(RiboSimParams, paint_radius), PT_F32},
    {"paint_grains", offsetof(RiboSimParams, paint_grains), PT_F32},
    {"rest_angle_mult", offsetof(RiboSimParams, rest_angle_mult), PT_F32},
    {"protease_power", offsetof(RiboSimParams, protease_power), PT_F32},
    {"lysis_yield", offsetof(RiboSimParams, lysis_yield), PT_F32},
    {"protease_direct", offsetof(RiboSimParams, protease_direct), PT_F32},
    {"joint_load", offsetof(RiboSimParams, joint_load), PT_F32},
    {"swim_grip", offsetof(RiboSimParams, swim_grip), PT_F32},
    {"swim_memory", offsetof(RiboSimParams, swim_memory), PT_F32},
    {"mouth_leak", offsetof(RiboSimParams, mouth_leak), PT_F32},
    {"leak_base", offsetof(RiboSimParams, leak_base), PT_F32},
    {"skin_uptake", offsetof(RiboSimParams, skin_uptake), PT_F32},
    {"death_metab", offsetof(RiboSimParams, death_metab), PT_F32},
    {"salvage", offsetof(RiboSimParams, salvage), PT_F32},
    {"bond_matter_share", offsetof(RiboSimParams, bond_matter_share), PT_F32},
    {"brownian_rot", offsetof(RiboSimParams, brownian_rot), PT_F32},
    {"transport_every", offsetof(RiboSimParams, transport_every), PT_U32},
    {"signal_crosstalk", offsetof(RiboSimParams, signal_crosstalk), PT_F32},
    {"flow_mass", offsetof(RiboSimParams, flow_mass), PT_F32},
    {"chemo_take", offsetof(RiboSimParams, chemo_take), PT_F32},
    {"vent_cycle_period", offsetof(RiboSimParams, vent_cycle_period), PT_F32},
    {"vent_cycle_amp", offsetof(RiboSimParams, vent_cycle_amp), PT_F32},
    {"sediment_cohesion", offsetof(RiboSimParams, sediment_cohesion), PT_F32},
    {"sediment_compaction", offsetof(RiboSimParams, sediment_compaction), PT_F32},
    {"rock_crumble", offsetof(RiboSimParams, rock_crumble), PT_F32},
    {"lysis_matter", offsetof(RiboSimParams, lysis_matter), PT_F32},
    {"scramble_conduction", offsetof(RiboSimParams, scramble_conduction), PT_U32},
    {"scramble_angles", offsetof(RiboSimParams, scramble_angles), PT_U32},
    {"scramble_response", offsetof(RiboSimParams, scramble_response), PT_U32},
    {"scramble_amount", offsetof(RiboSimParams, scramble_amount), PT_F32},
    {"pairing_energy_half", offsetof(RiboSimParams, pairing_energy_half), PT_F32},
    {"pairing_min_energy", offsetof(RiboSimParams, pairing_min_energy), PT_F32},
    {"uv_mutation", offsetof(RiboSimParams, uv_mutation), PT_F32},
    {"heat_mutation", offsetof(RiboSimParams, heat_mutation), PT_F32},
    {"birth_crowding", offsetof(RiboSimParams, birth_crowding), PT_F32},
    {"photo_damage", offsetof(RiboSimParams, photo_damage), PT_F32},
    {"naked_decay", offsetof(RiboSimParams, naked_decay), PT_F32},
    {"store_loss", offsetof(RiboSimParams, store_loss), PT_F32},
};

double param_get(const RiboSimParams& p, const ParamField& f) {
    const char* b = (const char*)&p + f.off;
    if (f.type == PT_U32) { uint32_t v; memcpy(&v, b, 4); return (double)v; }
    float v;
    memcpy(&v, b, 4);
    return (double)v;
}
void param_set(RiboSimParams& p, const ParamField& f, double v) {  // Rust `as` casts: saturating, NaN -> 0
    char* b = (char*)&p + f.off;
    if (f.type == PT_U32) {
        const uint32_t u = !(v > 0.0) ? 0u : (v >= 4294967295.0 ? 0xFFFFFFFFu : (uint32_t)v);
        memcpy(b, &u, 4);
    } else {
        const float x = (float)v;
        memcpy(b, &x, 4);
    }
}

// WorldSettings as serde names it (src/world/mod.rs:91-134); booleans are
// JSON booleans.
struct SettingField { const char* name; size_t off; bool boolean; };
const SettingField SETTING_FIELDS[] = {
    {"fluid_enabled", offsetof(RiboWorldSettings, fluid_enabled), true},
    {"fluid_substep", offsetof(RiboWorldSettings, fluid_substep), false},
    {"multigrid", offsetof(RiboWorldSettings, multigrid), true},
    {"mg_cycles", offsetof(RiboWorldSettings, mg_cycles), false},
    {"jacobi_iters", offsetof(RiboWorldSettings, jacobi_iters), false},
    {"light_rows_per_step", offsetof(RiboWorldSettings, light_rows_per_step), false},
    {"light_interval", offsetof(RiboWorldSettings, light_interval), false},
    {"terrain_enabled", offsetof(RiboWorldSettings, terrain_enabled), true},
    {"contact_enabled", offsetof(RiboWorldSettings, contact_enabled), true},
};
// WorldSettings::default() (src/world/mod.rs:117-134)
const uint32_t SETTING_DEFAULTS[] = {1u, 2u, 1u, 1u, 128u, 1u, 10u, 1u, 1u};

// The agent blocks of World::slot_bufs (src/world/snapshot.rs:212-236): a
// scene without agents holds each of them empty.
constexpr uint32_t AGENT_BYTES = 64;  // size_of::<Agent>() (src/params.rs:695-729)
const char* const AGENT_BLOCKS[] = {"agents", "genomes", "bodies", "organs", "rna_tails4", "bitten", "body_pos",
                                    "signals4", "joint_angle", "joint_base", "joint_state", "joint_active",
                                    "sensor_mem", "bonds2", "pair_mask"};
constexpr uint32_t SCENE_VERSION = 2;  // src/world/snapshot.rs:21
const char SCENE_MAGIC[8] = {'R', 'I', 'B', 'O', 'S', 'C', 'N', '1'};

// One block; false when it is too large for the u32 size in front (the
// format's limit: 4 GiB per block).
bool put_block(std::vector<uint8_t>& out, const std::string& name, const void* data, size_t bytes) {
    if (bytes > 0xFFFFFFFFull) return false;
    const std::vector<uint8_t> c = lz4_compress((const uint8_t*)data, bytes);
    const uint16_t nl = (uint16_t)name.size();
    const uint64_t cl = 4 + c.size();
    const uint32_t raw = (uint32_t)bytes;
    out.insert(out.end(), (const uint8_t*)&nl, (const uint8_t*)&nl + 2);
    out.insert(out.end(), name.begin(), name.end());
    out.insert(out.end(), (const uint8_t*)&cl, (const uint8_t*)&cl + 8);
    out.insert(out.end(), (const uint8_t*)&raw, (const uint8_t*)&raw + 4);
    out.insert(out.end(), c.begin(), c.end());
    return true;
}

struct GridSpec { const char* block; RiboBuffer buf; };
// The reference's world blocks, then this build's heat source (optional).
constexpr int NGRID = 7, NREQUIRED = 5;
const GridSpec GRIDS[NGRID] = {{"chem", RIBO_BUF_CHEM},         {"gamma", RIBO_BUF_GAMMA}, {"velocity", RIBO_BUF_VELOCITY},
                               {"pressure", RIBO_BUF_PRESSURE}, {"temperature", RIBO_BUF_TEMP}, {"redox", RIBO_BUF_REDOX},
                               {"heat_src", RIBO_BUF_HEAT_SRC}};
void** grid_ptr(RiboSceneGrids* g, int k) {
    void** p[NGRID] = {&g->chem, &g->gamma, &g->velocity, &g->pressure, &g->temperature, &g->redox, &g->heat_src};
    return p[k];
}

std::string cfg_json(const RiboWorldConfig& c) {  // cfg_json (src/world/snapshot.rs:187-195)
    return "{\"grid_size\": " + std::to_string(c.grid_size) + ", \"fluid_size\": " + std::to_string(c.fluid_size) +
           ", \"world_units_per_cell\": " + std::to_string(c.world_units_per_cell) +
           ", \"max_agents\": " + std::to_string(c.max_agents) + "}";
}

}  // namespace

namespace {


size_t lz4_bound_impl(size_t n) { return n + n / 255 + 16; }

int64_t lz4_compress_impl(const void* src, size_t n, void* dst, size_t dst_cap) {
    if ((!src && n) || !dst) return RIBO_ERR_ARG;
    const std::vector<uint8_t> c = lz4_compress((const uint8_t*)src, n);
    if (c.size() > dst_cap) return RIBO_ERR_ARG;
    memcpy(dst, c.data(), c.size());
    return (int64_t)c.size();
}

int64_t lz4_decompress_impl(const void* src, size_t n, void* dst, size_t dst_cap) {
    if ((!src && n) || (!dst && dst_cap)) return RIBO_ERR_ARG;
    const int64_t r = lz4_decompress((const uint8_t*)src, n, (uint8_t*)dst, dst_cap);
    return r < 0 ? RIBO_ERR_ARG : r;
}

int scene_write_agents_impl(const char* path, const RiboWorldConfig* cfg, const RiboSimParams* params,
                            const RiboWorldSettings* settings, const RiboSceneGrids* grids, uint32_t agents,
                            const RiboSceneBlock* blocks, uint32_t n_blocks);

int scene_write_impl(const char* path, const RiboWorldConfig* cfg, const RiboSimParams* params,
                     const RiboWorldSettings* settings, const RiboSceneGrids* grids) {
    return scene_write_agents_impl(path, cfg, params, settings, grids, 0, nullptr, 0);
}

int scene_write_agents_impl(const char* path, const RiboWorldConfig* cfg, const RiboSimParams* params,
                            const RiboWorldSettings* settings, const RiboSceneGrids* grids, uint32_t agents,
                            const RiboSceneBlock* blocks, uint32_t n_blocks) {
    if (!path || !cfg || !params || !settings || !grids || (n_blocks && !blocks)) return RIBO_ERR_ARG;
    if (agents > cfg->max_agents) return RIBO_ERR_ARG;
    // the caller's blocks: only the names the reference's scene carries
    // besides the world grids (slots, life_counters, the agent blocks), each once
    std::map<std::string, const RiboSceneBlock*> given;
    for (uint32_t i = 0; i < n_blocks; ++i) {
        if (!blocks[i].name || (blocks[i].bytes && !blocks[i].data)) return RIBO_ERR_ARG;
        const std::string nm(blocks[i].name);
        bool known = nm == "slots" || nm == "life_counters";
        for (const char* a : AGENT_BLOCKS) known = known || nm == a;
        if (!known || given.count(nm)) return RIBO_ERR_ARG;
        given[nm] = &blocks[i];
    }
    // what World::load_scene checks (src/world/snapshot.rs:403-405): one
    // slot word and one Agent (16 four-byte fields, src/params.rs:695-729)
    // per agent, 8 life counters
    auto given_bytes = [&](const char* nm) -> uint64_t {
        auto it = given.find(nm);
        return it == given.end() ? 0u : it->second->bytes;
    };
    if (given_bytes("slots") != 4ull * agents || given_bytes("agents") != (uint64_t)AGENT_BYTES * agents ||
        (given.count("life_counters") && given_bytes("life_counters") != 32u))
        return RIBO_ERR_ARG;
    // the header (src/world/snapshot.rs:325-343)
    std::string h = "{\n  \"versao\": " + std::to_string(SCENE_VERSION) + ",\n  \"gravado\": 0,\n  \"agentes\": " +
                    std::to_string(agents) + ",\n";
    h += "  \"cfg\": " + cfg_json(*cfg) + ",\n  \"params\": {";
    bool first = true;
    for (const ParamField& f : PARAM_FIELDS) {
        h += std::string(first ? "\n    \"" : ",\n    \"") + f.name + "\": " + num(param_get(*params, f));
        first = false;
    }
    h += "\n  },\n  \"settings\": {";
    first = true;
    for (const SettingField& f : SETTING_FIELDS) {
        uint32_t v;
        memcpy(&v, (const char*)settings + f.off, 4);
        h += std::string(first ? "\n    \"" : ",\n    \"") + f.name + "\": " +
             (f.boolean ? (v ? "true" : "false") : std::to_string(v));
        first = false;
    }
    h += "\n  },\n  \"fumarolas\": [],\n  \"ganho_fumarolas\": 1.0,\n  \"interface\": {}\n}";
    std::vector<uint8_t> out(SCENE_MAGIC, SCENE_MAGIC + 8);
    const uint32_t hl = (uint32_t)h.size();
    out.insert(out.end(), (const uint8_t*)&hl, (const uint8_t*)&hl + 4);
    out.insert(out.end(), h.begin(), h.end());
    uint32_t counters[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // life_counters of a world without agents
    auto put_given = [&](const char* name, const void* data, size_t bytes) {
        auto it = given.find(name);
        if (it != given.end()) return put_block(out, name, it->second->data, it->second->bytes);
        return put_block(out, name, data, bytes);
    };
    if (!put_given("slots", nullptr, 0) || !put_given("life_counters", counters, sizeof counters)) return RIBO_ERR_ARG;
    for (const char* name : AGENT_BLOCKS)
        if (!put_given(name, nullptr, 0)) return RIBO_ERR_ARG;
    RiboSceneGrids g = *grids;
    for (int k = 0; k < NGRID; ++k) {
        const void* data = *grid_ptr(&g, k);
        if (!data && k == NGRID - 1) continue;  // no heat source block
        if (!data || !put_block(out, GRIDS[k].block, data, ribo_buffer_bytes(cfg, GRIDS[k].buf))) return RIBO_ERR_ARG;
    }
    const std::string tmp = std::string(path) + ".tmp";
    FILE* f = fopen(tmp.c_str(), "wb");
    if (!f) return RIBO_ERR_ARG;
    const bool wrote = fwrite(out.data(), 1, out.size(), f) == out.size();
    if (fclose(f) != 0 || !wrote || rename(tmp.c_str(), path) != 0) {
        remove(tmp.c_str());
        return RIBO_ERR_ARG;
    }
    return RIBO_OK;
}

namespace {

// A scene file in memory: Scene::read (src/world/snapshot.rs:145-167).
struct SceneFile {
    std::vector<uint8_t> bytes;
    Json header;
    std::map<std::string, std::pair<size_t, size_t>> blocks;  // name -> (offset, length) of the compressed block

    bool open(const char* path) {
        FILE* f = fopen(path, "rb");
        if (!f) return false;
        uint8_t buf[1 << 16];
        size_t got;
        while ((got = fread(buf, 1, sizeof buf, f)) > 0) bytes.insert(bytes.end(), buf, buf + got);
        fclose(f);
        if (bytes.size() < 12 || memcmp(bytes.data(), SCENE_MAGIC, 8) != 0) return false;
        uint32_t hl;
        memcpy(&hl, bytes.data() + 8, 4);
        if ((size_t)hl > bytes.size() - 12) return false;
        JsonParser jp{(const char*)bytes.data() + 12, (const char*)bytes.data() + 12 + hl};
        header = jp.value();
        if (!jp.ok || header.kind != Json::OBJ) return false;
        size_t p = 12 + (size_t)hl;
        while (p < bytes.size()) {
            if (bytes.size() - p < 2) return false;
            uint16_t nl;
            memcpy(&nl, bytes.data() + p, 2);
            if (bytes.size() - p - 2 < (size_t)nl + 8) return false;
            const std::string name((const char*)bytes.data() + p + 2, nl);
            p += 2 + nl;
            uint64_t cl;
            memcpy(&cl, bytes.data() + p, 8);
            p += 8;
            if (cl > bytes.size() - p || cl < 4) return false;
            blocks[name] = {p, (size_t)cl};
            p += (size_t)cl;
        }
        return true;
    }
    // The size an LZ4 block says it holds (the 4 bytes lz4_flex prepends):
    // 0, 1 when absent, -1 when it exceeds what its compressed length can
    // expand to (an LZ4 sequence gives at most 255 bytes a byte, plus the
    // last literals), so a corrupt or crafted size is refused before any
    // allocation.
    int block_size(const std::string& name, uint32_t* raw) const {
        auto it = blocks.find(name);
        if (it == blocks.end()) return 1;
        memcpy(raw, bytes.data() + it->second.first, 4);
        const uint64_t most = 255ull * (uint64_t)(it->second.second - 4) + 16ull;
        return (uint64_t)*raw <= most ? 0 : -1;
    }
    // decompress_size_prepended of one block: 0, 1 when absent, -1 when malformed
    int block(const std::string& name, std::vector<uint8_t>* out) const {
        uint32_t raw = 0;
        const int sz = block_size(name, &raw);
        if (sz != 0) return sz;
        auto it = blocks.find(name);
        out->assign(raw, 0);
        const int64_t r = lz4_decompress(bytes.data() + it->second.first + 4, it->second.second - 4, out->data(), raw);
        return r == (int64_t)raw ? 0 : -1;
    }
};

}  // namespace

int scene_read_block_impl(const char* path, const char* name, void* dst, uint64_t cap, uint64_t* bytes) {
    if (!path || !name || !bytes || (cap && !dst)) return RIBO_ERR_ARG;
    SceneFile sf;
    if (!sf.open(path)) return RIBO_ERR_ARG;
    uint32_t raw = 0;
    if (sf.block_size(name, &raw) != 0) return RIBO_ERR_ARG;
    *bytes = raw;
    if (raw > cap) return cap == 0 ? RIBO_OK : RIBO_ERR_ARG;  // cap 0: the size query, nothing decompressed
    std::vector<uint8_t> data;
    const int rc = sf.block(name, &data);
    if (rc != 0) return RIBO_ERR_ARG;
    if (!data.empty()) memcpy(dst, data.data(), data.size());
    return RIBO_OK;
}

int scene_agents_impl(const char* path, uint32_t* agents) {
    if (!path || !agents) return RIBO_ERR_ARG;
    SceneFile sf;
    if (!sf.open(path)) return RIBO_ERR_ARG;
    const Json* a = sf.header.get("agentes");
    *agents = a && a->kind == Json::NUM && a->num > 0.0 ? (uint32_t)a->num : 0u;
    return RIBO_OK;
}

int scene_read_impl(const char* path, const RiboWorldConfig* cfg, RiboSimParams* params, RiboWorldSettings* settings,
                    RiboSceneGrids* grids) {
    if (!path || !cfg || !params || !settings || !grids) return RIBO_ERR_ARG;
    SceneFile sf;
    if (!sf.open(path)) return RIBO_ERR_ARG;
    const Json& header = sf.header;
    auto block = [&](const std::string& name, std::vector<uint8_t>* out) { return sf.block(name, out); };
    // World::load_scene's checks (src/world/snapshot.rs:370-400): the same
    // world size (a smaller max_agents in the file is accepted)
    const Json* c = header.get("cfg");
    if (!c) return RIBO_ERR_ARG;
    auto cu = [&](const char* k) -> double {
        const Json* v = c->get(k);
        return v && v->kind == Json::NUM ? v->num : -1.0;
    };
    if (cu("grid_size") != cfg->grid_size || cu("fluid_size") != cfg->fluid_size ||
        cu("world_units_per_cell") != cfg->world_units_per_cell || cu("max_agents") < 0.0 ||
        cu("max_agents") > (double)cfg->max_agents)
        return RIBO_ERR_ARG;
    std::vector<uint8_t> data[NGRID];
    for (int k = 0; k < NGRID; ++k) {
        const int rc = block(GRIDS[k].block, &data[k]);
        if (rc < 0) return RIBO_ERR_ARG;
        const size_t want = ribo_buffer_bytes(cfg, GRIDS[k].buf);
        if (k < NREQUIRED && (rc != 0 || data[k].size() != want)) return RIBO_ERR_ARG;
        if (k >= NREQUIRED && (rc != 0 || data[k].size() != want)) data[k].clear();  // optional: redox, heat
    }
    // parameters by name over SimParams::default() with the caller's seed
    RiboSimParams np;
    ribo_sim_params_fill_default(&np);
    np.seed = params->seed;
    if (const Json* pj = header.get("params"))
        for (const auto& kv : pj->obj)
            for (const ParamField& pf : PARAM_FIELDS)
                if (kv.first == pf.name && kv.second.kind == Json::NUM) param_set(np, pf, kv.second.num);
    // settings: serde(default) -- a missing field keeps WorldSettings::default()
    RiboWorldSettings ns = *settings;
    const Json* sj = header.get("settings");
    for (size_t k = 0; k < sizeof SETTING_FIELDS / sizeof SETTING_FIELDS[0]; ++k) {
        uint32_t v = SETTING_DEFAULTS[k];
        const Json* fv = sj ? sj->get(SETTING_FIELDS[k].name) : nullptr;
        if (fv && fv->kind == Json::BOOL) v = fv->b ? 1u : 0u;
        else if (fv && fv->kind == Json::NUM) v = fv->num > 0.0 ? (uint32_t)fv->num : 0u;
        memcpy((char*)&ns + SETTING_FIELDS[k].off, &v, 4);
    }
    for (int k = 0; k < NGRID; ++k) {
        void* dst = *grid_ptr(grids, k);
        if (!dst && !(k >= NREQUIRED && data[k].empty())) return RIBO_ERR_ARG;
    }
    for (int k = 0; k < NGRID; ++k)
        if (!data[k].empty()) memcpy(*grid_ptr(grids, k), data[k].data(), data[k].size());
    *params = np;
    *settings = ns;
    return RIBO_OK;
}

int world_save_scene_impl(RiboWorld* w, const char* path) {
    if (!w || !path) return RIBO_ERR_ARG;
    RiboWorldConfig cfg;
    RiboSimParams p;
    RiboWorldSettings s;
    int rc = ribo_world_get_config(w, &cfg);
    if (rc || (rc = ribo_world_get_params(w, &p)) || (rc = ribo_world_get_settings(w, &s))) return rc;
    std::vector<std::vector<uint8_t>> host(NGRID);
    RiboSceneGrids g{};
    for (int k = 0; k < NGRID; ++k) {
        host[k].resize(ribo_buffer_bytes(&cfg, GRIDS[k].buf));
        if ((rc = ribo_world_download(w, GRIDS[k].buf, host[k].data(), host[k].size()))) return rc;
        *grid_ptr(&g, k) = host[k].data();
    }
    return scene_write_impl(path, &cfg, &p, &s, &g);
}

int world_load_scene_impl(RiboWorld* w, const char* path) {
    if (!w || !path) return RIBO_ERR_ARG;
    RiboWorldConfig cfg;
    RiboSimParams p;
    RiboWorldSettings s;
    int rc = ribo_world_get_config(w, &cfg);
    if (rc || (rc = ribo_world_get_params(w, &p)) || (rc = ribo_world_get_settings(w, &s))) return rc;
    std::vector<std::vector<uint8_t>> host(NGRID);
    RiboSceneGrids g{};
    for (int k = 0; k < NGRID; ++k) {
        host[k].resize(ribo_buffer_bytes(&cfg, GRIDS[k].buf));
        *grid_ptr(&g, k) = host[k].data();
    }
    // keep the world's redox and heat source when the scene has none
    for (int k = NREQUIRED; k < NGRID; ++k)
        if ((rc = ribo_world_download(w, GRIDS[k].buf, host[k].data(), host[k].size()))) return rc;
    if ((rc = scene_read_impl(path, &cfg, &p, &s, &g))) return rc;
    if ((rc = ribo_world_set_params(w, &p)) || (rc = ribo_world_set_settings(w, &s))) return rc;
    // gamma first: the derived terrain follows it
    const int order[NGRID] = {1, 0, 2, 3, 4, 5, 6};
    for (int k : order)
        if ((rc = ribo_world_upload(w, GRIDS[k].buf, host[k].data(), host[k].size()))) return rc;
    return RIBO_OK;
}

}  // namespace

// Every entry point returns a status and never throws: a file whose sizes
// would need more memory than the host has gives RIBO_ERR_OOM, anything
// else unexpected RIBO_ERR_ARG.
#define RIBO_GUARD(expr)                                  \
    try {                                                 \
        return (expr);                                    \
    } catch (const std::bad_alloc&) {                     \
        return RIBO_ERR_OOM;                              \
    } catch (...) {                                       \
        return RIBO_ERR_ARG;                              \
    }

extern "C" {

size_t ribo_lz4_bound(size_t n) { return lz4_bound_impl(n); }
int64_t ribo_lz4_compress(const void* src, size_t n, void* dst, size_t dst_cap) {
    RIBO_GUARD(lz4_compress_impl(src, n, dst, dst_cap))
}
int64_t ribo_lz4_decompress(const void* src, size_t n, void* dst, size_t dst_cap) {
    RIBO_GUARD(lz4_decompress_impl(src, n, dst, dst_cap))
}
int ribo_scene_write(const char* path, const RiboWorldConfig* cfg, const RiboSimParams* params,
                     const RiboWorldSettings* settings, const RiboSceneGrids* grids) {
    RIBO_GUARD(scene_write_impl(path, cfg, params, settings, grids))
}
int ribo_scene_write_agents(const char* path, const RiboWorldConfig* cfg, const RiboSimParams* params,
                            const RiboWorldSettings* settings, const RiboSceneGrids* grids, uint32_t agents,
                            const RiboSceneBlock* blocks, uint32_t n_blocks) {
    RIBO_GUARD(scene_write_agents_impl(path, cfg, params, settings, grids, agents, blocks, n_blocks))
}
int ribo_scene_read_block(const char* path, const char* name, void* dst, uint64_t cap, uint64_t* bytes) {
    RIBO_GUARD(scene_read_block_impl(path, name, dst, cap, bytes))
}
int ribo_scene_agents(const char* path, uint32_t* agents) { RIBO_GUARD(scene_agents_impl(path, agents)) }
int ribo_scene_read(const char* path, const RiboWorldConfig* cfg, RiboSimParams* params, RiboWorldSettings* settings,
                    RiboSceneGrids* grids) {
    RIBO_GUARD(scene_read_impl(path, cfg, params, settings, grids))
}
int ribo_world_save_scene(RiboWorld* w, const char* path) { RIBO_GUARD(world_save_scene_impl(w, path)) }
int ribo_world_load_scene(RiboWorld* w, const char* path) { RIBO_GUARD(world_load_scene_impl(w, path)) }

}  // extern "C"

