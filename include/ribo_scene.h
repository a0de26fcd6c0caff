/*
 * ribo_scene.h - saved scenes (.ribo files) of the world step's state, in
 * the reference's format (src/world/snapshot.rs:1-11,99-185,274-556):
 *
 *   "RIBOSCN1", u32 header length, a JSON header, then blocks, each a u16
 *   name length, the name, a u64 compressed length and an LZ4 block with the
 *   uncompressed length in front (u32, the layout of lz4_flex's
 *   compress_prepend_size); every integer little-endian.
 *
 * The header carries the world configuration, SimParams by field name and
 * WorldSettings (serde field names).  The world blocks are chem, gamma,
 * velocity, pressure, temperature and redox (and heat_src, see below); a
 * scene written here holds no agents, with the agent blocks present and
 * empty, so the reference opens it as a world without agents.  Reading follows World::load_scene's rules:
 * the configuration must match (a smaller max_agents in the file is
 * accepted), parameters missing from the file keep SimParams::default()
 * (with the caller's seed), names the struct no longer has are ignored,
 * missing settings take WorldSettings::default().
 */
#ifndef RIBO_SCENE_H
#define RIBO_SCENE_H

#include "ribo_world.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Host buffers of a scene's world state; sizes as ribo_buffer_bytes of
 * RIBO_BUF_CHEM, GAMMA, VELOCITY, PRESSURE, TEMP, REDOX and HEAT_SRC. */
typedef struct RiboSceneGrids {
    void* chem;
    void* gamma;
    void* velocity;
    void* pressure;
    void* temperature;
    void* redox;  /* on read: left as it is when the scene has no redox block */
    void* heat_src; /* RIBO_BUF_HEAT_SRC, an extra block ("heat_src"): the
                       reference rebuilds its heat source from the fumaroles
                       and ignores the block; NULL on write leaves it out, on
                       read it is left as it is when the scene has none */
} RiboSceneGrids;

/* Writes a scene (World::save_scene without agents; the file appears
 * whole: it is written to "<path>.tmp" and renamed).  A block holds at most
 * 4 GiB - 1 (the format's u32 size): worlds of 16384^2 grid cells or more cannot be
 * saved (RIBO_ERR_ARG). */
RIBO_API int ribo_scene_write(const char* path, const RiboWorldConfig* cfg, const RiboSimParams* params,
                              const RiboWorldSettings* settings, const RiboSceneGrids* grids);
/* Reads a scene for a world of configuration cfg (Scene::read +
 * World::load_scene's checks): params->seed is the seed kept for parameters
 * the file lacks; the reference's settings fields are replaced, the
 * extension fields of *settings are kept.  RIBO_ERR_ARG on a malformed file,
 * another world size or wrong block sizes. */
RIBO_API int ribo_scene_read(const char* path, const RiboWorldConfig* cfg, RiboSimParams* params,
                             RiboWorldSettings* settings, RiboSceneGrids* grids);

/* A block a caller adds to a scene: one of the reference's blocks besides
 * the world grids ("slots", "life_counters" or an agent block: "agents",
 * "genomes", "bodies", "organs", "rna_tails4", "bitten", "body_pos",
 * "signals4", "joint_angle", "joint_base", "joint_state", "joint_active",
 * "sensor_mem", "bonds2", "pair_mask"), in the reference's own layout;
 * the library passes the bytes through. */
typedef struct RiboSceneBlock {
    const char* name;
    const void* data;
    uint64_t bytes;
} RiboSceneBlock;

/* ribo_scene_write with a caller's agents (World::save_scene,
 * src/world/snapshot.rs:274-363): the header's "agentes" count (at most
 * cfg->max_agents) and the blocks given; the blocks not given are written
 * empty (life_counters as eight zero words).  RIBO_ERR_ARG for an unknown or
 * repeated name, and for blocks the reference's load_scene would refuse
 * (snapshot.rs:403-405): "slots" must hold 4 bytes and "agents" 64 bytes
 * (one Agent) per agent, "life_counters" 32 bytes. */
RIBO_API int ribo_scene_write_agents(const char* path, const RiboWorldConfig* cfg, const RiboSimParams* params,
                                     const RiboWorldSettings* settings, const RiboSceneGrids* grids, uint32_t agents,
                                     const RiboSceneBlock* blocks, uint32_t n_blocks);
/* One block of a scene, decompressed (any name, the grids included):
 * *bytes receives its size; dst may be NULL with cap 0 to ask the size
 * (RIBO_OK, nothing decompressed).  RIBO_ERR_ARG for a malformed file (a
 * block whose stated size exceeds what its compressed bytes can expand to
 * is refused before any allocation), a missing block or 0 < cap < *bytes. */
RIBO_API int ribo_scene_read_block(const char* path, const char* name, void* dst, uint64_t cap, uint64_t* bytes);
/* The header's agent count ("agentes"). */
RIBO_API int ribo_scene_agents(const char* path, uint32_t* agents);

/* World::save_scene / Scene::read + World::load_scene on a world: its grids,
 * parameters and settings. */
RIBO_API int ribo_world_save_scene(RiboWorld* w, const char* path);
RIBO_API int ribo_world_load_scene(RiboWorld* w, const char* path);

/* The LZ4 block codec of the scene blocks (no size prefix): compress returns
 * the bytes written (dst_cap >= ribo_lz4_bound(n)), decompress the bytes
 * produced, or a negative RIBO_ERR_ARG for a malformed block or too small a
 * destination. */
RIBO_API size_t ribo_lz4_bound(size_t n);
RIBO_API int64_t ribo_lz4_compress(const void* src, size_t n, void* dst, size_t dst_cap);
RIBO_API int64_t ribo_lz4_decompress(const void* src, size_t n, void* dst, size_t dst_cap);

#ifdef __cplusplus
}
#endif

#endif
