/*
 * ribo_server.h - the serving form of the world step: scene requests are
 * admitted by a device-memory budget, their worlds stay resident in HBM,
 * each on its own HIP stream, and scheduler threads keep every admitted
 * world fed with chunks of up to 64 steps (ribo_world_encode_steps) as its
 * stream drains, so the work of scenes that arrive while others run
 * overlaps on the device.
 *
 * By default the server batches: scenes that fit the batch profile
 * (ribo_batch_accepts) and share its schedule are stepped together by one
 * scheduler thread, one launch per kernel for all of them (ribo_batch_*),
 * the batch rebuilt as scenes join and finish; the others keep their own
 * streams.  ribo_server_set_batching(s, 0) steps every scene on its own
 * stream.
 *
 * The reference has no counterpart: it steps one world from its main
 * thread (src/main.rs, src/world/mod.rs:1406-1632).  A scene here is what a
 * reference caller would build with World::new + uploads + encode_steps in
 * a loop; each scene's results are bit-identical to running its world
 * alone, as long as no tolerance-mode solve runs out of its iteration slots
 * (ribo_world_solver_stats counts those; the slots are chosen from the
 * world's own history either way).
 *
 * Threading: one caller thread at a time per server; the scheduler threads
 * are internal (each world is stepped by one of them only).  A scene's
 * buffers may be read once its state is DONE.
 */
#ifndef RIBO_SERVER_H
#define RIBO_SERVER_H

#include "ribo_world.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct RiboServer RiboServer;

/* One scene: a world configuration, its parameters and settings, the host
 * buffers to upload (NULL entries keep the world's zero state; GAMMA is
 * uploaded first, as the derived terrain depends on it), and the number of
 * steps to run.  The server copies the inputs at submission. */
typedef struct RiboScene {
    RiboWorldConfig cfg;
    uint32_t seed;
    uint32_t steps;
    RiboSimParams params;
    RiboWorldSettings settings;
    const void* input[RIBO_BUF_COUNT];
    size_t input_bytes[RIBO_BUF_COUNT];
} RiboScene;

enum { RIBO_SCENE_QUEUED = 0, RIBO_SCENE_RUNNING = 1, RIBO_SCENE_DONE = 2, RIBO_SCENE_FAILED = 3 };

typedef struct RiboSceneStatus {
    int32_t state;        /* RIBO_SCENE_* */
    int32_t error;        /* RIBO_ERR_* when FAILED, else 0 */
    uint32_t steps_done;  /* steps recorded so far */
    uint32_t _pad;
    double t_submit;      /* seconds since the server started */
    double t_admit;       /* 0 until admitted */
    double t_done;        /* 0 until every step has finished on the device */
    uint64_t footprint;   /* device bytes charged against the budget */
} RiboSceneStatus;

/* hbm_budget: device bytes the admitted scenes may hold together (each
 * scene is charged ribo_world_footprint); max_running: at most this many
 * scenes admitted at once (0 = no limit); threads: scheduler threads, each
 * stepping the scenes admitted to it (0 = half the host's hardware
 * threads, at most 8). */
RIBO_API int ribo_server_create(uint64_t hbm_budget, uint32_t max_running, uint32_t threads, RiboServer** out);
/* Stops the scheduler and frees every scene. */
RIBO_API int ribo_server_destroy(RiboServer* s);
/* on = 1 (the default): batch the scenes that fit a batch; 0: one stream
 * per scene.  Applies to scenes admitted afterwards. */
RIBO_API int ribo_server_set_batching(RiboServer* s, int on);
/* Queues a scene (first in, first admitted).  A scene larger than the whole
 * budget is refused with RIBO_ERR_OOM. */
RIBO_API int ribo_server_submit(RiboServer* s, const RiboScene* scene, uint64_t* ticket);
RIBO_API int ribo_server_status(RiboServer* s, uint64_t ticket, RiboSceneStatus* out);
/* The library's error message of a failed scene ("" otherwise), at most
 * cap - 1 bytes and a terminating zero. */
RIBO_API int ribo_server_error(RiboServer* s, uint64_t ticket, char* buf, size_t cap);
/* A finished scene's buffer (same ids and sizes as ribo_world_download). */
RIBO_API int ribo_server_download(RiboServer* s, uint64_t ticket, RiboBuffer which, void* dst, size_t bytes);
/* Frees a finished, failed or still queued scene and its share of the
 * budget; a running scene is refused (RIBO_ERR_ARG). */
RIBO_API int ribo_server_release(RiboServer* s, uint64_t ticket);
/* Counts of queued, running and finished (not yet released) scenes, and the
 * device bytes charged now. */
RIBO_API int ribo_server_stats(RiboServer* s, uint32_t* queued, uint32_t* running, uint32_t* finished, uint64_t* used);

#ifdef __cplusplus
}
#endif

#endif
