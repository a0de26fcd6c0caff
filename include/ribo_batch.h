/*
 * ribo_batch.h - many worlds stepped together, one launch per kernel.
 *
 * A batch takes worlds that the caller created and loaded as usual
 * (ribo_world_create, ribo_world_upload, ...) and records their steps with
 * one launch of each kernel over the blocks of every world: each workgroup
 * finds its world and block in a map built once, and runs the same kernel
 * body as the world's own launch with that world's arguments, so every
 * world computes exactly what it computes alone.  Small worlds, which are
 * bound by dispatch when each is stepped on its own stream, then share
 * the launches of the large ones.
 *
 * The reference has no counterpart (it steps one World from its main
 * thread, src/world/mod.rs:1406-1632); a batch is the one-launch form of a
 * loop of World::encode_steps over independent worlds.
 *
 * The profile a batch runs: the multigrid pressure solve, either the
 * reference V-cycle (mg_tol 0) or the tolerance mode's conjugate gradients
 * (mg_tol > 0, one precision per batch; each world keeps its own iteration
 * slots, chosen every 8 steps as when it steps alone), with a frozen
 * terrain and light field, no pending agent words and no solar heating
 * under a moving sun.  Every world must share fluid_substep,
 * transport_every, the solver (mg_cycles, or the tolerance mode's
 * precision), ledger_check and the device; parameters such as the seed, dt,
 * the tolerance and the sizes may differ.  Worlds keep their own epochs and
 * may join at different epochs when they are at the same phase of the
 * fluid / transport schedule (any epochs with both every step).  While a
 * world is in a batch its parameters, settings and terrain must not change,
 * and it may step on its own only as the others do (all checked at every
 * call).  The worlds stay owned by the caller; after a batch's steps their
 * buffers are read with ribo_world_download as usual.
 */
#ifndef RIBO_BATCH_H
#define RIBO_BATCH_H

#include "ribo_world.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct RiboBatch RiboBatch;

/* Builds the block maps and argument tables of `count` worlds (at most
 * 4096, all on the current device).  RIBO_ERR_UNSUPPORTED when a world is
 * outside the profile above, RIBO_ERR_ARG when the worlds disagree. */
RIBO_API int ribo_batch_create(RiboWorld* const* worlds, uint32_t count, RiboBatch** out);
RIBO_API int ribo_batch_destroy(RiboBatch* b);
/* Records min(steps, RIBO_MAX_STEPS_PER_CALL) steps of every world on the
 * batch's stream (ordered after each world's queued work, and before any
 * work queued on a world's stream afterwards); returns the count. */
RIBO_API int ribo_batch_encode_steps(RiboBatch* b, uint32_t steps);
/* ribo_batch_encode_steps in chunks of RIBO_MAX_STEPS_PER_CALL. */
RIBO_API int ribo_batch_run(RiboBatch* b, uint32_t steps);
RIBO_API int ribo_batch_sync(RiboBatch* b);
/* 1 when every step recorded so far has finished, 0 while work runs,
 * negative on error. */
RIBO_API int ribo_batch_idle(RiboBatch* b);
/* Kernel launches recorded so far (the world-by-world path would record
 * about `count` times as many). */
RIBO_API int ribo_batch_launches(const RiboBatch* b, uint64_t* launches);
/* 1 when the world is inside the profile a batch runs, 0 when it is not
 * (ribo_last_error says why), negative on a bad argument. */
RIBO_API int ribo_batch_accepts(const RiboWorld* w);

#ifdef __cplusplus
}
#endif
#endif
