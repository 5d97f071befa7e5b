/*
 * mpm3d_engine.h -- batched / served form of the MI355X step engine.
 *
 * The reference runs one problem per process (MPM3D.f90).  An engine holds
 * many independent instances in one device pool and advances all live
 * instances with one launch per phase per (shape, scheme) group; a step
 * sequence can be captured once as a HIP graph and replayed.  DT, time and
 * the end-of-run test live on the device, so instances of different sizes
 * and time steps stop on their own.  Instances can be added and removed
 * while others are resident (continuous batching).  None of these entry
 * points has a counterpart in the reference; they serve the ensemble and
 * serving drivers.
 */
#ifndef MPM3D_ENGINE_H
#define MPM3D_ENGINE_H

#include <stdint.h>
#include "mpm3d_problem.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mpm_engine mpm_engine;
typedef struct mpm_ctx mpm_ctx;

/* explicit capacities (particles, nodes, instances, material and body slots) */
int mpm_engine_create(long max_particles, long max_nodes, int max_instances, int max_materials, int max_bodies,
                      int precision, mpm_engine** out);
/* capacities derived from an HBM byte budget */
int mpm_engine_create_budget(double hbm_bytes, int max_instances, int precision, mpm_engine** out);
int mpm_engine_destroy(mpm_engine* e);
/* device bytes held by the pool */
long mpm_engine_device_bytes(mpm_engine* e);

/* admit an instance; MPM_ERR_NOMEM when it does not fit now.  With
 * stop_at_endtime the instance marks itself done once CurrentTime > endtime
 * (MPM3D.f90:152-158). */
int mpm_engine_add(mpm_engine* e, const mpm_problem* p, int stop_at_endtime, int* id);
int mpm_engine_remove(mpm_engine* e, int id);
/* stop an instance after max_steps steps (0: no limit), in addition to the
 * endtime test; a served request that asks for a fixed number of steps */
int mpm_engine_set_step_limit(mpm_engine* e, int id, int max_steps);
/* largest free particle / node ranges and free instance slots */
int mpm_engine_capacity(mpm_engine* e, long* free_particles, long* free_nodes, int* free_instances);

/* advance every live, unfinished instance by nsteps; with use_graph = 1 the
 * step sequence of graph_steps steps is captured once and replayed (captured
 * again after instances join or leave); use_graph = 2 (lazy) launches
 * plainly on the first call after such a change and captures only when the
 * next call finds the same instances (continuous batching) */
int mpm_engine_step(mpm_engine* e, int nsteps, int use_graph, int graph_steps);
int mpm_engine_sync(mpm_engine* e);
/* stable re-sort of every instance's particles by cell */
int mpm_engine_sort(mpm_engine* e);
/* steps advanced and whole-pool re-sorts run so far (either may be NULL) */
int mpm_engine_counters(mpm_engine* e, long* steps_done, long* sorts_done);
/* block sizes for particle and node kernels (64/128/256), re-sort period */
int mpm_engine_set_tuning(mpm_engine* e, int block_particles, int block_nodes, int sort_every);
/* interleaved (1) or contiguous (0) lane -> particle mapping in particle kernels */
int mpm_engine_set_interleave(mpm_engine* e, int enable);

/* Launch-shape settings.  sort_brick selects the cell order of the particle
 * sort ("cell bins" that keep a block's particles inside a compact box of
 * nodes): 0 the 3D Hilbert curve, B = 1, 2, 4, 8 brick-major order with
 * B^3-cell bricks.  Defaults: 256-particle and 256-node blocks, 2^3 bricks,
 * contiguous lanes, 256-particle scatter blocks, a re-sort every 64 steps.  step_ms / sort_ms / trials are filled by
 * mpm_engine_autotune and ignored on input. */
typedef struct mpm_tuning {
  int32_t block_particles;   /* 64, 128, 256 */
  int32_t block_nodes;       /* 64, 128, 256 */
  int32_t sort_every;        /* steps between re-sorts, 0 = never */
  int32_t sort_brick;        /* 0 (Hilbert), 1, 2, 4, 8 */
  int32_t interleave;        /* 0 / 1 */
  int32_t trials;            /* configurations timed by the autotuner */
  int32_t block_scatter;     /* particles per block of the particle-to-grid scatters (64, 128, 256; 0 keeps) */
  int32_t tile_nodes;        /* node capacity of the particle kernels' LDS tiles (8..512; 0 keeps) */
  double step_ms;            /* mean step time of the chosen shape, sort excluded */
  double sort_ms;            /* time of one re-sort */
} mpm_tuning;
int mpm_engine_get_tuning(mpm_engine* e, mpm_tuning* out);
int mpm_engine_apply_tuning(mpm_engine* e, const mpm_tuning* t);
/* Time launch shapes on the resident instances (trial_steps steps each,
 * from a device snapshot of the state that is restored
 * afterwards, so the simulation is unchanged), apply the fastest and
 * re-sort the pool in its cell order (a full re-sort period starts). */
int mpm_engine_autotune(mpm_engine* e, int trial_steps, mpm_tuning* best);

/* per-instance state (synchronising); layouts as in mpm3d_hip.h */
int mpm_engine_download_particles(mpm_engine* e, int id, double* pstate);
int mpm_engine_download_grid(mpm_engine* e, int id, double* gstate);
int mpm_engine_download_icell(mpm_engine* e, int id, int32_t* icell);
int mpm_engine_get_scalars(mpm_engine* e, int id, double* out8);
/* original particle index held by each device slot (the sort permutation) */
int mpm_engine_download_order(mpm_engine* e, int id, int32_t* orig);
/* per instance: particles whose stress update gave a negative volume
 * (a warning in the reference, Constitution.f90:135-137) and a non-finite
 * volume or stress, counted over all steps so far (-1: no such instance);
 * an instance's NaN stays in its own particle and node ranges, so the other
 * instances are unaffected (synchronising; either array may be NULL) */
int mpm_engine_numeric_flags(mpm_engine* e, int n, const int* ids, int* negvol, int* nonfinite);
/* done flag (1 finished, 0 running, -1 no such instance) and istep */
int mpm_engine_status(mpm_engine* e, int n, const int* ids, int* done, int* istep);
/* Non-blocking status: snapshot_status enqueues a copy of every instance's
 * time state behind the work queued so far and returns at once;
 * snapshot_read waits for that copy (not for work queued after it) and
 * returns per id the done flag and the 8 scalars of mpm_engine_get_scalars
 * as of the snapshot (out8 may be NULL).  A serving loop reads round k's
 * finished runs while round k+1 runs. */
int mpm_engine_snapshot_status(mpm_engine* e);
int mpm_engine_snapshot_read(mpm_engine* e, int n, const int* ids, int* done, double* out8);

/* Profiling: with enable != 0 every launch of the following steps is
 * bracketed by HIP events on the engine's stream; kernel_times returns the
 * accumulated milliseconds and launch counts per kernel class
 * (0 zero, 1 p2g, 2 bc, 3 stress, 4 integrate, 5 g2p, 6 finalize, 7 begin;
 * names from mpm_engine_kernel_name). */
int mpm_engine_profile(mpm_engine* e, int enable);
int mpm_engine_kernel_times(mpm_engine* e, double* total_ms8, long* launches8);
const char* mpm_engine_kernel_name(int kid);

/* the engine behind a single-problem context */
int mpm_ctx_engine(mpm_ctx* ctx, mpm_engine** out);

#ifdef __cplusplus
}
#endif
#endif
