/*
 * mpsd_serve.h -- C-ABI of the ground-state serving loop on one device.
 *
 * Independent ground-state solves (mpsd_gs_config: model, L, D, parameters, tolerance,
 * precision) arrive over time.  A request is admitted when a worker slot is free and its
 * HBM footprint (MPS + environments + eigensolver basis + scratch, mpsd_gs_footprint) fits in
 * the budget next to the solves already running; admitted solves keep everything on the
 * device and run concurrently, one HIP stream each.  With batching on, the matvecs that
 * concurrent solves issue are merged into grouped persistent launches
 * (mpsd_heff_apply_grouped) on a shared stream; every result is bit-identical to the same
 * solve run alone.  The reference has no counterpart (one solve per process).
 */
#ifndef MPSD_SERVE_H
#define MPSD_SERVE_H

#include <stdint.h>

#include "mpsd_gs.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mpsd_server mpsd_server;

typedef struct mpsd_serve_config {
  int device;
  int slots;          /* solves running at once (worker contexts) */
  int64_t hbm_budget; /* bytes the admitted solves may hold together (0: 80% of free HBM) */
  int batching;       /* 1: merge concurrent matvecs into grouped launches */
  int window_us;      /* longest wait of a matvec for others to join its group */
  int graph;          /* 1: the workers' eigensolver steps replayed from captured hipGraphs
                         (MPSD_OPT_GRAPH on every worker context; solves whose matvecs go
                         through the batcher keep stream launches) */
} mpsd_serve_config;

typedef struct mpsd_serve_result {
  int id;
  int status;        /* MPSD_OK or the solve's error code */
  double energy;
  double arrival_s;  /* submit time, seconds since the server started */
  double admit_s;    /* admission time */
  double finish_s;   /* completion time */
  int halfsweeps;
  int64_t matvecs;
  int64_t matvecs_lowp;
  int64_t footprint_bytes;
  double matvec_flops;  /* algorithmic flops of the solve's matvecs (mpsd_gs_result.matvec_flops) */
} mpsd_serve_result;

typedef struct mpsd_serve_stats {
  int64_t solves;
  int64_t batches;          /* grouped matvec launches */
  int64_t batched_matvecs;  /* matvecs that went through them */
  int64_t max_batch;
  int64_t peak_admitted_bytes;
  int peak_running;
  int64_t graph_replays;    /* captured eigensolver-step graphs launched by the workers */
} mpsd_serve_stats;

/* HBM footprint estimate of one solve (bytes). */
int64_t mpsd_gs_footprint(const mpsd_gs_config* cfg);

int mpsd_server_create(const mpsd_serve_config* cfg, mpsd_server** out);
int mpsd_server_destroy(mpsd_server* s);
const char* mpsd_server_last_error(const mpsd_server* s);
/* Queue one solve now; *id identifies it. */
int mpsd_server_submit(mpsd_server* s, const mpsd_gs_config* cfg, int* id);
/* Block until solve `id` has finished and return its result. */
int mpsd_server_wait(mpsd_server* s, int id, mpsd_serve_result* out);
/* Replay a trace: request i is submitted arrival_s[i] seconds after the call starts
 * (non-decreasing); blocks until all have finished; out[i] is request i's result. */
int mpsd_server_run_trace(mpsd_server* s, int n, const mpsd_gs_config* cfgs, const double* arrival_s,
                          mpsd_serve_result* out);
int mpsd_server_stats(const mpsd_server* s, mpsd_serve_stats* out);

#ifdef __cplusplus
}
#endif

#endif /* MPSD_SERVE_H */
