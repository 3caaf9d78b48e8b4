/*
 * ftl_server.h -- serving loop of the FTL step (part of libftl.so).
 *
 * No counterpart in the reference, which steps one groom per frame from the
 * render thread (Renderer::update -> Hair::update, src/core/renderer.cc:69-81,
 * src/fx/hair.cc:89-125).  Here many creatures stay resident on the device
 * and step requests arrive asynchronously:
 *
 *   - Admission by HBM budget: a creature added to the server is made
 *     resident (its positions, velocities and rest roots are copied to the
 *     device) only if the resident bytes stay within the budget; otherwise it
 *     waits, and is admitted, in order of arrival, when a removal frees memory.
 *   - Continuous batching: a dispatcher thread collects the pending step
 *     requests of resident creatures and launches them as one batched
 *     step (the kernels of ftl_batch_step), up to a per-batch cap
 *     on HBM traffic (64 B per control point).  Requests that arrive while
 *     batches run join the next batch as soon as one of the in-flight slots
 *     frees; up to `max_inflight` batches overlap on the device.
 *   - A creature has at most one step in flight; its requests run in the
 *     order they were submitted.  A step's inputs (dt, time, root transform,
 *     colliders) travel with the request, so a creature's results do not
 *     depend on which other creatures share its batches.
 *   - Latency of a request = time from ftl_server_submit to the host seeing
 *     its batch complete.  p50/p99 over the requests since the last reset.
 *
 * Threading: every ftl_server_* call may be made from any host thread; the
 * server serialises them internally.
 */
#ifndef FTL_SERVER_H_
#define FTL_SERVER_H_

#include "ftl.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ftl_server ftl_server;

typedef struct ftl_server_config {
  uint64_t hbm_budget_bytes; /* device bytes resident creatures may hold; 0 = 80% of free device memory */
  uint64_t batch_cap_bytes;  /* HBM traffic per batch (64 B per control point); 0 = 8 GB */
  int max_inflight;          /* batches on the device at once, 1..4; 0 = 2 */
  /* Coalescing window, microseconds (0 = none: a batch launches as soon as a
     slot is free).  A batch is launched only once its requests reach
     batch_target_bytes of traffic or its oldest request has waited
     coalesce_us: fewer and larger batches, traded against up to coalesce_us
     of added latency. */
  int coalesce_us;
  uint64_t batch_target_bytes; /* 0 = the per-batch cap */
} ftl_server_config;

/* Inputs of one step of one creature (the uniforms Hair::update sets before
   its dispatch, src/fx/hair.cc:100-112). */
typedef struct ftl_step_request {
  float dt;            /* seconds */
  float time;          /* wind time (ftl_groom_set_time) */
  int has_transform;   /* 1: `model` replaces the creature's root transform */
  float model[16];     /* column-major 4x4, as ftl_groom_set_root_transform */
  int ncolliders;      /* -1: keep the creature's colliders; 0..4: replace them */
  float spheres[16];   /* ncolliders x (centre xyz, radius) */
} ftl_step_request;

typedef struct ftl_server_stats {
  uint64_t submitted;     /* requests accepted since the last reset */
  uint64_t completed;     /* requests whose step has finished */
  uint64_t batches;       /* batches launched */
  uint64_t cps_stepped;   /* control-point updates completed */
  double p50_us, p99_us, max_us, mean_us; /* request latency */
  double mean_batch_requests;             /* requests per batch */
  double busy_s;          /* wall time with at least one batch on the device */
  double span_s;          /* wall time from the first submit to the last completion */
  uint64_t resident_bytes;
  int resident;           /* creatures on the device */
  int waiting;            /* creatures waiting for memory */
  uint64_t cps_finite;    /* control-point updates of completed steps that stayed finite */
  uint64_t dropped;       /* requests completed without a step (quarantined, removed or unadmitted creature) */
  uint64_t failed_admissions; /* waiting creatures dropped because they could not be made resident */
  int quarantined;        /* creatures quarantined now */
} ftl_server_stats;

ftl_status ftl_server_create(ftl_ctx* ctx, const ftl_server_config* cfg, ftl_server** out);
/* Waits for in-flight batches, frees every creature. */
void ftl_server_destroy(ftl_server* s);

/* Adds a creature: copies the caller's arrays (desc as ftl_groom_create,
   params as ftl_groom_set_params, root mode FTL_ROOT_REST when rest_roots is
   given, else FTL_ROOT_MODEL).  *id receives its handle; *resident (may be
   NULL) 1 if it was admitted at once, 0 if it waits for memory.  A creature
   larger than the whole budget is refused with FTL_ERR_BUDGET. */
ftl_status ftl_server_add(ftl_server* s, const ftl_groom_desc* desc, const ftl_params* params, int* id,
                          int* resident);
/* Removes a creature: waits for its step in flight, drops its pending
   requests (they complete unstepped: ftl_server_wait reports them with
   FTL_ERR_CANCELLED), frees its memory and admits waiting creatures that now
   fit.  A waiting creature that cannot be made resident for another reason
   than memory is dropped the same way when its turn comes. */
ftl_status ftl_server_remove(ftl_server* s, int id);
/* 1 when the creature is resident, 0 when it waits for memory. */
ftl_status ftl_server_resident(ftl_server* s, int id, int* resident);

/* Queues n step requests (ids[i], reqs[i]); tickets (may be NULL) receive
   their tickets, increasing in submission order.  Requests of a creature
   that waits for memory are kept until it is admitted. */
ftl_status ftl_server_submit(ftl_server* s, int n, const int* ids, const ftl_step_request* reqs,
                             uint64_t* tickets);
/* Waits until the request with this ticket has completed (timeout_ms < 0:
   no limit).  FTL_ERR_BUSY on timeout; FTL_ERR_NONFINITE or
   FTL_ERR_CANCELLED for a request completed without a step. */
ftl_status ftl_server_wait(ftl_server* s, uint64_t ticket, int timeout_ms);
/* Waits until every submitted request of a resident creature has completed
   (requests of creatures waiting for memory stay queued). */
ftl_status ftl_server_drain(ftl_server* s, int timeout_ms);
/* Copies a creature's state to the host (waits for its step in flight). */
ftl_status ftl_server_download(ftl_server* s, int id, float* pos, float* vel);
/* 1 when the creature is quarantined: one of its steps left NaN/Inf at a
   strand tip (the reference's normalize(0) on a zero-length segment,
   cs_simulation.glsl:114, propagates that way).  A quarantined creature is
   left out of batches; its queued and later requests complete without a step
   and ftl_server_wait reports them with FTL_ERR_NONFINITE. */
ftl_status ftl_server_status(ftl_server* s, int id, int* nonfinite);
/* Replaces a creature's state (float4 arrays as ftl_groom_upload; waits for
   its step in flight) and lifts its quarantine. */
ftl_status ftl_server_upload(ftl_server* s, int id, const float* pos, const float* vel);

ftl_status ftl_server_get_stats(ftl_server* s, ftl_server_stats* out);
/* Clears the latency samples and counters (not the creatures); requests
   submitted before the reset are left out of the statistics. */
ftl_status ftl_server_reset_stats(ftl_server* s);

#ifdef __cplusplus
}
#endif

#endif /* FTL_SERVER_H_ */
