// Internal host-side state of the C ABI objects.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <vector>

#include "ftl.h"
#include "ftl_classes.h"
#include "ftl_device.h"

struct ftl_ctx {
  int device = 0;
  int ncu = 0;
  int lanes[FTL_MAX_POINTS + 1];       // tuned lanes per strand by point count (0 = default rule)
  int lanes_wind[FTL_MAX_POINTS + 1];  // the same for plain launches with wind (0 = as lanes)
  // tuned launch class (tile shape included) by point count, -1 = none: the
  // class the tables pick among the base classes and their variants
  int pick[FTL_MAX_POINTS + 1];
  int pick_wind[FTL_MAX_POINTS + 1];
  float cls_ns_cp[FTL_NCLASSES] = {};  // measured ns per control point of each class (0 = not measured)
  std::mutex mu;
  uint64_t bytes = 0;  // device memory held by the context's objects
  int live_objects = 0;
  // scratch of ftl_batch_step, and the event recorded after the last batch
  // that reads it (on whatever stream that batch ran)
  std::mutex scratch_mu;
  hipEvent_t scratch_ev = nullptr;
  bool scratch_ev_recorded = false;
  void* scratch_steps = nullptr;
  size_t scratch_steps_bytes = 0;
  void* scratch_spans = nullptr;
  size_t scratch_spans_bytes = 0;
};

struct ftl_groom;

struct ftl_skin {
  ftl_ctx* ctx = nullptr;
  int nverts = 0, ntris = 0, nbones = 0;
  int mode = FTL_SKIN_DQ;
  bool has_no_joint = false;  // some joint index is FTL_NO_JOINT (LBS only)
  void* d_rest = nullptr;
  void* d_joints = nullptr;
  void* d_weights = nullptr;
  void* d_out = nullptr;
  void* d_tris = nullptr;
  void* d_dq = nullptr;    // 2 float4 per bone
  void* d_mats = nullptr;  // 3 float4 per bone (rows of the skinning matrix)
  void* d_nrest = nullptr; // normals (float4 per vertex), or null
  void* d_nout = nullptr;
  std::vector<float> h_dq, h_mats;
  bool dirty = false, mats_dirty = false;
  // grooms bound to this skin (ctx->mu): destroying the skin unbinds them
  std::vector<ftl_groom*> bound;
};

struct ftl_groom {
  ftl_ctx* ctx = nullptr;
  int nstrands = 0, npoints = 0;
  void* d_pos = nullptr;
  void* d_vel = nullptr;
  void* d_roots = nullptr;
  void* d_bind = nullptr;
  unsigned* d_status = nullptr;
  void* d_wind = nullptr;
  uint64_t wind_seed = 0;
  float wind_amp = 0.0f;
  ftl_params params;
  int ncol = 1;
  float sph[FTL_MAX_COLLIDERS][4];
  float model[16];
  int model_ident = 1;
  int root_mode = FTL_ROOT_MODEL;
  double time = 0.0;
  ftl_skin* skin = nullptr;
  int lanes = 0;
  // captured batches that step this groom (ctx->mu): destroying the groom
  // takes it out of them
  std::vector<ftl_batch*> batches;
};

// Takes g out of batch b's groom list (ctx->mu held): the batch's next
// launch is refused.
void ftl_batch_forget_groom_locked(ftl_batch* b, ftl_groom* g);

// Render-side expansion object (ftl_render.hip); its tangents are also
// written by the device-side initialisation (ftl_init.hip).
struct ftl_expand {
  ftl_ctx* ctx = nullptr;
  int nstrands = 0, npoints = 0, ntris = 0, nrandom = 0;
  int4* d_tris = nullptr;
  float4* d_tan = nullptr;
  float4* d_rand = nullptr;
};

ftl_status ftl_fail(ftl_status st, const char* fmt, ...);
// Validates simulation parameters (ranges, finiteness).
ftl_status ftl_check_params(const ftl_params* p);
// ftl_groom_create without a host state: zero state when desc->positions is
// NULL (filled on the device afterwards).
ftl_status ftl_groom_create_uninitialised(ftl_ctx* ctx, const ftl_groom_desc* d, ftl_groom** out);
ftl_status ftl_hip_fail(hipError_t e, const char* what);
ftl_status ftl_dev_alloc(ftl_ctx* ctx, void** p, size_t bytes, const char* what);
// Validates a caller's device pointer before a kernel uses `bytes` from it.
ftl_status ftl_check_device_range(const void* p, size_t bytes, const char* what);
void ftl_dev_free(ftl_ctx* ctx, void* p, size_t bytes);
// Fills the kernel descriptor of one step of g.
// may_alloc false: fail rather than allocate the groom's wind lattice.
ftl_status ftl_make_step(ftl_groom* g, float dt, FtlStep* A, bool may_alloc = true);
// Per-device set-up of the no-allocation batched step (its scratch address);
// called by ftl_init with the device current.
ftl_status ftl_noalloc_setup(int device);
// Launch class for npoints: requested lanes (0 = tuned/default), at most mmax
// points per lane.  -1 if none.
// Launch class of a groom: lanes = the groom's setting (0 = tuned); wind: a
// plain launch with wind (it runs the LDS-lattice form, tuned separately).
int ftl_class_for(ftl_ctx* ctx, int npoints, int lanes, int mmax, bool wind = false);
// The class of npoints without tuning tables (the built-in lanes rule).
int ftl_default_class(int npoints);

// Batched steps (ftl_batch_api.cpp), shared by the batch objects and the
// serving loop.
// One class's launch in a batch: spans [first, first + count) of the batch's
// span list, ntiles tiles in all, stepped by grid workgroups.
struct ClassRun {
  int cls, first, count, ntiles, grid;
  bool nt;  // non-temporal stores (the batch's state exceeds ftl_stream_store_bytes())
};
// Groups the grooms by class (one span per groom, in tile order within its
// class) and sizes the class grids.
ftl_status ftl_plan_tiles(ftl_ctx* ctx, ftl_groom* const* grooms, int n, std::vector<FtlSpan>& spans,
                          std::vector<ClassRun>& runs);
// Enqueues one tile-list kernel per class on s; with side streams (and their
// events) the classes after the first run on them, forked from and joined
// back to s.
hipError_t ftl_enqueue_runs(const std::vector<ClassRun>& runs, const FtlStep* d_steps, const FtlSpan* d_spans,
                            hipStream_t s, const std::vector<hipStream_t>* side, const std::vector<hipEvent_t>* side_ev,
                            hipEvent_t fork);
