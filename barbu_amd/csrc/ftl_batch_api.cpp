// Batched FTL steps: many grooms, uneven strand and point counts, stepped by
// one graph launch per frame.  There is no counterpart in the reference, which
// dispatches one workgroup per strand of its single groom (src/fx/hair.cc:112).
//
// A batch is planned once: each groom is cut into 64-lane tiles of its launch
// class (points per lane x lanes per strand), and the grooms are grouped by
// class, each a span of consecutive tiles of its class's launch (one record
// per groom: the kernel maps a tile to its groom by a binary search over the
// spans).  Every class has its own tile-list kernel, compiled with the register
// budget of that class (one kernel for all classes would carry the register
// count of the heaviest one).  A frame is [upload the per-groom descriptors
// from pinned memory -> one tile-list kernel per class present, as parallel
// branches], captured once into a HIP graph: one graph launch steps every
// groom.  Two graphs with their own pinned descriptor buffers alternate, so
// the host prepares frame k+1 while frame k runs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "ftl.h"
#include "ftl_classes.h"
#include "ftl_internal.h"
#include "ftl_launch.h"

struct ftl_batch {
  ftl_ctx* ctx = nullptr;
  std::vector<ftl_groom*> grooms;
  std::vector<ClassRun> runs;
  int ntiles = 0, grid = 0, nspans = 0;
  FtlSpan* d_spans = nullptr;
  FtlStep* d_steps[2] = {nullptr, nullptr};
  FtlStep* h_steps[2] = {nullptr, nullptr};
  hipStream_t cap = nullptr;
  std::vector<hipStream_t> side;
  std::vector<hipEvent_t> side_ev;
  hipEvent_t fork = nullptr;
  hipGraph_t graph[2] = {nullptr, nullptr};
  hipGraphExec_t exec[2] = {nullptr, nullptr};
  hipEvent_t done[2] = {nullptr, nullptr};
  int cur = 0;
};


// Time per control point of a launch class in ns for K iterations: measured
// by the autotuner (at 8 iterations) where available, otherwise estimated.
static double class_ns_per_cp(const ftl_ctx* ctx, int cls, int K) {
  const FtlClassInfo& k = kFtlClasses[cls];
  // per-lane work: M points through K + G - 1 pipeline steps, plus about
  // three steps' worth of loads, wind and stores
  const double steps = (double)(K + k.g + 2) / (double)(8 + k.g + 2);
  double ns = ctx->cls_ns_cp[cls];
  if (!(ns > 0.0)) ns = (k.guarded ? 0.028 : 0.022) * (k.g == 1 ? 1.0 : 1.0 + 0.1 * (k.g - 1));
  return ns * steps;
}

// Plans the tiles of a set of grooms, grouped by launch class (the class with
// the most expected work first; within a class, grooms with more iterations
// first), as one span per groom.  Linear in the number of grooms: the serving
// loop plans a batch per launch.
ftl_status ftl_plan_tiles(ftl_ctx* ctx, ftl_groom* const* grooms, int n, std::vector<FtlSpan>& spans,
                          std::vector<ClassRun>& runs) {
  spans.clear();
  runs.clear();
  struct GroomPlan {
    int idx, cls, cost, ntiles;
    double ns_tile;
  };
  std::vector<GroomPlan> plan;
  plan.reserve(n);
  std::vector<double> work(FTL_NCLASSES, 0.0);
  size_t total_tiles = 0, state_bytes = 0;
  for (int i = 0; i < n; ++i) {
    ftl_groom* g = grooms[i];
    if (!g) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "groom %d is NULL", i);
    if (g->ctx != ctx) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "groom %d belongs to another context", i);
    const int cls = ftl_class_for(ctx, g->npoints, g->lanes, FTL_MAX_POINTS);
    if (cls < 0) return ftl_fail(FTL_ERR_UNSUPPORTED, "no batched launch class for npoints=%d", g->npoints);
    const FtlClassInfo& k = kFtlClasses[cls];
    const int spt = FTL_WAVE / k.g;
    const int nt = (g->nstrands + spt - 1) / spt;
    const int cost = k.m * (g->params.iterations + k.g - 1);
    const double ns_tile = class_ns_per_cp(ctx, cls, g->params.iterations) * (double)FTL_WAVE * k.m;
    plan.push_back(GroomPlan{i, cls, cost, nt, ns_tile});
    work[cls] += ns_tile * nt;
    total_tiles += (size_t)nt;
    state_bytes += (size_t)g->nstrands * g->npoints * 2 * sizeof(float4);
  }
  // the class kernels of a batch run together: the batch's whole state decides
  const bool stream_stores = state_bytes > ftl_stream_store_bytes();
  if (total_tiles > (size_t)INT32_MAX) return ftl_fail(FTL_ERR_UNSUPPORTED, "too many tiles in one batch");
  {
    std::vector<ftl_groom*> seen(grooms, grooms + n);
    std::sort(seen.begin(), seen.end());
    if (std::adjacent_find(seen.begin(), seen.end()) != seen.end())
      return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "a groom appears twice in the batch");
  }
  // Classes with the most expected work first (they start first and run
  // longest); within a class, the costliest grooms first.
  std::stable_sort(plan.begin(), plan.end(), [&](const GroomPlan& a, const GroomPlan& b) {
    if (a.cls != b.cls) {
      if (work[a.cls] != work[b.cls]) return work[a.cls] > work[b.cls];
      return a.cls < b.cls;
    }
    return a.cost > b.cost;
  });
  spans.reserve(n);
  // One workgroup per tile: the classes' kernels run side by side and the
  // hardware dispatcher hands free wave slots to whichever has tiles left.
  for (size_t i = 0; i < plan.size();) {
    const int cls = plan[i].cls;
    const int first = (int)spans.size();
    int ntiles = 0;
    for (; i < plan.size() && plan[i].cls == cls; ++i) {
      const GroomPlan& gp = plan[i];
      if (gp.ntiles <= 0) continue;  // no strands: nothing to step
      spans.push_back(FtlSpan{gp.idx, ntiles, cls, gp.cost});
      ntiles += gp.ntiles;
    }
    const int count = (int)spans.size() - first;
    if (count > 0) runs.push_back(ClassRun{cls, first, count, ntiles, ntiles, stream_stores});
  }
  if (runs.size() > 64) return ftl_fail(FTL_ERR_UNSUPPORTED, "too many launch classes");
  return FTL_OK;
}

// Enqueues one tile-list kernel per class on s.  With side streams (graph
// capture, serving slots) the classes after the first are forked onto them
// and joined back, so they run side by side (parallel graph branches).
hipError_t ftl_enqueue_runs(const std::vector<ClassRun>& runs, const FtlStep* d_steps, const FtlSpan* d_spans,
                            hipStream_t s, const std::vector<hipStream_t>* side, const std::vector<hipEvent_t>* side_ev,
                            hipEvent_t fork) {
  hipError_t e;
  const bool par = side && runs.size() > 1;
  if (par && (e = hipEventRecord(fork, s)) != hipSuccess) return e;
  for (size_t r = 0; r < runs.size(); ++r) {
    hipStream_t rs = (par && r > 0) ? (*side)[r - 1] : s;
    if (par && r > 0 && (e = hipStreamWaitEvent(rs, fork, 0)) != hipSuccess) return e;
    const ClassRun& cr = runs[r];
    if ((e = ftl_launch_tiles(cr.cls, d_steps, d_spans + cr.first, cr.count, cr.ntiles, cr.grid, rs, cr.nt)) !=
        hipSuccess)
      return e;
    if (par && r > 0) {
      if ((e = hipEventRecord((*side_ev)[r - 1], rs)) != hipSuccess) return e;
      if ((e = hipStreamWaitEvent(s, (*side_ev)[r - 1], 0)) != hipSuccess) return e;
    }
  }
  return hipSuccess;
}

extern "C" ftl_status ftl_batch_step(ftl_ctx* ctx, ftl_groom* const* grooms, int n, const float* dt, void* stream) {
  if (!ctx || (!grooms && n > 0) || (!dt && n > 0) || n < 0) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "bad arguments");
  if (n == 0) return FTL_OK;
  std::vector<FtlSpan> spans;
  std::vector<ClassRun> runs;
  ftl_status st = ftl_plan_tiles(ctx, grooms, n, spans, runs);
  if (st != FTL_OK) return st;
  std::vector<FtlStep> steps(n);
  for (int i = 0; i < n; ++i)
    if ((st = ftl_make_step(grooms[i], dt[i], &steps[i])) != FTL_OK) return st;
  if (spans.empty()) return FTL_OK;
  hipError_t e;
  if ((e = hipSetDevice(ctx->device)) != hipSuccess) return ftl_hip_fail(e, "hipSetDevice");
  const size_t sb = steps.size() * sizeof(FtlStep), tb = spans.size() * sizeof(FtlSpan);
  // The scratch buffers may still be read by the previous batch step, which
  // may have run on another stream: wait for its kernels before the
  // descriptors are replaced (or the buffers freed to grow).
  std::lock_guard<std::mutex> scratch_lock(ctx->scratch_mu);
  if (!ctx->scratch_ev && (e = hipEventCreateWithFlags(&ctx->scratch_ev, hipEventDisableTiming)) != hipSuccess)
    return ftl_hip_fail(e, "hipEventCreate");
  if (ctx->scratch_ev_recorded && (e = hipEventSynchronize(ctx->scratch_ev)) != hipSuccess)
    return ftl_hip_fail(e, "hipEventSynchronize");
  if (sb > ctx->scratch_steps_bytes) {
    ftl_dev_free(ctx, ctx->scratch_steps, ctx->scratch_steps_bytes);
    ctx->scratch_steps = nullptr;
    ctx->scratch_steps_bytes = 0;
    if ((st = ftl_dev_alloc(ctx, &ctx->scratch_steps, sb, "batch descriptors")) != FTL_OK) return st;
    ctx->scratch_steps_bytes = sb;
  }
  if (tb > ctx->scratch_spans_bytes) {
    ftl_dev_free(ctx, ctx->scratch_spans, ctx->scratch_spans_bytes);
    ctx->scratch_spans = nullptr;
    ctx->scratch_spans_bytes = 0;
    if ((st = ftl_dev_alloc(ctx, &ctx->scratch_spans, tb, "batch spans")) != FTL_OK) return st;
    ctx->scratch_spans_bytes = tb;
  }
  if ((e = hipMemcpy(ctx->scratch_steps, steps.data(), sb, hipMemcpyHostToDevice)) != hipSuccess)
    return ftl_hip_fail(e, "hipMemcpy batch descriptors");
  if ((e = hipMemcpy(ctx->scratch_spans, spans.data(), tb, hipMemcpyHostToDevice)) != hipSuccess)
    return ftl_hip_fail(e, "hipMemcpy batch spans");
  if ((e = ftl_enqueue_runs(runs, (const FtlStep*)ctx->scratch_steps, (const FtlSpan*)ctx->scratch_spans,
                            (hipStream_t)stream, nullptr, nullptr, nullptr)) != hipSuccess)
    return ftl_hip_fail(e, "batched step launch");
  if ((e = hipEventRecord(ctx->scratch_ev, (hipStream_t)stream)) != hipSuccess)
    return ftl_hip_fail(e, "hipEventRecord");
  ctx->scratch_ev_recorded = true;
  return FTL_OK;
}

// ---------------------------------------------------------------------------
// No-allocation batched step: the descriptors and the span list live in a
// file-scope device array of fixed size (one copy per GPU, held until the
// process ends) instead of buffers the context allocates.  Host code never
// uses the array's name as an address: its device address is looked up once
// per device (ftl_noalloc_setup, from ftl_init) and kept in kNoalloc[dev].ptr.

static __device__ __align__(256) unsigned char scratch[FTL_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
struct NoallocDevice {
  std::mutex mu;            // one call at a time per device (host side)
  bool ready = false;
  void* ptr = nullptr;      // device address of `scratch` on this device
  hipEvent_t done = nullptr;  // recorded after the last call's kernels
  bool recorded = false;
};
NoallocDevice kNoalloc[kMaxDevices];
}  // namespace

ftl_status ftl_noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return ftl_fail(FTL_ERR_UNSUPPORTED, "device %d out of range", device);
  NoallocDevice& d = kNoalloc[device];
  std::lock_guard<std::mutex> lk(d.mu);
  if (d.ready) return FTL_OK;
  hipError_t e;
  void* p = nullptr;
  if ((e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch))) != hipSuccess || !p)
    return ftl_hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
  if ((e = hipEventCreateWithFlags(&d.done, hipEventDisableTiming)) != hipSuccess)
    return ftl_hip_fail(e, "hipEventCreate");
  d.ptr = p;
  d.ready = true;
  return FTL_OK;
}

extern "C" ftl_status ftl_batch_step_noalloc(ftl_ctx* ctx, ftl_groom* const* grooms, int n, const float* dt,
                                             void* stream) {
  if (!ctx || (!grooms && n > 0) || (!dt && n > 0) || n < 0) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "bad arguments");
  if (n == 0) return FTL_OK;
  if (ctx->device < 0 || ctx->device >= kMaxDevices || !kNoalloc[ctx->device].ready)
    return ftl_fail(FTL_ERR_NOT_INITIALIZED, "no-allocation scratch not set up on device %d", ctx->device);
  // At most one span per groom: the size is checked before any planning.
  const size_t sb = ((size_t)n * sizeof(FtlStep) + 255) & ~(size_t)255;
  if (sb + (size_t)n * sizeof(FtlSpan) > FTL_NOALLOC_SCRATCH_BYTES)
    return ftl_fail(FTL_ERR_INVALID_ARGUMENT,
                    "batch too large for the no-allocation scratch: %d grooms (%zu B of descriptors, %zu B of "
                    "spans) exceed its %u bytes; use ftl_batch_step, or split the batch",
                    n, sb, (size_t)n * sizeof(FtlSpan), (unsigned)FTL_NOALLOC_SCRATCH_BYTES);
  std::vector<FtlSpan> spans;
  std::vector<ClassRun> runs;
  ftl_status st = ftl_plan_tiles(ctx, grooms, n, spans, runs);
  if (st != FTL_OK) return st;
  const size_t tb = spans.size() * sizeof(FtlSpan);
  std::vector<FtlStep> steps(n);
  for (int i = 0; i < n; ++i)
    if ((st = ftl_make_step(grooms[i], dt[i], &steps[i], false)) != FTL_OK) return st;
  hipError_t e;
  if ((e = hipSetDevice(ctx->device)) != hipSuccess) return ftl_hip_fail(e, "hipSetDevice");
  NoallocDevice& d = kNoalloc[ctx->device];
  std::lock_guard<std::mutex> lk(d.mu);
  // The previous call's kernels (on any stream) must have read the scratch
  // before it is overwritten.
  if (d.recorded && (e = hipEventSynchronize(d.done)) != hipSuccess) return ftl_hip_fail(e, "hipEventSynchronize");
  unsigned char* base = (unsigned char*)d.ptr;
  if ((e = hipMemcpy(base, steps.data(), (size_t)n * sizeof(FtlStep), hipMemcpyHostToDevice)) != hipSuccess ||
      (tb > 0 && (e = hipMemcpy(base + sb, spans.data(), tb, hipMemcpyHostToDevice)) != hipSuccess))
    return ftl_hip_fail(e, "hipMemcpy to the no-allocation scratch");
  if ((e = ftl_enqueue_runs(runs, (const FtlStep*)base, (const FtlSpan*)(base + sb), (hipStream_t)stream, nullptr,
                            nullptr, nullptr)) != hipSuccess)
    return ftl_hip_fail(e, "batched step launch");
  if ((e = hipEventRecord(d.done, (hipStream_t)stream)) != hipSuccess) return ftl_hip_fail(e, "hipEventRecord");
  d.recorded = true;
  return FTL_OK;
}

static void batch_release(ftl_batch* b) {
  ftl_ctx* ctx = b->ctx;
  (void)hipSetDevice(ctx->device);
  (void)hipDeviceSynchronize();
  for (int k = 0; k < 2; ++k) {
    if (b->exec[k]) (void)hipGraphExecDestroy(b->exec[k]);
    if (b->graph[k]) (void)hipGraphDestroy(b->graph[k]);
    if (b->done[k]) (void)hipEventDestroy(b->done[k]);
    if (b->h_steps[k]) (void)hipHostFree(b->h_steps[k]);
    ftl_dev_free(ctx, b->d_steps[k], b->grooms.size() * sizeof(FtlStep));
  }
  ftl_dev_free(ctx, b->d_spans, (size_t)b->nspans * sizeof(FtlSpan));
  for (hipStream_t s : b->side) (void)hipStreamDestroy(s);
  for (hipEvent_t ev : b->side_ev) (void)hipEventDestroy(ev);
  if (b->fork) (void)hipEventDestroy(b->fork);
  if (b->cap) (void)hipStreamDestroy(b->cap);
}

extern "C" ftl_status ftl_batch_create(ftl_ctx* ctx, ftl_groom* const* grooms, int n, ftl_batch** out) {
  if (!ctx || !grooms || !out || n <= 0) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "bad arguments");
  *out = nullptr;
  std::vector<FtlSpan> spans;
  std::vector<ClassRun> runs;
  ftl_status st = ftl_plan_tiles(ctx, grooms, n, spans, runs);
  if (st != FTL_OK) return st;
  hipError_t e;
  if ((e = hipSetDevice(ctx->device)) != hipSuccess) return ftl_hip_fail(e, "hipSetDevice");
  ftl_batch* b = new ftl_batch();
  b->ctx = ctx;
  b->grooms.assign(grooms, grooms + n);
  b->nspans = (int)spans.size();
  b->runs = runs;
  for (const ClassRun& r : runs) {
    b->ntiles += r.ntiles;
    b->grid += r.grid;
  }
  const size_t sb = (size_t)n * sizeof(FtlStep), tb = spans.size() * sizeof(FtlSpan);
  auto bail = [&](ftl_status s) {
    batch_release(b);
    delete b;
    return s;
  };
  if (tb > 0) {
    if ((st = ftl_dev_alloc(ctx, (void**)&b->d_spans, tb, "batch spans")) != FTL_OK) return bail(st);
    if ((e = hipMemcpy(b->d_spans, spans.data(), tb, hipMemcpyHostToDevice)) != hipSuccess)
      return bail(ftl_hip_fail(e, "hipMemcpy batch spans"));
  }
  if ((e = hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking)) != hipSuccess)
    return bail(ftl_hip_fail(e, "hipStreamCreate"));
  if ((e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming)) != hipSuccess)
    return bail(ftl_hip_fail(e, "hipEventCreate"));
  for (size_t r = 1; r < runs.size(); ++r) {
    hipStream_t s2 = nullptr;
    hipEvent_t ev = nullptr;
    if ((e = hipStreamCreateWithFlags(&s2, hipStreamNonBlocking)) != hipSuccess)
      return bail(ftl_hip_fail(e, "hipStreamCreate"));
    b->side.push_back(s2);
    if ((e = hipEventCreateWithFlags(&ev, hipEventDisableTiming)) != hipSuccess)
      return bail(ftl_hip_fail(e, "hipEventCreate"));
    b->side_ev.push_back(ev);
  }
  for (int k = 0; k < 2; ++k) {
    if ((st = ftl_dev_alloc(ctx, (void**)&b->d_steps[k], sb, "batch descriptors")) != FTL_OK) return bail(st);
#ifndef FTL_KERNEL_UPLOAD
#define FTL_KERNEL_UPLOAD 1
#endif
    if ((e = hipHostMalloc((void**)&b->h_steps[k], sb, FTL_KERNEL_UPLOAD ? hipHostMallocCoherent : hipHostMallocDefault)) !=
        hipSuccess)
      return bail(ftl_hip_fail(e, "hipHostMalloc"));
    std::memset(b->h_steps[k], 0, sb);
    if ((e = hipEventCreateWithFlags(&b->done[k], hipEventDisableTiming)) != hipSuccess)
      return bail(ftl_hip_fail(e, "hipEventCreate"));
    if ((e = hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal)) != hipSuccess)
      return bail(ftl_hip_fail(e, "hipStreamBeginCapture"));
    hipError_t e1 = FTL_KERNEL_UPLOAD ? ftl_launch_upload(b->h_steps[k], b->d_steps[k], sb, b->cap)
                                      : hipMemcpyAsync(b->d_steps[k], b->h_steps[k], sb, hipMemcpyHostToDevice, b->cap);
    hipError_t e2 = e1 == hipSuccess
                        ? ftl_enqueue_runs(b->runs, b->d_steps[k], b->d_spans, b->cap, &b->side, &b->side_ev, b->fork)
                        : hipSuccess;
    hipError_t e3 = hipStreamEndCapture(b->cap, &b->graph[k]);
    if (e1 != hipSuccess) return bail(ftl_hip_fail(e1, "capture: descriptor upload"));
    if (e2 != hipSuccess) return bail(ftl_hip_fail(e2, "capture: tile kernels"));
    if (e3 != hipSuccess) return bail(ftl_hip_fail(e3, "hipStreamEndCapture"));
    if ((e = hipGraphInstantiate(&b->exec[k], b->graph[k], nullptr, nullptr, 0)) != hipSuccess)
      return bail(ftl_hip_fail(e, "hipGraphInstantiate"));
  }
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->live_objects++;
    for (ftl_groom* g : b->grooms) g->batches.push_back(b);
  }
  *out = b;
  return FTL_OK;
}

void ftl_batch_forget_groom_locked(ftl_batch* b, ftl_groom* g) {
  for (ftl_groom*& x : b->grooms)
    if (x == g) x = nullptr;
}

extern "C" void ftl_batch_destroy(ftl_batch* b) {
  if (!b) return;
  ftl_ctx* ctx = b->ctx;
  batch_release(b);
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->live_objects--;
    for (ftl_groom* g : b->grooms)
      if (g) g->batches.erase(std::remove(g->batches.begin(), g->batches.end(), b), g->batches.end());
  }
  delete b;
}

extern "C" ftl_status ftl_batch_launch(ftl_batch* b, const float* dt, void* stream) {
  if (!b || !dt) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "NULL argument");
  hipError_t e;
  if ((e = hipSetDevice(b->ctx->device)) != hipSuccess) return ftl_hip_fail(e, "hipSetDevice");
  const int k = b->cur;
  // The pinned descriptors of buffer k were last read by the frame before the
  // previous one.
  if ((e = hipEventSynchronize(b->done[k])) != hipSuccess) return ftl_hip_fail(e, "hipEventSynchronize");
  {
    std::lock_guard<std::mutex> lk(b->ctx->mu);
    for (size_t i = 0; i < b->grooms.size(); ++i)
      if (!b->grooms[i])
        return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "groom %zu of the batch was destroyed (destroy the batch)", i);
  }
  for (size_t i = 0; i < b->grooms.size(); ++i) {
    ftl_status st = ftl_make_step(b->grooms[i], dt[i], &b->h_steps[k][i]);
    if (st != FTL_OK) return st;
  }
  if ((e = hipGraphLaunch(b->exec[k], (hipStream_t)stream)) != hipSuccess) return ftl_hip_fail(e, "hipGraphLaunch");
  if ((e = hipEventRecord(b->done[k], (hipStream_t)stream)) != hipSuccess) return ftl_hip_fail(e, "hipEventRecord");
  b->cur ^= 1;
  return FTL_OK;
}

extern "C" ftl_status ftl_batch_info(ftl_batch* b, int* ntiles, int* grid) {
  if (!b) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "batch is NULL");
  if (ntiles) *ntiles = b->ntiles;
  if (grid) *grid = b->grid;
  return FTL_OK;
}
