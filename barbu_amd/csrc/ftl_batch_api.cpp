// Batched FTL steps: many grooms, uneven strand and point counts, one
// persistent launch per frame.  There is no counterpart in the reference,
// which dispatches one workgroup per strand of its single groom
// (src/fx/hair.cc:112).
//
// A batch is planned once: each groom is cut into 64-lane tiles of its launch
// class, and the tiles are ordered by decreasing cost (longest first) so the
// persistent grid drains the heavy tiles before the light ones.  A frame is
// [upload the per-groom descriptors from pinned memory -> reset the tile
// counter -> persistent kernel], captured once into a HIP graph.  Two graphs
// with their own pinned descriptor buffers alternate, so the host prepares
// frame k+1 while frame k runs.
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
  int ntiles = 0, grid = 0, mmax = 16;
  FtlTile* d_tiles = nullptr;
  FtlStep* d_steps[2] = {nullptr, nullptr};
  FtlStep* h_steps[2] = {nullptr, nullptr};
  unsigned* d_counter[2] = {nullptr, nullptr};
  hipStream_t cap = nullptr;
  hipGraph_t graph[2] = {nullptr, nullptr};
  hipGraphExec_t exec[2] = {nullptr, nullptr};
  hipEvent_t done[2] = {nullptr, nullptr};
  int cur = 0;
};

// Plans the tiles of a set of grooms.  Returns the largest points-per-lane.
static ftl_status plan_tiles(ftl_ctx* ctx, ftl_groom* const* grooms, int n, std::vector<FtlTile>& tiles, int* mmax) {
  tiles.clear();
  int mm = 0;
  for (int i = 0; i < n; ++i) {
    ftl_groom* g = grooms[i];
    if (!g) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "groom %d is NULL", i);
    if (g->ctx != ctx) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "groom %d belongs to another context", i);
    for (int j = 0; j < i; ++j)
      if (grooms[j] == g) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "groom %d appears twice in the batch", i);
    const int cls = ftl_class_for(ctx, g->npoints, g->lanes, 32);
    if (cls < 0) return ftl_fail(FTL_ERR_UNSUPPORTED, "no batched launch class for npoints=%d", g->npoints);
    const FtlClassInfo& k = kFtlClasses[cls];
    mm = std::max(mm, k.m);
    const int spt = FTL_WAVE / k.g;
    const int cost = k.m * (g->params.iterations + k.g - 1);
    for (int s0 = 0; s0 < g->nstrands; s0 += spt) tiles.push_back(FtlTile{i, s0, cls, cost});
  }
  std::stable_sort(tiles.begin(), tiles.end(), [](const FtlTile& a, const FtlTile& b) { return a.cost > b.cost; });
  *mmax = mm <= 16 ? 16 : 32;
  return FTL_OK;
}

static int batch_grid(ftl_ctx* ctx, int mmax, int ntiles) {
  int per_cu = ftl_batch_blocks_per_cu(mmax);
  if (per_cu <= 0) per_cu = 4;
  const long g = (long)ctx->ncu * per_cu;
  return (int)std::max(1L, std::min<long>(g, ntiles));
}

extern "C" ftl_status ftl_batch_step(ftl_ctx* ctx, ftl_groom* const* grooms, int n, const float* dt, void* stream) {
  if (!ctx || (!grooms && n > 0) || (!dt && n > 0) || n < 0) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "bad arguments");
  if (n == 0) return FTL_OK;
  std::vector<FtlTile> tiles;
  int mmax = 16;
  ftl_status st = plan_tiles(ctx, grooms, n, tiles, &mmax);
  if (st != FTL_OK) return st;
  std::vector<FtlStep> steps(n);
  for (int i = 0; i < n; ++i)
    if ((st = ftl_make_step(grooms[i], dt[i], &steps[i])) != FTL_OK) return st;
  hipError_t e;
  if ((e = hipSetDevice(ctx->device)) != hipSuccess) return ftl_hip_fail(e, "hipSetDevice");
  const size_t sb = steps.size() * sizeof(FtlStep), tb = tiles.size() * sizeof(FtlTile);
  // The scratch buffers may still be read by a previous batch step.
  if ((e = hipStreamSynchronize((hipStream_t)stream)) != hipSuccess) return ftl_hip_fail(e, "hipStreamSynchronize");
  if (sb > ctx->scratch_steps_bytes) {
    ftl_dev_free(ctx, ctx->scratch_steps, ctx->scratch_steps_bytes);
    ctx->scratch_steps = nullptr;
    ctx->scratch_steps_bytes = 0;
    if ((st = ftl_dev_alloc(ctx, &ctx->scratch_steps, sb, "batch descriptors")) != FTL_OK) return st;
    ctx->scratch_steps_bytes = sb;
  }
  if (tb > ctx->scratch_tiles_bytes) {
    ftl_dev_free(ctx, ctx->scratch_tiles, ctx->scratch_tiles_bytes);
    ctx->scratch_tiles = nullptr;
    ctx->scratch_tiles_bytes = 0;
    if ((st = ftl_dev_alloc(ctx, &ctx->scratch_tiles, tb, "batch tiles")) != FTL_OK) return st;
    ctx->scratch_tiles_bytes = tb;
  }
  if (!ctx->scratch_counter &&
      (st = ftl_dev_alloc(ctx, (void**)&ctx->scratch_counter, 256, "batch counter")) != FTL_OK)
    return st;
  if ((e = hipMemcpy(ctx->scratch_steps, steps.data(), sb, hipMemcpyHostToDevice)) != hipSuccess)
    return ftl_hip_fail(e, "hipMemcpy batch descriptors");
  if ((e = hipMemcpy(ctx->scratch_tiles, tiles.data(), tb, hipMemcpyHostToDevice)) != hipSuccess)
    return ftl_hip_fail(e, "hipMemcpy batch tiles");
  const int grid = batch_grid(ctx, mmax, (int)tiles.size());
  if ((e = ftl_launch_batch(mmax, (const FtlStep*)ctx->scratch_steps, (const FtlTile*)ctx->scratch_tiles,
                            (int)tiles.size(), ctx->scratch_counter, grid, (hipStream_t)stream)) != hipSuccess)
    return ftl_hip_fail(e, "batched step launch");
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
    ftl_dev_free(ctx, b->d_counter[k], 256);
  }
  ftl_dev_free(ctx, b->d_tiles, (size_t)b->ntiles * sizeof(FtlTile));
  if (b->cap) (void)hipStreamDestroy(b->cap);
}

extern "C" ftl_status ftl_batch_create(ftl_ctx* ctx, ftl_groom* const* grooms, int n, ftl_batch** out) {
  if (!ctx || !grooms || !out || n <= 0) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "bad arguments");
  *out = nullptr;
  std::vector<FtlTile> tiles;
  int mmax = 16;
  ftl_status st = plan_tiles(ctx, grooms, n, tiles, &mmax);
  if (st != FTL_OK) return st;
  hipError_t e;
  if ((e = hipSetDevice(ctx->device)) != hipSuccess) return ftl_hip_fail(e, "hipSetDevice");
  ftl_batch* b = new ftl_batch();
  b->ctx = ctx;
  b->grooms.assign(grooms, grooms + n);
  b->ntiles = (int)tiles.size();
  b->mmax = mmax;
  b->grid = batch_grid(ctx, mmax, b->ntiles);
  const size_t sb = (size_t)n * sizeof(FtlStep), tb = tiles.size() * sizeof(FtlTile);
  auto bail = [&](ftl_status s) {
    batch_release(b);
    delete b;
    return s;
  };
  if ((st = ftl_dev_alloc(ctx, (void**)&b->d_tiles, tb, "batch tiles")) != FTL_OK) return bail(st);
  if ((e = hipMemcpy(b->d_tiles, tiles.data(), tb, hipMemcpyHostToDevice)) != hipSuccess)
    return bail(ftl_hip_fail(e, "hipMemcpy batch tiles"));
  if ((e = hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking)) != hipSuccess)
    return bail(ftl_hip_fail(e, "hipStreamCreate"));
  for (int k = 0; k < 2; ++k) {
    if ((st = ftl_dev_alloc(ctx, (void**)&b->d_steps[k], sb, "batch descriptors")) != FTL_OK) return bail(st);
    if ((st = ftl_dev_alloc(ctx, (void**)&b->d_counter[k], 256, "batch counter")) != FTL_OK) return bail(st);
    if ((e = hipHostMalloc((void**)&b->h_steps[k], sb, hipHostMallocDefault)) != hipSuccess)
      return bail(ftl_hip_fail(e, "hipHostMalloc"));
    std::memset(b->h_steps[k], 0, sb);
    if ((e = hipEventCreateWithFlags(&b->done[k], hipEventDisableTiming)) != hipSuccess)
      return bail(ftl_hip_fail(e, "hipEventCreate"));
    if ((e = hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal)) != hipSuccess)
      return bail(ftl_hip_fail(e, "hipStreamBeginCapture"));
    hipError_t e1 = hipMemcpyAsync(b->d_steps[k], b->h_steps[k], sb, hipMemcpyHostToDevice, b->cap);
    hipError_t e2 = ftl_launch_batch(b->mmax, b->d_steps[k], b->d_tiles, b->ntiles, b->d_counter[k], b->grid, b->cap);
    hipError_t e3 = hipStreamEndCapture(b->cap, &b->graph[k]);
    if (e1 != hipSuccess) return bail(ftl_hip_fail(e1, "capture: descriptor upload"));
    if (e2 != hipSuccess) return bail(ftl_hip_fail(e2, "capture: batched kernel"));
    if (e3 != hipSuccess) return bail(ftl_hip_fail(e3, "hipStreamEndCapture"));
    if ((e = hipGraphInstantiate(&b->exec[k], b->graph[k], nullptr, nullptr, 0)) != hipSuccess)
      return bail(ftl_hip_fail(e, "hipGraphInstantiate"));
  }
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->live_objects++;
  }
  *out = b;
  return FTL_OK;
}

extern "C" void ftl_batch_destroy(ftl_batch* b) {
  if (!b) return;
  ftl_ctx* ctx = b->ctx;
  batch_release(b);
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->live_objects--;
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
