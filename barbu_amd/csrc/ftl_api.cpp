// C ABI of the FTL step (include/ftl.h): contexts, grooms, skins, batches.
//
// The reference drives its step from Hair::update (src/fx/hair.cc:89-125):
// bind the ping-pong SSBOs, set uTimeStep / uScaleFactor / uModel /
// uBoundingSphere, glDispatchCompute(nroots), barrier, copy the write buffer
// back.  Here a step is one kernel launch over device-resident state that is
// updated in place; parameters travel as kernel arguments (plain step) or as a
// device table uploaded inside a captured HIP graph (batched frames).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "ftl.h"
#include "ftl_classes.h"
#include "ftl_internal.h"
#include "ftl_launch.h"

// ---------------------------------------------------------------------------
// errors

static thread_local std::string t_err;

ftl_status ftl_fail(ftl_status st, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  t_err = buf;
  return st;
}

extern "C" const char* ftl_last_error(void) { return t_err.c_str(); }
extern "C" int ftl_abi_version(void) { return FTL_ABI_VERSION; }

ftl_status ftl_hip_fail(hipError_t e, const char* what) {
  return ftl_fail(e == hipErrorOutOfMemory ? FTL_ERR_OUT_OF_MEMORY : FTL_ERR_DEVICE, "%s: %s", what,
                  hipGetErrorString(e));
}

// ---------------------------------------------------------------------------
// memory

// A caller's device pointer: it must point into device, managed or pinned
// host memory, and its allocation must hold `bytes` from it, so a kernel
// launched on it stays in bounds.
ftl_status ftl_check_device_range(const void* p, size_t bytes, const char* what) {
  hipPointerAttribute_t a;
  if (hipPointerGetAttributes(&a, p) != hipSuccess || a.type == hipMemoryTypeUnregistered) {
    (void)hipGetLastError();
    return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "%s is not device memory", what);
  }
  hipDeviceptr_t base = nullptr;
  size_t size = 0;
  if (hipMemGetAddressRange(&base, &size, (hipDeviceptr_t)p) == hipSuccess && base) {
    const size_t avail = size - (size_t)((const char*)p - (const char*)base);
    if (bytes > avail)
      return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "%s: its allocation holds %zu bytes from the pointer, the call needs %zu",
                      what, avail, bytes);
  } else {
    (void)hipGetLastError();
  }
  return FTL_OK;
}

ftl_status ftl_dev_alloc(ftl_ctx* ctx, void** p, size_t bytes, const char* what) {
  *p = nullptr;
  if (bytes == 0) return FTL_OK;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    (void)hipGetLastError();
    return ftl_fail(e == hipErrorOutOfMemory ? FTL_ERR_OUT_OF_MEMORY : FTL_ERR_DEVICE,
                    "hipMalloc(%zu bytes) for %s: %s", bytes, what, hipGetErrorString(e));
  }
  std::lock_guard<std::mutex> lk(ctx->mu);
  ctx->bytes += bytes;
  return FTL_OK;
}

void ftl_dev_free(ftl_ctx* ctx, void* p, size_t bytes) {
  if (!p) return;
  (void)hipFree(p);
  std::lock_guard<std::mutex> lk(ctx->mu);
  ctx->bytes -= bytes;
}

// ---------------------------------------------------------------------------
// parameters

extern "C" void ftl_params_reference(ftl_params* p) {
  if (!p) return;
  std::memset(p, 0, sizeof *p);
  p->gravity[0] = 0.0f;
  p->gravity[1] = -9.81f;
  p->gravity[2] = 0.0f;
  p->force_scale = 20.0f;   // kForceCoeff (cs_simulation.glsl:72)
  p->damping = 0.80f;       // mix(0.80, 0.94, 0) on lane 0 (:100-102)
  p->length_scale = 1.45f;  // Hair::Parameters_t::render.lengthScale (hair.h:41)
  p->iterations = 8;        // kNumContraintsIteration (:197)
  p->velocity_mode = FTL_VELOCITY_SHIPPED;
  p->max_velocity = 100.0f;  // uMaxVelocity (:40)
  p->wind_amplitude = 0.0f;
  p->wind_frequency = 2.0f;
  p->wind_seed = 0;
}

ftl_status ftl_check_params(const ftl_params* p) {
  if (!p) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "params is NULL");
  if (p->iterations < 1 || p->iterations > 256)
    return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "iterations must be in [1, 256] (got %d)", p->iterations);
  if (p->velocity_mode != FTL_VELOCITY_SHIPPED && p->velocity_mode != FTL_VELOCITY_DFTL)
    return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "unknown velocity_mode %d", p->velocity_mode);
  const float vals[] = {p->gravity[0], p->gravity[1], p->gravity[2], p->force_scale, p->damping,
                        p->length_scale, p->max_velocity, p->wind_amplitude, p->wind_frequency,
                        p->wind_scroll[0], p->wind_scroll[1], p->wind_scroll[2]};
  for (float v : vals)
    if (!std::isfinite(v)) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "non-finite parameter");
  return FTL_OK;
}

// Wind lattice: FTL_WIND_GRID^3 float4 (xyz = amplitude * U[-1,1), w = 0),
// entry (z*G + y)*G + x, drawn from splitmix64 of the seed.
static uint64_t splitmix64(uint64_t x) {
  uint64_t z = x + 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

void ftl_wind_lattice(const ftl_params* p, float* out) {
  const int G = FTL_WIND_GRID;
  const float amp = p->wind_amplitude;
  for (int i = 0; i < G * G * G; ++i) {
    for (int c = 0; c < 3; ++c) {
      const uint64_t h = splitmix64(p->wind_seed * 0x632be59bd9b4e019ull + (uint64_t)(4 * i + c));
      const float u = (float)(h >> 40) * 0x1p-24f;  // [0, 1)
      out[4 * i + c] = amp * (2.0f * u - 1.0f);
    }
    out[4 * i + 3] = 0.0f;
  }
}

extern "C" ftl_status ftl_wind_table(const ftl_params* p, float* out, int* grid) {
  if (!p || !out) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "NULL argument");
  ftl_wind_lattice(p, out);
  if (grid) *grid = FTL_WIND_GRID;
  return FTL_OK;
}

// ---------------------------------------------------------------------------
// context

extern "C" ftl_status ftl_init(int device, ftl_ctx** out) {
  if (!out) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "out is NULL");
  *out = nullptr;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0) {
    (void)hipGetLastError();
    return ftl_fail(FTL_ERR_DEVICE, "no HIP device available (%s)", hipGetErrorString(e));
  }
  if (device < 0 || device >= n) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "device %d out of range [0,%d)", device, n);
  if ((e = hipSetDevice(device)) != hipSuccess) return ftl_hip_fail(e, "hipSetDevice");
  hipDeviceProp_t prop;
  if ((e = hipGetDeviceProperties(&prop, device)) != hipSuccess) return ftl_hip_fail(e, "hipGetDeviceProperties");
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return ftl_fail(FTL_ERR_UNSUPPORTED, "libftl is built for gfx950; device %d is %s", device, prop.gcnArchName);
  ftl_status st = ftl_noalloc_setup(device);
  if (st != FTL_OK) return st;
  ftl_ctx* c = new ftl_ctx();
  c->device = device;
  c->ncu = prop.multiProcessorCount;
  for (int i = 0; i <= FTL_MAX_POINTS; ++i) {
    c->lanes[i] = c->lanes_wind[i] = 0;
    c->pick[i] = c->pick_wind[i] = -1;
  }
  *out = c;
  return FTL_OK;
}

extern "C" ftl_status ftl_shutdown(ftl_ctx* ctx) {
  if (!ctx) return FTL_OK;
  (void)hipSetDevice(ctx->device);
  if (ctx->scratch_ev) {
    if (ctx->scratch_ev_recorded) (void)hipEventSynchronize(ctx->scratch_ev);
    (void)hipEventDestroy(ctx->scratch_ev);
  }
  if (ctx->scratch_steps) ftl_dev_free(ctx, ctx->scratch_steps, ctx->scratch_steps_bytes);
  if (ctx->scratch_spans) ftl_dev_free(ctx, ctx->scratch_spans, ctx->scratch_spans_bytes);
  if (ctx->live_objects != 0) {
    // Leaked handles keep their memory; the context itself is still released.
    ftl_fail(FTL_ERR_INVALID_ARGUMENT, "ftl_shutdown: %d objects still alive", ctx->live_objects);
    delete ctx;
    return FTL_ERR_INVALID_ARGUMENT;
  }
  delete ctx;
  return FTL_OK;
}

extern "C" ftl_status ftl_ctx_memory(ftl_ctx* ctx, uint64_t* bytes) {
  if (!ctx || !bytes) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "NULL argument");
  std::lock_guard<std::mutex> lk(ctx->mu);
  *bytes = ctx->bytes;
  return FTL_OK;
}

// Default lanes per strand for a point count when no tuning is stored.
static int default_lanes(int npoints) {
  switch (npoints) {
    case 16: return 2;
    case 32: return 2;
    case 64: return 4;
    default: return 1;
  }
}

int ftl_class_for(ftl_ctx* ctx, int npoints, int lanes, int mmax, bool wind) {
  // the tuned class of this point count (a base class or a tile-shape
  // variant), when it has the requested lanes
  if (ctx && npoints >= 0 && npoints <= FTL_MAX_POINTS) {
    const int p = (wind && ctx->pick_wind[npoints] >= 0) ? ctx->pick_wind[npoints] : ctx->pick[npoints];
    if (p >= 0 && p < FTL_NCLASSES && !kFtlClasses[p].guarded && kFtlClasses[p].m * kFtlClasses[p].g == npoints &&
        kFtlClasses[p].m <= mmax && (lanes == 0 || lanes == kFtlClasses[p].g))
      return p;
  }
  int G = lanes;
  if (G == 0 && ctx && npoints >= 0 && npoints <= FTL_MAX_POINTS)
    G = (wind && ctx->lanes_wind[npoints]) ? ctx->lanes_wind[npoints] : ctx->lanes[npoints];
  if (G == 0) G = default_lanes(npoints);
  // exact class with the requested lanes
  for (int c = 0; c < FTL_NCLASSES; ++c) {
    const FtlClassInfo& k = kFtlClasses[c];
    if (!k.guarded && k.g == G && k.m * k.g == npoints && k.m <= mmax) return c;
  }
  // any exact class within the register budget, fewest lanes first
  for (int g = 1; g <= 4; g *= 2)
    for (int c = 0; c < FTL_NCLASSES; ++c) {
      const FtlClassInfo& k = kFtlClasses[c];
      if (!k.guarded && k.g == g && k.m * k.g == npoints && k.m <= mmax) return c;
    }
  // guarded, two lanes per strand (unless one lane was asked for): the
  // smallest class whose first lane the strand fills (the least padding)
  if (lanes != 1) {
    int best2 = -1;
    for (int c = 0; c < FTL_NCLASSES; ++c) {
      const FtlClassInfo& k = kFtlClasses[c];
      if (k.guarded && k.g > 1 && (k.g - 1) * k.m < npoints && npoints <= k.g * k.m && k.m <= mmax &&
          (best2 < 0 || k.m < kFtlClasses[best2].m))
        best2 = c;
    }
    if (best2 >= 0) return best2;
  }
  // guarded, one lane per strand: the smallest that holds the strand
  int best = -1;
  for (int c = 0; c < FTL_NCLASSES; ++c) {
    const FtlClassInfo& k = kFtlClasses[c];
    if (k.guarded && k.g == 1 && k.m >= npoints && k.m <= mmax && (best < 0 || k.m < kFtlClasses[best].m)) best = c;
  }
  return best;
}

extern "C" int ftl_lanes_for(ftl_ctx* ctx, int npoints) {
  const int c = ftl_class_for(ctx, npoints, 0, FTL_MAX_POINTS);
  return c < 0 ? 0 : kFtlClasses[c].g;
}

extern "C" int ftl_lanes_for_wind(ftl_ctx* ctx, int npoints) {
  const int c = ftl_class_for(ctx, npoints, 0, FTL_MAX_POINTS, true);
  return c < 0 ? 0 : kFtlClasses[c].g;
}

int ftl_default_class(int npoints) { return ftl_class_for(nullptr, npoints, 0, FTL_MAX_POINTS); }

extern "C" int ftl_class_for_points(ftl_ctx* ctx, int npoints, int wind) {
  return ftl_class_for(ctx, npoints, 0, FTL_MAX_POINTS, wind != 0);
}

extern "C" ftl_status ftl_class_shape(int cls, int* points_per_lane, int* lanes, int* guarded, int* waves,
                                      int* chunk) {
  if (cls < 0 || cls >= FTL_NCLASSES) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "no launch class %d", cls);
  const FtlClassInfo& k = kFtlClasses[cls];
  if (points_per_lane) *points_per_lane = k.m;
  if (lanes) *lanes = k.g;
  if (guarded) *guarded = k.guarded;
  if (waves) *waves = k.waves;
  if (chunk) *chunk = k.chunk < k.m ? k.chunk : k.m;
  return FTL_OK;
}

// ---------------------------------------------------------------------------
// groom

static bool is_identity(const float m[16]) {
  for (int i = 0; i < 16; ++i)
    if (m[i] != ((i % 5 == 0) ? 1.0f : 0.0f)) return false;
  return true;
}

extern "C" ftl_status ftl_groom_create(ftl_ctx* ctx, const ftl_groom_desc* d, ftl_groom** out) {
  if (!ctx || !d || !out) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "NULL argument");
  if (!d->positions) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "positions is NULL");
  return ftl_groom_create_uninitialised(ctx, d, out);
}

// The groom of d; its state is copied from d->positions / d->velocities when
// given, zero otherwise (device-side initialisation fills it afterwards).
ftl_status ftl_groom_create_uninitialised(ftl_ctx* ctx, const ftl_groom_desc* d, ftl_groom** out) {
  if (!ctx || !d || !out) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "NULL argument");
  *out = nullptr;
  if (d->nstrands <= 0) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "nstrands must be > 0 (got %d)", d->nstrands);
  if (d->npoints < 2 || d->npoints > FTL_MAX_POINTS)
    return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "npoints must be in [2, %d] (got %d)", FTL_MAX_POINTS, d->npoints);
  const size_t n = (size_t)d->nstrands * (size_t)d->npoints;
  if (n > (size_t)INT32_MAX * 4) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "groom too large");
  hipError_t e;
  if ((e = hipSetDevice(ctx->device)) != hipSuccess) return ftl_hip_fail(e, "hipSetDevice");
  ftl_groom* g = new ftl_groom();
  g->ctx = ctx;
  g->nstrands = d->nstrands;
  g->npoints = d->npoints;
  ftl_params_reference(&g->params);
  g->ncol = 1;  // uBoundingSphere default (0,0,0,1) (cs_simulation.glsl:43)
  std::memset(g->sph, 0, sizeof g->sph);
  g->sph[0][3] = 1.0f;
  for (int i = 0; i < 16; ++i) g->model[i] = (i % 5 == 0) ? 1.0f : 0.0f;
  g->model_ident = 1;
  g->root_mode = FTL_ROOT_MODEL;
  const size_t bytes = n * sizeof(float4);
  ftl_status st;
  if ((st = ftl_dev_alloc(ctx, (void**)&g->d_pos, bytes, "positions")) != FTL_OK ||
      (st = ftl_dev_alloc(ctx, (void**)&g->d_vel, bytes, "velocities")) != FTL_OK ||
      (st = ftl_dev_alloc(ctx, (void**)&g->d_status, 256, "status")) != FTL_OK) {
    ftl_dev_free(ctx, g->d_pos, bytes);
    ftl_dev_free(ctx, g->d_vel, bytes);
    ftl_dev_free(ctx, g->d_status, 256);
    delete g;
    return st;
  }
  bool ok = d->positions ? hipMemcpy(g->d_pos, d->positions, bytes, hipMemcpyHostToDevice) == hipSuccess
                         : hipMemset(g->d_pos, 0, bytes) == hipSuccess;
  if (d->velocities)
    ok = ok && hipMemcpy(g->d_vel, d->velocities, bytes, hipMemcpyHostToDevice) == hipSuccess;
  else
    ok = ok && hipMemset(g->d_vel, 0, bytes) == hipSuccess;
  ok = ok && hipMemset(g->d_status, 0, 256) == hipSuccess;
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->live_objects++;
  }
  if (!ok) {
    ftl_groom_destroy(g);
    return ftl_fail(FTL_ERR_DEVICE, "upload of the initial state failed");
  }
  if (d->rest_roots) {
    if ((st = ftl_groom_set_rest_roots(g, d->rest_roots)) != FTL_OK) {
      ftl_groom_destroy(g);
      return st;
    }
  }
  *out = g;
  return FTL_OK;
}

// Takes g off its skin's list of bound grooms (ctx->mu held).
static void unbind_skin_locked(ftl_groom* g) {
  if (!g->skin) return;
  std::vector<ftl_groom*>& b = g->skin->bound;
  b.erase(std::remove(b.begin(), b.end(), g), b.end());
  g->skin = nullptr;
}

extern "C" void ftl_groom_destroy(ftl_groom* g) {
  if (!g) return;
  ftl_ctx* ctx = g->ctx;
  (void)hipSetDevice(ctx->device);
  (void)hipDeviceSynchronize();
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    unbind_skin_locked(g);
    for (ftl_batch* b : g->batches) ftl_batch_forget_groom_locked(b, g);
    g->batches.clear();
  }
  const size_t bytes = (size_t)g->nstrands * g->npoints * sizeof(float4);
  ftl_dev_free(ctx, g->d_pos, bytes);
  ftl_dev_free(ctx, g->d_vel, bytes);
  ftl_dev_free(ctx, g->d_status, 256);
  ftl_dev_free(ctx, g->d_roots, (size_t)g->nstrands * sizeof(float4));
  ftl_dev_free(ctx, g->d_bind, (size_t)g->nstrands * sizeof(int4));
  ftl_dev_free(ctx, g->d_wind, (size_t)FTL_WIND_GRID * FTL_WIND_GRID * FTL_WIND_GRID * 2 * sizeof(float4));
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->live_objects--;
  }
  delete g;
}

extern "C" ftl_status ftl_groom_set_params(ftl_groom* g, const ftl_params* p) {
  if (!g) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "groom is NULL");
  ftl_status st = ftl_check_params(p);
  if (st != FTL_OK) return st;
  g->params = *p;
  return FTL_OK;
}

extern "C" ftl_status ftl_groom_get_params(ftl_groom* g, ftl_params* p) {
  if (!g || !p) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "NULL argument");
  *p = g->params;
  return FTL_OK;
}

extern "C" ftl_status ftl_groom_set_colliders(ftl_groom* g, const float* s, int count) {
  if (!g) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "groom is NULL");
  if (count < 0 || count > FTL_MAX_COLLIDERS)
    return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "collider count must be in [0, %d] (got %d)", FTL_MAX_COLLIDERS, count);
  if (count > 0 && !s) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "spheres is NULL");
  for (int i = 0; i < 4 * count; ++i)
    if (!std::isfinite(s[i])) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "non-finite collider");
  g->ncol = count;
  std::memset(g->sph, 0, sizeof g->sph);
  for (int k = 0; k < count; ++k)
    for (int c = 0; c < 4; ++c) g->sph[k][c] = s[4 * k + c];
  return FTL_OK;
}

extern "C" ftl_status ftl_groom_set_root_transform(ftl_groom* g, const float m[16]) {
  if (!g || !m) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "NULL argument");
  if (m[3] != 0.0f || m[7] != 0.0f || m[11] != 0.0f || m[15] != 1.0f)
    return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "root transform must be affine (last row 0,0,0,1)");
  for (int i = 0; i < 16; ++i)
    if (!std::isfinite(m[i])) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "non-finite transform");
  std::memcpy(g->model, m, sizeof g->model);
  g->model_ident = is_identity(m) ? 1 : 0;
  return FTL_OK;
}

extern "C" ftl_status ftl_groom_set_root_mode(ftl_groom* g, int mode) {
  if (!g) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "groom is NULL");
  if (mode == FTL_ROOT_REST && !g->d_roots)
    return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "FTL_ROOT_REST needs rest roots (ftl_groom_set_rest_roots)");
  if (mode == FTL_ROOT_SKINNED && !g->skin)
    return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "FTL_ROOT_SKINNED needs ftl_groom_bind_skin");
  if (mode != FTL_ROOT_MODEL && mode != FTL_ROOT_REST && mode != FTL_ROOT_SKINNED)
    return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "unknown root mode %d", mode);
  g->root_mode = mode;
  return FTL_OK;
}

extern "C" ftl_status ftl_groom_set_rest_roots(ftl_groom* g, const float* roots) {
  if (!g || !roots) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "NULL argument");
  const size_t bytes = (size_t)g->nstrands * sizeof(float4);
  hipError_t e;
  if ((e = hipSetDevice(g->ctx->device)) != hipSuccess) return ftl_hip_fail(e, "hipSetDevice");
  if (!g->d_roots) {
    ftl_status st = ftl_dev_alloc(g->ctx, (void**)&g->d_roots, bytes, "rest roots");
    if (st != FTL_OK) return st;
  }
  if ((e = hipDeviceSynchronize()) != hipSuccess) return ftl_hip_fail(e, "hipDeviceSynchronize");
  if ((e = hipMemcpy(g->d_roots, roots, bytes, hipMemcpyHostToDevice)) != hipSuccess)
    return ftl_hip_fail(e, "hipMemcpy rest roots");
  return FTL_OK;
}

extern "C" ftl_status ftl_groom_set_time(ftl_groom* g, double seconds) {
  if (!g) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "groom is NULL");
  if (!std::isfinite(seconds)) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "non-finite time");
  g->time = seconds;
  return FTL_OK;
}

extern "C" ftl_status ftl_groom_set_lanes(ftl_groom* g, int lanes) {
  if (!g) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "groom is NULL");
  if (lanes != 0 && lanes != 1 && lanes != 2 && lanes != 4)
    return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "lanes per strand must be 0, 1, 2 or 4 (got %d)", lanes);
  g->lanes = lanes;
  return FTL_OK;
}

// Uploads the wind lattice when the wind parameters changed, in the device
// layout of the kernels: per node (v.x, v.y, v.z, d.x), (d.y, d.z, 0, 0) with
// d = v(x+1) - v(x) along the periodic x axis.
static ftl_status ensure_wind(ftl_groom* g, bool may_alloc) {
  const ftl_params& p = g->params;
  if (p.wind_amplitude == 0.0f) return FTL_OK;
  if (g->d_wind && g->wind_seed == p.wind_seed && g->wind_amp == p.wind_amplitude) return FTL_OK;
  const int G = FTL_WIND_GRID;
  const size_t cells = (size_t)G * G * G;
  const size_t bytes = cells * 2 * sizeof(float4);
  if (!g->d_wind && !may_alloc)
    return ftl_fail(FTL_ERR_INVALID_ARGUMENT,
                    "the groom's wind lattice is not on the device yet (a step that may allocate, such as "
                    "ftl_groom_step or ftl_batch_step, places it)");
  if (!g->d_wind) {
    ftl_status st = ftl_dev_alloc(g->ctx, (void**)&g->d_wind, bytes, "wind lattice");
    if (st != FTL_OK) return st;
  }
  std::vector<float> h(cells * 4), t(cells * 8, 0.0f);
  ftl_wind_lattice(&p, h.data());
  for (size_t n = 0; n < cells; ++n) {
    const size_t x = n % G, row = n - x;
    const float* a = &h[4 * n];
    const float* b = &h[4 * (row + (x + 1) % G)];
    float* o = &t[8 * n];
    o[0] = a[0];
    o[1] = a[1];
    o[2] = a[2];
    o[3] = b[0] - a[0];
    o[4] = b[1] - a[1];
    o[5] = b[2] - a[2];
  }
  hipError_t e;
  if ((e = hipDeviceSynchronize()) != hipSuccess) return ftl_hip_fail(e, "hipDeviceSynchronize");
  if ((e = hipMemcpy(g->d_wind, t.data(), bytes, hipMemcpyHostToDevice)) != hipSuccess)
    return ftl_hip_fail(e, "hipMemcpy wind lattice");
  g->wind_seed = p.wind_seed;
  g->wind_amp = p.wind_amplitude;
  return FTL_OK;
}

ftl_status ftl_make_step(ftl_groom* g, float dt, FtlStep* A, bool may_alloc) {
  if (!(dt > 0.0f) || !std::isfinite(dt)) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "dt must be finite and > 0");
  ftl_status st = ensure_wind(g, may_alloc);
  if (st != FTL_OK) return st;
  const ftl_params& p = g->params;
  std::memset(A, 0, sizeof *A);
  A->pos = (float4*)g->d_pos;
  A->vel = (float4*)g->d_vel;
  A->roots = (const float4*)g->d_roots;
  A->status = g->d_status;
  A->nstrands = g->nstrands;
  A->npoints = g->npoints;
  A->iterations = p.iterations;
  A->vmode = p.velocity_mode;
  A->rmode = g->root_mode;
  if (g->root_mode == FTL_ROOT_SKINNED) {
    if (!g->skin || !g->d_bind) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "skinned roots without a bound skin");
    A->bind = (const int4*)g->d_bind;
    A->tris = (const int4*)g->skin->d_tris;
    A->skinned = (const float4*)g->skin->d_out;
  }
  if (g->root_mode == FTL_ROOT_REST && !g->d_roots) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "no rest roots");
  A->ncol = g->ncol;
  A->model_ident = g->model_ident;
  A->dt = dt;
  A->dt2 = dt * dt;
  A->inv_dt = 1.0f / dt;
  A->sigma = p.length_scale;
  A->delta = p.damping;
  A->maxv = p.max_velocity;
  A->maxv2 = p.max_velocity * p.max_velocity;
  for (int c = 0; c < 3; ++c) A->force[c] = p.force_scale * p.gravity[c];
  for (int k = 0; k < g->ncol; ++k) {
    for (int c = 0; c < 4; ++c) A->sph[k][c] = g->sph[k][c];
    A->r2[k] = g->sph[k][3] * g->sph[k][3];
  }
  // column-major 4x4 -> row-major 3x4
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 4; ++c) A->model[4 * r + c] = g->model[4 * c + r];
  if (p.wind_amplitude != 0.0f) {
    A->wind = (const float4*)g->d_wind;
    A->wind_grid = FTL_WIND_GRID;
    A->wind_freq = p.wind_frequency;
    for (int c = 0; c < 3; ++c) A->wind_off[c] = (float)((double)p.wind_scroll[c] * g->time);
  }
  return FTL_OK;
}

static ftl_status groom_step(ftl_groom* g, float dt, void* stream, hipEvent_t start, hipEvent_t stop) {
  if (!g) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "groom is NULL");
  FtlStep A;
  ftl_status st = ftl_make_step(g, dt, &A);
  if (st != FTL_OK) return st;
  const int cls = ftl_class_for(g->ctx, g->npoints, g->lanes, FTL_MAX_POINTS, A.wind_grid != 0);
  if (cls < 0) return ftl_fail(FTL_ERR_UNSUPPORTED, "no launch class for npoints=%d", g->npoints);
  hipError_t e;
  if ((e = hipSetDevice(g->ctx->device)) != hipSuccess) return ftl_hip_fail(e, "hipSetDevice");
  const bool nt = (size_t)g->nstrands * g->npoints * 2 * sizeof(float4) > ftl_stream_store_bytes();
  if ((e = ftl_launch_plain(cls, A, (hipStream_t)stream, start, stop, nt)) != hipSuccess)
    return ftl_hip_fail(e, "FTL step launch");
  return FTL_OK;
}

extern "C" ftl_status ftl_groom_step(ftl_groom* g, float dt, void* stream) {
  return groom_step(g, dt, stream, nullptr, nullptr);
}

extern "C" ftl_status ftl_groom_step_timed(ftl_groom* g, float dt, void* stream, void* start_event,
                                           void* stop_event) {
  if (!start_event || !stop_event) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "NULL event");
  return groom_step(g, dt, stream, (hipEvent_t)start_event, (hipEvent_t)stop_event);
}

extern "C" ftl_status ftl_groom_device_buffers(ftl_groom* g, void** pos, void** vel) {
  if (!g) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "groom is NULL");
  if (pos) *pos = g->d_pos;
  if (vel) *vel = g->d_vel;
  return FTL_OK;
}

extern "C" ftl_status ftl_groom_download(ftl_groom* g, float* pos, float* vel) {
  if (!g) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "groom is NULL");
  const size_t bytes = (size_t)g->nstrands * g->npoints * sizeof(float4);
  hipError_t e;
  if ((e = hipSetDevice(g->ctx->device)) != hipSuccess) return ftl_hip_fail(e, "hipSetDevice");
  if ((e = hipDeviceSynchronize()) != hipSuccess) return ftl_hip_fail(e, "hipDeviceSynchronize");
  if (pos && (e = hipMemcpy(pos, g->d_pos, bytes, hipMemcpyDeviceToHost)) != hipSuccess)
    return ftl_hip_fail(e, "hipMemcpy positions");
  if (vel && (e = hipMemcpy(vel, g->d_vel, bytes, hipMemcpyDeviceToHost)) != hipSuccess)
    return ftl_hip_fail(e, "hipMemcpy velocities");
  return FTL_OK;
}

extern "C" ftl_status ftl_groom_upload(ftl_groom* g, const float* pos, const float* vel) {
  if (!g) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "groom is NULL");
  const size_t bytes = (size_t)g->nstrands * g->npoints * sizeof(float4);
  hipError_t e;
  if ((e = hipSetDevice(g->ctx->device)) != hipSuccess) return ftl_hip_fail(e, "hipSetDevice");
  if ((e = hipDeviceSynchronize()) != hipSuccess) return ftl_hip_fail(e, "hipDeviceSynchronize");
  if (pos && (e = hipMemcpy(g->d_pos, pos, bytes, hipMemcpyHostToDevice)) != hipSuccess)
    return ftl_hip_fail(e, "hipMemcpy positions");
  if (vel && (e = hipMemcpy(g->d_vel, vel, bytes, hipMemcpyHostToDevice)) != hipSuccess)
    return ftl_hip_fail(e, "hipMemcpy velocities");
  return FTL_OK;
}

extern "C" ftl_status ftl_groom_info(ftl_groom* g, int* nstrands, int* npoints) {
  if (!g) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "groom is NULL");
  if (nstrands) *nstrands = g->nstrands;
  if (npoints) *npoints = g->npoints;
  return FTL_OK;
}

extern "C" ftl_status ftl_groom_status(ftl_groom* g, int* nonfinite) {
  if (!g || !nonfinite) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "NULL argument");
  hipError_t e;
  if ((e = hipSetDevice(g->ctx->device)) != hipSuccess) return ftl_hip_fail(e, "hipSetDevice");
  if ((e = hipDeviceSynchronize()) != hipSuccess) return ftl_hip_fail(e, "hipDeviceSynchronize");
  unsigned v = 0;
  if ((e = hipMemcpy(&v, g->d_status, sizeof v, hipMemcpyDeviceToHost)) != hipSuccess)
    return ftl_hip_fail(e, "hipMemcpy status");
  if ((e = hipMemset(g->d_status, 0, sizeof v)) != hipSuccess) return ftl_hip_fail(e, "hipMemset status");
  *nonfinite = (v & 1u) ? 1 : 0;
  return FTL_OK;
}

// ---------------------------------------------------------------------------
// skin

extern "C" ftl_status ftl_skin_create(ftl_ctx* ctx, const ftl_skin_desc* d, ftl_skin** out) {
  if (!ctx || !d || !out) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "NULL argument");
  *out = nullptr;
  if (d->nverts <= 0 || d->ntris <= 0 || d->nbones <= 0 || d->nbones > 4096)
    return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "skin sizes out of range");
  if (!d->rest_positions || !d->joints || !d->weights || !d->triangles)
    return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "NULL skin array");
  bool no_joint = false;
  for (long i = 0; i < 4L * d->nverts; ++i) {
    if (d->joints[i] == FTL_NO_JOINT) {
      no_joint = true;
      continue;
    }
    if (d->joints[i] < 0 || d->joints[i] >= d->nbones)
      return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "joint index %d out of range at vertex %ld", d->joints[i], i / 4);
  }
  for (long i = 0; i < 3L * d->ntris; ++i)
    if (d->triangles[i] < 0 || d->triangles[i] >= d->nverts)
      return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "vertex index out of range in triangle %ld", i / 3);
  hipError_t e;
  if ((e = hipSetDevice(ctx->device)) != hipSuccess) return ftl_hip_fail(e, "hipSetDevice");
  ftl_skin* s = new ftl_skin();
  s->ctx = ctx;
  s->nverts = d->nverts;
  s->ntris = d->ntris;
  s->nbones = d->nbones;
  s->has_no_joint = no_joint;
  const size_t vb = (size_t)d->nverts * 16, tb = (size_t)d->ntris * 16, bb = (size_t)d->nbones * 32,
               mb = (size_t)d->nbones * 48;
  std::vector<int> tris4((size_t)d->ntris * 4, 0);
  for (int t = 0; t < d->ntris; ++t)
    for (int c = 0; c < 3; ++c) tris4[4 * (size_t)t + c] = d->triangles[3 * (size_t)t + c];
  ftl_status st = FTL_OK;
  if ((st = ftl_dev_alloc(ctx, (void**)&s->d_rest, vb, "skin rest")) != FTL_OK ||
      (st = ftl_dev_alloc(ctx, (void**)&s->d_joints, vb, "skin joints")) != FTL_OK ||
      (st = ftl_dev_alloc(ctx, (void**)&s->d_weights, vb, "skin weights")) != FTL_OK ||
      (st = ftl_dev_alloc(ctx, (void**)&s->d_out, vb, "skin out")) != FTL_OK ||
      (st = ftl_dev_alloc(ctx, (void**)&s->d_tris, tb, "skin tris")) != FTL_OK ||
      (st = ftl_dev_alloc(ctx, (void**)&s->d_dq, bb, "skin bones")) != FTL_OK ||
      (st = ftl_dev_alloc(ctx, (void**)&s->d_mats, mb, "skin matrices")) != FTL_OK) {
    {
      std::lock_guard<std::mutex> lk(ctx->mu);
      ctx->live_objects++;
    }
    ftl_skin_destroy(s);
    return st;
  }
  bool ok = hipMemcpy(s->d_rest, d->rest_positions, vb, hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(s->d_joints, d->joints, vb, hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(s->d_weights, d->weights, vb, hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(s->d_out, d->rest_positions, vb, hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(s->d_tris, tris4.data(), tb, hipMemcpyHostToDevice) == hipSuccess;
  // identity bones until they are set: dual quaternions (1, 0) and matrices [I | 0]
  s->h_dq.assign((size_t)d->nbones * 8, 0.0f);
  s->h_mats.assign((size_t)d->nbones * 12, 0.0f);
  for (int b = 0; b < d->nbones; ++b) {
    s->h_dq[8 * (size_t)b + 3] = 1.0f;
    for (int r = 0; r < 3; ++r) s->h_mats[12 * (size_t)b + 5 * r] = 1.0f;
  }
  ok = ok && hipMemcpy(s->d_dq, s->h_dq.data(), bb, hipMemcpyHostToDevice) == hipSuccess &&
       hipMemcpy(s->d_mats, s->h_mats.data(), mb, hipMemcpyHostToDevice) == hipSuccess;
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->live_objects++;
  }
  if (!ok) {
    ftl_skin_destroy(s);
    return ftl_fail(FTL_ERR_DEVICE, "skin upload failed");
  }
  *out = s;
  return FTL_OK;
}

extern "C" void ftl_skin_destroy(ftl_skin* s) {
  if (!s) return;
  ftl_ctx* ctx = s->ctx;
  (void)hipSetDevice(ctx->device);
  (void)hipDeviceSynchronize();
  {
    // Its grooms lose their binding: a step with skinned roots then fails
    // (no bound skin) instead of reading freed memory.
    std::lock_guard<std::mutex> lk(ctx->mu);
    for (ftl_groom* g : s->bound) g->skin = nullptr;
    s->bound.clear();
  }
  const size_t vb = (size_t)s->nverts * 16, tb = (size_t)s->ntris * 16, bb = (size_t)s->nbones * 32,
               mb = (size_t)s->nbones * 48;
  ftl_dev_free(ctx, s->d_rest, vb);
  ftl_dev_free(ctx, s->d_joints, vb);
  ftl_dev_free(ctx, s->d_weights, vb);
  ftl_dev_free(ctx, s->d_out, vb);
  ftl_dev_free(ctx, s->d_tris, tb);
  ftl_dev_free(ctx, s->d_dq, bb);
  ftl_dev_free(ctx, s->d_mats, mb);
  ftl_dev_free(ctx, s->d_nrest, vb);
  ftl_dev_free(ctx, s->d_nout, vb);
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->live_objects--;
  }
  delete s;
}

extern "C" ftl_status ftl_skin_set_mode(ftl_skin* s, int mode) {
  if (!s) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "skin is NULL");
  if (mode != FTL_SKIN_DQ && mode != FTL_SKIN_LBS)
    return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "unknown skinning mode %d", mode);
  if (mode == FTL_SKIN_DQ && s->has_no_joint)
    return ftl_fail(FTL_ERR_INVALID_ARGUMENT,
                    "dual-quaternion skinning needs a bone for every joint index (FTL_NO_JOINT is linear-blend only)");
  s->mode = mode;
  return FTL_OK;
}

extern "C" ftl_status ftl_skin_set_bones(ftl_skin* s, const float* dq, int nbones) {
  if (!s || !dq) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "NULL argument");
  if (nbones != s->nbones) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "expected %d bones, got %d", s->nbones, nbones);
  for (long i = 0; i < 8L * nbones; ++i)
    if (!std::isfinite(dq[i])) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "non-finite bone data");
  std::memcpy(s->h_dq.data(), dq, (size_t)nbones * 32);
  s->dirty = true;
  return FTL_OK;
}

// The unit quaternion of rotation R (rows R[r][c]), by the largest of
// w, x, y, z (Shepperd's method: the square root is taken of the largest
// diagonal combination, so no division by a small number), then the dual part
// 0.5 (t, 0) q.  Double precision, rounded to float at the end.
extern "C" ftl_status ftl_dualquat_from_matrix(const float* m, float* dq) {
  if (!m || !dq) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "NULL argument");
  double R[3][3], t[3];
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) R[r][c] = (double)m[4 * r + c];
    t[r] = (double)m[4 * r + 3];
  }
  for (int i = 0; i < 12; ++i)
    if (!std::isfinite(m[i])) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "non-finite matrix");
  double q[4];  // x, y, z, w
  const double tr = R[0][0] + R[1][1] + R[2][2];
  if (tr > 0.0) {
    const double s2 = std::sqrt(1.0 + tr), f = 0.5 / s2;
    q[3] = 0.5 * s2;
    q[0] = (R[2][1] - R[1][2]) * f;
    q[1] = (R[0][2] - R[2][0]) * f;
    q[2] = (R[1][0] - R[0][1]) * f;
  } else if (R[0][0] > R[1][1] && R[0][0] > R[2][2]) {
    const double s2 = std::sqrt(1.0 + R[0][0] - R[1][1] - R[2][2]), f = 0.5 / s2;
    q[0] = 0.5 * s2;
    q[1] = (R[0][1] + R[1][0]) * f;
    q[2] = (R[0][2] + R[2][0]) * f;
    q[3] = (R[2][1] - R[1][2]) * f;
  } else if (R[1][1] > R[2][2]) {
    const double s2 = std::sqrt(1.0 + R[1][1] - R[0][0] - R[2][2]), f = 0.5 / s2;
    q[1] = 0.5 * s2;
    q[0] = (R[0][1] + R[1][0]) * f;
    q[2] = (R[1][2] + R[2][1]) * f;
    q[3] = (R[0][2] - R[2][0]) * f;
  } else {
    const double s2 = std::sqrt(1.0 + R[2][2] - R[0][0] - R[1][1]), f = 0.5 / s2;
    q[2] = 0.5 * s2;
    q[0] = (R[0][2] + R[2][0]) * f;
    q[1] = (R[1][2] + R[2][1]) * f;
    q[3] = (R[1][0] - R[0][1]) * f;
  }
  // dual = 0.5 (t, 0) * q: w = -0.5 t.q_xyz, xyz = 0.5 (t q_w + t x q_xyz)
  const double d[4] = {0.5 * (t[0] * q[3] + t[1] * q[2] - t[2] * q[1]),
                       0.5 * (t[1] * q[3] + t[2] * q[0] - t[0] * q[2]),
                       0.5 * (t[2] * q[3] + t[0] * q[1] - t[1] * q[0]),
                       -0.5 * (t[0] * q[0] + t[1] * q[1] + t[2] * q[2])};
  for (int c = 0; c < 4; ++c) {
    dq[c] = (float)q[c];
    dq[4 + c] = (float)d[c];
  }
  return FTL_OK;
}

extern "C" ftl_status ftl_skin_set_matrices(ftl_skin* s, const float* m3x4, int nbones) {
  if (!s || !m3x4) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "NULL argument");
  if (nbones != s->nbones) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "expected %d bones, got %d", s->nbones, nbones);
  for (long i = 0; i < 12L * nbones; ++i)
    if (!std::isfinite(m3x4[i])) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "non-finite bone matrix");
  std::vector<float> dq((size_t)nbones * 8);
  for (int b = 0; b < nbones; ++b) {
    const ftl_status st = ftl_dualquat_from_matrix(m3x4 + 12 * (size_t)b, dq.data() + 8 * (size_t)b);
    if (st != FTL_OK) return st;
  }
  std::memcpy(s->h_mats.data(), m3x4, (size_t)nbones * 48);
  std::memcpy(s->h_dq.data(), dq.data(), (size_t)nbones * 32);
  s->dirty = s->mats_dirty = true;
  return FTL_OK;
}

extern "C" ftl_status ftl_skin_set_joint_matrices(ftl_skin* s, const float* global_pose, const float* inverse_bind,
                                                  int nbones) {
  if (!s || !global_pose) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "NULL argument");
  if (nbones != s->nbones) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "expected %d bones, got %d", s->nbones, nbones);
  std::vector<float> rows((size_t)nbones * 12);
  for (int b = 0; b < nbones; ++b) {
    const float* G = global_pose + 16 * (size_t)b;  // column-major: G[4 * col + row]
    const float* I = inverse_bind ? inverse_bind + 16 * (size_t)b : nullptr;
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 4; ++c) {
        double acc = 0.0;
        for (int k = 0; k < 4; ++k) {
          const double ik = I ? (double)I[4 * c + k] : (k == c ? 1.0 : 0.0);
          acc += (double)G[4 * k + r] * ik;
        }
        rows[12 * (size_t)b + 4 * r + c] = (float)acc;
      }
  }
  return ftl_skin_set_matrices(s, rows.data(), nbones);
}

extern "C" ftl_status ftl_skin_set_rest_normals(ftl_skin* s, const float* normals) {
  if (!s || !normals) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "NULL argument");
  for (long i = 0; i < 4L * s->nverts; ++i)
    if (!std::isfinite(normals[i])) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "non-finite normal");
  const size_t vb = (size_t)s->nverts * 16;
  hipError_t e;
  if ((e = hipSetDevice(s->ctx->device)) != hipSuccess) return ftl_hip_fail(e, "hipSetDevice");
  ftl_status st;
  if (!s->d_nrest && (st = ftl_dev_alloc(s->ctx, &s->d_nrest, vb, "skin rest normals")) != FTL_OK) return st;
  if (!s->d_nout && (st = ftl_dev_alloc(s->ctx, &s->d_nout, vb, "skin normals")) != FTL_OK) return st;
  if ((e = hipDeviceSynchronize()) != hipSuccess) return ftl_hip_fail(e, "hipDeviceSynchronize");
  if ((e = hipMemcpy(s->d_nrest, normals, vb, hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(s->d_nout, normals, vb, hipMemcpyHostToDevice)) != hipSuccess)
    return ftl_hip_fail(e, "hipMemcpy normals");
  return FTL_OK;
}

extern "C" ftl_status ftl_skin_update(ftl_skin* s, void* stream) {
  if (!s) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "skin is NULL");
  hipError_t e;
  if ((e = hipSetDevice(s->ctx->device)) != hipSuccess) return ftl_hip_fail(e, "hipSetDevice");
  if (s->dirty || s->mats_dirty) {
    // Synchronous copies: the bones are tiny and the host arrays are reused.
    if ((e = hipStreamSynchronize((hipStream_t)stream)) != hipSuccess) return ftl_hip_fail(e, "hipStreamSynchronize");
    if (s->dirty && (e = hipMemcpy(s->d_dq, s->h_dq.data(), (size_t)s->nbones * 32, hipMemcpyHostToDevice)) !=
                        hipSuccess)
      return ftl_hip_fail(e, "hipMemcpy bones");
    if (s->mats_dirty &&
        (e = hipMemcpy(s->d_mats, s->h_mats.data(), (size_t)s->nbones * 48, hipMemcpyHostToDevice)) != hipSuccess)
      return ftl_hip_fail(e, "hipMemcpy bone matrices");
    s->dirty = s->mats_dirty = false;
  }
  const float4* bones = (const float4*)(s->mode == FTL_SKIN_LBS ? s->d_mats : s->d_dq);
  if ((e = ftl_launch_skin(s->mode, s->nverts, bones, (const int4*)s->d_joints, (const float4*)s->d_weights,
                           (const float4*)s->d_rest, (const float4*)s->d_nrest, (float4*)s->d_out,
                           (float4*)s->d_nout, (hipStream_t)stream)) != hipSuccess)
    return ftl_hip_fail(e, "skinning launch");
  return FTL_OK;
}

extern "C" ftl_status ftl_skin_download(ftl_skin* s, float* verts) {
  if (!s || !verts) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "NULL argument");
  hipError_t e;
  if ((e = hipSetDevice(s->ctx->device)) != hipSuccess) return ftl_hip_fail(e, "hipSetDevice");
  if ((e = hipDeviceSynchronize()) != hipSuccess) return ftl_hip_fail(e, "hipDeviceSynchronize");
  if ((e = hipMemcpy(verts, s->d_out, (size_t)s->nverts * 16, hipMemcpyDeviceToHost)) != hipSuccess)
    return ftl_hip_fail(e, "hipMemcpy skinned vertices");
  return FTL_OK;
}

extern "C" ftl_status ftl_skin_download_normals(ftl_skin* s, float* normals) {
  if (!s || !normals) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "NULL argument");
  if (!s->d_nout) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "the skin has no normals (ftl_skin_set_rest_normals)");
  hipError_t e;
  if ((e = hipSetDevice(s->ctx->device)) != hipSuccess) return ftl_hip_fail(e, "hipSetDevice");
  if ((e = hipDeviceSynchronize()) != hipSuccess) return ftl_hip_fail(e, "hipDeviceSynchronize");
  if ((e = hipMemcpy(normals, s->d_nout, (size_t)s->nverts * 16, hipMemcpyDeviceToHost)) != hipSuccess)
    return ftl_hip_fail(e, "hipMemcpy skinned normals");
  return FTL_OK;
}

extern "C" ftl_status ftl_groom_root_frames(ftl_groom* g, float* roots, float* normals) {
  if (!g || !roots) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "NULL argument");
  if (!g->skin || !g->d_bind) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "the groom is not bound to a skin");
  ftl_skin* s = g->skin;
  if (normals && !s->d_nout)
    return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "normals asked for, but the skin has none (ftl_skin_set_rest_normals)");
  const size_t bytes = (size_t)g->nstrands * 16;
  hipError_t e;
  if ((e = hipSetDevice(g->ctx->device)) != hipSuccess) return ftl_hip_fail(e, "hipSetDevice");
  void *d_r = nullptr, *d_n = nullptr;
  ftl_status st;
  if ((st = ftl_dev_alloc(g->ctx, &d_r, bytes, "root frames")) != FTL_OK) return st;
  if (normals && (st = ftl_dev_alloc(g->ctx, &d_n, bytes, "root normals")) != FTL_OK) {
    ftl_dev_free(g->ctx, d_r, bytes);
    return st;
  }
  st = FTL_OK;
  if ((e = hipDeviceSynchronize()) != hipSuccess ||
      (e = ftl_launch_root_frames(g->nstrands, (const int4*)g->d_bind, (const int4*)s->d_tris,
                                  (const float4*)s->d_out, normals ? (const float4*)s->d_nout : nullptr,
                                  (float4*)d_r, (float4*)d_n, nullptr)) != hipSuccess ||
      (e = hipDeviceSynchronize()) != hipSuccess || (e = hipMemcpy(roots, d_r, bytes, hipMemcpyDeviceToHost)) != hipSuccess ||
      (normals && (e = hipMemcpy(normals, d_n, bytes, hipMemcpyDeviceToHost)) != hipSuccess))
    st = ftl_hip_fail(e, "root frames");
  ftl_dev_free(g->ctx, d_r, bytes);
  ftl_dev_free(g->ctx, d_n, bytes);
  return st;
}

extern "C" ftl_status ftl_groom_bind_skin(ftl_groom* g, ftl_skin* s, const int* tri, const float* bary) {
  if (!g || !s || !tri || !bary) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "NULL argument");
  if (s->ctx != g->ctx) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "skin and groom belong to different contexts");
  std::vector<int> b((size_t)g->nstrands * 4);
  for (int i = 0; i < g->nstrands; ++i) {
    if (tri[i] < 0 || tri[i] >= s->ntris)
      return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "triangle index %d out of range at strand %d", tri[i], i);
    if (!std::isfinite(bary[2 * (size_t)i]) || !std::isfinite(bary[2 * (size_t)i + 1]))
      return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "non-finite barycentric at strand %d", i);
    b[4 * (size_t)i] = tri[i];
    std::memcpy(&b[4 * (size_t)i + 1], &bary[2 * (size_t)i], 4);
    std::memcpy(&b[4 * (size_t)i + 2], &bary[2 * (size_t)i + 1], 4);
    b[4 * (size_t)i + 3] = 0;
  }
  const size_t bytes = (size_t)g->nstrands * sizeof(int4);
  hipError_t e;
  if ((e = hipSetDevice(g->ctx->device)) != hipSuccess) return ftl_hip_fail(e, "hipSetDevice");
  if (!g->d_bind) {
    ftl_status st = ftl_dev_alloc(g->ctx, (void**)&g->d_bind, bytes, "skin binding");
    if (st != FTL_OK) return st;
  }
  if ((e = hipDeviceSynchronize()) != hipSuccess) return ftl_hip_fail(e, "hipDeviceSynchronize");
  if ((e = hipMemcpy(g->d_bind, b.data(), bytes, hipMemcpyHostToDevice)) != hipSuccess)
    return ftl_hip_fail(e, "hipMemcpy binding");
  {
    std::lock_guard<std::mutex> lk(g->ctx->mu);
    if (g->skin != s) {
      unbind_skin_locked(g);
      g->skin = s;
      s->bound.push_back(g);
    }
  }
  g->root_mode = FTL_ROOT_SKINNED;
  return FTL_OK;
}

// ---------------------------------------------------------------------------
// diagnostics

extern "C" ftl_status ftl_diag_copy(ftl_ctx* ctx, const void* src, void* dst, size_t bytes, void* stream) {
  if (!ctx || !src || !dst || bytes % 16) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "bad copy arguments");
  hipError_t e;
  if ((e = hipSetDevice(ctx->device)) != hipSuccess) return ftl_hip_fail(e, "hipSetDevice");
  ftl_status st;
  if ((st = ftl_check_device_range(src, bytes, "src")) != FTL_OK ||
      (st = ftl_check_device_range(dst, bytes, "dst")) != FTL_OK)
    return st;
  if ((e = ftl_launch_copy(src, dst, bytes / 16, ctx->ncu, (hipStream_t)stream)) != hipSuccess)
    return ftl_hip_fail(e, "copy launch");
  return FTL_OK;
}

extern "C" ftl_status ftl_diag_alloc(ftl_ctx* ctx, size_t bytes, void** out) {
  if (!ctx || !out) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "NULL argument");
  hipError_t e;
  if ((e = hipSetDevice(ctx->device)) != hipSuccess) return ftl_hip_fail(e, "hipSetDevice");
  ftl_status st = ftl_dev_alloc(ctx, out, bytes, "diagnostic buffer");
  if (st != FTL_OK) return st;
  if ((e = hipMemset(*out, 0, bytes)) != hipSuccess) return ftl_hip_fail(e, "hipMemset");
  return FTL_OK;
}

extern "C" ftl_status ftl_diag_free(ftl_ctx* ctx, void* p, size_t bytes) {
  if (!ctx) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "ctx is NULL");
  (void)hipSetDevice(ctx->device);
  (void)hipDeviceSynchronize();
  ftl_dev_free(ctx, p, bytes);
  return FTL_OK;
}
