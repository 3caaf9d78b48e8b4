// Render-side strand expansion (include/ftl_render.h): the reference
// renderer's tessellation stream pass (src/fx/hair.cc:144-170, shaders
// src/shaders/hair/02_tess_stream/*) as one kernel.
//
// The GL pipeline runs vertex -> control -> evaluation -> geometry stages per
// (instance, patch) and captures line segments by transform feedback, two
// vertices each.  Here one thread owns one captured vertex position of a
// patch (scalp triangle, strand segment): end e of subsegment s.  It
// evaluates the three guide Hermite segments at u = (s + e) / n once, then
// blends them for every (instance, line) -- the only inputs that differ
// between those -- and writes one float4 per blend.  Consecutive threads take
// consecutive vertices of an isoline, so every store of a wave writes runs of
// 2 * nsub float4 (512 B at the reference's 16 subsegments): whole lines.
// At the settings where it measured faster (the reference's among them), a
// second form batches patches per wave instead: their guides are loaded once
// into LDS and each lane stores several vertices' blends
// (ftl_expand_batched_kernel).
#include <hip/hip_runtime.h>

#include <cstring>

#include "ftl_internal.h"
#include "ftl_render.h"

namespace {

using ftl_dev::f3;

struct ExpandArgs {
  const float4* __restrict__ pos;
  const float4* __restrict__ tan;
  const int4* __restrict__ tris;
  const float4* __restrict__ rnd;
  float4* __restrict__ out;
  int ntris, npoints, nseg, nsub, nlines, ninst;
  int nthreads;  // ntris * nseg * 2 * nsub
  float scale;
};

// Hermite basis at u (hermite_mix, inc_maths.glsl:210-228: vU * mHermit),
// one fixed evaluation order.  u = 0 gives (1, 0, 0, 0) and u = 1 gives
// (0, 1, 0, 0) exactly, so a segment's ends are its control points.
struct Basis {
  float h0, h1, h2, h3;
};
__device__ __forceinline__ Basis basis(float u) {
  const float u2 = u * u, u3 = u2 * u;
  const float a = 2.0f * u3, b = 3.0f * u2;
  Basis h;
  h.h0 = (a - b) + 1.0f;
  h.h1 = b - a;
  h.h2 = (u3 - 2.0f * u2) + u;
  h.h3 = u3 - u2;
  return h;
}

__device__ __forceinline__ float hmix(const Basis& h, float p0, float p1, float t0, float t1) {
  return __builtin_fmaf(h.h3, t1, __builtin_fmaf(h.h2, t0, __builtin_fmaf(h.h1, p1, h.h0 * p0)));
}

// smoothstep2(0, 1, j / N + u / N) (tes_stream_hair.glsl:64-65,
// inc_maths.glsl:263-266); maprange(0, 1, x) is clamp(x, 0, 1).
__device__ __forceinline__ float rel_pos(int j, float u, float n) {
  const float r = (float)j / n + u / n;
  const float t = fminf(fmaxf(r, 0.0f), 1.0f);
  return ((t * t) * t) * (10.0f + t * (-15.0f + 6.0f * t));
}

// Barycentric weights of (instance, line): the random entry the evaluation
// stage reads (tes_stream_hair.glsl:50-53) through sample_triangle2
// (inc_maths.glsl:107-115), whose statements run in order: the second reads
// the x the first just wrote, so y is left as it was.
__device__ __forceinline__ float3 bary(const float4* __restrict__ rnd, int inst, int line, int nlines) {
  const float v = (float)line / (float)nlines;
  const int idx = ((int)(v * 40.0f + (float)inst)) % 4096;
  const float4 r = rnd[idx];
  float sx = r.x, sy = r.y;
  if (sx + sy > 1.0f) {
    sx = fmaxf(sx, sy);
    sy = fminf(sx, sy);
    sx = 1.0f - sx;
  }
  return f3(sx, sy, 1.0f - (sx + sy));
}

__global__ __launch_bounds__(256) void ftl_expand_kernel(ExpandArgs A) {
  // The weights of every (instance, line), once per workgroup, in LDS: read
  // from there, they cost the store loop no wait on the vector-memory counter
  // (which also counts the stores, so a load inside the loop would make each
  // store wait for the one before it).
  extern __shared__ float4 wts[];
  // The thread's patch: its triangle is loaded before the weights so that the
  // two round trips overlap.  Threads past the end take the last vertex
  // (in bounds, never stored) so every thread reaches the barrier.
  const int tid = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  const bool live = tid < A.nthreads;
  const int t = live ? tid : A.nthreads - 1;
  // Vertex v of the patch's isoline: end e of subsegment s, at u = (s + e) / n.
  const int nv = 2 * A.nsub;
  const int v = t % nv;
  const int ts = t / nv;
  const int seg = ts % A.nseg;
  const int tri = ts / A.nseg;
  const int4 tv = A.tris[tri];
  const int ncombo = A.ninst * A.nlines;
  for (int c = (int)threadIdx.x; c < ncombo; c += (int)blockDim.x) {
    const float3 b = bary(A.rnd, c / A.nlines, c % A.nlines, A.nlines);
    wts[c] = make_float4(b.x, b.y, b.z, 0.0f);
  }
  const int root[3] = {tv.x, tv.y, tv.z};
  float4 P[3][2], Q[3][2];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const size_t base = (size_t)root[k] * A.npoints + seg;
    P[k][0] = A.pos[base];
    P[k][1] = A.pos[base + 1];
    Q[k][0] = A.tan[base];
    Q[k][1] = A.tan[base + 1];
  }
  __syncthreads();
  if (!live) return;

  const float u = (float)((v >> 1) + (v & 1)) / (float)A.nsub;
  const Basis h = basis(u);
  float3 g[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float4 p0 = P[k][0], p1 = P[k][1], q0 = Q[k][0], q1 = Q[k][1];
    // tangent * uScaleFactor (tcs_stream_hair.glsl:41)
    const float3 t0 = f3(q0.x * A.scale, q0.y * A.scale, q0.z * A.scale);
    const float3 t1 = f3(q1.x * A.scale, q1.y * A.scale, q1.z * A.scale);
    g[k] = f3(hmix(h, p0.x, p1.x, t0.x, t1.x), hmix(h, p0.y, p1.y, t0.y, t1.y), hmix(h, p0.z, p1.z, t0.z, t1.z));
  }
  const float w = rel_pos(seg, u, (float)A.npoints);

  // isoline (instance, triangle, segment, line) of the capture order
  long long iso = ((long long)tri * A.nseg + seg) * A.nlines;
  const long long inst_stride = (long long)A.ntris * A.nseg * A.nlines;
  for (int inst = 0; inst < A.ninst; ++inst, iso += inst_stride) {
    for (int line = 0; line < A.nlines; ++line) {
      const float4 c = wts[inst * A.nlines + line];
      // mat3x4(p0, p1, p2) * c
      const float4 o = make_float4(__builtin_fmaf(c.z, g[2].x, __builtin_fmaf(c.y, g[1].x, c.x * g[0].x)),
                                   __builtin_fmaf(c.z, g[2].y, __builtin_fmaf(c.y, g[1].y, c.x * g[0].y)),
                                   __builtin_fmaf(c.z, g[2].z, __builtin_fmaf(c.y, g[1].z, c.x * g[0].z)), w);
      A.out[(iso + line) * nv + v] = o;
    }
  }
}

// The same pass with more stores per load round trip: one wave per workgroup
// takes a batch of B patches.  It loads the batch's triangles, then their
// guide points and tangents (two round trips, as a thread of
// ftl_expand_kernel makes for its one patch, but at most six loads per lane
// instead of thirteen), into LDS, and then every lane evaluates B * NV / 64
// vertex positions of the batch and blends and stores each for every
// (instance, line), with no load in between: 12 stores per lane at the
// reference's tessellation (B = 4), against 6.  NV = 2 nsub positions per
// isoline is fixed per instantiation; the blend count (instances x lines) is
// a compile-time 6 at the reference's 3 x 2 x 16, so that loop unrolls, and
// a runtime count otherwise, read four weights at a time from LDS ahead of
// their stores (a wait on each weight before its store cost up to 8%).  Same operations per output vertex as
// ftl_expand_kernel.  (Measured at the reference's tessellation on the
// 100k-triangle scalp: B = 2, 4, 8, 16 ran 0.182, 0.160, 0.177, 0.175 ms.)
constexpr int kGuideF4 = 12;  // per patch: 3 guides x (point j, point j+1, tangent j, tangent j+1)

// Patches per batch for NV positions per isoline: the smallest multiple of 4
// whose B * NV positions fill whole rows of 64 lanes (4 at 16 subsegments,
// at most 32).
constexpr int expand_batch_of(int nv) {
  int b = 4;
  while ((b * nv) % 64 != 0) b += 4;
  return b;
}

__device__ __forceinline__ void expand_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Blends and stores batch b (patches b * B ...) from its guides in G0 (LDS)
// and the blend weights in wts (LDS).  GUARD: the pass's last batch, whose
// patches may end early.  NC > 0: the blend count, at compile time.
template <int NV, int NC, bool GUARD, int B>
__device__ __forceinline__ void expand_batch(const ExpandArgs& A, const float4* G0, const float4* wts, int b,
                                             int npatch, int lane) {
  const int nseg = A.nseg, nlines = A.nlines;
  const long long inst_stride = (long long)A.ntris * nseg * nlines;
  float4* __restrict__ out = A.out;
  const float fsub = (float)A.nsub, fpts = (float)A.npoints, scale = A.scale;
#pragma unroll
  for (int s = 0; s < B * NV / 64; ++s) {
    const int item = s * 64 + lane;
    const int i = item / NV, v = item % NV;
    const int patch = b * B + i;
    if (!GUARD || patch < npatch) {
      const float4* G = G0 + i * kGuideF4;
      const int tri = patch / nseg, seg = patch % nseg;
      const float u = (float)((v >> 1) + (v & 1)) / fsub;
      const Basis h = basis(u);
      float3 g[3];
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const float4 p0 = G[4 * q + 0], p1 = G[4 * q + 1], q0 = G[4 * q + 2], q1 = G[4 * q + 3];
        const float3 t0 = f3(q0.x * scale, q0.y * scale, q0.z * scale);
        const float3 t1 = f3(q1.x * scale, q1.y * scale, q1.z * scale);
        g[q] = f3(hmix(h, p0.x, p1.x, t0.x, t1.x), hmix(h, p0.y, p1.y, t0.y, t1.y), hmix(h, p0.z, p1.z, t0.z, t1.z));
      }
      const float w = rel_pos(seg, u, fpts);
      const long long iso = ((long long)tri * nseg + seg) * nlines;
      auto blend = [&](const float4 cw) {
        return make_float4(__builtin_fmaf(cw.z, g[2].x, __builtin_fmaf(cw.y, g[1].x, cw.x * g[0].x)),
                           __builtin_fmaf(cw.z, g[2].y, __builtin_fmaf(cw.y, g[1].y, cw.x * g[0].y)),
                           __builtin_fmaf(cw.z, g[2].z, __builtin_fmaf(cw.y, g[1].z, cw.x * g[0].z)), w);
      };
      if constexpr (NC > 0) {
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          const int inst = c / nlines, line = c % nlines;
          out[(iso + inst * inst_stride + line) * NV + v] = blend(wts[c]);
        }
      } else {
        // (instance, line) in the capture order, four blends at a time: their
        // weights are read from LDS together, ahead of the four stores
        const int nc = A.ninst * nlines;
        int inst = 0, line = 0;
        for (int c0 = 0; c0 < nc; c0 += 4) {
          float4 cw[4];
#pragma unroll
          for (int k = 0; k < 4; ++k) cw[k] = wts[min(c0 + k, nc - 1)];
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            if (c0 + k < nc) {
              out[(iso + inst * inst_stride + line) * NV + v] = blend(cw[k]);
              if (++line == nlines) {
                line = 0;
                ++inst;
              }
            }
          }
        }
      }
    }
  }
}

template <int NV, int NC, int B>
__global__ __launch_bounds__(64) void ftl_expand_batched_kernel(ExpandArgs A) {
  static_assert(NV % 2 == 0 && (B * NV) % 64 == 0 && B <= 64, "whole patches per wave-wide step");
  constexpr int kGuides = B * kGuideF4;
  extern __shared__ float4 sm[];
  const int nc = NC > 0 ? NC : A.ninst * A.nlines;
  float4* wts = sm;                     // [nc]
  float4* gd = sm + nc;                 // [B * kGuideF4]
  int4* td = (int4*)(gd + kGuides);     // [B]
  const int lane = (int)threadIdx.x;
  const int nseg = A.nseg, nlines = A.nlines;
  const int npatch = A.ntris * nseg, last = npatch - 1;
  const int b = (int)blockIdx.x;
  // The batch's triangles (patches clamped into the pass: in bounds, unused);
  // lanes past a run repeat another lane's load.
  const int ti = lane % B;
  const int4 tv0 = A.tris[min(b * B + ti, last) / nseg];
  for (int c = lane; c < nc; c += 64) {
    const float3 bw = bary(A.rnd, c / nlines, c % nlines, nlines);
    wts[c] = make_float4(bw.x, bw.y, bw.z, 0.0f);
  }
  td[ti] = tv0;
  expand_sync();
  // Guide float4 e of the batch: e = 12 i + 4 k + w (patch i, guide k; w: point
  // or tangent, j or j + 1); runs of 64, the last one repeated across the lanes.
  auto guide_index = [&](int r) { return kGuides - 64 * r >= 64 ? 64 * r + lane : 64 * r + lane % (kGuides - 64 * r); };
  auto guide_load = [&](int e) {
    const int i = e / kGuideF4, k = (e % kGuideF4) / 4, w = e % 4;
    const int4 tv = td[i];
    const int seg = min(b * B + i, last) % nseg;
    const int root = k == 0 ? tv.x : (k == 1 ? tv.y : tv.z);
    const size_t at = (size_t)root * A.npoints + seg + (w & 1);
    return (w >> 1) ? A.tan[at] : A.pos[at];
  };
  constexpr int kRuns = (kGuides + 63) / 64;
  float4 gv[kRuns];
#pragma unroll
  for (int r = 0; r < kRuns; ++r) gv[r] = guide_load(guide_index(r));
#pragma unroll
  for (int r = 0; r < kRuns; ++r) gd[guide_index(r)] = gv[r];
  expand_sync();
  if ((b + 1) * B <= npatch)
    expand_batch<NV, NC, false, B>(A, gd, wts, b, npatch, lane);
  else
    expand_batch<NV, NC, true, B>(A, gd, wts, b, npatch, lane);
}

// The batched pass at nsub = NSUB subsegments (nc blends).
template <int NSUB>
void launch_expand_batched(const ExpandArgs& A, int nc, hipStream_t stream) {
  constexpr int NV = 2 * NSUB, B = expand_batch_of(NV);
  static_assert(B <= 8 && B * NV / 64 <= 3, "the subsegment counts the batched pass takes");
  const int npatch = A.ntris * A.nseg;
  const int nbatch = (npatch + B - 1) / B;
  const size_t lds = (size_t)nc * sizeof(float4) + (size_t)B * kGuideF4 * sizeof(float4) + B * sizeof(int4);
  if constexpr (NSUB == 16) {
    if (nc == 6) {
      hipLaunchKernelGGL((ftl_expand_batched_kernel<NV, 6, B>), dim3(nbatch), dim3(64), lds, stream, A);
      return;
    }
  }
  hipLaunchKernelGGL((ftl_expand_batched_kernel<NV, 0, B>), dim3(nbatch), dim3(64), lds, stream, A);
}

// Launches the batched pass where it measured faster than ftl_expand_kernel
// on the 100k-triangle scalp (0 to -33% per pass): at most 10 blends and 8,
// 12 or 16 subsegments (batches of 4 or 8 patches), the reference's 3 x 2 x
// 16 among them.  Otherwise it launches nothing and returns false.  More
// blends per position already amortise the per-vertex kernel's loads (12 or
// 16 blends: 0 to +5% batched); fewer subsegments need batches of 16 or 32
// patches, which leave too few waves (3 x 2 x 3: +12%, 2 x 3 x 6: +8%,
// 3 x 3 x 7: +22%, 3 x 2 x 15: +13%), and 4 subsegments gain nothing
// (3 x 2 x 4: +2%).
bool launch_expand_batched_if_faster(const ExpandArgs& A, int nsub, int nc, hipStream_t st) {
  if (nc > 10) return false;
  switch (nsub) {
    case 8: launch_expand_batched<8>(A, nc, st); return true;
    case 12: launch_expand_batched<12>(A, nc, st); return true;
    case 16: launch_expand_batched<16>(A, nc, st); return true;
    default: return false;
  }
}

bool params_ok(const ftl_expand_params* p) {
  // GL_MAX_TESS_GEN_LEVEL is at least 64; the reference's levels are small.
  // Instances x lines weights live in LDS: at most 4096 (64 KB).
  return p && p->ninstances >= 1 && p->nlines >= 1 && p->nlines <= 64 && p->nsubsegments >= 1 &&
         p->nsubsegments <= 64 && (int64_t)p->ninstances * p->nlines <= 4096 &&
         p->tangent_scale == p->tangent_scale;
}

// Largest random index of the pass, with the device's float operations.
int max_random_index(const ftl_expand_params* p) {
  int mx = 0;
  for (int inst = 0; inst < p->ninstances; ++inst)
    for (int line = 0; line < p->nlines; ++line) {
      volatile float v = (float)line / (float)p->nlines;
      volatile float x = v * 40.0f;
      volatile float y = x + (float)inst;
      const int idx = ((int)y) % 4096;
      if (idx > mx) mx = idx;
    }
  return mx;
}

}  // namespace

extern "C" void ftl_expand_params_reference(ftl_expand_params* p) {
  if (!p) return;
  p->ninstances = 3;
  p->nlines = 2;
  p->nsubsegments = 16;
  p->tangent_scale = 1.0f;
}

extern "C" void ftl_expand_destroy(ftl_expand* e) {
  if (!e) return;
  ftl_ctx* ctx = e->ctx;
  (void)hipSetDevice(ctx->device);
  (void)hipDeviceSynchronize();
  ftl_dev_free(ctx, e->d_tris, (size_t)e->ntris * sizeof(int4));
  ftl_dev_free(ctx, e->d_tan, (size_t)e->nstrands * e->npoints * sizeof(float4));
  ftl_dev_free(ctx, e->d_rand, (size_t)e->nrandom * sizeof(float4));
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->live_objects--;
  }
  delete e;
}

extern "C" ftl_status ftl_expand_create(ftl_ctx* ctx, int nstrands, int npoints, const int* tris, int ntris,
                                        const float* tangents, const float* random, int nrandom, ftl_expand** out) {
  if (!ctx || !tris || !tangents || !random || !out) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "NULL argument");
  *out = nullptr;
  if (nstrands <= 0 || npoints < 2 || ntris <= 0 || nrandom <= 0)
    return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "ftl_expand_create: bad sizes (%d strands, %d points, %d tris, %d random)",
                    nstrands, npoints, ntris, nrandom);
  if ((int64_t)nstrands * npoints > (int64_t)INT32_MAX)
    return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "ftl_expand_create: too many control points");
  std::vector<int4> t4((size_t)ntris);
  for (int i = 0; i < ntris; ++i) {
    const int a = tris[3 * (size_t)i], b = tris[3 * (size_t)i + 1], c = tris[3 * (size_t)i + 2];
    if (a < 0 || a >= nstrands || b < 0 || b >= nstrands || c < 0 || c >= nstrands)
      return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "ftl_expand_create: triangle %d names a strand outside [0, %d)", i,
                      nstrands);
    t4[(size_t)i] = make_int4(a, b, c, 0);
  }
  hipError_t he;
  if ((he = hipSetDevice(ctx->device)) != hipSuccess) return ftl_hip_fail(he, "hipSetDevice");
  ftl_expand* e = new ftl_expand;
  e->ctx = ctx;
  e->nstrands = nstrands;
  e->npoints = npoints;
  e->ntris = ntris;
  e->nrandom = nrandom;
  const size_t tb = (size_t)ntris * sizeof(int4), nb = (size_t)nstrands * npoints * sizeof(float4),
               rb = (size_t)nrandom * sizeof(float4);
  ftl_status st;
  if ((st = ftl_dev_alloc(ctx, (void**)&e->d_tris, tb, "expansion triangles")) != FTL_OK ||
      (st = ftl_dev_alloc(ctx, (void**)&e->d_tan, nb, "expansion tangents")) != FTL_OK ||
      (st = ftl_dev_alloc(ctx, (void**)&e->d_rand, rb, "expansion random table")) != FTL_OK) {
    ftl_dev_free(ctx, e->d_tris, tb);
    ftl_dev_free(ctx, e->d_tan, nb);
    ftl_dev_free(ctx, e->d_rand, rb);
    delete e;
    return st;
  }
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->live_objects++;
  }
  if (hipMemcpy(e->d_tris, t4.data(), tb, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(e->d_tan, tangents, nb, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(e->d_rand, random, rb, hipMemcpyHostToDevice) != hipSuccess) {
    ftl_expand_destroy(e);
    return ftl_fail(FTL_ERR_DEVICE, "ftl_expand_create: upload failed");
  }
  *out = e;
  return FTL_OK;
}

extern "C" ftl_status ftl_expand_set_tangents(ftl_expand* e, const float* tangents) {
  if (!e || !tangents) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "NULL argument");
  hipError_t he;
  if ((he = hipSetDevice(e->ctx->device)) != hipSuccess) return ftl_hip_fail(he, "hipSetDevice");
  if ((he = hipMemcpy(e->d_tan, tangents, (size_t)e->nstrands * e->npoints * sizeof(float4), hipMemcpyHostToDevice)) !=
      hipSuccess)
    return ftl_hip_fail(he, "ftl_expand_set_tangents");
  return FTL_OK;
}

extern "C" ftl_status ftl_expand_get_tangents(ftl_expand* e, float* tangents) {
  if (!e || !tangents) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "NULL argument");
  hipError_t he;
  if ((he = hipSetDevice(e->ctx->device)) != hipSuccess) return ftl_hip_fail(he, "hipSetDevice");
  if ((he = hipDeviceSynchronize()) != hipSuccess) return ftl_hip_fail(he, "ftl_expand_get_tangents");
  if ((he = hipMemcpy(tangents, e->d_tan, (size_t)e->nstrands * e->npoints * sizeof(float4), hipMemcpyDeviceToHost)) !=
      hipSuccess)
    return ftl_hip_fail(he, "ftl_expand_get_tangents");
  return FTL_OK;
}

extern "C" int64_t ftl_expand_vertex_count(const ftl_expand* e, const ftl_expand_params* p) {
  if (!e || !params_ok(p)) return -1;
  return 2 * (int64_t)p->ninstances * e->ntris * (e->npoints - 1) * p->nlines * p->nsubsegments;
}

extern "C" ftl_status ftl_expand_run(ftl_expand* e, const ftl_expand_params* p, const float* positions, float* out,
                                     int64_t out_capacity, void* stream) {
  if (!e || !positions || !out) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "NULL argument");
  if (!params_ok(p))
    return ftl_fail(FTL_ERR_INVALID_ARGUMENT,
                    "ftl_expand_run: instances >= 1, lines and subsegments in [1, 64], finite tangent scale");
  const int64_t need = ftl_expand_vertex_count(e, p);
  if (out_capacity < need)
    return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "ftl_expand_run: output holds %lld vertices, the pass writes %lld",
                    (long long)out_capacity, (long long)need);
  const int ridx = max_random_index(p);
  if (ridx >= e->nrandom)
    return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "ftl_expand_run: random index %d, table holds %d entries", ridx,
                    e->nrandom);
  const int64_t nthreads = (int64_t)e->ntris * (e->npoints - 1) * 2 * p->nsubsegments;
  if (nthreads > (int64_t)INT32_MAX - 256)
    return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "ftl_expand_run: %lld patch vertices per pass exceed 2^31",
                    (long long)nthreads);
  hipError_t he;
  if ((he = hipSetDevice(e->ctx->device)) != hipSuccess) return ftl_hip_fail(he, "hipSetDevice");
  ftl_status st;
  if ((st = ftl_check_device_range(positions, (size_t)e->nstrands * e->npoints * sizeof(float4), "positions")) !=
          FTL_OK ||
      (st = ftl_check_device_range(out, (size_t)need * sizeof(float4), "out")) != FTL_OK)
    return st;
  ExpandArgs A;
  A.pos = (const float4*)positions;
  A.tan = e->d_tan;
  A.tris = e->d_tris;
  A.rnd = e->d_rand;
  A.out = (float4*)out;
  A.ntris = e->ntris;
  A.npoints = e->npoints;
  A.nseg = e->npoints - 1;
  A.nsub = p->nsubsegments;
  A.nlines = p->nlines;
  A.ninst = p->ninstances;
  A.nthreads = (int)nthreads;
  A.scale = p->tangent_scale;
  const size_t lds = (size_t)p->ninstances * p->nlines * sizeof(float4);
  const int nc = p->ninstances * p->nlines;
  if (!launch_expand_batched_if_faster(A, p->nsubsegments, nc, (hipStream_t)stream)) {
    const int blocks = (int)((nthreads + 255) / 256);
    hipLaunchKernelGGL(ftl_expand_kernel, dim3(blocks), dim3(256), lds, (hipStream_t)stream, A);
  }
  if ((he = hipGetLastError()) != hipSuccess) return ftl_hip_fail(he, "ftl_expand_kernel launch");
  return FTL_OK;
}
