// Groom initialisation on the device: Hair::init_simulation
// (src/fx/hair.cc:236-328) as two kernels, one thread per control point.
//
//   positions / velocities (hair.cc:258-288): strand j rooted at v_j with
//     normal n_j; r_j = (float)(1 + 0.1 (1 - 2 U_j)) with U_j the j-th draw of
//     the SplitMix64 stream of the seed (the reference draws rand()); offset_i
//     = ((float)i * (maxlength / N)) * r_j; pos_i = (v + offset_i n,
//     offset_i - offset_{i-1}), vel = 0.  SplitMix64 is counter based, so
//     thread (j, i) computes U_j directly: the result is bit-identical to the
//     host generator ftl_scn_reference_groom.
//   tangents (hair.cc:290-328): the "curly" style, with glm::simplex (the
//     2-D simplex noise of GLM's gtc/noise, a port of the webgl-noise
//     algorithm of McEwan, Sheets, Gustavson and Richardson, "Efficient
//     computational noise in GLSL", JGT 2012) restated below.
#include <hip/hip_runtime.h>

#include "ftl_internal.h"
#include "ftl_render.h"

namespace {

__device__ __forceinline__ uint64_t splitmix_at(uint64_t seed, uint64_t k) {
  // the (k+1)-th output of the stream s += golden; mix(s)
  uint64_t z = seed + (k + 1) * 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

struct InitArgs {
  const float* roots;    // float3 per strand
  const float* normals;  // float3 per strand
  float4* pos;
  float4* vel;
  int nroots, npoints;
  float maxlength;
  uint64_t seed;
};

__device__ __forceinline__ float strand_offset(int i, float scale_offset, float rv) {
  return (float)i * scale_offset * rv;
}

__global__ __launch_bounds__(256) void ftl_init_groom_kernel(InitArgs A) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)A.nroots * A.npoints) return;
  const int j = (int)(t / A.npoints), i = (int)(t % A.npoints);
  const double u = (double)(splitmix_at(A.seed, (uint64_t)j) >> 11) * 0x1p-53;
  const float rv = (float)(1.0 + 0.1 * (1.0 - 2.0 * u));
  const float scale_offset = A.maxlength / (float)A.npoints;
  const float off = strand_offset(i, scale_offset, rv);
  const float last = i > 0 ? strand_offset(i - 1, scale_offset, rv) : 0.0f;
  const float* v = A.roots + 3 * (size_t)j;
  const float* n = A.normals + 3 * (size_t)j;
  A.pos[t] = make_float4(v[0] + off * n[0], v[1] + off * n[1], v[2] + off * n[2], off - last);
  A.vel[t] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// --- glm::simplex(vec2), restated (see the file header) --------------------
__device__ __forceinline__ float mod289(float x) { return x - floorf(x * (1.0f / 289.0f)) * 289.0f; }
__device__ __forceinline__ float permute(float x) { return mod289(((x * 34.0f) + 1.0f) * x); }
__device__ __forceinline__ float fract(float x) { return x - floorf(x); }

__device__ float simplex2(float vx, float vy) {
  const float Cx = 0.211324865405187f;   // (3 - sqrt(3)) / 6
  const float Cy = 0.366025403784439f;   // (sqrt(3) - 1) / 2
  const float Cz = -0.577350269189626f;  // -1 + 2 Cx
  const float Cw = 0.024390243902439f;   // 1 / 41
  // first corner
  const float sk = vx * Cy + vy * Cy;
  float ix = floorf(vx + sk), iy = floorf(vy + sk);
  const float us = ix * Cx + iy * Cx;
  const float x0x = vx - ix + us, x0y = vy - iy + us;
  // other corners
  const float i1x = (x0x > x0y) ? 1.0f : 0.0f, i1y = (x0x > x0y) ? 0.0f : 1.0f;
  const float x12x = (x0x + Cx) - i1x, x12y = (x0y + Cx) - i1y;
  const float x12z = x0x + Cz, x12w = x0y + Cz;
  // permutations
  ix = ix - 289.0f * floorf(ix / 289.0f);
  iy = iy - 289.0f * floorf(iy / 289.0f);
  const float p0 = permute(permute(iy + 0.0f) + ix + 0.0f);
  const float p1 = permute(permute(iy + i1y) + ix + i1x);
  const float p2 = permute(permute(iy + 1.0f) + ix + 1.0f);
  float m0 = fmaxf(0.5f - (x0x * x0x + x0y * x0y), 0.0f);
  float m1 = fmaxf(0.5f - (x12x * x12x + x12y * x12y), 0.0f);
  float m2 = fmaxf(0.5f - (x12z * x12z + x12w * x12w), 0.0f);
  m0 = m0 * m0;
  m1 = m1 * m1;
  m2 = m2 * m2;
  m0 = m0 * m0;
  m1 = m1 * m1;
  m2 = m2 * m2;
  // gradients: 41 points on a line mapped onto a diamond
  const float gx0 = 2.0f * fract(p0 * Cw) - 1.0f, gx1 = 2.0f * fract(p1 * Cw) - 1.0f,
              gx2 = 2.0f * fract(p2 * Cw) - 1.0f;
  const float h0 = fabsf(gx0) - 0.5f, h1 = fabsf(gx1) - 0.5f, h2 = fabsf(gx2) - 0.5f;
  const float a0 = gx0 - floorf(gx0 + 0.5f), a1 = gx1 - floorf(gx1 + 0.5f), a2 = gx2 - floorf(gx2 + 0.5f);
  m0 *= 1.79284291400159f - 0.85373472095314f * (a0 * a0 + h0 * h0);
  m1 *= 1.79284291400159f - 0.85373472095314f * (a1 * a1 + h1 * h1);
  m2 *= 1.79284291400159f - 0.85373472095314f * (a2 * a2 + h2 * h2);
  const float g0 = a0 * x0x + h0 * x0y;
  const float g1 = a1 * x12x + h1 * x12y;
  const float g2 = a2 * x12z + h2 * x12w;
  return 130.0f * (m0 * g0 + m1 * g1 + m2 * g2);
}

struct TangentArgs {
  const float* normals;  // float3 per strand
  float4* tan;
  int nstrands, npoints, curly;
  float maxlength;
};

// Tangents of hair.cc:290-328 (loop-carried `curly` is reassigned before each
// use, so every point is independent).
__global__ __launch_bounds__(256) void ftl_init_tangents_kernel(TangentArgs A) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)A.nstrands * A.npoints) return;
  const int j = (int)(t / A.npoints), il = (int)(t % A.npoints);
  const float kPi = 3.14159265358979323846f;
  const float inv_nroots = 1.0f / (float)A.nstrands;
  const float scale_max_length = 0.125f * sqrtf(A.maxlength);
  const float dj = (float)(j + 1) * inv_nroots;
  const float* nr = A.normals + 3 * (size_t)j;
  const int last = A.npoints - 1;
  float4 out;
  if (il == 0) {
    out = make_float4(0.15f * nr[0], 0.15f * nr[1], 0.15f * nr[2], 0.0f);
  } else if (il == last) {
    float cx = 0.0f, cy = 0.0f, cz = 0.0f;
    if (A.curly) {
      const float n = 1.25f * simplex2(ftl_dev::sin_c(3.0f * dj), ftl_dev::cos_c(5.0f));
      cx = ftl_dev::cos_c(n * 4.0f * kPi);
      cy = -0.71f * n;
      cz = ftl_dev::sin_c(n * 2.7f * kPi);
    }
    out = make_float4(0.2f * (-nr[0] + cx), 0.2f * (-nr[1] + cy), 0.2f * (-nr[2] + cz), 0.0f);
  } else {
    const float dist_ab = (float)last;
    const float inv_dist = 1.0f / dist_ab;
    float cx = 0.0f, cy = 0.0f, cz = 0.0f;
    if (A.curly) {
      const float di = 10.0f * (float)(last - il) / (dist_ab - 1.0f);
      const float n = di * simplex2(ftl_dev::sin_c(43.0f * dj), ftl_dev::cos_c(5.0f * di));
      cx = -5.8f * (10.7f * ftl_dev::cos_c(n * kPi));
      cy = -5.8f * (-2.3f * n);
      cz = -5.8f * (20.5f * ftl_dev::sin_c(n * kPi));
    }
    const float s = 0.1f * (float)il * inv_dist * scale_max_length;
    out = make_float4(s * cx, s * cy, s * cz, s * 0.0f);
  }
  A.tan[t] = out;
}

}  // namespace

extern "C" ftl_status ftl_groom_create_reference(ftl_ctx* ctx, int nroots, const float* roots, const float* normals,
                                                 int npoints, float maxlength, uint64_t seed, ftl_groom** out) {
  if (!ctx || !roots || !normals || !out) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "NULL argument");
  *out = nullptr;
  if (nroots <= 0 || npoints < 2 || npoints > FTL_MAX_POINTS || !(maxlength > 0.0f))
    return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "ftl_groom_create_reference: bad sizes or length");
  // The groom is created zeroed (no host copy of the state), then filled on
  // the device.
  ftl_groom_desc d{nroots, npoints, nullptr, nullptr, nullptr};
  ftl_groom* g = nullptr;
  ftl_status st = ftl_groom_create_uninitialised(ctx, &d, &g);
  if (st != FTL_OK) return st;
  const size_t rb = (size_t)nroots * 3 * sizeof(float);
  void *d_roots = nullptr, *d_normals = nullptr;
  if ((st = ftl_dev_alloc(ctx, &d_roots, rb, "init roots")) != FTL_OK ||
      (st = ftl_dev_alloc(ctx, &d_normals, rb, "init normals")) != FTL_OK) {
    ftl_dev_free(ctx, d_roots, rb);
    ftl_groom_destroy(g);
    return st;
  }
  hipError_t he = hipMemcpy(d_roots, roots, rb, hipMemcpyHostToDevice);
  if (he == hipSuccess) he = hipMemcpy(d_normals, normals, rb, hipMemcpyHostToDevice);
  if (he == hipSuccess) {
    InitArgs A{(const float*)d_roots, (const float*)d_normals, (float4*)g->d_pos, (float4*)g->d_vel,
               nroots, npoints, maxlength, seed};
    const long long n = (long long)nroots * npoints;
    hipLaunchKernelGGL(ftl_init_groom_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, nullptr, A);
    he = hipGetLastError();
    if (he == hipSuccess) he = hipDeviceSynchronize();
  }
  ftl_dev_free(ctx, d_roots, rb);
  ftl_dev_free(ctx, d_normals, rb);
  if (he != hipSuccess) {
    ftl_groom_destroy(g);
    return ftl_hip_fail(he, "ftl_groom_create_reference");
  }
  *out = g;
  return FTL_OK;
}

extern "C" ftl_status ftl_expand_reference_tangents(ftl_expand* e, const float* normals, float maxlength, int curly) {
  if (!e || !normals) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "NULL argument");
  if (!(maxlength >= 0.0f)) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "ftl_expand_reference_tangents: bad length");
  ftl_ctx* ctx = e->ctx;
  const int nstrands = e->nstrands, npoints = e->npoints;
  float4* d_tan = e->d_tan;
  hipError_t he;
  if ((he = hipSetDevice(ctx->device)) != hipSuccess) return ftl_hip_fail(he, "hipSetDevice");
  const size_t nb = (size_t)nstrands * 3 * sizeof(float);
  void* d_normals = nullptr;
  ftl_status st;
  if ((st = ftl_dev_alloc(ctx, &d_normals, nb, "tangent normals")) != FTL_OK) return st;
  he = hipMemcpy(d_normals, normals, nb, hipMemcpyHostToDevice);
  if (he == hipSuccess) {
    TangentArgs A{(const float*)d_normals, d_tan, nstrands, npoints, curly ? 1 : 0, maxlength};
    const long long n = (long long)nstrands * npoints;
    hipLaunchKernelGGL(ftl_init_tangents_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, nullptr, A);
    he = hipGetLastError();
    if (he == hipSuccess) he = hipDeviceSynchronize();
  }
  ftl_dev_free(ctx, d_normals, nb);
  if (he != hipSuccess) return ftl_hip_fail(he, "ftl_expand_reference_tangents");
  return FTL_OK;
}
