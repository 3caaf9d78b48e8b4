// Small kernels and host dispatch: dual-quaternion skinning, the copy
// reference, and the class table -> per-class launcher map.
#include "ftl_classes.h"
#include "ftl_device.h"
#include "ftl_launch.h"

using namespace ftl_dev;

// ---------------------------------------------------------------------------
// Dual-quaternion blended skinning of mesh vertices
// (src/shaders/shared/inc_skinning.glsl:22-31 apply_skinning, :54-82 skinning_DQBS).

__device__ __forceinline__ float dot4(float4 a, float4 b) {
  return __builtin_fmaf(a.w, b.w, __builtin_fmaf(a.z, b.z, __builtin_fmaf(a.y, b.y, a.x * b.x)));
}
__device__ __forceinline__ float sign_c(float x) { return x > 0.f ? 1.f : (x < 0.f ? -1.f : 0.f); }
__device__ __forceinline__ float3 cross_c(float3 a, float3 b) {
  return f3(__builtin_fmaf(a.y, b.z, -(a.z * b.y)), __builtin_fmaf(a.z, b.x, -(a.x * b.z)),
            __builtin_fmaf(a.x, b.y, -(a.y * b.x)));
}

__global__ __launch_bounds__(256) void ftl_skin_kernel(int nverts, const float4* __restrict__ dq,
                                                       const int4* __restrict__ joints,
                                                       const float4* __restrict__ weights,
                                                       const float4* __restrict__ rest,
                                                       float4* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nverts) return;
  const float4 r = rest[i];
  float3 v = f3(r.x, r.y, r.z);
  const float4 wt = weights[i];
  if (wt.x <= 1e-6f) {  // Epsilon() (inc_constants.glsl:6)
    out[i] = r;
    return;
  }
  float w0 = wt.x, w1 = wt.y, w2 = wt.z;
  const float w3 = 1.0f - (w0 + w1 + w2);
  const int4 jt = joints[i];
  const float4 a0 = dq[2 * jt.x], a1 = dq[2 * jt.y], a2 = dq[2 * jt.z], a3 = dq[2 * jt.w];
  const float4 b0 = dq[2 * jt.x + 1], b1 = dq[2 * jt.y + 1], b2 = dq[2 * jt.z + 1], b3 = dq[2 * jt.w + 1];
  // Antipodality: weights.xyz *= sign(Ma[3] * mat3x4(Ma)).
  w0 = w0 * sign_c(dot4(a3, a0));
  w1 = w1 * sign_c(dot4(a3, a1));
  w2 = w2 * sign_c(dot4(a3, a2));
#define FTL_BL(a0, a1, a2, a3, c) \
  __builtin_fmaf(a3.c, w3, __builtin_fmaf(a2.c, w2, __builtin_fmaf(a1.c, w1, a0.c * w0)))
  float4 A = make_float4(FTL_BL(a0, a1, a2, a3, x), FTL_BL(a0, a1, a2, a3, y), FTL_BL(a0, a1, a2, a3, z),
                         FTL_BL(a0, a1, a2, a3, w));
  float4 B = make_float4(FTL_BL(b0, b1, b2, b3, x), FTL_BL(b0, b1, b2, b3, y), FTL_BL(b0, b1, b2, b3, z),
                         FTL_BL(b0, b1, b2, b3, w));
#undef FTL_BL
  const float inv = rsqrt_c(dot4(A, A));
  A = make_float4(A.x * inv, A.y * inv, A.z * inv, A.w * inv);
  B = make_float4(B.x * inv, B.y * inv, B.z * inv, B.w * inv);
  const float3 a = f3(A.x, A.y, A.z), b = f3(B.x, B.y, B.z);
  const float3 cv = cross_c(a, v);
  const float3 t = f3(__builtin_fmaf(A.w, v.x, cv.x), __builtin_fmaf(A.w, v.y, cv.y),
                      __builtin_fmaf(A.w, v.z, cv.z));
  const float3 rr = cross_c(a, t);
  v = f3(__builtin_fmaf(2.f, rr.x, v.x), __builtin_fmaf(2.f, rr.y, v.y), __builtin_fmaf(2.f, rr.z, v.z));
  const float3 ab = cross_c(a, b);
  const float3 tr = f3(__builtin_fmaf(A.w, b.x, __builtin_fmaf(-B.w, a.x, ab.x)),
                       __builtin_fmaf(A.w, b.y, __builtin_fmaf(-B.w, a.y, ab.y)),
                       __builtin_fmaf(A.w, b.z, __builtin_fmaf(-B.w, a.z, ab.z)));
  v = f3(__builtin_fmaf(2.f, tr.x, v.x), __builtin_fmaf(2.f, tr.y, v.y), __builtin_fmaf(2.f, tr.z, v.z));
  out[i] = make_float4(v.x, v.y, v.z, r.w);
}

// Streaming copy of float4s: the bandwidth reference for the step's traffic.
__global__ __launch_bounds__(256) void ftl_copy_kernel(const float4* __restrict__ src, float4* __restrict__ dst,
                                                       size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

// Upload of a small host array in pinned, coherent host memory by a kernel on
// the compute queue (a graph or a serving batch then needs no copy-engine
// node ahead of its kernels).
__global__ __launch_bounds__(256) void ftl_upload_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                         size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

// ---------------------------------------------------------------------------
// Host dispatch.

#define FTL_DECL(id, M, G, GU, W, CH)                                                                    \
  hipError_t ftl_launch_plain_##id(const FtlStep& A, hipStream_t stream);                           \
  hipError_t ftl_launch_tiles_##id(const FtlStep*, const FtlTile*, int, int, hipStream_t);          \
  int ftl_tiles_occ_##id();
FTL_CLASSES(FTL_DECL)
#undef FTL_DECL

int ftl_tiles_blocks_per_cu(int cls) {
  switch (cls) {
#define FTL_OC(id, M, G, GU, W, CH) \
  case id:                   \
    return ftl_tiles_occ_##id();
    FTL_CLASSES(FTL_OC)
#undef FTL_OC
  }
  return 0;
}

int ftl_num_classes() { return FTL_NCLASSES; }
int ftl_class_m(int cls) { return (cls >= 0 && cls < FTL_NCLASSES) ? kFtlClasses[cls].m : 0; }
int ftl_class_g(int cls) { return (cls >= 0 && cls < FTL_NCLASSES) ? kFtlClasses[cls].g : 0; }
int ftl_class_guarded(int cls) { return (cls >= 0 && cls < FTL_NCLASSES) ? kFtlClasses[cls].guarded : 0; }

hipError_t ftl_launch_plain(int cls, const FtlStep& A, hipStream_t stream) {
  switch (cls) {
#define FTL_LP(id, M, G, GU, W, CH) \
  case id:                   \
    return ftl_launch_plain_##id(A, stream);
    FTL_CLASSES(FTL_LP)
#undef FTL_LP
  }
  return hipErrorInvalidValue;
}

hipError_t ftl_launch_tiles(int cls, const FtlStep* grooms, const FtlTile* tiles, int ntiles, int grid,
                            hipStream_t stream) {
  switch (cls) {
#define FTL_LB(id, M, G, GU, W, CH) \
  case id:                   \
    return ftl_launch_tiles_##id(grooms, tiles, ntiles, grid, stream);
    FTL_CLASSES(FTL_LB)
#undef FTL_LB
  }
  return hipErrorInvalidValue;
}

hipError_t ftl_launch_skin(int nverts, const float4* dq, const int4* joints, const float4* weights,
                           const float4* rest, float4* out, hipStream_t stream) {
  if (nverts <= 0) return hipSuccess;
  hipLaunchKernelGGL(ftl_skin_kernel, dim3((nverts + 255) / 256), dim3(256), 0, stream, nverts, dq, joints,
                     weights, rest, out);
  return hipGetLastError();
}

static_assert(sizeof(FtlStep) % 16 == 0 && sizeof(FtlTile) % 16 == 0, "descriptors upload in 16-byte units");

hipError_t ftl_launch_upload(const void* host_src, void* dst, size_t bytes, hipStream_t stream) {
  if (bytes % 16 != 0) return hipErrorInvalidValue;
  const size_t n = bytes / 16;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(ftl_upload_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream,
                     (const uint4*)host_src, (uint4*)dst, n);
  return hipGetLastError();
}

hipError_t ftl_launch_copy(const void* src, void* dst, size_t n16, int ncu, hipStream_t stream) {
  if (n16 == 0) return hipSuccess;
  size_t blocks = (n16 + 255) / 256;
  const size_t cap = (size_t)ncu * 16;
  if (blocks > cap) blocks = cap;
  hipLaunchKernelGGL(ftl_copy_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, (const float4*)src, (float4*)dst,
                     n16);
  return hipGetLastError();
}
