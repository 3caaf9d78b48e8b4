// Small kernels and host dispatch: mesh skinning and root frames, the copy
// reference, and the class table -> per-class launcher map.
#include "ftl_classes.h"
#include "ftl_device.h"
#include "ftl_launch.h"

using namespace ftl_dev;

// ---------------------------------------------------------------------------
// Skinning of mesh vertices and their normals (src/shaders/shared/
// inc_skinning.glsl): apply_skinning's weight rule (:22-31), then
// dual-quaternion blending (skinning_DQBS, :54-82) or linear blending of 3x4
// joint matrices (skinning_LBS, :85-129).  One thread per vertex.

__device__ __forceinline__ float dot4(float4 a, float4 b) {
  return __builtin_fmaf(a.w, b.w, __builtin_fmaf(a.z, b.z, __builtin_fmaf(a.y, b.y, a.x * b.x)));
}
__device__ __forceinline__ float sign_c(float x) { return x > 0.f ? 1.f : (x < 0.f ? -1.f : 0.f); }
__device__ __forceinline__ float3 cross_c(float3 a, float3 b) {
  return f3(__builtin_fmaf(a.y, b.z, -(a.z * b.y)), __builtin_fmaf(a.z, b.x, -(a.x * b.z)),
            __builtin_fmaf(a.x, b.y, -(a.y * b.x)));
}
// x += 2 cross(a, cross(a, x) + aw x): the rotation of a unit quaternion.
__device__ __forceinline__ float3 dq_rotate(float3 a, float aw, float3 x) {
  const float3 cv = cross_c(a, x);
  const float3 t = f3(__builtin_fmaf(aw, x.x, cv.x), __builtin_fmaf(aw, x.y, cv.y), __builtin_fmaf(aw, x.z, cv.z));
  const float3 r = cross_c(a, t);
  return f3(__builtin_fmaf(2.f, r.x, x.x), __builtin_fmaf(2.f, r.y, x.y), __builtin_fmaf(2.f, r.z, x.z));
}

enum { kSkinDq = 0, kSkinLbs = 1 };
constexpr int kNoJoint = 0xff;  // inc_skinning.glsl:18

// Row r of joint j's 3x4 skinning matrix; kNoJoint is the identity (:89-93).
__device__ __forceinline__ float4 lbs_row(const float4* __restrict__ m, int j, int r) {
  if (j == kNoJoint) return make_float4(r == 0 ? 1.f : 0.f, r == 1 ? 1.f : 0.f, r == 2 ? 1.f : 0.f, 0.f);
  return m[3 * j + r];
}

// bones: MODE kSkinDq, two float4 per bone (real, dual); kSkinLbs, three
// float4 per bone (the rows of its skinning matrix, the transposed 3x4 of
// skeleton_controller.cc:251-256).  nrest / nout: normals, or null.
template <int MODE>
__global__ __launch_bounds__(256) void ftl_skin_kernel(int nverts, const float4* __restrict__ bones,
                                                       const int4* __restrict__ joints,
                                                       const float4* __restrict__ weights,
                                                       const float4* __restrict__ rest,
                                                       const float4* __restrict__ nrest, float4* __restrict__ out,
                                                       float4* __restrict__ nout) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nverts) return;
  const bool normals = nout != nullptr;
  const float4 r = rest[i];
  const float4 nr = normals ? nrest[i] : make_float4(0.f, 0.f, 0.f, 0.f);
  const float4 wt = weights[i];
  if (wt.x <= 1e-6f) {  // Epsilon() (inc_constants.glsl:6)
    out[i] = r;
    if (normals) nout[i] = nr;
    return;
  }
  float w0 = wt.x, w1 = wt.y, w2 = wt.z;
  const float w3 = 1.0f - (w0 + w1 + w2);
  const int4 jt = joints[i];
  if constexpr (MODE == kSkinDq) {
    const float4 a0 = bones[2 * jt.x], a1 = bones[2 * jt.y], a2 = bones[2 * jt.z], a3 = bones[2 * jt.w];
    const float4 b0 = bones[2 * jt.x + 1], b1 = bones[2 * jt.y + 1], b2 = bones[2 * jt.z + 1],
                 b3 = bones[2 * jt.w + 1];
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
    float3 v = dq_rotate(a, A.w, f3(r.x, r.y, r.z));
    const float3 ab = cross_c(a, b);
    const float3 tr = f3(__builtin_fmaf(A.w, b.x, __builtin_fmaf(-B.w, a.x, ab.x)),
                         __builtin_fmaf(A.w, b.y, __builtin_fmaf(-B.w, a.y, ab.y)),
                         __builtin_fmaf(A.w, b.z, __builtin_fmaf(-B.w, a.z, ab.z)));
    v = f3(__builtin_fmaf(2.f, tr.x, v.x), __builtin_fmaf(2.f, tr.y, v.y), __builtin_fmaf(2.f, tr.z, v.z));
    out[i] = make_float4(v.x, v.y, v.z, r.w);
    if (normals) {
      const float3 n = dq_rotate(a, A.w, f3(nr.x, nr.y, nr.z));
      nout[i] = make_float4(n.x, n.y, n.z, nr.w);
    }
  } else {
    const int j[4] = {jt.x, jt.y, jt.z, jt.w};
    float4 R[4][3];
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
      for (int c = 0; c < 3; ++c) R[k][c] = lbs_row(bones, j[k], c);
    // M[k] = (x, hw) * matrices[k]; result = M * weights (:108-118)
    auto blend = [&](float4 x, float hw) {
      const float4 xv = make_float4(x.x, x.y, x.z, hw);
      float o[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        float t[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) t[k] = dot4(xv, R[k][c]);
        o[c] = __builtin_fmaf(w3, t[3], __builtin_fmaf(w2, t[2], __builtin_fmaf(w1, t[1], w0 * t[0])));
      }
      return make_float4(o[0], o[1], o[2], x.w);
    };
    out[i] = blend(r, 1.0f);
    if (normals) nout[i] = blend(nr, 0.0f);
  }
}

// Root frame of each strand of a skinned groom: the barycentric root on its
// bound triangle (the step's root) and the unit normal
// normalize(w0 nA + w1 nB + w2 nC) there.
__global__ __launch_bounds__(256) void ftl_root_frames_kernel(int nstrands, const int4* __restrict__ bind,
                                                              const int4* __restrict__ tris,
                                                              const float4* __restrict__ verts,
                                                              const float4* __restrict__ normals,
                                                              float4* __restrict__ roots, float4* __restrict__ nout) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nstrands) return;
  const int4 b = bind[s];
  const int4 t = tris[b.x];
  const float w1 = __int_as_float(b.y), w2 = __int_as_float(b.z);
  const float w0 = (1.0f - w1) - w2;
  const float4 a = verts[t.x], bb = verts[t.y], c = verts[t.z];
  roots[s] = make_float4(__builtin_fmaf(w2, c.x, __builtin_fmaf(w1, bb.x, w0 * a.x)),
                         __builtin_fmaf(w2, c.y, __builtin_fmaf(w1, bb.y, w0 * a.y)),
                         __builtin_fmaf(w2, c.z, __builtin_fmaf(w1, bb.z, w0 * a.z)), 0.f);
  if (normals) {
    const float4 na = normals[t.x], nb = normals[t.y], nc = normals[t.z];
    const float3 n = f3(__builtin_fmaf(w2, nc.x, __builtin_fmaf(w1, nb.x, w0 * na.x)),
                        __builtin_fmaf(w2, nc.y, __builtin_fmaf(w1, nb.y, w0 * na.y)),
                        __builtin_fmaf(w2, nc.z, __builtin_fmaf(w1, nb.z, w0 * na.z)));
    const float is = rsqrt_c(dot3(n, n));
    nout[s] = make_float4(n.x * is, n.y * is, n.z * is, 0.f);
  }
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
  hipError_t ftl_launch_plain_##id(const FtlStep& A, hipStream_t stream, hipEvent_t, hipEvent_t, bool); \
  hipError_t ftl_launch_tiles_##id(const FtlStep*, const FtlSpan*, int, int, int, hipStream_t, bool);    \
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

size_t ftl_stream_store_bytes() { return (size_t)384 << 20; }

hipError_t ftl_launch_plain(int cls, const FtlStep& A, hipStream_t stream, hipEvent_t start, hipEvent_t stop,
                            bool nt) {
  switch (cls) {
#define FTL_LP(id, M, G, GU, W, CH) \
  case id:                   \
    return ftl_launch_plain_##id(A, stream, start, stop, nt);
    FTL_CLASSES(FTL_LP)
#undef FTL_LP
  }
  return hipErrorInvalidValue;
}

hipError_t ftl_launch_tiles(int cls, const FtlStep* grooms, const FtlSpan* spans, int nspans, int ntiles, int grid,
                            hipStream_t stream, bool nt) {
  switch (cls) {
#define FTL_LB(id, M, G, GU, W, CH) \
  case id:                   \
    return ftl_launch_tiles_##id(grooms, spans, nspans, ntiles, grid, stream, nt);
    FTL_CLASSES(FTL_LB)
#undef FTL_LB
  }
  return hipErrorInvalidValue;
}

hipError_t ftl_launch_skin(int mode, int nverts, const float4* bones, const int4* joints, const float4* weights,
                           const float4* rest, const float4* nrest, float4* out, float4* nout, hipStream_t stream) {
  if (nverts <= 0) return hipSuccess;
  const dim3 grid((nverts + 255) / 256), block(256);
  if (mode == kSkinLbs)
    hipLaunchKernelGGL(ftl_skin_kernel<kSkinLbs>, grid, block, 0, stream, nverts, bones, joints, weights, rest,
                       nrest, out, nout);
  else
    hipLaunchKernelGGL(ftl_skin_kernel<kSkinDq>, grid, block, 0, stream, nverts, bones, joints, weights, rest,
                       nrest, out, nout);
  return hipGetLastError();
}

hipError_t ftl_launch_root_frames(int nstrands, const int4* bind, const int4* tris, const float4* verts,
                                  const float4* normals, float4* roots, float4* nout, hipStream_t stream) {
  if (nstrands <= 0) return hipSuccess;
  hipLaunchKernelGGL(ftl_root_frames_kernel, dim3((nstrands + 255) / 256), dim3(256), 0, stream, nstrands, bind,
                     tris, verts, normals, roots, nout);
  return hipGetLastError();
}

static_assert(sizeof(FtlStep) % 16 == 0 && sizeof(FtlSpan) % 16 == 0, "descriptors upload in 16-byte units");

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
