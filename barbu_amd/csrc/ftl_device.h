// Device-side definitions shared by the FTL kernels: the per-groom step
// descriptor and the FP32 evaluation order of every formula of
// src/shaders/hair/01_simulation/cs_simulation.glsl.
//
// Every compound expression is spelled out as the IEEE operations it performs
// (mul, add, explicit fma) and the library is compiled with
// -ffp-contract=off, so the compiler can neither fuse nor reassociate.  The
// result of a step is therefore a fixed function of its inputs, independent of
// the launch shape (lanes per strand, tile size, batching).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include "ftl.h"

#define FTL_WAVE 64
#define FTL_NCOL 4

// One groom's step, by value in kernel arguments (plain launch) or in a device
// table (batched launch).  Mirrors the SSBO bindings and uniforms of
// cs_simulation.glsl:7-43 plus the constants it hard-codes.
struct FtlStep {
  float4* pos;                  // read_positions / write_positions (in place)
  float4* vel;                  // read_velocities / write_velocities (in place)
  const float4* roots;          // FTL_ROOT_REST: rest root per strand
  const int4* bind;             // FTL_ROOT_SKINNED: {triangle, w1 bits, w2 bits, 0}
  const int4* tris;             // FTL_ROOT_SKINNED: {a, b, c, 0} per triangle
  const float4* skinned;        // FTL_ROOT_SKINNED: skinned vertices
  const float4* wind;           // wind lattice, 2 float4 per node (layout at wind_fetch), or null
  unsigned* status;             // bit 0: a strand tip became non-finite
  int nstrands, npoints, iterations, vmode;
  int rmode, ncol, model_ident, wind_grid;
  float dt, dt2, inv_dt, sigma;
  float delta, maxv, maxv2, wind_freq;
  float force[4];               // kForceCoeff * gravity (w unused)
  float sph[FTL_NCOL][4];       // collider centre xyz, radius
  float r2[FTL_NCOL];           // radius * radius
  float model[12];              // row-major 3x4 root transform
  float wind_off[4];            // lattice offset of this step (w unused)
};

// One groom's share of a class's batched launch: tiles tile0, tile0 + 1, ...
// of the class's kernel step the groom's strands, 64 / G strands a tile.  A
// launch's spans are in tile order, and each workgroup finds its own by a
// binary search, so a batch uploads one record per groom, not per tile.
struct FtlSpan {
  int groom;  // index into the FtlStep table
  int tile0;  // first tile of the groom in its class's launch
  int cls;    // launch class (points per lane, lanes per strand)
  int cost;   // relative cost of a tile, for ordering
};

namespace ftl_dev {

__device__ __forceinline__ float3 f3(float x, float y, float z) { return make_float3(x, y, z); }
// c ? a : b per component (a select on float3 objects would select their
// addresses and keep arrays of them out of registers).
__device__ __forceinline__ float3 sel3(bool c, float3 a, float3 b) {
  return f3(c ? a.x : b.x, c ? a.y : b.y, c ? a.z : b.z);
}
__device__ __forceinline__ float3 sub3(float3 a, float3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ float3 mul3(float3 a, float s) { return f3(a.x * s, a.y * s, a.z * s); }
// fma(s, d, b) per component
__device__ __forceinline__ float3 fma3(float s, float3 d, float3 b) {
  return f3(__builtin_fmaf(s, d.x, b.x), __builtin_fmaf(s, d.y, b.y), __builtin_fmaf(s, d.z, b.z));
}
// dot(a, b) = fma(a.z, b.z, fma(a.y, b.y, a.x * b.x))
__device__ __forceinline__ float dot3(float3 a, float3 b) {
  return __builtin_fmaf(a.z, b.z, __builtin_fmaf(a.y, b.y, a.x * b.x));
}

// inversesqrt: integer seed, then three Newton steps y <- y * (1.5 - (x/2) y^2).
// Bit-reproducible (plain IEEE operations), <= 2 ulp.
__device__ __forceinline__ float rsqrt_c(float x) {
  unsigned i = __float_as_uint(x);
  i = 0x5f3759dfu - (i >> 1);
  float y = __uint_as_float(i);
  const float xh = 0.5f * x;
  y = y * __builtin_fmaf(-xh, y * y, 1.5f);
  y = y * __builtin_fmaf(-xh, y * y, 1.5f);
  y = y * __builtin_fmaf(-xh, y * y, 1.5f);
  return y;
}

// sin (cosine = 0) or cos (cosine = 1) of x, from plain IEEE operations so that
// the result is reproducible bit for bit: Cody-Waite reduction by pi/2 in three
// parts (k = rint(x * 2/pi), r = x - k pi/2 with one fma per part), then the
// Taylor polynomials of sin and cos on |r| <= pi/4 (terms to r^9 and r^10,
// Horner with fma) and the quadrant k + cosine.  Within 1.5 ulp of the exact
// value for |x| <= 200 (the range the groom generator uses).
__device__ __forceinline__ float trig_c(float x, int cosine) {
  const float k = __builtin_rintf(x * 0.636619747f);
  float r = __builtin_fmaf(-k, 1.57079637f, x);
  r = __builtin_fmaf(-k, -4.37113883e-08f, r);
  r = __builtin_fmaf(-k, -1.71512451e-15f, r);
  const int q = ((int)k + cosine) & 3;
  const float r2 = r * r;
  float ps = __builtin_fmaf(r2, 2.75573188e-06f, -1.98412701e-04f);
  ps = __builtin_fmaf(r2, ps, 8.33333377e-03f);
  ps = __builtin_fmaf(r2, ps, -1.66666672e-01f);
  const float s = __builtin_fmaf(r2 * r, ps, r);
  float pc = __builtin_fmaf(r2, -2.75573200e-07f, 2.48015876e-05f);
  pc = __builtin_fmaf(r2, pc, -1.38888892e-03f);
  pc = __builtin_fmaf(r2, pc, 4.16666679e-02f);
  pc = __builtin_fmaf(r2, pc, -0.5f);
  const float c = __builtin_fmaf(r2, pc, 1.0f);
  return q == 0 ? s : q == 1 ? c : q == 2 ? -s : -c;
}
__device__ __forceinline__ float sin_c(float x) { return trig_c(x, 0); }
__device__ __forceinline__ float cos_c(float x) { return trig_c(x, 1); }

// asin, acos, exp, log and pow from plain IEEE operations (reproducible bit for
// bit; within 2 ulp, pow 3 ulp, of the exact values on the Marschner tables'
// ranges):
// * asin on [0, 1/2]: a + a^3 P(a^2), P the Taylor series of asin to a^23
//   (coefficients (2n)! / (4^n (n!)^2 (2n+1)) rounded to float, Horner with
//   fma); above 1/2, asin a = pi/2 - 2 asin(sqrt((1 - a)/2)); acos from the same
//   core (pi/2 - asin, 2 asin(...), pi - 2 asin(...)), pi and pi/2 in two parts;
// * exp: Cody-Waite by ln 2 in two parts, the Taylor series of e^r to r^8,
//   then the scaling by 2^k (ldexp, one rounding);
// * log: x = 2^e m with m in [sqrt(1/2), sqrt(2)), log m = 2 atanh(s) with
//   s = (m - 1)/(m + 1) (series to s^11), plus e ln 2 in two parts;
// * pow(x, y) = exp(y log x) for x >= 0, pow(0, y) = 0.
__device__ __forceinline__ float asin_core_c(float a) {
  const float z = a * a;
  float p = __builtin_fmaf(z, 0.00731252600f, 0.00839033537f);
  p = __builtin_fmaf(z, p, 0.00976160914f);
  p = __builtin_fmaf(z, p, 0.0115518011f);
  p = __builtin_fmaf(z, p, 0.0139648439f);
  p = __builtin_fmaf(z, p, 0.0173527636f);
  p = __builtin_fmaf(z, p, 0.0223721582f);
  p = __builtin_fmaf(z, p, 0.0303819440f);
  p = __builtin_fmaf(z, p, 0.0446428582f);
  p = __builtin_fmaf(z, p, 0.0750000030f);
  p = __builtin_fmaf(z, p, 0.166666672f);
  return __builtin_fmaf(a * z, p, a);
}
__device__ __forceinline__ float asin_c(float x) {
  const float a = fabsf(x);
  float r;
  if (a <= 0.5f) {
    r = asin_core_c(a);
  } else {
    const float s = __builtin_sqrtf((1.0f - a) * 0.5f);
    r = __builtin_fmaf(-2.0f, asin_core_c(s), 1.57079637f) + -4.37113883e-08f;
  }
  return x < 0.0f ? -r : r;
}
__device__ __forceinline__ float acos_c(float x) {
  if (fabsf(x) <= 0.5f) return (1.57079637f - asin_core_c(fabsf(x)) * (x < 0.0f ? -1.0f : 1.0f)) + -4.37113883e-08f;
  if (x > 0.0f) return 2.0f * asin_core_c(__builtin_sqrtf((1.0f - x) * 0.5f));
  return __builtin_fmaf(-2.0f, asin_core_c(__builtin_sqrtf((1.0f + x) * 0.5f)), 3.14159274f) + -8.74227766e-08f;
}
__device__ __forceinline__ float exp_c(float x) {
  if (!(x < 89.0f)) return x != x ? x : __builtin_inff();
  if (x < -104.0f) return 0.0f;
  const float k = __builtin_rintf(x * 1.44269502f);
  float r = __builtin_fmaf(-k, 0.693147182f, x);
  r = __builtin_fmaf(-k, -1.90465421e-09f, r);
  float p = __builtin_fmaf(r, 2.48015876e-05f, 1.98412701e-04f);
  p = __builtin_fmaf(r, p, 1.38888892e-03f);
  p = __builtin_fmaf(r, p, 8.33333377e-03f);
  p = __builtin_fmaf(r, p, 4.16666679e-02f);
  p = __builtin_fmaf(r, p, 1.66666672e-01f);
  p = __builtin_fmaf(r, p, 0.5f);
  p = __builtin_fmaf(r, p, 1.0f);
  p = __builtin_fmaf(r, p, 1.0f);
  return __builtin_ldexpf(p, (int)k);
}
__device__ __forceinline__ float log_c(float x) {
  if (!(x > 0.0f) || x == __builtin_inff()) return x == 0.0f ? -__builtin_inff() : (x > 0.0f ? x : __builtin_nanf(""));
  int e = 0;
  if (x < 1.17549435e-38f) {
    x *= 16777216.0f;
    e = -24;
  }
  unsigned i = __float_as_uint(x);
  e += (int)((i >> 23) & 0xffu) - 127;
  float m = __uint_as_float((i & 0x007fffffu) | 0x3f800000u);
  if (m > 1.41421354f) {
    m *= 0.5f;
    e += 1;
  }
  const float f = m - 1.0f;
  const float s = f / (2.0f + f);
  const float s2 = s * s;
  float p = __builtin_fmaf(s2, 0.0909090936f, 0.111111112f);
  p = __builtin_fmaf(s2, p, 0.142857149f);
  p = __builtin_fmaf(s2, p, 0.200000003f);
  p = __builtin_fmaf(s2, p, 0.333333343f);
  const float lm = __builtin_fmaf(2.0f * s * s2, p, 2.0f * s);
  const float ef = (float)e;
  return __builtin_fmaf(ef, 0.693147182f, __builtin_fmaf(ef, -1.90465421e-09f, lm));
}
__device__ __forceinline__ float pow_c(float x, float y) { return x == 0.0f ? 0.0f : exp_c(y * log_c(x)); }

// One FTL point update, in place (DistanceConstraint, cs_simulation.glsl:110-116):
//   d = p - lead;  p = lead + (sl * inversesqrt(dot(d, d))) * d
// with the operations of dot3 / rsqrt_c / fma3 above, in the same order.
// Written out so the new position is produced in p's own registers (the
// compiler otherwise builds it in temporaries and copies it back every point).
// VALU instructions only; no memory access.
__device__ __forceinline__ void ftl_point(float3& p, const float3 lead, const float sl) {
  float d0, d1, d2, x, y, t, s;
  asm(
      "v_sub_f32 %[d0], %[p0], %[l0]\n\t"
      "v_sub_f32 %[d1], %[p1], %[l1]\n\t"
      "v_sub_f32 %[d2], %[p2], %[l2]\n\t"
      "v_mul_f32 %[x], %[d0], %[d0]\n\t"
      "v_fmac_f32 %[x], %[d1], %[d1]\n\t"
      "v_fmac_f32 %[x], %[d2], %[d2]\n\t"
      "v_lshrrev_b32 %[y], 1, %[x]\n\t"
      "v_sub_u32 %[y], 0x5f3759df, %[y]\n\t"
      "v_mul_f32 %[x], -0.5, %[x]\n\t"
      "v_mul_f32 %[t], %[y], %[y]\n\t"
      "v_fmaak_f32 %[t], %[x], %[t], 0x3fc00000\n\t"
      "v_mul_f32 %[y], %[y], %[t]\n\t"
      "v_mul_f32 %[t], %[y], %[y]\n\t"
      "v_fmaak_f32 %[t], %[x], %[t], 0x3fc00000\n\t"
      "v_mul_f32 %[y], %[y], %[t]\n\t"
      "v_mul_f32 %[t], %[y], %[y]\n\t"
      "v_fmaak_f32 %[t], %[x], %[t], 0x3fc00000\n\t"
      "v_mul_f32 %[y], %[y], %[t]\n\t"
      "v_mul_f32 %[s], %[sl], %[y]\n\t"
      "v_fma_f32 %[p0], %[s], %[d0], %[l0]\n\t"
      "v_fma_f32 %[p1], %[s], %[d1], %[l1]\n\t"
      "v_fma_f32 %[p2], %[s], %[d2], %[l2]"
      : [p0] "+v"(p.x), [p1] "+v"(p.y), [p2] "+v"(p.z), [d0] "=&v"(d0), [d1] "=&v"(d1), [d2] "=&v"(d2),
        [x] "=&v"(x), [y] "=&v"(y), [t] "=&v"(t), [s] "=&v"(s)
      : [l0] "v"(lead.x), [l1] "v"(lead.y), [l2] "v"(lead.z), [sl] "v"(sl));
}

// Root transform: row c = fma(m_c0, x, fma(m_c1, y, fma(m_c2, z, t_c))).
__device__ __forceinline__ float3 transform(const FtlStep& A, float3 x) {
  const float* m = A.model;
  return f3(__builtin_fmaf(m[0], x.x, __builtin_fmaf(m[1], x.y, __builtin_fmaf(m[2], x.z, m[3]))),
            __builtin_fmaf(m[4], x.x, __builtin_fmaf(m[5], x.y, __builtin_fmaf(m[6], x.z, m[7]))),
            __builtin_fmaf(m[8], x.x, __builtin_fmaf(m[9], x.y, __builtin_fmaf(m[10], x.z, m[11]))));
}

__device__ __forceinline__ float lerp_c(float a, float b, float t) { return __builtin_fmaf(t, b - a, a); }

// Wind acceleration at x: trilinear value noise over a periodic lattice with
// smoothstep weights (the force hook of CalculateForces, cs_simulation.glsl:70).
// Device layout of the lattice: two float4 per node n = (z*G + y)*G + x,
//   T[2n] = (v.x, v.y, v.z, d.x),  T[2n+1] = (d.y, d.z, 0, 0),
// with d = v(x+1, y, z) - v(x, y, z) rounded once when the table is built, so
// an x-edge is fma(s0, d, v): the same value as lerp(v, v(x+1), s0) computed
// with the subtraction at run time.
__device__ __forceinline__ float3 wind_edge(float4 e0, float2 e1, float s0) {
  return f3(__builtin_fmaf(s0, e0.w, e0.x), __builtin_fmaf(s0, e1.x, e0.y), __builtin_fmaf(s0, e1.y, e0.z));
}

// Wind parameters of one step, read once per tile.
// A pointer re-formed from scalar registers is a generic pointer; holding it
// as a global-address-space pointer keeps the gathers global loads (a flat
// load also counts against the LDS counter, and its wait waits for both).
#if defined(__HIP_DEVICE_COMPILE__)
typedef __attribute__((address_space(1))) const float4 GlobalFloat4;
#else
typedef const float4 GlobalFloat4;
#endif
struct WindArgs {
  GlobalFloat4* __restrict__ T;  // lattice (layout above), global memory
  float freq;                    // lattice cells per unit length
  float off[3];                  // lattice offset of this step
};

// The cell of x: smoothstep weights and the four x-edges, gathered together
// (one round trip).  wind_fetch issues the gathers; wind_blend consumes them,
// so a caller can place other memory operations in between.
struct WindTaps {
  float4 a[4];
  float2 b[4];
  float s0, s1, s2;
};

// T: the lattice in global memory (W.T) or a copy of it in LDS.
__device__ __forceinline__ WindTaps wind_fetch_from(const float4* __restrict__ T, const WindArgs& W, float3 x) {
  // The lattice side is the constant FTL_WIND_GRID (a power of two): corner
  // indices are masks and shifts, and every offset fits 32 bits, so the
  // gathers address the table from its scalar base.
  constexpr unsigned G = FTL_WIND_GRID, msk = FTL_WIND_GRID - 1;
  const float u0 = __builtin_fmaf(x.x, W.freq, W.off[0]);
  const float u1 = __builtin_fmaf(x.y, W.freq, W.off[1]);
  const float u2 = __builtin_fmaf(x.z, W.freq, W.off[2]);
  const float l0 = floorf(u0), l1 = floorf(u1), l2 = floorf(u2);
  const float f0 = u0 - l0, f1 = u1 - l1, f2 = u2 - l2;
  const unsigned i0 = (unsigned)(int)l0, i1 = (unsigned)(int)l1, i2 = (unsigned)(int)l2;
  const unsigned x0 = i0 & msk;
  const unsigned y0 = i1 & msk, y1 = (i1 + 1u) & msk;
  const unsigned z0 = i2 & msk, z1 = (i2 + 1u) & msk;
  WindTaps t;
  t.s0 = (f0 * f0) * __builtin_fmaf(-2.0f, f0, 3.0f);
  t.s1 = (f1 * f1) * __builtin_fmaf(-2.0f, f1, 3.0f);
  t.s2 = (f2 * f2) * __builtin_fmaf(-2.0f, f2, 3.0f);
  const unsigned n[4] = {(z0 * G + y0) * G + x0, (z0 * G + y1) * G + x0, (z1 * G + y0) * G + x0,
                         (z1 * G + y1) * G + x0};
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    t.a[e] = T[2u * n[e]];
    const float4 h = T[2u * n[e] + 1u];
    t.b[e] = make_float2(h.x, h.y);
  }
  return t;
}

__device__ __forceinline__ WindTaps wind_fetch(const WindArgs& W, float3 x) {
  return wind_fetch_from((const float4*)W.T, W, x);
}

// The same taps from a half lattice L that keeps each node's first float4,
// (v.x, v.y, v.z, d.x): d.y and d.z are formed here as v(x+1) - v(x), the
// single IEEE subtraction that built them in the full table.
__device__ __forceinline__ WindTaps wind_fetch_half(const float4* __restrict__ L, const WindArgs& W, float3 x) {
  constexpr unsigned G = FTL_WIND_GRID, msk = FTL_WIND_GRID - 1;
  const float u0 = __builtin_fmaf(x.x, W.freq, W.off[0]);
  const float u1 = __builtin_fmaf(x.y, W.freq, W.off[1]);
  const float u2 = __builtin_fmaf(x.z, W.freq, W.off[2]);
  const float l0 = floorf(u0), l1 = floorf(u1), l2 = floorf(u2);
  const float f0 = u0 - l0, f1 = u1 - l1, f2 = u2 - l2;
  const unsigned i0 = (unsigned)(int)l0, i1 = (unsigned)(int)l1, i2 = (unsigned)(int)l2;
  const unsigned x0 = i0 & msk, x1 = (i0 + 1u) & msk;
  const unsigned y0 = i1 & msk, y1 = (i1 + 1u) & msk;
  const unsigned z0 = i2 & msk, z1 = (i2 + 1u) & msk;
  WindTaps t;
  t.s0 = (f0 * f0) * __builtin_fmaf(-2.0f, f0, 3.0f);
  t.s1 = (f1 * f1) * __builtin_fmaf(-2.0f, f1, 3.0f);
  t.s2 = (f2 * f2) * __builtin_fmaf(-2.0f, f2, 3.0f);
  const unsigned r[4] = {(z0 * G + y0) * G, (z0 * G + y1) * G, (z1 * G + y0) * G, (z1 * G + y1) * G};
  float4 nb[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    t.a[e] = L[r[e] + x0];
    nb[e] = L[r[e] + x1];
  }
#pragma unroll
  for (int e = 0; e < 4; ++e) t.b[e] = make_float2(nb[e].y - t.a[e].y, nb[e].z - t.a[e].z);
  return t;
}

__device__ __forceinline__ float3 wind_blend(const WindTaps& t) {
  const float3 c00 = wind_edge(t.a[0], t.b[0], t.s0);
  const float3 c10 = wind_edge(t.a[1], t.b[1], t.s0);
  const float3 c01 = wind_edge(t.a[2], t.b[2], t.s0);
  const float3 c11 = wind_edge(t.a[3], t.b[3], t.s0);
  float3 r;
  r.x = lerp_c(lerp_c(c00.x, c10.x, t.s1), lerp_c(c01.x, c11.x, t.s1), t.s2);
  r.y = lerp_c(lerp_c(c00.y, c10.y, t.s1), lerp_c(c01.y, c11.y, t.s1), t.s2);
  r.z = lerp_c(lerp_c(c00.z, c10.z, t.s1), lerp_c(c01.z, c11.z, t.s1), t.s2);
  return r;
}

// The colliders of one step, read once per tile into scalar registers and
// made opaque there, so the per-point collision code never reloads them.
struct Spheres {
  float c[FTL_NCOL][3];
  float r[FTL_NCOL];
  float r2[FTL_NCOL];
};

__device__ __forceinline__ float pin_s(float x) {
  int i = __builtin_amdgcn_readfirstlane(__float_as_int(x));
  asm volatile("" : "+s"(i));
  return __int_as_float(i);
}

__device__ __forceinline__ int pin_i(int x) {
  int i = __builtin_amdgcn_readfirstlane(x);
  asm volatile("" : "+s"(i));
  return i;
}

template <typename T>
__device__ __forceinline__ T* pin_p(T* p) {
  unsigned long long u = (unsigned long long)p;
  int lo = __builtin_amdgcn_readfirstlane((int)(unsigned)u);
  int hi = __builtin_amdgcn_readfirstlane((int)(unsigned)(u >> 32));
  asm volatile("" : "+s"(lo), "+s"(hi));
  return (T*)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

__device__ __forceinline__ Spheres load_spheres(const FtlStep& A, unsigned cmask) {
  Spheres S;
#pragma unroll
  for (int k = 0; k < FTL_NCOL; ++k) {
    if (!(cmask & (1u << k))) {
      S.c[k][0] = S.c[k][1] = S.c[k][2] = S.r[k] = S.r2[k] = 0.0f;
      continue;
    }
    S.c[k][0] = pin_s(A.sph[k][0]);
    S.c[k][1] = pin_s(A.sph[k][1]);
    S.c[k][2] = pin_s(A.sph[k][2]);
    S.r[k] = pin_s(A.sph[k][3]);
    S.r2[k] = pin_s(A.r2[k]);
  }
  return S;
}

// CollideSphere(+1, c_k, r_k, p) for the colliders of `cmask`, in order
// (cs_simulation.glsl:129-153).  cmask is wave-uniform: the colliders some
// strand of the wave can reach (the others cannot change any result).
// REFLECT: also reflect u on each hit; only the last constraint iteration's
// velocity reaches the output.
// Test of bit k of the wave-uniform collider mask as a scalar branch.  The
// mask is made opaque at each test: hoisted out of the iteration loop, the
// test becomes a per-lane boolean that costs two vector instructions a use.
__device__ __forceinline__ bool has_collider(unsigned cmask, int k) {
  asm volatile("" : "+s"(cmask));
  return (cmask >> k) & 1u;
}

template <bool REFLECT>
__device__ __forceinline__ void collide(const Spheres& S, float3& p, float3& u, unsigned cmask) {
#pragma unroll
  for (int k = 0; k < FTL_NCOL; ++k) {
    if (!has_collider(cmask, k)) continue;
    const float3 c = f3(S.c[k][0], S.c[k][1], S.c[k][2]);
    const float3 w = sub3(p, c);
    const float dp = dot3(w, w);
    if (dp < S.r2[k]) {
      const float is = rsqrt_c(dp);
      const float3 n = mul3(w, is);
      p = fma3(S.r[k], n, c);
      if (REFLECT) {
        const float d2 = 2.0f * dot3(n, u);
        u = fma3(-d2, n, u);
      }
    }
  }
}

// Colliders a strand can touch during one step, as a bit mask.  After every
// sweep each point lies within the chain length L = sum |sigma*l| of its root
// p0 (segments are |sigma*l| long); collider j moves a point it contains to its surface, by
// at most r_j - |p0 - c_j| + D when D bounds the point's distance to p0, so
// D grows by that much for each collider that can touch the strand.  A
// collider farther than r_k + D (with a generous rounding margin) can never
// contain a point of the strand in this step.
__device__ __forceinline__ unsigned reachable_colliders(const FtlStep& A, float3 p0, float L) {
  float D = L * 1.001f + 1e-4f;
  const float a = sqrtf(dot3(p0, p0));
  unsigned mask = 0u;
  for (int k = 0; k < A.ncol; ++k) {
    const float3 c = f3(A.sph[k][0], A.sph[k][1], A.sph[k][2]);
    const float r = A.sph[k][3];
    const float dc = sqrtf(dot3(sub3(p0, c), sub3(p0, c)));
    const float tol = 1e-4f * (a + sqrtf(dot3(c, c)) + r + D + 1.0f);
    if (!(dc - r > D + tol)) {  // NaN keeps the collider
      mask |= 1u << k;
      D += fmaxf(0.0f, r - dc + D) + tol;
    }
  }
  return mask;
}

}  // namespace ftl_dev
