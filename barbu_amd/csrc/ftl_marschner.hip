// Marschner lookup tables on the device (include/ftl_render.h): the two
// compute passes of Marschner::generate (src/fx/marschner.cc:35-66),
//   M (cs_marschner_m.glsl:33-62): the longitudinal lobes R, TT, TRT and
//     cos(theta_d) over (sin theta_i, sin theta_r);
//   N (cs_marschner_n.glsl:32-84, generic branch): the azimuthal terms
//     Np(0), Np(1), Np(2) over (cos phi_d, cos theta_d)
//     (inc_marschner_n.glsl:28-131, inc_solver.glsl, inc_fresnel.glsl),
// one thread per texel, float4 texels (the reference stores RGBA16F).
// Integer powers are products: GLSL leaves pow(x, y) undefined for x < 0,
// which the reference reaches (cos theta_d < 0 on half the N table).
#include <hip/hip_runtime.h>

#include "ftl_internal.h"
#include "ftl_render.h"

namespace {

// inc_constants.glsl:6-8 (Pi() as the reference writes it)
constexpr float kEps = 1e-6f;
constexpr float kPiRef = 3.141564f;

struct Cubic {
  float r[3];
  int n;
};

// inc_solver.glsl:15-17
__device__ Cubic solve_linear(float a, float b) {
  Cubic c{{0.f, 0.f, 0.f}, 0};
  if (fabsf(a) > kEps) {
    c.r[0] = -b / a;
    c.n = 1;
  }
  return c;
}

// inc_solver.glsl:21-33 (w = 1 + step(eps, delta): a double root counts once)
__device__ Cubic solve_quadratic(float a, float b, float c) {
  if (fabsf(a) < kEps) return solve_linear(b, c);
  float delta = b * b - 4.0f * a * c;
  Cubic q{{0.f, 0.f, 0.f}, 0};
  if (delta < 0.0f) return q;
  delta = sqrtf(delta);
  q.r[0] = (-b + delta) / (2.0f * a);
  q.r[1] = (-b - delta) / (2.0f * a);
  q.n = 1 + (delta >= kEps ? 1 : 0);
  return q;
}

// inc_solver.glsl:48-83
__device__ Cubic solve_cubic_normalized(float a, float b, float c) {
  Cubic r{{0.f, 0.f, 0.f}, 0};
  if (fabsf(c) < kEps) {
    r = solve_quadratic(1.0f, a, b);
    r.r[r.n] = 0.0f;
    r.n += 1;
    return r;
  }
  const float Q = (3.0f * b - a * a) / 9.0f;
  const float R = (9.0f * a * b - 27.0f * c - 2.0f * a * a * a) / 54.0f;
  const float Q3 = Q * Q * Q;
  const float D = Q3 + R * R;
  const float third_a = a / 3.0f;
  if (D > 0.0f) {
    const float sqrtD = sqrtf(D);
    const float u = R + sqrtD, v = R - sqrtD;
    const float s = (u > 0.0f ? 1.0f : (u < 0.0f ? -1.0f : 0.0f)) * ftl_dev::pow_c(fabsf(u), 0.333f);
    const float t = (v > 0.0f ? 1.0f : (v < 0.0f ? -1.0f : 0.0f)) * ftl_dev::pow_c(fabsf(v), 0.333f);
    r.r[0] = s + t - third_a;
    r.n = 1;
  } else {
    const float theta = ftl_dev::acos_c(R * (1.0f / sqrtf(-Q3)));
    const float two_sqrt_q = 2.0f * sqrtf(-Q);
    r.r[0] = __builtin_fmaf(two_sqrt_q, ftl_dev::cos_c(theta / 3.0f), -third_a);
    r.r[1] = __builtin_fmaf(two_sqrt_q, ftl_dev::cos_c((theta + 2.0f * kPiRef) / 3.0f), -third_a);
    r.r[2] = __builtin_fmaf(two_sqrt_q, ftl_dev::cos_c((theta + 4.0f * kPiRef) / 3.0f), -third_a);
    r.n = 3;
  }
  return r;
}

// inc_solver.glsl:39-44
__device__ Cubic solve_cubic(float a, float b, float c, float d) {
  return (fabsf(a) < kEps) ? solve_quadratic(b, c, d) : solve_cubic_normalized(b / a, c / a, d / a);
}

// inc_fresnel.glsl:14-24
__device__ float fresnel_power_ratio(float eta_ratio, float nA, float nB, float cosI, float sinI) {
  const float e = eta_ratio * sinI;
  const float sinT2 = e * e;
  if (sinT2 > 1.0f) return 1.0f;
  const float cosT = sqrtf(1.0f - sinT2);
  const float A = nA * cosI, B = nB * cosT;
  const float R = (A - B) / (A + B);
  return fminf(1.0f, R * R);
}

// inc_fresnel.glsl:50-56: mix(reflective(eta_perp), transmittance(eta_par), 0.5)
__device__ float fresnel(float eta0, float eta_perp, float eta_par, float cosA, float sinA) {
  const float r = fresnel_power_ratio(eta0 / eta_perp, eta0, eta_perp, cosA, sinA);
  const float t = fresnel_power_ratio(eta0 / eta_par, eta_par, eta0, cosA, sinA);
  return r * (1.0f - 0.5f) + t * 0.5f;
}

// inc_marschner_n.glsl:64-92
__device__ float absorption_term(int p, float absorption, float eta_perp, float eta_par, float cosG, float sinG) {
  if (p == 0) return fresnel(1.0f, eta_perp, eta_par, cosG, sinG);
  const float sinT = sinG / eta_perp;
  const float gT = ftl_dev::asin_c(sinT);
  const float cosT = ftl_dev::cos_c(gT);
  const float fi = fresnel(1.0f, eta_perp, eta_par, cosG, sinG);
  const float ft = fresnel(1.0f, 1.0f / eta_perp, 1.0f / eta_par, cosT, sinT);
  const float t = ftl_dev::exp_c(-4.0f * absorption * (cosT * cosT));
  const float omf = 1.0f - fi;
  float a = omf * omf;
  for (int k = 0; k < p - 1; ++k) a *= ft;
  for (int k = 0; k < p; ++k) a *= t;
  return a;
}

// inc_marschner_n.glsl:97-131 (Np), with the coefficients of :28-39 and the
// derivative of :43-49
__device__ float Np(int p, float absorption, float eta_perp, float eta_par, float c, float phi) {
  const float kMinusEightOverPiCube = -0.25801227547f, kSixOverPi = 1.90985931710f;
  const float A = (float)p * c * kMinusEightOverPiCube;
  const float C = (float)p * c * kSixOverPi - 2.0f;
  const float D = (float)p * kPiRef - phi;
  const Cubic roots = solve_cubic(A, 0.0f, C, D);
  float L = 0.0f;
  for (int i = 0; i < roots.n; ++i) {
    const float g = roots.r[i];
    const float sinG = ftl_dev::sin_c(g), cosG = ftl_dev::cos_c(g);
    const float a = absorption_term(p, absorption, eta_perp, eta_par, cosG, sinG);
    const float d1 = (3.0f * A * (g * g) + C) / cosG;
    L += a * (1.0f / fabsf(2.0f * d1));
  }
  return fminf(L, 1.0f);
}

__device__ float gaussian(float sigma, float x_mu) {
  return ftl_dev::exp_c(-(x_mu * x_mu) / (2.0f * sigma * sigma)) / (2.5066282f * fabsf(sigma));
}

struct LutArgs {
  float4* m;
  float4* n;
  int res;
  float inv_res;
  float eta, absorption, ar, br;
};

__global__ __launch_bounds__(256) void ftl_marschner_kernel(LutArgs A) {
  const int idx = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (idx >= A.res * A.res) return;
  const int x = idx % A.res, y = idx / A.res;
  if (blockIdx.y == 0) {
    // M: cs_marschner_m.glsl:37-61
    const float sinI = 2.0f * (float)x * A.inv_res - 1.0f;
    const float sinR = 2.0f * (float)y * A.inv_res - 1.0f;
    const float thI = ftl_dev::asin_c(sinI), thR = ftl_dev::asin_c(sinR);
    const float thH = (thI + thR) / 2.0f, thD = (thI - thR) / 2.0f;
    const float deg = thH * (180.0f / 3.14159265358979323846f);
    const float lr = gaussian(1.0f * A.br, deg - 1.0f * A.ar);
    const float ltt = gaussian(0.5f * A.br, deg - (-0.5f * A.ar));
    const float ltrt = gaussian(2.0f * A.br, deg - (-1.5f * A.ar));
    A.m[idx] = make_float4(lr, ltt, ltrt, ftl_dev::cos_c(thD));
  } else {
    // N: cs_marschner_n.glsl:37-82
    const float cosPhiD = 2.0f * (float)x * A.inv_res - 1.0f;
    const float cosThD = 2.0f * (float)y * A.inv_res - 1.0f;
    const float sinThD2 = 1.0f - cosThD * cosThD;
    const float eta2 = A.eta * A.eta;
    const float eta_perp = sqrtf(eta2 - sinThD2) / cosThD;
    const float eta_par = eta2 / eta_perp;
    const float phiD = ftl_dev::acos_c(cosPhiD);
    const float c = ftl_dev::asin_c(1.0f / eta_perp);
    const float R = Np(0, A.absorption, eta_perp, eta_par, c, phiD);
    const float TT = Np(1, A.absorption, eta_perp, eta_par, c, phiD);
    const float TRT = Np(2, A.absorption, eta_perp, eta_par, c, phiD);
    A.n[idx] = make_float4(R, TT, TRT, 1.0f);
  }
}

}  // namespace

extern "C" void ftl_marschner_params_reference(ftl_marschner_params* p) {
  if (!p) return;
  p->eta = 1.55f;
  p->absorption = 0.20f;
  p->eccentricity = 0.85f;
  p->ar = -5.0f;
  p->br = 5.0f;
  p->glint_scale = 0.5f;
  p->azimuthal_width = 10.0f;
  p->delta_caustic = 0.2f;
  p->delta_hm = 0.5f;
}

extern "C" ftl_status ftl_marschner_luts(ftl_ctx* ctx, const ftl_marschner_params* p, int res, float* out_m,
                                         float* out_n, void* stream) {
  if (!ctx || !p || !out_m || !out_n) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "NULL argument");
  if (res < 2 || res > 8192) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "ftl_marschner_luts: resolution %d", res);
  hipError_t he;
  if ((he = hipSetDevice(ctx->device)) != hipSuccess) return ftl_hip_fail(he, "hipSetDevice");
  const size_t bytes = (size_t)res * res * sizeof(float4);
  ftl_status st;
  if ((st = ftl_check_device_range(out_m, bytes, "out_m")) != FTL_OK ||
      (st = ftl_check_device_range(out_n, bytes, "out_n")) != FTL_OK)
    return st;
  LutArgs A{(float4*)out_m, (float4*)out_n, res, 1.0f / (float)res, p->eta, p->absorption, p->ar, p->br};
  const int n = res * res;
  hipLaunchKernelGGL(ftl_marschner_kernel, dim3((unsigned)((n + 255) / 256), 2), dim3(256), 0, (hipStream_t)stream, A);
  if ((he = hipGetLastError()) != hipSuccess) return ftl_hip_fail(he, "ftl_marschner_kernel launch");
  return FTL_OK;
}
