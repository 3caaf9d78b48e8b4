// FTL hair-strand step: CPU oracle.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the product links, loads or calls this
// file.  The parity tests, the smoke check and the CPU-baseline leg of the
// benchmark use it as the checker; it is never the thing measured as the
// product, and the product never falls back to it.
//
// What it restates (paths relative to the tcoppex/barbu tree):
//   src/shaders/hair/01_simulation/cs_simulation.glsl   (whole file)
//     UnpackParticle/PackParticle        :53-66
//     CalculateForces                    :70-77   (+ a wind hook: value noise)
//     VelocityControl                    :79-84   (DFTL mode only)
//     DistanceConstraint                 :93-127  (serial sweep + damped shift)
//     CollideSphere / CollisionConstraint:129-153
//     SatisfyConstraints                 :155-161
//     main (predict / root / pack)       :170-208
//   src/shaders/hair/interop.h:8-30        (float4 pos with w = rest length, float4 vel)
//   src/fx/hair.cc:89-125                  (one update = one dispatch; the next
//                                            step reads this step's output)
//   src/shaders/shared/inc_skinning.glsl:22-129 (dual-quaternion and linear
//                                          blending of positions and normals,
//                                          kNoJoint; skinned roots)
//   src/shaders/hair/02_tess_stream/{vs,tcs,tes,gs}_stream_hair.glsl and
//   src/shaders/shared/inc_maths.glsl:107-115,210-228,239-241,263-266
//                                          (render-side strand expansion, the
//                                           stream pass of src/fx/hair.cc:144-170)
//   src/shaders/hair/marschner/*.glsl, src/shaders/shared/inc_{solver,fresnel}.glsl
//                                          (Marschner lookup tables, src/fx/marschner.cc:35-66)
//   src/fx/hair.cc:236-328                 (init_simulation: positions, velocities,
//                                           "curly" tangents; glm::simplex restated
//                                           from the published webgl-noise algorithm)
//
// The GLSL leaves rounding of compound expressions to the implementation:
// without `precise` a compiler may fuse a*b+c and apply a scalar factor once.
// The FP32 path below fixes ONE such permitted evaluation order, written out
// operation by operation, so that any other implementation that performs the
// same IEEE operations in the same order (the HIP kernels do) produces the same
// bits.  `inversesqrt` is left to the implementation by GLSL (2 ulp); the FP32
// path uses a bit-reproducible integer-seeded Newton iteration (<= 2 ulp,
// checked exhaustively over one binade pair by the tests).  sin, cos, asin,
// acos, exp, log and pow (the tangents and the Marschner tables) are likewise
// one FP32 definition each, built from their series (trig_f32, asin_f32, ...;
// within 2-3 ulp, checked against float64 by the tests).  The FP64 path
// evaluates the same formulas in double with 1/sqrt and the C library's
// functions, and serves as the accuracy reference for the FP32 path.
//
// Build: oracle/Makefile (g++ -O2 -ffp-contract=off, no fast-math).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>

#ifdef _OPENMP
#include <omp.h>
#endif

extern "C" {

// Mirrors the uniforms of cs_simulation.glsl plus the parameters the shipped
// shader hard-codes (force coefficient, iteration count, damping).
typedef struct oracle_params {
  int   npoints;          // HAIR_MAX_PARTICLE_PER_STRAND (interop.h:8)
  int   iterations;       // kNumContraintsIteration (cs_simulation.glsl:197)
  int   velocity_mode;    // 0 = shipped (stored velocity = damped correction), 1 = DFTL Eq. 9 (:201-204)
  int   root_mode;        // 0 = p0 = M * x0 (:190-191), 1 = p0 = M * roots[strand]
  float dt;               // uTimeStep
  float inv_dt;           // 1/dt, used by DFTL mode
  float length_scale;     // uScaleFactor (sigma)
  float damping;          // dftl_damp_scale (delta)
  float max_velocity;     // uMaxVelocity
  float force[3];         // kForceCoeff * gravity (already multiplied in float)
  int   ncolliders;       // 0..4 spheres, applied in order (uBoundingSphere is one)
  float spheres[4][4];    // xyz centre, w radius
  float model[12];        // row-major 3x4 affine applied to roots
  int   model_is_identity;
  int   wind_grid;        // 0 = no wind, else lattice side (power of two)
  float wind_frequency;   // lattice cells per metre
  float wind_offset[3];   // lattice-space offset for this step
  const float* wind_table;  // float4 [grid^3], index (z*grid + y)*grid + x
} oracle_params;

}  // extern "C"

namespace {

template <typename T>
struct V3 {
  T x, y, z;
};

template <typename T>
inline V3<T> sub(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }

// dot(a, b) evaluated as fma(a.z, b.z, fma(a.y, b.y, a.x * b.x)).
template <typename T>
inline T dot3(V3<T> a, V3<T> b) {
  return std::fma(a.z, b.z, std::fma(a.y, b.y, a.x * b.x));
}

// inversesqrt.  FP32: seed from the exponent/mantissa bits, then three Newton
// steps y <- y * (1.5 - (x/2) * y^2).  FP64: 1/sqrt.
inline float rsqrt_t(float x) {
  uint32_t i;
  std::memcpy(&i, &x, 4);
  i = 0x5f3759dfu - (i >> 1);
  float y;
  std::memcpy(&y, &i, 4);
  const float xh = 0.5f * x;
  y = y * std::fma(-xh, y * y, 1.5f);
  y = y * std::fma(-xh, y * y, 1.5f);
  y = y * std::fma(-xh, y * y, 1.5f);
  return y;
}
inline double rsqrt_t(double x) { return 1.0 / std::sqrt(x); }

template <typename T>
struct Particle {  // Particle_t (cs_simulation.glsl:47-51)
  V3<T> position;
  V3<T> velocity;
  T restlength;
};

template <typename T>
struct StepConsts {
  T dt, dt2, inv_dt, scale, damp, maxv, maxv2;
  V3<T> force;
  int n, iters, vmode, rmode, ncol;
  V3<T> c[4];
  T r[4], r2[4];
  T m[12];
  bool ident;
  int wgrid;
  T wfreq;
  V3<T> woff;
  const float* wtab;
};

template <typename T>
StepConsts<T> make_consts(const oracle_params* P) {
  StepConsts<T> k{};
  k.dt = (T)P->dt;
  k.dt2 = k.dt * k.dt;
  k.inv_dt = (T)P->inv_dt;
  k.scale = (T)P->length_scale;
  k.damp = (T)P->damping;
  k.maxv = (T)P->max_velocity;
  k.maxv2 = k.maxv * k.maxv;
  k.force = {(T)P->force[0], (T)P->force[1], (T)P->force[2]};
  k.n = P->npoints;
  k.iters = P->iterations;
  k.vmode = P->velocity_mode;
  k.rmode = P->root_mode;
  k.ncol = P->ncolliders < 0 ? 0 : (P->ncolliders > 4 ? 4 : P->ncolliders);
  for (int s = 0; s < 4; ++s) {
    k.c[s] = {(T)P->spheres[s][0], (T)P->spheres[s][1], (T)P->spheres[s][2]};
    k.r[s] = (T)P->spheres[s][3];
    k.r2[s] = k.r[s] * k.r[s];  // radius * radius (:133)
  }
  for (int i = 0; i < 12; ++i) k.m[i] = (T)P->model[i];
  k.ident = P->model_is_identity != 0;
  k.wgrid = P->wind_table ? P->wind_grid : 0;
  k.wfreq = (T)P->wind_frequency;
  k.woff = {(T)P->wind_offset[0], (T)P->wind_offset[1], (T)P->wind_offset[2]};
  k.wtab = P->wind_table;
  return k;
}

// Affine root transform p = M * (x, 1), row c evaluated as
// fma(m_c0, x, fma(m_c1, y, fma(m_c2, z, t_c))).
template <typename T>
V3<T> transform(const StepConsts<T>& k, V3<T> x) {
  if (k.ident) return x;
  const T* m = k.m;
  return {std::fma(m[0], x.x, std::fma(m[1], x.y, std::fma(m[2], x.z, m[3]))),
          std::fma(m[4], x.x, std::fma(m[5], x.y, std::fma(m[6], x.z, m[7]))),
          std::fma(m[8], x.x, std::fma(m[9], x.y, std::fma(m[10], x.z, m[11])))};
}

template <typename T>
inline T lerp_t(T a, T b, T t) { return std::fma(t, b - a, a); }

// Wind: trilinear value noise over a periodic lattice of random vectors, with
// smoothstep weights.  The shipped shader has no wind; CalculateForces (:70-77)
// takes the particle and ignores it, which is the hook this fills.
template <typename T>
V3<T> wind(const StepConsts<T>& k, V3<T> x) {
  if (k.wgrid == 0) return {T(0), T(0), T(0)};
  const int G = k.wgrid, msk = G - 1;
  T u[3] = {std::fma(x.x, k.wfreq, k.woff.x), std::fma(x.y, k.wfreq, k.woff.y),
            std::fma(x.z, k.wfreq, k.woff.z)};
  int i0[3], i1[3];
  T s[3];
  for (int c = 0; c < 3; ++c) {
    const T fl = std::floor(u[c]);
    const T f = u[c] - fl;
    const int ii = (int)fl;
    i0[c] = ii & msk;
    i1[c] = (ii + 1) & msk;
    s[c] = (f * f) * std::fma(T(-2), f, T(3));
  }
  V3<T> v[2][2][2];  // [dz][dy][dx]
  for (int dz = 0; dz < 2; ++dz)
    for (int dy = 0; dy < 2; ++dy)
      for (int dx = 0; dx < 2; ++dx) {
        const int ix = dx ? i1[0] : i0[0], iy = dy ? i1[1] : i0[1], iz = dz ? i1[2] : i0[2];
        const float* e = k.wtab + 4 * ((iz * G + iy) * G + ix);
        v[dz][dy][dx] = {(T)e[0], (T)e[1], (T)e[2]};
      }
  V3<T> r;
  T* out = &r.x;
  for (int c = 0; c < 3; ++c) {
    auto g = [&](int dz, int dy, int dx) { return (&v[dz][dy][dx].x)[c]; };
    const T c00 = lerp_t(g(0, 0, 0), g(0, 0, 1), s[0]);
    const T c10 = lerp_t(g(0, 1, 0), g(0, 1, 1), s[0]);
    const T c01 = lerp_t(g(1, 0, 0), g(1, 0, 1), s[0]);
    const T c11 = lerp_t(g(1, 1, 0), g(1, 1, 1), s[0]);
    const T c0 = lerp_t(c00, c10, s[1]);
    const T c1 = lerp_t(c01, c11, s[1]);
    out[c] = lerp_t(c0, c1, s[2]);
  }
  return r;
}

// CollideSphere(+1, centre, radius, p)  (cs_simulation.glsl:129-139).
template <typename T>
void collide_sphere(const StepConsts<T>& k, int s, Particle<T>& p) {
  const V3<T> pt = sub(p.position, k.c[s]);
  const T dp = dot3(pt, pt);
  if (dp < k.r2[s]) {
    const T is = rsqrt_t(dp);
    const V3<T> n = {pt.x * is, pt.y * is, pt.z * is};
    // center + radius * n
    p.position = {std::fma(k.r[s], n.x, k.c[s].x), std::fma(k.r[s], n.y, k.c[s].y),
                  std::fma(k.r[s], n.z, k.c[s].z)};
    // reflect(v, n) = v - 2 dot(n, v) n
    const T d2 = T(2) * dot3(n, p.velocity);
    p.velocity = {std::fma(-d2, n.x, p.velocity.x), std::fma(-d2, n.y, p.velocity.y),
                  std::fma(-d2, n.z, p.velocity.z)};
  }
}

// One strand, one step: main() of cs_simulation.glsl run by all N invocations
// of one workgroup, with DistanceConstraint's lane-0 loops over the shared array.
template <typename T, typename S>
void step_strand(const StepConsts<T>& k, S* pos, S* vel, const S* root) {
  const int N = k.n;
  Particle<T> P[64];
  V3<T> last[64];
  // UnpackParticle (:53-61)
  for (int i = 0; i < N; ++i) {
    P[i].position = {(T)pos[4 * i + 0], (T)pos[4 * i + 1], (T)pos[4 * i + 2]};
    P[i].restlength = (T)pos[4 * i + 3];
    P[i].velocity = {(T)vel[4 * i + 0], (T)vel[4 * i + 1], (T)vel[4 * i + 2]};
    last[i] = P[i].position;
  }
  // Integrate (:179-193)
  for (int i = 0; i < N; ++i) {
    Particle<T>& p = P[i];
    if (i > 0) {
      const V3<T> w = wind(k, p.position);
      const V3<T> F = {k.force.x + w.x, k.force.y + w.y, k.force.z + w.z};
      p.position = {std::fma(k.dt2, F.x, std::fma(k.dt, p.velocity.x, p.position.x)),
                    std::fma(k.dt2, F.y, std::fma(k.dt, p.velocity.y, p.position.y)),
                    std::fma(k.dt2, F.z, std::fma(k.dt, p.velocity.z, p.position.z))};
    } else {
      V3<T> src = last[0];
      if (k.rmode == 1) src = {(T)root[0], (T)root[1], (T)root[2]};
      p.position = transform(k, src);
      p.velocity = sub(p.position, last[0]);
    }
  }
  // SatisfyConstraints (:155-161)
  for (int it = 0; it < k.iters; ++it) {
    // DistanceConstraint, lane-0 body (:109-122).  Only lane 0's damping value
    // reaches the shift loop, so the per-index curve of :100-102 never applies.
    for (int i = 1; i < N; ++i) {
      const V3<T> p0 = P[i - 1].position;
      const V3<T> p1 = P[i].position;
      const V3<T> d = sub(p1, p0);
      const T s = (k.scale * P[i].restlength) * rsqrt_t(dot3(d, d));
      const V3<T> nb = {std::fma(s, d.x, p0.x), std::fma(s, d.y, p0.y), std::fma(s, d.z, p0.z)};
      P[i].position = nb;
      P[i].velocity = sub(nb, p1);
    }
    for (int i = 1; i < N - 1; ++i) {
      const V3<T> v = P[i + 1].velocity;
      P[i].velocity = {v.x * k.damp, v.y * k.damp, v.z * k.damp};
    }
    // CollisionConstraint (:149-153): every sphere in order, root excluded.
    for (int i = 1; i < N; ++i)
      for (int s = 0; s < k.ncol; ++s) collide_sphere(k, s, P[i]);
  }
  // DFTL Eq. 9 and VelocityControl (:200-204), disabled in the shipped shader.
  if (k.vmode == 1) {
    for (int i = 0; i < N; ++i) {
      Particle<T>& p = P[i];
      const V3<T> a = sub(p.position, last[i]);
      V3<T> v = {(a.x - p.velocity.x) * k.inv_dt, (a.y - p.velocity.y) * k.inv_dt,
                 (a.z - p.velocity.z) * k.inv_dt};
      const T dp = dot3(v, v);
      if (dp > k.maxv2) {
        const T f = k.maxv * rsqrt_t(dp);
        v = {v.x * f, v.y * f, v.z * f};
      }
      p.velocity = v;
    }
  }
  // PackParticle (:63-66)
  for (int i = 0; i < N; ++i) {
    pos[4 * i + 0] = (S)P[i].position.x;
    pos[4 * i + 1] = (S)P[i].position.y;
    pos[4 * i + 2] = (S)P[i].position.z;
    pos[4 * i + 3] = (S)P[i].restlength;
    vel[4 * i + 0] = (S)P[i].velocity.x;
    vel[4 * i + 1] = (S)P[i].velocity.y;
    vel[4 * i + 2] = (S)P[i].velocity.z;
    vel[4 * i + 3] = S(0);
  }
}

template <typename T, typename S>
int step_all(const oracle_params* P, int nstrands, S* pos, S* vel, const S* roots, int nthreads) {
  if (!P || P->npoints < 1 || P->npoints > 64 || nstrands < 0) return 1;
  if (P->root_mode == 1 && !roots) return 1;
  const StepConsts<T> k = make_consts<T>(P);
  const int N = P->npoints;
  (void)nthreads;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
#endif
  for (int s = 0; s < nstrands; ++s) {
    step_strand<T, S>(k, pos + (size_t)4 * N * s, vel + (size_t)4 * N * s,
                      roots ? roots + (size_t)4 * s : nullptr);
  }
  return 0;
}

// --- dual-quaternion skinning (inc_skinning.glsl:22-31, 54-82) ------------

template <typename T>
V3<T> cross_t(V3<T> a, V3<T> b) {
  return {std::fma(a.y, b.z, -(a.z * b.y)), std::fma(a.z, b.x, -(a.x * b.z)),
          std::fma(a.x, b.y, -(a.y * b.x))};
}

template <typename T>
T sign_t(T x) { return x > T(0) ? T(1) : (x < T(0) ? T(-1) : T(0)); }

template <typename T>
T dot4(const T* a, const T* b) {
  return std::fma(a[3], b[3], std::fma(a[2], b[2], std::fma(a[1], b[1], a[0] * b[0])));
}

// skinning_DQBS (inc_skinning.glsl:54-82) after apply_skinning's weight rule
// (:22-31).  normal / nout (optional, float4): the normal is rotated by the
// blended real part (:80-81), not renormalised.
template <typename T>
void skin_vertex(const float* dq, const int* joints, const float* weights, const float* rest,
                 float* out, const float* normal = nullptr, float* nout = nullptr) {
  V3<T> v = {(T)rest[0], (T)rest[1], (T)rest[2]};
  T w[4] = {(T)weights[0], (T)weights[1], (T)weights[2], T(0)};
  if (w[0] <= T(1e-6f)) {  // Epsilon() (inc_constants.glsl:6)
    out[0] = (float)v.x; out[1] = (float)v.y; out[2] = (float)v.z; out[3] = rest[3];
    if (nout) { nout[0] = normal[0]; nout[1] = normal[1]; nout[2] = normal[2]; nout[3] = normal[3]; }
    return;
  }
  w[3] = T(1) - (w[0] + w[1] + w[2]);
  T Ma[4][4], Mb[4][4];
  for (int j = 0; j < 4; ++j)
    for (int c = 0; c < 4; ++c) {
      Ma[j][c] = (T)dq[8 * joints[j] + c];
      Mb[j][c] = (T)dq[8 * joints[j] + 4 + c];
    }
  // weights.xyz *= sign(Ma[3] * mat3x4(Ma))
  for (int j = 0; j < 3; ++j) w[j] = w[j] * sign_t(dot4(Ma[3], Ma[j]));
  T A[4], B[4];
  for (int c = 0; c < 4; ++c) {
    A[c] = std::fma(Ma[3][c], w[3], std::fma(Ma[2][c], w[2], std::fma(Ma[1][c], w[1], Ma[0][c] * w[0])));
    B[c] = std::fma(Mb[3][c], w[3], std::fma(Mb[2][c], w[2], std::fma(Mb[1][c], w[1], Mb[0][c] * w[0])));
  }
  const T inv = rsqrt_t(dot4(A, A));
  for (int c = 0; c < 4; ++c) {
    A[c] = A[c] * inv;
    B[c] = B[c] * inv;
  }
  const V3<T> a = {A[0], A[1], A[2]}, b = {B[0], B[1], B[2]};
  // v += 2 cross(a, cross(a, v) + a.w v)
  const V3<T> cv = cross_t(a, v);
  const V3<T> t = {std::fma(A[3], v.x, cv.x), std::fma(A[3], v.y, cv.y), std::fma(A[3], v.z, cv.z)};
  const V3<T> r = cross_t(a, t);
  v = {std::fma(T(2), r.x, v.x), std::fma(T(2), r.y, v.y), std::fma(T(2), r.z, v.z)};
  // v += 2 (a.w b - b.w a + cross(a, b))
  const V3<T> ab = cross_t(a, b);
  const V3<T> tr = {std::fma(A[3], b.x, std::fma(-B[3], a.x, ab.x)),
                    std::fma(A[3], b.y, std::fma(-B[3], a.y, ab.y)),
                    std::fma(A[3], b.z, std::fma(-B[3], a.z, ab.z))};
  v = {std::fma(T(2), tr.x, v.x), std::fma(T(2), tr.y, v.y), std::fma(T(2), tr.z, v.z)};
  out[0] = (float)v.x; out[1] = (float)v.y; out[2] = (float)v.z; out[3] = rest[3];
  if (nout) {
    // n += 2 cross(a, cross(a, n) + a.w n), the rotation above
    V3<T> n = {(T)normal[0], (T)normal[1], (T)normal[2]};
    const V3<T> cn = cross_t(a, n);
    const V3<T> tn = {std::fma(A[3], n.x, cn.x), std::fma(A[3], n.y, cn.y), std::fma(A[3], n.z, cn.z)};
    const V3<T> rn = cross_t(a, tn);
    n = {std::fma(T(2), rn.x, n.x), std::fma(T(2), rn.y, n.y), std::fma(T(2), rn.z, n.z)};
    nout[0] = (float)n.x; nout[1] = (float)n.y; nout[2] = (float)n.z; nout[3] = normal[3];
  }
}

// skinning_LBS (inc_skinning.glsl:85-129) after apply_skinning's weight rule.
// mats: 12 floats per bone, the three rows of its skinning matrix (the
// transposed 3x4 the host uploads, skeleton_controller.cc:251-256); a joint
// index of 0xff (kNoJoint, :18) is the identity (:89-93).  Each bone's
// transform of (v, 1) and of (n, 0) is dot(vertex, row c) evaluated as
// fma(x3, r3, fma(x2, r2, fma(x1, r1, x0 * r0))), then blended as
// fma(w3, t3, fma(w2, t2, fma(w1, t1, w0 * t0))) (M * _weights, :117).
template <typename T>
void lbs_vertex(const float* mats, const int* joints, const float* weights, const float* rest, float* out,
                const float* normal = nullptr, float* nout = nullptr) {
  T w[4] = {(T)weights[0], (T)weights[1], (T)weights[2], T(0)};
  if (w[0] <= T(1e-6f)) {
    for (int c = 0; c < 4; ++c) out[c] = rest[c];
    if (nout)
      for (int c = 0; c < 4; ++c) nout[c] = normal[c];
    return;
  }
  w[3] = T(1) - (w[0] + w[1] + w[2]);
  T R[4][3][4];
  for (int k = 0; k < 4; ++k)
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 4; ++c)
        R[k][r][c] = joints[k] == 0xff ? T(r == c ? 1 : 0) : (T)mats[12 * joints[k] + 4 * r + c];
  auto apply = [&](const float* x, T hw, float* o, float ow) {
    const T xv[4] = {(T)x[0], (T)x[1], (T)x[2], hw};
    T t[4][3];
    for (int k = 0; k < 4; ++k)
      for (int r = 0; r < 3; ++r)
        t[k][r] = std::fma(xv[3], R[k][r][3], std::fma(xv[2], R[k][r][2], std::fma(xv[1], R[k][r][1],
                                                                                   xv[0] * R[k][r][0])));
    for (int r = 0; r < 3; ++r)
      o[r] = (float)std::fma(w[3], t[3][r], std::fma(w[2], t[2][r], std::fma(w[1], t[1][r], w[0] * t[0][r])));
    o[3] = ow;
  };
  apply(rest, T(1), out, rest[3]);
  if (nout) apply(normal, T(0), nout, normal[3]);
}

}  // namespace

extern "C" {

float oracle_rsqrt_f32(float x) { return rsqrt_t(x); }

// Fills out[i] = rsqrt(in[i]) for a whole array (exhaustive accuracy checks).
void oracle_rsqrt_f32_array(const float* in, float* out, int64_t n) {
  for (int64_t i = 0; i < n; ++i) out[i] = rsqrt_t(in[i]);
}

// One step of every strand, in place.  pos/vel: float4 [nstrands * npoints];
// roots: float4 [nstrands] (root_mode 1) or NULL.
int oracle_step_f32(const oracle_params* P, int nstrands, float* pos, float* vel, const float* roots) {
  return step_all<float, float>(P, nstrands, pos, vel, roots, 1);
}
int oracle_step_f32_mt(const oracle_params* P, int nstrands, float* pos, float* vel,
                       const float* roots, int nthreads) {
  return step_all<float, float>(P, nstrands, pos, vel, roots, nthreads);
}
int oracle_step_f64(const oracle_params* P, int nstrands, double* pos, double* vel, const double* roots) {
  return step_all<double, double>(P, nstrands, pos, vel, roots, 1);
}

// Wind acceleration at one point (the value CalculateForces adds to the force).
void oracle_wind_f32(const oracle_params* P, const float* x, float* out) {
  const StepConsts<float> k = make_consts<float>(P);
  const V3<float> w = wind(k, {x[0], x[1], x[2]});
  out[0] = w.x; out[1] = w.y; out[2] = w.z;
}

// Dual-quaternion blended positions.  dq: 8 floats per bone (real xyzw,
// dual xyzw, the glm::dualquat memory order); joints: int4 per vertex;
// weights: float4 per vertex (w ignored and derived); rest/out: float4.
void oracle_skin_f32(int nverts, const float* dq, const int* joints, const float* weights,
                     const float* rest, float* out) {
  for (int v = 0; v < nverts; ++v)
    skin_vertex<float>(dq, joints + 4 * v, weights + 4 * v, rest + 4 * v, out + 4 * v);
}
void oracle_skin_f64(int nverts, const float* dq, const int* joints, const float* weights,
                     const float* rest, float* out_as_f32) {
  for (int v = 0; v < nverts; ++v)
    skin_vertex<double>(dq, joints + 4 * v, weights + 4 * v, rest + 4 * v, out_as_f32 + 4 * v);
}

// Positions and normals (float4 each) of the dual-quaternion (mode 0) or
// linear-blend (mode 1, bones as 3x4 row matrices, 12 floats each) skinning.
// FP64 variant: the same formulas in double, rounded to float at the end.
void oracle_skin_full_f32(int mode, int nverts, const float* bones, const int* joints, const float* weights,
                          const float* rest, const float* normals, float* out, float* nout) {
  for (int v = 0; v < nverts; ++v) {
    if (mode == 0)
      skin_vertex<float>(bones, joints + 4 * v, weights + 4 * v, rest + 4 * v, out + 4 * v, normals + 4 * v,
                         nout + 4 * v);
    else
      lbs_vertex<float>(bones, joints + 4 * v, weights + 4 * v, rest + 4 * v, out + 4 * v, normals + 4 * v,
                        nout + 4 * v);
  }
}
void oracle_skin_full_f64(int mode, int nverts, const float* bones, const int* joints, const float* weights,
                          const float* rest, const float* normals, float* out, float* nout) {
  for (int v = 0; v < nverts; ++v) {
    if (mode == 0)
      skin_vertex<double>(bones, joints + 4 * v, weights + 4 * v, rest + 4 * v, out + 4 * v, normals + 4 * v,
                          nout + 4 * v);
    else
      lbs_vertex<double>(bones, joints + 4 * v, weights + 4 * v, rest + 4 * v, out + 4 * v, normals + 4 * v,
                         nout + 4 * v);
  }
}

// Root frame of each strand on its bound triangle: the root as
// oracle_bary_roots_f32, and the normal = normalize(w0 nA + w1 nB + w2 nC)
// with the same fma order and normalize(d) = d * rsqrt(dot(d, d)).
void oracle_root_frames_f32(int nstrands, const int* tris, const int* bind_tri, const float* bind_w,
                            const float* verts, const float* normals, float* out_roots, float* out_normals) {
  for (int s = 0; s < nstrands; ++s) {
    const int t = bind_tri[s];
    const float w1 = bind_w[2 * s], w2 = bind_w[2 * s + 1];
    const float w0 = (1.0f - w1) - w2;
    const int ia = tris[3 * t + 0], ib = tris[3 * t + 1], ic = tris[3 * t + 2];
    float n[3];
    for (int c = 0; c < 3; ++c) {
      out_roots[4 * s + c] = std::fma(w2, verts[4 * ic + c], std::fma(w1, verts[4 * ib + c], w0 * verts[4 * ia + c]));
      n[c] = std::fma(w2, normals[4 * ic + c], std::fma(w1, normals[4 * ib + c], w0 * normals[4 * ia + c]));
    }
    out_roots[4 * s + 3] = 0.0f;
    const float is = rsqrt_t(dot3<float>({n[0], n[1], n[2]}, {n[0], n[1], n[2]}));
    for (int c = 0; c < 3; ++c) out_normals[4 * s + c] = n[c] * is;
    out_normals[4 * s + 3] = 0.0f;
  }
}

// Root of each strand from its bound triangle: w0*A + w1*B + w2*C with
// w1, w2 the stored barycentrics and w0 = (1 - w1) - w2, evaluated as
// fma(w2, C, fma(w1, B, w0 * A)).  tris: int3 per triangle; bind: per strand
// {triangle index, w1, w2} as int + 2 floats; verts/out: float4.
void oracle_bary_roots_f32(int nstrands, const int* tris, const int* bind_tri,
                           const float* bind_w, const float* verts, float* out) {
  for (int s = 0; s < nstrands; ++s) {
    const int t = bind_tri[s];
    const float w1 = bind_w[2 * s], w2 = bind_w[2 * s + 1];
    const float w0 = (1.0f - w1) - w2;
    const float* A = verts + 4 * tris[3 * t + 0];
    const float* B = verts + 4 * tris[3 * t + 1];
    const float* C = verts + 4 * tris[3 * t + 2];
    for (int c = 0; c < 3; ++c) out[4 * s + c] = std::fma(w2, C[c], std::fma(w1, B[c], w0 * A[c]));
    out[4 * s + 3] = 0.0f;
  }
}

int oracle_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Render-side strand expansion (the tessellation stream pass).
//
// Patches: scalp triangle t (three strand indices), strand segment j; control
// points P[2k], P[2k+1] = positions of points j, j+1 of the strand at vertex k
// (hair.cc:402-411), tangents likewise, scaled by uScaleFactor
// (tcs_stream_hair.glsl:41).  Each patch is drawn ninst times; the
// tessellator makes nlines isolines (v = l / nlines) of nsub segments
// (u = s / nsub, ..., (s+1) / nsub), and the geometry stage captures each
// segment as two vertices (x, y, z, relPos) (gs_stream_hair.glsl:18-28), in
// the order instance, triangle, segment, line, subsegment.

namespace {

// Index of the random entry of (instance, line): tes_stream_hair.glsl:50-53,
// int(gl_TessCoord.y * 40 + inInstanceID) % HAIR_TF_RANDOMBUFFER_SIZE.
int expand_random_index(int inst, int line, int nlines) {
  const float v = (float)line / (float)nlines;
  const float x = v * 40.0f;
  const float y = x + (float)inst;
  return ((int)y) % 4096;
}

// sample_triangle2's weights (inc_maths.glsl:107-115), statement by statement:
// the second statement reads the x the first one wrote.
template <typename T>
void expand_weights(T rx, T ry, T c[3]) {
  T x = rx, y = ry;
  if (x + y > T(1)) {
    x = std::max(x, y);
    y = std::min(x, y);
    x = T(1) - x;
  }
  c[0] = x;
  c[1] = y;
  c[2] = T(1) - (x + y);
}

}  // namespace

extern "C" {

// FP32, in the evaluation order of include/ftl_render.h (the kernel's).
// pos, tan: float4 [nstrands*npoints]; tris: int3 [ntris]; rnd: float4
// [nrandom]; out: float4 [2*ninst*ntris*(npoints-1)*nlines*nsub].  Returns
// -1 when a random index or a triangle's strand index is out of range.
int oracle_expand_f32(const float* pos, const float* tan, int nstrands, int npoints, const int* tris, int ntris,
                      const float* rnd, int nrandom, int ninst, int nlines, int nsub, float scale, float* out) {
  for (int i = 0; i < 3 * ntris; ++i)
    if (tris[i] < 0 || tris[i] >= nstrands) return -1;
  for (int inst = 0; inst < ninst; ++inst)
    for (int line = 0; line < nlines; ++line)
      if (expand_random_index(inst, line, nlines) >= nrandom) return -1;
  const int nseg = npoints - 1;
  const float n = (float)nsub, nf = (float)npoints;
  for (int inst = 0; inst < ninst; ++inst)
    for (int tri = 0; tri < ntris; ++tri)
      for (int seg = 0; seg < nseg; ++seg)
        for (int line = 0; line < nlines; ++line) {
          const float* r = rnd + 4 * expand_random_index(inst, line, nlines);
          float c[3];
          expand_weights<float>(r[0], r[1], c);
          for (int s = 0; s < nsub; ++s)
            for (int e = 0; e < 2; ++e) {
              const float u = (float)(s + e) / n;
              const float u2 = u * u, u3 = u2 * u;
              const float a = 2.0f * u3, b = 3.0f * u2;
              const float h0 = (a - b) + 1.0f, h1 = b - a, h2 = (u3 - 2.0f * u2) + u, h3 = u3 - u2;
              float g[3][3];
              for (int k = 0; k < 3; ++k) {
                const size_t base = (size_t)tris[3 * tri + k] * npoints + seg;
                const float* p0 = pos + 4 * base;
                const float* p1 = pos + 4 * (base + 1);
                const float* q0 = tan + 4 * base;
                const float* q1 = tan + 4 * (base + 1);
                for (int d = 0; d < 3; ++d) {
                  const float t0 = q0[d] * scale, t1 = q1[d] * scale;
                  g[k][d] = std::fma(h3, t1, std::fma(h2, t0, std::fma(h1, p1[d], h0 * p0[d])));
                }
              }
              const float rp = (float)seg / nf + u / nf;
              const float t = std::min(std::max(rp, 0.0f), 1.0f);
              const float w = ((t * t) * t) * (10.0f + t * (-15.0f + 6.0f * t));
              const long long id =
                  2 * (((((long long)inst * ntris + tri) * nseg + seg) * nlines + line) * nsub + s) + e;
              float* o = out + 4 * id;
              for (int d = 0; d < 3; ++d) o[d] = std::fma(c[2], g[2][d], std::fma(c[1], g[1][d], c[0] * g[0][d]));
              o[3] = w;
            }
        }
  return 0;
}

// FP64, written as the GLSL states it: hermite_mix as the row vector
// (u^3, u^2, u, 1) times mHermit times B (inc_maths.glsl:210-228), the blend
// as mat3x4 * coords, smoothstep2 through maprange (:239-241, :263-266).  An
// independent evaluation of the same formulas: the accuracy reference of the
// FP32 path.  Output in double, same layout.
int oracle_expand_f64(const float* pos, const float* tan, int nstrands, int npoints, const int* tris, int ntris,
                      const float* rnd, int nrandom, int ninst, int nlines, int nsub, float scale, double* out) {
  for (int i = 0; i < 3 * ntris; ++i)
    if (tris[i] < 0 || tris[i] >= nstrands) return -1;
  for (int inst = 0; inst < ninst; ++inst)
    for (int line = 0; line < nlines; ++line)
      if (expand_random_index(inst, line, nlines) >= nrandom) return -1;
  // mHermit, column-major as the GLSL constructor lists its columns.
  const double Mh[4][4] = {{2, -3, 0, 1}, {-2, 3, 0, 0}, {1, -2, 1, 0}, {1, -1, 0, 0}};
  const int nseg = npoints - 1;
  for (int inst = 0; inst < ninst; ++inst)
    for (int tri = 0; tri < ntris; ++tri)
      for (int seg = 0; seg < nseg; ++seg)
        for (int line = 0; line < nlines; ++line) {
          const float* r = rnd + 4 * expand_random_index(inst, line, nlines);
          double c[3];
          expand_weights<double>(r[0], r[1], c);
          for (int s = 0; s < nsub; ++s)
            for (int e = 0; e < 2; ++e) {
              const double u = (double)(s + e) / (double)nsub;
              const double vU[4] = {u * u * u, u * u, u, 1.0};
              double h[4];  // vU * mHermit: column i of the product = dot(vU, column i)
              for (int i = 0; i < 4; ++i) {
                h[i] = 0.0;
                for (int k = 0; k < 4; ++k) h[i] += vU[k] * Mh[i][k];
              }
              double g[3][4];
              for (int k = 0; k < 3; ++k) {
                const size_t base = (size_t)tris[3 * tri + k] * npoints + seg;
                const float* p0 = pos + 4 * base;
                const float* p1 = pos + 4 * (base + 1);
                const float* q0 = tan + 4 * base;
                const float* q1 = tan + 4 * (base + 1);
                // B's columns: (p0.d, p1.d, t0.d, t1.d) for d = x, y, z, and (1, 1, 0, 0).
                for (int d = 0; d < 3; ++d) {
                  const double col[4] = {p0[d], p1[d], (double)q0[d] * scale, (double)q1[d] * scale};
                  g[k][d] = 0.0;
                  for (int i = 0; i < 4; ++i) g[k][d] += h[i] * col[i];
                }
                g[k][3] = h[0] + h[1];
              }
              const double x = (double)seg / npoints + u / npoints;
              const double t = std::min(std::max((x - 0.0) / (1.0 - 0.0), 0.0), 1.0);
              const long long id =
                  2 * (((((long long)inst * ntris + tri) * nseg + seg) * nlines + line) * nsub + s) + e;
              double* o = out + 4 * id;
              for (int d = 0; d < 3; ++d) o[d] = g[0][d] * c[0] + g[1][d] * c[1] + g[2][d] * c[2];
              o[3] = t * t * t * (10.0 + t * (-15.0 + 6.0 * t));
            }
        }
  return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Groom initialisation (Hair::init_simulation, hair.cc:236-328).
//
// rand() is replaced by SplitMix64 (Steele, Lea, Flood 2014): the j-th draw of
// the stream s_{k+1} = s_k + 0x9e3779b97f4a7c15 (s_0 = seed), mixed by two
// xor-shift-multiply rounds, as a double in [0, 1) from its top 53 bits.

namespace {

uint64_t init_draw(uint64_t seed, uint64_t j) {
  uint64_t s = seed;
  uint64_t z = 0;
  for (uint64_t k = 0; k <= j; ++k) {  // sequential on purpose: the stream, not a jump
    s += 0x9e3779b97f4a7c15ull;
    z = s;
  }
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

// sin and cos of the curly tangents (hair.cc:300-322 calls std::sin/std::cos,
// whose last bits differ between libraries): one FP32 definition the device
// shares, from the mathematics only.  Cody and Waite's reduction by pi/2 in
// three parts (k = rint(x * 2/pi), r = x - k pi/2, one fma per part; the
// parts are pi/2 rounded to float, then the rounded remainders), the Taylor
// series of sin (to r^9) and cos (to r^10) on |r| <= pi/4 with coefficients
// (-1)^n/n! rounded to float, evaluated by Horner with fma, and the quadrant.
float trig_f32(float x, int cosine) {
  const float k = std::nearbyint(x * 0.636619747f);
  float r = std::fma(-k, 1.57079637f, x);
  r = std::fma(-k, -4.37113883e-08f, r);
  r = std::fma(-k, -1.71512451e-15f, r);
  const int q = (static_cast<int>(k) + cosine) & 3;
  const float r2 = r * r;
  float ps = std::fma(r2, 2.75573188e-06f, -1.98412701e-04f);
  ps = std::fma(r2, ps, 8.33333377e-03f);
  ps = std::fma(r2, ps, -1.66666672e-01f);
  const float s = std::fma(r2 * r, ps, r);
  float pc = std::fma(r2, -2.75573200e-07f, 2.48015876e-05f);
  pc = std::fma(r2, pc, -1.38888892e-03f);
  pc = std::fma(r2, pc, 4.16666679e-02f);
  pc = std::fma(r2, pc, -0.5f);
  const float c = std::fma(r2, pc, 1.0f);
  return q == 0 ? s : q == 1 ? c : q == 2 ? -s : -c;
}
float trig_f32_sin(float x) { return trig_f32(x, 0); }
float trig_f32_cos(float x) { return trig_f32(x, 1); }

// asin, acos, exp, log and pow of the Marschner tables (cs_marschner_*.glsl),
// one FP32 definition the device shares, from the mathematics only:
// * asin on [0, 1/2]: a + a^3 P(a^2), P the Taylor series of asin to a^23
//   (coefficients (2n)! / (4^n (n!)^2 (2n+1)) rounded to float, Horner with
//   fma); above 1/2, asin a = pi/2 - 2 asin(sqrt((1 - a)/2)); acos from the same
//   core (pi/2 - asin, 2 asin(...), pi - 2 asin(...)), pi and pi/2 in two parts;
// * exp: Cody-Waite by ln 2 in two parts, the Taylor series of e^r to r^8,
//   then the scaling by 2^k (ldexp, one rounding);
// * log: x = 2^e m with m in [sqrt(1/2), sqrt(2)), log m = 2 atanh(s) with
//   s = (m - 1)/(m + 1) (series to s^11), plus e ln 2 in two parts;
// * pow(x, y) = exp(y log x) for x >= 0 (pow(0, y) = 0: y > 0 here).
// Within 2 ulp (pow: 3 ulp) of float64 on the tables' ranges.
float asin_core_f32(float a) {
  const float z = a * a;
  float p = std::fma(z, 0.00731252600f, 0.00839033537f);
  p = std::fma(z, p, 0.00976160914f);
  p = std::fma(z, p, 0.0115518011f);
  p = std::fma(z, p, 0.0139648439f);
  p = std::fma(z, p, 0.0173527636f);
  p = std::fma(z, p, 0.0223721582f);
  p = std::fma(z, p, 0.0303819440f);
  p = std::fma(z, p, 0.0446428582f);
  p = std::fma(z, p, 0.0750000030f);
  p = std::fma(z, p, 0.166666672f);
  return std::fma(a * z, p, a);
}
float asin_f32(float x) {
  const float a = std::fabs(x);
  float r;
  if (a <= 0.5f) {
    r = asin_core_f32(a);
  } else {
    const float s = std::sqrt((1.0f - a) * 0.5f);
    r = std::fma(-2.0f, asin_core_f32(s), 1.57079637f) + -4.37113883e-08f;
  }
  return x < 0.0f ? -r : r;
}
float acos_f32(float x) {
  if (std::fabs(x) <= 0.5f)
    return (1.57079637f - asin_core_f32(std::fabs(x)) * (x < 0.0f ? -1.0f : 1.0f)) + -4.37113883e-08f;
  if (x > 0.0f) return 2.0f * asin_core_f32(std::sqrt((1.0f - x) * 0.5f));
  return std::fma(-2.0f, asin_core_f32(std::sqrt((1.0f + x) * 0.5f)), 3.14159274f) + -8.74227766e-08f;
}
float exp_f32(float x) {
  if (!(x < 89.0f)) return x != x ? x : INFINITY;
  if (x < -104.0f) return 0.0f;
  const float k = std::nearbyint(x * 1.44269502f);
  float r = std::fma(-k, 0.693147182f, x);
  r = std::fma(-k, -1.90465421e-09f, r);
  float p = std::fma(r, 2.48015876e-05f, 1.98412701e-04f);
  p = std::fma(r, p, 1.38888892e-03f);
  p = std::fma(r, p, 8.33333377e-03f);
  p = std::fma(r, p, 4.16666679e-02f);
  p = std::fma(r, p, 1.66666672e-01f);
  p = std::fma(r, p, 0.5f);
  p = std::fma(r, p, 1.0f);
  p = std::fma(r, p, 1.0f);
  return std::ldexp(p, static_cast<int>(k));
}
float log_f32(float x) {
  if (!(x > 0.0f) || x == INFINITY) return x == 0.0f ? -INFINITY : (x > 0.0f ? x : NAN);
  int e = 0;
  if (x < 1.17549435e-38f) {
    x *= 16777216.0f;
    e = -24;
  }
  uint32_t i;
  std::memcpy(&i, &x, 4);
  e += static_cast<int>((i >> 23) & 0xffu) - 127;
  i = (i & 0x007fffffu) | 0x3f800000u;
  float m;
  std::memcpy(&m, &i, 4);
  if (m > 1.41421354f) {
    m *= 0.5f;
    e += 1;
  }
  const float f = m - 1.0f;
  const float s = f / (2.0f + f);
  const float s2 = s * s;
  float p = std::fma(s2, 0.0909090936f, 0.111111112f);
  p = std::fma(s2, p, 0.142857149f);
  p = std::fma(s2, p, 0.200000003f);
  p = std::fma(s2, p, 0.333333343f);
  const float lm = std::fma(2.0f * s * s2, p, 2.0f * s);
  const float ef = static_cast<float>(e);
  return std::fma(ef, 0.693147182f, std::fma(ef, -1.90465421e-09f, lm));
}
float pow_f32(float x, float y) { return x == 0.0f ? 0.0f : exp_f32(y * log_f32(x)); }

// glm::simplex(vec2) (glm/gtc/noise.inl; the GLM tree is not part of the
// reference), from the published algorithm: McEwan, Sheets, Gustavson,
// Richardson, "Efficient computational noise in GLSL", JGT 17(2), 2012
// (webgl-noise, snoise(vec2)), in single precision, operations in the order
// the GLSL writes them.
float noise_mod289(float x) { return x - std::floor(x * (1.0f / 289.0f)) * 289.0f; }
float noise_permute(float x) { return noise_mod289(((x * 34.0f) + 1.0f) * x); }

float simplex2_f32(float vx, float vy) {
  const float C[4] = {0.211324865405187f, 0.366025403784439f, -0.577350269189626f, 0.024390243902439f};
  const float skew = vx * C[1] + vy * C[1];
  float i[2] = {std::floor(vx + skew), std::floor(vy + skew)};
  const float unskew = i[0] * C[0] + i[1] * C[0];
  const float x0[2] = {vx - i[0] + unskew, vy - i[1] + unskew};
  const float i1[2] = {x0[0] > x0[1] ? 1.0f : 0.0f, x0[0] > x0[1] ? 0.0f : 1.0f};
  const float x12[4] = {(x0[0] + C[0]) - i1[0], (x0[1] + C[0]) - i1[1], x0[0] + C[2], x0[1] + C[2]};
  for (int k = 0; k < 2; ++k) i[k] = i[k] - 289.0f * std::floor(i[k] / 289.0f);
  const float oy[3] = {0.0f, i1[1], 1.0f}, ox[3] = {0.0f, i1[0], 1.0f};
  float p[3], m[3];
  for (int k = 0; k < 3; ++k) p[k] = noise_permute(noise_permute(i[1] + oy[k]) + i[0] + ox[k]);
  const float d[3] = {x0[0] * x0[0] + x0[1] * x0[1], x12[0] * x12[0] + x12[1] * x12[1],
                      x12[2] * x12[2] + x12[3] * x12[3]};
  for (int k = 0; k < 3; ++k) {
    m[k] = std::max(0.5f - d[k], 0.0f);
    m[k] = m[k] * m[k];
  }
  for (int k = 0; k < 3; ++k) m[k] = m[k] * m[k];
  float g[3];
  const float gx[3] = {x0[0], x12[0], x12[2]}, gy[3] = {x0[1], x12[1], x12[3]};
  for (int k = 0; k < 3; ++k) {
    const float fr = p[k] * C[3];
    const float x = 2.0f * (fr - std::floor(fr)) - 1.0f;
    const float h = std::fabs(x) - 0.5f;
    const float a0 = x - std::floor(x + 0.5f);
    m[k] *= 1.79284291400159f - 0.85373472095314f * (a0 * a0 + h * h);
    g[k] = a0 * gx[k] + h * gy[k];
  }
  return 130.0f * (m[0] * g[0] + m[1] * g[1] + m[2] * g[2]);
}

}  // namespace

extern "C" {

// roots, normals: float3 [nroots]; pos, vel: float4 [nroots*npoints].
void oracle_reference_groom_f32(int nroots, const float* roots, const float* normals, int npoints, float maxlength,
                                uint64_t seed, float* pos, float* vel) {
  const float scaleOffset = maxlength / static_cast<float>(npoints);
  for (int j = 0; j < nroots; ++j) {
    const double u = (double)(init_draw(seed, (uint64_t)j) >> 11) * 0x1p-53;
    const float r = static_cast<float>(1.0 + 0.1 * (1.0 - 2.0 * u));
    float last = 0.0f;
    for (int i = 0; i < npoints; ++i) {
      const float offset = static_cast<float>(i) * scaleOffset * r;
      float* P = pos + 4 * ((size_t)j * npoints + i);
      for (int c = 0; c < 3; ++c) P[c] = roots[3 * j + c] + offset * normals[3 * j + c];
      P[3] = offset - last;
      float* V = vel + 4 * ((size_t)j * npoints + i);
      V[0] = V[1] = V[2] = V[3] = 0.0f;
      last = offset;
    }
  }
}

float oracle_simplex2_f32(float x, float y) { return simplex2_f32(x, y); }
float oracle_sin_f32(float x) { return trig_f32(x, 0); }
float oracle_cos_f32(float x) { return trig_f32(x, 1); }
float oracle_asin_f32(float x) { return asin_f32(x); }
float oracle_acos_f32(float x) { return acos_f32(x); }
float oracle_exp_f32(float x) { return exp_f32(x); }
float oracle_log_f32(float x) { return log_f32(x); }
float oracle_pow_f32(float x, float y) { return pow_f32(x, y); }

// The tangents of hair.cc:290-328 (curly or not), float4 [nstrands*npoints].
void oracle_reference_tangents_f32(int nstrands, int npoints, const float* normals, float maxlength, int curly,
                                   float* tan) {
  const float inv_nroots = 1.0f / static_cast<float>(nstrands);
  const float kPi = 3.14159265358979323846f;
  const float scaleMaxLength = 0.125f * std::sqrt(maxlength);
  for (int j = 0; j < nstrands; ++j) {
    const int A = j * npoints, B = A + (npoints - 1);
    const float dj = static_cast<float>(j + 1) * inv_nroots;
    const float* n = normals + 3 * j;
    float c[3] = {0.0f, 0.0f, 0.0f};
    if (curly) {
      const float k = 1.25f * simplex2_f32(trig_f32_sin(3.0f * dj), trig_f32_cos(5.0f));
      c[0] = trig_f32_cos(k * 4.0f * kPi);
      c[1] = -0.71f * k;
      c[2] = trig_f32_sin(k * 2.7f * kPi);
    }
    for (int d = 0; d < 3; ++d) {
      tan[4 * (size_t)A + d] = 0.15f * n[d];
      tan[4 * (size_t)B + d] = 0.2f * (-n[d] + c[d]);
    }
    tan[4 * (size_t)A + 3] = 0.0f;
    tan[4 * (size_t)B + 3] = 0.0f;
    const float dist = static_cast<float>(B - A);
    const float inv_dist = 1.0f / dist;
    for (int i = A + 1; i < B; ++i) {
      if (curly) {
        const float di = 10.0f * static_cast<float>(B - i) / (dist - 1.0f);
        const float k = di * simplex2_f32(trig_f32_sin(43.0f * dj), trig_f32_cos(5.0f * di));
        c[0] = -5.8f * (10.7f * trig_f32_cos(k * kPi));
        c[1] = -5.8f * (-2.3f * k);
        c[2] = -5.8f * (20.5f * trig_f32_sin(k * kPi));
      }
      const float s = 0.1f * static_cast<float>(i - A) * inv_dist * scaleMaxLength;
      for (int d = 0; d < 3; ++d) tan[4 * (size_t)i + d] = s * c[d];
      tan[4 * (size_t)i + 3] = s * 0.0f;
    }
  }
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Marschner lookup tables (cs_marschner_m.glsl, cs_marschner_n.glsl with
// inc_marschner_n.glsl, inc_solver.glsl, inc_fresnel.glsl), templated on the
// precision: float (the FP32 restatement) and double (the accuracy check).
// Integer powers are products; Pi() is the reference's 3.141564.

namespace {

// The transcendentals of the tables: the shared FP32 definitions above for
// the FP32 restatement, the C library's for the FP64 accuracy check.
float m_sin(float x) { return trig_f32(x, 0); }
float m_cos(float x) { return trig_f32(x, 1); }
float m_asin(float x) { return asin_f32(x); }
float m_acos(float x) { return acos_f32(x); }
float m_exp(float x) { return exp_f32(x); }
float m_pow(float x, float y) { return pow_f32(x, y); }
double m_sin(double x) { return std::sin(x); }
double m_cos(double x) { return std::cos(x); }
double m_asin(double x) { return std::asin(x); }
double m_acos(double x) { return std::acos(x); }
double m_exp(double x) { return std::exp(x); }
double m_pow(double x, double y) { return std::pow(x, y); }

template <typename T>
struct MRoots {
  T r[3];
  int n;
};

template <typename T>
MRoots<T> m_linear(T a, T b) {
  MRoots<T> c{{0, 0, 0}, 0};
  if (std::fabs(a) > T(1e-6f)) {
    c.r[0] = -b / a;
    c.n = 1;
  }
  return c;
}

template <typename T>
MRoots<T> m_quadratic(T a, T b, T c) {
  if (std::fabs(a) < T(1e-6f)) return m_linear(b, c);
  T delta = b * b - T(4) * a * c;
  MRoots<T> q{{0, 0, 0}, 0};
  if (delta < T(0)) return q;
  delta = std::sqrt(delta);
  q.r[0] = (-b + delta) / (T(2) * a);
  q.r[1] = (-b - delta) / (T(2) * a);
  q.n = 1 + (delta >= T(1e-6f) ? 1 : 0);
  return q;
}

template <typename T>
MRoots<T> m_cubic(T A, T B, T C, T Dd) {
  if (std::fabs(A) < T(1e-6f)) return m_quadratic(B, C, Dd);
  const T a = B / A, b = C / A, c = Dd / A;
  MRoots<T> r{{0, 0, 0}, 0};
  if (std::fabs(c) < T(1e-6f)) {
    r = m_quadratic(T(1), a, b);
    r.r[r.n] = T(0);
    r.n += 1;
    return r;
  }
  const T Q = (T(3) * b - a * a) / T(9);
  const T R = (T(9) * a * b - T(27) * c - T(2) * a * a * a) / T(54);
  const T Q3 = Q * Q * Q;
  const T D = Q3 + R * R;
  const T third = a / T(3);
  const T pi = T(3.141564f);
  if (D > T(0)) {
    const T sd = std::sqrt(D);
    const T u = R + sd, v = R - sd;
    const T s = (u > T(0) ? T(1) : (u < T(0) ? T(-1) : T(0))) * m_pow(std::fabs(u), T(0.333f));
    const T t = (v > T(0) ? T(1) : (v < T(0) ? T(-1) : T(0))) * m_pow(std::fabs(v), T(0.333f));
    r.r[0] = s + t - third;
    r.n = 1;
  } else {
    const T theta = m_acos(R * (T(1) / std::sqrt(-Q3)));
    const T tq = T(2) * std::sqrt(-Q);
    r.r[0] = std::fma(tq, m_cos(theta / T(3)), -third);
    r.r[1] = std::fma(tq, m_cos((theta + T(2) * pi) / T(3)), -third);
    r.r[2] = std::fma(tq, m_cos((theta + T(4) * pi) / T(3)), -third);
    r.n = 3;
  }
  return r;
}

template <typename T>
T m_power_ratio(T ratio, T nA, T nB, T cosI, T sinI) {
  const T e = ratio * sinI;
  const T s2 = e * e;
  if (s2 > T(1)) return T(1);
  const T cosT = std::sqrt(T(1) - s2);
  const T A = nA * cosI, B = nB * cosT;
  const T R = (A - B) / (A + B);
  return std::min(T(1), R * R);
}

template <typename T>
T m_fresnel(T e0, T eperp, T epar, T c, T s) {
  return m_power_ratio(e0 / eperp, e0, eperp, c, s) * (T(1) - T(0.5f)) + m_power_ratio(e0 / epar, epar, e0, c, s) * T(0.5f);
}

template <typename T>
T m_absorption(int p, T absorption, T eperp, T epar, T cosG, T sinG) {
  if (p == 0) return m_fresnel(T(1), eperp, epar, cosG, sinG);
  const T sinT = sinG / eperp;
  const T cosT = m_cos(m_asin(sinT));
  const T fi = m_fresnel(T(1), eperp, epar, cosG, sinG);
  const T ft = m_fresnel(T(1), T(1) / eperp, T(1) / epar, cosT, sinT);
  const T t = m_exp(T(-4) * absorption * (cosT * cosT));
  T a = (T(1) - fi) * (T(1) - fi);
  for (int k = 0; k < p - 1; ++k) a *= ft;
  for (int k = 0; k < p; ++k) a *= t;
  return a;
}

template <typename T>
T m_np(int p, T absorption, T eperp, T epar, T c, T phi) {
  const T A = T(p) * c * T(-0.25801227547f);
  const T C = T(p) * c * T(1.90985931710f) - T(2);
  const T D = T(p) * T(3.141564f) - phi;
  const MRoots<T> roots = m_cubic(A, T(0), C, D);
  T L = T(0);
  for (int i = 0; i < roots.n; ++i) {
    const T g = roots.r[i];
    const T a = m_absorption(p, absorption, eperp, epar, m_cos(g), m_sin(g));
    const T d1 = (T(3) * A * (g * g) + C) / m_cos(g);
    L += a * (T(1) / std::fabs(T(2) * d1));
  }
  return std::min(L, T(1));
}

template <typename T>
T m_gauss(T sigma, T xmu) {
  return m_exp(-(xmu * xmu) / (T(2) * sigma * sigma)) / (T(2.5066282f) * std::fabs(sigma));
}

template <typename T>
void m_luts(int res, float eta_f, float absorption_f, float ar_f, float br_f, T* M, T* N) {
  const T inv = T(1.0f / (float)res);
  const T eta = eta_f, absorption = absorption_f, ar = ar_f, br = br_f;
  for (int y = 0; y < res; ++y)
    for (int x = 0; x < res; ++x) {
      T* m = M + 4 * ((size_t)y * res + x);
      const T sI = T(2) * T(x) * inv - T(1), sR = T(2) * T(y) * inv - T(1);
      const T tI = m_asin(sI), tR = m_asin(sR);
      const T tH = (tI + tR) / T(2), tD = (tI - tR) / T(2);
      const T deg = tH * (T(180.0f) / T(3.14159265358979323846f));
      m[0] = m_gauss(T(1) * br, deg - T(1) * ar);
      m[1] = m_gauss(T(0.5f) * br, deg - T(-0.5f) * ar);
      m[2] = m_gauss(T(2) * br, deg - T(-1.5f) * ar);
      m[3] = m_cos(tD);
      T* n = N + 4 * ((size_t)y * res + x);
      const T cPhi = T(2) * T(x) * inv - T(1), cTh = T(2) * T(y) * inv - T(1);
      const T s2 = T(1) - cTh * cTh;
      const T eta2 = eta * eta;
      const T eperp = std::sqrt(eta2 - s2) / cTh;
      const T epar = eta2 / eperp;
      const T phi = m_acos(cPhi);
      const T c = m_asin(T(1) / eperp);
      n[0] = m_np(0, absorption, eperp, epar, c, phi);
      n[1] = m_np(1, absorption, eperp, epar, c, phi);
      n[2] = m_np(2, absorption, eperp, epar, c, phi);
      n[3] = T(1);
    }
}

}  // namespace

extern "C" {

void oracle_marschner_f32(int res, float eta, float absorption, float ar, float br, float* M, float* N) {
  m_luts<float>(res, eta, absorption, ar, br, M, N);
}
void oracle_marschner_f64(int res, float eta, float absorption, float ar, float br, double* M, double* N) {
  m_luts<double>(res, eta, absorption, ar, br, M, N);
}
// Np(p) at one (eta_perp, eta_par, c, phi), and the roots of the cubic
// A g^3 + C g + D (count returned; roots into r[3]), for the tests.
float oracle_marschner_np_f32(int p, float absorption, float eperp, float epar, float c, float phi) {
  return m_np<float>(p, absorption, eperp, epar, c, phi);
}
int oracle_cubic_f64(double A, double B, double C, double D, double* r) {
  const MRoots<double> m = m_cubic<double>(A, B, C, D);
  for (int i = 0; i < 3; ++i) r[i] = m.r[i];
  return m.n;
}
float oracle_fresnel_f32(float eperp, float epar, float c, float s) { return m_fresnel<float>(1.0f, eperp, epar, c, s); }

}  // extern "C"
