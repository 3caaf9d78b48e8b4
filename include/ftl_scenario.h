/*
 * ftl_scenario.h -- seeded generators of groom state, skinned meshes and bone
 * motion for libftl.so (CPU only; no device is touched).
 *
 * The reference builds its initial state on the host from a scalp mesh with
 * an unseeded rand() (Hair::init_simulation, src/fx/hair.cc:236-288).  These
 * generators produce the same layout (include/ftl.h) deterministically from a
 * seed.
 */
#ifndef FTL_SCENARIO_H_
#define FTL_SCENARIO_H_

#include <stdint.h>

#include "ftl.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Strands rooted on a sphere: roots uniform on the sphere (seeded), direction
   = normalize(normal + jitter * N(0,1)^3), `length` split into npoints-1 equal
   segments, zero velocity.  coherent = 1 orders the strands along a Morton
   curve of their roots (spatially coherent, as the vertex order of a scalp
   mesh is); 0 keeps generation order. */
typedef struct ftl_scn_sphere {
  int nstrands;
  int npoints;
  float center[3];
  float radius;
  float length;
  float jitter;
  uint64_t seed;
  int coherent;
} ftl_scn_sphere;

/* pos, vel: float4 [nstrands*npoints]; rest_roots: float4 [nstrands] or NULL. */
ftl_status ftl_scn_sphere_groom(const ftl_scn_sphere* d, float* pos, float* vel, float* rest_roots);

/* Hair::init_simulation (hair.cc:258-288) with rand() replaced by a seeded
   stream: for root j with normal n_j and a per-strand factor
   r_j = 1 + 0.1 (1 - 2 U), offset_i = i * (maxlength / npoints) * r_j,
   pos_i = (root_j + offset_i n_j, offset_i - offset_{i-1}), vel = 0.
   roots, normals: float3 [nroots]. */
ftl_status ftl_scn_reference_groom(int nroots, const float* roots, const float* normals, int npoints,
                                   float maxlength, uint64_t seed, float* pos, float* vel);

/* Skinned emitter: a latitude/longitude sphere mesh of about `ntris` triangles,
   `nbones` bones at seeded pivots inside it, 4 influences per vertex (nearest
   pivots, inverse-square weights, largest first), and `nstrands` strands bound
   to area-sampled triangles with seeded barycentrics. */
typedef struct ftl_scn_skinned {
  int nstrands;
  int npoints;
  int ntris;
  int nbones;
  float radius;
  float length;
  float jitter;
  float max_angle;   /* radians: bone rotations are max_angle * sin(w t + phi) */
  uint64_t seed;
} ftl_scn_skinned;

ftl_status ftl_scn_skinned_sizes(const ftl_scn_skinned* d, int* nverts, int* ntris);
/* rest: float4 [nverts]; joints: int4 [nverts]; weights: float4 [nverts];
   tris: int3 [ntris]. */
ftl_status ftl_scn_skinned_mesh(const ftl_scn_skinned* d, float* rest, int* joints, float* weights, int* tris);
/* Strands bound to the mesh: tri [nstrands], bary float2 [nstrands];
   pos, vel: float4 [nstrands*npoints] hanging from the rest-pose roots. */
ftl_status ftl_scn_skinned_groom(const ftl_scn_skinned* d, const float* rest, const int* tris, int* tri,
                                 float* bary, float* pos, float* vel);
/* Bone dual quaternions at time t: 8 floats per bone (real xyzw, dual xyzw),
   each a rotation of max_angle * sin(w_b t + phi_b) about a seeded axis
   through the bone's pivot. */
ftl_status ftl_scn_bones(const ftl_scn_skinned* d, double t, float* dq);

/* Per-point tangents of strands (float4, w = 0) by central differences:
   t_i = (p_{i+1} - p_{i-1}) / 2, one-sided at the ends.  A stand-in for the
   reference's "curly" initial tangents (hair.cc:290-328), which need GLM's
   simplex noise (not part of the reference tree). */
ftl_status ftl_scn_strand_tangents(const float* pos, int nstrands, int npoints, float* tangents);

/* The renderer's random table (RandomBuffer::generate_values,
   src/memory/random_buffer.cc:22-36: uniform [0, 1) values) from a seed
   instead of std::random_device: float4 [n], xyz uniform [0, 1), w = 0. */
ftl_status ftl_scn_random_table(uint64_t seed, int n, float* out);

#ifdef __cplusplus
}
#endif

#endif /* FTL_SCENARIO_H_ */
