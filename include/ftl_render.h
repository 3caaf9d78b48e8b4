/*
 * ftl_render.h -- the render-side consumer of the simulated strands in
 * libftl.so: expansion of the guide strands into interpolated render strands
 * (line segments), on the device.
 *
 * It replaces the reference renderer's "tessellation stream" pass
 * (src/fx/hair.cc:144-170): an instanced draw of isoline patches, one patch per
 * (scalp triangle, strand segment) with the six control points of the three
 * guide strands rooted at the triangle's vertices (hair.cc:398-411), through
 *   vs_stream_hair.glsl:19-24   position, tangent, relative position j / N
 *   tcs_stream_hair.glsl:30-42  tangent * uScaleFactor, tess levels
 *                               (outer[0] = lines, outer[1] = segments per line)
 *   tes_stream_hair.glsl:27-72  Hermite segment of each guide at u, a seeded
 *                               barycentric blend of the three, smoothstep2
 *                               of the relative position
 *   gs_stream_hair.glsl:18-28   each line segment captured as two vertices
 *                               (xyz, relPos) by transform feedback
 * and the fixed 64 MiB capture buffer (hair.cc:424-427), whose overflow the
 * reference does not check: here the caller sizes the output with
 * ftl_expand_vertex_count and a short buffer is an error.
 *
 * The arithmetic is one fixed FP32 evaluation order (below), shared with the
 * CPU oracle; the GLSL leaves fma contraction to the driver.
 */
#ifndef FTL_RENDER_H_
#define FTL_RENDER_H_

#include <stdint.h>

#include "ftl.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Tessellation settings of the stream pass (Hair::Parameters_t::tess,
   src/fx/hair.h:32-36, and uScaleFactor = render.lengthScale). */
typedef struct ftl_expand_params {
  int ninstances;       /* instanced draws of every patch (uNumInstances) */
  int nlines;           /* isolines per patch, gl_TessLevelOuter[0] (uNumLines) */
  int nsubsegments;     /* segments per isoline, gl_TessLevelOuter[1] (uNumSubSegments) */
  float tangent_scale;  /* uScaleFactor: tangents are scaled before the Hermite segment */
} ftl_expand_params;

/* The reference defaults: 3 instances, 2 lines, 16 subsegments, scale 1. */
void ftl_expand_params_reference(ftl_expand_params* p);

typedef struct ftl_expand ftl_expand;

/* Builds the patch set of a scalp: tris = 3 strand indices per triangle (the
   strands rooted at the triangle's vertices, hair.cc:402-411), every index in
   [0, nstrands); tangents = float4 [nstrands*npoints] (xyz used; the
   reference's third simulation attribute, interop.h:15-18, which its step
   never rewrites); random = float4 [nrandom] (xy used; the std430 vec3 array
   RandomBuffer of tes_stream_hair.glsl:21-24). All are copied to the device. */
ftl_status ftl_expand_create(ftl_ctx* ctx, int nstrands, int npoints, const int* tris, int ntris,
                             const float* tangents, const float* random, int nrandom, ftl_expand** out);
void ftl_expand_destroy(ftl_expand* e);
/* Replaces the tangents (host float4 [nstrands*npoints]). */
ftl_status ftl_expand_set_tangents(ftl_expand* e, const float* tangents);
/* Copies the tangents to the host (float4 [nstrands*npoints]); waits for the device. */
ftl_status ftl_expand_get_tangents(ftl_expand* e, float* tangents);

/* The reference's initial tangents (Hair::init_simulation, hair.cc:290-328),
   computed on the device into e's tangents: normals = float3 per strand (the
   roots' normals), maxlength = params_.sim.maxlength.  Per strand j (dj =
   (j + 1) / nstrands): the root gets 0.15 n; the tip 0.2 (-n + c) with, when
   curly, c = (cos(4 pi k), -0.71 k, sin(2.7 pi k)), k = 1.25 simplex(sin 3dj,
   cos 5); inner point i gets s_i c_i with s_i = ((0.1 i) / (N-1)) * 0.125
   sqrt(maxlength) and, when curly, c_i = -5.8 (10.7 cos(pi k_i), -2.3 k_i,
   20.5 sin(pi k_i)), k_i = d_i simplex(sin 43dj, cos 5d_i), d_i = 10 (N-1-i) /
   (N-2) (0 otherwise).  simplex is GLM's 2-D simplex noise (glm/gtc/noise),
   restated from its published algorithm. */
ftl_status ftl_expand_reference_tangents(ftl_expand* e, const float* normals, float maxlength, int curly);

/* Vertices one run writes: 2 * ninstances * ntris * (npoints-1) * nlines *
   nsubsegments (two per line segment); -1 on invalid parameters. */
int64_t ftl_expand_vertex_count(const ftl_expand* e, const ftl_expand_params* p);

/* One stream pass over positions (device float4 [nstrands*npoints], e.g. a
   groom's buffer from ftl_groom_device_buffers) into out (device float4,
   capacity in vertices), asynchronous on stream.  Vertex pairs are in the
   capture order: instance, triangle, segment, line, subsegment; each vertex
   is (x, y, z, relPos).  Per patch (triangle t, segment j), line l and
   subsegment s, with P_k, T_k the points j, j+1 of guide k and T scaled:
     u = s / n (and (s+1) / n), IEEE division;
     a = 2 u^3, b = 3 u^2, h0 = (a - b) + 1, h1 = b - a, h2 = (u^3 - 2 u^2) + u,
     h3 = u^3 - u^2;  g_k = fma(h3, T_k1, fma(h2, T_k0, fma(h1, P_k1, h0 P_k0)));
     r = random[int((l / nlines) * 40 + instance) % 4096].xy, folded as the
     reference's sample_triangle2 (inc_maths.glsl:107-115) executes it:
     if r.x + r.y > 1 then r.x = 1 - max(r.x, r.y) (r.y unchanged);
     c = (r.x, r.y, 1 - (r.x + r.y));  xyz = fma(c2, g_2, fma(c1, g_1, c0 g_0));
     relPos = smoothstep2(j / N + u / N) (inc_maths.glsl:263-266),
     = ((t t) t) (10 + t (-15 + 6 t)) with t = clamp(j / N + u / N, 0, 1).
   Errors: a short output buffer, or a random index >= nrandom. */
ftl_status ftl_expand_run(ftl_expand* e, const ftl_expand_params* p, const float* positions, float* out,
                          int64_t out_capacity, void* stream);

/* -- Marschner lookup tables (src/fx/marschner.{h,cc}) ----------------------- */

/* Marschner::ShadingParameters_t (marschner.h:31-46): fibre (eta,
   absorption, eccentricity), surface (ar = longitudinal shift, br = width,
   degrees) and glint settings. */
typedef struct ftl_marschner_params {
  float eta, absorption, eccentricity;
  float ar, br;
  float glint_scale, azimuthal_width, delta_caustic, delta_hm;
} ftl_marschner_params;

/* The reference defaults: 1.55, 0.20, 0.85, -5, 5, 0.5, 10, 0.2, 0.5. */
void ftl_marschner_params_reference(ftl_marschner_params* p);

/* Marschner::generate (marschner.cc:35-66): the two res x res tables into
   out_m and out_n (device float4 [res*res] each, texel (x, y) at y*res + x),
   asynchronous on stream:
     M (cs_marschner_m.glsl:33-62) over sin theta_i = 2x/res - 1, sin theta_r =
       2y/res - 1: the R, TT, TRT longitudinal gaussians of the half angle
       (shifts (1, -0.5, -1.5) ar, widths (1, 0.5, 2) br) and cos theta_d;
     N (cs_marschner_n.glsl:32-84, the generic branch) over cos phi_d = 2x/res
       - 1, cos theta_d = 2y/res - 1: Np(0), Np(1), Np(2) of
       inc_marschner_n.glsl:97-131 (cubic roots of inc_solver.glsl, Fresnel of
       inc_fresnel.glsl), and 1.
   FP32 texels (the reference stores RGBA16F).  Integer powers are products
   (GLSL's pow is undefined for negative bases, which the reference reaches
   where cos theta_d < 0); Pi is the reference's 3.141564. */
ftl_status ftl_marschner_luts(ftl_ctx* ctx, const ftl_marschner_params* p, int res, float* out_m, float* out_n,
                              void* stream);

#ifdef __cplusplus
}
#endif

#endif /* FTL_RENDER_H_ */
