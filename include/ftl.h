/*
 * ftl.h -- C ABI of the MI355X FTL hair-strand step (libftl.so).
 *
 * Drop-in for the simulation half of tcoppex/barbu's `class Hair`
 * (src/fx/hair.h:24-121): the GLSL pass src/shaders/hair/01_simulation/
 * cs_simulation.glsl driven by Hair::update (src/fx/hair.cc:89-125).
 *
 * Plain pointers and sizes only.  A `stream` argument is a hipStream_t passed
 * as void* (NULL = the null stream).  Every call returns an ftl_status; on
 * failure ftl_last_error() holds a message for the calling thread.  Nothing
 * here calls exit(): the reference's fatal GL error path (src/core/
 * graphics.cc:565-574) becomes FTL_ERR_DEVICE.
 *
 * Data layout (kept from src/shaders/hair/interop.h:13-30 and
 * src/fx/hair.cc:265-288): strand-major arrays of float4,
 *   pos[s*npoints + i] = (x, y, z, rest length of the segment ending at i)
 *                        (w = 0 at the root, i = 0)
 *   vel[s*npoints + i] = (x, y, z, 0)
 * One handle's buffers are updated in place: there is no read/write pair and
 * no copy-back (the reference's PingPongBuffer::swap, src/memory/
 * pingpong_buffer.cc:73-84, has no counterpart).
 *
 * Threading: one host thread per handle at a time (as the reference, which runs
 * on the GL thread).  Handles of one context may be used from different threads
 * if each handle has a single user.
 */
#ifndef FTL_H_
#define FTL_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FTL_ABI_VERSION 2
#define FTL_MAX_COLLIDERS 4
#define FTL_MAX_POINTS 64
#define FTL_WIND_GRID 8    /* side of the periodic wind lattice */

typedef int ftl_status;
enum {
  FTL_OK = 0,
  FTL_ERR_INVALID_ARGUMENT = 1,
  FTL_ERR_DEVICE = 2,          /* HIP runtime error (launch, copy, allocation state) */
  FTL_ERR_OUT_OF_MEMORY = 3,   /* device allocation failed */
  FTL_ERR_BUDGET = 4,          /* server admission refused: HBM budget */
  FTL_ERR_NOT_INITIALIZED = 5, /* handle has no state (Hair::initialized() == false) */
  FTL_ERR_NONFINITE = 6,       /* a step produced NaN/Inf; see ftl_groom_status */
  FTL_ERR_UNSUPPORTED = 7,
  FTL_ERR_BUSY = 8,            /* server: the creature already has a step in flight */
  FTL_ERR_CANCELLED = 9        /* server: the request was dropped unstepped (its creature was removed, or
                                  could not be made resident) */
};

/* Velocity stored after a step. */
enum {
  /* cs_simulation.glsl as shipped: the stored velocity is the damped FTL
     correction of the last constraint iteration (Eq. 9 is commented out,
     :200-204). */
  FTL_VELOCITY_SHIPPED = 0,
  /* DFTL Eq. 9 with the clamp of VelocityControl (:79-84, :201-204):
     v = ((p - x) - u) / dt, |v| <= max_velocity. */
  FTL_VELOCITY_DFTL = 1
};

/* Where the root (i = 0) of each strand comes from at the start of a step. */
enum {
  /* p0 = M * x0, x0 = the root's current position (cs_simulation.glsl:190-191;
     the shipped shader fixes M = identity). */
  FTL_ROOT_MODEL = 0,
  /* p0 = M * r_s, r_s = rest root of strand s (the "object space positions kept
     in another buffer" that the comment at :186-188 asks for). */
  FTL_ROOT_REST = 1,
  /* p0 = barycentric point of a triangle of a dual-quaternion skinned mesh
     (the "post skinning" root of :188; blending from inc_skinning.glsl:54-82). */
  FTL_ROOT_SKINNED = 2
};

/* Simulation parameters (uniforms and hard-coded constants of
   cs_simulation.glsl; Hair::Parameters_t, src/fx/hair.h:27-51). */
typedef struct ftl_params {
  float gravity[3];       /* (0,-9.81,0)                                 :71 */
  float force_scale;      /* kForceCoeff, 20 in the shipped shader        :72 */
  float damping;          /* FTL damping delta; 0.80 as shipped (lane 0's
                             value of the curve at :100-102)               */
  float length_scale;     /* uScaleFactor = render.lengthScale (1.45)     :41, hair.cc:108 */
  int iterations;         /* kNumContraintsIteration = 8                  :197 */
  int velocity_mode;      /* FTL_VELOCITY_*                                    */
  float max_velocity;     /* uMaxVelocity = 100 (DFTL mode only)          :40 */
  float wind_amplitude;   /* m/s^2; 0 disables wind (the shipped shader has none) */
  float wind_frequency;   /* lattice cells per metre of the value-noise field  */
  float wind_scroll[3];   /* lattice cells per second the field drifts by       */
  uint64_t wind_seed;     /* seed of the lattice                                */
} ftl_params;

/* Initial state of a groom.  Copied at creation; the caller keeps ownership. */
typedef struct ftl_groom_desc {
  int nstrands;                 /* > 0 */
  int npoints;                  /* control points per strand, 2..FTL_MAX_POINTS */
  const float* positions;       /* float4 [nstrands*npoints] (layout above)      */
  const float* velocities;      /* float4 [nstrands*npoints], or NULL for zero   */
  const float* rest_roots;      /* float4 [nstrands] for FTL_ROOT_REST, or NULL  */
} ftl_groom_desc;

typedef struct ftl_ctx ftl_ctx;
typedef struct ftl_groom ftl_groom;
typedef struct ftl_batch ftl_batch;
typedef struct ftl_skin ftl_skin;

/* -- library / context ---------------------------------------------------- */

/* Message of the last failed call on this thread ("" if none). */
const char* ftl_last_error(void);
int ftl_abi_version(void);

/* Fills *p with the shipped shader's constants: gravity (0,-9.81,0), force
   scale 20, damping 0.80, length scale 1.45 (Hair::Parameters_t default),
   8 iterations, shipped velocity mode, max velocity 100, no wind. */
void ftl_params_reference(ftl_params* p);

/* Binds a HIP device.  ≙ Hair::init (hair.cc:26-40): allocates the shared
   resources (no program compile is needed: kernels are built ahead of time). */
ftl_status ftl_init(int device, ftl_ctx** out);
/* Frees the context.  Grooms, batches and skins must be destroyed first. */
ftl_status ftl_shutdown(ftl_ctx* ctx);
/* Device memory the context's objects hold, in bytes. */
ftl_status ftl_ctx_memory(ftl_ctx* ctx, uint64_t* bytes);

/* -- one groom (≙ one Hair) ------------------------------------------------ */

/* ≙ Hair::setup (hair.cc:42-64) + init_simulation upload (:330-345), with the
   initial state supplied by the caller.  Parameters start at
   ftl_params_reference, root mode FTL_ROOT_MODEL with M = identity, and one
   collider: the shader's default uBoundingSphere (0,0,0,1) (:43). */
ftl_status ftl_groom_create(ftl_ctx* ctx, const ftl_groom_desc* desc, ftl_groom** out);
/* ≙ Hair::setup + init_simulation's positions and velocities
   (hair.cc:236-288), computed on the device: strand j rooted at roots[j]
   (float3) along normals[j] (float3), offset_i = ((float)i * (maxlength /
   npoints)) * r_j with the per-strand factor r_j = (float)(1 + 0.1 (1 - 2 U_j))
   from the seed's SplitMix64 stream in place of rand(); w = offset_i -
   offset_{i-1}; zero velocities.  Bit-identical to ftl_scn_reference_groom
   (ftl_scenario.h), without the host-side state or its upload. */
ftl_status ftl_groom_create_reference(ftl_ctx* ctx, int nroots, const float* roots, const float* normals,
                                      int npoints, float maxlength, uint64_t seed, ftl_groom** out);
/* ≙ Hair::deinit (hair.cc:66-87).  NULL is ignored. */
void ftl_groom_destroy(ftl_groom* g);

ftl_status ftl_groom_set_params(ftl_groom* g, const ftl_params* p);
ftl_status ftl_groom_get_params(ftl_groom* g, ftl_params* p);

/* Sphere colliders applied in order at every constraint iteration: float4
   (centre xyz, radius) x count, count <= FTL_MAX_COLLIDERS.  count = 1
   ≙ Hair::set_bounding_sphere (hair.h:72-74) / uBoundingSphere. */
ftl_status ftl_groom_set_colliders(ftl_groom* g, const float* spheres, int count);

/* Root transform M, column-major 4x4 (glm::mat4 / uModel layout; the last row
   must be (0,0,0,1)).  Used by FTL_ROOT_MODEL and FTL_ROOT_REST. */
ftl_status ftl_groom_set_root_transform(ftl_groom* g, const float m[16]);
ftl_status ftl_groom_set_root_mode(ftl_groom* g, int mode);
/* Replaces the rest roots (float4 [nstrands]). */
ftl_status ftl_groom_set_rest_roots(ftl_groom* g, const float* roots);

/* Wind-field time: the lattice is offset by wind_scroll * time.  Set before a
   step; a plain step does not advance it. */
ftl_status ftl_groom_set_time(ftl_groom* g, double seconds);

/* One simulation step of dt seconds, enqueued on `stream`.  ≙ Hair::update
   (hair.cc:89-125), without the copy-back.  Asynchronous. */
ftl_status ftl_groom_step(ftl_groom* g, float dt, void* stream);
/* ftl_groom_step with its kernel bracketed by two caller-created hipEvent_t
   (passed as void*) that take the kernel's own start and end timestamps
   (hipExtLaunchKernelGGL): hipEventElapsedTime(start, stop) is the launch's
   duration without host or queue gaps.  Diagnostics; no counterpart in the
   reference. */
ftl_status ftl_groom_step_timed(ftl_groom* g, float dt, void* stream, void* start_event, void* stop_event);

/* Device pointers of the state (float4 arrays above), for zero-copy consumers
   such as a renderer (the reference's VAO reads them, hair.cc:363-391). */
ftl_status ftl_groom_device_buffers(ftl_groom* g, void** pos, void** vel);
/* Synchronous copies of the whole state.  Either pointer may be NULL. */
ftl_status ftl_groom_download(ftl_groom* g, float* pos, float* vel);
ftl_status ftl_groom_upload(ftl_groom* g, const float* pos, const float* vel);
ftl_status ftl_groom_info(ftl_groom* g, int* nstrands, int* npoints);
/* Reads, then clears, the groom's non-finite flag (set by a step whose output
   holds NaN/Inf at a strand tip).  Synchronises the context's device. */
ftl_status ftl_groom_status(ftl_groom* g, int* nonfinite);

/* Launch shape of plain steps: lanes per strand (1, 2 or 4; 0 = automatic).
   A point count that no class of that shape holds exactly runs a guarded
   class: one lane per strand if 1 was asked for, else the pipelined guarded
   class with the fewest points per lane.  Results do not depend on it. */
ftl_status ftl_groom_set_lanes(ftl_groom* g, int lanes_per_strand);

/* -- dual-quaternion skinned roots (FTL_ROOT_SKINNED) ----------------------- */

typedef struct ftl_skin_desc {
  int nverts;
  int ntris;
  int nbones;
  const float* rest_positions;  /* float4 [nverts] bind-pose positions          */
  const int* joints;            /* int4 [nverts] bone indices                   */
  const float* weights;         /* float4 [nverts]; w is derived as 1-(x+y+z)   */
  const int* triangles;         /* int3 [ntris] vertex indices                  */
} ftl_skin_desc;

/* Blending of the skinning stage: the uSkinning subroutine of
   inc_skinning.glsl:15-16 (SkinComponent's SkinningMode, skin.cc:63-74). */
enum {
  FTL_SKIN_DQ = 0,  /* dual-quaternion blending, skinning_DQBS (:54-82) */
  FTL_SKIN_LBS = 1  /* linear blending of 3x4 joint matrices, skinning_LBS (:85-129) */
};
/* Joint index meaning "no joint": the identity matrix in linear blending
   (kNoJoint, inc_skinning.glsl:18,89-93). */
#define FTL_NO_JOINT 0xff

/* Joint indices must be in [0, nbones), or FTL_NO_JOINT (linear blending
   only).  A new skin blends dual quaternions, with identity bones. */
ftl_status ftl_skin_create(ftl_ctx* ctx, const ftl_skin_desc* desc, ftl_skin** out);
void ftl_skin_destroy(ftl_skin* s);
/* FTL_SKIN_DQ or FTL_SKIN_LBS (dual quaternions are refused while a joint
   index is FTL_NO_JOINT). */
ftl_status ftl_skin_set_mode(ftl_skin* s, int mode);
/* Bone dual quaternions for the next skinning: 8 floats per bone, real part
   xyzw then dual part xyzw (glm::dualquat memory order, uploaded as a float4
   texture buffer by SkinComponent::updateSkinningBuffer, skin.cc:46-84). */
ftl_status ftl_skin_set_bones(ftl_skin* s, const float* dq, int nbones);
/* Bone skinning matrices for the next skinning, in the layout the reference
   uploads for linear blending: 12 floats per bone, the three rows of the
   affine skinning matrix (glm::mat3x4(transpose(skin_matrix)),
   skeleton_controller.cc:251-256).  Also sets the bones' dual quaternions,
   converted from the same matrices (glm::dualquat(mat3x4),
   skeleton_controller.cc:259-264), so either mode can be used. */
ftl_status ftl_skin_set_matrices(ftl_skin* s, const float* m3x4, int nbones);
/* SkeletonController::generate_skinning_datas (skeleton_controller.cc:248-265):
   skin_i = global_pose_i * inverse_bind_i (column-major 4x4 each, glm::mat4
   memory order; inverse_bind may be NULL for identities), then as
   ftl_skin_set_matrices. */
ftl_status ftl_skin_set_joint_matrices(ftl_skin* s, const float* global_pose, const float* inverse_bind,
                                       int nbones);
/* Rest normals (float4 [nverts]) skinned with the positions from now on:
   rotated by the blended real part (:80-81), or blended as (n, 0) by the
   joint matrices (:128).  Not renormalised, as in the reference. */
ftl_status ftl_skin_set_rest_normals(ftl_skin* s, const float* normals);
/* Skins every vertex (and normal) on `stream` (inc_skinning.glsl:22-31 and
   the mode's subroutine). */
ftl_status ftl_skin_update(ftl_skin* s, void* stream);
/* Synchronous copy of the skinned vertices (float4 [nverts]). */
ftl_status ftl_skin_download(ftl_skin* s, float* verts);
/* Synchronous copy of the skinned normals (float4 [nverts]); needs rest normals. */
ftl_status ftl_skin_download_normals(ftl_skin* s, float* normals);
/* The dual quaternion of a rigid 3x4 transform (12 floats: rows of [R | t]):
   real part q = the rotation's quaternion, dual part = 0.5 (t, 0) q; 8
   floats as ftl_skin_set_bones.  Host only.  glm::dualquat(mat3x4). */
ftl_status ftl_dualquat_from_matrix(const float* m3x4, float* dq);
/* Binds each strand's root to a triangle: tri[s] in [0,ntris), bary[2s..2s+1]
   = (w1, w2) with w0 = 1 - w1 - w2.  Sets the groom's root mode to
   FTL_ROOT_SKINNED.  Destroying the skin unbinds its grooms: their next step
   with skinned roots fails with FTL_ERR_INVALID_ARGUMENT until they are
   bound again or take another root mode. */
ftl_status ftl_groom_bind_skin(ftl_groom* g, ftl_skin* s, const int* tri, const float* bary);
/* The emitter frame of each strand of a skinned groom at the skin's current
   pose: its root on the bound triangle (the root the next step uses) and,
   when the skin has normals, the unit normal there, normalize(w0 nA + w1 nB
   + w2 nC) (float4 [nstrands] each, w = 0; normals may be NULL).  These are
   the roots and normals Hair::init_simulation grows strands from
   (hair.cc:258-288; ftl_groom_create_reference takes them), i.e. the "post
   skinning" roots cs_simulation.glsl:186-188 asks for.  Synchronous. */
ftl_status ftl_groom_root_frames(ftl_groom* g, float* roots, float* normals);

/* -- batched steps --------------------------------------------------------- */

/* Steps n grooms of one context as one batch: the grooms are cut into 64-lane
   tiles, and one kernel per launch class present steps the tiles of its
   grooms (the classes run side by side).  dt: one value per groom.  For repeated frames over the same
   grooms, ftl_batch_create / ftl_batch_launch avoid re-planning. */
ftl_status ftl_batch_step(ftl_ctx* ctx, ftl_groom* const* grooms, int n, const float* dt,
                          void* stream);

/* Device scratch of ftl_batch_step_noalloc: per GPU, held until the process
   ends.  It holds the batch's descriptors (288 B per groom, rounded up to
   256 B) and its span list (16 B per groom): about 13,800 grooms. */
#define FTL_NOALLOC_SCRATCH_BYTES (4u << 20)
/* ftl_batch_step without any allocation, for callers that cannot allocate
   (code under HIP graph capture of their own, tight serving loops, harnesses
   that forbid hipMalloc).  No counterpart in the reference.  The descriptors
   and spans go to a fixed device array instead of context-owned buffers;
   batches that do not fit it (FTL_NOALLOC_SCRATCH_BYTES) are refused with
   FTL_ERR_INVALID_ARGUMENT, and a groom whose wind lattice is not on the
   device yet (it is placed by a first ftl_groom_step or ftl_batch_step) is
   refused too.  Results are identical to ftl_batch_step.  Every call on one
   GPU shares the array: a call waits (on the host) for the previous call's
   kernels before it overwrites the array, and calls made concurrently from
   several host threads on one GPU are serialised; the array is not meant to
   be read by anything else. */
ftl_status ftl_batch_step_noalloc(ftl_ctx* ctx, ftl_groom* const* grooms, int n, const float* dt, void* stream);

/* A fixed set of grooms stepped as one frame, captured once into a HIP graph
   (parameter upload + one kernel per launch class) and replayed by
   ftl_batch_launch. */
ftl_status ftl_batch_create(ftl_ctx* ctx, ftl_groom* const* grooms, int n, ftl_batch** out);
void ftl_batch_destroy(ftl_batch* b);
/* One frame: per-groom dt (n values).  Colliders, transforms, params and time
   are read from the grooms at call time.  Asynchronous on `stream`.  Refused
   with FTL_ERR_INVALID_ARGUMENT once one of the batch's grooms has been
   destroyed. */
ftl_status ftl_batch_launch(ftl_batch* b, const float* dt, void* stream);
/* Tiles (64-lane work units) in the batch and the total number of
   workgroups its kernels launch. */
ftl_status ftl_batch_info(ftl_batch* b, int* ntiles, int* grid);

/* -- helpers -------------------------------------------------------------- */

/* The wind lattice the parameters select: FTL_WIND_GRID^3 float4
   (xyz acceleration, w = 0), entry (z*G + y)*G + x.  CPU only. */
ftl_status ftl_wind_table(const ftl_params* p, float* out, int* grid);

/* Diagnostics: copies `bytes` (a multiple of 16) from src to dst, both device
   pointers, with a plain streaming kernel on `stream`.  The bandwidth
   reference the step's HBM traffic is compared with.  No counterpart in the
   reference. */
ftl_status ftl_diag_copy(ftl_ctx* ctx, const void* src, void* dst, size_t bytes, void* stream);
/* Diagnostics: device buffer helpers for the copy above. */
ftl_status ftl_diag_alloc(ftl_ctx* ctx, size_t bytes, void** out);
ftl_status ftl_diag_free(ftl_ctx* ctx, void* p, size_t bytes);

/* -- launch-shape tuning ---------------------------------------------------- */

/* Times every exact launch class of each tuned point count (4, 8, 16, 32,
   64) on the current device: the lanes-per-strand choices (1, 2, 4) and, for
   some, tile-shape variants (another load/store chunk or occupancy target),
   without and with wind, and stores the fastest class per point count in the
   context (used when lanes = 0; plain steps with wind run a different kernel
   form, so they take the second table).  If cache_path is not NULL the tables
   are read from it when present and written to it after tuning (JSON). */
ftl_status ftl_autotune(ftl_ctx* ctx, const char* cache_path, int force);
/* The chosen lanes per strand for npoints (1, 2 or 4): batched and served
   steps, and plain steps without wind. */
int ftl_lanes_for(ftl_ctx* ctx, int npoints);
/* The same for plain steps with wind.  No counterpart in the reference (its
   workgroup shape is fixed). */
int ftl_lanes_for_wind(ftl_ctx* ctx, int npoints);
/* The launch class the context picks for npoints (a plain step with wind
   when wind != 0; batched and served steps use wind = 0), -1 if none.  A class
   is a tile shape: points per lane, lanes per strand, the occupancy target the
   registers are allocated for and the load/store chunk; the tuning tables may
   pick a variant of a base class. */
int ftl_class_for_points(ftl_ctx* ctx, int npoints, int wind);
/* The shape of launch class cls (any pointer may be NULL): points per lane,
   lanes per strand, guarded (1: ragged point counts), waves per SIMD the
   register allocation targets, points per lane of a load/store chunk. */
ftl_status ftl_class_shape(int cls, int* points_per_lane, int* lanes, int* guarded, int* waves, int* chunk);

#ifdef __cplusplus
}
#endif

#endif /* FTL_H_ */
