// Tile function of the FTL step kernels (included by the per-class kernel
// translation units).
//
// Semantics: src/shaders/hair/01_simulation/cs_simulation.glsl (tcoppex/barbu)
//   predict / root          :179-193
//   SatisfyConstraints      :155-161 (K x {DistanceConstraint; CollisionConstraint})
//   DistanceConstraint      :93-127  (serial FTL sweep + damped velocity shift)
//   CollideSphere           :129-147
//   Eq. 9 / VelocityControl :79-84, :200-204 (DFTL mode)
//
// Mapping (the reference runs one 4-lane workgroup per strand with a serial
// sweep on lane 0, :88-127):
//   * A tile is one wave64.  Each strand is owned by G consecutive lanes
//     (G = 1, 2 or 4); lane g holds control points [g*M, g*M + M) in VGPRs for
//     the whole step (M = npoints / G).
//   * Load: the tile's strands are contiguous in HBM; they are read in chunks of
//     C points per lane with fully coalesced float4 loads, integrated (predict /
//     root) in that layout, and transposed through an XOR-swizzled LDS buffer.
//   * Constraints: each lane sweeps its own points.  With G > 1 the lanes of a
//     strand form a pipeline across iterations: lane g runs iteration t - g at
//     time step t and takes its leader point (lane g-1's last point, before
//     collisions) through a cross-lane shuffle.  K iterations take K + G - 1
//     time steps.
//   * Only the last iteration's velocity reaches the output (every iteration
//     overwrites u_i for i >= 1 before reading it), so velocity work, the
//     damped shift and the reflections are done in the last iteration only.
//   * Store: the inverse of the load, through the same LDS buffer, so every
//     global store writes whole 128-byte lines.  With G = 1, velocities leave
//     chunk by chunk during the last sweep and positions after it.  With G > 1
//     the lanes of a strand finish at different time steps, so velocities are
//     written by each lane as they become final; positions leave coalesced
//     after the pipeline drains.
//   * In place: a tile owns its strands, so no second buffer and no copy-back.

#pragma once

#include <type_traits>

#include "ftl_classes.h"
#include "ftl_device.h"
#ifndef FTL_NO_DFTL
#define FTL_NO_DFTL 0
#endif
#ifndef FTL_WIND_ON
#define FTL_WIND_ON 1
#endif
// Rows of the load phase whose loads are in flight ahead of the row being
// integrated (default of the forms with at most four waves per SIMD).
#ifndef FTL_LOAD_AHEAD
#define FTL_LOAD_AHEAD 3
#endif

using namespace ftl_dev;

namespace ftl_tile_impl {

enum { kVelShipped = 0, kVelDftl = 1 };
enum { kRootModel = 0, kRootRest = 1, kRootSkinned = 2 };

template <int M, int G, bool GUARD>
struct Tile {
  static constexpr int C = ftl_chunk_of(M, G, GUARD);  // points per lane per load/store chunk
  static constexpr int SPT = FTL_WAVE / G;   // strands per tile
};

// Pipelined forms stage their final velocities through LDS (whole-line
// stores) when a lane holds at most two chunks; with more, the flushes inside
// the unrolled sweep cost more registers than the partial-line stores they
// replace, and each lane stores its velocities directly.
template <int M, int G>
constexpr bool kStagedVel = (M / Tile<M, G, false>::C) <= 2;

// LDS slot of (owner lane r, point j of the chunk): XOR swizzle, conflict-free
// for both the per-lane and the coalesced access order.
template <int C>
__device__ __forceinline__ int slot(int r, int j) {
  return r * C + (j ^ (r & (C - 1)));
}

// Rest lengths of the lane's points.  With one lane per strand they live in
// LDS (point-major, so a wave's reads are conflict-free), leaving the VGPRs to
// the positions; the pipelined forms keep them in registers.
template <int M>
struct LenRegs {
  float v[M];
  __device__ __forceinline__ float get(int j) const { return v[j]; }
  __device__ __forceinline__ void set(int j, float x) { v[j] = x; }
};
struct LenLds {
  float* base;  // lane's column: element j at base[j * FTL_WAVE]
  __device__ __forceinline__ float get(int j) const { return base[j * FTL_WAVE]; }
  __device__ __forceinline__ void set(int j, float x) { base[j * FTL_WAVE] = x; }
};

// Parameters of the final iteration's velocities, read once per tile into
// scalar registers (reloading them point by point costs a scalar-memory wait
// each).
struct VelArgs {
  int dftl;  // DFTL Eq. 9 instead of the shipped correction
  float inv_dt, maxv, maxv2, delta;
};

__device__ __forceinline__ VelArgs vel_args(const FtlStep& A) {
  VelArgs V;
  V.dftl = pin_i(!FTL_NO_DFTL && A.vmode == kVelDftl);
  V.inv_dt = pin_s(A.inv_dt);
  V.maxv = pin_s(A.maxv);
  V.maxv2 = pin_s(A.maxv2);
  V.delta = pin_s(A.delta);
  return V;
}

// Stored velocity of a final point: the damped, reflected correction u as
// shipped, or DFTL Eq. 9 with VelocityControl (:79-84, :201-204).  In DFTL
// mode x is re-read from memory, where the point's position is still the
// step's input (positions are written after all velocities).
__device__ __forceinline__ float3 out_velocity(const VelArgs& V, const float4* __restrict__ pp, int j, float3 p,
                                               float3 u) {
  if (FTL_NO_DFTL || !V.dftl) return u;
  const float4 x = pp[j];
  float3 v = f3(((p.x - x.x) - u.x) * V.inv_dt, ((p.y - x.y) - u.y) * V.inv_dt, ((p.z - x.z) - u.z) * V.inv_dt);
  const float dp = dot3(v, v);
  if (dp > V.maxv2) v = mul3(v, V.maxv * rsqrt_c(dp));
  return v;
}

// Root source of strand s for the rest and skinned root modes (the model
// mode's source is the root's own input position).
__device__ __forceinline__ float3 root_source(const FtlStep& A, int s) {
  if (A.rmode == kRootRest) {
    const float4 r = A.roots[s];
    return f3(r.x, r.y, r.z);
  }
  const int4 b = A.bind[s];
  const int4 t = A.tris[b.x];
  const float w1 = __int_as_float(b.y), w2 = __int_as_float(b.z);
  const float w0 = (1.0f - w1) - w2;
  const float4 a = A.skinned[t.x], bb = A.skinned[t.y], c = A.skinned[t.z];
  return f3(__builtin_fmaf(w2, c.x, __builtin_fmaf(w1, bb.x, w0 * a.x)),
            __builtin_fmaf(w2, c.y, __builtin_fmaf(w1, bb.y, w0 * a.y)),
            __builtin_fmaf(w2, c.z, __builtin_fmaf(w1, bb.z, w0 * a.z)));
}

// Orders the tile's LDS stage accesses across the lanes of its wave.  A tile
// is one wave and only that wave touches its stage, so no workgroup barrier
// is needed (several tiles may share a workgroup): LDS instructions of one
// wave execute in issue order, and the fences keep the compiler from moving
// stage accesses across this point.
__device__ __forceinline__ void tile_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Writes chunk c of the tile (C points per lane, staged in LDS) to dst in
// the coalesced order of the load.  Called by the whole wave.  gsel >= 0
// writes only the points owned by lanes with g == gsel (the lanes of a
// strand finish at different time steps in the pipelined forms).  NT: the
// stores carry the non-temporal hint (launches whose state is too large for
// the next step to find any of it in the caches).
typedef float ftl_nt_f4 __attribute__((ext_vector_type(4)));
template <int M, int G, bool GUARD, bool NT>
__device__ __forceinline__ void flush_chunk(float4* __restrict__ dst, int strand0, int nstr, int N, int c,
                                            const float4* __restrict__ stage, int lane, int gsel = -1) {
  constexpr int C = Tile<M, G, GUARD>::C;
  // Row k holds, in lane ln, point ln % C of the chunk of owner lane
  // r = k * (64 / C) + ln / C.  64 / C is a multiple of G and of C, so from row
  // to row only the strand moves, by RS strands, and the LDS slot by 64: the
  // lane's address and slot are formed once and stepped.
  static_assert((FTL_WAVE / C) % G == 0 && (FTL_WAVE / C) % C == 0, "rows of a chunk hold whole strands");
  constexpr int RS = FTL_WAVE / C / G;
  // Recomputed here (opaque lane): keeping the load phase's addresses alive
  // until the stores would cost registers for the whole step.
  int ln = lane;
  asm volatile("" : "+v"(ln));
  const int r0 = ln / C, j = ln % C;
  const int s0 = strand0 + r0 / G;
  const int cp = (r0 % G) * M + c * C + j;
  const bool mine = (!GUARD || cp < N) && (gsel < 0 || r0 % G == gsel);
  float4* __restrict__ d0 = dst + ((size_t)s0 * N + cp);
  const float4* __restrict__ t0 = stage + slot<C>(r0, j);
  tile_sync();
#pragma unroll
  for (int k = 0; k < C; ++k)
    if (mine && s0 + k * RS < nstr) {
      const float4 t = t0[k * FTL_WAVE];
      if constexpr (NT)
        __builtin_nontemporal_store(ftl_nt_f4{t.x, t.y, t.z, t.w}, (ftl_nt_f4*)&d0[(size_t)(k * RS) * N]);
      else
        d0[(size_t)(k * RS) * N] = t;
    }
  tile_sync();
}

// Lane-predicated collisions of one point: lanes with `on` false are left
// untouched; `reflect` also reflects u (last iteration).
__device__ __forceinline__ void collide_pred(const Spheres& S, float3& p, float3& u, bool on, bool reflect,
                                             unsigned cmask) {
#pragma unroll
  for (int k = 0; k < FTL_NCOL; ++k) {
    if (!has_collider(cmask, k)) continue;
    const float3 c = f3(S.c[k][0], S.c[k][1], S.c[k][2]);
    const float3 w = sub3(p, c);
    const float dp = dot3(w, w);
    if (on && dp < S.r2[k]) {
      const float is = rsqrt_c(dp);
      const float3 n = mul3(w, is);
      p = fma3(S.r[k], n, c);
      if (reflect) {
        const float d2 = 2.0f * dot3(n, u);
        u = fma3(-d2, n, u);
      }
    }
  }
}

// Collisions of a sweep's points (not its last iteration), after all of the
// sweep's constraints: a constraint reads its leader before the leader's
// collision, so the collisions may follow the whole sweep.  Colliders outside,
// points inside: each collider's centre and squared radius are copied to
// vector registers once per sweep, since a vector instruction with a scalar
// operand issues at half the rate of one with vector operands only when
// several waves share a SIMD.  Each point still meets the colliders in order
// with the same operations.  Point j takes part when on && j < jend && (j > 0
// || on0).
template <int M>
__device__ __forceinline__ void collide_sweep(const Spheres& S, float3 (&p)[M], bool on, bool on0, int jend,
                                              unsigned cmask) {
#pragma unroll
  for (int k = 0; k < FTL_NCOL; ++k) {
    if (!has_collider(cmask, k)) continue;
    float cx = S.c[k][0], cy = S.c[k][1], cz = S.c[k][2], r2 = S.r2[k];
    asm volatile("" : "+v"(cx), "+v"(cy), "+v"(cz), "+v"(r2));
    const float3 c = f3(cx, cy, cz);
#pragma unroll
    for (int j = 0; j < M; ++j) {
      if (j >= jend) continue;
      const float3 w = sub3(p[j], c);
      const float dp = dot3(w, w);
      if (on && (j > 0 || on0) && dp < r2) {
        const float is = rsqrt_c(dp);
        const float3 n = mul3(w, is);
        p[j] = fma3(S.r[k], n, c);
      }
    }
  }
}

// One sweep of DistanceConstraint over this lane's points (G == 1), followed
// by the collisions of its points.  FINAL: last iteration, which
// also forms the velocities, stages them and flushes every completed chunk;
// the lane's last point is left to the caller (its velocity is the raw
// correction and its collision comes after the sweep).
template <int M, bool GUARD, bool FINAL, bool NT>
__device__ __forceinline__ void sweep(const FtlStep& A, const VelArgs& V, const Spheres& S, float3 (&p)[M],
                                      const LenLds& l, int nl,
                                      const float4* __restrict__ pp, float3& last_pre, float3& clast,
                                      float4* __restrict__ stage, int strand0, int N, int lane, unsigned cm) {
  constexpr int C = Tile<M, 1, GUARD>::C;
  // Opaque per sweep: keeps the compiler from hoisting sigma * l[j] out of the
  // iteration loop into M extra registers (one multiply per point instead).
  float sg = A.sigma;
  asm volatile("" : "+v"(sg));
  const float dl = V.delta;
#pragma unroll
  for (int j = 1; j < M; ++j) {
    // GUARD: slots j >= nl hold copies of the last point; they are swept like
    // the others (finite, never stored) so the loop keeps a fixed shape.
    // The leader p[j-1] has not been collided yet: its collision follows.
    const float3 q = p[j];
    ftl_point(p[j], p[j - 1], sg * l.get(j));
    const float3 pn = p[j];
    if (FINAL) {
      const float3 c = sub3(pn, q);
      if (!GUARD || j < nl) {
        // Point m = j-1 is final: u = delta * c_j.  The root keeps its own,
        // p0 - x0 (:192), formed here from x0, which is still in memory.
        const int m = j - 1;
        float3 pm = p[m];
        float3 u = mul3(c, dl);
        if (m == 0) {
          const float4 x0 = pp[0];
          u = sub3(pm, f3(x0.x, x0.y, x0.z));
        } else {
          collide<true>(S, pm, u, cm);
        }
        p[m] = pm;
        const float3 v = out_velocity(V, pp, m, pm, u);
        stage[slot<C>(lane, m % C)] = make_float4(v.x, v.y, v.z, 0.0f);
        if (m % C == C - 1) flush_chunk<M, 1, GUARD, NT>(A.vel, strand0, A.nstrands, N, m / C, stage, lane);
      }
      if (GUARD ? (j == nl - 1) : (j == M - 1)) {
        clast = c;
        last_pre = pn;
      }
    } else if (GUARD ? (j == nl - 1) : (j == M - 1)) {
      last_pre = pn;
    }
  }
  // Collisions of points 1 .. last (the root never collides; with GUARD the
  // slots past the last point are not collided at all).
  if (!FINAL) collide_sweep<M>(S, p, true, false, GUARD ? nl : M, cm);
}

// The value of the previous lane of the same row of 16 lanes (DPP row_shr:1,
// one vector instruction; the first lane of a row reads 0).  The lanes of a
// strand (G = 2 or 4) never straddle a row, so lane g > 0 of a strand reads
// lane g - 1, and lane 0 of a strand, which never uses the value, reads
// another strand's lane or 0.
__device__ __forceinline__ float from_prev_lane(float x) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x111, 0xf, 0xf, true));
}

// One time step of the lane pipeline (G > 1): every lane runs the same code;
// `active` lanes perform iteration t - g, `fin` lanes are in their last
// iteration and also write their final velocities (lanes finish at different
// time steps, so these go straight to memory; positions leave coalesced after
// the pipeline drains).  FINS: some lane may be in its last
// iteration (the last G time steps); the earlier steps skip the velocity work.
// GUARD (strands of N = NS points, the last lane of a strand holding fewer
// than M): tip is that lane's slot of the strand's last point (-1 on the other
// lanes).  The slots after it hold copies of the last input point; they are
// swept (finite, never stored) but take no velocity work, and the tip itself
// is finished at its own slot in the last iteration (its velocity is its own
// correction, as for the last lane of a full strand).
template <int M, int G, bool FINS, bool GUARD, bool NT>
__device__ __forceinline__ void sweep_pipe(const FtlStep& A, const VelArgs& V, const Spheres& S, float3 (&p)[M],
                                           const LenRegs<M>& l,
                                           float3 lead, int g, bool active, bool fin, int gfin,
                                           const float4* __restrict__ pp, float4* __restrict__ stage, int strand0,
                                           int lane, float3 u0, float3& last_pre, float3& cfirst, float3& clast,
                                           unsigned cm, int NS = M * G, int tip = -1, bool* tbad = nullptr) {
  constexpr int C = Tile<M, G, GUARD>::C;
  // The final steps' velocity constants in vector registers (vector-only
  // operands issue at twice the rate of a scalar operand).
  float dl = V.delta;
  VelArgs Vv = V;
  if constexpr (FINS) asm volatile("" : "+v"(dl), "+v"(Vv.inv_dt), "+v"(Vv.maxv2));
  if (!FINS) fin = false;
#pragma unroll
  for (int j = 0; j < M; ++j) {
    const float3 q = p[j];
    const bool upd = active && (j != 0 || g != 0);  // lane 0's slot 0 is the root
    // The point is updated in its own registers under the lane mask (a select
    // would be three vector instructions with a scalar mask operand); its
    // leader is the previous point before that point's collision.
    if (upd) ftl_point(p[j], j == 0 ? lead : p[j - 1], l.get(j));  // l holds sigma * rest length
    const float3 pn = p[j];
    if (FINS) {
      const float3 c = sub3(pn, q);
      if (j == 0) cfirst = sel3(fin, c, cfirst);
      if (j >= 1) {
        const int m = j - 1;
        const bool root = (m == 0) && (g == 0);
        // GUARD: m is a real point before the tip (or any point of a full lane)
        const bool real = !GUARD || tip < 0 || m < tip;
        const bool skip = GUARD && fin && m == tip;  // finished at its own slot
        float3 pm = p[m];
        float3 u = mul3(c, dl);
        if (m == 0) u = sel3(g == 0, u0, u);
        collide_pred(S, pm, u, active && !root && !skip, fin, cm);
        p[m] = pm;
        // Final velocities go through the LDS stage; a completed chunk of the
        // finishing lanes leaves coalesced.  The last chunk waits for the
        // lane's last point (after the pipeline drains).
        if constexpr (kStagedVel<M, G>) {
          if (fin && real) {
            const float3 v = out_velocity(Vv, pp, m, pm, u);
            stage[slot<C>(lane, m % C)] = make_float4(v.x, v.y, v.z, 0.0f);
          }
          if (m % C == C - 1 && m / C < M / C - 1)
            flush_chunk<M, G, GUARD, NT>(A.vel, strand0, A.nstrands, NS, m / C, stage, lane, gfin);
        } else if (fin && real && strand0 + lane / G < A.nstrands) {
          const float3 v = out_velocity(Vv, pp, m, pm, u);
          A.vel[(size_t)(strand0 + lane / G) * NS + g * M + m] = make_float4(v.x, v.y, v.z, 0.0f);
        }
      }
      if constexpr (GUARD) {
        if (fin && j == tip) {
          // the strand's last point: its own correction, collided and reflected
          float3 pt = pn, ut = c;
          collide_pred(S, pt, ut, true, true, cm);
          p[j] = pt;
          const float3 v = out_velocity(Vv, pp, j, pt, ut);
          if constexpr (kStagedVel<M, G>) {
            stage[slot<C>(lane, j % C)] = make_float4(v.x, v.y, v.z, 0.0f);
          } else if (strand0 + lane / G < A.nstrands) {
            A.vel[(size_t)(strand0 + lane / G) * NS + g * M + j] = make_float4(v.x, v.y, v.z, 0.0f);
          }
          *tbad = !(isfinite(pt.x) && isfinite(pt.y) && isfinite(pt.z) && isfinite(v.x) && isfinite(v.y) &&
                    isfinite(v.z));
        }
      }
      if (j == M - 1) clast = sel3(fin, c, clast);
    }
    if (j == M - 1) {
      // An inactive lane's value is read only by the next lane while that lane
      // is inactive too (fill and drain shift by one step per lane).
      last_pre = pn;
      if (FINS) {
        float3 dummy = f3(0.f, 0.f, 0.f);
        collide_pred(S, p[j], dummy, active && !fin, false, cm);  // deferred when final
      }
    }
  }
  // Before the final steps every point collides after the whole sweep (the
  // root of lane 0 excepted).
  if (!FINS) collide_sweep<M>(S, p, active, g != 0, M, cm);
}

// Shared memory of one tile: the load/store stage, and the rest lengths of
// the forms with one lane per strand (an empty base for the others, so a
// pipelined tile takes exactly its stage).
template <int M, bool ON>
struct TileLens {
  float lens[FTL_WAVE * M];
};
template <int M>
struct TileLens<M, false> {};

template <int M, int G, bool GUARD>
struct TileLds : TileLens<M, G == 1> {
  float4 stage[FTL_WAVE * Tile<M, G, GUARD>::C];
};

// PD: rows of the load phase whose loads are in flight ahead of the row being
// integrated (each costs 7 VGPRs; the forms with a high occupancy target use 1).
// LW: where the wind gathers read the groom's wind lattice: 0 global memory,
// 1 wl = the whole lattice copied to LDS by the workgroup, 2 wl = its half
// lattice (first float4 of each node) in LDS.
template <int M, int G, bool GUARD, bool NT, int PD_ = FTL_LOAD_AHEAD, int LW = 0>
__device__ __forceinline__ void ftl_tile(const FtlStep& A, int strand0, TileLds<M, G, GUARD>& sh,
                                         const float4* wl = nullptr) {
  float4* __restrict__ stage = sh.stage;
  constexpr int C = Tile<M, G, GUARD>::C;
  // Opaque: in the tile-list kernel (a workgroup may step several tiles)
  // nothing lane-derived is hoisted out of the tile loop (it would stay live, in registers, across every tile).
  int tid = (int)threadIdx.x;
  asm volatile("" : "+v"(tid));
  const int lane = tid & (FTL_WAVE - 1);
  const int g = (G == 1) ? 0 : (lane % G);
  const int N = GUARD ? A.npoints : M * G;
  // points held by this lane: with G > 1 and GUARD the strand's last lane
  // holds N - (G - 1) * M of its M slots (the guarded pipelined classes take
  // (G - 1) * M < N <= G * M)
  const int nl = GUARD ? (G == 1 ? A.npoints : min(max(A.npoints - g * M, 0), M)) : M;
  const int nstr = A.nstrands;
  const int strand = strand0 + lane / G;
  const bool valid = strand < nstr;
  const int slast = nstr - 1;

  float3 p[M];
  using Len = typename std::conditional<G == 1, LenLds, LenRegs<M>>::type;
  Len l;
  if constexpr (G == 1) l.base = sh.lens + lane;

  // ---- load + integrate, through LDS ------------------------------------
  // Row i of the tile is point i % C of chunk i / C of every owner lane, read
  // in the coalesced order (64 consecutive float4 per instruction).  Loads run
  // PD rows ahead of the row being integrated, and a row's wind gathers are
  // issued before the loads of row i + PD: loads return in issue order, so
  // waiting for the gathers never waits for HBM.  Loads are clamped into the
  // groom (out-of-range lanes read valid data that is never stored).
  const float dt = pin_s(A.dt), dt2 = pin_s(A.dt2);
  const float3 F0 = f3(pin_s(A.force[0]), pin_s(A.force[1]), pin_s(A.force[2]));
  const bool wind = FTL_WIND_ON && A.wind_grid;
  WindArgs W;
  W.T = (GlobalFloat4*)pin_p(A.wind);
  W.freq = pin_s(A.wind_freq);
  W.off[0] = pin_s(A.wind_off[0]);
  W.off[1] = pin_s(A.wind_off[1]);
  W.off[2] = pin_s(A.wind_off[2]);
  // The root source of the lane's strand, fetched first so that its latency
  // overlaps the row loads.  The root's row slot carries its input position
  // through the transpose; the root is formed from both after it.
  const bool owns_root = (G == 1 || g == 0);
  float3 rsrc = f3(0.f, 0.f, 0.f);
  if (owns_root && A.rmode != kRootModel) rsrc = root_source(A, min(strand, slast));
  __builtin_amdgcn_sched_barrier(0);

  constexpr int PD = PD_ < M ? PD_ : M;
  float4 xr[PD + 1], vr[PD + 1];
  auto row_index = [&](int i) -> size_t {
    const int q = (i % C) * FTL_WAVE + lane;
    const int r = q / C;
    const int s = min(strand0 + r / G, slast);
    int cp = (r % G) * M + (i / C) * C + (q % C);
    if (GUARD) cp = min(cp, N - 1);
    return (size_t)s * N + cp;
  };
  auto load_row = [&](int i) {
    const size_t idx = row_index(i);
    xr[i % (PD + 1)] = A.pos[idx];
    vr[i % (PD + 1)] = A.vel[idx];
  };
#pragma unroll
  for (int i = 0; i < PD; ++i) load_row(i);
  float3 u0 = f3(0.f, 0.f, 0.f);
#pragma unroll
  for (int i = 0; i < M; ++i) {
    const int c = i / C;
    const int q = (i % C) * FTL_WAVE + lane;
    const int r = q / C;
    const int j = q % C;
    const int cp = (r % G) * M + c * C + j;
    const float4 x = xr[i % (PD + 1)], v = vr[i % (PD + 1)];
    WindTaps wt;
    if (wind) {
      if constexpr (LW == 2)
        wt = wind_fetch_half(wl, W, f3(x.x, x.y, x.z));
      else if constexpr (LW == 1)
        wt = wind_fetch_from(wl, W, f3(x.x, x.y, x.z));
      else
        wt = wind_fetch(W, f3(x.x, x.y, x.z));
    }
    __builtin_amdgcn_sched_barrier(0);
    if (i + PD < M) load_row(i + PD);
    __builtin_amdgcn_sched_barrier(0);
    float3 F = F0;
    if (wind) {
      const float3 w = wind_blend(wt);
      F = f3(F0.x + w.x, F0.y + w.y, F0.z + w.z);
    }
    float3 pr = f3(__builtin_fmaf(dt2, F.x, __builtin_fmaf(dt, v.x, x.x)),
                   __builtin_fmaf(dt2, F.y, __builtin_fmaf(dt, v.y, x.y)),
                   __builtin_fmaf(dt2, F.z, __builtin_fmaf(dt, v.z, x.z)));
    if (cp == 0) pr = f3(x.x, x.y, x.z);
    stage[slot<C>(r, j)] = make_float4(pr.x, pr.y, pr.z, x.w);
    __builtin_amdgcn_sched_barrier(0);
    if (i % C == C - 1) {
      tile_sync();
#pragma unroll
      for (int k = 0; k < C; ++k) {
        const float4 t = stage[slot<C>(lane, k)];
        p[c * C + k] = f3(t.x, t.y, t.z);
        l.set(c * C + k, t.w);
      }
      tile_sync();
      if (c == 0 && owns_root) {
        // root: p0 = M * source (:190-193); the pipelined forms keep u0 = p0 - x0
        const float3 x0 = p[0];
        const float3 src = A.rmode == kRootModel ? x0 : rsrc;
        p[0] = A.model_ident ? src : transform(A, src);
        if (G > 1) u0 = sub3(p[0], x0);
      }
    }
  }

  // ---- rest lengths scaled once per tile (pipelined forms) ---------------
  // Every sweep uses sigma * l[j]: formed here once, the same product each
  // iteration formed before.  With sigma == 1 it is l[j] itself; otherwise
  // the position store re-reads the rest lengths from the step's input.  (The
  // one-lane forms keep their lengths in LDS and multiply per point: the
  // store's re-read costs them registers they do not have.)
  const bool unit_sigma = G == 1 || A.sigma == 1.0f;
  if (!unit_sigma) {
    const float sg = A.sigma;
#pragma unroll
    for (int j = 0; j < M; ++j) l.set(j, sg * l.get(j));
  }

  // ---- colliders the wave can reach this step ----------------------------
  unsigned cm;
  {
    float L = 0.0f;
#pragma unroll
    for (int j = 0; j < M; ++j)
      // |sigma * l|: a negative length scale or rest length still places a
      // point that far from its leader.
      if (j >= (G == 1 ? 1 : 0) && (!GUARD || j < nl)) L += fabsf(G == 1 ? A.sigma * l.get(j) : l.get(j));
    float3 root = p[0];
    if (G > 1) {
      // strand total and root, from the lanes of the strand
      for (int o = 1; o < G; o <<= 1) L += __shfl_xor(L, o, G);
      root = f3(__shfl(root.x, 0, G), __shfl(root.y, 0, G), __shfl(root.z, 0, G));
    }
    const unsigned mine = valid ? reachable_colliders(A, root, L) : 0u;
    unsigned any = 0u;
    for (int k = 0; k < A.ncol; ++k)
      if (__ballot((mine >> k) & 1u) != 0ull) any |= 1u << k;
    cm = __builtin_amdgcn_readfirstlane(any);
  }
  const Spheres S = load_spheres(A, cm);
  const VelArgs V = vel_args(A);

  // ---- constraints ------------------------------------------------------
  // Lane's own points, for the DFTL re-read of x (never stored through).
  const float4* __restrict__ pp = A.pos + (size_t)(valid ? strand : 0) * N + (size_t)g * M;
  const int K = A.iterations;
  float3 last_pre = p[0];
  float3 cfirst = f3(0.f, 0.f, 0.f), clast = f3(0.f, 0.f, 0.f);
  float3 u;
  bool bad = false;
  if constexpr (G == 1) {
    for (int it = 0; it < K - 1; ++it)
      sweep<M, GUARD, false, NT>(A, V, S, p, l, nl, pp, last_pre, clast, stage, strand0, N, lane, cm);
    sweep<M, GUARD, true, NT>(A, V, S, p, l, nl, pp, last_pre, clast, stage, strand0, N, lane, cm);
    // The last point: velocity = raw correction, collisions deferred to here.
    u = clast;
#pragma unroll
    for (int j = 1; j < M; ++j) {
      if (GUARD ? (j == nl - 1) : (j == M - 1)) {
        float3 pm = p[j];
        collide<true>(S, pm, u, cm);
        p[j] = pm;
        const float3 v = out_velocity(V, pp, j, pm, u);
        bad = !(isfinite(pm.x) && isfinite(pm.y) && isfinite(pm.z) && isfinite(v.x) && isfinite(v.y) &&
                isfinite(v.z));
        stage[slot<C>(lane, j % C)] = make_float4(v.x, v.y, v.z, 0.0f);
        flush_chunk<M, G, GUARD, NT>(A.vel, strand0, nstr, N, j / C, stage, lane);
      }
    }
  } else {
    // No lane is in its last iteration before time step K - 1.
    // GUARD: the last lane of a strand finishes its tip (slot nl - 1) in its
    // last sweep; the other lanes keep the post-drain path below.
    static_assert(16 % G == 0, "a strand's lanes lie in one row of 16 lanes");
    const int tip = (GUARD && g == G - 1) ? nl - 1 : -1;
    bool tbad = false;
    for (int t = 0; t < K - 1; ++t) {
      const float3 li = f3(from_prev_lane(last_pre.x), from_prev_lane(last_pre.y), from_prev_lane(last_pre.z));
      // (lane 0 of a strand never reads its lead: its slot 0 is the root)
      const float3 lead = li;
      sweep_pipe<M, G, false, GUARD, NT>(A, V, S, p, l, lead, g, t >= g, false, -1, pp, stage, strand0, lane, u0,
                                     last_pre, cfirst, clast, cm, N, tip, &tbad);
    }
    for (int t = K - 1; t < K + G - 1; ++t) {
      const float3 li = f3(from_prev_lane(last_pre.x), from_prev_lane(last_pre.y), from_prev_lane(last_pre.z));
      const float3 lead = li;
      const int it = t - g;
      sweep_pipe<M, G, true, GUARD, NT>(A, V, S, p, l, lead, g, it >= 0 && it < K, it == K - 1, t - (K - 1), pp,
                                    stage, strand0, lane, u0, last_pre, cfirst, clast, cm, N, tip, &tbad);
    }
    // The lane's last point: its velocity needs the next lane's correction.
    const float3 cn = f3(__shfl_down(cfirst.x, 1, G), __shfl_down(cfirst.y, 1, G),
                         __shfl_down(cfirst.z, 1, G));
    u = sel3(g == G - 1, clast, mul3(cn, V.delta));
    float3 pm = p[M - 1];
    const bool own = !GUARD || tip < 0;  // GUARD: the tip lane is done
    if constexpr (GUARD)
      collide_pred(S, pm, u, own, true, cm);
    else
      collide<true>(S, pm, u, cm);
    p[M - 1] = pm;
    // (GUARD: the tip lane's slot M - 1 may lie past the strand; it reads the
    // strand's first point instead, and the result is dropped)
    const float3 v = out_velocity(V, pp, own ? M - 1 : 0, pm, u);
    if constexpr (kStagedVel<M, G>) {
      if (own) stage[slot<C>(lane, (M - 1) % C)] = make_float4(v.x, v.y, v.z, 0.0f);
      flush_chunk<M, G, GUARD, NT>(A.vel, strand0, nstr, N, M / C - 1, stage, lane);
    } else if (valid && own) {
      A.vel[(size_t)strand * N + g * M + (M - 1)] = make_float4(v.x, v.y, v.z, 0.0f);
    }
    bad = own ? !(isfinite(pm.x) && isfinite(pm.y) && isfinite(pm.z) && isfinite(v.x) && isfinite(v.y) &&
                  isfinite(v.z))
              : tbad;
  }
  // positions, after every velocity (DFTL reads the step's input positions)
#pragma unroll
  for (int c = 0; c < M / C; ++c) {
    if (GUARD && G == 1 && c * C >= nl) continue;
    float w[C];
    if (unit_sigma) {
#pragma unroll
      for (int j = 0; j < C; ++j) w[j] = l.get(c * C + j);
    } else {
      // the unscaled rest lengths, from the input (positions are written
      // below); a guarded lane's slots past the strand read its last point
#pragma unroll
      for (int j = 0; j < C; ++j) w[j] = pp[GUARD ? min(c * C + j, nl - 1) : c * C + j].w;
    }
#pragma unroll
    for (int j = 0; j < C; ++j) {
      const float3 q = p[c * C + j];
      stage[slot<C>(lane, j)] = make_float4(q.x, q.y, q.z, w[j]);
    }
    flush_chunk<M, G, GUARD, NT>(A.pos, strand0, nstr, N, c, stage, lane);
  }
  bad = bad && valid;
  const unsigned long long b = __ballot(bad);
  // A plain (system-scope) store of 1, not a read-modify-write: the word may
  // live in pinned host memory (the serving loop's per-request flags), and
  // every writer stores the same value.
  if (b != 0ull && lane == (int)(__ffsll((long long)b) - 1))
    __hip_atomic_store(A.status, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

}  // namespace ftl_tile_impl
