// Host-side entry points of the kernels (ftl_plain.hip, ftl_misc.hip), used by the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include "ftl_device.h"

// Launch classes: (points per lane M, lanes per strand G, guarded).
int ftl_num_classes();
int ftl_class_m(int cls);
int ftl_class_g(int cls);
int ftl_class_guarded(int cls);

// State size (positions and velocities of every strand a launch steps)
// above which the step kernels store with the non-temporal hint.  A step
// rewrites all of its state; when that exceeds the device's last-level cache
// (256 MB Infinity Cache) by enough that none of it is still cached when the
// next step reads it back, keeping the stores out of the caches measured
// faster (single grooms of 537 MB to 2.1 GB: -2 to -4%; the 64-groom frame,
// 1.9 GB: -2.5%); a 268 MB groom measured 4-7% slower with the hint.
size_t ftl_stream_store_bytes();
// start / stop: optional events bracketing the kernel itself (timing).
// nt: non-temporal stores (state above ftl_stream_store_bytes()).
hipError_t ftl_launch_plain(int cls, const FtlStep& A, hipStream_t stream, hipEvent_t start = nullptr,
                            hipEvent_t stop = nullptr, bool nt = false);
// Tile-list kernel of one class over the ntiles tiles of spans[0..nspans)
// (every groom of that class, in tile order): workgroup b steps tiles b,
// b + grid, ...
hipError_t ftl_launch_tiles(int cls, const FtlStep* grooms, const FtlSpan* spans, int nspans, int ntiles, int grid,
                            hipStream_t stream, bool nt);
// Resident tile-list workgroups per CU for a class.
int ftl_tiles_blocks_per_cu(int cls);
// Skinning of mesh vertices (and normals when nout is not null): mode 0 dual
// quaternions (2 float4 per bone), mode 1 linear blending (3 float4 rows per
// bone, joint 0xff = identity).
hipError_t ftl_launch_skin(int mode, int nverts, const float4* bones, const int4* joints, const float4* weights,
                           const float4* rest, const float4* nrest, float4* out, float4* nout, hipStream_t stream);
// Per strand of a skinned groom: the root on its bound triangle and, when
// normals is not null, the unit normal there.
hipError_t ftl_launch_root_frames(int nstrands, const int4* bind, const int4* tris, const float4* verts,
                                  const float4* normals, float4* roots, float4* nout, hipStream_t stream);
// Copies bytes (a multiple of 16) from pinned, coherent host memory to the
// device with a kernel on `stream`.
hipError_t ftl_launch_upload(const void* host_src, void* dst, size_t bytes, hipStream_t stream);
hipError_t ftl_launch_copy(const void* src, void* dst, size_t n16, int ncu, hipStream_t stream);
