// Host-side entry points of the kernels (ftl_kernels.hip), used by the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include "ftl_device.h"

// Launch classes: (points per lane M, lanes per strand G, guarded).
int ftl_num_classes();
int ftl_class_m(int cls);
int ftl_class_g(int cls);
int ftl_class_guarded(int cls);

hipError_t ftl_launch_plain(int cls, const FtlStep& A, hipStream_t stream);
// Resets the tile counter, then runs the persistent kernel compiled for
// classes with at most `mmax` points per lane (16 or 32).
hipError_t ftl_launch_batch(int mmax, const FtlStep* grooms, const FtlTile* tiles, int ntiles,
                            unsigned* counter, int grid, hipStream_t stream);
hipError_t ftl_launch_skin(int nverts, const float4* dq, const int4* joints, const float4* weights,
                           const float4* rest, float4* out, hipStream_t stream);
// Resident persistent-kernel workgroups per CU for the given register budget.
int ftl_batch_blocks_per_cu(int mmax);
