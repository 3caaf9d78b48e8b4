// Batched persistent FTL step kernel: a grid sized to the device pulls 64-lane
// tiles of any groom from a work list (atomic counter) until it is drained.
// Built once per register budget with -DFTL_BATCH_MMAX=16|32: the classes with
// at most that many points per lane are compiled in.
#include "ftl_classes.h"
#include "ftl_launch.h"
#include "ftl_tile.h"

#ifndef FTL_BATCH_MMAX
#error "build with -DFTL_BATCH_MMAX=16|32"
#endif

namespace {

constexpr int kMmax = FTL_BATCH_MMAX;
constexpr int kWavesPerEu = kMmax <= 16 ? 3 : 2;

__global__ __launch_bounds__(FTL_WAVE) __attribute__((amdgpu_waves_per_eu(kWavesPerEu, 8)))
void ftl_batch_kernel(const FtlStep* __restrict__ grooms, const FtlTile* __restrict__ tiles, int ntiles,
                      unsigned* __restrict__ counter) {
  __shared__ float4 stage[FTL_WAVE * 8];
  __shared__ float4 u0s[FTL_WAVE];
  __shared__ int s_next;
  for (;;) {
    if (threadIdx.x == 0) s_next = (int)atomicAdd(counter, 1u);
    __syncthreads();
    const int t = __builtin_amdgcn_readfirstlane(s_next);
    __syncthreads();
    if (t >= ntiles) break;  // every wave reaches this once the list is drained
    const FtlTile td = tiles[t];
    const int cls = __builtin_amdgcn_readfirstlane(td.cls);
    const int gi = __builtin_amdgcn_readfirstlane(td.groom);
    const int s0 = __builtin_amdgcn_readfirstlane(td.strand0);
    const FtlStep& A = grooms[gi];
    switch (cls) {
#define FTL_CASE(id, M, G, GU)                                      \
  case id:                                                          \
    if constexpr (M <= kMmax) ftl_tile_impl::ftl_tile<M, G, GU>(A, s0, stage, u0s); \
    break;
      FTL_CLASSES(FTL_CASE)
#undef FTL_CASE
      default:
        break;
    }
  }
}

}  // namespace

#define FTL_PASTE2(a, b) a##b
#define FTL_PASTE(a, b) FTL_PASTE2(a, b)

hipError_t FTL_PASTE(ftl_launch_batch_m, FTL_BATCH_MMAX)(const FtlStep* grooms, const FtlTile* tiles, int ntiles,
                                                          unsigned* counter, int grid, hipStream_t stream) {
  hipLaunchKernelGGL(ftl_batch_kernel, dim3(grid), dim3(FTL_WAVE), 0, stream, grooms, tiles, ntiles, counter);
  return hipGetLastError();
}

int FTL_PASTE(ftl_batch_occ_m, FTL_BATCH_MMAX)() {
  int blocks = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, ftl_batch_kernel, FTL_WAVE, 0) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return blocks;
}
