// Plain (one groom) FTL step kernels for one launch class, selected at build
// time with -DFTL_CLASS_ID=<id>: one wave64 workgroup per tile of strands, or,
// for launches with wind, workgroups of four tiles sharing an LDS copy of the
// wind lattice; and the tile-list kernel of batched frames.  Each kernel comes
// in two forms, NT (non-temporal stores, for launches whose state exceeds
// ftl_stream_store_bytes()) and not.
#include <hip/hip_ext.h>

#include "ftl_classes.h"
#include "ftl_launch.h"
#include "ftl_tile.h"

#ifndef FTL_CLASS_ID
#error "build with -DFTL_CLASS_ID=<launch class>"
#endif
#ifndef FTL_WIND_LDS
#define FTL_WIND_LDS 1
#endif

namespace {

template <int ID>
struct ClassOf;
#define FTL_CO(id, M_, G_, GU_, W_, CH_)         \
  template <>                                \
  struct ClassOf<id> {                       \
    static constexpr int M = M_, G = G_, W = W_; \
    static constexpr bool GU = GU_;          \
  };
FTL_CLASSES(FTL_CO)
#undef FTL_CO

using CL = ClassOf<FTL_CLASS_ID>;

// Occupancy target of the class (waves per SIMD the register allocator aims at).
#ifdef FTL_WPE_OVERRIDE
constexpr int kWavesPerEu = FTL_WPE_OVERRIDE;
#else
constexpr int kWavesPerEu = CL::W;
#endif
// Load-phase prefetch depth: one row for the forms that target six or more
// waves per SIMD (their register budget), FTL_LOAD_AHEAD for the others.
constexpr int kLoadAhead = kWavesPerEu >= 6 ? 1 : FTL_LOAD_AHEAD;

template <bool NT>
__global__ __launch_bounds__(FTL_WAVE) __attribute__((amdgpu_waves_per_eu(kWavesPerEu, 8)))
void ftl_plain_kernel(FtlStep arg) {
  (void)arg;
  // The descriptor is read in place from the kernel-argument segment (scalar
  // loads, runtime-indexed collider array) instead of through a private copy
  // of the by-value argument.
#if defined(__HIP_DEVICE_COMPILE__)
  const FtlStep& A = *(const FtlStep*)__builtin_amdgcn_kernarg_segment_ptr();
#else
  const FtlStep& A = arg;
#endif
  __shared__ ftl_tile_impl::TileLds<CL::M, CL::G, CL::GU> sh;
  ftl_tile_impl::ftl_tile<CL::M, CL::G, CL::GU, NT, kLoadAhead>(A, blockIdx.x * (FTL_WAVE / CL::G), sh);
}

// With wind: four tiles (waves) per workgroup share one LDS copy of the wind
// lattice, so the load phase's gathers are LDS reads (their waits are
// independent of the HBM row loads in flight) instead of global gathers.
// Wave w of workgroup b steps tile 4b + w.  The copy is the half lattice (each
// node's first float4, 8 KB): four stages plus it make 40 KB, so the four
// workgroups of a CU at four waves per SIMD fit its 160 KB.  (Eight-tile
// workgroups with the whole lattice also fit, but a workgroup holds its LDS
// until its slowest tile ends.)  Used by the classes that run at most four
// waves per SIMD.
#ifndef FTL_WIND_WAVES
#define FTL_WIND_WAVES 4
#endif
constexpr int kWindWaves = FTL_WIND_WAVES;
// eight-wave groups copy the whole lattice, four-wave groups its half lattice
constexpr int kWindHalf = kWindWaves == 4;
constexpr int kLatticeF4 = (kWindHalf ? 1 : 2) * FTL_WIND_GRID * FTL_WIND_GRID * FTL_WIND_GRID;
constexpr size_t kWindGroupLds =
    kWindWaves * sizeof(ftl_tile_impl::TileLds<CL::M, CL::G, CL::GU>) + kLatticeF4 * sizeof(float4);
// workgroups per CU at the class's occupancy target (4 waves per SIMD for the
// classes that reach it)
constexpr int kWindGroupsPerCu = (kWavesPerEu == 2 ? 8 : 16) / kWindWaves;
constexpr bool kWindLds = (kWavesPerEu == 4 || kWavesPerEu == 2) && kWindGroupsPerCu * kWindGroupLds <= 160 * 1024;

template <bool NT>
__global__ __launch_bounds__(FTL_WAVE * kWindWaves) __attribute__((amdgpu_waves_per_eu(kWavesPerEu, 8)))
void ftl_plain_wind_kernel(FtlStep arg) {
  (void)arg;
#if defined(__HIP_DEVICE_COMPILE__)
  const FtlStep& A = *(const FtlStep*)__builtin_amdgcn_kernarg_segment_ptr();
#else
  const FtlStep& A = arg;
#endif
  // (a one-element placeholder for the classes that do not use this form)
  __shared__ float4 lattice[kWindLds ? kLatticeF4 : 1];
  __shared__ ftl_tile_impl::TileLds<CL::M, CL::G, CL::GU> sh[kWindLds ? kWindWaves : 1];
  if constexpr (kWindLds) {
    for (int i = (int)threadIdx.x; i < kLatticeF4; i += FTL_WAVE * kWindWaves)
      lattice[i] = A.wind[kWindHalf ? 2 * i : i];
    __syncthreads();
    constexpr int spt = FTL_WAVE / CL::G;
    const int ntiles = (A.nstrands + spt - 1) / spt;
    const int w = __builtin_amdgcn_readfirstlane((int)threadIdx.x / FTL_WAVE);
    const int t = (int)blockIdx.x * kWindWaves + w;
    if (t < ntiles)
      ftl_tile_impl::ftl_tile<CL::M, CL::G, CL::GU, NT, kLoadAhead, kWindHalf ? 2 : 1>(A, t * spt, sh[w], lattice);
  }
}

}  // namespace

#define FTL_PASTE2(a, b) a##b
#define FTL_PASTE(a, b) FTL_PASTE2(a, b)

// start / stop (may be null): events that take the kernel's own start and end
// timestamps (hipExtLaunchKernelGGL), for per-launch timing.
hipError_t FTL_PASTE(ftl_launch_plain_, FTL_CLASS_ID)(const FtlStep& A, hipStream_t stream, hipEvent_t start,
                                                      hipEvent_t stop, bool nt) {
  constexpr int spt = FTL_WAVE / CL::G;
  const int nblocks = (A.nstrands + spt - 1) / spt;
  if (nblocks <= 0) return hipSuccess;
  if (kWindLds && FTL_WIND_LDS && A.wind_grid) {
    const dim3 grid((nblocks + kWindWaves - 1) / kWindWaves), block(FTL_WAVE * kWindWaves);
    if (nt)
      hipExtLaunchKernelGGL(ftl_plain_wind_kernel<true>, grid, block, 0, stream, start, stop, 0, A);
    else
      hipExtLaunchKernelGGL(ftl_plain_wind_kernel<false>, grid, block, 0, stream, start, stop, 0, A);
    return hipGetLastError();
  }
  if (nt)
    hipExtLaunchKernelGGL(ftl_plain_kernel<true>, dim3(nblocks), dim3(FTL_WAVE), 0, stream, start, stop, 0, A);
  else
    hipExtLaunchKernelGGL(ftl_plain_kernel<false>, dim3(nblocks), dim3(FTL_WAVE), 0, stream, start, stop, 0, A);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Tile-list form of the same class for batched frames: workgroup b steps the
// tiles b, b + grid, b + 2 grid, ... of the 64-lane tiles of any grooms of
// this class, each groom a span of consecutive tiles; a workgroup finds its
// tile's groom by a binary search over the spans' first tiles (scalar loads,
// a few per tile).  With grid = number of tiles (the batched default)
// each workgroup takes one tile and the hardware dispatcher balances the
// load; a smaller grid makes the workgroups persistent.  (A shared atomic
// work counter was measured 2-5x slower: every claim contends on one address
// across the eight XCDs.)

namespace {

template <bool NT>
__global__ __launch_bounds__(FTL_WAVE) __attribute__((amdgpu_waves_per_eu(kWavesPerEu, 8)))
void ftl_tiles_kernel(const FtlStep* __restrict__ grooms, const FtlSpan* __restrict__ spans, int nspans,
                      int ntiles) {
  __shared__ ftl_tile_impl::TileLds<CL::M, CL::G, CL::GU> sh;
  constexpr int spt = FTL_WAVE / CL::G;
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    // the last span whose first tile is at or before t (spans[0].tile0 == 0)
    int lo = 0, hi = nspans - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (__builtin_amdgcn_readfirstlane(spans[mid].tile0) <= t)
        lo = mid;
      else
        hi = mid - 1;
    }
    const int gi = __builtin_amdgcn_readfirstlane(spans[lo].groom);
    const int s0 = (t - __builtin_amdgcn_readfirstlane(spans[lo].tile0)) * spt;
    ftl_tile_impl::ftl_tile<CL::M, CL::G, CL::GU, NT, kLoadAhead>(grooms[gi], s0, sh);
  }
}

}  // namespace

hipError_t FTL_PASTE(ftl_launch_tiles_, FTL_CLASS_ID)(const FtlStep* grooms, const FtlSpan* spans, int nspans,
                                                      int ntiles, int grid, hipStream_t stream, bool nt) {
  if (ntiles <= 0 || nspans <= 0) return hipSuccess;
  if (nt)
    hipLaunchKernelGGL(ftl_tiles_kernel<true>, dim3(grid), dim3(FTL_WAVE), 0, stream, grooms, spans, nspans, ntiles);
  else
    hipLaunchKernelGGL(ftl_tiles_kernel<false>, dim3(grid), dim3(FTL_WAVE), 0, stream, grooms, spans, nspans, ntiles);
  return hipGetLastError();
}

int FTL_PASTE(ftl_tiles_occ_, FTL_CLASS_ID)() {
  int blocks = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, ftl_tiles_kernel<false>, FTL_WAVE, 0) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return blocks;
}
