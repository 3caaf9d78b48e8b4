// Launch classes of the FTL step: (points per lane M, lanes per strand G,
// guarded).  Exact classes need npoints == M * G; guarded classes accept any
// npoints <= M (G = 1) or (G - 1) * M < npoints <= G * M (G = 2 or 4).  The class id indexes the per-class kernels and the
// `cls` field of batched tiles.
#pragma once

// X(id, points per lane, lanes per strand, guarded, waves per SIMD targeted
// by the register allocator: the highest at which neither the plain nor the
// tile-list kernel of the class spills, points per lane of a load/store
// chunk: 8 float4 make one 128-byte line, so each line of the groom is read
// by one request; 4 halves the LDS stage where a bigger one would cost
// occupancy, at the price of two requests per line)
#define FTL_CLASSES(X)        \
  X(0, 4, 1, false, 8, 4)     \
  X(1, 8, 1, false, 6, 4)     \
  X(2, 16, 1, false, 4, 4)    \
  X(3, 32, 1, false, 2, 8)    \
  X(4, 8, 2, false, 4, 8)     \
  X(5, 16, 2, false, 2, 8)    \
  X(6, 32, 2, false, 2, 8)    \
  X(7, 4, 4, false, 6, 4)     \
  X(8, 8, 4, false, 4, 8)     \
  X(9, 16, 4, false, 3, 8)    \
  X(10, 4, 1, true, 6, 4)     \
  X(11, 8, 1, true, 6, 4)     \
  X(12, 16, 1, true, 4, 4)    \
  X(13, 32, 1, true, 2, 4)    \
  X(14, 64, 1, true, 1, 8)    \
  X(15, 4, 2, false, 7, 4)     \
  X(16, 48, 1, true, 1, 8)     \
  X(17, 8, 2, true, 4, 8)      \
  X(18, 16, 2, true, 2, 8)     \
  X(19, 32, 2, true, 2, 8)     \
  X(20, 12, 2, true, 4, 4)     \
  X(21, 24, 2, true, 3, 8)     \
  X(22, 16, 4, true, 3, 8)     \
  FTL_VARIANTS(X)

// Tile-shape variants of exact classes: the same (M, G) with another load /
// store chunk or occupancy target.  ftl_class_for never picks them on its
// own; the autotuner times them beside the base classes and its tables
// select them per point count.
#define FTL_VARIANTS(X)        \
  X(23, 16, 2, false, 2, 4)    \
  X(24, 16, 4, false, 3, 4)    \
  X(25, 8, 2, false, 4, 4)     \
  X(26, 8, 4, false, 6, 8)     \
  X(27, 32, 1, false, 2, 4)    \
  X(28, 16, 1, false, 3, 4)

#define FTL_NCLASSES 29

struct FtlClassInfo {
  int m, g, guarded, waves, chunk;
};

static constexpr FtlClassInfo kFtlClasses[FTL_NCLASSES] = {
#define FTL_CI(id, M, G, GU, W, CH) {M, G, GU ? 1 : 0, W, CH},
    FTL_CLASSES(FTL_CI)
#undef FTL_CI
};

// Load/store chunk of a tile of M points per lane.  A kernel translation unit
// compiles one class (FTL_CLASS_ID), whose chunk it takes; elsewhere, the
// first class with these parameters.
constexpr int ftl_chunk_of(int m, int g, bool guarded) {
#ifdef FTL_CLASS_ID
  (void)g;
  (void)guarded;
  return kFtlClasses[FTL_CLASS_ID].chunk < m ? kFtlClasses[FTL_CLASS_ID].chunk : m;
#else
  for (int i = 0; i < FTL_NCLASSES; ++i)
    if (kFtlClasses[i].m == m && kFtlClasses[i].g == g && (kFtlClasses[i].guarded != 0) == guarded)
      return kFtlClasses[i].chunk < m ? kFtlClasses[i].chunk : m;
  return m < 4 ? m : 4;
#endif
}
