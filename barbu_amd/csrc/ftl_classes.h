// Launch classes of the FTL step: (points per lane M, lanes per strand G,
// guarded).  Exact classes need npoints == M * G; guarded classes (G = 1)
// accept any npoints <= M.  The class id indexes the per-class kernels and the
// `cls` field of batched tiles.
#pragma once

// X(id, points per lane, lanes per strand, guarded, waves per SIMD targeted
// by the register allocator: the highest at which neither the plain nor the
// persistent kernel of the class spills)
#define FTL_CLASSES(X)     \
  X(0, 4, 1, false, 8)     \
  X(1, 8, 1, false, 6)     \
  X(2, 16, 1, false, 4)    \
  X(3, 32, 1, false, 2)    \
  X(4, 8, 2, false, 4)     \
  X(5, 16, 2, false, 2)    \
  X(6, 32, 2, false, 2)    \
  X(7, 4, 4, false, 6)     \
  X(8, 8, 4, false, 4)     \
  X(9, 16, 4, false, 3)    \
  X(10, 4, 1, true, 6)     \
  X(11, 8, 1, true, 6)     \
  X(12, 16, 1, true, 4)    \
  X(13, 32, 1, true, 2)    \
  X(14, 64, 1, true, 2)

#define FTL_NCLASSES 15

struct FtlClassInfo {
  int m, g, guarded, waves;
};

static constexpr FtlClassInfo kFtlClasses[FTL_NCLASSES] = {
#define FTL_CI(id, M, G, GU, W) {M, G, GU ? 1 : 0, W},
    FTL_CLASSES(FTL_CI)
#undef FTL_CI
};
