// Launch classes of the FTL step: (points per lane M, lanes per strand G,
// guarded).  Exact classes need npoints == M * G; guarded classes (G = 1)
// accept any npoints <= M.  The class id indexes the per-class kernels and the
// `cls` field of batched tiles.
#pragma once

#define FTL_CLASSES(X)  \
  X(0, 4, 1, false)     \
  X(1, 8, 1, false)     \
  X(2, 16, 1, false)    \
  X(3, 32, 1, false)    \
  X(4, 8, 2, false)     \
  X(5, 16, 2, false)    \
  X(6, 32, 2, false)    \
  X(7, 4, 4, false)     \
  X(8, 8, 4, false)     \
  X(9, 16, 4, false)    \
  X(10, 4, 1, true)     \
  X(11, 8, 1, true)     \
  X(12, 16, 1, true)    \
  X(13, 32, 1, true)    \
  X(14, 64, 1, true)

#define FTL_NCLASSES 15

struct FtlClassInfo {
  int m, g, guarded;
};

static constexpr FtlClassInfo kFtlClasses[FTL_NCLASSES] = {
#define FTL_CI(id, M, G, GU) {M, G, GU ? 1 : 0},
    FTL_CLASSES(FTL_CI)
#undef FTL_CI
};
