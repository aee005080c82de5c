// Launch-grid arithmetic of the executor, host-only and free of HIP types so
// that it can be checked by a plain C++ program.
//
// An AQL dispatch packet counts work-items per dimension in 32 bits: a
// one-dimensional grid of n blocks x b threads must keep n * b < 2^32.
//  - Per-item kernels (one block per list item) use a 2D grid of at most
//    kGridX blocks per row; a block's item index is y * gridDim.x + x, and the
//    kernel returns for indices >= n.
//  - Element-wise kernels (k_zero, k_zero_slots) loop with a grid stride over
//    a grid capped at kZeroBlocks blocks.
#pragma once
#include <stdint.h>

namespace dfm {

constexpr unsigned kGridX = 32768;
constexpr int64_t kZeroBlocks = 65536;

struct Grid2 {
  unsigned x, y;
};

// 2D grid covering n >= 1 items (n < 2^31): x * y >= n and x * y - n < x.
inline Grid2 itemGrid2(int64_t n) {
  if (n < 1) n = 1;
  const int64_t gx = n < (int64_t)kGridX ? n : (int64_t)kGridX;
  Grid2 g;
  g.x = (unsigned)gx;
  g.y = (unsigned)((n + gx - 1) / gx);
  return g;
}

// Blocks of b threads for a grid-stride loop over n elements: at least one,
// at most kZeroBlocks.
inline unsigned zeroGrid(int64_t n, int b) {
  int64_t nb = (n + b - 1) / b;
  if (nb < 1) nb = 1;
  if (nb > kZeroBlocks) nb = kZeroBlocks;
  return (unsigned)nb;
}

}  // namespace dfm
