// Checks the launch-grid arithmetic of the executor (defmm_amd/csrc/grid.hpp)
// on the host: every item of a list is covered exactly once by the 2D item
// grid, and no grid dimension reaches 2^32 work-items, for list sizes up to
// 2^31 - 1 items and element counts far above 2^32.  Prints "ok" or the first
// failure.
#include <cstdio>
#include <cstdint>

#include "../../defmm_amd/csrc/grid.hpp"

int main() {
  using namespace dfm;
  const int64_t two32 = (int64_t)1 << 32;
  const int64_t items[] = {1, 2, 255, 32767, 32768, 32769, 65536, 1000003, 16777216, 16777217,
                           (int64_t)1 << 30, 2147483647};
  const int threads[] = {64, 128, 256, 384, 768, 1024};
  for (int64_t n : items)
    for (int b : threads) {
      const Grid2 g = itemGrid2(n);
      const int64_t cover = (int64_t)g.x * g.y;
      if (cover < n || cover - n >= (int64_t)g.x) {
        std::printf("item grid %lld: %u x %u does not cover exactly\n", (long long)n, g.x, g.y);
        return 1;
      }
      if ((int64_t)g.x * b >= two32 || (int64_t)g.y >= two32) {
        std::printf("item grid %lld x %d threads: a dimension reaches 2^32\n", (long long)n, b);
        return 1;
      }
      // the largest block index y * x + x - 1 must fit the kernels' int index
      if ((int64_t)(g.y - 1) * g.x + (g.x - 1) > 2147483647LL) {
        std::printf("item grid %lld: block index overflows int\n", (long long)n);
        return 1;
      }
    }
  // element-wise kernels: n up to 64 right-hand sides x 2^31 elements
  const int64_t elems[] = {0, 1, 255, 256, 257, two32 - 1, two32, two32 + 1, 5 * two32 + 7,
                           (int64_t)64 << 31};
  for (int64_t n : elems) {
    const unsigned nb = zeroGrid(n, 256);
    if (nb < 1 || (int64_t)nb > kZeroBlocks || (int64_t)nb * 256 >= two32) {
      std::printf("zero grid %lld: %u blocks\n", (long long)n, nb);
      return 1;
    }
    // a grid-stride loop reaches every element: stride * iterations >= n
    const int64_t stride = (int64_t)nb * 256;
    const int64_t iters = (n + stride - 1) / stride;
    if (iters * stride < n || (n > 0 && (iters - 1) * stride >= n)) {
      std::printf("zero grid %lld: stride loop does not cover\n", (long long)n);
      return 1;
    }
  }
  std::printf("ok\n");
  return 0;
}
