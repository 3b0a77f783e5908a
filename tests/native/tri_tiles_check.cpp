// Host check of the lower-triangle tiling of the two-sided pass (tri_tiles,
// tri_up, tri_ranges in csrc/necg_two.hpp and csrc/necg_pair2.hpp), with the
// per-tile ket ranges and pair rules of pair2_tile: every unordered pair of
// functions {k, l} is evaluated exactly once, each function's bra and ket
// partials are summed over exactly the tiles that wrote them, and no tile is
// empty.  Prints "ok" and exits 0, or names the first failure.
#include <cstdio>
#include <set>
#include <vector>

#include "../../nbo_correlatedgaussians_amd/csrc/necg_two.hpp"

using namespace necg;

static int check(int K, int kt) {
  std::vector<unsigned char> seen((size_t)K * K, 0);
  std::vector<std::set<int>> bra_tiles(K), ket_tiles(K);
  const int up = tri_up(K, kt);
  int bad = 0;
  tri_tiles(K, kt, [&](int bx, int by) {
    const int k0 = 64 * bx, l0 = by * kt;
    const int lup = k0 < up ? up : K;
    const int l1 = l0 < lup ? std::min(std::min(K, l0 + kt), k0 + 64) : std::min(K, l0 + kt);
    if (l1 <= l0 || k0 >= K) {
      std::printf("K=%d kt=%d: empty tile (%d, %d)\n", K, kt, bx, by);
      ++bad;
      return;
    }
    for (int k = k0; k < std::min(K, k0 + 64); ++k) bra_tiles[k].insert(by);
    for (int l = l0; l < l1; ++l) ket_tiles[l].insert(bx);
    for (int k = k0; k < std::min(K, k0 + 64); ++k)
      for (int l = l0; l < l1; ++l)
        if (k >= l || l >= lup) {
          const int a = k > l ? k : l, b = k > l ? l : k;
          if (seen[(size_t)a * K + b]++) {
            std::printf("K=%d kt=%d: pair (%d, %d) twice\n", K, kt, a, b);
            ++bad;
          }
        }
  });
  for (int a = 0; a < K; ++a)
    for (int b = 0; b <= a; ++b)
      if (seen[(size_t)a * K + b] != 1) {
        std::printf("K=%d kt=%d: pair (%d, %d) seen %d times\n", K, kt, a, b,
                    seen[(size_t)a * K + b]);
        return bad + 1;
      }
  for (int k = 0; k < K; ++k) {
    const TriRanges r = tri_ranges(K, kt, k);
    std::set<int> yb, xk;
    for (int y = r.yb0; y < r.yb1; ++y) yb.insert(y);
    for (int y = r.yu0; y < r.yu1; ++y)
      if (!yb.insert(y).second) ++bad;
    for (int x = r.x0; x < r.x1; ++x) xk.insert(x);
    if (yb != bra_tiles[k] || xk != ket_tiles[k]) {
      std::printf("K=%d kt=%d: function %d reduces the wrong tiles\n", K, kt, k);
      return bad + 1;
    }
  }
  if (tri_count(K, kt) <= 0) ++bad;
  return bad;
}

int main() {
  const int kts[] = {1, 2, 3, 4, 5, 7, 8, 12, 16, 32, 64, 100};
  int bad = 0;
  for (int K = 1; K <= 300 && !bad; ++K)
    for (int kt : kts) bad += check(K, kt);
  const int big[] = {511, 513, 1000, 1024, 1025, 2048, 2049};
  for (int K : big)
    for (int kt : kts) bad += check(K, kt);
  std::printf(bad ? "FAILED\n" : "ok\n");
  return bad ? 1 : 0;
}
