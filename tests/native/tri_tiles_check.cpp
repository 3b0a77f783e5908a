// Host check of the lower-triangle tiling of the two-sided pass (tri_tiles,
// tri_up, tri_ranges in csrc/necg_two.hpp and csrc/necg_pair2.hpp), with the
// per-tile ket ranges and pair rules of pair2_tile: every unordered pair of
// functions {k, l} is evaluated exactly once, each function's bra and ket
// partials are summed over exactly the tiles that wrote them, and no tile is
// empty.  The same for the lists the kernels take (tri_tile_list, decoded by
// tri_tile), whose last round may be split into half tiles: each function's
// bra partials then also come from the extra slots xr names.  Prints "ok" and
// exits 0, or names the first failure.
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

// The kernels' list for ncu CUs; counts the lists that have split tiles.
static int check_list(int K, int kt, int ncu, int &nsplit_lists) {
  std::vector<int> list, xr;
  const int r = tri_tile_list(K, kt, ncu, list, xr);
  long long nt, nfull;
  tri_counts(K, kt, nt, nfull);
  if (nt != tri_count(K, kt) || (long long)list.size() != 2 * (nt + r) ||
      r != tri_split_count(nt, nfull, kt, ncu)) {
    std::printf("K=%d kt=%d ncu=%d: list size\n", K, kt, ncu);
    return 1;
  }
  if (r) ++nsplit_lists;
  std::vector<unsigned char> seen((size_t)K * K, 0);
  std::vector<std::set<int>> bra_y(K), bra_e(K), ket_x(K);
  std::vector<int> e_use(r, 0);
  for (size_t i = 0; i < list.size() / 2; ++i) {
    const TriTile t = tri_tile(list.data(), (int)i, K, kt);
    if (t.l1 <= t.l0 || t.k0 >= K || t.e >= r) {
      std::printf("K=%d kt=%d ncu=%d: bad entry %zu\n", K, kt, ncu, i);
      return 1;
    }
    if (t.e >= 0) ++e_use[t.e];
    for (int k = t.k0; k < std::min(K, t.k0 + 64); ++k) {
      if (t.e >= 0) bra_e[k].insert(t.e);
      else bra_y[k].insert(t.by);
    }
    for (int l = t.l0; l < t.l1; ++l) ket_x[l].insert(t.bx);
    for (int k = t.k0; k < std::min(K, t.k0 + 64); ++k)
      for (int l = t.l0; l < t.l1; ++l)
        if (k >= l || l >= t.lup) {
          const int a = k > l ? k : l, b = k > l ? l : k;
          if (seen[(size_t)a * K + b]++) {
            std::printf("K=%d kt=%d ncu=%d: pair (%d, %d) twice\n", K, kt, ncu, a, b);
            return 1;
          }
        }
  }
  for (int e = 0; e < r; ++e)
    if (e_use[e] != 1) {
      std::printf("K=%d kt=%d ncu=%d: extra slot %d used %d times\n", K, kt, ncu, e, e_use[e]);
      return 1;
    }
  for (int a = 0; a < K; ++a)
    for (int b = 0; b <= a; ++b)
      if (seen[(size_t)a * K + b] != 1) {
        std::printf("K=%d kt=%d ncu=%d: pair (%d, %d) missed\n", K, kt, ncu, a, b);
        return 1;
      }
  for (int k = 0; k < K; ++k) {
    const TriRanges g = tri_ranges(K, kt, k);
    std::set<int> yb, ex, xk;
    for (int y = g.yb0; y < g.yb1; ++y) yb.insert(y);
    for (int y = g.yu0; y < g.yu1; ++y) yb.insert(y);
    for (int e = xr[2 * (k >> 6)]; e < xr[2 * (k >> 6) + 1]; ++e) ex.insert(e);
    for (int x = g.x0; x < g.x1; ++x) xk.insert(x);
    if (yb != bra_y[k] || ex != bra_e[k] || xk != ket_x[k]) {
      std::printf("K=%d kt=%d ncu=%d: function %d reduces the wrong partials\n", K, kt, ncu, k);
      return 1;
    }
  }
  return 0;
}

int main() {
  const int kts[] = {1, 2, 3, 4, 5, 7, 8, 12, 16, 32, 64, 100};
  int bad = 0;
  for (int K = 1; K <= 300 && !bad; ++K)
    for (int kt : kts) bad += check(K, kt);
  const int big[] = {511, 513, 1000, 1024, 1025, 2048, 2049};
  for (int K : big)
    for (int kt : kts) bad += check(K, kt);
  // the kernels' lists, with and without split last rounds
  int nsplit_lists = 0;
  const int cus[] = {1, 2, 3, 5, 8, 9, 20, 21, 40, 50, 80, 256};
  for (int K = 1; K <= 200 && !bad; K += 7)
    for (int kt : kts)
      for (int ncu : cus) bad += check_list(K, kt, ncu, nsplit_lists);
  for (int K : big)
    for (int kt : {2, 4, 8, 16})
      for (int ncu : {200, 256}) bad += check_list(K, kt, ncu, nsplit_lists);
  {
    std::vector<int> l, x;
    if (tri_tile_list(2048, 8, 256, l, x) != 128) {   // LiH K = 2048: 4,224 tiles on 256 CUs
      std::printf("K=2048 kt=8: expected 128 split tiles\n");
      ++bad;
    }
  }
  if (nsplit_lists < 50) {
    std::printf("only %d lists with split tiles\n", nsplit_lists);
    ++bad;
  }
  std::printf(bad ? "FAILED\n" : "ok\n");
  return bad ? 1 : 0;
}
