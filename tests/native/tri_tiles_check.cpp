// Host check of the lower-triangle tiling of the two-sided pass (tri_tiles,
// tri_up, tri_ranges in csrc/necg_two.hpp and csrc/necg_pair2.hpp), with the
// per-tile ket ranges and pair rules of pair2_tile and pair2_fold_tile: every
// unordered pair of functions {k, l} is evaluated exactly once (the
// diagonals that fold tiles leave counted in the diagonal tiles), each
// function's bra and ket partials are summed over exactly the tiles that
// wrote them, and no tile is empty.  The same for the lists the kernels take
// (tri_tile_list, decoded by tri_tile), whose last round may be split into
// half tiles: each function's bra partials then also come from the extra
// slots xr names; and the lists' whole regular tiles stand in XCD order
// (tri_xcd_order).  Folded and unfolded tilings.  Prints "ok" and exits 0, or
// names the first failure.
#include <cstdio>
#include <set>
#include <vector>

#include "../../nbo_correlatedgaussians_amd/csrc/necg_two.hpp"

using namespace necg;

// The pairs and partial slots of one tile: bra partial slot by for the bras
// [k0, k0 + 64), a ket record for bra tile bx per ket; returns false on a
// pair seen twice.
struct Cover {
  int K;
  std::vector<unsigned char> seen;
  std::vector<std::set<int>> bra_tiles, ket_tiles;
  explicit Cover(int K_) : K(K_), seen((size_t)K_ * K_, 0), bra_tiles(K_), ket_tiles(K_) {}
  bool pair(int k, int l) {
    const int a = k > l ? k : l, b = k > l ? l : k;
    return seen[(size_t)a * K + b]++ == 0;
  }
  // a regular tile: bras [k0, k0+64) x kets [l0, l1), pairs k >= l or l >= lup
  bool regular(int bx, int y, int l0, int l1, int lup) {
    const int k0 = 64 * bx;
    bool ok = true;
    for (int k = k0; k < std::min(K, k0 + 64); ++k) bra_tiles[k].insert(y);
    for (int l = l0; l < l1; ++l) ket_tiles[l].insert(bx);
    for (int k = k0; k < std::min(K, k0 + 64); ++k)
      for (int l = l0; l < l1; ++l)
        if (k >= l || l >= lup) ok &= pair(k, l);
    return ok;
  }
  // a fold tile: sweeps [s0, s1) of bra tile bx's diagonal block
  bool fold(int bx, int y, int s0, int s1) {
    const int k0 = 64 * bx;
    bool ok = true;
    for (int k = k0; k < k0 + 64; ++k) bra_tiles[k].insert(y);
    for (int s = s0; s < s1; ++s) {
      ket_tiles[k0 + s].insert(bx);
      ket_tiles[k0 + s + 32].insert(bx);
      for (int j = 0; j < 64; ++j) ok &= pair(k0 + j, k0 + fold_ket(s, j));
    }
    return ok;
  }
  // the diagonal tile of a folded bra tile: diagonals of its upper 32
  // functions, bra partial slot y for all 64
  bool diag(int bx, int y) {
    const int k0 = 64 * bx;
    bool ok = true;
    for (int k = k0; k < k0 + 64; ++k) bra_tiles[k].insert(y);
    for (int j = 32; j < 64; ++j) ok &= pair(k0 + j, k0 + j);
    return ok;
  }
  int missed() const {
    for (int a = 0; a < K; ++a)
      for (int b = 0; b <= a; ++b)
        if (seen[(size_t)a * K + b] != 1) return a * K + b;
    return -1;
  }
};

static int check(int K, int kt, int fold) {
  Cover c(K);
  const int up = tri_up(K, kt);
  int bad = 0;
  tri_tiles(K, kt, fold, [&](int bx, int by, int f) {
    const int k0 = 64 * bx;
    if (f) {
      const int g = by - k0 / kt;
      if (g * kt == 32) {
        if (!c.diag(bx, by)) ++bad;
      } else if (!c.fold(bx, by, g * kt, g * kt + kt)) {
        ++bad;
      }
      return;
    }
    const int l0 = by * kt;
    const int lup = k0 < up ? up : K;
    const int l1 = l0 < lup ? std::min(std::min(K, l0 + kt), k0 + 64) : std::min(K, l0 + kt);
    if (l1 <= l0 || k0 >= K) {
      std::printf("K=%d kt=%d: empty tile (%d, %d)\n", K, kt, bx, by);
      ++bad;
      return;
    }
    if (!c.regular(bx, by, l0, l1, lup)) ++bad;
  });
  if (bad) {
    std::printf("K=%d kt=%d fold=%d: a pair twice\n", K, kt, fold);
    return bad;
  }
  const int m = c.missed();
  if (m >= 0) {
    std::printf("K=%d kt=%d fold=%d: pair (%d, %d) not seen once\n", K, kt, fold, m / K, m % K);
    return 1;
  }
  for (int k = 0; k < K; ++k) {
    const TriRanges r = tri_ranges(K, kt, k, fold);
    std::set<int> yb, xk;
    for (int y = r.yb0; y < r.yb1; ++y) yb.insert(y);
    for (int y = r.yu0; y < r.yu1; ++y)
      if (!yb.insert(y).second) ++bad;
    for (int x = r.x0; x < r.x1; ++x) xk.insert(x);
    if (yb != c.bra_tiles[k] || xk != c.ket_tiles[k]) {
      std::printf("K=%d kt=%d fold=%d: function %d reduces the wrong tiles\n", K, kt, fold, k);
      return bad + 1;
    }
  }
  if (tri_count(K, kt, fold) <= 0) ++bad;
  return bad;
}

// The XCD order of a tile list's whole regular tiles (tri_xcd_order): the
// section after the leading fold and diagonal tiles holds the whole regular
// tiles that were not split, in tri_tiles order within each class by mod 8,
// and list position i takes class i mod 8 while that class has tiles left
// (else the next class that has).  Returns the number of positions that got
// their own class i mod 8, or -1 on a mismatch.
static long long check_xcd(const std::vector<int> &list, int K, int kt, int fold, int r,
                           long long nfull) {
  std::vector<int> t;   // the tiles in tri_tiles order, as tri_tile_list sees them
  tri_tiles(K, kt, fold, [&](int bx, int by, int f) {
    t.push_back((int)((unsigned)bx | (f ? 3u << 30 : 0u)));
    t.push_back(by);
  });
  size_t nf = 0;
  while (2 * nf < t.size() && ((unsigned)t[2 * nf] >> 30) == 3u) ++nf;
  const size_t n0 = nf + (size_t)r, n1 = nf + (size_t)nfull;   // the section's tiles in t
  std::vector<std::vector<int>> q(kXcds);
  for (size_t i = n0; i < n1; ++i) {
    q[(size_t)(t[2 * i + 1] % kXcds)].push_back(t[2 * i]);
    q[(size_t)(t[2 * i + 1] % kXcds)].push_back(t[2 * i + 1]);
  }
  std::vector<size_t> head(kXcds, 0);
  long long own = 0;
  for (size_t i = nf; i < nf + (n1 - n0); ++i) {
    int c = (int)(i % kXcds);
    while (head[(size_t)c] >= q[(size_t)c].size()) c = (c + 1) % kXcds;
    if (list[2 * i] != q[(size_t)c][head[(size_t)c]] ||
        list[2 * i + 1] != q[(size_t)c][head[(size_t)c] + 1])
      return -1;
    head[(size_t)c] += 2;
    own += c == (int)(i % kXcds);
  }
  return own;
}

// The kernels' list for ncu CUs; counts the lists that have split tiles.
static int check_list(int K, int kt, int ncu, int fold, int &nsplit_lists) {
  std::vector<int> list, xr;
  const int r = tri_tile_list(K, kt, ncu, fold, list, xr);
  long long nt, nfull;
  tri_counts(K, kt, fold, nt, nfull);
  if (nt != tri_count(K, kt, fold) || (long long)list.size() != 2 * (nt + r) ||
      r != tri_split_count(nt - tri_fold_tiles(K, kt, fold), nfull, kt, ncu)) {
    std::printf("K=%d kt=%d ncu=%d fold=%d: list size\n", K, kt, ncu, fold);
    return 1;
  }
  if (check_xcd(list, K, kt, fold, r, nfull) < 0) {
    std::printf("K=%d kt=%d ncu=%d fold=%d: whole regular tiles not in XCD order\n", K, kt,
                ncu, fold);
    return 1;
  }
  if (r) ++nsplit_lists;
  Cover c(K);
  std::vector<std::set<int>> bra_e(K);
  std::vector<int> e_use(r, 0);
  for (size_t i = 0; i < list.size() / 2; ++i) {
    const TriTile t = tri_tile(list.data(), (int)i, K, kt);
    if (t.l1 <= t.l0 || t.k0 >= K || t.e >= r || (t.fold && (t.e >= 0 || t.k0 + 64 > K))) {
      std::printf("K=%d kt=%d ncu=%d fold=%d: bad entry %zu\n", K, kt, ncu, fold, i);
      return 1;
    }
    bool ok;
    if (t.fold == 2) {
      ok = c.diag(t.bx, t.by);
    } else if (t.fold) {
      ok = c.fold(t.bx, t.by, t.l0, t.l1);
    } else if (t.e >= 0) {
      ++e_use[t.e];
      for (int k = t.k0; k < std::min(K, t.k0 + 64); ++k) bra_e[k].insert(t.e);
      // the second half's bra partials go to extra slot e, not to slot by
      ok = c.regular(t.bx, -1, t.l0, t.l1, t.lup);
      for (int k = t.k0; k < std::min(K, t.k0 + 64); ++k) c.bra_tiles[k].erase(-1);
    } else {
      ok = c.regular(t.bx, t.by, t.l0, t.l1, t.lup);
    }
    if (!ok) {
      std::printf("K=%d kt=%d ncu=%d fold=%d: a pair twice\n", K, kt, ncu, fold);
      return 1;
    }
  }
  for (int e = 0; e < r; ++e)
    if (e_use[e] != 1) {
      std::printf("K=%d kt=%d ncu=%d: extra slot %d used %d times\n", K, kt, ncu, e, e_use[e]);
      return 1;
    }
  const int m = c.missed();
  if (m >= 0) {
    std::printf("K=%d kt=%d ncu=%d fold=%d: pair (%d, %d) missed\n", K, kt, ncu, fold, m / K,
                m % K);
    return 1;
  }
  for (int k = 0; k < K; ++k) {
    const TriRanges g = tri_ranges(K, kt, k, fold);
    std::set<int> yb, ex, xk;
    for (int y = g.yb0; y < g.yb1; ++y) yb.insert(y);
    for (int y = g.yu0; y < g.yu1; ++y) yb.insert(y);
    for (int e = xr[2 * (k >> 6)]; e < xr[2 * (k >> 6) + 1]; ++e) ex.insert(e);
    for (int x = g.x0; x < g.x1; ++x) xk.insert(x);
    if (yb != c.bra_tiles[k] || ex != bra_e[k] || xk != c.ket_tiles[k]) {
      std::printf("K=%d kt=%d ncu=%d fold=%d: function %d reduces the wrong partials\n", K, kt,
                  ncu, fold, k);
      return 1;
    }
  }
  return 0;
}

// The tiling model (tri_choose) against the measured two-sided pair pass:
// at each size below, a sweep over forced tilings (ket tile 4..64, term
// groups 1..16, pair pass time over 5-10 launches) found the expected
// tiling fastest, or within 0.2 % of the fastest (LiH K = 700: 16 x 4 took
// 0.7528 ms, 8 x 2 0.7536; K = 3000: 8 x 1 12.349 ms, 16 x 1 12.365; K = 2048,
// four alternating runs: 16 x 1 5.720-5.744 ms, 8 x 1 5.735-5.743).  With a
// byte budget below every tiling's partials the smallest is taken.
int check_choose(bool quick) {
  struct Want { int n, nst, K, kt, ts; };
  const Want want[] = {
      {5, 24, 256, 8, 4},  {5, 24, 400, 8, 8},   {5, 24, 700, 8, 2},  {5, 24, 1300, 8, 1},
      {5, 24, 1600, 32, 2}, {5, 24, 2048, 16, 1}, {5, 24, 3000, 16, 1}, {5, 24, 4096, 32, 1},
      {2, 2, 300, 4, 1},   {2, 2, 1024, 16, 1},  {2, 2, 2048, 32, 1}, {2, 2, 4096, 32, 1},
      {3, 2, 300, 4, 1},   {3, 2, 1024, 32, 1},  {3, 2, 2048, 32, 1}, {3, 2, 4096, 32, 1}};
  int bad = 0;
  for (const Want &w : want) {
    if (quick && w.K > 2048) continue;
    const TriChoice c = tri_choose(w.K, w.nst, w.n, 4, 1, 256, 0, 0, 62, 0);
    if (c.kt != w.kt || c.ts != w.ts) {
      std::printf("tri_choose n=%d K=%d: %d x %d, measured fastest %d x %d\n", w.n, w.K, c.kt,
                  c.ts, w.kt, w.ts);
      ++bad;
    }
  }
  // a budget: the pick's partials fit it; below every tiling, the smallest
  const TriChoice a = tri_choose(400, 24, 5, 4, 1, 256, 0, 0, 62, (size_t)24 << 20);
  std::vector<int> l, x;
  const int na = tri_tile_list(400, a.kt, 256, tri_fold(a.kt, 1) ? 1 : 0, l, x);
  if (tri_grad_bytes(400, a.kt, a.ts, na, 62) > ((size_t)24 << 20)) {
    std::printf("tri_choose K=400 with a 24 MiB budget: %d x %d does not fit\n", a.kt, a.ts);
    ++bad;
  }
  const TriChoice b = tri_choose(400, 24, 5, 4, 1, 256, 0, 0, 62, 1);
  if (b.ts != 1) {
    std::printf("tri_choose K=400 with no room: %d x %d, not the smallest\n", b.kt, b.ts);
    ++bad;
  }
  return bad;
}

int main(int argc, char **argv) {
  // "quick": a smaller sweep (for builds with sanitizers, several times slower)
  const bool quick = argc > 1 && argv[1][0] == 'q';
  const int kts[] = {1, 2, 3, 4, 5, 7, 8, 12, 16, 32, 64, 100};
  int bad = 0;
  for (int fold = 0; fold <= 1; ++fold) {
    for (int K = 1; K <= (quick ? 140 : 300) && !bad; ++K)
      for (int kt : kts) bad += check(K, kt, fold);
    const int big[] = {511, 513, 1000, 1024, 1025, 2048, 2049};
    for (int K : big)
      if (!quick || K < 1100)
        for (int kt : kts) bad += check(K, kt, fold);
  }
  // the kernels' lists, with and without split last rounds
  int nsplit_lists = 0;
  const int cus[] = {1, 2, 3, 5, 8, 9, 20, 21, 40, 50, 80, 256};
  for (int fold = 0; fold <= 1; ++fold) {
    for (int K = 1; K <= 200 && !bad; K += (quick ? 13 : 7))
      for (int kt : kts)
        for (int ncu : cus) bad += check_list(K, kt, ncu, fold, nsplit_lists);
    const int big[] = {511, 513, 1000, 1024, 1025, 2048, 2049};
    for (int K : big)
      for (int kt : {2, 4, 8, 16})
        for (int ncu : {200, 256}) bad += check_list(K, kt, ncu, fold, nsplit_lists);
  }
  {
    std::vector<int> l, x;
    if (tri_tile_list(2048, 8, 256, 0, l, x) != 128) {   // LiH K = 2048: 4,224 tiles
      std::printf("K=2048 kt=8: expected 128 split tiles\n");
      ++bad;
    }
    if (tri_count(2048, 8, 1) != 4128 || tri_tile_list(2048, 8, 256, 1, l, x) != 0) {
      std::printf("K=2048 kt=8 folded: expected 4,096 + 32 diagonal tiles, none split\n");
      ++bad;
    }
    // LiH K = 2048 at ket tile 16, folded: 1,984 whole regular tiles, and all
    // but the last few positions take their own XCD's class
    long long nt, nfull;
    tri_counts(2048, 16, 1, nt, nfull);
    const int r16 = tri_tile_list(2048, 16, 256, 1, l, x);
    const long long own = check_xcd(l, 2048, 16, 1, r16, nfull);
    if (nfull != 1984 || own < 1984 * 95 / 100) {
      std::printf("K=2048 kt=16 folded: %lld whole tiles, %lld in their own XCD class\n", nfull,
                  own);
      ++bad;
    }
  }
  bad += check_choose(quick);
  if (nsplit_lists < (quick ? 20 : 50)) {
    std::printf("only %d lists with split tiles\n", nsplit_lists);
    ++bad;
  }
  std::printf(bad ? "FAILED\n" : "ok\n");
  return bad ? 1 : 0;
}
