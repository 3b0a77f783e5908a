// necg_two.hpp -- host-side setup of the two-sided pair form (shared by the
// single-system context and the batch): whether a system's terms allow it,
// and the per-(term, lane) target-slot words of the ket-side sums.
#pragma once

#include <algorithm>
#include <functional>
#include <queue>
#include <utility>
#include <vector>

#include "../../include/necg.h"
#include "necg_pair2.hpp"

namespace necg {

// The two-sided pair form applies when every term is a particle
// permutation P (sml[p] a permutation matrix, sms[p] = kron(P^T, I3)) that
// leaves masses and charges unchanged, and P^T is a term of the same weight.
// Fills the per-(term, lane) words of target slots: for chunk c, byte c of
// lane j's word is the slot of the ket's own accumulator record that receives
// value 16 c + (j >> 2) of the two-sided term's ket-side values (X_H, X_S
// interleaved per packed (a, b), Y_H, Y_S per (a, x), sigma_H, sigma_S), with
// X' -> P X' P^T and Y' -> P Y' applied; 255 where there is no value.
inline bool two_sided_table(const necg_system *sys, std::vector<unsigned> &tab) {
  const int n = sys->n, nst = sys->nst, n3 = 3 * n;
  if (n > 5) return false;
  auto P = [&](int p, int i, int j) { return sys->sml[((size_t)p * n + i) * n + j]; };
  std::vector<int> pinv((size_t)nst * n);
  for (int p = 0; p < nst; ++p) {
    for (int i = 0; i < n; ++i) {
      int hits = 0, col = -1;
      for (int j = 0; j < n; ++j) {
        const double v = P(p, i, j);
        if (v != 0.0 && v != 1.0) return false;
        if (v == 1.0) {
          ++hits;
          col = j;
        }
      }
      if (hits != 1) return false;
      pinv[(size_t)p * n + col] = i;  // P[i][pi(i)] = 1, pinv = pi^-1
    }
    for (int j = 0; j < n; ++j) {
      int hits = 0;
      for (int i = 0; i < n; ++i) hits += P(p, i, j) == 1.0;
      if (hits != 1) return false;
    }
    for (int r = 0; r < n3; ++r)
      for (int q = 0; q < n3; ++q) {
        const double want = (r % 3 == q % 3) ? P(p, q / 3, r / 3) : 0.0;
        if (sys->sms[((size_t)p * n3 + r) * n3 + q] != want) return false;
      }
    // masses and charges unchanged: M[pi(i)][pi(j)] = M[i][j], q[pi(i)] = q[i]
    for (int i = 0; i < n; ++i) {
      int pii = -1;
      for (int j = 0; j < n; ++j)
        if (P(p, i, j) == 1.0) pii = j;
      if (sys->ncv[pii + 1] != sys->ncv[i + 1]) return false;
      for (int j = 0; j < n; ++j) {
        int pij = -1;
        for (int m = 0; m < n; ++m)
          if (P(p, j, m) == 1.0) pij = m;
        if (sys->mass[pii * n + pij] != sys->mass[i * n + j]) return false;
      }
    }
    // P^T is a term with the same weight
    bool found = false;
    for (int q = 0; q < nst && !found; ++q) {
      bool same = sys->scvv[q] == sys->scvv[p];
      for (int i = 0; i < n && same; ++i)
        for (int j = 0; j < n && same; ++j) same = P(q, i, j) == P(p, j, i);
      found = same;
    }
    if (!found) return false;
  }
  const int ns = n * (n + 1) / 2;
  const int a_xh = 0, a_yh = ns, a_xs = ns + n3, a_ys = 2 * ns + n3, a_sh = 2 * ns + 2 * n3;
  const int nv = a_sh + 2;
  if (nv > 64) return false;
  tab.assign((size_t)nst * 64, 0u);
  for (int p = 0; p < nst; ++p) {
    const int *iv = &pinv[(size_t)p * n];
    for (int lane = 0; lane < 64; ++lane) {
      unsigned word = 0;
      for (int c = 0; c < 4; ++c) {
        const int s = 16 * c + (lane >> 2);
        int t = 255;
        if (s < 2 * ns) {
          const int pk = s / 2;
          int a = 0;
          while ((a + 1) * (a + 2) / 2 <= pk) ++a;
          const int b = pk - a * (a + 1) / 2;
          t = ((s & 1) ? a_xs : a_xh) + sx(iv[a], iv[b]);
        } else if (s < 2 * ns + 2 * n3) {
          const int u = (s - 2 * ns) / 2;
          t = (((s - 2 * ns) & 1) ? a_ys : a_yh) + 3 * iv[u / 3] + u % 3;
        } else if (s < nv) {
          t = s - 2 * ns - 2 * n3 + a_sh;
        }
        word |= (unsigned)t << (8 * c);
      }
      tab[(size_t)p * 64 + lane] = word;
    }
  }
  return true;
}

// Lower-triangle tiles of the two-sided and element-only passes, in launch
// order (full tiles first): emit(bx, by, f) for bra tile bx (64 functions)
// and ket tile by (kt functions), f = 1 for the tiles of a folded diagonal
// block (tri_fold: by = 64 bx / kt + g, sweeps [g kt, g kt + kt) for
// g < 32 / kt, the block's diagonal tile for g = 32 / kt).
// Bra tile bx covers its kets l <= its last function.  When K is not a
// multiple of 64 and kt divides 64, the functions of the last, partial bra
// tile P are paired with the full tiles' lanes instead (upper pairs, kets >=
// tri_up(K, kt)), and tile P covers only its own diagonal block: fewer idle
// lanes.  With fold set (and kt dividing 32) every full bra tile's diagonal
// block is folded.
template <class Emit>
void tri_tiles(int K, int kt, int fold, Emit emit) {
  const int nbt = (K + 63) / 64, nkt = (K + kt - 1) / kt;
  const int up = tri_up(K, kt);
  const bool fd = tri_fold(kt, fold);
  std::vector<int> part;
  auto add = [&](int bx, int by) {
    const int l0 = by * kt, k0 = 64 * bx;
    const int lup = k0 < up ? up : K;
    int l1 = std::min(K, l0 + kt);
    if (l0 < lup) l1 = std::min(l1, k0 + 64);
    if (l1 - l0 == kt && std::min(64, K - k0) == 64) emit(bx, by, 0);
    else { part.push_back(bx); part.push_back(by); }
  };
  // folded blocks first, their diagonal tiles (short) then their fold tiles
  // (1.1 ordinary tiles): the longer tiles start before the ordinary ones
  const int nfb = fd ? K / 64 : 0;
  for (int bx = 0; bx < nfb; ++bx) emit(bx, 64 * bx / kt + 32 / kt, 1);
  for (int bx = 0; bx < nfb; ++bx)
    for (int g = 0; g < 32 / kt; ++g) emit(bx, 64 * bx / kt + g, 1);
  for (int bx = 0; bx < nbt; ++bx) {
    const int k0 = 64 * bx;
    const int lo = (k0 >= up) ? up / kt : 0;           // tile P alone: its own block
    const int nyl = std::min(nkt, (k0 + 63) / kt + 1);
    if (bx < nfb) {
      for (int by = lo; by < k0 / kt; ++by) add(bx, by);   // below the diagonal block
    } else {
      for (int by = lo; by < nyl; ++by) add(bx, by);
    }
    if (k0 < up && up < K)                             // upper pairs with tile P's functions
      for (int by = up / kt; by < nkt; ++by) add(bx, by);
  }
  for (size_t i = 0; i < part.size(); i += 2) emit(part[i], part[i + 1], 0);
}

inline long long tri_count(int K, int kt, int fold) {
  long long n = 0;
  tri_tiles(K, kt, fold, [&](int, int, int) { ++n; });
  return n;
}

// Tiles of the tri_tiles list of (K, kt), and how many whole regular tiles
// (kt kets, 64 bras, not fold tiles) come before its first partial tile.
inline void tri_counts(int K, int kt, int fold, long long &nt, long long &nfull) {
  nt = nfull = 0;
  bool partial = false;
  tri_tiles(K, kt, fold, [&](int bx, int by, int f) {
    ++nt;
    if (f) return;
    int l0, l1, lup;
    tri_extent(K, kt, bx, by, l0, l1, lup);
    if (l1 - l0 == kt && K - 64 * bx >= 64 && !partial) ++nfull;
    else partial = true;
  });
}

// Bra tiles whose diagonal block is folded (each has a diagonal tile).
inline int tri_fold_tiles(int K, int kt, int fold) {
  return tri_fold(kt, fold) ? K / 64 : 0;
}

// Whether a tri_tiles list of nt tiles (nfull whole-kt tiles first) on ncu
// CUs, one workgroup per CU, has its whole-tile last round split: a last
// round of r = nt mod ncu <= ncu / 2 workgroups becomes 2 r half tiles of
// kt / 2 kets, one round of half length.  Returns r, or 0 for no split.
inline int tri_split_count(long long nt, long long nfull, int kt, int ncu) {
  if (kt < 2 || (kt & 1) || ncu < 2 || nt < 2LL * ncu) return 0;
  const long long r = nt % ncu;
  if (r == 0 || 2 * r > ncu || r > nfull || r > 0x3fff) return 0;   // e has 14 bits
  return (int)r;
}

// Dispatch order for L2 locality.  The hardware hands the consecutive
// workgroups of a launch to the device's XCDs in turn (8 on MI355X, each with
// its own L2), and a tile (bx, by) reads the ket records of ket tile by.
// Tiles [t0, t1) of a list (word pairs) are reordered so that the tile at
// list position p has ket tile by = p mod 8 while that class lasts (else the
// next class that has tiles left), each class in its previous order: each
// XCD then streams the ket records of one eighth of the ket tiles (LiH
// K = 2048: 3.1 MB of the 64 fields the two-sided term reads, within a 4 MB
// L2) rather than a quarter or more of them.
constexpr int kXcds = 8;

inline void tri_xcd_order(std::vector<int> &list, size_t t0, size_t t1) {
  if (t1 <= t0 + 1) return;
  std::vector<std::vector<int>> q(kXcds);
  for (size_t i = t0; i < t1; ++i) {
    std::vector<int> &c = q[(size_t)(list[2 * i + 1] % kXcds)];
    c.push_back(list[2 * i]);
    c.push_back(list[2 * i + 1]);
  }
  std::vector<size_t> head(kXcds, 0);
  for (size_t i = t0; i < t1; ++i) {
    int c = (int)(i % kXcds);
    for (int s = 0; s < kXcds && head[(size_t)c] >= q[(size_t)c].size(); ++s) c = (c + 1) % kXcds;
    list[2 * i] = q[(size_t)c][head[(size_t)c]];
    list[2 * i + 1] = q[(size_t)c][head[(size_t)c] + 1];
    head[(size_t)c] += 2;
  }
}

// The tile list of a lower-triangle pass as the kernels take it (TriTile
// words): the tri_tiles list with its first r whole regular tiles split into
// halves (tri_split_count) and moved to the end, first and second half of
// each tile side by side, and the other whole regular tiles in XCD order
// (tri_xcd_order).  The first tiles belong to the bra tiles with the
// fewest ket tiles, whose reduction (tile_reduce2) has the least to sum:
// their extra slots leave the longest reduction, the last bra tile's, as it
// was.  xr[2 bx], xr[2 bx + 1] = the extra-slot range of bra tile bx's split
// tiles (contiguous: the split tiles are in bra-tile order).  Returns r.
inline int tri_tile_list(int K, int kt, int ncu, int fold, std::vector<int> &list,
                         std::vector<int> &xr) {
  std::vector<int> t;
  long long nt, nfull;
  tri_counts(K, kt, fold, nt, nfull);
  t.reserve((size_t)2 * nt);
  tri_tiles(K, kt, fold, [&](int bx, int by, int f) {
    t.push_back((int)((unsigned)bx | (f ? 3u << 30 : 0u)));
    t.push_back(by);
  });
  const int nbt = (K + 63) / 64;
  xr.assign((size_t)2 * nbt, 0);
  // the fold and diagonal tiles lead the list; the whole regular tiles follow
  size_t nf = 0;
  while (nf < (size_t)nt && ((unsigned)t[2 * nf] >> 30) == 3u) ++nf;
  // the diagonal tiles (short, dispatched first) take no slot of a round
  const int r = tri_split_count(nt - tri_fold_tiles(K, kt, fold), nfull, kt, ncu);
  if (r == 0) {
    list.swap(t);
    tri_xcd_order(list, nf, nf + (size_t)nfull);
    return 0;
  }
  list.clear();
  list.reserve((size_t)2 * (nt + r));
  std::vector<int> split;   // the first r whole regular tiles, in list order
  for (size_t i = 0; i < t.size(); i += 2) {
    const bool regular = ((unsigned)t[i] >> 30) == 0u;
    if (regular && (int)split.size() < 2 * r) {
      split.push_back(t[i]);
      split.push_back(t[i + 1]);
    } else {
      list.push_back(t[i]);
      list.push_back(t[i + 1]);
    }
  }
  // the whole regular tiles left after the r split ones
  tri_xcd_order(list, nf, nf + (size_t)(nfull - r));
  for (int e = 0; e < r; ++e) {
    const int bx = split[(size_t)2 * e], by = split[(size_t)2 * e + 1];
    if (xr[(size_t)2 * bx + 1] == 0) xr[(size_t)2 * bx] = e;
    xr[(size_t)2 * bx + 1] = e + 1;
    for (unsigned h = 1; h <= 2; ++h) {
      list.push_back((int)((unsigned)bx | ((unsigned)e << 16) | (h << 30)));
      list.push_back(by);
    }
  }
  return r;
}

// Cost of one workgroup of a lower-triangle launch, in term evaluations per
// lane, for tile t (tri_tile) with tp terms and w waves (waves take
// alternate kets or sweeps): a regular tile of nk kets ceil(nk / w) (tp +
// 0.06) + ovh, a fold tile of kt sweeps ceil(kt / w) (1.14 tp + 0.12) + ovh
// (two kets' wave sums per sweep), a diagonal tile 0.2 tp + ovh.  ovh is the
// tile start and end (tri_overhead); 0.06 is a ket's record flush.
inline double tri_tile_cost(const TriTile &t, int kt, int tp, int w, double ovh) {
  if (t.fold == 2) return 0.2 * tp + ovh;
  if (t.fold == 1) return (double)((kt + w - 1) / w) * (1.14 * tp + 0.12) + ovh;
  return (double)((t.l1 - t.l0 + w - 1) / w) * (tp + 0.06) + ovh;
}

// Per-workgroup overhead of tri_tile_cost for n pseudo-particles, in term
// evaluations.  Fitted with the constants of tri_tile_cost to the measured
// two-sided pair pass over forced tilings (ket tile 4..64, term groups
// 1..16): LiH (n = 5, 24 terms) at K = 256..4096, H2 (n = 3) and He (n = 2,
// 2 terms each) at K = 300..4096.  With these values the choice of
// tri_choose is the fastest measured tiling or within 0.3 % of it at every
// measured size.
inline double tri_overhead(int n) { return n == 3 ? 1.25 : 0.75; }

// Makespan of a lower-triangle launch of (K, kt) with ts term groups of tps
// terms (grid x = the tri_tile_list tiles, y = term groups, x fastest), on
// ncu CUs that each run one workgroup at a time and take the next workgroup
// in launch order when one finishes (list scheduling): the end of the last
// workgroup, in the units of tri_tile_cost.  nsplit = the list's split
// tiles (tri_tile_list).
inline double tri_makespan(int K, int kt, int ts, int tps, int fold, int ncu, int nst, int w,
                           double ovh, int &nsplit) {
  std::vector<int> list, xr;
  nsplit = tri_tile_list(K, kt, ncu, fold, list, xr);
  const int n = (int)(list.size() / 2);
  std::vector<TriTile> tiles((size_t)n);
  for (int i = 0; i < n; ++i) tiles[(size_t)i] = tri_tile(list.data(), i, K, kt);
  std::priority_queue<double, std::vector<double>, std::greater<double>> free_at;
  for (int i = 0; i < std::max(1, ncu); ++i) free_at.push(0.0);
  double end = 0.0;
  for (int z = 0; z < ts; ++z) {
    const int tp = std::min(nst, (z + 1) * tps) - z * tps;
    for (const TriTile &t : tiles) {
      const double e = free_at.top() + tri_tile_cost(t, kt, tp, w, ovh);
      free_at.pop();
      free_at.push(e);
      end = std::max(end, e);
    }
  }
  return end;
}

// Device bytes of the two-sided gradient pass's partials for (K, kt, ts):
// bra partials of every ket tile and split half, ket partials of every bra
// tile, the element partials of H and S (nacc accumulator slots per
// function).
inline size_t tri_grad_bytes(int K, int kt, int ts, int nsplit, int nacc) {
  const size_t nkt = (size_t)((K + kt - 1) / kt), nbt = (size_t)((K + 63) / 64);
  return 8 * ((nkt + (size_t)nsplit) * ts * nacc * (size_t)K + nbt * ts * (size_t)K * nacc +
              2 * (size_t)ts * K * K);
}

// Ket tile and term groups of a lower-triangle pass: of the ket tiles
// 2..64 (divisors of 64, for the upper pairs) and the term groups ts = 1, 2,
// 4, ... (as many as give distinct terms per group, while the launch has at
// most 8 workgroups per CU), the pair with the shortest tri_makespan; of
// those with its term groups and within 0.3 % of it, the one with the
// smallest partials.
// req_kt / req_ts > 0 fix one or both.  budget > 0: tilings whose partials
// (tri_grad_bytes) exceed it are passed over unless none fits, and then the
// smallest is taken.  Ket tiles whose list would exceed 2^18 tiles are passed
// over when a larger one is a candidate (the list is walked on the host), and
// so are ket tiles smaller than the wave count (idle waves) once the list has
// more than 16 rounds.
struct TriChoice {
  int kt = 8, ts = 1;
  double cost = -1.0;
};

inline TriChoice tri_choose(int K, int nst, int n, int w, int fold_mode, int ncu, int req_kt,
                            int req_ts, int nacc, size_t budget) {
  std::vector<int> kts, tss;
  if (req_kt > 0) kts.push_back(req_kt);
  else kts = {2, 4, 8, 16, 32, 64};
  if (req_ts > 0) tss.push_back(std::min(req_ts, nst));
  else for (int t = 1; t <= nst; t *= 2) tss.push_back(t);
  ncu = std::max(1, ncu);
  const double ovh = tri_overhead(n);
  const long long nbt = (K + 63) / 64;
  TriChoice small;
  size_t small_bytes = 0;
  std::vector<std::pair<TriChoice, size_t>> cands;   // (tiling, partial bytes) within budget
  for (int q : kts) {
    const long long approx = nbt * (nbt + 1) / 2 * ((64 + q - 1) / q);
    if (req_kt <= 0 && q < 64 && approx > (1LL << 18)) continue;
    if (req_kt <= 0 && q < w && approx > 16LL * ncu) continue;   // idle waves, many rounds
    const int fold = tri_fold(q, fold_mode) ? 1 : 0;
    long long nt, nfull;
    tri_counts(K, q, fold, nt, nfull);
    int last_tps = -1;
    for (int tq : tss) {
      const int tps = (nst + tq - 1) / tq, ts = (nst + tps - 1) / tps;
      if (tps == last_tps) continue;
      last_tps = tps;
      if (ts > 1 && req_ts <= 0 && nt * ts > 8LL * ncu) break;
      int nsplit = 0;
      const double c = tri_makespan(K, q, ts, tps, fold, ncu, nst, w, ovh, nsplit);
      const size_t bytes = tri_grad_bytes(K, q, ts, nsplit, nacc);
      if (small.cost < 0.0 || bytes < small_bytes) {
        small = {q, ts, c};
        small_bytes = bytes;
      }
      if (budget > 0 && bytes > budget) continue;
      cands.push_back({{q, ts, c}, bytes});
    }
  }
  if (cands.empty()) return small;
  // of the tilings with the shortest one's term groups and within 0.3 % of
  // it (the model's resolution), the smallest partials: less to write, and
  // less for the reduction after the pass
  const std::pair<TriChoice, size_t> *best = &cands[0];
  for (const auto &e : cands)
    if (e.first.cost < best->first.cost) best = &e;
  const std::pair<TriChoice, size_t> *pick = best;
  for (const auto &e : cands)
    if (e.first.ts == best->first.ts && e.first.cost <= 1.003 * best->first.cost &&
        e.second < pick->second)
      pick = &e;
  return pick->first;
}

}  // namespace necg
