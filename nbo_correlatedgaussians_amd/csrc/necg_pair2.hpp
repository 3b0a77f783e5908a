// necg_pair2.hpp -- the two-sided pair term: one evaluation of the (bra k,
// ket l, term P) element gives the bra-side gradient (w.r.t. L_k, s_k) and
// the ket-side gradient (w.r.t. L_l, s_l) together, so every unordered pair
// {k, l} is evaluated once per term instead of once per ordering.
//
// Why this is the reference's gradient.  make_grad takes, for every ordered
// pair (l, k), the bra derivative of sum_P c_P <l|H|P k> (python/make_grad.py:
// 34-55, python/gradel.py:27).  When every term P is a particle permutation
// (sms = kron(P^T, I3)) under which masses and charges are invariant, and P^-1
// is a term with c_{P^-1} = c_P, then <l|H|P k> = <k|H|P^-1 l>, so
// sum_P c_P d_bra <l|H|P k> = sum_P c_P d_ket <k|H|P l>: the ket derivative of
// the (k, l) evaluation replaces the separate (l, k) evaluation.
//
// The element is symmetric under (A, S) <-> (A', S'), so the ket-side
// derivative w.r.t. (A', S') has the bra-side form with A <-> A',
// dif1 <-> dif2 and 12 A'MA' <-> 12 A M A in Gam.  With a1 = A dif1
// (A' dif2 = -a1) and Fc = Lb st - efs q e_z^T, B^-1 A' = I - B^-1 A gives
//   Phi_ket = B^-1 (Fc - 8 A M a1),
//   Phi_bra = B^-1 (Fc + 8 A'M a1) = Phi_ket + 8 M a1,
// and the kinetic trace 3 tr(M A B^-1 A') = 3 (tr(M A) - <A M A, B^-1>).
// The ket-side (X', Y') are in the frame of A' = P^T A_l P, S' = P^T S_l; in
// the ket's own frame they are P X' P^T and P Y', a relabelling of slots that
// the per-term target table applies when the ket-side sums are added up.
//
// Cross-lane sums.  Lanes are 64 bras of one ket, so the ket side is summed
// over the wave: 16 values at a time by a halving exchange (permlane32 swap,
// permlane16 swap, row mirror, half-row mirror), then the quads of the four
// chunks together by a transposing exchange, after which every lane holds a
// different value's full sum.  Every sum is formed in a fixed order, so
// results are bitwise reproducible.
#pragma once

#include "necg_pair.hpp"

namespace necg {

// Bra-side data of the two-sided term: per field f, the tile's 64 bras in
// LDS ([f][lane], stride 64 doubles), so that no part of the record has to
// stay in registers across the term.  Fields: A | S | AS | A M A | i1 pk C tr(MA).
template <int N>
struct BraLds {
  static constexpr int F_A = 0;
  static constexpr int F_S = F_A + Dim<N>::NS;
  static constexpr int F_AS = F_S + Dim<N>::N3;
  static constexpr int F_AMA = F_AS + Dim<N>::N3;
  static constexpr int F_I1 = F_AMA + Dim<N>::NS;
  static constexpr int F_PK = F_I1 + 1;
  static constexpr int F_C = F_PK + 1;
  static constexpr int F_TRMA = F_C + 1;
  static constexpr int NF = F_TRMA + 1;
  const double *p;  // this lane's field 0
  __device__ __forceinline__ double f(int t) const { return p[64 * t]; }
  __device__ __forceinline__ double A(int t) const { return f(F_A + t); }
  __device__ __forceinline__ double S(int t) const { return f(F_S + t); }
  __device__ __forceinline__ double AS(int t) const { return f(F_AS + t); }
  __device__ __forceinline__ double AMA(int t) const { return f(F_AMA + t); }
};

// Where field f of the tile's bra table comes from in the bra record.
template <int N>
__device__ __forceinline__ int bra_lds_src(int f) {
  using D = Dim<N>;
  using L = BraLds<N>;
  if (f < L::F_S) return D::B_A + f;
  if (f < L::F_AS) return D::B_S + (f - L::F_S);
  if (f < L::F_AMA) return D::B_AS + (f - L::F_AS);
  if (f < L::F_I1) return D::B_AMA + (f - L::F_AMA);
  if (f == L::F_I1) return D::B_I1;
  if (f == L::F_PK) return D::B_PK;
  if (f == L::F_C) return D::B_C;
  return D::B_TRMA;
}

// First ket of the upper pairs of a lower-triangle pass: when K is not a
// multiple of 64 and kt divides 64, the functions of the last bra tile
// P = (K-1)/64 are taken as kets of the full bra tiles (pairs k < l); else K
// (no upper pairs).
__host__ __device__ inline int tri_up(int K, int kt) {
  return ((K & 63) != 0 && kt > 0 && (64 % kt) == 0) ? 64 * ((K - 1) / 64) : K;
}

// Ket range [l0, l1) of lower-triangle tile (bra tile bx, ket tile by) and
// the first ket of its upper pairs, lup.
__host__ __device__ inline void tri_extent(int K, int kt, int bx, int by, int &l0, int &l1,
                                           int &lup) {
  const int k0 = 64 * bx, up = tri_up(K, kt);
  lup = k0 < up ? up : K;
  l0 = by * kt;
  l1 = l0 + kt < K ? l0 + kt : K;
  if (l0 < lup && k0 + 64 < l1) l1 = k0 + 64;
}

// Folded diagonal blocks.  The diagonal block of a full bra tile (64
// functions, bras = kets) holds 64 * 65 / 2 pairs; one wave per ket leaves
// half of its lanes idle there.  With folding, sweep s (0 <= s < 32) of the
// block serves two kets, s and s + 32, with every lane its own bra: lanes
// s .. s + 32 (lane offset d = (lane - s) mod 64 <= 32) pair with ket s,
// the diagonal at d = 0 included; the other 31 lanes pair with ket s + 32
// (offset 1 .. 31 from it).  Every off-diagonal pair {a, b} of the block is
// then evaluated exactly once, by ket b when (a - b) mod 64 is in its range,
// else by ket a; a pair with bra < ket is as valid as its mirror image in
// the two-sided form.  The 32 sweeps leave the diagonals of kets 32 .. 63,
// which one diagonal tile per block evaluates (bra half only).  Folding
// applies when kt divides 32: the block's 64 / kt ket tiles become 32 / kt
// fold tiles of kt sweeps (bra partial slots 64 bx / kt + g) and a diagonal
// tile (slot 64 bx / kt + 32 / kt).
__host__ __device__ inline bool tri_fold(int kt, int fold) {
  return fold && kt >= 1 && kt <= 32 && (32 % kt) == 0;
}

// The ket offset (s or s + 32) that lane `lane` pairs with in sweep s.
__host__ __device__ inline int fold_ket(int s, int lane) {
  return ((lane - s) & 63) <= 32 ? s : s + 32;
}

// One entry of a lower-triangle tile list: word 0 = bx | e << 16 | h << 30,
// word 1 = by.  h = 0: the whole tile; h = 1, 2: its first and second half of
// kets (kt even), the second half writing its bra partials to extra slot e;
// h = 3: fold tile g = by - 64 bx / kt of bra tile bx, sweeps [l0, l1) =
// [g kt, g kt + kt) of its diagonal block (fold = 1), or for g = 32 / kt its
// diagonal tile (fold = 2).
struct TriTile {
  int bx, by, k0, l0, l1, lup, e, fold;
};

__host__ __device__ inline TriTile tri_tile(const int *tiles, int i, int K, int kt) {
  const unsigned w = (unsigned)tiles[2 * i];
  TriTile t;
  t.bx = (int)(w & 0xffffu);
  t.by = tiles[2 * i + 1];
  t.k0 = 64 * t.bx;
  t.e = -1;
  t.fold = 0;
  const unsigned h = w >> 30;
  if (h == 3u) {
    const int g = t.by - t.k0 / kt;
    t.fold = g * kt == 32 ? 2 : 1;
    t.l0 = g * kt;
    t.l1 = t.l0 + kt;
    t.lup = K;
    return t;
  }
  tri_extent(K, kt, t.bx, t.by, t.l0, t.l1, t.lup);
  if (h == 1u) {
    t.l1 = t.l0 + kt / 2;
  } else if (h == 2u) {
    t.l0 += kt / 2;
    t.e = (int)((w >> 16) & 0x3fffu);
  }
  return t;
}

// Ket-side values per evaluation: X_H, X_S interleaved per packed (i, j),
// then Y_H, Y_S interleaved per (i, x), then sigma_H, sigma_S; padded to
// whole chunks of 16.
template <int N>
struct Two {
  static constexpr int NV = Dim<N>::NACC;
  static constexpr int NCH = (NV + 15) / 16;
  static constexpr int KV = NCH * 16;
  // ket-record fields the two-sided term reads (A', S', A'S', A'MA', i2, pl,
  // scvv C, scvv): the record's leading fields, one per lane when staged
  static constexpr int KF = Dim<N>::K_SC + 1;
};

__device__ __forceinline__ double lane_swap32_add(double x, double y) {
  // v_permlane32_swap: the upper 32 lanes of x trade places with the lower
  // 32 lanes of y.  Lanes 0-31 then hold (x, partner's x), lanes 32-63
  // (partner's y, y): the sum is x summed over lanes i, i^32 in the lower
  // half and y summed likewise in the upper half.
  const unsigned long long ux = __builtin_bit_cast(unsigned long long, x);
  const unsigned long long uy = __builtin_bit_cast(unsigned long long, y);
  const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)ux, (unsigned)uy, false, false);
  const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)(ux >> 32), (unsigned)(uy >> 32),
                                                   false, false);
  const double nx = __builtin_bit_cast(double, ((unsigned long long)hi[0] << 32) | lo[0]);
  const double ny = __builtin_bit_cast(double, ((unsigned long long)hi[1] << 32) | lo[1]);
  return nx + ny;
}

__device__ __forceinline__ double lane_swap16_add(double x, double y) {
  // v_permlane16_swap: odd 16-lane rows of x trade places with even rows of
  // y; even rows then sum x over rows (2r, 2r+1), odd rows sum y.
  const unsigned long long ux = __builtin_bit_cast(unsigned long long, x);
  const unsigned long long uy = __builtin_bit_cast(unsigned long long, y);
  const auto lo = __builtin_amdgcn_permlane16_swap((unsigned)ux, (unsigned)uy, false, false);
  const auto hi = __builtin_amdgcn_permlane16_swap((unsigned)(ux >> 32), (unsigned)(uy >> 32),
                                                   false, false);
  const double nx = __builtin_bit_cast(double, ((unsigned long long)hi[0] << 32) | lo[0]);
  const double ny = __builtin_bit_cast(double, ((unsigned long long)hi[1] << 32) | lo[1]);
  return nx + ny;
}

// v from the DPP partner lane (CTRL: 0x140 row mirror, 0x141 half-row
// mirror, quad permutations below 0x100); every lane has a partner and
// every row and bank is enabled, so no lane keeps a previous value (the
// move needs no initialised destination).
template <int CTRL>
__device__ __forceinline__ double lane_dpp(double v) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = __builtin_amdgcn_mov_dpp((int)(unsigned)u, CTRL, 0xf, 0xf, false);
  const unsigned hi = __builtin_amdgcn_mov_dpp((int)(unsigned)(u >> 32), CTRL, 0xf, 0xf, false);
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// One halving step with a DPP involution that pairs lanes with bit b clear
// and set: lanes with the bit clear keep x summed over the pair, lanes with
// it set keep y.
template <int CTRL>
__device__ __forceinline__ double lane_xchg_add(double x, double y, bool upper) {
  const double keep = upper ? y : x;
  const double send = upper ? x : y;
  return keep + lane_dpp<CTRL>(send);
}

// The halving exchange of 16 values over the wave: wave_sum16_a across lane
// bits 5 and 4 (16 values to 4 per lane), wave_sum16_rows within each row of
// 16 lanes (4 values to 1: a partial of value lane >> 2, to be summed over
// the lane's quad).
__device__ __forceinline__ void wave_sum16_a(const double *v, double *u) {
  double w[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) w[j] = lane_swap32_add(v[j], v[j + 8]);   // lane bit 5
#pragma unroll
  for (int j = 0; j < 4; ++j) u[j] = lane_swap16_add(w[j], w[j + 4]);   // lane bit 4
}

// The row steps after wave_sum16_a: each lane then holds a partial of value
// lane >> 2 of the chunk, to be summed over the four lanes of its quad.
__device__ __forceinline__ double wave_sum16_rows(const double *u, int lane) {
  double t[2];
  const bool b3 = (lane & 8) != 0, b2 = (lane & 4) != 0;
#pragma unroll
  for (int j = 0; j < 2; ++j) t[j] = lane_xchg_add<0x140>(u[j], u[j + 2], b3);  // row mirror
  return lane_xchg_add<0x141>(t[0], t[1], b2);                         // half-row mirror
}

// Four chunks' quad partials r[0..3] summed over each quad with a
// transposing exchange: lane q of a quad (q = lane & 3) returns the full sum
// of r[q], so the four lanes hold four different chunks' values.
__device__ __forceinline__ double quad_sum4(const double *r, int lane) {
  const bool b1 = (lane & 2) != 0, b0 = (lane & 1) != 0;
  const double s0 = lane_xchg_add<0x4E>(r[0], r[2], b1);   // quad_perm [2,3,0,1]
  const double s1 = lane_xchg_add<0x4E>(r[1], r[3], b1);
  return lane_xchg_add<0xB1>(s0, s1, b0);                  // quad_perm [1,0,3,2]
}

// One (bra, ket-under-term) evaluation, both sides.  kd: the ket record;
// mb, mk: 1 where this lane's bra-side / ket-side contribution counts, else
// 0.  Adds sc*ov*alpha to hel and sc*ov to sel and the bra-side gradient
// (weight scvv_P C_l) to the lane's LDS slots, and writes this lane's
// ket-side values (weight scvv_P C_k) to kv[0 .. Two<N>::KV), which ket_sums
// adds up over the wave.  R1: the mass matrix is diag(mdiag) + mc 1 1^T
// (SysConst::mrank1), so M a1 takes one column sum instead of a full product.
// FOLD (a sweep of a folded diagonal block): the bra side's two weight sums,
// sigma_H and sigma_S, are added to sig[0], sig[1] (registers) instead of
// their LDS slots, which the fold tile gives to the second ket.
template <int N, bool R1, bool FOLD = false>
__device__ __forceinline__ bool pair_term2(const BraLds<N> &b, const double *__restrict__ kd,
                                           const SysConst &sc_, double mb, double mk,
                                           double &hel, double &sel, LdsSink &sink, double *kv,
                                           double *sig = nullptr) {
  using D = Dim<N>;
  constexpr int NS = D::NS, N3 = D::N3;
  const double *__restrict__ Ap = kd + D::K_AP;
  const double *__restrict__ Sp = kd + D::K_SP;
  const double *__restrict__ ApSp = kd + D::K_APSP;
  const double *__restrict__ Np = kd + D::K_NP;
  const double i2 = kd[D::K_I2], pl = kd[D::K_PL], scv = kd[D::K_SC];
  const double wb = mb * kd[D::K_WC];      // scvv_P C_l

  // Cholesky of B = A + A' (lower, packed), rd[j] = 1 / R_jj.
  double R[NS], rd[N];
  bool ok = true;
#pragma unroll
  for (int j = 0; j < N; ++j) {
    double s = b.A(sx(j, j)) + Ap[sx(j, j)];
#pragma unroll
    for (int m = 0; m < j; ++m) s = __builtin_fma(-R[sx(j, m)], R[sx(j, m)], s);
    ok = ok && (s > 0.0);
    const double r = rsq(s);
    rd[j] = r;
    R[sx(j, j)] = s * r;
#pragma unroll
    for (int i = j + 1; i < N; ++i) {
      double t = b.A(sx(i, j)) + Ap[sx(i, j)];
#pragma unroll
      for (int m = 0; m < j; ++m) t = __builtin_fma(-R[sx(i, m)], R[sx(j, m)], t);
      R[sx(i, j)] = t * r;
    }
  }
  double Ri[NS];
#pragma unroll
  for (int j = 0; j < N; ++j) {
    Ri[sx(j, j)] = rd[j];
#pragma unroll
    for (int i = j + 1; i < N; ++i) {
      double t = 0.0;
#pragma unroll
      for (int m = j; m < i; ++m) t = __builtin_fma(R[sx(i, m)], Ri[sx(m, j)], t);
      Ri[sx(i, j)] = -t * rd[i];
    }
  }
  double Bi[NS];
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j <= i; ++j) {
      double t = 0.0;
#pragma unroll
      for (int m = i; m < N; ++m) t = __builtin_fma(Ri[sx(m, i)], Ri[sx(m, j)], t);
      Bi[sx(i, j)] = t;
    }
  double rp = rd[0];  // det(B)^-1/2 = prod 1 / R_jj; det(B)^-3/2 its cube
#pragma unroll
  for (int j = 1; j < N; ++j) rp *= rd[j];
  const double det32 = rp * rp * rp;

  // E, st = B^-1 E, i3 (python/matel.py:125-133)
  double st[N3];
  double i3 = 0.0;
  {
    double Em[N3];
#pragma unroll
    for (int t = 0; t < N3; ++t) Em[t] = b.AS(t) + ApSp[t];
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int x = 0; x < 3; ++x) {
        double t = 0.0;
#pragma unroll
        for (int j = 0; j < N; ++j) t = __builtin_fma(Bi[sx(i, j)], Em[3 * j + x], t);
        st[3 * i + x] = t;
        i3 = __builtin_fma(Em[3 * i + x], t, i3);
      }
  }
  const double ov = b.f(BraLds<N>::F_PK) * pl * det32 * exp(i3 - b.f(BraLds<N>::F_I1) - i2);

  // Kinetic (python/matel.py:144-155): a1 = A dif1, Ma1 = M a1,
  // <A dif1, M A' dif2> = -<a1, M a1>; 3 tr(M A B^-1 A') = 3 (tr(MA) - <AMA, B^-1>).
  double Ma1[N3];
  double i3k = 0.0;
  {
    double a1[N3];
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int x = 0; x < 3; ++x) {
        double v = 0.0;
#pragma unroll
        for (int j = 0; j < N; ++j)
          v = __builtin_fma(b.A(sx(i, j)), st[3 * j + x] - b.S(3 * j + x), v);
        a1[3 * i + x] = v;
      }
    if constexpr (R1) {
      double cs[3];
#pragma unroll
      for (int x = 0; x < 3; ++x) {
        double t = a1[x];
#pragma unroll
        for (int j = 1; j < N; ++j) t += a1[3 * j + x];
        cs[x] = sc_.mc * t;
      }
#pragma unroll
      for (int i = 0; i < N; ++i)
#pragma unroll
        for (int x = 0; x < 3; ++x) {
          const double t = __builtin_fma(sc_.mdiag[i], a1[3 * i + x], cs[x]);
          Ma1[3 * i + x] = t;
          i3k = __builtin_fma(-a1[3 * i + x], t, i3k);
        }
    } else {
#pragma unroll
      for (int i = 0; i < N; ++i)
#pragma unroll
        for (int x = 0; x < 3; ++x) {
          double t = 0.0;
#pragma unroll
          for (int j = 0; j < N; ++j) t = __builtin_fma(sc_.msym[sx(i, j)], a1[3 * j + x], t);
          Ma1[3 * i + x] = t;
          i3k = __builtin_fma(-a1[3 * i + x], t, i3k);
        }
    }
  }
  double ab = 0.0;  // <AMA, B^-1> over the full symmetric matrices
  double abo = 0.0;  // the strictly lower part, counted twice below
#pragma unroll
  for (int i = 0; i < N; ++i) {
    ab = __builtin_fma(b.AMA(sx(i, i)), Bi[sx(i, i)], ab);
#pragma unroll
    for (int j = 0; j < i; ++j) abo = __builtin_fma(b.AMA(sx(i, j)), Bi[sx(i, j)], abo);
  }
  ab = __builtin_fma(2.0, abo, ab);
  const double tfac = 2.0 * i3k + 3.0 * (b.f(BraLds<N>::F_TRMA) - ab);

  // Coulomb (python/matel.py:162-197; gradients python/gradel.py:266-337).
  // Per particle pair (p, q) (p = 0 the fixed particle): with
  // h = 2 c_pq ri^3 / sqrt(pi), the coefficients of d(i2p)/dB^-1 are
  // ca = f h + 2 F0' h z (La) and cb = 2 F0' h (Lb).  The diagonal entry of
  // particle a sums the pairs that contain a (the first of them, (0, a),
  // sets it); the off-diagonal entry (a, b) is minus the pair's value.
  double csum2 = 0.0;  // twice the Coulomb sum
  // La and Gl = Lb st: Lb has the pair's cb on both diagonal entries and -cb
  // off the diagonal, so a particle pair (a, b) adds cb d and -cb d (d = st_a
  // - st_b) to rows a and b of Lb st, and a pair (0, b) sets row b to cb st_b
  double La[NS], Gl[N3];
  {
    int pi = 0;
#pragma unroll
    for (int p = 0; p <= N; ++p)
#pragma unroll
      for (int q = p + 1; q <= N; ++q, ++pi) {
        const int a = p - 1, bq = q - 1;
        double i2p, dx, dy, dz;
        if (p == 0) {
          i2p = Bi[sx(bq, bq)];
          dx = st[3 * bq];
          dy = st[3 * bq + 1];
          dz = st[3 * bq + 2];
        } else {
          i2p = __builtin_fma(-2.0, Bi[sx(a, bq)], Bi[sx(a, a)] + Bi[sx(bq, bq)]);
          dx = st[3 * a] - st[3 * bq];
          dy = st[3 * a + 1] - st[3 * bq + 1];
          dz = st[3 * a + 2] - st[3 * bq + 2];
        }
        const double ri = rsq(i2p);
        const double inv2 = ri * ri;
        const double z = __builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx)) * inv2;
        double f, f2p;
        boys_f0_line<false>(z, f, f2p);
        const double cr = sc_.cpair2[pi] * ri;   // 2 c / sqrt(pi i2p)
        csum2 = __builtin_fma(cr, f, csum2);
        const double h = cr * inv2;
        const double cb = f2p * h;
        const double ca = __builtin_fma(cb, z, f * h);
        if (p == 0) {
          La[sx(bq, bq)] = ca;
          Gl[3 * bq] = cb * dx;
          Gl[3 * bq + 1] = cb * dy;
          Gl[3 * bq + 2] = cb * dz;
        } else {
          La[sx(a, a)] += ca;
          La[sx(bq, bq)] += ca;
          La[sx(bq, a)] = -ca;
          Gl[3 * a] = __builtin_fma(cb, dx, Gl[3 * a]);
          Gl[3 * a + 1] = __builtin_fma(cb, dy, Gl[3 * a + 1]);
          Gl[3 * a + 2] = __builtin_fma(cb, dz, Gl[3 * a + 2]);
          Gl[3 * bq] = __builtin_fma(-cb, dx, Gl[3 * bq]);
          Gl[3 * bq + 1] = __builtin_fma(-cb, dy, Gl[3 * bq + 1]);
          Gl[3 * bq + 2] = __builtin_fma(-cb, dz, Gl[3 * bq + 2]);
        }
      }
  }
  // Stark (python/matel.py:205-208)
  double zsum = 0.0;
#pragma unroll
  for (int i = 0; i < N; ++i) zsum = __builtin_fma(sc_.q[i + 1], st[3 * i + 2], zsum);
  const double alpha = 2.0 * tfac + csum2 - sc_.efs * zsum;
  hel = __builtin_fma(scv * ov, alpha, hel);
  sel = __builtin_fma(scv, ov, sel);

  // Shared gradient piece: Phk = B^-1 (Lb st - efs q e_z^T - 8 A M a1), the
  // ket side's Phi; the bra side's is Phk + 8 M a1.
  double Phk[N3];
  {
    double G[N3];
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int x = 0; x < 3; ++x) {
        double t = Gl[3 * i + x], u = 0.0;
#pragma unroll
        for (int m = 0; m < N; ++m) u = __builtin_fma(b.A(sx(i, m)), Ma1[3 * m + x], u);
        if (x == 2) t = __builtin_fma(-sc_.efs, sc_.q[i + 1], t);
        G[3 * i + x] = __builtin_fma(-8.0, u, t);
      }
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int x = 0; x < 3; ++x) {
        double t = 0.0;
#pragma unroll
        for (int m = 0; m < N; ++m) t = __builtin_fma(Bi[sx(i, m)], G[3 * m + x], t);
        Phk[3 * i + x] = t;
      }
  }

  // Bra side: Phi = Phk + 8 M a1, G = B^-1 U B^-1 with U = 12 A'MA' + La.
  // With e = 2 wo dif1 (the Y_S values) the slots are
  //   X_S = -3 wo B^-1 - sum_x e_i d_j,   X_H = alpha X_S + wo (G - dp),
  //   Y_H = alpha e + wo Phi,             Y_S = e,
  // dp_ij = <d_i, Phi_j> + <Phi_i, d_j>; G row by row as (B^-1 U)_i B^-1.
  {
    double Phi[N3], d1[N3], e1[N3];
    const double wo = wb * ov;
    const double wo2 = 2.0 * wo;
#pragma unroll
    for (int t = 0; t < N3; ++t) {
      Phi[t] = __builtin_fma(8.0, Ma1[t], Phk[t]);
      d1[t] = st[t] - b.S(t);
      e1[t] = wo2 * d1[t];
    }
    // the shift and weight slots first: their LDS adds then drain under the
    // matrix products below instead of in front of the ket side's LDS reads
#pragma unroll
    for (int t = 0; t < N3; ++t) {
      sink.add(D::A_YH + t, __builtin_fma(alpha, e1[t], wo * Phi[t]));
      sink.add(D::A_YS + t, e1[t]);
    }
    const double woa = wo * alpha;
    if constexpr (FOLD) {
      sig[0] += woa;
      sig[1] += wo;
    } else {
      sink.add(D::A_SH, woa);
      sink.add(D::A_SS, wo);
    }
    double U[NS];
#pragma unroll
    for (int t = 0; t < NS; ++t) U[t] = __builtin_fma(12.0, Np[t], La[t]);
    const double m3wo = -3.0 * wo;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      double V[N];   // row i of B^-1 U
#pragma unroll
      for (int c = 0; c < N; ++c) {
        double v = Bi[sx(i, 0)] * U[sx(0, c)];
#pragma unroll
        for (int m = 1; m < N; ++m) v = __builtin_fma(Bi[sx(i, m)], U[sx(m, c)], v);
        V[c] = v;
      }
#pragma unroll
      for (int j = 0; j <= i; ++j) {
        double g = V[0] * Bi[sx(0, j)];
#pragma unroll
        for (int m = 1; m < N; ++m) g = __builtin_fma(V[m], Bi[sx(m, j)], g);
        double xs = m3wo * Bi[sx(i, j)];
#pragma unroll
        for (int x = 0; x < 3; ++x) {
          g = __builtin_fma(-d1[3 * i + x], Phi[3 * j + x], g);
          g = __builtin_fma(-Phi[3 * i + x], d1[3 * j + x], g);
          xs = __builtin_fma(-e1[3 * i + x], d1[3 * j + x], xs);
        }
        const int t = sx(i, j);
        sink.add(D::A_XH + t, __builtin_fma(alpha, xs, wo * g));
        sink.add(D::A_XS + t, xs);
      }
    }
  }

  // Ket side: Phi' = Phk, G' = B^-1 U' B^-1 with U' = 12 A M A + La,
  // dif2 = st - S', the slots as on the bra side, into kv (summed over the
  // wave by ket_sums).
  {
    const double wk = mk * scv * b.f(BraLds<N>::F_C);   // scvv_P C_k
    const double wo = wk * ov;
    const double wo2 = 2.0 * wo;
    double d2[N3], e2[N3];
#pragma unroll
    for (int t = 0; t < N3; ++t) {
      d2[t] = st[t] - Sp[t];
      e2[t] = wo2 * d2[t];
    }
    double U[NS];
#pragma unroll
    for (int t = 0; t < NS; ++t) U[t] = __builtin_fma(12.0, b.AMA(t), La[t]);
    const double m3wo = -3.0 * wo;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      double V[N];   // row i of B^-1 U'
#pragma unroll
      for (int c = 0; c < N; ++c) {
        double v = Bi[sx(i, 0)] * U[sx(0, c)];
#pragma unroll
        for (int m = 1; m < N; ++m) v = __builtin_fma(Bi[sx(i, m)], U[sx(m, c)], v);
        V[c] = v;
      }
#pragma unroll
      for (int j = 0; j <= i; ++j) {
        double g = V[0] * Bi[sx(0, j)];
#pragma unroll
        for (int m = 1; m < N; ++m) g = __builtin_fma(V[m], Bi[sx(m, j)], g);
        double xs = m3wo * Bi[sx(i, j)];
#pragma unroll
        for (int x = 0; x < 3; ++x) {
          g = __builtin_fma(-d2[3 * i + x], Phk[3 * j + x], g);
          g = __builtin_fma(-Phk[3 * i + x], d2[3 * j + x], g);
          xs = __builtin_fma(-e2[3 * i + x], d2[3 * j + x], xs);
        }
        const int t = sx(i, j);
        kv[2 * t] = __builtin_fma(alpha, xs, wo * g);
        kv[2 * t + 1] = xs;
      }
    }
#pragma unroll
    for (int t = 0; t < N3; ++t) {
      kv[2 * NS + 2 * t] = __builtin_fma(alpha, e2[t], wo * Phk[t]);
      kv[2 * NS + 2 * t + 1] = e2[t];
    }
    kv[2 * NS + 2 * N3] = wo * alpha;
    kv[2 * NS + 2 * N3 + 1] = wo;
#pragma unroll
    for (int t = Two<N>::NV; t < Two<N>::KV; ++t) kv[t] = 0.0;
  }
  return ok;
}

// The ket side of one term summed over the wave into the ket accumulator.
// Each chunk of 16 values is halved across lane bits 5 and 4 and then within
// each row of 16 lanes, leaving a partial per lane that its quad sums; the
// quads are summed with a transposing exchange over the (at most four)
// chunks, after which lane j holds value 16 (j & 3) + (j >> 2): its target
// slot is byte j & 3 of pw (this lane's four target slots, one byte per chunk,
// 255 = none).  Every lane adds its own value with one LDS add, so that no
// branch splits the caller's loop body.  A lane without a target holds the
// sum of a padding value (kv zero there, so the sum is +0.0 exactly, or a
// chunk past NCH, also +0.0) and adds it to the spare slot 63, which is never
// read.  FOLD: the lane's values go to kacc where mA = 1 and to kaccB where
// mA = 0 (two sums per chunk).
template <int N, bool FOLD = false>
__device__ __forceinline__ void ket_sums(const double *kv, unsigned pw, double *kacc, int lane,
                                         double mA = 1.0, double *kaccB = nullptr) {
  static_assert(Dim<N>::NACC < 64, "slot 63 is the spare");
  static_assert(Two<N>::NCH <= 4, "one transposing quad sum");
  const unsigned tg = (pw >> (8 * (lane & 3))) & 255u;
  const int slot = tg != 255u ? (int)tg : 63;
  double r[4], rc[4];
#pragma unroll
  for (int ch = 0; ch < 4; ++ch) {
    r[ch] = rc[ch] = 0.0;
    if (ch >= Two<N>::NCH) continue;
    double u[4];
    if constexpr (FOLD) {
      // split by ket (exact: mA is 0 or 1)
      double a[16], c[16];
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        a[i] = kv[16 * ch + i] * mA;
        c[i] = kv[16 * ch + i] - a[i];
      }
      wave_sum16_a(a, u);
      r[ch] = wave_sum16_rows(u, lane);
      wave_sum16_a(c, u);
      rc[ch] = wave_sum16_rows(u, lane);
    } else {
      wave_sum16_a(kv + 16 * ch, u);
      r[ch] = wave_sum16_rows(u, lane);
    }
  }
  __hip_atomic_fetch_add(kacc + slot, quad_sum4(r, lane), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_WAVEFRONT);
  if constexpr (FOLD)
    __hip_atomic_fetch_add(kaccB + slot, quad_sum4(rc, lane), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_WAVEFRONT);
}

// One two-sided pair tile: bras [k0, k0+64) x kets [l0, l1) x terms [p0, p1),
// run by the pair_waves(N) waves of a workgroup (each takes alternate kets).
// The ket fields of each (ket, term) reach the term through the wave's LDS
// buffer kst_s (64 doubles per wave): the record of the next (ket, term) is
// loaded, one field per lane, while the current term runs and stored to the
// buffer after it, so the term's reads are LDS broadcasts with no wait on
// scalar memory.
// A pair contributes when bra k >= ket l, or when l >= lup (the upper pairs
// of the last, partial bra tile's functions): the bra side of every such
// pair and the ket side of those with k != l.  Writes the element partials
// to hp/sp at [min(k, l)][max(k, l)], the tile's bra accumulators to bpt
// ([slot][k], NACC x K) and, per ket, its accumulator record summed over the
// tile's bras to kpt ([l][slot], K x NACC).
template <int N, bool R1>
__device__ __forceinline__ void pair2_tile(const double *__restrict__ bra,
                                           const double *__restrict__ ket, int K, int nst, int k0,
                                           int l0, int l1, int lup, int p0, int p1,
                                           const SysConst &sc, const unsigned *__restrict__ tperm,
                                           double *__restrict__ bpt, double *__restrict__ kpt,
                                           double *__restrict__ hp, double *__restrict__ sp,
                                           int *__restrict__ flag, double *acc_s,
                                           double *kacc_s, double *bra_s, double *kst_s) {
  using D = Dim<N>;
  constexpr int W = pair_waves(N);
  constexpr int KF = Two<N>::KF;
  static_assert(KF <= 64, "one staged ket field per lane");
  constexpr int NA = D::NACC * 64;
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int k = k0 + lane;
  const bool active = k < K;
  const int kk = active ? k : K - 1;
  // The wave's share of the bra table (fields wv, wv + W, ...) and its first
  // ket record: every load issued before the first store, so the tile start
  // waits for one memory round trip, not one per field.
  constexpr int NF = BraLds<N>::NF, FW = (NF + W - 1) / W;
  double bv[FW];
#pragma unroll
  for (int i = 0; i < FW; ++i) {
    const int f = wv + i * W;
    bv[i] = f < NF ? bra[(size_t)kk * D::BS + bra_lds_src<N>(f)] : 0.0;
  }
  double first = 0.0;
  if (l0 + wv < l1 && lane < KF) first = ket[((size_t)(l0 + wv) * nst + p0) * D::KS + lane];
  for (int v = 0; v < D::NACC; ++v) acc_s[wv * NA + v * 64 + lane] = 0.0;
  double *kacc = kacc_s + wv * 64;
  kacc[lane] = 0.0;
#pragma unroll
  for (int i = 0; i < FW; ++i) {
    const int f = wv + i * W;
    if (f < NF) bra_s[64 * f + lane] = bv[i];
  }
  double *kst = kst_s + wv * 64;
  if (lane < KF) kst[lane] = first;
  __syncthreads();
  const BraLds<N> b{bra_s + lane};
  LdsSink sink{acc_s + wv * NA + lane};
  bool ok = true;
  for (int l = l0 + wv; l < l1; l += W) {
    const bool upper = l >= lup;
    const bool valid = active && (k >= l || upper);
    const double mb = valid ? 1.0 : 0.0;
    const double mk = (valid && k != l) ? 1.0 : 0.0;
    double hel = 0.0, sel = 0.0;
    for (int p = p0; p < p1; ++p) {
      // the record after (l, p): (l, p + 1), else the wave's next ket at p0
      const int pn = p + 1 < p1 ? p + 1 : p0;
      const int ln = p + 1 < p1 ? l : l + W;
      double nx = 0.0;
      if (ln < l1 && lane < KF) nx = ket[((size_t)ln * nst + pn) * D::KS + lane];
      const unsigned pw = tperm[p * 64 + lane];
      double kv[Two<N>::KV];
      ok &= pair_term2<N, R1>(b, kst, sc, mb, mk, hel, sel, sink, kv);
      ket_sums<N>(kv, pw, kacc, lane);
      // after every read of the term (the wave's LDS operations complete in order)
      if (lane < KF) kst[lane] = nx;
    }
    if (valid) {
      const size_t o = upper ? (size_t)k * K + l : (size_t)l * K + k;
      hp[o] = hel;
      sp[o] = sel;
    }
    // this ket's record, summed over the tile's bras and terms: read and
    // clear (the wave's own LDS operations complete in order)
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    if (lane < D::NACC) {
      const double v = __hip_atomic_exchange(kacc + lane, 0.0, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_WAVEFRONT);
      kpt[(size_t)l * D::NACC + lane] = v;
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  }
  if (active && !ok) atomicOr(flag, (int)F_NOT_PD);
  __syncthreads();
  if (!active) return;
  // The wave's slots (wv, wv + W, ...) summed over the W waves' records: all
  // LDS reads first, then all stores, so the tile ends after one LDS and one
  // memory round trip rather than one per slot.
  // (Only the last round of slots can pass NACC - 1: its reads are clamped
  // into range, with no branch, and its store is skipped.)
  constexpr int VW = (D::NACC + W - 1) / W;
  double t[VW];
#pragma unroll
  for (int i = 0; i < VW; ++i) {
    const int v = min(wv + i * W, D::NACC - 1);
    double a = 0.0;
#pragma unroll
    for (int u = 0; u < W; ++u) a += acc_s[u * NA + v * 64 + lane];
    t[i] = a;
  }
  double *o = bpt + k;
#pragma unroll
  for (int i = 0; i < VW; ++i) {
    const int v = wv + i * W;
    if ((i + 1) * W <= D::NACC || v < D::NACC) o[(size_t)v * K] = t[i];
  }
}

// One fold tile: sweeps [s0, s1) of the folded diagonal block of the full
// bra tile at k0 (tri_fold), terms [p0, p1), the pair_waves(N) waves taking
// alternate sweeps.  Sweep s pairs every lane (its own bra k0 + lane) with
// ket k0 + fold_ket(s, lane): two kets per sweep, each with its own staged
// record and ket accumulator.  LDS is the regular tile's: the accumulators
// keep NACC - 2 slots per wave (sigma_H and sigma_S are summed in registers),
// and the 4 KB this frees at the end of acc_s holds the second ket's
// accumulator and staging buffer per wave.  Writes the element partials of
// each lane's pair at [min(k, l)][max(k, l)], the tile's bra accumulators to
// bpt ([slot][k]) and both kets' records to kpt ([l][slot]).
template <int N, bool R1>
__device__ __forceinline__ void pair2_fold_tile(const double *__restrict__ bra,
                                                const double *__restrict__ ket, int K, int nst,
                                                int k0, int s0, int s1, int p0, int p1,
                                                const SysConst &sc,
                                                const unsigned *__restrict__ tperm,
                                                double *__restrict__ bpt, double *__restrict__ kpt,
                                                double *__restrict__ hp, double *__restrict__ sp,
                                                int *__restrict__ flag, double *acc_s,
                                                double *kacc_s, double *bra_s, double *kst_s) {
  using D = Dim<N>;
  constexpr int W = pair_waves(N);
  constexpr int KF = Two<N>::KF;
  constexpr int NV = D::NACC - 2;   // LDS slots: every slot but sigma_H, sigma_S
  constexpr int NA = NV * 64;
  static_assert(D::A_SH == NV && D::A_SS == NV + 1, "the two weight slots are the last two");
  static_assert((D::NACC * 64 - NA) * W == 2 * W * 64,
                "the two freed slots per wave hold the second ket's buffers");
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int k = k0 + lane;          // a full bra tile: every lane is a bra
  constexpr int NF = BraLds<N>::NF, FW = (NF + W - 1) / W;
  double bv[FW];
#pragma unroll
  for (int i = 0; i < FW; ++i) {
    const int f = wv + i * W;
    bv[i] = f < NF ? bra[(size_t)k * D::BS + bra_lds_src<N>(f)] : 0.0;
  }
  double firstA = 0.0, firstB = 0.0;
  if (s0 + wv < s1 && lane < KF) {
    firstA = ket[((size_t)(k0 + s0 + wv) * nst + p0) * D::KS + lane];
    firstB = ket[((size_t)(k0 + s0 + wv + 32) * nst + p0) * D::KS + lane];
  }
  for (int v = 0; v < NV; ++v) acc_s[wv * NA + v * 64 + lane] = 0.0;
  double *tail = acc_s + W * NA;    // W * 128 doubles: kaccB, then kstB, per wave
  double *kaccA = kacc_s + wv * 64, *kaccB = tail + wv * 64;
  kaccA[lane] = 0.0;
  kaccB[lane] = 0.0;
#pragma unroll
  for (int i = 0; i < FW; ++i) {
    const int f = wv + i * W;
    if (f < NF) bra_s[64 * f + lane] = bv[i];
  }
  double *kstA = kst_s + wv * 64, *kstB = tail + W * 64 + wv * 64;
  if (lane < KF) {
    kstA[lane] = firstA;
    kstB[lane] = firstB;
  }
  __syncthreads();
  const BraLds<N> b{bra_s + lane};
  LdsSink sink{acc_s + wv * NA + lane};
  double sig[2] = {0.0, 0.0};
  bool ok = true;
  for (int s = s0 + wv; s < s1; s += W) {
    const int d = (lane - s) & 63;
    const bool first = d <= 32;                 // ket s (its diagonal at d = 0), else s + 32
    const int l = k0 + (first ? s : s + 32);
    const double mk = d == 0 ? 0.0 : 1.0;       // the diagonal takes the bra half only
    const double mA = first ? 1.0 : 0.0;
    const double *kd = first ? kstA : kstB;
    double hel = 0.0, sel = 0.0;
    for (int p = p0; p < p1; ++p) {
      const int pn = p + 1 < p1 ? p + 1 : p0;
      const int sn = p + 1 < p1 ? s : s + W;
      double nxA = 0.0, nxB = 0.0;
      if (sn < s1 && lane < KF) {
        nxA = ket[((size_t)(k0 + sn) * nst + pn) * D::KS + lane];
        nxB = ket[((size_t)(k0 + sn + 32) * nst + pn) * D::KS + lane];
      }
      const unsigned pw = tperm[p * 64 + lane];
      double kv[Two<N>::KV];
      ok &= pair_term2<N, R1, true>(b, kd, sc, 1.0, mk, hel, sel, sink, kv, sig);
      ket_sums<N, true>(kv, pw, kaccA, lane, mA, kaccB);
      if (lane < KF) {   // after every read of the term (in-order LDS operations)
        kstA[lane] = nxA;
        kstB[lane] = nxB;
      }
    }
    const size_t o = k < l ? (size_t)k * K + l : (size_t)l * K + k;
    hp[o] = hel;
    sp[o] = sel;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    if (lane < D::NACC) {
      const double va = __hip_atomic_exchange(kaccA + lane, 0.0, __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_WAVEFRONT);
      const double vb = __hip_atomic_exchange(kaccB + lane, 0.0, __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_WAVEFRONT);
      kpt[(size_t)(k0 + s) * D::NACC + lane] = va;
      kpt[(size_t)(k0 + s + 32) * D::NACC + lane] = vb;
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  }
  if (!ok) atomicOr(flag, (int)F_NOT_PD);
  __syncthreads();
  // the weight sums: each wave's registers into the freed tail, then every
  // slot summed over the W waves in a fixed order
  tail[(2 * wv) * 64 + lane] = sig[0];
  tail[(2 * wv + 1) * 64 + lane] = sig[1];
  __syncthreads();
  double *o = bpt + k;
  for (int v = wv; v < D::NACC; v += W) {
    double t = 0.0;
    if (v < NV) {
#pragma unroll
      for (int u = 0; u < W; ++u) t += acc_s[u * NA + v * 64 + lane];
    } else {
#pragma unroll
      for (int u = 0; u < W; ++u) t += tail[(2 * u + (v - NV)) * 64 + lane];
    }
    o[(size_t)v * K] = t;
  }
}

// The diagonal tile of a folded bra tile: the pairs (k, k) of the upper 32
// functions of the full bra tile at k0, terms [p0, p1), bra half only; the
// pair_waves(N) waves take alternate terms and every lane's own ket record is
// read from the ket table directly.  Writes the element partials at [k][k]
// and the bra accumulators of all 64 functions (zero for the lower 32) to bpt
// ([slot][k]), each the sum over the waves in a fixed order.  LDS is the
// regular tile's.
template <int N, bool R1>
__device__ __forceinline__ void pair2_diag_tile(const double *__restrict__ bra,
                                                const double *__restrict__ ket, int K, int nst,
                                                int k0, int p0, int p1, const SysConst &sc,
                                                const unsigned *__restrict__ tperm,
                                                double *__restrict__ bpt, double *__restrict__ hp,
                                                double *__restrict__ sp, int *__restrict__ flag,
                                                double *acc_s, double *kacc_s, double *bra_s,
                                                double *kst_s) {
  using D = Dim<N>;
  constexpr int W = pair_waves(N);
  constexpr int NF = BraLds<N>::NF;
  constexpr int NA = D::NACC * 64;
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int k = k0 + lane;
  for (int f = wv; f < NF; f += W)
    bra_s[64 * f + lane] = bra[(size_t)k * D::BS + bra_lds_src<N>(f)];
  for (int v = 0; v < D::NACC; ++v) acc_s[wv * NA + v * 64 + lane] = 0.0;
  double *kacc = kacc_s + wv * 64;   // the ket side is zero (mk = 0)
  kacc[lane] = 0.0;
  __syncthreads();
  const BraLds<N> b{bra_s + lane};
  LdsSink sink{acc_s + wv * NA + lane};
  const double mb = lane >= 32 ? 1.0 : 0.0;
  double hel = 0.0, sel = 0.0;
  bool ok = true;
  for (int p = p0 + wv; p < p1; p += W) {
    double kv[Two<N>::KV];   // the ket side is zero here and not summed
    ok &= pair_term2<N, R1>(b, ket + ((size_t)k * nst + p) * D::KS, sc, mb, 0.0, hel, sel, sink,
                            kv);
  }
  if (lane >= 32 && !ok) atomicOr(flag, (int)F_NOT_PD);
  kst_s[wv * 64 + lane] = hel;   // the staging buffers are free here
  kacc[lane] = sel;
  __syncthreads();
  if (wv == 0 && lane >= 32) {
    double h = 0.0, s2 = 0.0;
#pragma unroll
    for (int u = 0; u < W; ++u) {
      h += kst_s[u * 64 + lane];
      s2 += kacc_s[u * 64 + lane];
    }
    hp[(size_t)k * K + k] = h;
    sp[(size_t)k * K + k] = s2;
  }
  double *o = bpt + k;
  for (int v = wv; v < D::NACC; v += W) {
    double t = 0.0;
#pragma unroll
    for (int u = 0; u < W; ++u) t += acc_s[u * NA + v * 64 + lane];
    o[(size_t)v * K] = t;
  }
}

// The partials that function k sums in a lower-triangle pass (tri_tiles):
// as a bra, those of ket tiles [yb0, yb1) and [yu0, yu1) of its bra tile; as
// a ket, those of bra tiles [x0, x1).
struct TriRanges {
  int yb0, yb1, yu0, yu1, x0, x1;
};

__host__ __device__ inline TriRanges tri_ranges(int K, int kt, int k, int fold) {
  const int bx = k >> 6;
  const int nkt = (K + kt - 1) / kt, nbt = (K + 63) / 64;
  const int up = tri_up(K, kt);
  TriRanges r;
  r.yb0 = (64 * bx >= up) ? up / kt : 0;   // tile P alone: its own block
  r.yb1 = (64 * bx + 63) / kt + 1 < nkt ? (64 * bx + 63) / kt + 1 : nkt;
  if (tri_fold(kt, fold) && 64 * bx + 64 <= K)   // fold tiles and the diagonal slot
    r.yb1 = 64 * bx / kt + 32 / kt + 1;
  r.yu0 = r.yu1 = 0;
  if (64 * bx < up && up < K) {             // upper pairs with tile P's functions
    r.yu0 = up / kt;
    r.yu1 = nkt;
  }
  r.x0 = bx;
  r.x1 = nbt;
  if (up < K) {
    if (k < up) r.x1 = nbt - 1;            // tile P takes no kets below its block
    else r.x0 = 0;                         // every full tile takes tile P's functions
  }
  return r;
}

// Sum of the two-sided partials of function k, slot v: bra partials at
// bpart[((y * ts + z) * nacc + v) * K + k], ket partials at
// kpart[((x * ts + z) * K + k) * nacc + v], over tri_ranges.  Part q of np
// (np lanes share one function) sums every np-th partial from the q-th on;
// the caller adds the np part sums in order q = 0, 1, ...: a fixed order.
// With a split tile list (xr not null), bra tile bx's split tiles have their
// second halves' bra partials at xpart[((e * ts + z) * nacc + v) * K + k],
// e in [xr[2 bx], xr[2 bx + 1]).
__device__ __forceinline__ double tile_reduce2_part(const double *__restrict__ bpart,
                                                    const double *__restrict__ kpart, int K,
                                                    int kt, int ts, int nacc, int v, int k,
                                                    int q, int np,
                                                    const double *__restrict__ xpart,
                                                    const int *__restrict__ xr, int fold) {
  const TriRanges r = tri_ranges(K, kt, k, fold);
  double s = 0.0;
  for (int t = r.yb0 * ts + q; t < r.yb1 * ts; t += np)
    s += bpart[((size_t)t * nacc + v) * K + k];
  for (int t = r.yu0 * ts + q; t < r.yu1 * ts; t += np)
    s += bpart[((size_t)t * nacc + v) * K + k];
  if (xr) {                                 // second halves of split tiles
    const int bx = k >> 6;
    for (int t = xr[2 * bx] * ts + q; t < xr[2 * bx + 1] * ts; t += np)
      s += xpart[((size_t)t * nacc + v) * K + k];
  }
  for (int t = r.x0 * ts + q; t < r.x1 * ts; t += np)
    s += kpart[((size_t)t * K + k) * nacc + v];
  return s;
}

// The reduction of slot v for the 64 functions [k0, k0 + 64) by the 4 waves
// of a 256-lane workgroup: wave q sums part q, wave 0 adds the parts.
__device__ __forceinline__ void tile_reduce2_block(const double *__restrict__ bpart,
                                                   const double *__restrict__ kpart, int K,
                                                   int kt, int ts, int nacc, int v, int k0,
                                                   double *__restrict__ red, double *sh,
                                                   const double *__restrict__ xpart,
                                                   const int *__restrict__ xr, int fold) {
  const int lane = threadIdx.x & 63, q = threadIdx.x >> 6;
  const int k = k0 + lane;
  sh[threadIdx.x] =
      k < K ? tile_reduce2_part(bpart, kpart, K, kt, ts, nacc, v, k, q, 4, xpart, xr, fold)
            : 0.0;
  __syncthreads();
  if (q == 0 && k < K)
    red[(size_t)v * K + k] = ((sh[lane] + sh[64 + lane]) + sh[128 + lane]) + sh[192 + lane];
}

}  // namespace necg
