// necg_pair.hpp -- the per (bra, ket, symmetry term) FP64 algebra of the ECG
// energy path, written for one lane of a CDNA4 wavefront.
//
// Notation (n-space: every 3n x 3n operator kron(X, I3) of the reference acts
// as the n x n matrix X on an n x 3 matrix whose row i is the xyz of
// pseudo-particle i):
//   A = L L^T (bra), A' = T^T A_l T and S' = sms s_l (ket under term P,
//   python/matel.py:76-87), B = A + A', E = A S + A' S', st = B^-1 E,
//   dif1 = st - S, dif2 = st - S', M = inverse-mass tensor, q = charges.
//
// Elements (python/matel.py:92-208):
//   ov  = pk pl det(B)^-3/2 exp(i3 - i1 - i2)            (:125-136)
//   tk  = 2 ov (2 <A dif1, M A' dif2> + 3 tr(M A B^-1 A')) (:144-155)
//   ert = 2 ov sum_{p<q} q_p q_q F0(|st^T w|^2 / w^T B^-1 w) / sqrt(pi w^T B^-1 w)
//         with w = e_q (p = 0) or e_p - e_q                    (:162-197)
//   zht = ov efs sum_i q_i st[i, z]                            (:205-208)
// so that tk + ert - zht = ov * alpha with
//   alpha = 2 tfac + 2 sum c f r0 - efs zsum,  tfac = 2 i3k + T2.
//
// Bra gradient (python/gradel.py:108-390).  Every term of gradel has the form
// vech(X L) for the vech(L) block and A Y for the shift block, so the gradient
// of one term is carried as a symmetric n x n X and an n x 3 Y:
//   S part:  X_S = ov (1.5 A^-1 + X0),  Y_S = 2 ov dif1,
//            X0  = -3 B^-1 - 2 dif1 dif1^T          (ogl :114-148, ogs :108)
//   H part:  X_H = alpha ov (1.5 A^-1 + X0) + ov B^-1 Gam B^-1 - (dif1 Phi^T + Phi dif1^T)
//            Y_H = 2 alpha ov dif1 + Phi
//            Gam = 12 A' M A' + La,  Phi = ov B^-1 F^T,
//            F^T = 4 A' M (A dif1 - A' dif2) + Lb st - efs q e_z^T
//   where La = sum_c c (2 f + 4 f' z) r0 / i2pq  w w^T and Lb = sum_c c 4 f' r0 / i2pq w w^T.
// These identities follow from gradel's own terms: the six kcm products
// (:233-254) collapse to -V dif1^T with V = B^-1 A' M (A dif1 - A' dif2), the
// two "6 ov" products (:189-197) to 6 B^-1 A' M A' B^-1, the Coulomb rank-one
// updates (:296-335) to B^-1 La B^-1 and dif1 st^T Lb B^-1 (+ transpose), and
// the Stark updates (:346-370) to the efs q e_z^T column of F.  The 1.5 A^-1
// part is kept as a scalar: vech(A^-1 L) = vech(L^-T) = diag(1 / L_ii).
#pragma once

#include <hip/hip_runtime.h>

namespace necg {

constexpr double kPi = 3.14159265358979311599796346854;  // python/shared.py:71
constexpr double kZcut = 16.3578;                         // python/utils.py:85
// Pade constants of F0 (python/shared.py:76-86)
constexpr double kC1 = 0.564691197e-04, kC2 = 0.758433197e-03, kC3 = 0.769838037e-02,
                 kC4 = 0.629344460e-01, kC5 = 0.213271302e+00, kC6 = 0.720266520e-04,
                 kC7 = 0.955528842e-03, kC8 = 0.101431553e-01, kC9 = 0.738522953e-01,
                 kC10 = 0.338450368e+00, kC11 = 0.879937801e+00;

constexpr int kMaxN = 6;


__host__ __device__ constexpr int sx(int i, int j) {
  return i >= j ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i;
}

template <int N>
struct Dim {
  static constexpr int NS = N * (N + 1) / 2;  // packed symmetric / vech length
  static constexpr int N3 = 3 * N;            // shift block length
  static constexpr int BLK = NS + N3;         // parameters per basis function
  static constexpr int NP = N * (N + 1) / 2;  // Coulomb pairs incl. reference
  // ket-table record for one (ket, term): A' | S' | A'S' | A'MA' | i2 pl wc sc | A'M |
  // tr(M A'), padded to whole 128-byte lines; the two-sided term reads the
  // fields before A'M only (4 lines for n = 5)
  static constexpr int K_AP = 0;
  static constexpr int K_SP = K_AP + NS;
  static constexpr int K_APSP = K_SP + N3;
  static constexpr int K_NP = K_APSP + N3;
  static constexpr int K_I2 = K_NP + NS;
  static constexpr int K_PL = K_I2 + 1;
  static constexpr int K_WC = K_PL + 1;  // scvv_P * C_l
  static constexpr int K_SC = K_WC + 1;  // scvv_P
  static constexpr int K_APM = K_SC + 1;
  static constexpr int K_TRMP = K_APM + N * N;  // tr(M A')
  static constexpr int KS = (K_TRMP + 1 + 15) & ~15;
  // bra-table record: A | S | AS | i1 pk C | L (vech) | 1/L_ii | A M A | tr(M A)
  static constexpr int B_A = 0;
  static constexpr int B_S = B_A + NS;
  static constexpr int B_AS = B_S + N3;
  static constexpr int B_I1 = B_AS + N3;
  static constexpr int B_PK = B_I1 + 1;
  static constexpr int B_C = B_PK + 1;
  static constexpr int B_L = B_C + 1;
  static constexpr int B_LD = B_L + NS;
  static constexpr int B_AMA = B_LD + N;
  static constexpr int B_TRMA = B_AMA + NS;
  static constexpr int BS = (B_TRMA + 1 + 1) & ~1;
  // accumulator record of one bra: X_H Y_H X_S Y_S sigma_H sigma_S
  static constexpr int A_XH = 0;
  static constexpr int A_YH = A_XH + NS;
  static constexpr int A_XS = A_YH + N3;
  static constexpr int A_YS = A_XS + NS;
  static constexpr int A_SH = A_YS + N3;
  static constexpr int A_SS = A_SH + 1;
  static constexpr int NACC = A_SS + 1;
};

// Kernel-argument block: the physics constants of one system.
struct SysConst {
  int n, nst;
  double efs;
  double mass[kMaxN * kMaxN];             // row-major (ket-table products)
  double msym[kMaxN * (kMaxN + 1) / 2];   // packed symmetric copy of mass
  double q[kMaxN + 1];                    // ncv
  double cpair[kMaxN * (kMaxN + 1) / 2];  // q_p q_q in pair order
  double cpair2[kMaxN * (kMaxN + 1) / 2]; // 2 q_p q_q / sqrt(pi) in pair order
  // mass = diag(mdiag) + mc 1 1^T when every off-diagonal entry equals mc
  // (relative coordinates about a reference particle); then mrank1 = 1
  int mrank1;
  double mc;
  double mdiag[kMaxN];
};

// The mass constants of SysConst from the row-major n x n mass matrix.
inline void set_mass(SysConst &sc, const double *mass, int n) {
  for (int i = 0; i < n * n; ++i) sc.mass[i] = mass[i];
  for (int i = 0; i < n; ++i)
    for (int j = 0; j <= i; ++j) sc.msym[sx(i, j)] = mass[i * n + j];
  sc.mc = n > 1 ? mass[1] : 0.0;
  sc.mrank1 = 1;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j)
      if (i != j && mass[i * n + j] != sc.mc) sc.mrank1 = 0;
  for (int i = 0; i < n; ++i) sc.mdiag[i] = mass[i * n + i] - sc.mc;
}

// Position of particle pair (p, q), 0 <= p < q <= n (0 the fixed particle),
// in the pair order p = 0, 1, ...; q = p + 1, ..., n.
__host__ __device__ constexpr int pair_index(int p, int q, int n) {
  return p * n - p * (p - 1) / 2 + (q - p - 1);
}

// The pair constants of SysConst from the charges ncv[0..n].
inline void set_pair_charges(SysConst &sc, const double *ncv, int n) {
  for (int p = 0; p <= n; ++p)
    for (int q = p + 1; q <= n; ++q) {
      const double c = ncv[p] * ncv[q];
      sc.cpair[pair_index(p, q, n)] = c;
      sc.cpair2[pair_index(p, q, n)] = c * (2.0 * 0.56418958354775628695);
    }
}

// 1/x: hardware seed + two Newton steps (relative error ~1 ulp for the
// positive normal arguments that reach here).
__device__ __forceinline__ double rcp(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = __builtin_fma(-x, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-x, r, 1.0);
  return __builtin_fma(r, e, r);
}

// 1/sqrt(x) for positive normal x: the hardware seed y (relative error
// below 2^-25) and one third-order step from the series of (1 - e)^-1/2,
// e = 1 - x y^2: y (1 + e/2 + 3 e^2/8).  The truncation error is O(e^3),
// far below rounding; over 4M log-uniform x in [e^-20, e^20] the result is
// within 0.994 ulp of 1/sqrt(x) (two Newton steps: 0.99 ulp, one: 20 ulp).
__device__ __forceinline__ double rsq(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  const double e = __builtin_fma(-x * y, y, 1.0);
  return __builtin_fma(y, e * __builtin_fma(0.375, e, 0.5), y);
}

constexpr double kSqrtPiHalf = 0.88622692545275801365;   // sqrt(pi)/2
constexpr double kInvSqrtPi = 0.56418958354775628695;    // 1/sqrt(pi)

// F0 and 2 dF0/dz (python/utils.py:78-106), written so that the
// 15 evaluations of a term form one straight-line block: the Pade part is
// evaluated at min(z, 16.3578) (finite, and not used above the cut), each
// Horner pass yields the polynomial and its derivative together (only the
// reference's 11 constants appear, as addends), and the branch is a select
// of factors.  v = (P Q^3)^-1/2 below the cut gives F0 = P Q v and
// 2 dF0 = (Q P' - P Q') v; above it v = z^-1/2, F0 = sqrt(pi)/2 v and
// 2 dF0 = -F0 v^2.  FENCE: an empty asm on the derivative's numerator, so
// that the compiler forms it on every lane rather than branching around it.
// The per-ordering term runs faster with it (LiH K=2048: 9.87 ms against
// 10.33 ms), the two-sided term without it (6.55 ms against 6.79 ms), where
// the whole term then schedules as one block.
template <bool FENCE>
__device__ __forceinline__ void boys_f0_line(double z, double &f, double &f2p) {
  const bool big = z > kZcut;
  const double zc = __builtin_fmin(z, kZcut);
  double P = kC1, dP = kC1;
  P = __builtin_fma(P, zc, kC2);
  dP = __builtin_fma(dP, zc, P);
  P = __builtin_fma(P, zc, kC3);
  dP = __builtin_fma(dP, zc, P);
  P = __builtin_fma(P, zc, kC4);
  dP = __builtin_fma(dP, zc, P);
  P = __builtin_fma(P, zc, kC5);
  dP = __builtin_fma(dP, zc, P);
  P = __builtin_fma(P, zc, 1.0);
  double Q = kC6, dQ = kC6;
  Q = __builtin_fma(Q, zc, kC7);
  dQ = __builtin_fma(dQ, zc, Q);
  Q = __builtin_fma(Q, zc, kC8);
  dQ = __builtin_fma(dQ, zc, Q);
  Q = __builtin_fma(Q, zc, kC9);
  dQ = __builtin_fma(dQ, zc, Q);
  Q = __builtin_fma(Q, zc, kC10);
  dQ = __builtin_fma(dQ, zc, Q);
  Q = __builtin_fma(Q, zc, kC11);
  dQ = __builtin_fma(dQ, zc, Q);
  Q = __builtin_fma(Q, zc, 1.0);
  const double PQ = P * Q;
  const double v = rsq(big ? z : PQ * (Q * Q));
  f = (big ? kSqrtPiHalf : PQ) * v;
  double num = __builtin_fma(Q, dP, -P * dQ);
  if constexpr (FENCE) asm volatile("" : "+v"(num));
  f2p = (big ? -f * v : num) * v;
}

// F0 alone (python/utils.py:78-90) for the element modes: with
// v = (P Q^3)^-1/2 below the cut F0 = P Q v, above it F0 = sqrt(pi)/2 z^-1/2.
__device__ __forceinline__ double boys_f0_value(double z) {
  const bool big = z > kZcut;
  const double zc = __builtin_fmin(z, kZcut);
  const double P = __builtin_fma(__builtin_fma(__builtin_fma(__builtin_fma(__builtin_fma(kC1, zc, kC2), zc, kC3), zc, kC4), zc, kC5), zc, 1.0);
  const double Q = __builtin_fma(__builtin_fma(__builtin_fma(__builtin_fma(__builtin_fma(__builtin_fma(kC6, zc, kC7), zc, kC8), zc, kC9), zc, kC10), zc, kC11), zc, 1.0);
  const double PQ = P * Q;
  return (big ? kSqrtPiHalf : PQ) * rsq(big ? z : PQ * (Q * Q));
}

// Bra-side data of one basis function, held in registers.
template <int N>
struct BraReg {
  double a[Dim<N>::B_PK + 1];
  __device__ __forceinline__ double A(int t) const { return a[Dim<N>::B_A + t]; }
  __device__ __forceinline__ double S(int t) const { return a[Dim<N>::B_S + t]; }
  __device__ __forceinline__ double AS(int t) const { return a[Dim<N>::B_AS + t]; }
  __device__ __forceinline__ double i1() const { return a[Dim<N>::B_I1]; }
  __device__ __forceinline__ double pk() const { return a[Dim<N>::B_PK]; }
};

template <int N>
__device__ __forceinline__ void load_bra(BraReg<N> &b, const double *__restrict__ rec) {
#pragma unroll
  for (int t = 0; t <= Dim<N>::B_PK; ++t) b.a[t] = rec[t];
}


// Accumulator sinks.  Slot layout (Dim<N>::A_*): X_H | Y_H | X_S | Y_S | sH | sS.
// MODE 0 keeps the H and S parts apart (E not yet known); MODE 1 folds them
// with the known energy (X = X_H - E X_S, ...) into the H slots.
template <int N>
struct RegSink {  // per-lane registers (query kernels)
  double a[Dim<N>::NACC];
  __device__ __forceinline__ void zero() {
#pragma unroll
    for (int t = 0; t < Dim<N>::NACC; ++t) a[t] = 0.0;
  }
  __device__ __forceinline__ void add(int t, double v) { a[t] += v; }
};

struct LdsSink {  // this lane's slots in LDS, stride 64 lanes (pair kernels)
  double *p;
  // ds_add_f64: a read-modify-write in the LDS unit that holds no register
  // for the old value; the slots are private to the lane
  __device__ __forceinline__ void add(int t, double v) {
    __hip_atomic_fetch_add(p + 64 * t, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
  }
};

// Separate element parts for the per-term query entry points.
struct Parts {
  double ov, tk, ert, zht;
};

// MODE_SPLIT: elements + gradient with H and S parts apart; MODE_COMBINED:
// gradient with the energy folded in; MODE_ELEMENTS: H and S only;
// MODE_TV: the kinetic and potential matrices (python/necg.py:43-50) in place
// of H and S.
enum { MODE_SPLIT = 0, MODE_COMBINED = 1, MODE_ELEMENTS = 2, MODE_TV = 3 };

__host__ __device__ constexpr bool mode_has_grad(int mode) {
  return mode == MODE_SPLIT || mode == MODE_COMBINED;
}

// One (bra, ket-under-term) contribution.  kd points at the ket-table record.
// Adds sc*ov*alpha to hel and sc*ov to sel; for the gradient modes adds the
// wc*ov-weighted X/Y parts to the sink (wc = scvv_P C_l; eshift = E in
// MODE 1).  Returns false if B is not positive definite.
//
// Two identities shorten the reference's sequence: A' dif2 = -A dif1 (because
// A dif1 + A' dif2 = B st - E = 0), so the kinetic quadratic is
// <A dif1, M A' dif2> = -<a1, M a1> with a1 = A dif1, and the vector
// A dif1 - A' dif2 of the gradient is 2 a1.
template <int N, int MODE, bool R1 = false, class Bra, class Sink>
__device__ __forceinline__ bool pair_term(const Bra &b, const double *__restrict__ kd,
                                          const SysConst &sc_, double eshift, double &hel,
                                          double &sel, Sink &sink, Parts *parts) {
  using D = Dim<N>;
  constexpr int NS = D::NS;
  const double *__restrict__ Ap = kd + D::K_AP;
  const double *__restrict__ ApSp = kd + D::K_APSP;
  const double *__restrict__ Np = kd + D::K_NP;
  const double *__restrict__ ApM = kd + D::K_APM;
  const double i2 = kd[D::K_I2], pl = kd[D::K_PL], wc = kd[D::K_WC], scv = kd[D::K_SC];

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
  // R^-1 (lower, packed) and B^-1 = R^-T R^-1 (symmetric, packed).
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
  double Em[3 * N], st[3 * N];
#pragma unroll
  for (int t = 0; t < 3 * N; ++t) Em[t] = b.AS(t) + ApSp[t];
  double i3 = 0.0;
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
  const double ov = b.pk() * pl * det32 * exp(i3 - b.i1() - i2);

  // Kinetic (python/matel.py:144-155).  a1 = A dif1; the gradient needs
  // only u = A'M a1 from it (8 u enters F), formed here so that dif1, a1 and
  // A'M are dead before the Coulomb loop.
  double u1[3 * N];
  double i3k = 0.0;  // -<a1, M a1>
  {
    double d1[3 * N], a1[3 * N];
#pragma unroll
    for (int t = 0; t < 3 * N; ++t) d1[t] = st[t] - b.S(t);
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int x = 0; x < 3; ++x) {
        double v = 0.0;
#pragma unroll
        for (int j = 0; j < N; ++j) v = __builtin_fma(b.A(sx(i, j)), d1[3 * j + x], v);
        a1[3 * i + x] = v;
      }
    double cs[3];  // R1: the constant part of M a1
#pragma unroll
    for (int x = 0; x < 3; ++x) {
      double t = a1[x];
#pragma unroll
      for (int j = 1; j < N; ++j) t += a1[3 * j + x];
      cs[x] = sc_.mc * t;
    }
    (void)cs;
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int x = 0; x < 3; ++x) {
        double t = 0.0, v = 0.0;
        if constexpr (R1) t = __builtin_fma(sc_.mdiag[i], a1[3 * i + x], cs[x]);
#pragma unroll
        for (int j = 0; j < N; ++j) {
          if constexpr (!R1) t = __builtin_fma(sc_.msym[sx(i, j)], a1[3 * j + x], t);
          if (mode_has_grad(MODE)) v = __builtin_fma(ApM[i * N + j], a1[3 * j + x], v);
        }
        i3k = __builtin_fma(-a1[3 * i + x], t, i3k);
        u1[3 * i + x] = v;
      }
  }
  // 3 tr(M A B^-1 A') = 3 (tr(M A') - <A'MA', B^-1>), since A B^-1 = I - A' B^-1
  double ab = 0.0;
  double abo = 0.0;  // the strictly lower part, counted twice below
#pragma unroll
  for (int i = 0; i < N; ++i) {
    ab = __builtin_fma(Np[sx(i, i)], Bi[sx(i, i)], ab);
#pragma unroll
    for (int j = 0; j < i; ++j) abo = __builtin_fma(Np[sx(i, j)], Bi[sx(i, j)], abo);
  }
  ab = __builtin_fma(2.0, abo, ab);
  const double tfac = 2.0 * i3k + 3.0 * (kd[D::K_TRMP] - ab);

  // Coulomb (python/matel.py:162-197; gradients python/gradel.py:266-337).
  // Per particle pair (p, q): with h = 2 c_pq ri^3 / sqrt(pi), the
  // coefficients of d(i2p)/dB^-1 are ca = f h + 2 F0' h z (La) and
  // cb = 2 F0' h (Lb); the diagonal entry of particle a sums the pairs that
  // contain a (the first, (0, a), sets it), the off-diagonal entry (a, b) is
  // minus the pair's value.
  double csum2 = 0.0;  // twice sum_c c f r0
  // La and Gl = Lb st (Lb: the pair's cb on both diagonal entries, -cb off the
  // diagonal; a particle pair (a, b) adds cb d and -cb d, d = st_a - st_b, to
  // rows a and b of Lb st, and a pair (0, b) sets row b to cb st_b)
  double La[NS], Gl[3 * N];
  (void)La;
  (void)Gl;
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
        const double ri = rsq(i2p);  // 1/sqrt(i2p)
        const double inv2 = ri * ri;
        const double z = __builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx)) * inv2;
        const double cr = sc_.cpair2[pi] * ri;  // 2 c / sqrt(pi i2p)
        if (mode_has_grad(MODE)) {
          double f, f2p;
          boys_f0_line<true>(z, f, f2p);
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
        } else {
          csum2 = __builtin_fma(cr, boys_f0_value(z), csum2);
        }
      }
  }
  // Stark (python/matel.py:205-208)
  double zsum = 0.0;
#pragma unroll
  for (int i = 0; i < N; ++i) zsum = __builtin_fma(sc_.q[i + 1], st[3 * i + 2], zsum);
  const double alpha = 2.0 * tfac + csum2 - sc_.efs * zsum;
  if (MODE == MODE_TV) {
    // Kinetic += scvv tk, Potential += scvv ert (the Stark term is in H only)
    hel = __builtin_fma(scv * ov, 2.0 * tfac, hel);
    sel = __builtin_fma(scv * ov, csum2, sel);
  } else {
    hel = __builtin_fma(scv * ov, alpha, hel);
    sel = __builtin_fma(scv, ov, sel);
  }
  if (parts) {
    parts->ov = ov;
    parts->tk = 2.0 * ov * tfac;
    parts->ert = ov * csum2;
    parts->zht = ov * zsum * sc_.efs;
  }
  if (!mode_has_grad(MODE)) return ok;

  // Gradient.  F^T = 4 A'M (A dif1 - A' dif2) + Lb st - efs q e_z^T
  //                = 8 A'M a1 + Lb st - efs q e_z^T;  Phi = B^-1 F^T.
  double Ft[3 * N];
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int x = 0; x < 3; ++x) {
      double t = __builtin_fma(8.0, u1[3 * i + x], Gl[3 * i + x]);
      if (x == 2) t = __builtin_fma(-sc_.efs, sc_.q[i + 1], t);
      Ft[3 * i + x] = t;
    }
  double Phi[3 * N];
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int x = 0; x < 3; ++x) {
      double t = 0.0;
#pragma unroll
      for (int m = 0; m < N; ++m) t = __builtin_fma(Bi[sx(i, m)], Ft[3 * m + x], t);
      Phi[3 * i + x] = t;
    }
  // G = B^-1 U B^-1 with U = 12 A'MA' + La (symmetric), row by row as
  // (B^-1 U)_i B^-1.  With e = 2 wo dif1 (the Y_S values) the slots are
  //   X_S = -3 wo B^-1 - sum_x e_i d_j,   X_H = am X_S + wo (G - dp),
  //   Y_H = am e + wo Phi,                Y_S = e,
  // dp_ij = <d_i, Phi_j> + <Phi_i, d_j>.
  double d1[3 * N], e1[3 * N];
  const double wo = wc * ov;
  const double wo2 = 2.0 * wo;
#pragma unroll
  for (int t = 0; t < 3 * N; ++t) {
    d1[t] = st[t] - b.S(t);
    e1[t] = wo2 * d1[t];
  }
  const double am = (MODE == MODE_COMBINED) ? (alpha - eshift) : alpha;
  const double woa = wo * am;
  double U[Dim<N>::NS];
#pragma unroll
  for (int t = 0; t < Dim<N>::NS; ++t) U[t] = __builtin_fma(12.0, Np[t], La[t]);
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
      sink.add(D::A_XH + t, __builtin_fma(am, xs, wo * g));
      if (MODE == MODE_SPLIT) sink.add(D::A_XS + t, xs);
    }
  }
#pragma unroll
  for (int t = 0; t < 3 * N; ++t) {
    sink.add(D::A_YH + t, __builtin_fma(am, e1[t], wo * Phi[t]));
    if (MODE == MODE_SPLIT) sink.add(D::A_YS + t, e1[t]);
  }
  sink.add(D::A_SH, woa);
  if (MODE == MODE_SPLIT) sink.add(D::A_SS, wo);
  return ok;
}

// latched device errors; necg_status reports the root cause first:
// bad input, then a singular pair matrix, then C^T S C <= 0, then NaN results
enum FlagBits { F_NOT_PD = 1, F_NONFINITE = 2, F_CSC = 4, F_BAD_INPUT = 8 };

// Waves per pair-kernel workgroup: they work on the same 64 bras and each
// owns a private set of accumulator slots in LDS, so the sums are formed in a
// fixed order (bitwise reproducible for a given launch shape).  LDS per
// workgroup: W*NACC*512 B = 124 KB for n = 5, W = 4.
__host__ __device__ constexpr int pair_waves(int n) { return n <= 5 ? 4 : 2; }

template <int N, int MODE>
struct PairLds {
  static constexpr int NA = mode_has_grad(MODE) ? Dim<N>::NACC * 64 : 1;
  static constexpr int ACC = pair_waves(N) * NA;
};

// One pair tile: the 64 bras [k0, k0+64) x kets [l0, l1) x terms [p0, p1),
// run by the pair_waves(N) waves of a workgroup.  Each lane holds its bra
// record in registers; each wave owns a set of accumulator slots in LDS
// ([slot][lane]) and takes alternate kets; ket records are wave-uniform
// (scalar loads, or LDS broadcasts through a staging buffer with STAGE: the
// per-ordering pass of LiH K = 2048 takes 8.65 ms staged, 9.91 ms not).
// Writes the element partials of bra >= ket to hp/sp ([l][k], K x K) and the
// tile's summed accumulators to pt ([slot][k], NACC x K).  Kets l >= lup
// (element modes of the lower-triangle pass: the last, partial bra tile's
// functions) pair with every bra of the tile, written at [k][l].
template <int N, int MODE, bool R1 = false, bool STAGE = false>
__device__ __forceinline__ void pair_tile(const double *__restrict__ bra,
                                          const double *__restrict__ ket, int K, int nst, int k0,
                                          int l0, int l1, int p0, int p1, const SysConst &sc,
                                          double eshift, double *__restrict__ pt,
                                          double *__restrict__ hp, double *__restrict__ sp,
                                          int *__restrict__ flag, double *acc_s,
                                          int lup = 0x7fffffff, double *kst_s = nullptr) {
  using D = Dim<N>;
  constexpr int W = pair_waves(N);
  constexpr int NA = PairLds<N, MODE>::NA;
  const int lane = threadIdx.x & 63;
  // wave index: uniform across the wave, so that the ket index and every
  // ket-record address are scalar (s_load into SGPRs)
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int k = k0 + lane;
  const bool active = k < K;
  const int kk = active ? k : K - 1;
  if (mode_has_grad(MODE))
    for (int v = 0; v < D::NACC; ++v) acc_s[wv * NA + v * 64 + lane] = 0.0;
  __syncthreads();
  // The bra record is held in registers for the whole tile.  (Re-reading it
  // from an LDS table for every term, interleaved with the LDS accumulator
  // adds and the scalar ket loads, gave wrong element values in the gradient
  // modes on gfx950; that layout is not used.)
  BraReg<N> b;
  load_bra<N>(b, bra + (size_t)kk * D::BS);
  LdsSink sink{acc_s + wv * NA + lane};
  bool ok = true;
  // STAGE: the ket record of each (ket, term) comes through the wave's LDS
  // buffer kst_s (KS doubles per wave), loaded by the lanes (fields lane and
  // 64 + lane) while the previous term runs and stored after it
  double *kst = STAGE ? kst_s + wv * D::KS : nullptr;
  const int f0 = lane < D::KS ? lane : D::KS - 1;
  const int f1 = 64 + lane < D::KS ? 64 + lane : D::KS - 1;
  if (STAGE && l0 + wv < l1) {
    const double *r = ket + ((size_t)(l0 + wv) * nst + p0) * D::KS;
    const double v0 = r[f0], v1 = r[f1];
    kst[f0] = v0;
    kst[f1] = v1;
  }
  for (int l = l0 + wv; l < l1; l += W) {
    double hel = 0.0, sel = 0.0;
    const double *kl = ket + (size_t)l * nst * D::KS;
    for (int p = p0; p < p1; ++p) {
      if constexpr (STAGE) {
        // the record after (l, p): (l, p + 1), else the wave's next ket at p0
        // (past the wave's last record: its first record again, unused)
        const int pn = p + 1 < p1 ? p + 1 : p0;
        const int ln0 = p + 1 < p1 ? l : l + W;
        const int ln = ln0 < l1 ? ln0 : l0 + wv;
        const double *r = ket + ((size_t)ln * nst + (ln0 < l1 ? pn : p0)) * D::KS;
        const double v0 = r[f0], v1 = r[f1];
        ok &= pair_term<N, MODE, R1>(b, kst, sc, eshift, hel, sel, sink, nullptr);
        kst[f0] = v0;   // after every read of the term (the wave's LDS operations
        kst[f1] = v1;   // complete in order)
      } else {
        ok &= pair_term<N, MODE, R1>(b, kl + (size_t)p * D::KS, sc, eshift, hel, sel, sink,
                                     nullptr);
      }
    }
    if (MODE != MODE_COMBINED && active && (k >= l || l >= lup)) {
      const size_t o = (l >= lup && k < l) ? (size_t)k * K + l : (size_t)l * K + k;
      hp[o] = hel;
      sp[o] = sel;
    }
  }
  if (active && !ok) atomicOr(flag, (int)F_NOT_PD);
  __syncthreads();
  if (!mode_has_grad(MODE) || !active) return;
  double *o = pt + k;
  for (int v = wv; v < D::NACC; v += W) {
    if (MODE == MODE_COMBINED && v >= D::A_XS && v != D::A_SH) continue;
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < W; ++w) t += acc_s[w * NA + v * 64 + lane];
    o[(size_t)v * K] = t;
  }
}

// vech(X L) + 1.5 s diag(1/L_ii) for symmetric X (packed) and lower L (vech
// order, column-major): the vech(L) block of one gradient.
template <int N>
__device__ __forceinline__ void grad_L(const double *X, double s, const double *Lv,
                                       const double *Ld, double *out) {
  // Lv is vech order: column j, rows j..N-1.
  int k = 0;
#pragma unroll
  for (int j = 0; j < N; ++j) {
    const int c0 = j * N - j * (j - 1) / 2;  // start of column j in vech order
#pragma unroll
    for (int i = j; i < N; ++i, ++k) {
      double t = 0.0;
#pragma unroll
      for (int m = j; m < N; ++m) t = __builtin_fma(X[sx(i, m)], Lv[c0 + (m - j)], t);
      if (i == j) t = __builtin_fma(1.5 * s, Ld[j], t);
      out[k] = t;
    }
  }
}

// A Y for symmetric A (packed) and Y (n x 3): the shift block of one gradient.
template <int N>
__device__ __forceinline__ void grad_S(const double *A, const double *Y, double *out) {
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int x = 0; x < 3; ++x) {
      double t = 0.0;
#pragma unroll
      for (int m = 0; m < N; ++m) t = __builtin_fma(A[sx(i, m)], Y[3 * m + x], t);
      out[3 * i + x] = t;
    }
}

}  // namespace necg
