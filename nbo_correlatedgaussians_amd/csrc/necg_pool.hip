// necg_pool.hip -- single-function replacement trials against resident
// systems (serving).  A pool references the resident state of several
// contexts, or of every system of a batch (bra tables, H, S, HC, SC, cHc, cSc
// of their last evaluation).  A
// trial replaces function k of system s by new parameters: its bra and ket
// records are built, the new row of H and S is formed (bra j of the resident
// table against the new ket, the diagonal from the new bra; the matel loop of
// python/outprod.py:49-54 restricted to one row), and the Rayleigh quotient
// with the resident C is updated in O(K):
//   C^T H' C = cHc - 2 C_k (HC)_k + C_k^2 H_kk + 2 C_k r.C - C_k^2 r_k.
// All per-call buffers are the caller's, so calls on different streams may
// overlap; each trial's numbers depend only on that trial (bitwise the same
// alone or in a batch).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <new>
#include <vector>

#include "necg_ctx.hpp"
#include "necg_pool_impl.hpp"

using namespace necg;

namespace {

#define POOL_SWITCH_N(n, call)                    \
  switch (n) {                                    \
    case 1: { constexpr int NN = 1; call; break; } \
    case 2: { constexpr int NN = 2; call; break; } \
    case 3: { constexpr int NN = 3; call; break; } \
    case 4: { constexpr int NN = 4; call; break; } \
    case 5: { constexpr int NN = 5; call; break; } \
    default: { constexpr int NN = 6; call; break; } \
  }

// The record of trial t's system and the replaced function k, or null when
// tsys[t] or tfun[t] is out of range (t_prep latches the bad-input status).
__device__ __forceinline__ const PoolSys *trial_sys(const PoolSys *__restrict__ sys, int nsys,
                                                    const int32_t *__restrict__ tsys,
                                                    const int32_t *__restrict__ tfun, int t,
                                                    int &k) {
  const int s = tsys[t];
  if (s < 0 || s >= nsys) return nullptr;
  k = tfun[t];
  if (k < 0 || k >= sys[s].K) return nullptr;
  return sys + s;
}

// Workgroup per trial: the new bra record, then its ket record under every term.
__global__ void __launch_bounds__(256) t_prep(const PoolSys *__restrict__ sys,
                                              const int32_t *__restrict__ tsys,
                                              const int32_t *__restrict__ tfun, int nsys,
                                              const double *__restrict__ newp, int pstride,
                                              double *__restrict__ ws, long long wstride,
                                              int *__restrict__ flag) {
  const int t = blockIdx.x;
  int k = 0;
  const PoolSys *bp = trial_sys(sys, nsys, tsys, tfun, t, k);
  if (!bp) {
    if (threadIdx.x == 0) atomicOr(flag, (int)F_BAD_INPUT);
    return;
  }
  const PoolSys &b = *bp;
  double *w = ws + (size_t)t * wstride;
  if (threadIdx.x == 0) {
    POOL_SWITCH_N(b.n, (make_bra<NN>(newp + (size_t)t * pstride,
                                     b.bra[(size_t)k * Dim<NN>::BS + Dim<NN>::B_C], true, b.sc,
                                     w, flag)));
  }
  __syncthreads();
  for (int p = threadIdx.x; p < b.nst; p += blockDim.x) {
    POOL_SWITCH_N(b.n, (make_ket<NN>(w, b.sml + (size_t)p * NN * NN,
                                     b.sms + (size_t)p * 9 * NN * NN, b.scvv[p], b.sc,
                                     w + Dim<NN>::BS + (size_t)p * Dim<NN>::KS)));
  }
}

template <int N>
__device__ __forceinline__ void trial_row(const PoolSys &b, int k, const double *__restrict__ w,
                                          int j, double *rh, double *rs, int *flag) {
  using D = Dim<N>;
  BraReg<N> br;
  load_bra<N>(br, j == k ? w : b.bra + (size_t)j * D::BS);
  RegSink<N> none;
  double h = 0.0, s = 0.0;
  bool ok = true;
  if (b.sc.mrank1) {
    for (int p = 0; p < b.nst; ++p)
      ok &= pair_term<N, MODE_ELEMENTS, true>(br, w + D::BS + (size_t)p * D::KS, b.sc, 0.0, h, s,
                                              none, nullptr);
  } else {
    for (int p = 0; p < b.nst; ++p)
      ok &= pair_term<N, MODE_ELEMENTS, false>(br, w + D::BS + (size_t)p * D::KS, b.sc, 0.0, h, s,
                                               none, nullptr);
  }
  if (!ok) atomicOr(flag, (int)F_NOT_PD);
  *rh = h;
  *rs = s;
}

// Grid (ceil(Kmax/64), ntr), 64 lanes: lane j = bra j of the new row.
__global__ void __launch_bounds__(64) t_rows(const PoolSys *__restrict__ sys,
                                             const int32_t *__restrict__ tsys,
                                             const int32_t *__restrict__ tfun, int nsys,
                                             const double *__restrict__ ws, long long wstride,
                                             double *__restrict__ rowH, double *__restrict__ rowS,
                                             int rstride, int *__restrict__ flag) {
  const int t = blockIdx.y;
  int k = 0;
  const PoolSys *bp = trial_sys(sys, nsys, tsys, tfun, t, k);
  if (!bp) return;
  const PoolSys &b = *bp;
  const int j = blockIdx.x * 64 + threadIdx.x;
  if (j >= b.K) return;
  double *rh = rowH + (size_t)t * rstride + j, *rs = rowS + (size_t)t * rstride + j;
  POOL_SWITCH_N(b.n, trial_row<NN>(b, k, ws + (size_t)t * wstride, j, rh, rs, flag));
}

// Workgroup of 256 per trial: the updated quotient.
__global__ void __launch_bounds__(256) t_energy(const PoolSys *__restrict__ sys,
                                                const int32_t *__restrict__ tsys,
                                                const int32_t *__restrict__ tfun, int nsys,
                                                const double *__restrict__ rowH,
                                                const double *__restrict__ rowS, int rstride,
                                                double *__restrict__ E, int *__restrict__ flag) {
  __shared__ double red[8];
  const int t = blockIdx.x;
  int k = 0;
  const PoolSys *bp = trial_sys(sys, nsys, tsys, tfun, t, k);
  if (!bp) {
    if (threadIdx.x == 0) E[t] = __builtin_nan("");
    return;
  }
  const PoolSys &b = *bp;
  int bs = 0, bc = 0;
  POOL_SWITCH_N(b.n, (bs = Dim<NN>::BS, bc = Dim<NN>::B_C));
  const double *rh = rowH + (size_t)t * rstride, *rs = rowS + (size_t)t * rstride;
  double a = 0.0, c = 0.0;
  for (int j = threadIdx.x; j < b.K; j += 256) {
    const double cj = b.bra[(size_t)j * bs + bc];
    a = __builtin_fma(rh[j], cj, a);
    c = __builtin_fma(rs[j], cj, c);
  }
  a = block_sum256(a, red);
  c = block_sum256(c, red + 4);
  if (threadIdx.x == 0) {
    const double ck = b.bra[(size_t)k * bs + bc];
    const double hkk = b.H[(size_t)k * b.K + k], skk = b.S[(size_t)k * b.K + k];
    const double cHc = b.scal[2] - 2.0 * ck * b.HC[k] + ck * ck * hkk + 2.0 * ck * a -
                       ck * ck * rh[k];
    const double cSc = b.scal[3] - 2.0 * ck * b.SC[k] + ck * ck * skk + 2.0 * ck * c -
                       ck * ck * rs[k];
    E[t] = cHc / cSc;
    if (!(cSc > 0.0)) atomicOr(flag, (int)F_CSC);
  }
}

// v from the DPP partner lane (CTRL: 0xB1 quad_perm [1,0,3,2], 0x4E
// quad_perm [2,3,0,1], 0x141 half-row mirror, 0x140 row mirror); every lane
// has a partner and every row and bank is enabled.
template <int CTRL>
__device__ __forceinline__ double dpp_partner(double v) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = __builtin_amdgcn_mov_dpp((int)(unsigned)u, CTRL, 0xf, 0xf, false);
  const unsigned hi = __builtin_amdgcn_mov_dpp((int)(unsigned)(u >> 32), CTRL, 0xf, 0xf, false);
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// v plus its value in the other 16-lane row of the pair (permlane16 swap,
// W = 16) or the other 32-lane half (permlane32 swap, W = 32).
template <int W>
__device__ __forceinline__ double swap_add(double v) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
  const unsigned l = (unsigned)u, h = (unsigned)(u >> 32);
  unsigned l0, l1, h0, h1;
  if constexpr (W == 16) {
    const auto a = __builtin_amdgcn_permlane16_swap(l, l, false, false);
    const auto b = __builtin_amdgcn_permlane16_swap(h, h, false, false);
    l0 = a[0]; l1 = a[1]; h0 = b[0]; h1 = b[1];
  } else {
    const auto a = __builtin_amdgcn_permlane32_swap(l, l, false, false);
    const auto b = __builtin_amdgcn_permlane32_swap(h, h, false, false);
    l0 = a[0]; l1 = a[1]; h0 = b[0]; h1 = b[1];
  }
  return __builtin_bit_cast(double, ((unsigned long long)h0 << 32) | l0) +
         __builtin_bit_cast(double, ((unsigned long long)h1 << 32) | l1);
}

// Sum of v over the 64 lanes, the same bits in every lane: each step adds
// the partner's partial sum, and a + b = b + a exactly.
__device__ __forceinline__ double wave_allsum(double v) {
  v += dpp_partner<0xB1>(v);
  v += dpp_partner<0x4E>(v);
  v += dpp_partner<0x141>(v);
  v += dpp_partner<0x140>(v);
  v = swap_add<16>(v);
  return swap_add<32>(v);
}

// Sum of four values over a 256-lane workgroup in a fixed order; every lane
// receives the totals.  red: 32 doubles, used in two halves by turns (buf),
// so one barrier per call suffices: a half is written again only two calls
// later, after every lane has passed the barrier of the call between.
__device__ __forceinline__ void block_sum4(double v[4], double *red, int &buf) {
  double *r = red + 16 * buf;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const double a = wave_allsum(v[q]);
    if ((threadIdx.x & 63) == 0) r[(threadIdx.x >> 6) * 4 + q] = a;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 4; ++q) v[q] = (r[q] + r[4 + q]) + (r[8 + q] + r[12 + q]);
  buf ^= 1;
}

constexpr int kProjSlices = 8;  // i-range slices of the projection (partial rows)

// Grid (ceil(Kmax/256), ntr, kProjSlices), 256 lanes.  For trial t (function
// k of system s replaced): the changes of row/column k, w = h' - H[:, k] and
// v = s' - S[:, k] with their k-th entries halved (so that
// H' = H + e_k w^T + w e_k^T), projected on the resident factors:
// beta = V^T w, gamma = V^T v, as kProjSlices partial sums over i-ranges.
// bg[t]: for each slice, beta (kmax) then gamma (kmax).
__global__ void __launch_bounds__(256) t_proj(const PoolSys *__restrict__ sys,
                                              const int32_t *__restrict__ tsys,
                                              const int32_t *__restrict__ tfun, int nsys,
                                              const double *__restrict__ rowH,
                                              const double *__restrict__ rowS, int rstride,
                                              double *__restrict__ bg, int kmax) {
  const int t = blockIdx.y;
  int k = 0;
  const PoolSys *bp = trial_sys(sys, nsys, tsys, tfun, t, k);
  if (!bp) return;
  const PoolSys &b = *bp;
  const int K = b.K;
  if (blockIdx.x * 256 >= K || !b.V) return;  // uniform over the workgroup
  const int j = blockIdx.x * 256 + threadIdx.x;
  const int per = (K + kProjSlices - 1) / kProjSlices;
  const int i0 = blockIdx.z * per, i1 = min(K, i0 + per);
  const double *rh = rowH + (size_t)t * rstride, *rs = rowS + (size_t)t * rstride;
  const double *hk = b.H + (size_t)k * K, *sk = b.S + (size_t)k * K;
  double be = 0.0, ga = 0.0;
  if (j < K) {
    for (int i = i0; i < i1; ++i) {
      double w = rh[i] - hk[i], v = rs[i] - sk[i];
      if (i == k) {
        w *= 0.5;
        v *= 0.5;
      }
      const double vij = b.V[(size_t)i * K + j];
      be = __builtin_fma(vij, w, be);
      ga = __builtin_fma(vij, v, ga);
    }
    double *o = bg + ((size_t)t * kProjSlices + blockIdx.z) * 2 * kmax;
    o[j] = be;
    o[kmax + j] = ga;
  }
}

// beta_j and gamma_j of trial t from the projection's partial rows.
__device__ __forceinline__ void proj_sum(const double *__restrict__ bg, int kmax, int j,
                                         double &be, double &ga) {
  be = 0.0;
  ga = 0.0;
#pragma unroll
  for (int z = 0; z < kProjSlices; ++z) {
    be += bg[(size_t)z * 2 * kmax + j];
    ga += bg[(size_t)z * 2 * kmax + kmax + j];
  }
}

// The count of eigenvalues of the trial pencil (H', S') below x.  In the
// resident factor basis, V^T (H' - x S') V = D + a b^T + b a^T with
// D = diag(lam - x), a = V^T e_k (row k of V), b = beta - x gamma.  By
// Sylvester's law (S' positive definite) the count is the negative inertia of
// that matrix, and by Haynsworth's inertia additivity on the bordered matrix
// [[D, U], [U^T, -C]] with U = [a b], C = [[0, 1], [1, 0]] (C^-1 = C):
//   nu(D + U C U^T) = nu(D) + nu(N) - nu(-C),  N = -C - U^T D^-1 U,  nu(-C) = 1.
// det(N) changes sign at every eigenvalue between two poles lam_j, which the
// root finder uses.
struct TrialCount {
  int count;    // eigenvalues of (H', S') below x
  int below;    // resident eigenvalues lam_j below x
  double detN;  // det N(x)
};

__device__ TrialCount trial_count(const double *__restrict__ lam, const double *__restrict__ a,
                                  const double *__restrict__ beta,
                                  const double *__restrict__ gam, int K, double x,
                                  double *red, int &buf) {
  double v[4] = {0.0, 0.0, 0.0, 0.0};  // below, aa, ab, bb
  for (int j = threadIdx.x; j < K; j += 256) {
    double d = lam[j] - x;
    v[0] += d < 0.0 ? 1.0 : 0.0;
    if (d == 0.0) d = 1e-300;
    const double r = 1.0 / d;
    const double bj = beta[j] - x * gam[j];
    v[1] = __builtin_fma(a[j] * a[j], r, v[1]);
    v[2] = __builtin_fma(a[j] * bj, r, v[2]);
    v[3] = __builtin_fma(bj * bj, r, v[3]);
  }
  block_sum4(v, red, buf);
  const double p = -v[1], q = -1.0 - v[2], rr = -v[3];
  TrialCount c;
  c.detN = p * rr - q * q;
  const int nuN = c.detN < 0.0 ? 1 : ((p + rr) < 0.0 ? 2 : 0);
  c.below = (int)v[0];
  c.count = c.below + nuN - 1;
  return c;
}

// One workgroup (256) per trial: the lowest eigenvalue mu of the trial pencil.
// Bracket: mu <= lam_2 (a rank-2 change moves the lowest eigenvalue at most
// past lam_2: nu(D) = 2 just above it) and mu <= E, the trial's Rayleigh
// quotient; below, step down until the count is 0.  While a pole lam_j lies
// inside the bracket, the count is taken just above (or below) the lowest
// such pole, which moves an end past it or onto it.  In a bracket free of
// poles, between lam_{m-1} and lam_m, det N has simple poles at both (its
// double-pole terms cancel), so Illinois regula falsi runs on the smooth
// g(x) = det N(x) (x - lam_{m-1}) (lam_m - x), which changes sign where the
// count does; a step that would land within an ulp of an end takes the end's
// neighbour instead.  Down to adjacent doubles: 9-14 counts per trial where
// bisection to the pole took 26-49.  NaN if the system has no resident
// factors or S' is not positive definite by a margin of 1e-10 (relative to S).
__global__ void __launch_bounds__(256) t_eig(const PoolSys *__restrict__ sys,
                                             const int32_t *__restrict__ tsys,
                                             const int32_t *__restrict__ tfun, int nsys,
                                             const double *__restrict__ bg, int kmax,
                                             const double *__restrict__ E,
                                             double *__restrict__ bgs,
                                             double *__restrict__ Eeig, int *__restrict__ flag) {
  __shared__ double red[32];
  int buf = 0;
  const int t = blockIdx.x;
  int k = 0;
  const PoolSys *bp = trial_sys(sys, nsys, tsys, tfun, t, k);
  if (!bp) {
    if (threadIdx.x == 0) Eeig[t] = __builtin_nan("");
    return;
  }
  const PoolSys &b = *bp;
  const int K = b.K;
  if (!b.V || !b.lam) {
    if (threadIdx.x == 0) {
      Eeig[t] = __builtin_nan("");
      atomicOr(flag, (int)F_BAD_INPUT);
    }
    return;
  }
  // beta, gamma from the partial rows, into this trial's summed rows
  double *beta = bgs + (size_t)t * 2 * kmax, *gam = beta + kmax;
  const double *part = bg + (size_t)t * kProjSlices * 2 * kmax;
  const double *a = b.V + (size_t)k * K;
  // V^T S' V = I + a g^T + g a^T has the smallest eigenvalue 1 + a.g - |a| |g|
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  for (int j = threadIdx.x; j < K; j += 256) {
    double be, ga;
    proj_sum(part, kmax, j, be, ga);
    beta[j] = be;
    gam[j] = ga;
    v[1] = __builtin_fma(a[j], a[j], v[1]);
    v[2] = __builtin_fma(a[j], ga, v[2]);
    v[3] = __builtin_fma(ga, ga, v[3]);
  }
  block_sum4(v, red, buf);  // also orders the beta/gamma stores before the reads below
  const double smin = 1.0 + v[2] - sqrt(v[1]) * sqrt(v[3]);
  double hi = E[t];
  if (!(smin > 1e-10) || !isfinite(hi)) {
    if (threadIdx.x == 0) Eeig[t] = __builtin_nan("");
    return;
  }
  const double scale = fmax(1.0, fabs(hi));
  if (K >= 2) hi = fmin(hi, b.lam[1]);
  double step = 1e-12 * scale;
  hi += step;
  TrialCount ch = trial_count(b.lam, a, beta, gam, K, hi, red, buf);
  for (int it = 0; ch.count < 1 && it < 64; ++it) {
    step *= 2.0;
    hi += step;
    ch = trial_count(b.lam, a, beta, gam, K, hi, red, buf);
  }
  step = 1e-3 * scale;
  double lo = fmin(hi, b.lam[0]) - step;
  TrialCount cl = trial_count(b.lam, a, beta, gam, K, lo, red, buf);
  for (int it = 0; cl.count >= 1 && it < 200; ++it) {
    step *= 2.0;
    lo -= step;
    cl = trial_count(b.lam, a, beta, gam, K, lo, red, buf);
  }
  // (x - lam_{m-1}) (lam_m - x): the factor that removes the poles at the
  // ends of a bracket with m resident eigenvalues below it
  auto gfac = [&](double x, int m) {
    double f = 1.0;
    if (m >= 1) f *= x - b.lam[m - 1];
    if (m < K) f *= b.lam[m] - x;
    return f;
  };
  const double eps = 1e-12 * scale;
  bool have_g = false;
  double glo = 0.0, ghi = 0.0;
  int side = 0;  // last end moved: -1 lo, +1 hi (Illinois)
  for (int it = 0; it < 200; ++it) {
    const double mid = 0.5 * (lo + hi);
    if (!(mid > lo && mid < hi)) break;
    double x = mid;
    if (cl.below != ch.below) {   // a pole inside: the count beside the lowest one
      const double pole = b.lam[cl.below];
      x = pole + eps;
      if (!(x > lo && x < hi)) {
        x = pole - eps;
        if (!(x > lo && x < hi)) x = mid;
      }
      have_g = false;
    } else {
      if (!have_g) {
        glo = cl.detN * gfac(lo, cl.below);
        ghi = ch.detN * gfac(hi, cl.below);
        side = 0;
        have_g = true;
      }
      if ((glo < 0.0) != (ghi < 0.0)) {
        const double xf = hi - ghi * (hi - lo) / (ghi - glo);
        if (xf > lo && xf < hi) x = xf;
        else if (xf <= lo) x = nextafter(lo, hi);
        else x = nextafter(hi, lo);
        if (!(x > lo && x < hi)) x = mid;
      }
    }
    const TrialCount cx = trial_count(b.lam, a, beta, gam, K, x, red, buf);
    if (cx.count >= 1) {
      hi = x;
      ch = cx;
      if (have_g) {
        ghi = cx.detN * gfac(x, cx.below);
        if (side == 1) glo *= 0.5;
        side = 1;
      }
    } else {
      lo = x;
      cl = cx;
      if (have_g) {
        glo = cx.detN * gfac(x, cx.below);
        if (side == -1) ghi *= 0.5;
        side = -1;
      }
    }
    if (cx.detN == 0.0) break;
  }
  if (threadIdx.x == 0) Eeig[t] = hi;
}

int pool_ws_stride(const std::vector<PoolSys> &v) {
  long long m = 0;
  for (const PoolSys &b : v) {
    long long w = 0;
    POOL_SWITCH_N(b.n, w = Dim<NN>::BS + (long long)b.nst * Dim<NN>::KS);
    m = std::max(m, w);
  }
  return (int)((m + 1) & ~1LL);
}

}  // namespace

struct necg_pool {
  int device = 0;
  std::vector<necg_ctx *> ctx;  // empty for a batch-backed pool
  std::vector<PoolSys> hsys;
  PoolSys *d_sys = nullptr;
  int *flag = nullptr;
  int kmax = 0, wstride = 0;
};

namespace {

// The resident state of every context, taken under the context's lock; false
// if a context holds no consistent state (no full evaluation yet, or a later
// call rebuilt its tables or matrices without forming the quotient again).
bool snapshot(necg_pool *P, std::vector<PoolSys> &out) {
  out.clear();
  for (necg_ctx *c : P->ctx) {
    std::lock_guard<std::mutex> lk(c->mu);
    if (c->last_K < 1 || !c->bra || !c->H || !c->HC || !c->scal) return false;
    if (c->full_gen != c->gen) return false;
    PoolSys b{};
    b.sc = c->sc;
    b.n = c->n;
    b.nst = c->nst;
    b.K = c->last_K;
    b.bra = c->bra;
    b.H = c->H;
    b.S = c->S;
    b.HC = c->HC;
    b.SC = c->SC;
    b.scal = c->scal;
    b.sml = c->d_sml;
    b.sms = c->d_sms;
    b.scvv = c->d_scvv;
    out.push_back(b);
  }
  return true;
}

// A context-backed pool: whether the contexts still hold the state the pool
// was made from (a batch-backed pool always does).
bool pool_current(necg_pool *P) {
  if (P->ctx.empty()) return true;
  std::vector<PoolSys> now;
  if (!snapshot(P, now)) return false;
  for (size_t i = 0; i < now.size(); ++i)
    if (now[i].bra != P->hsys[i].bra || now[i].H != P->hsys[i].H || now[i].K != P->hsys[i].K)
      return false;
  return true;
}

// The pool's device current for the scope of a call, the caller's restored.
struct PoolDevice {
  int prev = -1, dev;
  explicit PoolDevice(int d) : dev(d) {
    (void)hipGetDevice(&prev);
    if (prev != dev) (void)hipSetDevice(dev);
  }
  ~PoolDevice() {
    if (prev >= 0 && prev != dev) (void)hipSetDevice(prev);
  }
};

// The trial kernels: new records, rows of H' and S', the quotient.
void launch_trials(necg_pool *P, int ntr, const int32_t *tsys, const int32_t *tfun,
                   const double *newp, int pstride, double *ws, double *E, double *rowH,
                   double *rowS, int rstride, hipStream_t st) {
  const int nsys = (int)P->hsys.size();
  hipLaunchKernelGGL(t_prep, dim3(ntr), dim3(256), 0, st, P->d_sys, tsys, tfun, nsys, newp,
                     pstride, ws, (long long)P->wstride, P->flag);
  hipLaunchKernelGGL(t_rows, dim3((P->kmax + 63) / 64, ntr), dim3(64), 0, st, P->d_sys, tsys, tfun,
                     nsys, ws, (long long)P->wstride, rowH, rowS, rstride, P->flag);
  hipLaunchKernelGGL(t_energy, dim3(ntr), dim3(256), 0, st, P->d_sys, tsys, tfun, nsys, rowH,
                     rowS, rstride, E, P->flag);
}

// The eigenvalue kernels: projection on the resident factors, the root.
void launch_eig(necg_pool *P, int ntr, const int32_t *tsys, const int32_t *tfun,
                const double *rowH, const double *rowS, int rstride, const double *E,
                double *bg, double *Eeig, hipStream_t st) {
  const int nsys = (int)P->hsys.size();
  double *bgs = bg + (size_t)ntr * kProjSlices * 2 * P->kmax;
  hipLaunchKernelGGL(t_proj, dim3((P->kmax + 255) / 256, ntr, kProjSlices), dim3(256), 0, st,
                     P->d_sys, tsys, tfun, nsys, rowH, rowS, rstride, bg, P->kmax);
  hipLaunchKernelGGL(t_eig, dim3(ntr), dim3(256), 0, st, P->d_sys, tsys, tfun, nsys, bg, P->kmax,
                     E, bgs, Eeig, P->flag);
}

}  // namespace

int necg::pool_create_from(int device, const std::vector<PoolSys> &recs, necg_pool **out) {
  if (recs.empty() || !out) return NECG_ERR_ARG;
  necg_pool *P = new (std::nothrow) necg_pool();
  if (!P) return NECG_ERR_OOM;
  P->device = device;
  P->hsys = recs;
  for (const PoolSys &b : P->hsys) P->kmax = std::max(P->kmax, b.K);
  P->wstride = pool_ws_stride(P->hsys);
  int prev = -1;
  (void)hipGetDevice(&prev);
  (void)hipSetDevice(device);
  int rc = NECG_OK;
  if (hipMalloc((void **)&P->d_sys, sizeof(PoolSys) * P->hsys.size()) != hipSuccess ||
      hipMalloc((void **)&P->flag, sizeof(int)) != hipSuccess ||
      hipMemcpy(P->d_sys, P->hsys.data(), sizeof(PoolSys) * P->hsys.size(),
                hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(P->flag, 0, sizeof(int)) != hipSuccess)
    rc = NECG_ERR_DEVICE;
  if (prev >= 0) (void)hipSetDevice(prev);
  if (rc) {
    if (P->d_sys) (void)hipFree(P->d_sys);
    if (P->flag) (void)hipFree(P->flag);
    delete P;
    return rc;
  }
  *out = P;
  return NECG_OK;
}

extern "C" {

int necg_pool_destroy(necg_pool *P) {
  if (!P) return NECG_ERR_ARG;
  int prev = -1;
  (void)hipGetDevice(&prev);
  (void)hipSetDevice(P->device);
  if (P->d_sys) (void)hipFree(P->d_sys);
  if (P->flag) (void)hipFree(P->flag);
  if (prev >= 0) (void)hipSetDevice(prev);
  delete P;
  return NECG_OK;
}

int necg_pool_create(necg_ctx *const *ctxs, int32_t nctx, necg_pool **out) {
  if (!ctxs || nctx < 1 || !out) return NECG_ERR_ARG;
  *out = nullptr;
  necg_pool probe;
  for (int i = 0; i < nctx; ++i) {
    if (!ctxs[i] || ctxs[i]->device != ctxs[0]->device) return NECG_ERR_ARG;
    probe.ctx.push_back(ctxs[i]);
  }
  std::vector<PoolSys> recs;
  if (!snapshot(&probe, recs)) return NECG_ERR_ARG;  // every context needs one evaluation first
  int rc = pool_create_from(ctxs[0]->device, recs, out);
  if (rc == NECG_OK) (*out)->ctx = probe.ctx;
  return rc;
}

int necg_pool_sizes(necg_pool *P, int32_t *kmax, int32_t *ws_stride) {
  if (!P || !kmax || !ws_stride) return NECG_ERR_ARG;
  *kmax = P->kmax;
  *ws_stride = P->wstride;
  return NECG_OK;
}

int necg_pool_trials(necg_pool *P, int32_t ntr, const int32_t *tsys, const int32_t *tfun,
                     const double *newp, int32_t pstride, double *ws, double *E, double *rowH,
                     double *rowS, int32_t rstride, void *stream) {
  if (!P || ntr < 0 || (ntr > 0 && (!tsys || !tfun || !newp || !ws || !E || !rowH || !rowS)) ||
      rstride < P->kmax || pstride < 1)
    return NECG_ERR_ARG;
  if (ntr == 0) return NECG_OK;
  NECG_REQUIRE_DEVICE((const void *)tsys, (const void *)tfun, (const void *)newp,
                      (const void *)ws, (const void *)E, (const void *)rowH, (const void *)rowS);
  if (!pool_current(P)) return NECG_ERR_ARG;
  PoolDevice pd(P->device);
  launch_trials(P, ntr, tsys, tfun, newp, pstride, ws, E, rowH, rowS, rstride,
                (hipStream_t)stream);
  return hipGetLastError() == hipSuccess ? NECG_OK : NECG_ERR_DEVICE;
}

int necg_pool_trials_io(necg_pool *P, int32_t ntr, const void *in_h, void *in_d,
                        int32_t pstride, double *ws, double *rowH, double *rowS, int32_t rstride,
                        double *bg, double *out_d, double *out_h, void *stream) {
  if (!P || ntr < 0 ||
      (ntr > 0 && (!in_h || !in_d || !ws || !rowH || !rowS || !out_d || !out_h)) ||
      rstride < P->kmax || pstride < 1)
    return NECG_ERR_ARG;
  if (ntr == 0) return NECG_OK;
  NECG_REQUIRE_DEVICE((const void *)in_d, (const void *)ws, (const void *)rowH,
                      (const void *)rowS, (const void *)bg, (const void *)out_d);
  if (((uintptr_t)in_d & 7u) != 0) return NECG_ERR_ARG;
  if (!pool_current(P)) return NECG_ERR_ARG;
  PoolDevice pd(P->device);
  hipStream_t st = (hipStream_t)stream;
  const size_t off = ((size_t)8 * ntr + 7) & ~(size_t)7;
  const size_t bytes = off + (size_t)8 * ntr * pstride;
  if (hipMemcpyAsync(in_d, in_h, bytes, hipMemcpyHostToDevice, st) != hipSuccess)
    return NECG_ERR_DEVICE;
  const int32_t *tsys = (const int32_t *)in_d;
  const int32_t *tfun = tsys + ntr;
  const double *newp = (const double *)((const char *)in_d + off);
  launch_trials(P, ntr, tsys, tfun, newp, pstride, ws, out_d, rowH, rowS, rstride, st);
  if (bg) launch_eig(P, ntr, tsys, tfun, rowH, rowS, rstride, out_d, bg, out_d + ntr, st);
  if (hipGetLastError() != hipSuccess) return NECG_ERR_DEVICE;
  const size_t nout = (size_t)(bg ? 2 : 1) * ntr;
  if (hipMemcpyAsync(out_h, out_d, nout * sizeof(double), hipMemcpyDeviceToHost, st) !=
      hipSuccess)
    return NECG_ERR_DEVICE;
  return NECG_OK;
}

int necg_pool_set_factors(necg_pool *P, int32_t s, const double *V, const double *lam) {
  if (!P || s < 0 || s >= (int)P->hsys.size() || (!V) != (!lam)) return NECG_ERR_ARG;
  NECG_REQUIRE_DEVICE((const void *)V, (const void *)lam);
  P->hsys[s].V = V;
  P->hsys[s].lam = lam;
  int prev = -1;
  (void)hipGetDevice(&prev);
  if (prev != P->device) (void)hipSetDevice(P->device);
  const hipError_t e =
      hipMemcpy(P->d_sys + s, &P->hsys[s], sizeof(PoolSys), hipMemcpyHostToDevice);
  if (prev >= 0 && prev != P->device) (void)hipSetDevice(prev);
  return e == hipSuccess ? NECG_OK : NECG_ERR_DEVICE;
}

int necg_pool_eig(necg_pool *P, int32_t ntr, const int32_t *tsys, const int32_t *tfun,
                  const double *rowH, const double *rowS, int32_t rstride, const double *E,
                  double *bg, double *Eeig, void *stream) {
  if (!P || ntr < 0 || (ntr > 0 && (!tsys || !tfun || !rowH || !rowS || !E || !bg || !Eeig)) ||
      rstride < P->kmax)
    return NECG_ERR_ARG;
  if (ntr == 0) return NECG_OK;
  NECG_REQUIRE_DEVICE((const void *)tsys, (const void *)tfun, (const void *)rowH,
                      (const void *)rowS, (const void *)E, (const void *)bg, (const void *)Eeig);
  PoolDevice pd(P->device);
  launch_eig(P, ntr, tsys, tfun, rowH, rowS, rstride, E, bg, Eeig, (hipStream_t)stream);
  return hipGetLastError() == hipSuccess ? NECG_OK : NECG_ERR_DEVICE;
}

int necg_pool_status(necg_pool *P, void *stream) {
  if (!P) return NECG_ERR_ARG;
  if (hipStreamSynchronize((hipStream_t)stream) != hipSuccess) return NECG_ERR_DEVICE;
  int flags = 0;
  if (hipMemcpy(&flags, P->flag, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
    return NECG_ERR_DEVICE;
  if (flags && hipMemset(P->flag, 0, sizeof(int)) != hipSuccess) return NECG_ERR_DEVICE;
  if (flags & F_BAD_INPUT) return NECG_ERR_NONFINITE;
  if (flags & F_NOT_PD) return NECG_ERR_NOT_PD;
  if (flags & F_CSC) return NECG_ERR_CSC;
  if (flags & F_NONFINITE) return NECG_ERR_NONFINITE;
  return NECG_OK;
}

}  // extern "C"
