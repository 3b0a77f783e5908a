// necg_props.hip -- the properties pass of the driver: the doubly symmetrised
// overlap s and position matrices r_c of every pseudo-particle component
// (python/rrms.py:21-120, source/rrms.f90), contracted with C into the centres
// <r_c> = C^T r_c C / C^T s C and the dipole moment sum_i ncv[i+1] <r_i>
// (python/necg.py:200-232, source/necg.f90:387-446).
//
// n-space form.  Record of function k under term p (bra and ket alike; the
// reference transforms both sides): A' = P^T A_k P, s' = T s_k (T = sms[p],
// an n x 3 matrix after the transform), u = A' s', i1 = s'.u and
// f = 8^(n/4) |L_k|^(3/2) C_k scvv_p.  A term pair (k p, l q) then needs
//   M = A'_kp + A'_lq,  e = u_kp + u_lq,  st = M^-1 e,
//   w ov = f_kp f_lq det(M)^(-3/2) exp(e.st - i1_kp - i1_lq),
// which is scvv_p scvv_q C_k C_l times the reference's ov
// (exp(-(i1+i2)+i3) sqrt(8^n) sqrt((detak detal / dett)^3)), and adds w ov to s
// and w ov st to r.  Pairs k > l count twice (the reference mirrors its lower
// triangle before contracting with C).
//
// Symmetry reduction.  When every term is a permutation (sml[p] = P, sms[p] =
// kron(P^T, I3)) and every product Q P^T is again a term, the transformed
// pair (P k, Q l) is P applied to the pair (k, R l) with R = Q P^T: the same
// ov and det, and st(P k, Q l) = P^T st(k, R l).  The nst^2 term pairs then
// collapse to nst: for each term r, w_r = sum of scvv_p scvv_q over the (p, q)
// with Q P^T = R_r, and W_r = the same sum of c_p c_q P^T, so that the pair
// (k, l) adds w_r ov and ov W_r st(k, R_r l) (bra k untransformed).  Terms
// that do not close keep the full double sum.
//
// Launch: one wave per (64 consecutive (k, p) records, ket tile of functions);
// the bra record sits in registers, the ket records (l, q) are wave-uniform
// loads; each wave folds its lanes into one partial row (fixed butterfly), and
// one workgroup sums the partial rows in a fixed order.  Deterministic.
#include <algorithm>

#include "necg_ctx.hpp"

namespace {

using namespace necg;

template <int N>
struct PR {
  static constexpr int NS = N * (N + 1) / 2;
  static constexpr int A = 0;            // A' packed, sx(i, j)
  static constexpr int U = NS;           // u, i * 3 + xyz
  static constexpr int I1 = NS + 3 * N;  // s'.u
  static constexpr int F = I1 + 1;       // 8^(n/4) |L|^(3/2) C scvv
  static constexpr int RS = F + 1;
  static constexpr int NV = 1 + 3 * N;   // s, r_c
  static constexpr int BLK = NS + 3 * N;
};

template <int N>
__global__ void __launch_bounds__(256)
    prop_prep(const double *__restrict__ x, int K, int nst, const double *__restrict__ scvv,
              const double *__restrict__ sml, const double *__restrict__ sms,
              double *__restrict__ rec, int *__restrict__ flag) {
  using R = PR<N>;
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (long long)K * nst) return;
  const int k = (int)(g / nst), p = (int)(g - (long long)k * nst);
  const double *iv = x + (size_t)k * R::BLK;
  double L[N][N];
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) L[i][j] = 0.0;
  {
    int t = 0;
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = i; j < N; ++j) L[j][i] = iv[t++];
  }
  double A[N][N];
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      double a = 0.0;
#pragma unroll
      for (int m = 0; m < N; ++m) a += L[i][m] * L[j][m];
      A[i][j] = a;
    }
  const double *P = sml + (size_t)p * N * N;  // P[a N + b] = reference sml(a, b, p)
  double AP[N][N];
#pragma unroll
  for (int a = 0; a < N; ++a)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      double v = 0.0;
#pragma unroll
      for (int b = 0; b < N; ++b) v += A[a][b] * P[b * N + j];
      AP[a][j] = v;
    }
  double Ap[N][N];
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      double v = 0.0;
#pragma unroll
      for (int a = 0; a < N; ++a) v += P[a * N + i] * AP[a][j];
      Ap[i][j] = v;
    }
  const double *T = sms + (size_t)p * 9 * N * N;
  double sp[3 * N];
#pragma unroll
  for (int r = 0; r < 3 * N; ++r) {
    double v = 0.0;
    for (int c = 0; c < 3 * N; ++c) v += T[r * 3 * N + c] * iv[R::NS + c];
    sp[r] = v;
  }
  double *o = rec + (size_t)g * R::RS;
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j <= i; ++j) o[R::A + sx(i, j)] = Ap[i][j];
  double i1 = 0.0;
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      double u = 0.0;
#pragma unroll
      for (int j = 0; j < N; ++j) u += Ap[i][j] * sp[j * 3 + c];
      o[R::U + i * 3 + c] = u;
      i1 += sp[i * 3 + c] * u;
    }
  o[R::I1] = i1;
  double det = 1.0;
#pragma unroll
  for (int i = 0; i < N; ++i) det *= L[i][i];
  const double dk = fabs(det);
  const double Ck = x[(size_t)K * R::BLK + k];
  o[R::F] = exp2(0.75 * N) * dk * sqrt(dk) * Ck * (scvv ? scvv[p] : 1.0);
  if (!isfinite(i1) || !isfinite(dk) || !isfinite(Ck)) atomicOr(flag, (int)F_BAD_INPUT);
  else if (!(dk > 0.0)) atomicOr(flag, (int)F_NOT_PD);  // A_k singular (a zero L diagonal)
}

// One term pair: bra record b (registers), ket record r (wave-uniform);
// returns w ov and st = M^-1 e.
template <int N>
__device__ __forceinline__ double prop_core(const double *b, const double *__restrict__ r,
                                            double mult, double (&st)[N][3], bool &pd) {
  using R = PR<N>;
  double M[R::NS];
#pragma unroll
  for (int s = 0; s < R::NS; ++s) M[s] = b[R::A + s] + r[R::A + s];
  double e[N][3];
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int c = 0; c < 3; ++c) e[i][c] = b[R::U + i * 3 + c] + r[R::U + i * 3 + c];
  // M = L D L^T, L unit lower triangular; W = L D
  double Lm[N][N], W[N][N], invD[N];
  double invdet = 1.0;
#pragma unroll
  for (int j = 0; j < N; ++j) {
    double d = M[sx(j, j)];
#pragma unroll
    for (int m = 0; m < j; ++m) d -= Lm[j][m] * W[j][m];
    pd = pd && (d > 0.0);
    invD[j] = rcp(d);
    invdet *= invD[j];
#pragma unroll
    for (int i = j + 1; i < N; ++i) {
      double w = M[sx(i, j)];
#pragma unroll
      for (int m = 0; m < j; ++m) w -= W[i][m] * Lm[j][m];
      W[i][j] = w;
      Lm[i][j] = w * invD[j];
    }
  }
  // st = M^-1 e (three right-hand sides), i3 = e.st
  double y[N][3];
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      double v = e[i][c];
#pragma unroll
      for (int m = 0; m < i; ++m) v -= Lm[i][m] * y[m][c];
      y[i][c] = v;
    }
#pragma unroll
  for (int i = N - 1; i >= 0; --i)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      double v = y[i][c] * invD[i];
#pragma unroll
      for (int m = i + 1; m < N; ++m) v -= Lm[m][i] * st[m][c];
      st[i][c] = v;
    }
  double i3 = 0.0;
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int c = 0; c < 3; ++c) i3 += e[i][c] * st[i][c];
  return mult * b[R::F] * r[R::F] * (invdet * sqrt(invdet)) * exp(i3 - b[R::I1] - r[R::I1]);
}

template <int N>
__device__ __forceinline__ void prop_term(const double *b, const double *__restrict__ r,
                                          double mult, double *acc, bool &pd) {
  double st[N][3];
  const double wov = prop_core<N>(b, r, mult, st, pd);
  acc[0] += wov;
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int c = 0; c < 3; ++c) acc[1 + i * 3 + c] += wov * st[i][c];
}

// The reduced form: tb = [w_r, W_r (n x n row-major)] of the ket's term.
template <int N>
__device__ __forceinline__ void prop_term_grp(const double *b, const double *__restrict__ r,
                                              const double *__restrict__ tb, double mult,
                                              double *acc, bool &pd) {
  double st[N][3];
  const double ov = prop_core<N>(b, r, mult, st, pd);
  acc[0] = __builtin_fma(ov, tb[0], acc[0]);
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      double v = 0.0;
#pragma unroll
      for (int j = 0; j < N; ++j) v = __builtin_fma(tb[1 + i * N + j], st[j][c], v);
      acc[1 + i * 3 + c] = __builtin_fma(ov, v, acc[1 + i * 3 + c]);
    }
}

template <int N>
__global__ void __launch_bounds__(64)
    prop_pair(const double *__restrict__ rec, int K, int nst, int ket_tile,
              double *__restrict__ part, int *__restrict__ flag) {
  using R = PR<N>;
  const int lane = threadIdx.x;
  const int nb = blockIdx.x, t = blockIdx.y;
  const long long G = (long long)K * nst;
  const long long g = (long long)nb * 64 + lane;
  const bool valid = g < G;
  const int k = valid ? (int)(g / nst) : -1;
  const long long glast = min((long long)nb * 64 + 63, G - 1);
  const int kmax = (int)(glast / nst);  // largest bra function of the wave
  const int l0 = t * ket_tile;
  const int l1 = min(min(K, l0 + ket_tile), kmax + 1);
  double acc[R::NV];
#pragma unroll
  for (int v = 0; v < R::NV; ++v) acc[v] = 0.0;
  bool pd = true;
  if (l0 < l1) {
    double b[R::RS];
    const double *br = rec + (size_t)(valid ? g : 0) * R::RS;
#pragma unroll
    for (int v = 0; v < R::RS; ++v) b[v] = br[v];
    for (int l = l0; l < l1; ++l) {
      const double mult = l < k ? 2.0 : 1.0;
      const bool on = valid && l <= k;
      const double *kr = rec + (size_t)l * nst * R::RS;
      for (int q = 0; q < nst; ++q, kr += R::RS) {
        if (on) prop_term<N>(b, kr, mult, acc, pd);
      }
    }
  }
#pragma unroll
  for (int v = 0; v < R::NV; ++v) {
    double a = acc[v];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o, 64);
    acc[v] = a;
  }
  if (lane == 0) {
    double *out = part + ((size_t)t * gridDim.x + nb) * R::NV;
#pragma unroll
    for (int v = 0; v < R::NV; ++v) out[v] = acc[v];
  }
  if (!pd) atomicOr(flag, (int)F_NOT_PD);
}

// Reduced form: lanes are bra functions k (record of the identity term), the
// ket records (l, r) and the table rows of r are wave-uniform.
template <int N>
__global__ void __launch_bounds__(64)
    prop_pair_grp(const double *__restrict__ rec, const double *__restrict__ tab, int K, int nst,
                  int ident, int ket_tile, double *__restrict__ part, int *__restrict__ flag) {
  using R = PR<N>;
  constexpr int TS = 1 + N * N;
  const int lane = threadIdx.x;
  const int nb = blockIdx.x, t = blockIdx.y;
  const int k0 = nb * 64;
  const int k = k0 + lane;
  const bool valid = k < K;
  const int kmax = min(k0 + 63, K - 1);
  const int l0 = t * ket_tile;
  const int l1 = min(min(K, l0 + ket_tile), kmax + 1);
  double acc[R::NV];
#pragma unroll
  for (int v = 0; v < R::NV; ++v) acc[v] = 0.0;
  bool pd = true;
  if (l0 < l1) {
    double b[R::RS];
    const double *br = rec + ((size_t)(valid ? k : 0) * nst + ident) * R::RS;
#pragma unroll
    for (int v = 0; v < R::RS; ++v) b[v] = br[v];
    for (int l = l0; l < l1; ++l) {
      const double mult = l < k ? 2.0 : 1.0;
      const bool on = valid && l <= k;
      const double *kr = rec + (size_t)l * nst * R::RS;
      const double *tb = tab;
      for (int r = 0; r < nst; ++r, kr += R::RS, tb += TS) {
        if (on) prop_term_grp<N>(b, kr, tb, mult, acc, pd);
      }
    }
  }
#pragma unroll
  for (int v = 0; v < R::NV; ++v) {
    double a = acc[v];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o, 64);
    acc[v] = a;
  }
  if (lane == 0) {
    double *out = part + ((size_t)t * gridDim.x + nb) * R::NV;
#pragma unroll
    for (int v = 0; v < R::NV; ++v) out[v] = acc[v];
  }
  if (!pd) atomicOr(flag, (int)F_NOT_PD);
}

// out = [C^T s C, <r_c> (3n), dipole (3)]
template <int N>
__global__ void __launch_bounds__(256)
    prop_reduce(const double *__restrict__ part, int nitems, SysConst sc,
                double *__restrict__ out, int *__restrict__ flag) {
  using R = PR<N>;
  __shared__ double red4[4];
  __shared__ double tot[R::NV];
  for (int v = 0; v < R::NV; ++v) {
    double a = 0.0;
    for (int i = threadIdx.x; i < nitems; i += 256) a += part[(size_t)i * R::NV + v];
    a = block_sum256(a, red4);
    if (threadIdx.x == 0) tot[v] = a;
  }
  if (threadIdx.x == 0) {
    const double s = tot[0];
    out[0] = s;
    double dip[3] = {0.0, 0.0, 0.0};
    bool fin = isfinite(s);
    for (int c = 0; c < 3 * N; ++c) {
      const double r = tot[1 + c] / s;
      out[1 + c] = r;
      dip[c % 3] += sc.q[c / 3 + 1] * r;
      fin = fin && isfinite(r);
    }
    for (int a = 0; a < 3; ++a) out[1 + 3 * N + a] = dip[a];
    if (!(s > 0.0)) atomicOr(flag, (int)F_CSC);
    if (!fin) atomicOr(flag, (int)F_NONFINITE);
  }
}

int hip_rc(hipError_t e) {
  if (e == hipSuccess) return NECG_OK;
  return e == hipErrorOutOfMemory ? NECG_ERR_OOM : NECG_ERR_DEVICE;
}

int grow(double **p, size_t &cap, size_t count) {
  if (count <= cap) return NECG_OK;
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  cap = 0;
  int rc = hip_rc(hipMalloc((void **)p, count * sizeof(double)));
  if (rc == NECG_OK) cap = count;
  return rc;
}

struct DevSwitch {
  int prev = -1;
  explicit DevSwitch(int dev) {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    if (prev != dev) (void)hipSetDevice(dev);
  }
  ~DevSwitch() {
    if (prev >= 0) (void)hipSetDevice(prev);
  }
};

// Ket tile (functions) so that about 8192 (bra block, ket tile) waves carry
// work (about half of the grid: only ket tiles at or below the block's bras).
int prop_ket_tile(int K, int nblocks) {
  const long long t = ((long long)nblocks * K) / (2LL * 8192);
  return (int)std::max(1LL, std::min<long long>(t, K));
}

// Analyse the terms once per context: permutations with sms = kron(P^T, I3)
// whose products Q P^T are all terms give the reduced form and its table
// [w_r, W_r] per term r.
int prop_analyse(necg_ctx *c) {
  if (c->prop_analysed) return NECG_OK;
  const int n = c->n, nst = c->nst, n3 = 3 * n;
  std::vector<double> sml((size_t)nst * n * n), sms((size_t)nst * n3 * n3), w(nst);
  int rc;
  if ((rc = hip_rc(hipMemcpy(sml.data(), c->d_sml, sizeof(double) * sml.size(),
                             hipMemcpyDeviceToHost))) ||
      (rc = hip_rc(hipMemcpy(sms.data(), c->d_sms, sizeof(double) * sms.size(),
                             hipMemcpyDeviceToHost))) ||
      (rc = hip_rc(hipMemcpy(w.data(), c->d_scvv, sizeof(double) * nst, hipMemcpyDeviceToHost))))
    return rc;
  auto P = [&](int p, int i, int j) { return sml[((size_t)p * n + i) * n + j]; };
  bool ok = true;
  for (int p = 0; p < nst && ok; ++p) {
    for (int i = 0; i < n && ok; ++i) {
      int rows = 0, cols = 0;
      for (int j = 0; j < n; ++j) {
        const double a = P(p, i, j), b = P(p, j, i);
        if ((a != 0.0 && a != 1.0) || (b != 0.0 && b != 1.0)) ok = false;
        rows += a == 1.0;
        cols += b == 1.0;
      }
      ok = ok && rows == 1 && cols == 1;
    }
    for (int r = 0; r < n3 && ok; ++r)
      for (int q = 0; q < n3 && ok; ++q) {
        const double want = (r % 3 == q % 3) ? P(p, q / 3, r / 3) : 0.0;
        ok = sms[((size_t)p * n3 + r) * n3 + q] == want;
      }
  }
  int ident = -1;
  std::vector<double> tab;
  if (ok) {
    auto find = [&](const std::vector<double> &m) {
      for (int r = 0; r < nst; ++r)
        if (std::equal(m.begin(), m.end(), sml.begin() + (size_t)r * n * n)) return r;
      return -1;
    };
    std::vector<double> eye((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) eye[(size_t)i * n + i] = 1.0;
    ident = find(eye);
    const int ts = 1 + n * n;
    tab.assign((size_t)nst * ts, 0.0);
    std::vector<double> m((size_t)n * n);
    for (int p = 0; p < nst && ok; ++p)
      for (int q = 0; q < nst && ok; ++q) {
        for (int i = 0; i < n; ++i)  // Q P^T
          for (int j = 0; j < n; ++j) {
            double v = 0.0;
            for (int t = 0; t < n; ++t) v += P(q, i, t) * P(p, j, t);
            m[(size_t)i * n + j] = v;
          }
        const int r = find(m);
        if (r < 0) {
          ok = false;
          break;
        }
        const double cw = w[p] * w[q];
        tab[(size_t)r * ts] += cw;
        for (int i = 0; i < n; ++i)
          for (int j = 0; j < n; ++j) tab[(size_t)r * ts + 1 + i * n + j] += cw * P(p, j, i);
      }
    ok = ok && ident >= 0;
  }
  if (ok) {
    size_t cap = 0;  // the table is made once per context
    if ((rc = grow(&c->prop_tab, cap, tab.size()))) return rc;
    if ((rc = hip_rc(hipMemcpy(c->prop_tab, tab.data(), sizeof(double) * tab.size(),
                               hipMemcpyHostToDevice))))
      return rc;
  }
  c->prop_grp_ok = ok ? 1 : 0;
  c->prop_ident = ident;
  c->prop_analysed = 1;
  return NECG_OK;
}

template <int N>
int do_dipole(necg_ctx *c, int K, const double *x, double *out, hipStream_t st) {
  using R = PR<N>;
  int rc;
  if ((rc = prop_analyse(c))) return rc;
  const bool grp = c->prop_mode == 1 || (c->prop_mode == 0 && c->prop_grp_ok);
  if (grp && !c->prop_grp_ok) return NECG_ERR_UNSUPPORTED;
  const long long G = (long long)K * c->nst;
  const int nb = grp ? (K + 63) / 64 : (int)((G + 63) / 64);
  const int tile = prop_ket_tile(K, nb);
  const int nt = (K + tile - 1) / tile;
  if ((rc = grow(&c->prop_rec, c->prop_rec_cap, (size_t)G * R::RS))) return rc;
  if ((rc = grow(&c->prop_part, c->prop_part_cap, (size_t)nb * nt * R::NV))) return rc;
  hipLaunchKernelGGL((prop_prep<N>), dim3((unsigned)((G + 255) / 256)), dim3(256), 0, st, x, K,
                     c->nst, grp ? nullptr : c->d_scvv, c->d_sml, c->d_sms, c->prop_rec,
                     c->d_flag);
  if (grp)
    hipLaunchKernelGGL((prop_pair_grp<N>), dim3(nb, nt), dim3(64), 0, st, c->prop_rec,
                       c->prop_tab, K, c->nst, c->prop_ident, tile, c->prop_part, c->d_flag);
  else
    hipLaunchKernelGGL((prop_pair<N>), dim3(nb, nt), dim3(64), 0, st, c->prop_rec, K, c->nst,
                       tile, c->prop_part, c->d_flag);
  hipLaunchKernelGGL((prop_reduce<N>), dim3(1), dim3(256), 0, st, c->prop_part, nb * nt, c->sc,
                     out, c->d_flag);
  return hip_rc(hipGetLastError());
}

}  // namespace

extern "C" int necg_set_dipole_mode(necg_ctx *c, int32_t mode) {
  if (!c || mode < 0 || mode > 2) return NECG_ERR_ARG;
  std::lock_guard<std::mutex> lk(c->mu);
  c->prop_mode = mode;
  return NECG_OK;
}

extern "C" int necg_dipole(necg_ctx *c, int32_t K, const double *x, double *out, void *stream) {
  if (!c || K < 1 || !x || !out) return NECG_ERR_ARG;
  if ((long long)K * c->nst > (1LL << 31) - 64) return NECG_ERR_ARG;
  NECG_REQUIRE_DEVICE((const void *)x, (const void *)out);
  std::lock_guard<std::mutex> lk(c->mu);
  DevSwitch dg(c->device);
  const hipStream_t st = (hipStream_t)stream;
  switch (c->n) {
    case 1: return do_dipole<1>(c, K, x, out, st);
    case 2: return do_dipole<2>(c, K, x, out, st);
    case 3: return do_dipole<3>(c, K, x, out, st);
    case 4: return do_dipole<4>(c, K, x, out, st);
    case 5: return do_dipole<5>(c, K, x, out, st);
    case 6: return do_dipole<6>(c, K, x, out, st);
    default: return NECG_ERR_UNSUPPORTED;
  }
}
