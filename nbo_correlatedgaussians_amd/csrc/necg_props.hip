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
  o[R::F] = exp2(0.75 * N) * dk * sqrt(dk) * Ck * scvv[p];
  if (!isfinite(i1) || !isfinite(dk) || !isfinite(Ck)) atomicOr(flag, (int)F_BAD_INPUT);
  else if (!(dk > 0.0)) atomicOr(flag, (int)F_NOT_PD);  // A_k singular (a zero L diagonal)
}

// One term pair: bra record b (registers), ket record r (wave-uniform).
template <int N>
__device__ __forceinline__ void prop_term(const double *b, const double *__restrict__ r,
                                          double mult, double *acc, bool &pd) {
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
  double st[N][3];
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
  const double wov =
      mult * b[R::F] * r[R::F] * (invdet * sqrt(invdet)) * exp(i3 - b[R::I1] - r[R::I1]);
  acc[0] += wov;
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int c = 0; c < 3; ++c) acc[1 + i * 3 + c] += wov * st[i][c];
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

// Ket tile (functions) so that about 8192 (bra block, ket tile) waves carry work.
int prop_ket_tile(int K, int nblocks) {
  const long long t = ((long long)nblocks * K) / (2LL * 8192);
  return (int)std::max(1LL, std::min<long long>(t, K));
}

template <int N>
int do_dipole(necg_ctx *c, int K, const double *x, double *out, hipStream_t st) {
  using R = PR<N>;
  const long long G = (long long)K * c->nst;
  const int nb = (int)((G + 63) / 64);
  const int tile = prop_ket_tile(K, nb);
  const int nt = (K + tile - 1) / tile;
  int rc;
  if ((rc = grow(&c->prop_rec, c->prop_rec_cap, (size_t)G * R::RS))) return rc;
  if ((rc = grow(&c->prop_part, c->prop_part_cap, (size_t)nb * nt * R::NV))) return rc;
  hipLaunchKernelGGL((prop_prep<N>), dim3((unsigned)((G + 255) / 256)), dim3(256), 0, st, x, K,
                     c->nst, c->d_scvv, c->d_sml, c->d_sms, c->prop_rec, c->d_flag);
  hipLaunchKernelGGL((prop_pair<N>), dim3(nb, nt), dim3(64), 0, st, c->prop_rec, K, c->nst, tile,
                     c->prop_part, c->d_flag);
  hipLaunchKernelGGL((prop_reduce<N>), dim3(1), dim3(256), 0, st, c->prop_part, nb * nt, c->sc,
                     out, c->d_flag);
  return hip_rc(hipGetLastError());
}

}  // namespace

extern "C" int necg_dipole(necg_ctx *c, int32_t K, const double *x, double *out, void *stream) {
  if (!c || K < 1 || !x || !out) return NECG_ERR_ARG;
  if ((long long)K * c->nst > (1LL << 31) - 64) return NECG_ERR_ARG;
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
