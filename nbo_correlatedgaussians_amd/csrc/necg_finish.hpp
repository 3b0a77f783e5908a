// necg_finish.hpp -- per-function preparation and per-system finishing steps
// of one energy+gradient evaluation, as device functions shared by the
// single-system kernels (necg.hip) and the segmented batch kernels
// (necg_batch.hip).
#pragma once

#include "necg_pair.hpp"

namespace necg {

// Bra record of one basis function from its parameter block p (vech L | s)
// and coefficient c (python/matel.py:56-70, :125-136), with A M A and
// tr(M A) for the two-sided pair term.
template <int N>
__device__ __forceinline__ void make_bra(const double *__restrict__ p, double c, bool with_c,
                                         const SysConst &sc, double *__restrict__ r,
                                         int *__restrict__ flag) {
  using D = Dim<N>;
  double L[N][N];
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) L[i][j] = 0.0;
  int t = 0;
  bool bad = false, pd = true;
#pragma unroll
  for (int j = 0; j < N; ++j)
#pragma unroll
    for (int i = j; i < N; ++i, ++t) {
      L[i][j] = p[t];
      bad = bad || !isfinite(p[t]);
    }
  double detL = 1.0;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    detL *= L[i][i];
    pd = pd && (L[i][i] != 0.0);
    r[D::B_LD + i] = 1.0 / L[i][i];
  }
#pragma unroll
  for (int v = 0; v < D::NS; ++v) r[D::B_L + v] = p[v];
  double A[D::NS];
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j <= i; ++j) {
      double s = 0.0;
#pragma unroll
      for (int m = 0; m <= j; ++m) s = __builtin_fma(L[i][m], L[j][m], s);
      A[sx(i, j)] = s;
      r[D::B_A + sx(i, j)] = s;
    }
  double S[D::N3];
#pragma unroll
  for (int v = 0; v < D::N3; ++v) {
    S[v] = p[D::NS + v];
    bad = bad || !isfinite(S[v]);
    r[D::B_S + v] = S[v];
  }
  double i1 = 0.0;
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int x = 0; x < 3; ++x) {
      double s = 0.0;
#pragma unroll
      for (int j = 0; j < N; ++j) s = __builtin_fma(A[sx(i, j)], S[3 * j + x], s);
      r[D::B_AS + 3 * i + x] = s;
      i1 = __builtin_fma(S[3 * i + x], s, i1);
    }
  r[D::B_I1] = i1;
  // A M A (symmetric, packed) and tr(M A)
  double AM[N][N];
  double trma = 0.0;
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      double s = 0.0;
#pragma unroll
      for (int m = 0; m < N; ++m) s = __builtin_fma(A[sx(i, m)], sc.mass[m * N + j], s);
      AM[i][j] = s;
      if (i == j) trma += s;
    }
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j <= i; ++j) {
      double s = 0.0;
#pragma unroll
      for (int m = 0; m < N; ++m) s = __builtin_fma(AM[i][m], A[sx(m, j)], s);
      r[D::B_AMA + sx(i, j)] = s;
    }
  r[D::B_TRMA] = trma;
  // pk = 8^(n/4) |det L|^(3/2): the bra half of
  // sqrt(8^n) sqrt((|det L_k| |det L_l| / det B)^3) (python/matel.py:135)
  const double ad = fabs(detL);
  r[D::B_PK] = pow(8.0, 0.25 * N) * ad * sqrt(ad);
  if (with_c) bad = bad || !isfinite(c);
  r[D::B_C] = with_c ? c : 0.0;
  if (bad) atomicOr(flag, (int)F_BAD_INPUT);
  if (!pd) atomicOr(flag, (int)F_NOT_PD);
}

// Ket record of function l (bra record b) under term P with ket transform T
// (n x n) and shift transform Sm (3n x 3n): A' = T^T A T, S' = Sm s
// (python/matel.py:76-87), A'S', A'MA', A'M, tr(M A'), i2, the ket prefactor,
// and the weights scvv_P C_l and scvv_P.
template <int N>
__device__ __forceinline__ void make_ket(const double *__restrict__ b,
                                         const double *__restrict__ T,
                                         const double *__restrict__ Sm, double scvv_p,
                                         const SysConst &sc, double *__restrict__ r) {
  using D = Dim<N>;
  double AT[N][N], Ap[N][N];
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      double s = 0.0;
#pragma unroll
      for (int m = 0; m < N; ++m) s = __builtin_fma(b[D::B_A + sx(i, m)], T[m * N + j], s);
      AT[i][j] = s;
    }
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      double s = 0.0;
#pragma unroll
      for (int m = 0; m < N; ++m) s = __builtin_fma(T[m * N + i], AT[m][j], s);
      Ap[i][j] = s;
    }
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j <= i; ++j) r[D::K_AP + sx(i, j)] = 0.5 * (Ap[i][j] + Ap[j][i]);
  double Sp[D::N3];
#pragma unroll
  for (int u = 0; u < D::N3; ++u) {
    double s = 0.0;
#pragma unroll
    for (int v = 0; v < D::N3; ++v) s = __builtin_fma(Sm[u * D::N3 + v], b[D::B_S + v], s);
    Sp[u] = s;
    r[D::K_SP + u] = s;
  }
  double i2 = 0.0;
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int x = 0; x < 3; ++x) {
      double s = 0.0;
#pragma unroll
      for (int j = 0; j < N; ++j) s = __builtin_fma(Ap[i][j], Sp[3 * j + x], s);
      r[D::K_APSP + 3 * i + x] = s;
      i2 = __builtin_fma(Sp[3 * i + x], s, i2);
    }
  double ApM[N][N];
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      double s = 0.0;
#pragma unroll
      for (int m = 0; m < N; ++m) s = __builtin_fma(Ap[i][m], sc.mass[m * N + j], s);
      ApM[i][j] = s;
      r[D::K_APM + i * N + j] = s;
    }
  double trm = 0.0;
#pragma unroll
  for (int i = 0; i < N; ++i) trm += ApM[i][i];
  r[D::K_TRMP] = trm;
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j <= i; ++j) {
      double s = 0.0, s2 = 0.0;
#pragma unroll
      for (int m = 0; m < N; ++m) {
        s = __builtin_fma(ApM[i][m], Ap[m][j], s);
        s2 = __builtin_fma(ApM[j][m], Ap[m][i], s2);
      }
      r[D::K_NP + sx(i, j)] = 0.5 * (s + s2);
    }
  r[D::K_I2] = i2;
  r[D::K_PL] = b[D::B_PK];
  r[D::K_WC] = scvv_p * b[D::B_C];
  r[D::K_SC] = scvv_p;
}

// Sum of a value over the 256 lanes of a workgroup (4 waves): wave
// reduction by shuffles, then one LDS slot per wave.
__device__ __forceinline__ double block_sum256(double v, double *red4) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  if ((threadIdx.x & 63) == 0) red4[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0) t = (red4[0] + red4[1]) + (red4[2] + red4[3]);
  __syncthreads();
  return t;
}

// Row i, columns [256 chunk, 256 chunk + 256) of H and S: sum the term-group
// partials of the lower triangle (bra >= ket), mirror (python/outprod.py:57-60),
// and the chunk's partial sums of (H C)_i and (S C)_i.  256 lanes.
__device__ __forceinline__ void hs_row_chunk(const double *__restrict__ hpart,
                                             const double *__restrict__ spart, int K, int ts,
                                             const double *__restrict__ bra, int bs, int bc,
                                             int i, int chunk, double *__restrict__ H,
                                             double *__restrict__ S, double *__restrict__ hcp,
                                             double *__restrict__ scp, double *red8) {
  const int j = chunk * 256 + threadIdx.x;
  double hc = 0.0, scv = 0.0;
  if (j < K) {
    const int lo = min(i, j), hi = max(i, j);
    double h = 0.0, s = 0.0;
    for (int z = 0; z < ts; ++z) {
      const size_t o = ((size_t)z * K + lo) * K + hi;
      h += hpart[o];
      s += spart[o];
    }
    H[(size_t)i * K + j] = h;
    S[(size_t)i * K + j] = s;
    const double c = bra[(size_t)j * bs + bc];
    hc = h * c;
    scv = s * c;
  }
  hc = block_sum256(hc, red8);
  scv = block_sum256(scv, red8 + 4);
  if (threadIdx.x == 0) {
    hcp[(size_t)chunk * K + i] = hc;
    scp[(size_t)chunk * K + i] = scv;
  }
}

// Full symmetric H and S for one pair of 64 x 64 blocks (I, J), I >= J, and
// their H C / S C chunk partials, by a 256-lane workgroup.  The term-group
// partials hold element (i, j) at [min(i, j)][max(i, j)] (the reference
// mirrors the lower triangle, python/outprod.py:57-60), so block [J][I] of
// the partials is read row by row (coalesced), summed over the term groups,
// written to H[J][I] as read and, transposed through LDS, to H[I][J].
// hcp[c * K + i] = sum over the 64 columns j of chunk c of H[i][j] C_j; each
// (chunk, row) has one writer, and every sum runs in a fixed order.
// bh, bsl: 64 x 65 LDS blocks; cc: 128 LDS doubles.
__device__ __forceinline__ void hs_tile(const double *__restrict__ hpart,
                                        const double *__restrict__ spart, int K, int ts,
                                        const double *__restrict__ bra, int bs, int bc, int I,
                                        int J, double *__restrict__ H, double *__restrict__ S,
                                        double *__restrict__ hcp, double *__restrict__ scp,
                                        double *bh, double *bsl, double *cc) {
  constexpr int LD = 65;
  const int tid = threadIdx.x, c = tid & 63, r0 = tid >> 6;
  const int i0 = 64 * I, j0 = 64 * J;
  const bool diag = I == J;
  if (tid < 128) {
    const int j = (tid < 64 ? i0 : j0) + (tid & 63);
    cc[tid] = j < K ? bra[(size_t)j * bs + bc] : 0.0;
  }
#pragma unroll 4
  for (int r = r0; r < 64; r += 4) {
    const int row = j0 + r, col = i0 + c;
    double h = 0.0, s = 0.0;
    if (row < K && col < K && (!diag || r <= c)) {
      for (int z = 0; z < ts; ++z) {
        const size_t o = ((size_t)z * K + row) * K + col;
        h += hpart[o];
        s += spart[o];
      }
    }
    bh[r * LD + c] = h;
    bsl[r * LD + c] = s;
  }
  __syncthreads();
#pragma unroll 4
  for (int r = r0; r < 64; r += 4) {
    const int row = j0 + r, col = i0 + c;
    if (row < K && col < K) {
      const int a = diag ? min(r, c) : r, b = diag ? max(r, c) : c;
      H[(size_t)row * K + col] = bh[a * LD + b];
      S[(size_t)row * K + col] = bsl[a * LD + b];
    }
    const int row2 = i0 + r, col2 = j0 + c;
    if (!diag && row2 < K && col2 < K) {
      H[(size_t)row2 * K + col2] = bh[c * LD + r];
      S[(size_t)row2 * K + col2] = bsl[c * LD + r];
    }
  }
  // lanes 0-127: rows j0 + r over chunk I (H, then S); lanes 128-255: rows
  // i0 + r over chunk J (off the diagonal only)
  const int q = tid >> 6, r = tid & 63;
  const double *blk = (q & 1) ? bsl : bh;
  double t = 0.0;
  if (q < 2) {
    if (j0 + r < K) {
      for (int k = 0; k < 64; ++k) {
        const int a = diag ? min(r, k) : r, b = diag ? max(r, k) : k;
        t = __builtin_fma(blk[a * LD + b], cc[k], t);
      }
      ((q & 1) ? scp : hcp)[(size_t)I * K + j0 + r] = t;
    }
  } else if (!diag && i0 + r < K) {
    for (int k = 0; k < 64; ++k) t = __builtin_fma(blk[k * LD + r], cc[64 + k], t);
    ((q & 1) ? scp : hcp)[(size_t)J * K + i0 + r] = t;
  }
}

// The 64 ket records [idx0, idx0 + 64) of a system (one lane per (ket l,
// term p), idx = l nst + p), built in the LDS rows rec (64 x (KS + 1)) and
// written out by the workgroup of 64 lanes as one contiguous run (coalesced).
template <int N>
__device__ __forceinline__ void ket_records_block(const double *__restrict__ bra, int K,
                                                  int nst, const double *__restrict__ sml,
                                                  const double *__restrict__ sms,
                                                  const double *__restrict__ scvv,
                                                  const SysConst &sc, int idx0,
                                                  double *__restrict__ ket, double *rec) {
  using D = Dim<N>;
  constexpr int RS = D::KS + 1;  // odd row stride: the lanes' rows start in different banks
  const int nk = K * nst, idx = idx0 + (int)threadIdx.x;
  double *r = rec + threadIdx.x * RS;
  if (idx < nk) {
    const int l = idx / nst, p = idx - l * nst;
    make_ket<N>(bra + (size_t)l * D::BS, sml + (size_t)p * N * N,
                sms + (size_t)p * D::N3 * D::N3, scvv[p], sc, r);
  }
  for (int f = D::K_TRMP + 1; f < D::KS; ++f) r[f] = 0.0;  // padding
  __syncthreads();
  const int n = min(64, nk - idx0) * D::KS;
  double *out = ket + (size_t)idx0 * D::KS;
  for (int i = threadIdx.x; i < n; i += 64) out[i] = rec[(i / D::KS) * RS + i % D::KS];
}

// H C and S C for the 64 rows [i0, i0 + 64) from the 64-column chunk
// partials, by a 256-lane workgroup: wave q sums the chunks c = q mod 4, the
// four parts are added in order.  ph, ps: 256 LDS doubles each.
__device__ __forceinline__ void hc_rows(const double *__restrict__ hcp,
                                        const double *__restrict__ scp, int nchunk, int K,
                                        int i0, double *__restrict__ HC, double *__restrict__ SC,
                                        double *ph, double *ps) {
  const int lane = threadIdx.x & 63, q = threadIdx.x >> 6;
  const int i = i0 + lane;
  double a = 0.0, b = 0.0;
  if (i < K) {
#pragma unroll 4
    for (int c = q; c < nchunk; c += 4) {
      a += hcp[(size_t)c * K + i];
      b += scp[(size_t)c * K + i];
    }
  }
  ph[threadIdx.x] = a;
  ps[threadIdx.x] = b;
  __syncthreads();
  if (q == 0 && i < K) {
    HC[i] = ((ph[lane] + ph[64 + lane]) + ph[128 + lane]) + ph[192 + lane];
    SC[i] = ((ps[lane] + ps[64 + lane]) + ps[128 + lane]) + ps[192 + lane];
  }
}

// HC, SC from the chunk partials (not stored where HC, SC are null), then
// C^T H C, C^T S C, E = cHc / cSc and 1/cSc (python/outprod.py:65-68).  One workgroup of 1024 lanes per system;
// rh/rs are 1024-entry LDS arrays.
__device__ __forceinline__ void quotient_block(const double *__restrict__ hcp,
                                               const double *__restrict__ scp, int nchunk,
                                               const double *__restrict__ bra, int bs, int bc,
                                               int K, double *__restrict__ HC,
                                               double *__restrict__ SC, double *__restrict__ scal,
                                               double *__restrict__ fout, int *__restrict__ flag,
                                               double *rh, double *rs) {
  double h = 0.0, s = 0.0;
  for (int i = threadIdx.x; i < K; i += blockDim.x) {
    double a = 0.0, b = 0.0;
#pragma unroll 8
    for (int c = 0; c < nchunk; ++c) {
      a += hcp[(size_t)c * K + i];
      b += scp[(size_t)c * K + i];
    }
    if (HC) HC[i] = a;
    if (SC) SC[i] = b;
    const double ci = bra[(size_t)i * bs + bc];
    h = __builtin_fma(ci, a, h);
    s = __builtin_fma(ci, b, s);
  }
  rh[threadIdx.x] = h;
  rs[threadIdx.x] = s;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      rh[threadIdx.x] += rh[threadIdx.x + w];
      rs[threadIdx.x] += rs[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double cHc = rh[0], cSc = rs[0];
    const double E = cHc / cSc;
    scal[0] = E;
    scal[1] = 1.0 / cSc;
    scal[2] = cHc;
    scal[3] = cSc;
    if (fout) fout[0] = E;
    if (!(cSc > 0.0)) atomicOr(flag, (int)F_CSC);
    if (!isfinite(E)) atomicOr(flag, (int)F_NONFINITE);
  }
}

// Sum over the pair tiles of accumulator slot v of bra k.
__device__ __forceinline__ void tile_reduce_one(const double *__restrict__ part, int ntiles,
                                                int K, int nacc, int v, int k,
                                                double *__restrict__ red) {
  double a = 0.0;
  for (int t = 0; t < ntiles; ++t) a += part[((size_t)t * nacc + v) * K + k];
  red[(size_t)v * K + k] = a;
}

// Gradient of function k from its summed accumulators:
// g_theta = norm (a - E b) (MODE_SPLIT) or norm a (MODE_COMBINED, E folded),
// a = 2 C_k [vech(X_H L) + 1.5 sigma_H diag(1/L) | A Y_H]
// (python/make_grad.py:61-79); g_C = 2 norm (HC - E SC) (:85).  gb scales
// the whole gradient (vector-Jacobian product).  g points at the system's x
// layout; the function's parameter block goes to gk when given (a staging
// row), else to its place in g.
template <int N, int MODE>
__device__ __forceinline__ void grad_one(const double *__restrict__ red, int K, int k,
                                         const double *__restrict__ bra,
                                         const double *__restrict__ scal,
                                         const double *__restrict__ HC,
                                         const double *__restrict__ SC, double gb,
                                         double *__restrict__ g, double *gk = nullptr) {
  using D = Dim<N>;
  double a[D::NACC];
#pragma unroll
  for (int v = 0; v < D::NACC; ++v)
    a[v] = (MODE == MODE_SPLIT || v < D::A_XS || v == D::A_SH) ? red[(size_t)v * K + k] : 0.0;
  const double *r = bra + (size_t)k * D::BS;
  const double E = scal[0], nrm = scal[1];
  const double fac = gb * nrm * 2.0 * r[D::B_C];
  double oh[D::BLK];
  grad_L<N>(a + D::A_XH, a[D::A_SH], r + D::B_L, r + D::B_LD, oh);
  grad_S<N>(r + D::B_A, a + D::A_YH, oh + D::NS);
  if (!gk) gk = g + (size_t)k * D::BLK;
  if (MODE == MODE_SPLIT) {
    double os[D::BLK];
    grad_L<N>(a + D::A_XS, a[D::A_SS], r + D::B_L, r + D::B_LD, os);
    grad_S<N>(r + D::B_A, a + D::A_YS, os + D::NS);
#pragma unroll
    for (int v = 0; v < D::BLK; ++v) gk[v] = fac * (oh[v] - E * os[v]);
  } else {
#pragma unroll
    for (int v = 0; v < D::BLK; ++v) gk[v] = fac * oh[v];
  }
  g[(size_t)K * D::BLK + k] = gb * 2.0 * nrm * (HC[k] - E * SC[k]);
}

// The gradient blocks of the 64 functions [k0, k0 + 64), formed in the LDS
// rows go (64 x (BLK + 1)) and written out by the workgroup of 64 lanes as
// one contiguous run (coalesced); g_C goes straight to its place.
template <int N, int MODE>
__device__ __forceinline__ void grad_block(const double *__restrict__ red, int K, int k0,
                                           const double *__restrict__ bra,
                                           const double *__restrict__ scal,
                                           const double *__restrict__ HC,
                                           const double *__restrict__ SC, double gb,
                                           double *__restrict__ g, double *go) {
  using D = Dim<N>;
  constexpr int RS = D::BLK + 1;  // odd row stride: the lanes' rows start in different banks
  const int k = k0 + (int)threadIdx.x;
  if (k < K) grad_one<N, MODE>(red, K, k, bra, scal, HC, SC, gb, g, go + threadIdx.x * RS);
  __syncthreads();
  const int n = min(64, K - k0) * D::BLK;
  double *out = g + (size_t)k0 * D::BLK;
  for (int i = threadIdx.x; i < n; i += 64) out[i] = go[(i / D::BLK) * RS + i % D::BLK];
}

}  // namespace necg
