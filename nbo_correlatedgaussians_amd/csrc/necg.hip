// necg.hip -- HIP kernels (gfx950) and the C ABI of include/necg.h.
//
// Evaluation pipeline for one parameter vector x on one stream:
//   prep_bra   one lane per basis function: L, A = L L^T, S, A S, i1, prefactor
//   prep_ket   one lane per (function, term): the ket under term P
//              (A' = T^T A T, S' = sms s, A'S', A'MA', A'M, i2, weights)
//   pair       the hot kernel: lane = bra k, loop over a tile of kets x terms,
//              ket records read wave-uniformly; H/S elements of the lower
//              triangle and the per-bra gradient accumulators (deterministic
//              partial sums per tile, no atomics)
//   hs_finish  sum term-group partials, mirror to full symmetric H and S
//   quotient   HC, SC, C^T H C, C^T S C -> E, 1/C^T S C (one workgroup)
//   grad_final sum tile partials, vech(X L) / A Y transforms, contraction
//              with C -> g (python/make_grad.py:61-90)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "../../include/necg.h"
#include "necg_ctx.hpp"
#include "necg_pair2.hpp"
#include "necg_two.hpp"

using namespace necg;

namespace {


// ------------------------------------------------------------------ prep ----

// One lane per basis function k: the bra record.  (Staging the parameter
// blocks and records through LDS, as prep_ket does, measured slower here:
// the work per lane, not the access pattern, sets this kernel's time.)
template <int N>
__global__ void prep_bra(const double *__restrict__ x, int K, int with_c, SysConst sc,
                         double *__restrict__ bra, int *__restrict__ flag) {
  using D = Dim<N>;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= K) return;
  make_bra<N>(x + (size_t)k * D::BLK, with_c ? x[(size_t)K * D::BLK + k] : 0.0, with_c != 0,
              sc, bra + (size_t)k * D::BS, flag);
}

// One lane per (ket l, term p): the transformed ket record, built in LDS and
// written out by the wave as one contiguous run of 64 records (coalesced).
template <int N>
__global__ void __launch_bounds__(64) prep_ket(const double *__restrict__ bra, int K, int nst,
                                               const double *__restrict__ sml,
                                               const double *__restrict__ sms,
                                               const double *__restrict__ scvv, SysConst sc,
                                               double *__restrict__ ket) {
  __shared__ double rec[64 * (Dim<N>::KS + 1)];
  ket_records_block<N>(bra, K, nst, sml, sms, scvv, sc, blockIdx.x * 64, ket, rec);
}

// ------------------------------------------------------------------ pair ----

// Lane = bra k.  blockIdx.y = ket tile, blockIdx.z = term group.
// part layout: part[((tile * NACC) + v) * K + k], tile = y * TS + z.
// hpart/spart layout: [z][l][k] for bra k >= ket l.

// One wave per SIMD: the term needs ~300 live FP64 values; at this
// occupancy the compiler keeps the overflow in AGPRs instead of scratch.
#ifndef NECG_MIN_WAVES
#define NECG_MIN_WAVES 1
#endif

// Lane = bra k.  blockIdx.y = ket tile, blockIdx.z = term group.
// part layout: part[((tile * NACC) + v) * K + k], tile = y * TS + z.
// hpart/spart layout: [z][l][k] for bra k >= ket l.
template <int N, int MODE, bool R1>
__global__ void __launch_bounds__(64 * pair_waves(N), NECG_MIN_WAVES) pair_kernel(
    const double *__restrict__ bra, const double *__restrict__ ket, int K, int nst, int kt,
    int tps, SysConst sc, const double *__restrict__ scal, double *__restrict__ part,
    double *__restrict__ hpart, double *__restrict__ spart, int *__restrict__ flag) {
  using D = Dim<N>;
  __shared__ double acc_s[PairLds<N, MODE>::ACC];
  __shared__ double kst_s[pair_waves(N) * D::KS];
  const int l0 = blockIdx.y * kt;
  const int p0 = blockIdx.z * tps;
  const size_t tile = (size_t)blockIdx.y * gridDim.z + blockIdx.z;
  pair_tile<N, MODE, R1, true>(bra, ket, K, nst, blockIdx.x * 64, l0, min(K, l0 + kt), p0,
                               min(nst, p0 + tps), sc, (MODE == MODE_COMBINED) ? scal[0] : 0.0,
                               part + tile * D::NACC * K, hpart + (size_t)blockIdx.z * K * K,
                               spart + (size_t)blockIdx.z * K * K, flag, acc_s, 0x7fffffff,
                               kst_s);
}

// Two-sided pair pass over the lower triangle.  blockIdx.x indexes the tile
// list (TriTile words: bra tile, ket tile, ket start <= bra tile end, whole
// or half); blockIdx.y = term group z of gridDim.y.
// bpart layout: [(ket tile * ts + z) * NACC + v] * K + k, the second half of
// a split tile at xpart [(e * ts + z) * NACC + v] * K + k instead;
// kpart layout: [((bra tile * ts + z) * K + l) * NACC + v].
template <int N, bool R1>
__global__ void __launch_bounds__(64 * pair_waves(N), NECG_MIN_WAVES) pair2_kernel(
    const double *__restrict__ bra, const double *__restrict__ ket, int K, int nst, int kt,
    int tps, SysConst sc, const int *__restrict__ tiles, const unsigned *__restrict__ tperm,
    double *__restrict__ bpart, double *__restrict__ xpart, double *__restrict__ kpart,
    double *__restrict__ hpart, double *__restrict__ spart, int *__restrict__ flag) {
  using D = Dim<N>;
  static_assert(Two<N>::NCH <= 4, "two-sided form: at most 64 accumulator slots");
  __shared__ double acc_s[PairLds<N, MODE_SPLIT>::ACC];
  __shared__ double kacc_s[pair_waves(N) * 64];
  __shared__ double bra_s[BraLds<N>::NF * 64];
  __shared__ double kst_s[pair_waves(N) * 64];
  const TriTile t = tri_tile(tiles, blockIdx.x, K, kt);
  const int z = blockIdx.y, ts = gridDim.y;
  const int p0 = z * tps;
  double *bp = t.e < 0 ? bpart + (size_t)(t.by * ts + z) * D::NACC * K
                       : xpart + (size_t)(t.e * ts + z) * D::NACC * K;
  double *kp = kpart + (size_t)(t.bx * ts + z) * K * D::NACC;
  double *hp = hpart + (size_t)z * K * K, *sp = spart + (size_t)z * K * K;
  if (t.fold == 1)   // uniform over the workgroup
    pair2_fold_tile<N, R1>(bra, ket, K, nst, t.k0, t.l0, t.l1, p0, min(nst, p0 + tps), sc,
                           tperm, bp, kp, hp, sp, flag, acc_s, kacc_s, bra_s, kst_s);
  else if (t.fold == 2)
    pair2_diag_tile<N, R1>(bra, ket, K, nst, t.k0, p0, min(nst, p0 + tps), sc, tperm, bp, hp,
                           sp, flag, acc_s, kacc_s, bra_s, kst_s);
  else
    pair2_tile<N, R1>(bra, ket, K, nst, t.k0, t.l0, t.l1, t.lup, p0, min(nst, p0 + tps), sc,
                      tperm, bp, kp, hp, sp, flag, acc_s, kacc_s, bra_s, kst_s);
}

// Element-only pass (MODE_ELEMENTS / MODE_TV) over the same lower-triangle
// tile list: pairs with bra k >= ket l only.
template <int N, int MODE, bool R1>
__global__ void __launch_bounds__(64 * pair_waves(N), NECG_MIN_WAVES) pair_tri_kernel(
    const double *__restrict__ bra, const double *__restrict__ ket, int K, int nst, int kt,
    int tps, SysConst sc, const int *__restrict__ tiles, double *__restrict__ hpart,
    double *__restrict__ spart, int *__restrict__ flag) {
  static_assert(!mode_has_grad(MODE), "element modes only");
  __shared__ double acc_s[PairLds<N, MODE>::ACC];
  __shared__ double kst_s[pair_waves(N) * Dim<N>::KS];
  const TriTile t = tri_tile(tiles, blockIdx.x, K, kt);
  const int z = blockIdx.y;
  const int p0 = z * tps;
  // a fold tile of sweeps [l0, l1) covers kets k0 + [l0, l1) and k0 + 32 + [l0, l1) of the
  // diagonal block: here, as two ordinary runs of kets (bra >= ket), which
  // include the diagonals; a diagonal tile has nothing left to do
  if (t.fold == 2) return;
  const int runs = t.fold ? 2 : 1;
  const int base = t.fold ? t.k0 : 0;
  for (int r = 0; r < runs; ++r)
    pair_tile<N, MODE, R1, true>(bra, ket, K, nst, t.k0, base + 32 * r + t.l0,
                                 base + 32 * r + t.l1, p0, min(nst, p0 + tps), sc, 0.0, nullptr,
                                 hpart + (size_t)z * K * K, spart + (size_t)z * K * K, flag,
                                 acc_s, t.lup, kst_s);
}

// Two-sided partial sums per (function, slot).  Grid (ceil(K/64), nv),
// 256 lanes: 64 functions per workgroup, four waves over the partials.
__global__ void __launch_bounds__(256) tile_reduce2(const double *__restrict__ bpart,
                                                    const double *__restrict__ kpart, int K,
                                                    int kt, int ts, int nacc,
                                                    double *__restrict__ red,
                                                    const double *__restrict__ xpart,
                                                    const int *__restrict__ xr, int fold) {
  __shared__ double sh[256];
  tile_reduce2_block(bpart, kpart, K, kt, ts, nacc, blockIdx.y, blockIdx.x * 64, red, sh, xpart,
                     xr, fold);
}

// Full symmetric H and S and the per-chunk (64 columns) partials of H C and
// S C.  One 256-lane workgroup per block pair (I, J), I >= J: blockIdx.x
// enumerates them row by row, (0,0), (1,0), (1,1), (2,0), ...
__global__ void __launch_bounds__(256) hs_finish(
    const double *__restrict__ hpart, const double *__restrict__ spart, int K, int ts,
    const double *__restrict__ bra, int bs, int bc, double *__restrict__ H,
    double *__restrict__ S, double *__restrict__ hcp, double *__restrict__ scp) {
  __shared__ double bh[64 * 65], bsl[64 * 65], cc[128];
  const int t = blockIdx.x;
  int I = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
  while ((I + 1) * (I + 2) / 2 <= t) ++I;
  while (I * (I + 1) / 2 > t) --I;
  const int J = t - I * (I + 1) / 2;
  hs_tile(hpart, spart, K, ts, bra, bs, bc, I, J, H, S, hcp, scp, bh, bsl, cc);
}

// H C and S C from the 64-column chunk partials: 64 rows per 256-lane
// workgroup, wave q summing the chunks c = q mod 4, the four parts added in
// order.
__global__ void __launch_bounds__(256) hc_sum(const double *__restrict__ hcp,
                                              const double *__restrict__ scp, int nchunk, int K,
                                              double *__restrict__ HC, double *__restrict__ SC) {
  __shared__ double ph[256], ps[256];
  hc_rows(hcp, scp, nchunk, K, blockIdx.x * 64, HC, SC, ph, ps);
}

__global__ void __launch_bounds__(1024) quotient(
    const double *__restrict__ hcp, const double *__restrict__ scp, int nchunk,
    const double *__restrict__ bra, int bs, int bc, int K, double *__restrict__ HC,
    double *__restrict__ SC, double *__restrict__ scal, double *__restrict__ fout,
    int *__restrict__ flag) {
  __shared__ double rh[1024], rs[1024];
  quotient_block(hcp, scp, nchunk, bra, bs, bc, K, HC, SC, scal, fout, flag, rh, rs);
}

// Grid (ceil(K/256), slots), 256 lanes; coalesced over k.
__global__ void __launch_bounds__(256) tile_reduce(const double *__restrict__ part, int ntiles,
                                                   int K, int nacc, double *__restrict__ red) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < K) tile_reduce_one(part, ntiles, K, nacc, blockIdx.y, k, red);
}

// One lane per function; the 64 blocks of a workgroup are formed in LDS and
// written out as one contiguous run (coalesced).
template <int N, int MODE>
__global__ void __launch_bounds__(64) grad_final(const double *__restrict__ red, int K,
                                                 const double *__restrict__ bra,
                                                 const double *__restrict__ scal,
                                                 const double *__restrict__ HC,
                                                 const double *__restrict__ SC,
                                                 const double *__restrict__ gbar,
                                                 double *__restrict__ g) {
  __shared__ double go[64 * (Dim<N>::BLK + 1)];
  grad_block<N, MODE>(red, K, blockIdx.x * 64, bra, scal, HC, SC, gbar ? gbar[0] : 1.0, g, go);
}

// ----------------------------------------------------------- queries -----

// Per-term matel / gradel for (stf, i, j) triples (python/matel.py:28,
// python/gradel.py:27).  The ket record is rebuilt in registers per query
// with unit weights.
template <int N>
__global__ void query_kernel(const double *__restrict__ bra, int K, int nst,
                             const double *__restrict__ sml, const double *__restrict__ sms,
                             SysConst sc, int nq, const int32_t *__restrict__ q,
                             double *__restrict__ out4, double *__restrict__ tgvl,
                             double *__restrict__ tgvls, int *__restrict__ flag) {
  using D = Dim<N>;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nq) return;
  const int p = q[3 * t], i = q[3 * t + 1], j = q[3 * t + 2];
  if (p < 0 || p >= nst || i < 0 || i >= K || j < 0 || j >= K) {
    atomicOr(flag, (int)F_BAD_INPUT);
    return;
  }
  double kr[D::KS];
  make_ket<N>(bra + (size_t)j * D::BS, sml + (size_t)p * N * N,
              sms + (size_t)p * D::N3 * D::N3, 1.0, sc, kr);
  kr[D::K_WC] = 1.0;
  BraReg<N> b;
  load_bra<N>(b, bra + (size_t)i * D::BS);
  RegSink<N> acc;
  acc.zero();
  double hel = 0.0, sel = 0.0;
  Parts parts;
  bool ok;
  if (tgvl)
    ok = pair_term<N, MODE_SPLIT>(b, kr, sc, 0.0, hel, sel, acc, &parts);
  else
    ok = pair_term<N, MODE_ELEMENTS>(b, kr, sc, 0.0, hel, sel, acc, &parts);
  if (!ok) atomicOr(flag, (int)F_NOT_PD);
  if (out4) {
    out4[4 * t + 0] = parts.ov;
    out4[4 * t + 1] = parts.tk;
    out4[4 * t + 2] = parts.ert;
    out4[4 * t + 3] = parts.zht;
  }
  if (tgvl) {
    const double *r = bra + (size_t)i * D::BS;
    const double two = (i == j) ? 2.0 : 1.0;  // python/gradel.py:388-390
    double o[D::BLK];
    grad_L<N>(acc.a + D::A_XH, acc.a[D::A_SH], r + D::B_L, r + D::B_LD, o);
    grad_S<N>(r + D::B_A, acc.a + D::A_YH, o + D::NS);
#pragma unroll
    for (int v = 0; v < D::BLK; ++v) tgvl[(size_t)t * D::BLK + v] = two * o[v];
    grad_L<N>(acc.a + D::A_XS, acc.a[D::A_SS], r + D::B_L, r + D::B_LD, o);
    grad_S<N>(r + D::B_A, acc.a + D::A_YS, o + D::NS);
#pragma unroll
    for (int v = 0; v < D::BLK; ++v) tgvls[(size_t)t * D::BLK + v] = two * o[v];
  }
}

// Materialised DH/DS (python/make_grad.py:34-55): lane per ordered (l, k),
// all terms, bra l.  DH[(l*blk + r)*K + k].
template <int N>
__global__ void dhds_kernel(const double *__restrict__ bra, const double *__restrict__ ket,
                            int K, int nst, SysConst sc, double *__restrict__ DH,
                            double *__restrict__ DS, int *__restrict__ flag) {
  using D = Dim<N>;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;  // ket (coalesced writes)
  const int l = blockIdx.y;                              // bra
  if (k >= K) return;
  BraReg<N> b;
  load_bra<N>(b, bra + (size_t)l * D::BS);
  RegSink<N> acc;
  acc.zero();
  double hel = 0.0, sel = 0.0;
  bool ok = true;
  for (int p = 0; p < nst; ++p) {
    // weight scvv_P only (the ket table carries scvv_P C_k in K_WC)
    const double *kr = ket + ((size_t)k * nst + p) * D::KS;
    double rec[D::KS];
#pragma unroll
    for (int v = 0; v < D::KS; ++v) rec[v] = kr[v];
    rec[D::K_WC] = rec[D::K_SC];
    ok &= pair_term<N, MODE_SPLIT>(b, rec, sc, 0.0, hel, sel, acc, nullptr);
  }
  if (!ok) atomicOr(flag, (int)F_NOT_PD);
  const double *r = bra + (size_t)l * D::BS;
  const double two = (l == k) ? 2.0 : 1.0;
  double o[D::BLK];
  grad_L<N>(acc.a + D::A_XH, acc.a[D::A_SH], r + D::B_L, r + D::B_LD, o);
  grad_S<N>(r + D::B_A, acc.a + D::A_YH, o + D::NS);
#pragma unroll
  for (int v = 0; v < D::BLK; ++v) DH[((size_t)l * D::BLK + v) * K + k] = two * o[v];
  grad_L<N>(acc.a + D::A_XS, acc.a[D::A_SS], r + D::B_L, r + D::B_LD, o);
  grad_S<N>(r + D::B_A, acc.a + D::A_YS, o + D::NS);
#pragma unroll
  for (int v = 0; v < D::BLK; ++v) DS[((size_t)l * D::BLK + v) * K + k] = two * o[v];
}

}  // namespace

// =================================================================== host ===

// necg_ctx is defined in necg_ctx.hpp

namespace {

int hip_status(hipError_t e) {
  if (e == hipSuccess) return NECG_OK;
  if (e == hipErrorOutOfMemory) return NECG_ERR_OOM;
  return NECG_ERR_DEVICE;
}

#define HIP_TRY(expr)                         \
  do {                                        \
    int rc__ = hip_status((expr));            \
    if (rc__ != NECG_OK) return rc__;         \
  } while (0)

int dev_alloc(double **p, size_t count) {
  if (*p) {
    (void)hipFree(*p);
    *p = nullptr;
  }
  if (count == 0) return NECG_OK;
  return hip_status(hipMalloc((void **)p, count * sizeof(double)));
}

// Workspace buffers of the context: device allocations, or in
// allocation-free mode slices of the arena (256-byte aligned; nothing is
// freed, the previous pointers belong to the context's own workspace).
int ws_alloc(necg_ctx *c, void **p, size_t bytes) {
  if (!c->arena) return dev_alloc((double **)p, (bytes + sizeof(double) - 1) / sizeof(double));
  *p = nullptr;
  const size_t take = (bytes + 255) & ~(size_t)255;
  if (c->arena_used + take > c->arena_cap) return NECG_ERR_SCRATCH;
  *p = c->arena + c->arena_used;
  c->arena_used += take;
  return NECG_OK;
}
int ws_alloc(necg_ctx *c, double **p, size_t count) {
  return ws_alloc(c, (void **)p, count * sizeof(double));
}

int bra_stride(int n) {
  switch (n) {
    case 1: return Dim<1>::BS;
    case 2: return Dim<2>::BS;
    case 3: return Dim<3>::BS;
    case 4: return Dim<4>::BS;
    case 5: return Dim<5>::BS;
    default: return Dim<6>::BS;
  }
}
int ket_stride(int n) {
  switch (n) {
    case 1: return Dim<1>::KS;
    case 2: return Dim<2>::KS;
    case 3: return Dim<3>::KS;
    case 4: return Dim<4>::KS;
    case 5: return Dim<5>::KS;
    default: return Dim<6>::KS;
  }
}
int bra_c_off(int n) {
  switch (n) {
    case 1: return Dim<1>::B_C;
    case 2: return Dim<2>::B_C;
    case 3: return Dim<3>::B_C;
    case 4: return Dim<4>::B_C;
    case 5: return Dim<5>::B_C;
    default: return Dim<6>::B_C;
  }
}

// The capacity fields drop to 0 before any buffer they cover is freed and
// rise only after every allocation succeeded, so a failed allocation leaves a
// context whose next call reallocates instead of launching on null buffers.
int ensure_K(necg_ctx *c, int K) {
  if (K <= c->Kcap) return NECG_OK;
  if (c->pins && !c->arena) return NECG_ERR_ARG;   // a captured graph holds these addresses
  int rc;
  const size_t k = (size_t)K;
  c->Kcap = 0;
  if ((rc = ws_alloc(c, &c->bra, k * bra_stride(c->n)))) return rc;
  if ((rc = ws_alloc(c, &c->ket, k * c->nst * ket_stride(c->n)))) return rc;
  if ((rc = ws_alloc(c, &c->H, k * k))) return rc;
  if ((rc = ws_alloc(c, &c->S, k * k))) return rc;
  if ((rc = ws_alloc(c, &c->HC, k))) return rc;
  if ((rc = ws_alloc(c, &c->SC, k))) return rc;
  if ((rc = ws_alloc(c, &c->scal, 8))) return rc;
  const size_t nchunk = (k + 63) / 64;  // H C / S C partials per 64 columns
  if ((rc = ws_alloc(c, &c->hcp, nchunk * k))) return rc;
  if ((rc = ws_alloc(c, &c->scp, nchunk * k))) return rc;
  if ((rc = ws_alloc(c, &c->red, (size_t)Dim<kMaxN>::NACC * k))) return rc;
  c->Kcap = K;
  return NECG_OK;
}

int ensure_part(necg_ctx *c, size_t count) {
  if (count <= c->part_cap) return NECG_OK;
  if (c->pins && !c->arena) return NECG_ERR_ARG;
  c->part_cap = 0;
  int rc = ws_alloc(c, &c->part, count);
  if (rc) return rc;
  c->part_cap = count;
  return NECG_OK;
}
int ensure_hpart(necg_ctx *c, size_t count) {
  if (count <= c->hpart_cap) return NECG_OK;
  if (c->pins && !c->arena) return NECG_ERR_ARG;
  c->hpart_cap = 0;
  int rc = ws_alloc(c, &c->hpart, count);
  if (rc) return rc;
  rc = ws_alloc(c, &c->spart, count);
  if (rc) return rc;
  c->hpart_cap = count;
  return NECG_OK;
}

// Launch shape: lanes are bras (64 per wave); a work item covers `kt` kets
// and `tps` terms.  Aim for a few thousand waves so that all 256 CUs stay
// busy without a long tail.
void choose_tiling(necg_ctx *c, int K, int &kt, int &ts, int &w) {
  const int nst = c->nst;
  w = pair_waves(c->n);
  const long bra_waves = (K + 63) / 64;
  const long target = 4096;
  if (c->req_kt > 0) {
    kt = c->req_kt;
  } else {
    long ket_tiles = std::max<long>(1, target / bra_waves);
    kt = (int)std::max<long>(1, (K + ket_tiles - 1) / ket_tiles);
  }
  if (c->req_ts > 0) {
    ts = std::min(c->req_ts, nst);
  } else {
    long waves = bra_waves * ((K + kt - 1) / kt);
    ts = 1;
    while (waves * ts < target && ts < nst) ts *= 2;
    ts = std::min(ts, nst);
  }
}

template <int N>
int run_prep(necg_ctx *c, int K, const double *x, int with_c, hipStream_t st) {
  c->gen++;
  hipLaunchKernelGGL((prep_bra<N>), dim3((K + 255) / 256), dim3(256), 0, st, x, K, with_c,
                     c->sc, c->bra, c->d_flag);
  const int nk = K * c->nst;
  hipLaunchKernelGGL((prep_ket<N>), dim3((nk + 63) / 64), dim3(64), 0, st, c->bra, K, c->nst,
                     c->d_sml, c->d_sms, c->d_scvv, c->sc, c->ket);
  return hip_status(hipGetLastError());
}

// Runs the pair kernel in MODE and the finishing kernels.  f, H, S, g may be
// null where the mode does not produce them.
// Launch shape and workspace of the per-ordering pair pass in MODE.
template <int N, int MODE>
int reserve_pairs(necg_ctx *c, int K, int &kt, int &ts, int &w, int &tps) {
  choose_tiling(c, K, kt, ts, w);
  tps = (c->nst + ts - 1) / ts;
  ts = (c->nst + tps - 1) / tps;
  const int ntiles = ((K + kt - 1) / kt) * ts;
  int rc;
  if (mode_has_grad(MODE)) {
    if ((rc = ensure_part(c, (size_t)ntiles * Dim<N>::NACC * K))) return rc;
  }
  if (MODE != MODE_COMBINED) {
    if ((rc = ensure_hpart(c, (size_t)ts * K * K))) return rc;
  }
  return NECG_OK;
}

template <int N, int MODE>
int run_pairs(necg_ctx *c, int K, hipStream_t st, int &ntiles, int &ts_out) {
  int kt, ts, w, tps, rc;
  if ((rc = reserve_pairs<N, MODE>(c, K, kt, ts, w, tps))) return rc;
  const int kts = (K + kt - 1) / kt;
  const int threads = 64 * w;
  dim3 grid((K + 63) / 64, kts, ts);
  ntiles = kts * ts;
  ts_out = ts;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (c->timing) {
    if ((size_t)(2 * c->nev + 2) > c->ev.size()) {
      hipEvent_t a, b;
      HIP_TRY(hipEventCreate(&a));
      HIP_TRY(hipEventCreate(&b));
      c->ev.push_back(a);
      c->ev.push_back(b);
    }
    e0 = c->ev[2 * c->nev];
    e1 = c->ev[2 * c->nev + 1];
    c->nev++;
    HIP_TRY(hipEventRecord(e0, st));
  }
  if (c->sc.mrank1)
    hipLaunchKernelGGL((pair_kernel<N, MODE, true>), grid, dim3(threads), 0, st, c->bra, c->ket,
                       K, c->nst, kt, tps, c->sc, c->scal, c->part, c->hpart, c->spart,
                       c->d_flag);
  else
    hipLaunchKernelGGL((pair_kernel<N, MODE, false>), grid, dim3(threads), 0, st, c->bra, c->ket,
                       K, c->nst, kt, tps, c->sc, c->scal, c->part, c->hpart, c->spart,
                       c->d_flag);
  if (c->timing) HIP_TRY(hipEventRecord(e1, st));
  c->last_kt = kt;
  c->last_ts = ts;
  c->last_w = w;
  return hip_status(hipGetLastError());
}

// Full symmetric matrices from the element partials, into H and S (the
// context's own by default).
int finish_hs(necg_ctx *c, int K, int ts, hipStream_t st, double *H = nullptr,
              double *S = nullptr) {
  const int bs = bra_stride(c->n), bc = bra_c_off(c->n);
  const int nb = (K + 63) / 64;
  hipLaunchKernelGGL(hs_finish, dim3(nb * (nb + 1) / 2), dim3(256), 0, st, c->hpart, c->spart,
                     K, ts, c->bra, bs, bc, H ? H : c->H, S ? S : c->S, c->hcp, c->scp);
  return hip_status(hipGetLastError());
}

int finish_quotient(necg_ctx *c, int K, double *f, hipStream_t st) {
  const int bs = bra_stride(c->n), bc = bra_c_off(c->n);
  hipLaunchKernelGGL(hc_sum, dim3((K + 63) / 64), dim3(256), 0, st, c->hcp, c->scp,
                     (K + 63) / 64, K, c->HC, c->SC);
  // the quotient over the summed vectors (one chunk: H C and S C themselves)
  hipLaunchKernelGGL(quotient, dim3(1), dim3(1024), 0, st, c->HC, c->SC, 1, c->bra, bs, bc, K,
                     nullptr, nullptr, c->scal, f, c->d_flag);
  return hip_status(hipGetLastError());
}

template <int N, int MODE>
int finish_grad(necg_ctx *c, int K, int ntiles, const double *gbar, double *g, hipStream_t st) {
  using D = Dim<N>;
  const int nv = (MODE == MODE_SPLIT) ? D::NACC : D::A_SH + 1;  // MODE 1 uses slots < A_XS and A_SH
  hipLaunchKernelGGL(tile_reduce, dim3((K + 255) / 256, nv), dim3(256), 0, st, c->part, ntiles, K,
                     D::NACC, c->red);
  hipLaunchKernelGGL((grad_final<N, MODE>), dim3((K + 63) / 64), dim3(64), 0, st, c->red, K,
                     c->bra, c->scal, c->HC, c->SC, gbar, g);
  return hip_status(hipGetLastError());
}

// ----------------------------------------------------- lower triangle ----

// Accumulator slots per function of the two-sided pass for n particles.
int nacc_of(int n) {
  switch (n) {
    case 1: return Dim<1>::NACC;
    case 2: return Dim<2>::NACC;
    case 3: return Dim<3>::NACC;
    case 4: return Dim<4>::NACC;
    case 5: return Dim<5>::NACC;
    default: return Dim<6>::NACC;
  }
}

// Ket tile and term groups of the lower-triangle passes (tri_choose: the
// tiling with the shortest modelled launch).  In an allocation-free call the
// two-sided pass's partials must fit what the arena has left after the
// per-K buffers (the same tiling as an ordinary call whenever that fits).
// The choice is kept for the next evaluation with the same inputs (walking
// the candidates' tile lists costs about a millisecond of host time at
// K = 2048).
void choose_tiling2(necg_ctx *c, int K, int &kt, int &ts) {
  const int ncu = std::max(1, c->ncu);
  size_t budget = 0;
  if (c->arena) {   // what is left after the per-K buffers, less the tile list
    if (c->arena_budget == 0) {
      const size_t left = c->arena_cap > c->arena_used ? c->arena_cap - c->arena_used : 0;
      const size_t list = (size_t)16 * ((K + 63) / 64) * ((K + 63) / 64 + 1) * 32 + 4096;
      c->arena_budget = left > list ? left - list : 1;
    }
    budget = c->arena_budget;
  }
  if (c->tc_K == K && c->tc_req_kt == c->req_kt && c->tc_req_ts == c->req_ts &&
      c->tc_ncu == ncu && c->tc_fold == c->fold_mode && c->tc_budget == budget) {
    kt = c->tc_kt;
    ts = c->tc_ts;
    return;
  }
  const TriChoice ch = tri_choose(K, c->nst, c->n, pair_waves(c->n), c->fold_mode, ncu,
                                  c->req_kt, c->req_ts, nacc_of(c->n), budget);
  kt = ch.kt;
  ts = ch.ts;
  c->tc_K = K;
  c->tc_req_kt = c->req_kt;
  c->tc_req_ts = c->req_ts;
  c->tc_ncu = ncu;
  c->tc_fold = c->fold_mode;
  c->tc_budget = budget;
  c->tc_kt = kt;
  c->tc_ts = ts;
}

// The tile list for (K, kt) on the device (tri_tile_list: tri_tiles order,
// the last round's tiles split when that leaves CUs idle), then the
// per-bra-tile extra-slot ranges.
int ensure_tiles(necg_ctx *c, int K, int kt, hipStream_t st) {
  const int fold = tri_fold(kt, c->fold_mode) ? 1 : 0;
  if (c->tiles_K == K && c->tiles_kt == kt && c->tiles_fold == fold) return NECG_OK;
  if (c->pins && !c->arena) return NECG_ERR_ARG;   // a captured graph reads the current list
  if ((K + 63) / 64 > 0xffff) return NECG_ERR_UNSUPPORTED;   // bra tile index has 16 bits
  std::vector<int> full, xr;
  const int nsplit = tri_tile_list(K, kt, std::max(1, c->ncu), fold, full, xr);
  const size_t ntile_words = full.size();
  full.insert(full.end(), xr.begin(), xr.end());
  c->tiles_K = 0;   // no valid list until the copy below completes
  if (c->arena) {
    // the list goes into the arena on the call's stream, after the kernels of
    // earlier calls on that stream that read the arena; its host copy is kept
    // on the context so that the pending copy's source stays valid
    int *d = nullptr;
    int rc = ws_alloc(c, (void **)&d, full.size() * sizeof(int));
    if (rc) return rc;
    const long long key = ((long long)K << 32) ^ ((long long)kt << 16) ^
                          ((long long)fold << 15) ^ (long long)c->ncu;
    std::vector<int> &keep = c->arena_lists[key];   // the same list for the same key
    if (keep.empty()) keep = full;
    HIP_TRY(hipMemcpyAsync(d, keep.data(), keep.size() * sizeof(int), hipMemcpyHostToDevice, st));
    c->d_tiles = d;
    c->tiles_cap = keep.size();
    c->tiles_K = K;
    c->tiles_kt = kt;
    c->tiles_fold = fold;
    c->ntri = (int)(ntile_words / 2);
    c->nsplit = nsplit;
    c->d_xr = c->d_tiles + ntile_words;
    return NECG_OK;
  }
  if (full.size() > c->tiles_cap) {
    if (c->d_tiles) (void)hipFree(c->d_tiles);
    c->d_tiles = nullptr;
    c->tiles_cap = 0;
    HIP_TRY(hipMalloc((void **)&c->d_tiles, full.size() * sizeof(int)));
    c->tiles_cap = full.size();
  }
  HIP_TRY(hipMemcpy(c->d_tiles, full.data(), full.size() * sizeof(int), hipMemcpyHostToDevice));
  c->tiles_K = K;
  c->tiles_kt = kt;
  c->tiles_fold = fold;
  c->ntri = (int)(ntile_words / 2);
  c->nsplit = nsplit;
  c->d_xr = c->d_tiles + ntile_words;
  return NECG_OK;
}

int ensure_kpart(necg_ctx *c, size_t count) {
  if (count <= c->kpart_cap) return NECG_OK;
  if (c->pins && !c->arena) return NECG_ERR_ARG;
  c->kpart_cap = 0;
  int rc = ws_alloc(c, &c->kpart, count);
  if (rc) return rc;
  c->kpart_cap = count;
  return NECG_OK;
}

// A pinned context (necg_pin_workspace) runs the lower-triangle passes only
// at the K whose tile list is on the device: checked before any launch, so a
// refused call leaves the captured state untouched.
bool pinned_elsewhere(const necg_ctx *c, int K) {
  return c->pins && (K > c->Kcap || K != c->tiles_K);
}

// Timing events around one pair launch (necg_timing).
int timing_begin(necg_ctx *c, hipStream_t st, hipEvent_t &e1) {
  e1 = nullptr;
  if (!c->timing) return NECG_OK;
  if ((size_t)(2 * c->nev + 2) > c->ev.size()) {
    hipEvent_t a, b;
    HIP_TRY(hipEventCreate(&a));
    HIP_TRY(hipEventCreate(&b));
    c->ev.push_back(a);
    c->ev.push_back(b);
  }
  hipEvent_t e0 = c->ev[2 * c->nev];
  e1 = c->ev[2 * c->nev + 1];
  c->nev++;
  HIP_TRY(hipEventRecord(e0, st));
  return NECG_OK;
}

// Two-sided gradient pair pass (MODE_SPLIT partials) over the lower triangle.
// Launch shape, tile list and workspace of the two-sided pair pass.
template <int N>
int reserve_pairs2(necg_ctx *c, int K, hipStream_t st, int &kt, int &ts, int &tps) {
  using D = Dim<N>;
  choose_tiling2(c, K, kt, ts);
  tps = (c->nst + ts - 1) / ts;
  ts = (c->nst + tps - 1) / tps;
  const int nkt = (K + kt - 1) / kt, nbt = (K + 63) / 64;
  int rc;
  if ((rc = ensure_tiles(c, K, kt, st))) return rc;
  if ((rc = ensure_part(c, (size_t)(nkt + c->nsplit) * ts * D::NACC * K))) return rc;
  if ((rc = ensure_kpart(c, (size_t)nbt * ts * K * D::NACC))) return rc;
  return ensure_hpart(c, (size_t)ts * K * K);
}

template <int N>
int run_pairs2(necg_ctx *c, int K, hipStream_t st, int &kt, int &ts) {
  using D = Dim<N>;
  int tps, rc;
  if ((rc = reserve_pairs2<N>(c, K, st, kt, ts, tps))) return rc;
  const int nkt = (K + kt - 1) / kt;
  double *xpart = c->part + (size_t)nkt * ts * D::NACC * K;
  hipEvent_t e1;
  if ((rc = timing_begin(c, st, e1))) return rc;
  if (c->sc.mrank1)
    hipLaunchKernelGGL((pair2_kernel<N, true>), dim3(c->ntri, ts), dim3(64 * pair_waves(N)), 0,
                       st, c->bra, c->ket, K, c->nst, kt, tps, c->sc, c->d_tiles, c->d_tperm,
                       c->part, xpart, c->kpart, c->hpart, c->spart, c->d_flag);
  else
    hipLaunchKernelGGL((pair2_kernel<N, false>), dim3(c->ntri, ts), dim3(64 * pair_waves(N)), 0,
                       st, c->bra, c->ket, K, c->nst, kt, tps, c->sc, c->d_tiles, c->d_tperm,
                       c->part, xpart, c->kpart, c->hpart, c->spart, c->d_flag);
  if (c->timing) HIP_TRY(hipEventRecord(e1, st));
  c->last_kt = kt;
  c->last_ts = ts;
  c->last_w = pair_waves(N);
  return hip_status(hipGetLastError());
}

// Element-only pass (MODE_ELEMENTS or MODE_TV) over the lower triangle.
template <int N, int MODE>
int run_tri(necg_ctx *c, int K, hipStream_t st, int &ts) {
  int kt;
  choose_tiling2(c, K, kt, ts);
  const int tps = (c->nst + ts - 1) / ts;
  ts = (c->nst + tps - 1) / tps;
  int rc;
  if ((rc = ensure_tiles(c, K, kt, st))) return rc;
  if ((rc = ensure_hpart(c, (size_t)ts * K * K))) return rc;
  hipEvent_t e1;
  if ((rc = timing_begin(c, st, e1))) return rc;
  if (c->sc.mrank1)
    hipLaunchKernelGGL((pair_tri_kernel<N, MODE, true>), dim3(c->ntri, ts),
                       dim3(64 * pair_waves(N)), 0, st, c->bra, c->ket, K, c->nst, kt, tps, c->sc,
                       c->d_tiles, c->hpart, c->spart, c->d_flag);
  else
    hipLaunchKernelGGL((pair_tri_kernel<N, MODE, false>), dim3(c->ntri, ts),
                       dim3(64 * pair_waves(N)), 0, st, c->bra, c->ket, K, c->nst, kt, tps, c->sc,
                       c->d_tiles, c->hpart, c->spart, c->d_flag);
  if (c->timing) HIP_TRY(hipEventRecord(e1, st));
  c->last_kt = kt;
  c->last_ts = ts;
  c->last_w = pair_waves(N);
  return hip_status(hipGetLastError());
}

template <int N>
int finish_grad2(necg_ctx *c, int K, int kt, int ts, const double *gbar, double *g,
                 hipStream_t st) {
  using D = Dim<N>;
  const int nkt = (K + kt - 1) / kt;
  hipLaunchKernelGGL(tile_reduce2, dim3((K + 63) / 64, D::NACC), dim3(256), 0, st, c->part,
                     c->kpart, K, kt, ts, D::NACC, c->red,
                     c->part + (size_t)nkt * ts * D::NACC * K, c->nsplit ? c->d_xr : nullptr,
                     c->tiles_fold);
  hipLaunchKernelGGL((grad_final<N, MODE_SPLIT>), dim3((K + 63) / 64), dim3(64), 0, st, c->red,
                     K, c->bra, c->scal, c->HC, c->SC, gbar, g);
  return hip_status(hipGetLastError());
}

// Whether the gradient pair pass takes the two-sided form.
bool use_two(const necg_ctx *c) {
  return c->two_ok && c->pair_mode != 2 && c->n <= 5;
}

// Gradient pair pass of x (bra/ket tables ready): H/S partials and the
// gradient partials in either form; records the form for the finish.
template <int N>
int grad_pass(necg_ctx *c, int K, hipStream_t st, int &ntiles, int &ts, int &kt) {
  int rc;
  if constexpr (N <= 5) {
    if (use_two(c)) {
      if ((rc = run_pairs2<N>(c, K, st, kt, ts))) return rc;
      ntiles = 0;
      c->last_two = 1;
      return NECG_OK;
    }
  }
  if ((rc = run_pairs<N, MODE_SPLIT>(c, K, st, ntiles, ts))) return rc;
  kt = c->last_kt;
  c->last_two = 0;
  return NECG_OK;
}

// Everything the gradient pass of K needs, reserved before any launch (the
// allocation-free entry point checks the whole workspace up front).
template <int N>
int reserve_grad_pass(necg_ctx *c, int K, hipStream_t st) {
  int kt, ts, w, tps;
  if constexpr (N <= 5) {
    if (use_two(c)) return reserve_pairs2<N>(c, K, st, kt, ts, tps);
  }
  return reserve_pairs<N, MODE_SPLIT>(c, K, kt, ts, w, tps);
}

template <int N>
int grad_finish(necg_ctx *c, int K, bool two, int ntiles, int kt, int ts, const double *gbar,
                double *g, hipStream_t st) {
  if constexpr (N <= 5) {
    if (two) return finish_grad2<N>(c, K, kt, ts, gbar, g, st);
  }
  return finish_grad<N, MODE_SPLIT>(c, K, ntiles, gbar, g, st);
}

template <int N>
int do_outprod(necg_ctx *c, int K, const double *x, double *f, double *g, hipStream_t st) {
  int rc, ntiles, ts, kt;
  if ((rc = run_prep<N>(c, K, x, 1, st))) return rc;
  if ((rc = grad_pass<N>(c, K, st, ntiles, ts, kt))) return rc;
  if ((rc = finish_hs(c, K, ts, st))) return rc;
  if ((rc = finish_quotient(c, K, f, st))) return rc;
  return grad_finish<N>(c, K, c->last_two != 0, ntiles, kt, ts, nullptr, g, st);
}

// Energy with the gradient's pair partials kept for a later backward
// (necg_energy_bwd): the gradient pair pass, H/S and the quotient.
template <int N>
int do_energy_fwd(necg_ctx *c, int K, const double *x, double *f, hipStream_t st) {
  int rc, ntiles, ts, kt;
  if ((rc = run_prep<N>(c, K, x, 1, st))) return rc;
  if ((rc = grad_pass<N>(c, K, st, ntiles, ts, kt))) return rc;
  if ((rc = finish_hs(c, K, ts, st))) return rc;
  if ((rc = finish_quotient(c, K, f, st))) return rc;
  c->tape_K = K;
  c->tape_ntiles = ntiles;
  c->tape_two = c->last_two;
  c->tape_kt = kt;
  c->tape_ts = ts;
  return NECG_OK;
}

template <int N>
int do_energy(necg_ctx *c, int K, const double *x, double *f, double *H, double *S,
              hipStream_t st) {
  int rc, ts;
  if ((rc = run_prep<N>(c, K, x, 1, st))) return rc;
  if ((rc = run_tri<N, MODE_ELEMENTS>(c, K, st, ts))) return rc;
  if ((rc = finish_hs(c, K, ts, st))) return rc;
  if ((rc = finish_quotient(c, K, f, st))) return rc;
  if (H) HIP_TRY(hipMemcpyAsync(H, c->H, sizeof(double) * K * K, hipMemcpyDeviceToDevice, st));
  if (S) HIP_TRY(hipMemcpyAsync(S, c->S, sizeof(double) * K * K, hipMemcpyDeviceToDevice, st));
  return NECG_OK;
}

template <int N>
int do_vjp(necg_ctx *c, int K, const double *x, const double *gbar, double *gx,
           hipStream_t st) {
  int rc, ntiles, ts;
  // the bra/ket tables of x are rebuilt: x may have been rewritten in place
  if ((rc = run_prep<N>(c, K, x, 1, st))) return rc;
  if ((rc = run_pairs<N, MODE_COMBINED>(c, K, st, ntiles, ts))) return rc;
  return finish_grad<N, MODE_COMBINED>(c, K, ntiles, gbar, gx, st);
}

template <int N>
int do_matrices(necg_ctx *c, int K, const double *iv, double *H, double *S, hipStream_t st) {
  int rc, ts;
  if ((rc = run_prep<N>(c, K, iv, 0, st))) return rc;
  if ((rc = run_tri<N, MODE_ELEMENTS>(c, K, st, ts))) return rc;
  if ((rc = finish_hs(c, K, ts, st))) return rc;
  if (H) HIP_TRY(hipMemcpyAsync(H, c->H, sizeof(double) * K * K, hipMemcpyDeviceToDevice, st));
  if (S) HIP_TRY(hipMemcpyAsync(S, c->S, sizeof(double) * K * K, hipMemcpyDeviceToDevice, st));
  return NECG_OK;
}

template <int N>
int do_tv(necg_ctx *c, int K, const double *x, double *T, double *V, hipStream_t st) {
  int rc, ts;
  if ((rc = run_prep<N>(c, K, x, 1, st))) return rc;
  if ((rc = run_tri<N, MODE_TV>(c, K, st, ts))) return rc;
  return finish_hs(c, K, ts, st, T, V);
}

template <int N>
int do_query(necg_ctx *c, int K, const double *iv, int nq, const int32_t *q, double *out4,
             double *tgvl, double *tgvls, hipStream_t st) {
  int rc;
  if ((rc = run_prep<N>(c, K, iv, 0, st))) return rc;
  if (nq == 0) return NECG_OK;
  hipLaunchKernelGGL((query_kernel<N>), dim3((nq + 63) / 64), dim3(64), 0, st, c->bra, K,
                     c->nst, c->d_sml, c->d_sms, c->sc, nq, q, out4, tgvl, tgvls, c->d_flag);
  return hip_status(hipGetLastError());
}

template <int N>
int do_dhds(necg_ctx *c, int K, const double *iv, double *DH, double *DS, hipStream_t st) {
  int rc;
  if ((rc = run_prep<N>(c, K, iv, 0, st))) return rc;
  hipLaunchKernelGGL((dhds_kernel<N>), dim3((K + 63) / 64, K), dim3(64), 0, st, c->bra, c->ket,
                     K, c->nst, c->sc, DH, DS, c->d_flag);
  return hip_status(hipGetLastError());
}

#define DISPATCH_N(n, call)                 \
  switch (n) {                              \
    case 1: { constexpr int NN = 1; return call; } \
    case 2: { constexpr int NN = 2; return call; } \
    case 3: { constexpr int NN = 3; return call; } \
    case 4: { constexpr int NN = 4; return call; } \
    case 5: { constexpr int NN = 5; return call; } \
    case 6: { constexpr int NN = 6; return call; } \
    default: return NECG_ERR_UNSUPPORTED;   \
  }

struct DeviceGuard {
  int prev = -1;
  explicit DeviceGuard(int dev) {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    if (prev != dev) (void)hipSetDevice(dev);
  }
  ~DeviceGuard() {
    if (prev >= 0) (void)hipSetDevice(prev);
  }
};

}  // namespace

// ================================================ allocation-free scratch ===

// Device memory of the allocation-free entry point (necg_outprod_noalloc):
// every workspace buffer of an evaluation is carved out of this array, so the
// call makes no device allocation.  Each GPU the library runs on holds its
// own copy until the process ends.  In host code `scratch` is a host-side
// stand-in, not a device address: only the address that hipGetSymbolAddress
// returns for the current device (scratch_base) is handed to kernels and
// copies.
static constexpr size_t kScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
char *g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on `device` (current), looked up on the
// first call per device and kept.
int scratch_base(int device, char **out) {
  if (device < 0 || device >= kMaxDevices) return NECG_ERR_DEVICE;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return NECG_ERR_DEVICE;
    g_scratch_ptr[device] = (char *)p;
  }
  *out = g_scratch_ptr[device];
  return NECG_OK;
}

// Points the context's workspace at the arena for one call and puts its own
// workspace, tile list and resident-state bookkeeping back afterwards: the
// context's buffers and the state a pool or a tape reads are untouched.
struct ArenaScope {
  necg_ctx *c;
  int Kcap, tiles_K, tiles_kt, tiles_fold, ntri, nsplit, last_K, last_kt, last_ts, last_w,
      last_two;
  double *bra, *ket, *H, *S, *HC, *SC, *scal, *part, *hpart, *spart, *hcp, *scp, *red, *kpart;
  size_t part_cap, hpart_cap, kpart_cap, tiles_cap;
  int *d_tiles, *d_xr;
  long long gen, full_gen;
  bool timing;
  ArenaScope(necg_ctx *cc, char *base, size_t cap) : c(cc) {
    Kcap = c->Kcap; tiles_K = c->tiles_K; tiles_kt = c->tiles_kt; tiles_fold = c->tiles_fold;
    ntri = c->ntri;
    nsplit = c->nsplit; last_K = c->last_K; last_kt = c->last_kt; last_ts = c->last_ts;
    last_w = c->last_w; last_two = c->last_two;
    bra = c->bra; ket = c->ket; H = c->H; S = c->S; HC = c->HC; SC = c->SC; scal = c->scal;
    part = c->part; hpart = c->hpart; spart = c->spart; hcp = c->hcp; scp = c->scp;
    red = c->red; kpart = c->kpart;
    part_cap = c->part_cap; hpart_cap = c->hpart_cap; kpart_cap = c->kpart_cap;
    tiles_cap = c->tiles_cap; d_tiles = c->d_tiles; d_xr = c->d_xr;
    gen = c->gen; full_gen = c->full_gen; timing = c->timing;
    c->Kcap = 0; c->tiles_K = 0; c->tiles_kt = 0;
    c->bra = c->ket = c->H = c->S = c->HC = c->SC = c->scal = nullptr;
    c->part = c->hpart = c->spart = c->hcp = c->scp = c->red = c->kpart = nullptr;
    c->part_cap = c->hpart_cap = c->kpart_cap = c->tiles_cap = 0;
    c->d_tiles = c->d_xr = nullptr;
    c->timing = false;
    c->arena = base;
    c->arena_cap = cap;
    c->arena_used = 0;
    c->arena_budget = 0;
  }
  ~ArenaScope() {
    c->Kcap = Kcap; c->tiles_K = tiles_K; c->tiles_kt = tiles_kt; c->tiles_fold = tiles_fold;
    c->ntri = ntri;
    c->nsplit = nsplit; c->last_K = last_K; c->last_kt = last_kt; c->last_ts = last_ts;
    c->last_w = last_w; c->last_two = last_two;
    c->bra = bra; c->ket = ket; c->H = H; c->S = S; c->HC = HC; c->SC = SC; c->scal = scal;
    c->part = part; c->hpart = hpart; c->spart = spart; c->hcp = hcp; c->scp = scp;
    c->red = red; c->kpart = kpart;
    c->part_cap = part_cap; c->hpart_cap = hpart_cap; c->kpart_cap = kpart_cap;
    c->tiles_cap = tiles_cap; c->d_tiles = d_tiles; c->d_xr = d_xr;
    c->gen = gen; c->full_gen = full_gen; c->timing = timing;
    c->arena = nullptr;
    c->arena_cap = c->arena_used = c->arena_budget = 0;
  }
};

}  // namespace

// ================================================================= C ABI ===

extern "C" {

const char *necg_strerror(int code) {
  switch (code) {
    case NECG_OK: return "ok";
    case NECG_ERR_ARG: return "invalid argument";
    case NECG_ERR_UNSUPPORTED: return "unsupported system";
    case NECG_ERR_DEVICE: return "HIP runtime error";
    case NECG_ERR_OOM: return "device out of memory";
    case NECG_ERR_NOT_PD: return "pair matrix not positive definite (zero Cholesky diagonal?)";
    case NECG_ERR_CSC: return "C^T S C <= 0";
    case NECG_ERR_NONFINITE: return "NaN or Inf in input or result";
    case NECG_ERR_STALE: return "no forward tape for this backward (another evaluation ran since)";
    case NECG_ERR_SCRATCH:
      return "workspace larger than the allocation-free entry point's 64 MiB scratch";
    default: return "unknown status";
  }
}

int necg_create(const necg_system *sys, int device, necg_ctx **out) {
  if (!sys || !out) return NECG_ERR_ARG;
  *out = nullptr;
  const int n = sys->n, nst = sys->nst;
  if (n < 1 || n > kMaxN) return NECG_ERR_UNSUPPORTED;
  if (nst < 1 || !sys->mass || !sys->ncv || !sys->scvv || !sys->sml || !sys->sms)
    return NECG_ERR_ARG;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j)
      if (sys->mass[i * n + j] != sys->mass[j * n + i]) return NECG_ERR_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
    return NECG_ERR_DEVICE;
  necg_ctx *c = new (std::nothrow) necg_ctx();
  if (!c) return NECG_ERR_OOM;
  c->device = device;
  c->n = n;
  c->nst = nst;
  c->sc.n = n;
  c->sc.nst = nst;
  c->sc.efs = sys->efs;
  set_mass(c->sc, sys->mass, n);
  for (int i = 0; i <= n; ++i) c->sc.q[i] = sys->ncv[i];
  set_pair_charges(c->sc, sys->ncv, n);
  DeviceGuard g(device);
  int rc = NECG_OK;
  const size_t nsml = (size_t)nst * n * n, nsms = (size_t)nst * 9 * n * n;
  if ((rc = dev_alloc(&c->d_scvv, nst)) || (rc = dev_alloc(&c->d_sml, nsml)) ||
      (rc = dev_alloc(&c->d_sms, nsms))) {
    necg_destroy(c);
    return rc;
  }
  if ((rc = hip_status(hipMemcpy(c->d_scvv, sys->scvv, sizeof(double) * nst,
                                 hipMemcpyHostToDevice))) ||
      (rc = hip_status(hipMemcpy(c->d_sml, sys->sml, sizeof(double) * nsml,
                                 hipMemcpyHostToDevice))) ||
      (rc = hip_status(hipMemcpy(c->d_sms, sys->sms, sizeof(double) * nsms,
                                 hipMemcpyHostToDevice))) ||
      (rc = hip_status(hipMalloc((void **)&c->d_flag, sizeof(int)))) ||
      (rc = hip_status(hipMemset(c->d_flag, 0, sizeof(int))))) {
    necg_destroy(c);
    return rc;
  }
  int ncu = 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess &&
      ncu > 0)
    c->ncu = c->dev_ncu = ncu;
  std::vector<unsigned> tab;
  if (two_sided_table(sys, tab)) {
    if ((rc = hip_status(hipMalloc((void **)&c->d_tperm, tab.size() * sizeof(unsigned)))) ||
        (rc = hip_status(hipMemcpy(c->d_tperm, tab.data(), tab.size() * sizeof(unsigned),
                                   hipMemcpyHostToDevice)))) {
      necg_destroy(c);
      return rc;
    }
    c->two_ok = 1;
  }
  *out = c;
  return NECG_OK;
}

int necg_destroy(necg_ctx *c) {
  if (!c) return NECG_ERR_ARG;
  {
    DeviceGuard g(c->device);
    double *bufs[] = {c->d_scvv, c->d_sml, c->d_sms, c->bra, c->ket, c->H, c->S,
                      c->HC, c->SC, c->scal, c->part, c->hpart, c->spart, c->hcp, c->scp, c->red,
                      c->prop_rec, c->prop_part, c->prop_tab};
    for (double *p : bufs)
      if (p) (void)hipFree(p);
    if (c->kpart) (void)hipFree(c->kpart);
    if (c->d_tiles) (void)hipFree(c->d_tiles);
    if (c->d_tperm) (void)hipFree(c->d_tperm);
    if (c->d_flag) (void)hipFree(c->d_flag);
    for (hipEvent_t e : c->ev) (void)hipEventDestroy(e);
  }
  delete c;
  return NECG_OK;
}

int necg_outprod(necg_ctx *c, int32_t K, const double *x, double *f, double *g, void *stream) {
  if (!c || K < 1 || !x || !g) return NECG_ERR_ARG;
  NECG_REQUIRE_DEVICE((const void *)x, (const void *)f, (const void *)g);
  std::lock_guard<std::mutex> lk(c->mu);
  if (pinned_elsewhere(c, K)) return NECG_ERR_ARG;
  c->tape_K = 0;
  DeviceGuard dg(c->device);
  int rc = ensure_K(c, K);
  if (rc) return rc;
  c->last_K = K;
  auto run = [&]() -> int { DISPATCH_N(c->n, (do_outprod<NN>(c, K, x, f, g, (hipStream_t)stream))); };
  if ((rc = run()) == NECG_OK) c->full_gen = c->gen;
  return rc;
}

int necg_outprod_noalloc(necg_ctx *c, int32_t K, const double *x, double *f, double *g,
                         void *stream) {
  if (!c || K < 1 || !x || !g) return NECG_ERR_ARG;
  NECG_REQUIRE_DEVICE((const void *)x, (const void *)f, (const void *)g);
  std::lock_guard<std::mutex> lk(c->mu);
  DeviceGuard dg(c->device);
  char *base = nullptr;
  int rc = scratch_base(c->device, &base);
  if (rc) return rc;
  c->tape_K = 0;
  ArenaScope arena(c, base, kScratchBytes);
  hipStream_t st = (hipStream_t)stream;
  // the whole workspace is carved before the first launch: a call that does
  // not fit is refused with nothing enqueued except, possibly, the tile list
  if ((rc = ensure_K(c, K))) return rc;
  auto reserve = [&]() -> int { DISPATCH_N(c->n, (reserve_grad_pass<NN>(c, K, st))); };
  if ((rc = reserve())) return rc;
  DISPATCH_N(c->n, (do_outprod<NN>(c, K, x, f, g, st)));
}

int necg_energy(necg_ctx *c, int32_t K, const double *x, double *f, double *H, double *S,
                void *stream) {
  if (!c || K < 1 || !x) return NECG_ERR_ARG;
  NECG_REQUIRE_DEVICE((const void *)x, (const void *)f, (const void *)H, (const void *)S);
  std::lock_guard<std::mutex> lk(c->mu);
  if (pinned_elsewhere(c, K)) return NECG_ERR_ARG;
  c->tape_K = 0;
  DeviceGuard dg(c->device);
  int rc = ensure_K(c, K);
  if (rc) return rc;
  c->last_K = K;
  auto run = [&]() -> int { DISPATCH_N(c->n, (do_energy<NN>(c, K, x, f, H, S, (hipStream_t)stream))); };
  if ((rc = run()) == NECG_OK) c->full_gen = c->gen;
  return rc;
}

int necg_energy_fwd(necg_ctx *c, int32_t K, const double *x, double *f, int64_t *tape,
                    void *stream) {
  if (!c || K < 1 || !x || !tape) return NECG_ERR_ARG;
  NECG_REQUIRE_DEVICE((const void *)x, (const void *)f);
  std::lock_guard<std::mutex> lk(c->mu);
  if (pinned_elsewhere(c, K)) return NECG_ERR_ARG;
  c->tape_K = 0;
  *tape = 0;
  DeviceGuard dg(c->device);
  int rc = ensure_K(c, K);
  if (rc) return rc;
  c->last_K = K;
  auto run = [&]() -> int { DISPATCH_N(c->n, (do_energy_fwd<NN>(c, K, x, f, (hipStream_t)stream))); };
  if ((rc = run()) == NECG_OK) {
    *tape = c->tape_id = ++c->tape_seq;
    c->full_gen = c->gen;
  }
  return rc;
}

int necg_energy_bwd(necg_ctx *c, int32_t K, int64_t tape, const double *gbar, double *gx,
                    void *stream) {
  if (!c || K < 1 || !gx) return NECG_ERR_ARG;
  NECG_REQUIRE_DEVICE((const void *)gbar, (const void *)gx);
  std::lock_guard<std::mutex> lk(c->mu);
  if (c->tape_K != K || tape == 0 || tape != c->tape_id) return NECG_ERR_STALE;
  DeviceGuard dg(c->device);
  DISPATCH_N(c->n, (grad_finish<NN>(c, K, c->tape_two != 0, c->tape_ntiles, c->tape_kt,
                                    c->tape_ts, gbar, gx, (hipStream_t)stream)));
}

int necg_energy_vjp(necg_ctx *c, int32_t K, const double *x, const double *gbar, double *gx,
                    void *stream) {
  if (!c || K < 1 || !x || !gx) return NECG_ERR_ARG;
  NECG_REQUIRE_DEVICE((const void *)x, (const void *)gbar, (const void *)gx);
  std::lock_guard<std::mutex> lk(c->mu);
  if (pinned_elsewhere(c, K)) return NECG_ERR_ARG;
  c->tape_K = 0;
  if (c->last_K != K || K > c->Kcap) return NECG_ERR_ARG;  // needs necg_energy first
  DeviceGuard dg(c->device);
  DISPATCH_N(c->n, (do_vjp<NN>(c, K, x, gbar, gx, (hipStream_t)stream)));
}

int necg_matrices(necg_ctx *c, int32_t K, const double *iv, double *H, double *S, void *stream) {
  if (!c || K < 1 || !iv) return NECG_ERR_ARG;
  NECG_REQUIRE_DEVICE((const void *)iv, (const void *)H, (const void *)S);
  std::lock_guard<std::mutex> lk(c->mu);
  if (pinned_elsewhere(c, K)) return NECG_ERR_ARG;
  c->tape_K = 0;
  DeviceGuard dg(c->device);
  int rc = ensure_K(c, K);
  if (rc) return rc;
  DISPATCH_N(c->n, (do_matrices<NN>(c, K, iv, H, S, (hipStream_t)stream)));
}

int necg_kinetic_potential(necg_ctx *c, int32_t K, const double *x, double *T, double *V,
                           void *stream) {
  if (!c || K < 1 || !x || !T || !V) return NECG_ERR_ARG;
  NECG_REQUIRE_DEVICE((const void *)x, (const void *)T, (const void *)V);
  std::lock_guard<std::mutex> lk(c->mu);
  if (pinned_elsewhere(c, K)) return NECG_ERR_ARG;
  c->tape_K = 0;
  DeviceGuard dg(c->device);
  int rc = ensure_K(c, K);
  if (rc) return rc;
  DISPATCH_N(c->n, (do_tv<NN>(c, K, x, T, V, (hipStream_t)stream)));
}

int necg_matel(necg_ctx *c, int32_t K, const double *iv, int32_t nq, const int32_t *q,
               double *out, void *stream) {
  if (!c || K < 1 || !iv || nq < 0 || (nq > 0 && (!q || !out))) return NECG_ERR_ARG;
  NECG_REQUIRE_DEVICE((const void *)iv, (const void *)q, (const void *)out);
  std::lock_guard<std::mutex> lk(c->mu);
  c->tape_K = 0;
  DeviceGuard dg(c->device);
  int rc = ensure_K(c, K);
  if (rc) return rc;
  DISPATCH_N(c->n, (do_query<NN>(c, K, iv, nq, q, out, nullptr, nullptr, (hipStream_t)stream)));
}

int necg_gradel(necg_ctx *c, int32_t K, const double *iv, int32_t nq, const int32_t *q,
                double *tgvl, double *tgvls, void *stream) {
  if (!c || K < 1 || !iv || nq < 0 || (nq > 0 && (!q || !tgvl || !tgvls))) return NECG_ERR_ARG;
  NECG_REQUIRE_DEVICE((const void *)iv, (const void *)q, (const void *)tgvl,
                      (const void *)tgvls);
  std::lock_guard<std::mutex> lk(c->mu);
  c->tape_K = 0;
  DeviceGuard dg(c->device);
  int rc = ensure_K(c, K);
  if (rc) return rc;
  DISPATCH_N(c->n,
             (do_query<NN>(c, K, iv, nq, q, nullptr, tgvl, tgvls, (hipStream_t)stream)));
}

int necg_dhds(necg_ctx *c, int32_t K, const double *iv, double *DH, double *DS, void *stream) {
  if (!c || K < 1 || !iv || !DH || !DS) return NECG_ERR_ARG;
  NECG_REQUIRE_DEVICE((const void *)iv, (const void *)DH, (const void *)DS);
  std::lock_guard<std::mutex> lk(c->mu);
  c->tape_K = 0;
  DeviceGuard dg(c->device);
  int rc = ensure_K(c, K);
  if (rc) return rc;
  DISPATCH_N(c->n, (do_dhds<NN>(c, K, iv, DH, DS, (hipStream_t)stream)));
}

int necg_device_visible(const void *p) { return p && necg::device_visible({p}) ? 1 : 0; }

int necg_status(necg_ctx *c, void *stream) {
  if (!c) return NECG_ERR_ARG;
  std::lock_guard<std::mutex> lk(c->mu);
  DeviceGuard dg(c->device);
  HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
  int flags = 0;
  HIP_TRY(hipMemcpy(&flags, c->d_flag, sizeof(int), hipMemcpyDeviceToHost));
  if (flags) HIP_TRY(hipMemset(c->d_flag, 0, sizeof(int)));
  if (flags & F_BAD_INPUT) return NECG_ERR_NONFINITE;
  if (flags & F_NOT_PD) return NECG_ERR_NOT_PD;
  if (flags & F_CSC) return NECG_ERR_CSC;
  if (flags & F_NONFINITE) return NECG_ERR_NONFINITE;
  return NECG_OK;
}

int necg_set_tiling(necg_ctx *c, int32_t ket_tile, int32_t term_split, int32_t waves) {
  if (!c || ket_tile < 0 || term_split < 0) return NECG_ERR_ARG;
  if (waves != 0 && waves != pair_waves(c->n)) return NECG_ERR_ARG;  // fixed per n
  std::lock_guard<std::mutex> lk(c->mu);
  if (c->pins) return NECG_ERR_ARG;
  c->req_kt = ket_tile;
  c->req_ts = term_split;
  c->req_w = waves;
  return NECG_OK;
}

int necg_timing(necg_ctx *c, int32_t mode, double *ms, int64_t *launches) {
  if (!c || mode < 0 || mode > 2) return NECG_ERR_ARG;
  std::lock_guard<std::mutex> lk(c->mu);
  DeviceGuard dg(c->device);
  if (mode == 1) {
    c->timing = true;
    c->nev = 0;
    return NECG_OK;
  }
  if (mode == 0) {
    c->timing = false;
    return NECG_OK;
  }
  double tot = 0.0;
  for (int i = 0; i < c->nev; ++i) {
    HIP_TRY(hipEventSynchronize(c->ev[2 * i + 1]));
    float t = 0.f;
    HIP_TRY(hipEventElapsedTime(&t, c->ev[2 * i], c->ev[2 * i + 1]));
    tot += t;
  }
  if (ms) *ms = tot;
  if (launches) *launches = c->nev;
  return NECG_OK;
}

int necg_set_pair_mode(necg_ctx *c, int32_t mode) {
  if (!c || mode < 0 || mode > 2) return NECG_ERR_ARG;
  std::lock_guard<std::mutex> lk(c->mu);
  if (c->pins) return NECG_ERR_ARG;
  if (mode == 1 && !c->two_ok) return NECG_ERR_UNSUPPORTED;
  c->pair_mode = mode;
  c->tape_K = 0;
  return NECG_OK;
}

int necg_pair_form(necg_ctx *c, int32_t *available, int32_t *last) {
  if (!c || !available || !last) return NECG_ERR_ARG;
  std::lock_guard<std::mutex> lk(c->mu);
  *available = c->two_ok;
  *last = c->last_two;
  return NECG_OK;
}

int necg_set_launch_cus(necg_ctx *c, int32_t ncu) {
  if (!c || ncu < 0) return NECG_ERR_ARG;
  std::lock_guard<std::mutex> lk(c->mu);
  if (c->pins) return NECG_ERR_ARG;
  c->ncu = ncu > 0 ? ncu : c->dev_ncu;
  c->tiles_K = 0;   // the tile list depends on it
  return NECG_OK;
}

int necg_set_fold(necg_ctx *c, int32_t fold) {
  if (!c || fold < 0 || fold > 1) return NECG_ERR_ARG;
  std::lock_guard<std::mutex> lk(c->mu);
  if (c->pins) return NECG_ERR_ARG;
  c->fold_mode = fold;
  c->tape_K = 0;
  return NECG_OK;
}

int necg_pin_workspace(necg_ctx *c, int32_t pin) {
  if (!c) return NECG_ERR_ARG;
  std::lock_guard<std::mutex> lk(c->mu);
  if (pin) {
    c->pins++;
  } else {
    if (c->pins == 0) return NECG_ERR_ARG;
    c->pins--;
  }
  return NECG_OK;
}

int necg_get_tiling(necg_ctx *c, int32_t *ket_tile, int32_t *term_split, int32_t *waves) {
  if (!c || !ket_tile || !term_split || !waves) return NECG_ERR_ARG;
  std::lock_guard<std::mutex> lk(c->mu);
  *ket_tile = c->last_kt;
  *term_split = c->last_ts;
  *waves = c->last_w;
  return NECG_OK;
}

}  // extern "C"
