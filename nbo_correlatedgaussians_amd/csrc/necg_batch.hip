// necg_batch.hip -- many independent systems (uneven n, nst and K) in one
// evaluation: segmented preparation and finishing kernels over all systems,
// and one persistent pair-kernel launch per particle count n whose
// workgroups walk a cost-sorted list of (system, bra tile, ket tile, term
// group) items.  The whole evaluation is a fixed sequence of launches with no
// host synchronisation and no memset, so a caller can capture it in a
// hipGraph and replay it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <mutex>
#include <new>
#include <numeric>
#include <vector>

#include "../../include/necg.h"
#include "necg_abi.hpp"
#include "necg_finish.hpp"
#include "necg_pair2.hpp"
#include "necg_pool_impl.hpp"
#include "necg_two.hpp"

using namespace necg;

namespace {

// Per-system record, resident on the device.
struct BatchSys {
  SysConst sc;
  int n, nst, K, ts, tps, nchunk, ntiles, nacc;
  long long x_off;                   // x / g (doubles)
  long long bra_off, ket_off;        // bra / ket tables
  long long hs_off;                  // H, S: K*K
  long long hp_off;                  // element partials: ts*K*K
  long long part_off;                // gradient partials: ntiles*nacc*K
  long long red_off;                 // summed accumulators: nacc*K
  long long vec_off;                 // HC, SC: K
  long long hcp_off;                 // chunk partials: nchunk*K
  long long sml_off, sms_off, scvv_off;
  // two-sided pair form (lower-triangle items): enabled, ket tile, folded
  // diagonal blocks (tri_fold), the system's target-slot table and its ket
  // partials ((K+63)/64 * ts * K * nacc)
  int two, kt, fold;
  long long tperm_off, kpart_off;
};

struct Item {  // one pair tile (fold: sweeps [l0, l1) of a folded diagonal block)
  int sys, k0, l0, l1, p0, p1, tile, fold;
};

// Segment s with prefix[s] <= idx < prefix[s+1] (prefix has nseg+1 entries).
__device__ __forceinline__ int seg_find(const long long *__restrict__ prefix, int nseg,
                                        long long idx) {
  int lo = 0, hi = nseg - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (prefix[mid] <= idx) lo = mid; else hi = mid - 1;
  }
  return lo;
}

#define SWITCH_N(n, call)                         \
  switch (n) {                                    \
    case 1: { constexpr int NN = 1; call; break; } \
    case 2: { constexpr int NN = 2; call; break; } \
    case 3: { constexpr int NN = 3; call; break; } \
    case 4: { constexpr int NN = 4; call; break; } \
    case 5: { constexpr int NN = 5; call; break; } \
    default: { constexpr int NN = 6; call; break; } \
  }

// one lane per (system, function)
// active (nullable): per-system flags; systems with active[s] == 0 are
// skipped by every kernel of the evaluation and keep their previous state.
__global__ void b_prep_bra(const BatchSys *__restrict__ sys, int nsys,
                           const long long *__restrict__ fpre, const int32_t *__restrict__ active,
                           const double *__restrict__ x, double *__restrict__ bra,
                           int *__restrict__ flag) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= fpre[nsys]) return;
  const int s = seg_find(fpre, nsys, idx);
  if (active && !active[s]) return;
  const BatchSys &b = sys[s];
  const int k = (int)(idx - fpre[s]);
  const double *xs = x + b.x_off;
  SWITCH_N(b.n, (make_bra<NN>(xs + (size_t)k * Dim<NN>::BLK, xs[(size_t)b.K * Dim<NN>::BLK + k],
                              true, b.sc, bra + b.bra_off + (size_t)k * Dim<NN>::BS, flag)));
}

// one 64-lane workgroup per (system, 64 (ket, term) records): the records
// built in LDS and written out as one contiguous run
__global__ void __launch_bounds__(64) b_prep_ket(const BatchSys *__restrict__ sys, int nsys,
                                                 const long long *__restrict__ kwpre,
                                                 const int32_t *__restrict__ active,
                                                 const double *__restrict__ sml,
                                                 const double *__restrict__ sms,
                                                 const double *__restrict__ scvv,
                                                 const double *__restrict__ bra,
                                                 double *__restrict__ ket) {
  __shared__ double rec[64 * (Dim<kMaxN>::KS + 1)];
  const long long blk = blockIdx.x;
  const int s = seg_find(kwpre, nsys, blk);
  if (active && !active[s]) return;  // uniform over the workgroup
  const BatchSys &b = sys[s];
  const int idx0 = (int)(blk - kwpre[s]) * 64;
  SWITCH_N(b.n, (ket_records_block<NN>(bra + b.bra_off, b.K, b.nst, sml + b.sml_off,
                                       sms + b.sms_off, scvv + b.scvv_off, b.sc, idx0,
                                       ket + b.ket_off, rec)));
}

// Persistent pair kernel for one particle count: each workgroup walks the
// item list with stride gridDim.x (items sorted by decreasing cost).
template <int N>
__global__ void __launch_bounds__(64 * pair_waves(N), 1) b_pair(
    const BatchSys *__restrict__ sys, const Item *__restrict__ items, int nitems,
    const int32_t *__restrict__ active, const double *__restrict__ bra,
    const double *__restrict__ ket, double *__restrict__ part, double *__restrict__ hpart,
    double *__restrict__ spart, int *__restrict__ flag) {
  using D = Dim<N>;
  __shared__ double acc_s[PairLds<N, MODE_SPLIT>::ACC];
  __shared__ double kst_s[pair_waves(N) * D::KS];
  for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
    const Item w = items[it];
    if (active && !active[w.sys]) continue;  // uniform over the workgroup
    const BatchSys &b = sys[w.sys];
    const int K = b.K;
    const int z = w.tile % b.ts;
    pair_tile<N, MODE_SPLIT, false, true>(
        bra + b.bra_off, ket + b.ket_off, K, b.nst, w.k0, w.l0, w.l1, w.p0, w.p1, b.sc, 0.0,
        part + b.part_off + (size_t)w.tile * D::NACC * K, hpart + b.hp_off + (size_t)z * K * K,
        spart + b.hp_off + (size_t)z * K * K, flag, acc_s, 0x7fffffff, kst_s);
    __syncthreads();  // the next item overwrites the LDS tables
  }
}

// Two-sided persistent pair kernel: the lower-triangle items (tri_tiles) of
// every system whose terms allow the two-sided form, all particle counts
// n <= 5 in one launch (items sorted by decreasing cost, so the small
// systems' items fill the tail).  LDS is sized for n = 5.  Workgroup g takes
// item g first; with a work counter (ctr, zero at launch) each workgroup then
// takes the next unclaimed item as it finishes one (the longest items first,
// the tail filled by whichever workgroup is free); without one it strides by
// the grid.  Every item writes its own partial slots, so the results do not
// depend on which workgroup took it.
__global__ void __launch_bounds__(256, 1) b_pair2_all(
    const BatchSys *__restrict__ sys, const Item *__restrict__ items, int nitems,
    const int32_t *__restrict__ active, const double *__restrict__ bra,
    const double *__restrict__ ket, const unsigned *__restrict__ tperm,
    double *__restrict__ part, double *__restrict__ kpart, double *__restrict__ hpart,
    double *__restrict__ spart, int *__restrict__ flag, int *__restrict__ ctr) {
  static_assert(pair_waves(5) == 4 && pair_waves(1) == 4, "one workgroup shape for n <= 5");
  __shared__ double acc_s[PairLds<5, MODE_SPLIT>::ACC];
  __shared__ double kacc_s[4 * 64];
  __shared__ double bra_s[BraLds<5>::NF * 64];
  __shared__ double kst_s[4 * 64];
  for (int it = blockIdx.x; it < nitems;) {
    const Item w = items[it];
    if (active && !active[w.sys]) {  // uniform over the workgroup
      it += gridDim.x;
      continue;
    }
    const BatchSys &b = sys[w.sys];
    const int K = b.K;
    const int z = w.tile % b.ts, bx = w.k0 >> 6;
    const int up = tri_up(K, b.kt), lup = w.k0 < up ? up : K;
    const double *br = bra + b.bra_off, *kt = ket + b.ket_off;
    const unsigned *tp = tperm + b.tperm_off;
    double *hp = hpart + b.hp_off + (size_t)z * K * K, *sp = spart + b.hp_off + (size_t)z * K * K;
    double *bp = part + b.part_off + (size_t)w.tile * b.nacc * K;
    double *kp = kpart + b.kpart_off + (size_t)(bx * b.ts + z) * K * b.nacc;
    switch (b.n) {
#define NECG_B2(NN)                                                                          \
  case NN:                                                                                   \
    if (w.fold == 2) {                                                                       \
      if (b.sc.mrank1)                                                                       \
        pair2_diag_tile<NN, true>(br, kt, K, b.nst, w.k0, w.p0, w.p1, b.sc, tp, bp, hp, sp,  \
                                  flag, acc_s, kacc_s, bra_s, kst_s);                        \
      else                                                                                   \
        pair2_diag_tile<NN, false>(br, kt, K, b.nst, w.k0, w.p0, w.p1, b.sc, tp, bp, hp, sp, \
                                   flag, acc_s, kacc_s, bra_s, kst_s);                       \
    } else if (w.fold) {                                                                     \
      if (b.sc.mrank1)                                                                       \
        pair2_fold_tile<NN, true>(br, kt, K, b.nst, w.k0, w.l0, w.l1, w.p0, w.p1, b.sc, tp,  \
                                  bp, kp, hp, sp, flag, acc_s, kacc_s, bra_s, kst_s);        \
      else                                                                                   \
        pair2_fold_tile<NN, false>(br, kt, K, b.nst, w.k0, w.l0, w.l1, w.p0, w.p1, b.sc, tp, \
                                   bp, kp, hp, sp, flag, acc_s, kacc_s, bra_s, kst_s);       \
    } else if (b.sc.mrank1) {                                                                \
      pair2_tile<NN, true>(br, kt, K, b.nst, w.k0, w.l0, w.l1, lup, w.p0, w.p1, b.sc, tp, bp, \
                           kp, hp, sp, flag, acc_s, kacc_s, bra_s, kst_s);                   \
    } else {                                                                                 \
      pair2_tile<NN, false>(br, kt, K, b.nst, w.k0, w.l0, w.l1, lup, w.p0, w.p1, b.sc, tp,   \
                            bp, kp, hp, sp, flag, acc_s, kacc_s, bra_s, kst_s);              \
    }                                                                                        \
    break;
      NECG_B2(1)
      NECG_B2(2)
      NECG_B2(3)
      NECG_B2(4)
      NECG_B2(5)
#undef NECG_B2
      default: break;
    }
    __syncthreads();  // the next item overwrites the LDS tables
    if (ctr) {
      // the claim passes through kacc_s[0], which the next tile clears only
      // after the second barrier
      if (threadIdx.x == 0)
        kacc_s[0] = __builtin_bit_cast(double, (long long)(atomicAdd(ctr, 1) + (int)gridDim.x));
      __syncthreads();
      it = (int)__builtin_bit_cast(long long, kacc_s[0]);
      __syncthreads();
    } else {
      it += gridDim.x;
    }
  }
}

// Clears the work counter of b_pair2_all (a kernel, not a memset: the batch
// evaluation is captured in HIP graphs).
__global__ void b_clear_counter(int *__restrict__ ctr) { *ctr = 0; }

// one 256-lane workgroup per (system, pair of 64 x 64 blocks I >= J)
__global__ void __launch_bounds__(256) b_hs_finish(const BatchSys *__restrict__ sys, int nsys,
                                                   const long long *__restrict__ hpre,
                                                   const int32_t *__restrict__ active,
                                                   const double *__restrict__ hpart,
                                                   const double *__restrict__ spart,
                                                   const double *__restrict__ bra,
                                                   double *__restrict__ H, double *__restrict__ S,
                                                   double *__restrict__ hcp,
                                                   double *__restrict__ scp) {
  __shared__ double bh[64 * 65], bsl[64 * 65], cc[128];
  const long long blk = blockIdx.x;
  const int s = seg_find(hpre, nsys, blk);
  if (active && !active[s]) return;  // uniform over the workgroup
  const BatchSys &b = sys[s];
  const int t = (int)(blk - hpre[s]);
  int I = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
  while ((I + 1) * (I + 2) / 2 <= t) ++I;
  while (I * (I + 1) / 2 > t) --I;
  const int J = t - I * (I + 1) / 2;
  int bs = 0, bc = 0;
  SWITCH_N(b.n, (bs = Dim<NN>::BS, bc = Dim<NN>::B_C));
  hs_tile(hpart + b.hp_off, spart + b.hp_off, b.K, b.ts, bra + b.bra_off, bs, bc, I, J,
          H + b.hs_off, S + b.hs_off, hcp + b.hcp_off, scp + b.hcp_off, bh, bsl, cc);
}

// one 256-lane workgroup per (system, 64 rows): H C and S C from the chunk
// partials
__global__ void __launch_bounds__(256) b_hc_sum(const BatchSys *__restrict__ sys, int nsys,
                                                const long long *__restrict__ fwpre,
                                                const int32_t *__restrict__ active,
                                                const double *__restrict__ hcp,
                                                const double *__restrict__ scp,
                                                double *__restrict__ HC, double *__restrict__ SC) {
  __shared__ double ph[256], ps[256];
  const long long blk = blockIdx.x;
  const int s = seg_find(fwpre, nsys, blk);
  if (active && !active[s]) return;  // uniform over the workgroup
  const BatchSys &b = sys[s];
  hc_rows(hcp + b.hcp_off, scp + b.hcp_off, b.nchunk, b.K, (int)(blk - fwpre[s]) * 64,
          HC + b.vec_off, SC + b.vec_off, ph, ps);
}

// one workgroup of 1024 lanes per system
__global__ void __launch_bounds__(1024) b_quotient(const BatchSys *__restrict__ sys,
                                                   const int32_t *__restrict__ active,
                                                   const double *__restrict__ bra,
                                                   const double *__restrict__ HC,
                                                   const double *__restrict__ SC,
                                                   double *__restrict__ scal,
                                                   double *__restrict__ f, int *__restrict__ flag) {
  __shared__ double rh[1024], rs[1024];
  const int s = blockIdx.x;
  if (active && !active[s]) return;
  const BatchSys &b = sys[s];
  int bs = 0, bc = 0;
  SWITCH_N(b.n, (bs = Dim<NN>::BS, bc = Dim<NN>::B_C));
  // over the summed H C and S C (b_hc_sum): one chunk, nothing stored back
  quotient_block(HC + b.vec_off, SC + b.vec_off, 1, bra + b.bra_off, bs, bc, b.K, nullptr,
                 nullptr, scal + 8 * (size_t)s, f ? f + s : nullptr, flag, rh, rs);
}

// one 256-lane workgroup per (system, slot, 64 functions): the two-sided
// reduction over four waves, or (per-ordering systems) one lane per function
__global__ void __launch_bounds__(256) b_tile_reduce(const BatchSys *__restrict__ sys, int nsys,
                                                     const long long *__restrict__ r2pre,
                                                     const int32_t *__restrict__ active,
                                                     const double *__restrict__ part,
                                                     const double *__restrict__ kpart,
                                                     double *__restrict__ red) {
  __shared__ double sh[256];
  const long long blk = blockIdx.x;
  const int s = seg_find(r2pre, nsys, blk);
  if (active && !active[s]) return;  // uniform over the workgroup
  const BatchSys &b = sys[s];
  const int r = (int)(blk - r2pre[s]);
  const int nch = (b.K + 63) / 64;
  const int v = r / nch, k0 = (r - v * nch) * 64;
  if (b.two) {
    tile_reduce2_block(part + b.part_off, kpart + b.kpart_off, b.K, b.kt, b.ts, b.nacc, v, k0,
                       red + b.red_off, sh, nullptr, nullptr, b.fold);
  } else if (threadIdx.x < 64 && k0 + (int)threadIdx.x < b.K) {
    tile_reduce_one(part + b.part_off, b.ntiles, b.K, b.nacc, v, k0 + threadIdx.x,
                    red + b.red_off);
  }
}

// one 64-lane workgroup per (system, 64 functions)
__global__ void __launch_bounds__(64) b_grad_final(const BatchSys *__restrict__ sys, int nsys,
                                                   const long long *__restrict__ fwpre,
                                                   const int32_t *__restrict__ active,
                                                   const double *__restrict__ red,
                                                   const double *__restrict__ bra,
                                                   const double *__restrict__ scal,
                                                   const double *__restrict__ HC,
                                                   const double *__restrict__ SC,
                                                   double *__restrict__ g) {
  __shared__ double go[64 * (Dim<kMaxN>::BLK + 1)];
  const long long blk = blockIdx.x;
  const int s = seg_find(fwpre, nsys, blk);
  if (active && !active[s]) return;  // uniform over the workgroup
  const BatchSys &b = sys[s];
  const int k0 = (int)(blk - fwpre[s]) * 64;
  SWITCH_N(b.n, (grad_block<NN, MODE_SPLIT>(red + b.red_off, b.K, k0, bra + b.bra_off,
                                            scal + 8 * (size_t)s, HC + b.vec_off,
                                            SC + b.vec_off, 1.0, g + b.x_off, go)));
}

template <int N>
void launch_pair(int grid, hipStream_t st, const BatchSys *sys, const Item *items, int nitems,
                 const int32_t *active, const double *bra, const double *ket, double *part,
                 double *hpart, double *spart, int *flag) {
  hipLaunchKernelGGL((b_pair<N>), dim3(grid), dim3(64 * pair_waves(N)), 0, st, sys, items,
                     nitems, active, bra, ket, part, hpart, spart, flag);
}

void launch_pair2(int grid, hipStream_t st, const BatchSys *sys, const Item *items, int nitems,
                  const int32_t *active, const double *bra, const double *ket,
                  const unsigned *tperm, double *part, double *kpart, double *hpart,
                  double *spart, int *flag, int *ctr) {
  if (ctr) hipLaunchKernelGGL(b_clear_counter, dim3(1), dim3(1), 0, st, ctr);
  hipLaunchKernelGGL(b_pair2_all, dim3(grid), dim3(256), 0, st, sys, items, nitems, active, bra,
                     ket, tperm, part, kpart, hpart, spart, flag, ctr);
}

int hip_rc(hipError_t e) {
  if (e == hipSuccess) return NECG_OK;
  if (e == hipErrorOutOfMemory) return NECG_ERR_OOM;
  return NECG_ERR_DEVICE;
}

template <class T>
int dalloc(T **p, size_t count) {
  if (*p) {
    (void)hipFree(*p);
    *p = nullptr;
  }
  if (count == 0) return NECG_OK;
  return hip_rc(hipMalloc((void **)p, count * sizeof(T)));
}

// relative cost of one unordered (pair, term) of the two-sided form: it takes
// about 1.15 times one ordered (pair, term) of the per-ordering form
// (LiH K = 2048: 7.5 ms for K^2/2 against 13.0 ms for K^2)
double term_cost2(int n);

// relative cost of one ordered (pair, term): F_E + F_G of the FLOP model
double term_cost(int n) {
  switch (n) {
    case 1: return 200.0;
    case 2: return 1321.0;
    case 3: return 3005.0;
    case 4: return 6000.0;
    case 5: return 9672.0;
    default: return 15000.0;
  }
}

double term_cost2(int n) { return 1.15 * term_cost(n); }

int dim_bs(int n) { int r = 0; SWITCH_N(n, r = Dim<NN>::BS); return r; }
int dim_ks(int n) { int r = 0; SWITCH_N(n, r = Dim<NN>::KS); return r; }
int dim_blk(int n) { int r = 0; SWITCH_N(n, r = Dim<NN>::BLK); return r; }
int dim_nacc(int n) { int r = 0; SWITCH_N(n, r = Dim<NN>::NACC); return r; }

}  // namespace

struct necg_batch {
  int device = 0, nsys = 0, ncu = 0;
  std::vector<BatchSys> hsys;
  // host copies of the prefixes: functions, 64-function groups, 64-record
  // groups, (64 x 64 block pair) workgroups, (slot, 64 functions) workgroups
  std::vector<long long> fpre, fwpre, kwpre, hpre, r2pre;
  BatchSys *d_sys = nullptr;
  long long *d_fpre = nullptr, *d_fwpre = nullptr, *d_kwpre = nullptr, *d_hpre = nullptr,
            *d_r2pre = nullptr;
  double *sml = nullptr, *sms = nullptr, *scvv = nullptr;
  double *bra = nullptr, *ket = nullptr, *H = nullptr, *S = nullptr, *hpart = nullptr,
         *spart = nullptr, *part = nullptr, *red = nullptr, *HC = nullptr, *SC = nullptr,
         *hcp = nullptr, *scp = nullptr, *scal = nullptr, *kpart = nullptr;
  unsigned *tperm = nullptr;  // target-slot tables of the two-sided systems
  int *flag = nullptr;   // [0]: latched status; [1]: the pair kernel's work counter
  // pair work lists per (n, form)
  struct Cls {
    int n = 0;
    bool two = false;
    Item *d_items = nullptr;
    int nitems = 0;
    int grid = 0;
  };
  std::vector<Cls> cls;
  double item_flops = 0.0;  // target cost of one pair tile (model FLOP)
  int grid_per_cu = 0;      // persistent workgroups per CU (0: occupancy)
  int launch_cus = 0;       // CUs the persistent grids are sized for (0: the device's)
  long long nx = 0;
  int pins = 0;  // necg_batch_pin: captured graphs hold the work lists
  std::mutex mu;
  std::vector<int> in_n, in_nst, in_K;
  std::vector<std::vector<double>> in_sml, in_sms, in_scvv;
};

namespace {

// Tiles every system into items of about `target` model FLOP and builds the
// per-n work lists, heaviest items first.
int plan(necg_batch *B) {
  int rc;
  const double target = B->item_flops > 0 ? B->item_flops : 6.0e7;
  // lists[n]: the per-ordering items of the systems with n particles;
  // lists[0]: the two-sided items of all systems that take that form
  std::vector<std::vector<Item>> lists(kMaxN + 1);
  std::vector<std::vector<double>> costs(kMaxN + 1);
  long long part_tot = 0, hp_tot = 0, kp_tot = 0;
  const int w = 4;  // waves per pair workgroup for n <= 5 (pair_waves)
  for (int s = 0; s < B->nsys; ++s) {
    BatchSys &b = B->hsys[s];
    const int K = b.K, nst = b.nst;
    const int nb = std::min(K, 64);
    const double tc = b.two ? term_cost2(b.n) : term_cost(b.n);
    const double per_ket = nb * nst * tc;
    int kt = (int)std::max(1.0, std::floor(target / per_ket));
    // two-sided: a divisor of 32 (for the upper pairs and the folded
    // diagonal blocks), at least one ket per wave
    if (b.two) {
      int q = w;
      while (q * 2 <= std::min(kt, 32)) q *= 2;
      kt = q;
    }
    kt = std::min(kt, K);
    int tps = nst;
    if (kt == 1) {  // a single ket is already above target: split the terms
      tps = (int)std::max(1.0, std::ceil(target / (nb * tc)));
      tps = std::min(tps, nst);
    }
    const int ts = (nst + tps - 1) / tps;
    const int kts = (K + kt - 1) / kt;
    b.kt = kt;
    b.fold = (b.two && tri_fold(kt, 1)) ? 1 : 0;
    b.ts = ts;
    b.tps = tps;
    b.ntiles = kts * ts;
    b.part_off = part_tot;
    part_tot += (long long)b.ntiles * b.nacc * K;
    b.hp_off = hp_tot;
    hp_tot += (long long)ts * K * K;
    b.kpart_off = kp_tot;
    if (b.two) kp_tot += (long long)((K + 63) / 64) * ts * K * b.nacc;
    const int li = b.two ? 0 : b.n;
    auto add = [&](int bt, int y, int f) {
      for (int z = 0; z < ts; ++z) {
        Item it;
        it.sys = s;
        it.k0 = bt * 64;
        it.fold = f;
        if (f) {   // sweeps of the folded diagonal block (two kets each), or its diagonal tile
          it.l0 = (y - it.k0 / kt) * kt;
          it.l1 = it.l0 + kt;
          if (it.l0 == 32) it.fold = 2;
        } else {
          it.l0 = y * kt;
          it.l1 = std::min(K, it.l0 + kt);
          if (b.two) {
            const int up = tri_up(K, kt);
            if (it.l0 < (it.k0 < up ? up : K)) it.l1 = std::min(it.l1, it.k0 + 64);
          }
        }
        it.p0 = z * tps;
        it.p1 = std::min(nst, it.p0 + tps);
        it.tile = y * ts + z;
        // fold sweeps at 1.12 ket sweeps; a diagonal tile first of all (it is
        // short but long-latency)
        const double c = it.fold == 2 ? 1e30
                                      : (double)std::min(64, K - it.k0) * (it.l1 - it.l0) *
                                            (f ? 1.12 : 1.0) * (it.p1 - it.p0) * tc;
        lists[li].push_back(it);
        costs[li].push_back(c);
      }
    };
    if (b.two) {
      tri_tiles(K, kt, b.fold, add);
    } else {
      for (int bt = 0; bt < (K + 63) / 64; ++bt)
        for (int y = 0; y < kts; ++y) add(bt, y, 0);
    }
  }

  if ((rc = dalloc(&B->part, (size_t)part_tot))) return rc;
  if ((rc = dalloc(&B->hpart, (size_t)hp_tot))) return rc;
  if ((rc = dalloc(&B->spart, (size_t)hp_tot))) return rc;
  if ((rc = dalloc(&B->kpart, (size_t)kp_tot))) return rc;
  for (auto &c : B->cls)
    if (c.d_items) (void)hipFree(c.d_items);
  B->cls.clear();
  for (int li = 0; li <= kMaxN; ++li) {
    if (lists[li].empty()) continue;
    const int n = li;
    const bool two = li == 0;
    std::vector<size_t> ord(lists[li].size());
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(),
                     [&](size_t a, size_t b) { return costs[li][a] > costs[li][b]; });
    std::vector<Item> sorted;
    for (size_t i : ord) sorted.push_back(lists[li][i]);
    necg_batch::Cls c;
    c.n = n;
    c.two = two;
    c.nitems = (int)sorted.size();
    if ((rc = dalloc(&c.d_items, sorted.size()))) return rc;
    if ((rc = hip_rc(hipMemcpy(c.d_items, sorted.data(), sorted.size() * sizeof(Item),
                               hipMemcpyHostToDevice))))
      return rc;
    int per_cu = B->grid_per_cu;
    if (per_cu <= 0) {
      int occ = 1;
      const void *fn = nullptr;
      if (two) {
        fn = (const void *)&b_pair2_all;
      } else {
        SWITCH_N(n, fn = (const void *)&b_pair<NN>);
      }
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, two ? 256 : 64 * pair_waves(n),
                                                       0) !=
              hipSuccess ||
          occ < 1)
        occ = 1;
      per_cu = occ;
    }
    const int cus = B->launch_cus > 0 ? std::min(B->launch_cus, B->ncu) : B->ncu;
    c.grid = std::max(1, std::min(c.nitems, per_cu * cus));
    B->cls.push_back(c);
  }
  if ((rc = hip_rc(hipMemcpy(B->d_sys, B->hsys.data(), B->hsys.size() * sizeof(BatchSys),
                             hipMemcpyHostToDevice))))
    return rc;
  return NECG_OK;
}

struct DevGuard {
  int prev = -1;
  explicit DevGuard(int d) {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    if (prev != d) (void)hipSetDevice(d);
  }
  ~DevGuard() {
    if (prev >= 0) (void)hipSetDevice(prev);
  }
};

int upload_prefix(long long **d, const std::vector<long long> &h) {
  int rc = dalloc(d, h.size());
  if (rc) return rc;
  return hip_rc(hipMemcpy(*d, h.data(), h.size() * sizeof(long long), hipMemcpyHostToDevice));
}

}  // namespace

extern "C" {

int necg_batch_destroy(necg_batch *B) {
  if (!B) return NECG_ERR_ARG;
  {
    DevGuard g(B->device);
    double *bufs[] = {B->sml, B->sms, B->scvv, B->bra, B->ket, B->H, B->S, B->hpart, B->spart,
                      B->part, B->red, B->HC, B->SC, B->hcp, B->scp, B->scal, B->kpart};
    for (double *p : bufs)
      if (p) (void)hipFree(p);
    long long *pre[] = {B->d_fpre, B->d_fwpre, B->d_kwpre, B->d_hpre, B->d_r2pre};
    for (long long *p : pre)
      if (p) (void)hipFree(p);
    if (B->d_sys) (void)hipFree(B->d_sys);
    if (B->flag) (void)hipFree(B->flag);
    if (B->tperm) (void)hipFree(B->tperm);
    for (auto &c : B->cls)
      if (c.d_items) (void)hipFree(c.d_items);
  }
  delete B;
  return NECG_OK;
}

int necg_batch_create(const necg_system *systems, int32_t nsys, const int32_t *Ks, int device,
                      necg_batch **out) {
  if (!systems || !Ks || !out || nsys < 1) return NECG_ERR_ARG;
  *out = nullptr;
  for (int s = 0; s < nsys; ++s) {
    const necg_system &y = systems[s];
    if (y.n < 1 || y.n > kMaxN) return NECG_ERR_UNSUPPORTED;
    if (y.nst < 1 || Ks[s] < 1 || !y.mass || !y.ncv || !y.scvv || !y.sml || !y.sms)
      return NECG_ERR_ARG;
    for (int i = 0; i < y.n; ++i)
      for (int j = 0; j < y.n; ++j)
        if (y.mass[i * y.n + j] != y.mass[j * y.n + i]) return NECG_ERR_ARG;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
    return NECG_ERR_DEVICE;
  necg_batch *B = new (std::nothrow) necg_batch();
  if (!B) return NECG_ERR_OOM;
  B->device = device;
  B->nsys = nsys;
  DevGuard g(device);
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
    necg_batch_destroy(B);
    return NECG_ERR_DEVICE;
  }
  B->ncu = prop.multiProcessorCount;
  // host-side system records and offsets
  std::vector<double> h_sml, h_sms, h_scvv;
  std::vector<unsigned> h_tperm;
  long long xo = 0, bo = 0, ko = 0, hso = 0, ro = 0, vo = 0, hco = 0;
  B->fpre.assign(1, 0);
  B->fwpre.assign(1, 0);
  B->kwpre.assign(1, 0);
  B->hpre.assign(1, 0);
  B->r2pre.assign(1, 0);
  for (int s = 0; s < nsys; ++s) {
    const necg_system &y = systems[s];
    const int n = y.n, K = Ks[s];
    BatchSys b{};
    b.sc.n = n;
    b.sc.nst = y.nst;
    b.sc.efs = y.efs;
    set_mass(b.sc, y.mass, n);
    for (int i = 0; i <= n; ++i) b.sc.q[i] = y.ncv[i];
    set_pair_charges(b.sc, y.ncv, n);
    b.n = n;
    b.nst = y.nst;
    b.K = K;
    b.nacc = dim_nacc(n);
    b.nchunk = (K + 63) / 64;  // H C / S C partials per 64 columns
    b.x_off = xo;
    xo += (long long)K * (dim_blk(n) + 1);
    b.bra_off = bo;
    bo += (long long)K * dim_bs(n);
    b.ket_off = ko;
    ko += (long long)K * y.nst * dim_ks(n);
    b.hs_off = hso;
    hso += (long long)K * K;
    b.red_off = ro;
    ro += (long long)b.nacc * K;
    b.vec_off = vo;
    vo += K;
    b.hcp_off = hco;
    hco += (long long)b.nchunk * K;
    b.sml_off = (long long)h_sml.size();
    h_sml.insert(h_sml.end(), y.sml, y.sml + (size_t)y.nst * n * n);
    b.sms_off = (long long)h_sms.size();
    h_sms.insert(h_sms.end(), y.sms, y.sms + (size_t)y.nst * 9 * n * n);
    b.scvv_off = (long long)h_scvv.size();
    h_scvv.insert(h_scvv.end(), y.scvv, y.scvv + y.nst);
    std::vector<unsigned> tab;
    b.two = two_sided_table(&y, tab) ? 1 : 0;
    b.tperm_off = (long long)h_tperm.size();
    h_tperm.insert(h_tperm.end(), tab.begin(), tab.end());
    B->hsys.push_back(b);
    B->fpre.push_back(B->fpre.back() + K);
    B->fwpre.push_back(B->fwpre.back() + (K + 63) / 64);
    B->kwpre.push_back(B->kwpre.back() + ((long long)K * y.nst + 63) / 64);
    B->hpre.push_back(B->hpre.back() + (long long)b.nchunk * (b.nchunk + 1) / 2);
    B->r2pre.push_back(B->r2pre.back() + (long long)b.nacc * ((K + 63) / 64));
  }
  B->nx = xo;
  int rc = NECG_OK;
  if ((rc = dalloc(&B->sml, h_sml.size())) || (rc = dalloc(&B->sms, h_sms.size())) ||
      (rc = dalloc(&B->scvv, h_scvv.size())) || (rc = dalloc(&B->bra, (size_t)bo)) ||
      (rc = dalloc(&B->ket, (size_t)ko)) || (rc = dalloc(&B->H, (size_t)hso)) ||
      (rc = dalloc(&B->S, (size_t)hso)) || (rc = dalloc(&B->red, (size_t)ro)) ||
      (rc = dalloc(&B->HC, (size_t)vo)) || (rc = dalloc(&B->SC, (size_t)vo)) ||
      (rc = dalloc(&B->hcp, (size_t)hco)) || (rc = dalloc(&B->scp, (size_t)hco)) ||
      (rc = dalloc(&B->scal, (size_t)8 * nsys)) || (rc = dalloc(&B->d_sys, (size_t)nsys)) ||
      (rc = dalloc(&B->flag, 2)) || (rc = upload_prefix(&B->d_fpre, B->fpre)) ||
      (rc = upload_prefix(&B->d_fwpre, B->fwpre)) ||
      (rc = upload_prefix(&B->d_kwpre, B->kwpre)) || (rc = upload_prefix(&B->d_hpre, B->hpre)) ||
      (rc = upload_prefix(&B->d_r2pre, B->r2pre))) {
    necg_batch_destroy(B);
    return rc;
  }
  if ((rc = hip_rc(hipMemcpy(B->sml, h_sml.data(), h_sml.size() * 8, hipMemcpyHostToDevice))) ||
      (rc = hip_rc(hipMemcpy(B->sms, h_sms.data(), h_sms.size() * 8, hipMemcpyHostToDevice))) ||
      (rc = hip_rc(hipMemcpy(B->scvv, h_scvv.data(), h_scvv.size() * 8,
                             hipMemcpyHostToDevice))) ||
      (rc = hip_rc(hipMemset(B->flag, 0, 2 * sizeof(int))))) {
    necg_batch_destroy(B);
    return rc;
  }
  if (!h_tperm.empty() &&
      ((rc = hip_rc(hipMalloc((void **)&B->tperm, h_tperm.size() * sizeof(unsigned)))) ||
       (rc = hip_rc(hipMemcpy(B->tperm, h_tperm.data(), h_tperm.size() * sizeof(unsigned),
                              hipMemcpyHostToDevice))))) {
    necg_batch_destroy(B);
    return rc;
  }
  if ((rc = plan(B))) {
    necg_batch_destroy(B);
    return rc;
  }
  *out = B;
  return NECG_OK;
}

int necg_batch_layout(necg_batch *B, int64_t *x_offsets) {
  if (!B || !x_offsets) return NECG_ERR_ARG;
  for (int s = 0; s < B->nsys; ++s) x_offsets[s] = B->hsys[s].x_off;
  x_offsets[B->nsys] = B->nx;
  return NECG_OK;
}

int necg_batch_set_tiling(necg_batch *B, double item_flops, int32_t grid_per_cu) {
  if (!B || item_flops < 0 || grid_per_cu < 0) return NECG_ERR_ARG;
  std::lock_guard<std::mutex> lk(B->mu);
  if (B->pins) return NECG_ERR_ARG;  // a captured evaluation uses the current lists
  DevGuard g(B->device);
  B->item_flops = item_flops;
  B->grid_per_cu = grid_per_cu;
  return plan(B);
}

int necg_batch_set_launch_cus(necg_batch *B, int32_t ncu) {
  if (!B || ncu < 0) return NECG_ERR_ARG;
  std::lock_guard<std::mutex> lk(B->mu);
  if (B->pins) return NECG_ERR_ARG;
  DevGuard g(B->device);
  B->launch_cus = ncu;
  return plan(B);
}

int necg_batch_pin(necg_batch *B, int32_t pin) {
  if (!B) return NECG_ERR_ARG;
  std::lock_guard<std::mutex> lk(B->mu);
  if (pin) {
    B->pins++;
  } else {
    if (B->pins == 0) return NECG_ERR_ARG;
    B->pins--;
  }
  return NECG_OK;
}

int necg_batch_outprod_masked(necg_batch *B, const int32_t *active, const double *x, double *f,
                              double *g, void *stream) {
  if (!B || !x || !g) return NECG_ERR_ARG;
  NECG_REQUIRE_DEVICE((const void *)active, (const void *)x, (const void *)f, (const void *)g);
  std::lock_guard<std::mutex> lk(B->mu);
  DevGuard dg(B->device);
  hipStream_t st = (hipStream_t)stream;
  const int n = B->nsys;
  const int tb = 256;
  auto blocks = [](long long count, int t) { return (unsigned)((count + t - 1) / t); };
  hipLaunchKernelGGL(b_prep_bra, dim3(blocks(B->fpre[n], tb)), dim3(tb), 0, st, B->d_sys, n,
                     B->d_fpre, active, x, B->bra, B->flag);
  hipLaunchKernelGGL(b_prep_ket, dim3((unsigned)B->kwpre[n]), dim3(64), 0, st, B->d_sys, n,
                     B->d_kwpre, active, B->sml, B->sms, B->scvv, B->bra, B->ket);
  for (const auto &c : B->cls) {
    if (c.two) {
      // the work counter serves whole-batch evaluations only: masked ones may
      // run concurrently on other streams and keep to the fixed stride
      launch_pair2(c.grid, st, B->d_sys, c.d_items, c.nitems, active, B->bra, B->ket, B->tperm,
                   B->part, B->kpart, B->hpart, B->spart, B->flag,
                   active ? nullptr : B->flag + 1);
    } else {
      SWITCH_N(c.n, launch_pair<NN>(c.grid, st, B->d_sys, c.d_items, c.nitems, active, B->bra,
                                    B->ket, B->part, B->hpart, B->spart, B->flag));
    }
  }
  hipLaunchKernelGGL(b_hs_finish, dim3((unsigned)B->hpre[n]), dim3(256), 0, st, B->d_sys, n,
                     B->d_hpre, active, B->hpart, B->spart, B->bra, B->H, B->S, B->hcp, B->scp);
  hipLaunchKernelGGL(b_hc_sum, dim3((unsigned)B->fwpre[n]), dim3(256), 0, st, B->d_sys, n,
                     B->d_fwpre, active, B->hcp, B->scp, B->HC, B->SC);
  hipLaunchKernelGGL(b_quotient, dim3(n), dim3(1024), 0, st, B->d_sys, active, B->bra, B->HC,
                     B->SC, B->scal, f, B->flag);
  hipLaunchKernelGGL(b_tile_reduce, dim3((unsigned)B->r2pre[n]), dim3(256), 0, st, B->d_sys, n,
                     B->d_r2pre, active, B->part, B->kpart, B->red);
  hipLaunchKernelGGL(b_grad_final, dim3((unsigned)B->fwpre[n]), dim3(64), 0, st, B->d_sys, n,
                     B->d_fwpre, active, B->red, B->bra, B->scal, B->HC, B->SC, g);
  return hip_rc(hipGetLastError());
}

int necg_batch_outprod(necg_batch *B, const double *x, double *f, double *g, void *stream) {
  return necg_batch_outprod_masked(B, nullptr, x, f, g, stream);
}

int necg_batch_pool(necg_batch *B, necg_pool **out) {
  if (!B || !out) return NECG_ERR_ARG;
  std::vector<PoolSys> recs;
  for (const BatchSys &b : B->hsys) {
    PoolSys r{};
    r.sc = b.sc;
    r.n = b.n;
    r.nst = b.nst;
    r.K = b.K;
    r.bra = B->bra + b.bra_off;
    r.H = B->H + b.hs_off;
    r.S = B->S + b.hs_off;
    r.HC = B->HC + b.vec_off;
    r.SC = B->SC + b.vec_off;
    r.scal = B->scal + 8 * (size_t)(&b - B->hsys.data());
    r.sml = B->sml + b.sml_off;
    r.sms = B->sms + b.sms_off;
    r.scvv = B->scvv + b.scvv_off;
    recs.push_back(r);
  }
  return pool_create_from(B->device, recs, out);
}

int necg_batch_matrices(necg_batch *B, int32_t s, double *H, double *S, void *stream) {
  if (!B || s < 0 || s >= B->nsys) return NECG_ERR_ARG;
  NECG_REQUIRE_DEVICE((const void *)H, (const void *)S);
  std::lock_guard<std::mutex> lk(B->mu);
  DevGuard dg(B->device);
  const BatchSys &b = B->hsys[s];
  const size_t bytes = sizeof(double) * (size_t)b.K * b.K;
  hipStream_t st = (hipStream_t)stream;
  if (H && hipMemcpyAsync(H, B->H + b.hs_off, bytes, hipMemcpyDeviceToDevice, st) != hipSuccess)
    return NECG_ERR_DEVICE;
  if (S && hipMemcpyAsync(S, B->S + b.hs_off, bytes, hipMemcpyDeviceToDevice, st) != hipSuccess)
    return NECG_ERR_DEVICE;
  return NECG_OK;
}

int necg_batch_status(necg_batch *B, void *stream) {
  if (!B) return NECG_ERR_ARG;
  std::lock_guard<std::mutex> lk(B->mu);
  DevGuard dg(B->device);
  if (hipStreamSynchronize((hipStream_t)stream) != hipSuccess) return NECG_ERR_DEVICE;
  int flags = 0;
  if (hipMemcpy(&flags, B->flag, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
    return NECG_ERR_DEVICE;
  if (flags && hipMemset(B->flag, 0, sizeof(int)) != hipSuccess) return NECG_ERR_DEVICE;
  if (flags & F_BAD_INPUT) return NECG_ERR_NONFINITE;
  if (flags & F_NOT_PD) return NECG_ERR_NOT_PD;
  if (flags & F_CSC) return NECG_ERR_CSC;
  if (flags & F_NONFINITE) return NECG_ERR_NONFINITE;
  return NECG_OK;
}

}  // extern "C"
