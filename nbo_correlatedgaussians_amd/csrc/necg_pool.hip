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

// Workgroup per trial: the new bra record, then its ket record under every term.
__global__ void __launch_bounds__(256) t_prep(const PoolSys *__restrict__ sys,
                                              const int32_t *__restrict__ tsys,
                                              const int32_t *__restrict__ tfun,
                                              const double *__restrict__ newp, int pstride,
                                              double *__restrict__ ws, long long wstride,
                                              int *__restrict__ flag) {
  const int t = blockIdx.x;
  const PoolSys &b = sys[tsys[t]];
  const int k = tfun[t];
  double *w = ws + (size_t)t * wstride;
  if (threadIdx.x == 0) {
    POOL_SWITCH_N(b.n, (make_bra<NN>(newp + (size_t)t * pstride,
                                     b.bra[(size_t)k * Dim<NN>::BS + Dim<NN>::B_C], true, w,
                                     flag)));
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
  for (int p = 0; p < b.nst; ++p)
    ok &= pair_term<N, MODE_ELEMENTS>(br, w + D::BS + (size_t)p * D::KS, b.sc, 0.0, h, s, none,
                                      nullptr);
  if (!ok) atomicOr(flag, (int)F_NOT_PD);
  *rh = h;
  *rs = s;
}

// Grid (ceil(Kmax/64), ntr), 64 lanes: lane j = bra j of the new row.
__global__ void __launch_bounds__(64) t_rows(const PoolSys *__restrict__ sys,
                                             const int32_t *__restrict__ tsys,
                                             const int32_t *__restrict__ tfun,
                                             const double *__restrict__ ws, long long wstride,
                                             double *__restrict__ rowH, double *__restrict__ rowS,
                                             int rstride, int *__restrict__ flag) {
  const int t = blockIdx.y;
  const PoolSys &b = sys[tsys[t]];
  const int j = blockIdx.x * 64 + threadIdx.x;
  if (j >= b.K) return;
  const int k = tfun[t];
  double *rh = rowH + (size_t)t * rstride + j, *rs = rowS + (size_t)t * rstride + j;
  POOL_SWITCH_N(b.n, trial_row<NN>(b, k, ws + (size_t)t * wstride, j, rh, rs, flag));
}

// Workgroup of 256 per trial: the updated quotient.
__global__ void __launch_bounds__(256) t_energy(const PoolSys *__restrict__ sys,
                                                const int32_t *__restrict__ tsys,
                                                const int32_t *__restrict__ tfun,
                                                const double *__restrict__ rowH,
                                                const double *__restrict__ rowS, int rstride,
                                                double *__restrict__ E, int *__restrict__ flag) {
  __shared__ double red[8];
  const int t = blockIdx.x;
  const PoolSys &b = sys[tsys[t]];
  const int k = tfun[t];
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

bool snapshot(necg_pool *P, std::vector<PoolSys> &out) {
  out.clear();
  for (necg_ctx *c : P->ctx) {
    if (c->last_K < 1 || !c->bra || !c->H || !c->HC || !c->scal) return false;
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
  // a context-backed pool: the contexts must still hold the state the pool
  // was made from
  if (!P->ctx.empty()) {
    std::vector<PoolSys> now;
    if (!snapshot(P, now)) return NECG_ERR_ARG;
    for (size_t i = 0; i < now.size(); ++i)
      if (now[i].bra != P->hsys[i].bra || now[i].H != P->hsys[i].H || now[i].K != P->hsys[i].K)
        return NECG_ERR_ARG;
  }
  int prev = -1;
  (void)hipGetDevice(&prev);
  if (prev != P->device) (void)hipSetDevice(P->device);
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(t_prep, dim3(ntr), dim3(256), 0, st, P->d_sys, tsys, tfun, newp, pstride, ws,
                     (long long)P->wstride, P->flag);
  hipLaunchKernelGGL(t_rows, dim3((P->kmax + 63) / 64, ntr), dim3(64), 0, st, P->d_sys, tsys, tfun,
                     ws, (long long)P->wstride, rowH, rowS, rstride, P->flag);
  hipLaunchKernelGGL(t_energy, dim3(ntr), dim3(256), 0, st, P->d_sys, tsys, tfun, rowH, rowS,
                     rstride, E, P->flag);
  const int rc = hipGetLastError() == hipSuccess ? NECG_OK : NECG_ERR_DEVICE;
  if (prev >= 0 && prev != P->device) (void)hipSetDevice(prev);
  return rc;
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
