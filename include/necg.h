/*
 * necg.h -- C ABI of the MI355X FP64 energy-and-gradient engine for
 * non-Born-Oppenheimer explicitly correlated Gaussians.
 *
 * Every entry point is plain C: integer sizes and raw pointers.  Array
 * arguments marked "device" are HIP device pointers (e.g. torch CUDA tensors'
 * data_ptr()); "stream" is a hipStream_t passed as void* (NULL = the null
 * stream).  Calls are asynchronous on that stream unless stated otherwise, so
 * they can be captured in a hipGraph.  Errors that can only be found on the
 * device (a pair matrix that is not positive definite, C^T S C <= 0, NaN/Inf)
 * are latched in the context and returned by necg_status().
 *
 * Parameter vector x (length K*(blk+1), blk = n(n+1)/2 + 3n): K blocks of
 * [vech(L_k) column-major lower incl. diagonal | s_k particle-major xyz]
 * followed by the K linear coefficients C -- the layout of the reference's
 * outprod(x) (python/outprod.py:40-41, python/matel.py:56-87).
 */
#ifndef NECG_H
#define NECG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct necg_ctx necg_ctx;

enum necg_status_code {
  NECG_OK = 0,
  NECG_ERR_ARG = 1,         /* bad pointer, size or dimension                  */
  NECG_ERR_UNSUPPORTED = 2, /* n outside 1..6, or a system the kernels refuse  */
  NECG_ERR_DEVICE = 3,      /* HIP runtime error                               */
  NECG_ERR_OOM = 4,         /* device allocation failed                        */
  NECG_ERR_NOT_PD = 5,      /* A_k + A_l' not positive definite / L_ii == 0    */
  NECG_ERR_CSC = 6,         /* C^T S C <= 0 (reference: ZeroDivisionError)     */
  NECG_ERR_NONFINITE = 7,   /* NaN or Inf in x or in a result                  */
  NECG_ERR_STALE = 8,       /* necg_energy_bwd without its necg_energy_fwd     */
  NECG_ERR_SCRATCH = 9      /* workspace above necg_outprod_noalloc's scratch  */
};

/* The physics block of an input file: the module globals filled by
 * python/read_input.py:16-157 (python/shared.py:15-66). */
typedef struct {
  int32_t n;          /* pseudo-particles (reference: ne), 1..6            */
  int32_t nst;        /* symmetry terms                                     */
  double efs;         /* electric-field strength                            */
  const double *mass; /* n*n row-major inverse-mass tensor                  */
  const double *ncv;  /* n+1 charges, ncv[0] = reference particle           */
  const double *scvv; /* nst weights ALREADY divided by nst
                         (python/read_input.py:141)                         */
  const double *sml;  /* nst*n*n, sml[p*n*n + i*n + j] = reference sml(i,j,p) */
  const double *sms;  /* nst*(3n)*(3n), same convention                     */
} necg_system;

/* Create a context bound to HIP device `device` for one system.  Replaces the
 * module state that read_input fills (python/read_input.py:16). Host pointers. */
int necg_create(const necg_system *sys, int device, necg_ctx **out);

/* Release the context and all device memory it holds. */
int necg_destroy(necg_ctx *ctx);

/* Energy and gradient: *f = C^T H C / C^T S C and g = dE/dx (length of x).
 * Replaces outprod(x) -> (f, g) (python/outprod.py:20-76, with make_grad,
 * python/make_grad.py:18-90) and the Fortran SFUN outprod(n, x, f, g)
 * (source/outprod.f90:2).  x, f, g: device. Async. */
int necg_outprod(necg_ctx *ctx, int32_t K, const double *x, double *f, double *g, void *stream);

/* necg_outprod without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP-graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same kernels as necg_outprod; every workspace buffer (bra/ket
 * tables, H, S, partials, tile list) is carved out of one static 64 MiB
 * device array per GPU.  Where necg_outprod's tiling fits the array the
 * results are bitwise the same; otherwise the call takes the fastest tiling
 * whose partials fit (the same sums in other groups: equal to rounding), and
 * returns NECG_ERR_SCRATCH if none does (LiH, n = 5 with 24 terms: K = 900
 * fits, K = 1200 does not).  The context's own workspace, resident
 * state and pins are untouched (its tape is dropped, as by any evaluation).
 * All calls on one GPU share that array, which keeps the last call's
 * contents: calls on one stream are ordered, but calls that overlap on one
 * GPU (other streams, other threads, other contexts) race.  x, f, g: device.
 * Async. */
int necg_outprod_noalloc(necg_ctx *ctx, int32_t K, const double *x, double *f, double *g,
                         void *stream);

/* Forward half of outprod: H, S (K*K row-major, symmetric) and *f = E.
 * Replaces the matel loop + quotient (python/outprod.py:49-68).  H and S may be
 * NULL (kept inside the context).  x, f, H, S: device.  Async. */
int necg_energy(necg_ctx *ctx, int32_t K, const double *x, double *f, double *H, double *S,
                void *stream);

/* Forward with a tape: *f = E as necg_energy, running the gradient pair
 * kernel of outprod (python/outprod.py:49-68 and the gradel pair loop of
 * python/make_grad.py:34-55) and keeping its per-function partials on the
 * context for necg_energy_bwd; *tape (host) receives the tape's id.  Any other
 * evaluation on the context drops the tape.  x, f: device.  Async. */
int necg_energy_fwd(necg_ctx *ctx, int32_t K, const double *x, double *f, int64_t *tape,
                    void *stream);

/* Backward of necg_energy_fwd: gx = gbar * dE/dx from the tape (the
 * contraction of make_grad, python/make_grad.py:57-90: GRADA from the pair
 * partials, GRADC from H C and S C).  NECG_ERR_STALE unless the last
 * evaluation on the context was the necg_energy_fwd that returned `tape`, at
 * this K.  gbar: device scalar (NULL = 1); gx: device.  Async. */
int necg_energy_bwd(necg_ctx *ctx, int32_t K, int64_t tape, const double *gbar, double *gx,
                    void *stream);

/* Backward half: gx = gbar * dE/dx for the x of the last necg_energy call on
 * this context (the make_grad contraction, python/make_grad.py:37-90, as an
 * adjoint: H-bar = C C^T / C^T S C, S-bar = -E C C^T / C^T S C).  gbar is a
 * device scalar (NULL = 1).  x, gx: device.  Async. */
int necg_energy_vjp(necg_ctx *ctx, int32_t K, const double *x, const double *gbar, double *gx,
                    void *stream);

/* H and S only (python/necg.py:36-58 _build_matrices; python/outprod.py:49-60).
 * iv has K*blk entries (no C).  Device.  Async. */
int necg_matrices(necg_ctx *ctx, int32_t K, const double *iv, double *H, double *S, void *stream);

/* Kinetic and potential matrices (K*K row-major, symmetric):
 * T = sum_P scvv tk, V = sum_P scvv ert (python/necg.py:43-57; the Stark
 * term enters H only).  x is the full parameter vector.  Device.  Async. */
/* The dipole pass of the driver (python/necg.py:200-232 over python/rrms.py:21-120;
 * source/necg.f90:387-446 over source/rrms.f90): both bra and ket symmetrised,
 * weights scvv[stf] scvv[stfe], contracted with C.  out (device, 3n + 4):
 * out[0] = C^T s C; out[1 + 3 i + a] = <r_i,a> (the reference's d3 of
 * pseudo-particle i, component a); out[3n + 1 + a] = sum_i ncv[i+1] <r_i,a>
 * (d7, d8, d9; times 2.541765 for Debye).  x is the full parameter vector.
 * Async. */
int necg_dipole(necg_ctx *ctx, int32_t K, const double *x, double *out, void *stream);

/* How necg_dipole sums the nst^2 term pairs: 0 (default) uses the reduced
 * form when every term is a permutation with sms = kron(sml^T, I3) and every
 * product sml[q] sml[p]^T is a term (nst term pairs per basis pair, with
 * per-term weights summed over the (p, q) that map to it), else the full
 * double sum; 1 forces the reduced form (NECG_ERR_UNSUPPORTED from
 * necg_dipole if the terms do not allow it); 2 forces the full double sum.
 * No counterpart in the reference, which always forms the double sum. */
int necg_set_dipole_mode(necg_ctx *ctx, int32_t mode);

/* Form of the gradient pair pass of necg_outprod / necg_energy_fwd.  The
 * reference evaluates every ordered pair (l, k) and takes bra derivatives
 * (python/make_grad.py:34-55); when every term is a particle permutation
 * (sms = kron(P^T, I3)) that leaves masses and charges unchanged and P^T is a
 * term of the same weight, the two-sided form evaluates each unordered pair
 * once per term and takes both the bra and the ket derivative (n <= 5).
 * mode 0 (default) uses it whenever it applies; 1 forces it
 * (NECG_ERR_UNSUPPORTED here if it does not apply); 2 forces the per-ordering
 * form.  No counterpart in the reference. */
int necg_set_pair_mode(necg_ctx *ctx, int32_t mode);

/* *available = 1 if the two-sided form applies to the context's system;
 * *last = 1 if the last gradient pair pass used it.  Host pointers. */
int necg_pair_form(necg_ctx *ctx, int32_t *available, int32_t *last);

/* The FP64 vector FMA rate device `device` sustains, in TFLOP/s (*tflops,
 * host pointer): eight independent FMA chains per lane, `iters` steps each,
 * eight 256-lane workgroups per CU, best of three timed launches.  A
 * measurement aid for setting the pair kernels' rates beside the device's own
 * peak; no counterpart in the reference. */
int necg_fp64_fma_rate(int32_t device, int32_t iters, double *tflops);

int necg_kinetic_potential(necg_ctx *ctx, int32_t K, const double *x, double *T, double *V,
                           void *stream);

/* Per-term elements: for query t, (stf, i, j) = q[3t..3t+2] (0-based term,
 * bra, ket) gives out[4t..4t+3] = (ov, tk, ert, zht).  Replaces
 * matel(stf, ii, jj) (python/matel.py:28, 1-based there).  Device.  Async. */
int necg_matel(necg_ctx *ctx, int32_t K, const double *iv, int32_t nq, const int32_t *q,
               double *out, void *stream);

/* Per-term bra gradients: tgvl[blk*t ..], tgvls[blk*t ..] for query t.
 * Replaces gradel(stf, ii, jj) (python/gradel.py:27).  Device.  Async. */
int necg_gradel(necg_ctx *ctx, int32_t K, const double *iv, int32_t nq, const int32_t *q,
                double *tgvl, double *tgvls, void *stream);

/* The derivative matrices of make_grad, materialised:
 * DH[(l*blk + r)*K + k] = sum_stf scvv * gradel(stf, l, k)[r], DS likewise
 * (python/make_grad.py:34-55).  Device, K*blk*K each.  Async. */
int necg_dhds(necg_ctx *ctx, int32_t K, const double *iv, double *DH, double *DS, void *stream);

/* 1 if the GPU can address p (device, managed or pinned host memory), else
 * 0.  Every entry point applies this check to its pointer arguments and
 * returns NECG_ERR_ARG for plain (pageable) host memory.  No counterpart in
 * the reference, whose arrays are host NumPy/Fortran arrays. */
int necg_device_visible(const void *p);

/* Synchronise `stream` and return the first latched error (NECG_OK if none);
 * clears the latch.  Blocking. */
int necg_status(necg_ctx *ctx, void *stream);

/* Launch shape of the pair kernels: kets per workgroup and term groups (0 keeps
 * the built-in choice).  The waves per workgroup are fixed per n (4 for n <= 5,
 * 2 for n = 6): `waves` must be 0 or that value.  Used by the autotuner. */
int necg_set_tiling(necg_ctx *ctx, int32_t ket_tile, int32_t term_split, int32_t waves);

/* Compute units that the pair passes size their launches for (0: the
 * device's own count).  A smaller count suits a device shared with other
 * work; it also changes which tiles of the lower-triangle passes are split
 * into halves for the last round.  No counterpart in the reference. */
int necg_set_launch_cus(necg_ctx *ctx, int32_t ncu);

/* Folded diagonal blocks in the lower-triangle passes (fold = 1, the
 * default, or 0): when the ket tile divides 32, each full 64-function bra
 * tile's diagonal block is evaluated in 32 sweeps of two kets each (every
 * lane busy) instead of 64 half-idle ones, and a small diagonal pass adds
 * the diagonals of the block's upper half.  Results agree to rounding.  No
 * counterpart in the reference. */
int necg_set_fold(necg_ctx *ctx, int32_t fold);

/* Workspace pin for HIP-graph capture: pin = 1 adds a pin, pin = 0 removes
 * one.  A captured evaluation holds the addresses of the context's workspace
 * and reads its lower-triangle tile list, which belong to the K of the
 * capture.  While any pin is held, a call that would reallocate a workspace
 * buffer or rebuild the tile list (an evaluation at another K, or one whose
 * pass needs more workspace than the captured one) returns NECG_ERR_ARG
 * instead, and so do necg_set_tiling, necg_set_pair_mode and
 * necg_set_launch_cus.  Evaluations that fit the pinned workspace run as
 * usual.  No counterpart in the reference, which has no graph capture. */
int necg_pin_workspace(necg_ctx *ctx, int32_t pin);

/* Launch shape that the last pair-kernel launch used (for reports). */
int necg_get_tiling(necg_ctx *ctx, int32_t *ket_tile, int32_t *term_split, int32_t *waves);

/* Pair-kernel timing with HIP events recorded on the launch stream around
 * every pair-kernel launch.  mode 1: reset and start recording; mode 0: stop;
 * mode 2: read.  *ms = summed milliseconds, *launches = launch count since the
 * last reset (read synchronises the events).  Not for use under graph capture. */
int necg_timing(necg_ctx *ctx, int32_t mode, double *ms, int64_t *launches);

/* ---- batches: many independent systems in one evaluation ---------------- */

typedef struct necg_batch necg_batch;

/* A batch of nsys systems (uneven n, nst, K).  Systems are laid out back to
 * back in one x vector (and one g vector of the same layout); see
 * necg_batch_layout.  No counterpart in the reference, which evaluates one
 * system per process; each system's numbers are those of outprod. */
int necg_batch_create(const necg_system *systems, int32_t nsys, const int32_t *K, int device,
                      necg_batch **out);
int necg_batch_destroy(necg_batch *batch);

/* x_offsets[s] = start of system s in x and g; x_offsets[nsys] = total. */
int necg_batch_layout(necg_batch *batch, int64_t *x_offsets);

/* Energy and gradient of every system: f[s] = E_s, g = dE_s/dx_s per segment.
 * A fixed sequence of launches (no host sync, no memset): capturable in a
 * hipGraph.  Whole-batch evaluations of one batch must not overlap (they
 * share the pair pass's work counter as well as every result).  x, f, g:
 * device.  Async. */
int necg_batch_outprod(necg_batch *batch, const double *x, double *f, double *g, void *stream);

/* As necg_batch_outprod for the systems s with active[s] != 0 only (active:
 * device, nsys entries); the others are skipped and keep their results and
 * resident state.  Evaluations with disjoint active sets may run concurrently
 * on different streams.  Async. */
int necg_batch_outprod_masked(necg_batch *batch, const int32_t *active, const double *x,
                              double *f, double *g, void *stream);

/* H and S (K_s x K_s) of system s from the last evaluation.  Device.  Async. */
int necg_batch_matrices(necg_batch *batch, int32_t s, double *H, double *S, void *stream);

/* Pair-tile cost target in model FLOP (0: default 6e7) and persistent
 * workgroups per CU (0: occupancy limit).  Re-plans the work lists; used by
 * the autotuner.  Blocking. */
int necg_batch_set_tiling(necg_batch *batch, double item_flops, int32_t grid_per_cu);

/* Size the batch's persistent pair launches for ncu CUs (0: the device's
 * count), leaving the others free of pair workgroups while one runs: a
 * serving loop keeps CUs for its trial kernels this way.  No counterpart in
 * the reference.  Re-plans; NECG_ERR_ARG while a pin is held.  Blocking. */
int necg_batch_set_launch_cus(necg_batch *batch, int32_t ncu);

/* Work-list pin for HIP-graph capture (pin = 1 adds, 0 removes one): a
 * captured batch evaluation holds the addresses of the batch's work lists
 * and partial buffers, which necg_batch_set_tiling replaces.  While any pin
 * is held, necg_batch_set_tiling returns NECG_ERR_ARG.  No counterpart in the
 * reference. */
int necg_batch_pin(necg_batch *batch, int32_t pin);

/* As necg_status, for a batch.  Blocking. */
int necg_batch_status(necg_batch *batch, void *stream);

/* ---- serving: single-function replacement trials on resident systems ---- */

typedef struct necg_pool necg_pool;

/* A pool over contexts that have each completed one evaluation (necg_outprod,
 * necg_energy or necg_energy_fwd) at their working K; the pool reads their
 * resident H, S, H C, S C, C^T H C and C^T S C.  Trials are refused
 * (NECG_ERR_ARG) if a context was re-dimensioned since, or if a later call
 * (necg_matrices, necg_matel, necg_energy_vjp, ...) rebuilt its tables or
 * matrices without a full evaluation after it.  No counterpart in the reference: it has
 * no serving mode; a trial's row is the matel loop (python/outprod.py:49-54)
 * restricted to one row and its energy the quotient of python/outprod.py:65-68. */
int necg_pool_create(necg_ctx *const *ctxs, int32_t nctx, necg_pool **out);
int necg_pool_destroy(necg_pool *pool);

/* Largest K in the pool and the per-trial workspace stride (doubles). */
int necg_pool_sizes(necg_pool *pool, int32_t *kmax, int32_t *ws_stride);

/* ntr trials: trial t replaces function tfun[t] (0-based) of pool system
 * tsys[t] by the block newp[t*pstride ..] = [vech L | s] and yields the new
 * row of H and S (rowH/rowS[t*rstride + j], rstride >= kmax) and
 * E[t] = C^T H' C / C^T S' C with the resident C.  The resident state is not
 * modified.  All pointers are device buffers owned by the caller (ws:
 * ntr*ws_stride doubles), so concurrent calls on different streams are safe
 * while no evaluation of the same contexts runs.  Async. */
int necg_pool_trials(necg_pool *pool, int32_t ntr, const int32_t *tsys, const int32_t *tfun,
                     const double *newp, int32_t pstride, double *ws, double *E, double *rowH,
                     double *rowS, int32_t rstride, void *stream);

/* Resident factors of pool system s (device, owned by the caller, valid while
 * the pool is used): V (K x K row-major) with V^T S V = I, V^T H V = diag(lam)
 * and lam ascending -- the generalised eigenproblem that initialises C in the
 * driver (python/necg.py:129-134, scipy.linalg.eigh; DSYGV at
 * source/necg.f90:175).  V = lam = NULL removes them.  Blocking. */
int necg_pool_set_factors(necg_pool *pool, int32_t s, const double *V, const double *lam);

/* After necg_pool_trials on the same stream (same tsys/tfun/rowH/rowS/E):
 * Eeig[t] = the lowest eigenvalue of the trial's pencil (H', S'), i.e. the
 * energy with C re-optimised for the replaced basis, from the resident factors
 * in O(K^2) per trial.  NaN where the system has no factors (and the
 * bad-input status is latched) or where S' is not positive definite by a
 * margin (smallest eigenvalue of S' relative to S below 1e-10; no status).
 * bg: device workspace of ntr * 18 * kmax doubles (kmax: necg_pool_sizes).
 * Async. */
int necg_pool_eig(necg_pool *pool, int32_t ntr, const int32_t *tsys, const int32_t *tfun,
                  const double *rowH, const double *rowS, int32_t rstride, const double *E,
                  double *bg, double *Eeig, void *stream);

/* necg_pool_trials and (bg not NULL) necg_pool_eig in one call, with the
 * copies in and out: no counterpart in the reference; it is for serving
 * loops, where every host-side call adds latency.  On `stream`: copies in_h
 * (host, pinned for an asynchronous copy) to in_d (device, 8-byte aligned),
 * laid out [tsys (ntr int32) | tfun (ntr int32) | newp (ntr x pstride
 * doubles, from byte (8 ntr + 7) & ~7)]; runs the trials with E into out_d[0,
 * ntr) and Eeig into out_d[ntr, 2 ntr); copies out_d's ntr (or 2 ntr) values
 * to out_h (host).  Results as the two calls; the caller synchronises on
 * `stream` before reading out_h.  ws, rowH, rowS, bg as there.  Async. */
int necg_pool_trials_io(necg_pool *pool, int32_t ntr, const void *in_h, void *in_d,
                        int32_t pstride, double *ws, double *rowH, double *rowS, int32_t rstride,
                        double *bg, double *out_d, double *out_h, void *stream);

/* A pool over the resident state of every system of a batch (after one full
 * evaluation of the batch).  The batch must outlive the pool. */
int necg_batch_pool(necg_batch *batch, necg_pool **out);

/* As necg_status, for the pool's trials.  Blocking. */
int necg_pool_status(necg_pool *pool, void *stream);

/* Human-readable text of a status code. */
const char *necg_strerror(int code);

#ifdef __cplusplus
}
#endif

#endif /* NECG_H */
