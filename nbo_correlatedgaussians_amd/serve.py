"""Serving: energy and gradient requests against systems resident in HBM.

Each resident system keeps, on the device, its parameters and the state of
its last evaluation (bra/ket tables, H, S, H C, S C, C^T H C, C^T S C).  Two
request kinds arrive over time:

* ``trial``: replace basis function k of system s by new (L_k, s_k); the new
  row/column of H and S, the trial energy C^T H' C / C^T S' C (resident C) and,
  from the resident factors of (H, S), the lowest eigenvalue of the trial
  pencil (H', S') (C re-optimised) are computed; the resident state is not
  modified.
* ``full``: energy and analytic gradient of system s at its resident x
  (``outprod``, python/outprod.py:20).
* ``accept``: make a replacement resident: function k of system s becomes the
  new (L_k, s_k) in the resident x, the system is evaluated again (energy and
  gradient at the new x, returned like a full request, with H, S, H C, S C
  and the quotient refreshed) and the factors of its pencil are formed again,
  so that later trials on the system start from the new basis.

The scheduler does continuous batching: on every pass it collects the
requests that have arrived, admits as many trials as the free memory budget
allows into one batched trial launch on a free trial stream, and admits the
waiting full evaluations (one per system) into one masked batch evaluation on
a free evaluation stream; launches on different streams overlap on the device
and complete independently, and a new pass admits whatever arrived meanwhile.
A system with a full evaluation or an accept in flight takes no new work
until it ends, a full evaluation or an accept waits for its system's trials
in flight, and a trial that arrived after a waiting full evaluation or accept
of its system waits behind it.
Latency is measured from a request's arrival to its results being on the
host.  The loop polls (event queries, arrivals) and sleeps only while it is
idle and the next arrival is more than a millisecond away.
"""

from __future__ import annotations

import ctypes as ct
import math
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .engine import Engine
from .system import generate_parameters


def _torch():
    import torch
    return torch


@dataclass
class Request:
    rid: int
    kind: str                 # "trial", "full" or "accept"
    sys: int
    arrival: float            # seconds from the start of the run
    func: int = -1            # trial, accept: replaced function (0-based)
    newp: np.ndarray = None   # trial, accept: new [vech L | s] block
    done: float = None
    E: float = None
    Eeig: float = None        # trial: lowest eigenvalue of (H', S')
    g: np.ndarray = None
    pairs: int = 0            # unordered basis pairs of the work

    @property
    def latency(self):
        return None if self.done is None else self.done - self.arrival


@dataclass
class _Launch:
    kind: str
    slot: int
    event: object
    reqs: list
    nbytes: int
    systems: set
    host: dict = field(default_factory=dict)


class Server:
    """Resident systems on one GPU with continuous batching of requests.

    All systems live in one ``Batch`` (their parameters, bra/ket tables, H, S,
    H C, S C and quotients resident in HBM) together with the factors of
    their pencils (V with V^T S V = I, V^T H V = diag(lam)); trials run
    through a pool over that resident state, and the full evaluations admitted
    in one scheduling pass run as one masked batch evaluation (only their
    systems active)."""

    def __init__(self, systems, xs, device=None, mem_budget: int = 4 << 30,
                 max_trials: int = 512, trial_streams: int = 2, eval_streams: int = 2,
                 item_flops: float = 0.0, tune: bool = False, factors: bool = True,
                 reserve_cus: int = 0):
        torch = _torch()
        from .batch import Batch, autotune
        self.device = torch.device(device if device is not None else "cuda:0")
        self.systems = list(systems)
        self.xs = [np.array(x, dtype=np.float64) for x in xs]   # own copies: accepts change them
        Ks = []
        for s, x in zip(self.systems, self.xs):
            if x.size % (s.blk + 1):
                raise ValueError("x length is not a multiple of (blk + 1)")
            Ks.append(x.size // (s.blk + 1))
        self.Ks = Ks
        self.batch = Batch(self.systems, Ks, self.device)
        if item_flops:
            self.batch.set_tiling(item_flops, 0)
        self.xb = torch.as_tensor(self.batch.pack(self.xs), device=self.device)
        if tune:                                # pair-tile size timed on these systems
            autotune(self.batch, self.xb.cpu().numpy(), reps=2)
        if reserve_cus > 0:                     # CUs that full evaluations leave to trials
            ncu = torch.cuda.get_device_properties(self.device).multi_processor_count
            self.batch.set_launch_cus(max(1, ncu - int(reserve_cus)))
        self.batch.outprod_t(self.xb)           # resident state of every system
        torch.cuda.synchronize(self.device)
        self.batch.check("initial evaluation")
        self.offsets = self.batch.offsets
        self.lib = _lib.load()
        self._pool = self.batch.pool()
        kmax, wst = ct.c_int32(), ct.c_int32()
        self.lib.necg_pool_sizes(self._pool, ct.byref(kmax), ct.byref(wst))
        self.kmax, self.wstride = kmax.value, wst.value
        self.pmax = max(s.blk for s in systems)
        self.factors = factors
        self._fac = []                          # (V, lam) per system, kept alive here
        if factors:
            self._make_factors()
        self.max_trials = max_trials
        self.mem_budget = mem_budget
        nsys = len(self.systems)
        f64 = dict(dtype=torch.float64, device=self.device)
        i32 = dict(dtype=torch.int32, device=self.device)
        self._tbuf = []
        for _ in range(trial_streams):
            T = max_trials
            # inputs packed as [tsys (n int32) | tfun (n int32) | newp (n x pmax
            # doubles, 8-byte aligned)] and outputs as [E (n) | Eeig (n)]: one
            # copy each way per launch
            nin = 8 * T + 8 * T * self.pmax + 8
            self._tbuf.append(dict(
                stream=torch.cuda.Stream(device=self.device),
                d_in=torch.empty(nin, dtype=torch.uint8, device=self.device),
                h_in=torch.empty(nin, dtype=torch.uint8).pin_memory(),
                ws=torch.empty(T * self.wstride, **f64),
                rowH=torch.empty(T * self.kmax, **f64), rowS=torch.empty(T * self.kmax, **f64),
                bg=torch.empty(T * 18 * self.kmax, **f64),
                d_out=torch.empty(2 * T, **f64),
                h_out=torch.empty(2 * T, dtype=torch.float64).pin_memory(),
                ev=torch.cuda.Event()))   # reused: a slot runs one launch at a time
        self._ebuf = []
        for _ in range(eval_streams):
            self._ebuf.append(dict(
                stream=torch.cuda.Stream(device=self.device),
                active=torch.zeros(nsys, **i32), f=torch.empty(nsys, **f64),
                g=torch.empty(self.batch.nx, **f64),
                h_active=torch.zeros(nsys, dtype=torch.int32).pin_memory(),
                h_f=torch.empty(nsys, dtype=torch.float64).pin_memory(),
                h_g=torch.empty(self.batch.nx, dtype=torch.float64).pin_memory(),
                ev=torch.cuda.Event()))
        self._engines = None
        # bytes a request holds while in flight (admission budget)
        self.trial_bytes = 8 * (self.pmax + self.wstride + 2 + 20 * self.kmax)
        self.full_bytes = [self._eval_bytes(s, K) for s, K in zip(self.systems, self.Ks)]

    def _make_factors(self):
        """V, lam of every system's pencil (H, S) from the resident matrices."""
        torch = _torch()
        self.factor_bytes = 0
        for s, K in enumerate(self.Ks):
            self._fac.append(None)
            self._factor(s)
            self.factor_bytes += 8 * (K * K + K)
        torch.cuda.synchronize(self.device)

    def _factor(self, s: int):
        """(Re)forms the factors of system s from its resident H and S, on the
        current stream, and hands them to the pool (which synchronises)."""
        torch = _torch()
        from .driver import generalized_eigh
        K = self.Ks[s]
        H = torch.empty(K, K, dtype=torch.float64, device=self.device)
        S = torch.empty(K, K, dtype=torch.float64, device=self.device)
        st = ct.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)
        _lib.raise_for(self.lib.necg_batch_matrices(self.batch._h, s, ct.c_void_p(H.data_ptr()),
                                                    ct.c_void_p(S.data_ptr()), st),
                       "necg_batch_matrices")
        lam, V = generalized_eigh(H, S)
        del H, S
        torch.cuda.current_stream(self.device).synchronize()
        _lib.raise_for(self.lib.necg_pool_set_factors(self._pool, s, ct.c_void_p(V.data_ptr()),
                                                      ct.c_void_p(lam.data_ptr())),
                       "necg_pool_set_factors")
        self._fac[s] = (V, lam)     # the previous factors are released only now

    @staticmethod
    def _eval_bytes(system, K):
        """Device bytes one evaluation of this system works through: H, S and
        the element partials (4 K^2), the ket table (K nst KS), the gradient
        partials (about K/16 tiles of NACC x K) and the H C / S C partials of
        the 64-column chunks."""
        n, nst = system.n, system.nst
        nacc = n * (n + 1) + 6 * n + 2
        ks = (n * (n + 1) + 6 * n + 4 + n * n + 1 + 15) // 16 * 16   # ket record, whole lines
        return 8 * (4 * K * K + K * nst * ks + max(1, K // 16) * nacc * K
                    + 2 * ((K + 63) // 64) * K)

    def close(self):
        if getattr(self, "_pool", None) and self._pool.value:
            self.lib.necg_pool_destroy(self._pool)
            self._pool = ct.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- launches ---------------------------------------------------------
    def _launch_trials(self, slot, reqs):
        b = self._tbuf[slot]
        n = len(reqs)
        P = self.pmax
        hb = b["h_in"].numpy()
        off_p = (8 * n + 7) & ~7                      # newp: 8-byte aligned after the indices
        nbytes = off_p + 8 * n * P
        idx = hb[:8 * n].view(np.int32)
        idx[:n] = [r.sys for r in reqs]
        idx[n:] = [r.func for r in reqs]
        newp = hb[off_p:nbytes].view(np.float64)
        for i, r in enumerate(reqs):
            newp[i * P:i * P + r.newp.size] = r.newp
        st = ct.c_void_p(b["stream"].cuda_stream)
        # copy in, the trial and eigenvalue kernels, copy out: one call
        rc = self.lib.necg_pool_trials_io(
            self._pool, n, ct.c_void_p(b["h_in"].data_ptr()), ct.c_void_p(b["d_in"].data_ptr()),
            P, ct.c_void_p(b["ws"].data_ptr()), ct.c_void_p(b["rowH"].data_ptr()),
            ct.c_void_p(b["rowS"].data_ptr()), self.kmax,
            ct.c_void_p(b["bg"].data_ptr()) if self.factors else None,
            ct.c_void_p(b["d_out"].data_ptr()), ct.c_void_p(b["h_out"].data_ptr()), st)
        _lib.raise_for(rc, "necg_pool_trials_io")
        b["ev"].record(b["stream"])
        return _Launch("trial", slot, b["ev"], reqs, self.trial_bytes * n, {r.sys for r in reqs})

    def _launch_full(self, slot, reqs):
        """One masked batch evaluation of the requests' systems (distinct);
        an accept request first writes its block into the resident x."""
        torch = _torch()
        b = self._ebuf[slot]
        ha = b["h_active"].numpy()
        ha[:] = 0
        for r in reqs:
            ha[r.sys] = 1
        with torch.cuda.stream(b["stream"]):
            for r in reqs:
                if r.kind == "accept":
                    blk = self.systems[r.sys].blk
                    lo = int(self.offsets[r.sys]) + r.func * blk
                    newp = torch.as_tensor(np.asarray(r.newp, dtype=np.float64))
                    self.xb[lo:lo + blk].copy_(newp.pin_memory(), non_blocking=True)
                    self.xs[r.sys][r.func * blk:(r.func + 1) * blk] = r.newp
            b["active"].copy_(b["h_active"], non_blocking=True)
            self.batch.outprod_t(self.xb, b["f"], b["g"], active=b["active"])
            b["h_f"].copy_(b["f"], non_blocking=True)
            for r in reqs:
                lo, hi = int(self.offsets[r.sys]), int(self.offsets[r.sys + 1])
                b["h_g"][lo:hi].copy_(b["g"][lo:hi], non_blocking=True)
            b["ev"].record(b["stream"])
        nbytes = sum(self.full_bytes[r.sys] for r in reqs)
        return _Launch("full", slot, b["ev"], list(reqs), nbytes, {r.sys for r in reqs})

    # -- the loop ---------------------------------------------------------
    def run(self, requests, on_pass=None):
        """Serve the requests (sorted by arrival, times in seconds from now);
        fills in done/E/g of each and returns them."""
        torch = _torch()
        reqs = sorted(requests, key=lambda r: r.arrival)
        for r in reqs:
            self._validate(r)
            K = self.Ks[r.sys]
            r.pairs = K if r.kind == "trial" else K * (K + 1) // 2
        total = len(reqs)
        ptr = 0
        ready_t, ready_f = deque(), deque()
        inflight = []
        free_t = list(range(len(self._tbuf)))
        free_e = list(range(len(self._ebuf)))
        full_busy = set()       # systems with a full evaluation in flight
        trial_busy = {}         # system -> trials in flight
        used = 0
        finished = 0
        t0 = time.perf_counter()
        while finished < total:
            now = time.perf_counter() - t0
            while ptr < total and reqs[ptr].arrival <= now:
                r = reqs[ptr]
                (ready_t if r.kind == "trial" else ready_f).append(r)
                ptr += 1
            # completions
            still = []
            for L in inflight:
                if not L.event.query():
                    still.append(L)
                    continue
                tdone = time.perf_counter() - t0
                if L.kind == "trial":
                    b = self._tbuf[L.slot]
                    out = b["h_out"].numpy()
                    nl = len(L.reqs)
                    for i, r in enumerate(L.reqs):
                        r.E = float(out[i])
                        r.Eeig = float(out[nl + i]) if self.factors else None
                        r.done = tdone
                        trial_busy[r.sys] -= 1
                    free_t.append(L.slot)
                else:
                    b = self._ebuf[L.slot]
                    fh, gh = b["h_f"].numpy(), b["h_g"].numpy()
                    for r in L.reqs:
                        r.E = float(fh[r.sys])
                        r.g = gh[self.offsets[r.sys]:self.offsets[r.sys + 1]].copy()
                        if r.kind == "accept" and self.factors:
                            with torch.cuda.stream(b["stream"]):
                                self._factor(r.sys)     # the new pencil's factors
                        r.done = time.perf_counter() - t0 if r.kind == "accept" else tdone
                        full_busy.discard(r.sys)
                    free_e.append(L.slot)
                used -= L.nbytes
                finished += len(L.reqs)
            inflight = still
            # a trial that arrived after a waiting full evaluation of its
            # system waits behind it (per-system arrival order)
            first_full = {}
            for r in ready_f:
                first_full.setdefault(r.sys, r.arrival)
            # admit trials: one batched launch per free trial stream
            while free_t and ready_t:
                take = []
                keep = deque()
                while ready_t and len(take) < self.max_trials:
                    r = ready_t.popleft()
                    if r.sys in full_busy or r.arrival > first_full.get(r.sys, math.inf):
                        keep.append(r)
                        continue
                    if used + self.trial_bytes * (len(take) + 1) > self.mem_budget:
                        keep.append(r)              # budget full: stop admitting
                        break
                    take.append(r)
                ready_t.extendleft(reversed(keep))
                if not take:
                    break
                for r in take:
                    trial_busy[r.sys] = trial_busy.get(r.sys, 0) + 1
                L = self._launch_trials(free_t.pop(), take)
                used += L.nbytes
                inflight.append(L)
            # admit full evaluations: one masked batch launch per free stream
            while free_e and ready_f:
                take, keep, chosen, nb = [], deque(), set(), 0
                while ready_f:
                    r = ready_f.popleft()
                    if (r.sys in full_busy or r.sys in chosen or trial_busy.get(r.sys, 0) > 0
                            or used + nb + self.full_bytes[r.sys] > self.mem_budget):
                        keep.append(r)
                        continue
                    take.append(r)
                    chosen.add(r.sys)
                    nb += self.full_bytes[r.sys]
                ready_f.extendleft(reversed(keep))
                if not take:
                    break
                full_busy |= chosen
                L = self._launch_full(free_e.pop(), take)
                used += L.nbytes
                inflight.append(L)
            if on_pass:
                on_pass(len(inflight), len(ready_t) + len(ready_f))
            if not inflight and ptr < total and not ready_t and not ready_f:
                # idle: sleep only while the next arrival is more than 1 ms
                # away, and wake 0.5 ms early; the rest is polled (a sleep
                # overshoots by tens of microseconds, which a trial of about
                # 60 us on the device would pay in full)
                wait = reqs[ptr].arrival - (time.perf_counter() - t0)
                if wait > 1e-3:
                    time.sleep(wait - 5e-4)
        self.batch.check("serving")
        _lib.raise_for(self.lib.necg_pool_status(self._pool, ct.c_void_p(
            self._tbuf[0]["stream"].cuda_stream)), "serving trials")
        return reqs

    def _validate(self, r: Request):
        if r.kind not in ("trial", "full", "accept"):
            raise ValueError(f"request {r.rid}: unknown kind {r.kind!r}")
        if not 0 <= r.sys < len(self.systems):
            raise ValueError(f"request {r.rid}: no system {r.sys}")
        if r.kind in ("trial", "accept"):
            if not 0 <= r.func < self.Ks[r.sys]:
                raise ValueError(f"request {r.rid}: system {r.sys} has no function {r.func}")
            blk = self.systems[r.sys].blk
            if r.newp is None or np.asarray(r.newp).size != blk:
                raise ValueError(f"request {r.rid}: new block must have {blk} numbers")

    def full_alone(self, s: int):
        """(E, g) of system s from a masked evaluation with only s active
        (reference for the in-mix result)."""
        L = self._launch_full(0, [Request(0, "full", s, 0.0)])
        L.event.synchronize()
        self.batch.check("full evaluation")
        b = self._ebuf[0]
        return float(b["h_f"][s]), b["h_g"].numpy()[self.offsets[s]:self.offsets[s + 1]].copy()

    def engines(self):
        """Per-system engines (the plain per-system path), built on first use."""
        if self._engines is None:
            self._engines = [Engine(s, self.device) for s in self.systems]
        return self._engines

    def run_plain(self, requests):
        """The same requests one at a time, each launched alone and waited
        for (no batching, no overlap): trials one by one through the pool,
        full evaluations through each system's own engine (``outprod``).
        Returns (seconds, pairs)."""
        torch = _torch()
        engines = self.engines()
        xd = [torch.as_tensor(x, device=self.device) for x in self.xs]
        f = torch.empty(1, dtype=torch.float64, device=self.device)
        g = torch.empty(max(x.numel() for x in xd), dtype=torch.float64, device=self.device)
        torch.cuda.synchronize(self.device)
        t0 = time.perf_counter()
        pairs = 0
        for r in requests:
            K = self.Ks[r.sys]
            if r.kind == "trial":
                L = self._launch_trials(0, [Request(r.rid, "trial", r.sys, 0.0, r.func, r.newp)])
                L.event.synchronize()
                pairs += K
            else:
                nx = xd[r.sys].numel()
                engines[r.sys].outprod_t(xd[r.sys], f, g[:nx])
                E, gh = float(f.item()), g[:nx].cpu()
                pairs += K * (K + 1) // 2
        torch.cuda.synchronize(self.device)
        dt = time.perf_counter() - t0
        for e in engines:
            e.check("plain serving")
        return dt, pairs

    def trial_alone(self, r: Request):
        """One trial launched by itself (reference for the in-mix result):
        (E, (row of H', row of S'), lowest eigenvalue of (H', S') or None)."""
        q = Request(r.rid, "trial", r.sys, 0.0, r.func, r.newp)
        self._validate(q)
        L = self._launch_trials(0, [q])
        L.event.synchronize()
        _lib.raise_for(self.lib.necg_pool_status(self._pool, ct.c_void_p(
            self._tbuf[0]["stream"].cuda_stream)), "trial")
        b = self._tbuf[0]
        return float(b["h_out"][0]), (b["rowH"][:self.Ks[r.sys]].cpu().numpy(),
                                      b["rowS"][:self.Ks[r.sys]].cpu().numpy()), (
            float(b["h_out"][1]) if self.factors else None)


def poisson_trace(systems, Ks, count: int, rate: float, trial_frac: float = 0.7,
                  seed: int = 0):
    """count requests with exponential inter-arrival times (rate per second);
    trial_frac of them are single-function replacement trials on a uniformly
    chosen system and function, with a new block drawn from the parameter
    distributions of generate_parameters; the rest are full evaluations."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(count):
        t += rng.exponential(1.0 / rate)
        s = int(rng.integers(len(systems)))
        if rng.random() < trial_frac:
            k = int(rng.integers(Ks[s]))
            iv, _ = generate_parameters(systems[s], 1, int(rng.integers(1 << 31)))
            out.append(Request(i, "trial", s, t, k, iv))
        else:
            out.append(Request(i, "full", s, t))
    return out


def latency_stats(reqs):
    lat = np.array([r.latency for r in reqs if r.latency is not None])
    span = max(r.done for r in reqs) - min(r.arrival for r in reqs)
    pairs = sum(r.pairs for r in reqs)
    out = {"requests": len(reqs), "p50_ms": float(np.percentile(lat, 50) * 1e3),
           "p99_ms": float(np.percentile(lat, 99) * 1e3), "mean_ms": float(lat.mean() * 1e3),
           "span_s": span, "pairs_per_s": pairs / span}
    for kind in ("trial", "full"):
        lk = np.array([r.latency for r in reqs if r.kind == kind and r.latency is not None])
        if lk.size:
            out[f"{kind}_p50_ms"] = float(np.percentile(lk, 50) * 1e3)
            out[f"{kind}_p99_ms"] = float(np.percentile(lk, 99) * 1e3)
    return out
