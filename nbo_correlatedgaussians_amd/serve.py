"""Serving: energy and gradient requests against systems resident in HBM.

Each resident system keeps, on the device, its parameters and the state of
its last evaluation (bra/ket tables, H, S, H C, S C, C^T H C, C^T S C).  Two
request kinds arrive over time:

* ``trial``: replace basis function k of system s by new (L_k, s_k); the new
  row/column of H and S and the trial energy C^T H' C / C^T S' C (resident C)
  are computed; the resident state is not modified.
* ``full``: energy and analytic gradient of system s at its resident x
  (``outprod``, python/outprod.py:20).

The scheduler does continuous batching: on every pass it collects the trials
that have arrived, admits as many as the free memory budget allows into one
batched trial launch on a free trial stream, and starts full evaluations on
free evaluation streams; launches on different streams overlap on the device
and complete independently, and a new pass admits whatever arrived meanwhile.
A system with a full evaluation in flight takes no new work until it ends.
Latency is measured from a request's arrival to its results being on the
host.
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
    kind: str                 # "trial" or "full"
    sys: int
    arrival: float            # seconds from the start of the run
    func: int = -1            # trial: replaced function (0-based)
    newp: np.ndarray = None   # trial: new [vech L | s] block
    done: float = None
    E: float = None
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
    """Resident systems on one GPU with continuous batching of requests."""

    def __init__(self, systems, xs, device=None, mem_budget: int = 4 << 30,
                 max_trials: int = 512, trial_streams: int = 2, eval_streams: int = 2):
        torch = _torch()
        self.device = torch.device(device if device is not None else "cuda:0")
        self.systems = list(systems)
        self.engines = []
        self.x = []
        for s, x in zip(systems, xs):
            eng = Engine(s, self.device)
            xd = torch.as_tensor(np.asarray(x, dtype=np.float64), device=self.device)
            eng.outprod_t(xd)           # resident state of x
            self.engines.append(eng)
            self.x.append(xd)
        torch.cuda.synchronize(self.device)
        for e in self.engines:
            e.check("initial evaluation")
        self.Ks = [e.K_of(x.numel()) for e, x in zip(self.engines, self.x)]
        self.lib = _lib.load()
        ctxs = (ct.c_void_p * len(self.engines))(*[e._ctx.value for e in self.engines])
        h = ct.c_void_p()
        _lib.raise_for(self.lib.necg_pool_create(ctxs, len(self.engines), ct.byref(h)),
                       "necg_pool_create")
        self._pool = h
        kmax, wst = ct.c_int32(), ct.c_int32()
        self.lib.necg_pool_sizes(h, ct.byref(kmax), ct.byref(wst))
        self.kmax, self.wstride = kmax.value, wst.value
        self.pmax = max(s.blk for s in systems)
        self.max_trials = max_trials
        self.mem_budget = mem_budget
        f64 = dict(dtype=torch.float64, device=self.device)
        i32 = dict(dtype=torch.int32, device=self.device)
        self._tbuf = []
        for _ in range(trial_streams):
            T = max_trials
            self._tbuf.append(dict(
                stream=torch.cuda.Stream(device=self.device),
                tsys=torch.empty(T, **i32), tfun=torch.empty(T, **i32),
                newp=torch.empty(T * self.pmax, **f64), ws=torch.empty(T * self.wstride, **f64),
                E=torch.empty(T, **f64), rowH=torch.empty(T * self.kmax, **f64),
                rowS=torch.empty(T * self.kmax, **f64),
                h_in=torch.empty(T * (2 + self.pmax), dtype=torch.float64).pin_memory(),
                h_E=torch.empty(T, dtype=torch.float64).pin_memory()))
        xmax = max(x.numel() for x in self.x)
        self._ebuf = []
        for _ in range(eval_streams):
            self._ebuf.append(dict(
                stream=torch.cuda.Stream(device=self.device),
                f=torch.empty(1, **f64), g=torch.empty(xmax, **f64),
                h_f=torch.empty(1, dtype=torch.float64).pin_memory(),
                h_g=torch.empty(xmax, dtype=torch.float64).pin_memory()))
        # bytes a request holds while in flight (admission budget)
        self.trial_bytes = 8 * (self.pmax + self.wstride + 1 + 2 * self.kmax)
        self.full_bytes = [self._eval_bytes(e, K) for e, K in zip(self.engines, self.Ks)]

    @staticmethod
    def _eval_bytes(eng, K):
        """Device bytes one evaluation of this system works through: H, S and
        the element partials (4 K^2), the ket table (K nst KS), and the
        gradient partials (about K/16 tiles of NACC x K)."""
        n, nst = eng.system.n, eng.system.nst
        nacc = n * (n + 1) + 6 * n + 2
        ks = n * (n + 1) + 6 * n + n * n + 6
        return 8 * (4 * K * K + K * nst * ks + max(1, K // 16) * nacc * K)

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
        torch = _torch()
        b = self._tbuf[slot]
        n = len(reqs)
        hin = b["h_in"].numpy()
        hin[:n] = [r.sys for r in reqs]
        hin[n:2 * n] = [r.func for r in reqs]
        P = self.pmax
        for i, r in enumerate(reqs):
            hin[2 * n + i * P: 2 * n + i * P + r.newp.size] = r.newp
        with torch.cuda.stream(b["stream"]):
            staged = b["h_in"][:2 * n + n * P].to(self.device, non_blocking=True)
            b["tsys"][:n].copy_(staged[:n].to(torch.int32))
            b["tfun"][:n].copy_(staged[n:2 * n].to(torch.int32))
            b["newp"][:n * P].copy_(staged[2 * n:])
            st = ct.c_void_p(b["stream"].cuda_stream)
            rc = self.lib.necg_pool_trials(
                self._pool, n, ct.c_void_p(b["tsys"].data_ptr()), ct.c_void_p(b["tfun"].data_ptr()),
                ct.c_void_p(b["newp"].data_ptr()), P, ct.c_void_p(b["ws"].data_ptr()),
                ct.c_void_p(b["E"].data_ptr()), ct.c_void_p(b["rowH"].data_ptr()),
                ct.c_void_p(b["rowS"].data_ptr()), self.kmax, st)
            _lib.raise_for(rc, "necg_pool_trials")
            b["h_E"][:n].copy_(b["E"][:n], non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(b["stream"])
        return _Launch("trial", slot, ev, reqs, self.trial_bytes * n, {r.sys for r in reqs})

    def _launch_full(self, slot, r):
        torch = _torch()
        b = self._ebuf[slot]
        eng, x = self.engines[r.sys], self.x[r.sys]
        nx = x.numel()
        with torch.cuda.stream(b["stream"]):
            eng.outprod_t(x, b["f"], b["g"][:nx])
            b["h_f"].copy_(b["f"], non_blocking=True)
            b["h_g"][:nx].copy_(b["g"][:nx], non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(b["stream"])
        return _Launch("full", slot, ev, [r], self.full_bytes[r.sys], {r.sys})

    # -- the loop ---------------------------------------------------------
    def run(self, requests, on_pass=None):
        """Serve the requests (sorted by arrival, times in seconds from now);
        fills in done/E/g of each and returns them."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        for r in reqs:
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
                    E = b["h_E"].numpy()
                    for i, r in enumerate(L.reqs):
                        r.E = float(E[i])
                        r.done = tdone
                        trial_busy[r.sys] -= 1
                    free_t.append(L.slot)
                else:
                    b = self._ebuf[L.slot]
                    r = L.reqs[0]
                    r.E = float(b["h_f"][0])
                    r.g = b["h_g"][:self.x[r.sys].numel()].numpy().copy()
                    r.done = tdone
                    full_busy.discard(r.sys)
                    free_e.append(L.slot)
                used -= L.nbytes
                finished += len(L.reqs)
            inflight = still
            # admit trials: one batched launch per free trial stream
            while free_t and ready_t:
                take = []
                keep = deque()
                while ready_t and len(take) < self.max_trials:
                    r = ready_t.popleft()
                    if r.sys in full_busy:          # wait for that evaluation
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
            # admit full evaluations
            skipped = deque()
            while free_e and ready_f:
                r = ready_f.popleft()
                if (r.sys in full_busy or trial_busy.get(r.sys, 0) > 0 or
                        used + self.full_bytes[r.sys] > self.mem_budget):
                    skipped.append(r)
                    continue
                full_busy.add(r.sys)
                L = self._launch_full(free_e.pop(), r)
                used += L.nbytes
                inflight.append(L)
            ready_f.extendleft(reversed(skipped))
            if on_pass:
                on_pass(len(inflight), len(ready_t) + len(ready_f))
            if not inflight and ptr < total and not ready_t and not ready_f:
                wait = reqs[ptr].arrival - (time.perf_counter() - t0)
                if wait > 0:
                    time.sleep(min(wait, 1e-3))
        for e in self.engines:
            e.check("serving")
        return reqs

    def run_plain(self, requests):
        """The same requests one at a time, each launched alone and waited
        for (no batching, no overlap): the per-request GPU path that the
        scheduler is compared with.  Returns (seconds, pairs)."""
        torch = _torch()
        t0 = time.perf_counter()
        pairs = 0
        for r in requests:
            K = self.Ks[r.sys]
            if r.kind == "trial":
                L = self._launch_trials(0, [Request(r.rid, "trial", r.sys, 0.0, r.func, r.newp)])
                pairs += K
            else:
                L = self._launch_full(0, Request(r.rid, "full", r.sys, 0.0))
                pairs += K * (K + 1) // 2
            L.event.synchronize()
        torch.cuda.synchronize(self.device)
        return time.perf_counter() - t0, pairs

    def trial_alone(self, r: Request):
        """One trial launched by itself (reference for the in-mix result)."""
        torch = _torch()
        q = Request(r.rid, "trial", r.sys, 0.0, r.func, r.newp)
        L = self._launch_trials(0, [q])
        L.event.synchronize()
        _lib.raise_for(self.lib.necg_pool_status(self._pool, ct.c_void_p(
            self._tbuf[0]["stream"].cuda_stream)), "trial")
        return float(self._tbuf[0]["h_E"][0]), (
            self._tbuf[0]["rowH"][:self.Ks[r.sys]].cpu().numpy(),
            self._tbuf[0]["rowS"][:self.Ks[r.sys]].cpu().numpy())


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
