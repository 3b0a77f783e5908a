"""Many independent systems evaluated together on one GPU.

``Batch`` packs systems of uneven n, nst and K back to back (one x vector,
one g vector) and evaluates all energies and gradients with one fixed
sequence of launches: segmented preparation kernels, one persistent pair
kernel per particle count walking a cost-sorted list of pair tiles, and
segmented finishing kernels.  ``BatchGraph`` captures that sequence in a HIP
graph for replay.  ``autotune`` picks the pair-tile size and the persistent
grid by timing candidates on the batch itself.

Each system's numbers are those of ``outprod`` (python/outprod.py:20) for that
system alone.
"""

from __future__ import annotations

import ctypes as ct
import math
import time

import numpy as np

from . import _lib
from .system import System, generate_parameters, he_system, h2_system, lih_system, pack_x


def _torch():
    import torch
    return torch


class Batch:
    def __init__(self, systems, Ks, device=None):
        torch = _torch()
        if not torch.cuda.is_available():
            raise RuntimeError("nbo_correlatedgaussians_amd needs a HIP GPU; none is visible")
        if len(systems) != len(Ks) or not systems:
            raise ValueError("systems and Ks must be non-empty and of equal length")
        self.systems = list(systems)
        self.Ks = [int(k) for k in Ks]
        self.device = torch.device(device if device is not None else "cuda:0")
        self.index = self.device.index if self.device.index is not None else 0
        self.device = torch.device("cuda", self.index)
        self.lib = _lib.load()
        self._keep = []
        descs = (_lib.NecgSystem * len(systems))()
        for i, s in enumerate(systems):
            s.validate()
            arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in
                    (s.mass, s.ncv, s.scvv, s.sml, s.sms)]
            self._keep.append(arrs)
            m, q, c, l, t = arrs
            descs[i] = _lib.NecgSystem(s.n, s.nst, float(s.efs), m.ctypes.data, q.ctypes.data,
                                       c.ctypes.data, l.ctypes.data, t.ctypes.data)
        self._descs = descs
        Karr = (ct.c_int32 * len(Ks))(*self.Ks)
        h = ct.c_void_p()
        with torch.cuda.device(self.index):
            torch.cuda.init()
            rc = self.lib.necg_batch_create(descs, len(systems), Karr, self.index, ct.byref(h))
        _lib.raise_for(rc, "necg_batch_create")
        self._h = h
        offs = (ct.c_int64 * (len(systems) + 1))()
        self.lib.necg_batch_layout(h, offs)
        self.offsets = np.array(list(offs), dtype=np.int64)
        self.nx = int(self.offsets[-1])
        self.tiling = (0.0, 0)

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            self.lib.necg_batch_destroy(self._h)
            self._h = ct.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        return len(self.systems)

    @property
    def pairs(self) -> int:
        """Unordered basis pairs of one evaluation of every system."""
        return int(sum(K * (K + 1) // 2 for K in self.Ks))

    def _stream(self):
        torch = _torch()
        return ct.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def pack(self, xs):
        """Concatenate per-system x vectors into the batch layout."""
        out = np.empty(self.nx)
        for i, x in enumerate(xs):
            x = np.asarray(x, dtype=np.float64)
            a, b = self.offsets[i], self.offsets[i + 1]
            if x.size != b - a:
                raise ValueError(f"system {i}: x has {x.size} numbers, expected {b - a}")
            out[a:b] = x
        return out

    def unpack(self, g):
        return [g[self.offsets[i]:self.offsets[i + 1]] for i in range(len(self))]

    def set_tiling(self, item_flops: float = 0.0, grid_per_cu: int = 0):
        _lib.raise_for(self.lib.necg_batch_set_tiling(self._h, float(item_flops), int(grid_per_cu)),
                       "necg_batch_set_tiling")
        self.tiling = (float(item_flops), int(grid_per_cu))

    def set_launch_cus(self, ncu: int = 0):
        """Size the persistent pair launches for ncu CUs (0: the device's),
        leaving the others free of pair workgroups while one runs."""
        _lib.raise_for(self.lib.necg_batch_set_launch_cus(self._h, int(ncu)),
                       "necg_batch_set_launch_cus")

    def check(self, where: str = ""):
        _lib.raise_for(self.lib.necg_batch_status(self._h, self._stream()), where)

    def outprod_t(self, x, f=None, g=None, active=None):
        """All energies f (nsys,) and gradients g (batch layout) for a device
        x in the batch layout.  With ``active`` (device int32, one flag per
        system) only the flagged systems are evaluated; the entries of f and
        the segments of g of the others are left as they were.  Asynchronous."""
        torch = _torch()
        if x.numel() != self.nx or x.dtype != torch.float64 or not x.is_contiguous():
            raise ValueError("x must be a contiguous float64 tensor in the batch layout")
        f = torch.empty(len(self), dtype=torch.float64, device=self.device) if f is None else f
        g = torch.empty(self.nx, dtype=torch.float64, device=self.device) if g is None else g
        if active is None:
            rc = self.lib.necg_batch_outprod(self._h, ct.c_void_p(x.data_ptr()),
                                             ct.c_void_p(f.data_ptr()), ct.c_void_p(g.data_ptr()),
                                             self._stream())
        else:
            if (active.numel() != len(self) or active.dtype != torch.int32
                    or active.device != self.device):
                raise ValueError("active must be an int32 device tensor with one flag per system")
            rc = self.lib.necg_batch_outprod_masked(
                self._h, ct.c_void_p(active.data_ptr()), ct.c_void_p(x.data_ptr()),
                ct.c_void_p(f.data_ptr()), ct.c_void_p(g.data_ptr()), self._stream())
        _lib.raise_for(rc, "necg_batch_outprod")
        return f, g

    def pool(self):
        """Handle of a trial pool over this batch's resident state (valid
        after one full evaluation; the batch must outlive it)."""
        h = ct.c_void_p()
        _lib.raise_for(self.lib.necg_batch_pool(self._h, ct.byref(h)), "necg_batch_pool")
        return h

    def outprod(self, xs):
        """[(E_s, g_s)] for a list of per-system x vectors (synchronous)."""
        torch = _torch()
        xa = self.pack(xs)
        if not np.all(np.isfinite(xa)):
            raise FloatingPointError("x contains NaN or Inf")
        xd = torch.as_tensor(xa, device=self.device)
        f, g = self.outprod_t(xd)
        self.check("batch outprod")
        fh = f.cpu().numpy()
        return [(float(fh[i]), gi.copy()) for i, gi in enumerate(self.unpack(g.cpu().numpy()))]

    def matrices(self, s: int):
        torch = _torch()
        K = self.Ks[s]
        H = torch.empty(K, K, dtype=torch.float64, device=self.device)
        S = torch.empty(K, K, dtype=torch.float64, device=self.device)
        _lib.raise_for(self.lib.necg_batch_matrices(self._h, s, ct.c_void_p(H.data_ptr()),
                                                    ct.c_void_p(S.data_ptr()), self._stream()),
                       "necg_batch_matrices")
        self.check("batch matrices")
        return H.cpu().numpy(), S.cpu().numpy()


class BatchGraph:
    """The batch evaluation captured in a HIP graph: write new parameters into
    ``x`` (in place), call ``replay()``, read ``f`` and ``g``.  The graph holds
    the addresses of the batch's work lists and partial buffers, so it pins
    them (``necg_batch_pin``): until ``release()`` (or garbage collection),
    ``batch.set_tiling`` and ``autotune`` raise instead of replacing them."""

    def __init__(self, batch: Batch, x0):
        self._pinned = False
        torch = _torch()
        self.batch = batch
        self.x = torch.as_tensor(np.asarray(x0, dtype=np.float64), device=batch.device).clone()
        self.f = torch.empty(len(batch), dtype=torch.float64, device=batch.device)
        self.g = torch.empty(batch.nx, dtype=torch.float64, device=batch.device)
        s = torch.cuda.Stream(device=batch.device)
        s.wait_stream(torch.cuda.current_stream(batch.device))
        with torch.cuda.stream(s):  # warm-up outside the capture
            batch.outprod_t(self.x, self.f, self.g)
        torch.cuda.current_stream(batch.device).wait_stream(s)
        batch.check("graph warm-up")
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            batch.outprod_t(self.x, self.f, self.g)
        _lib.raise_for(batch.lib.necg_batch_pin(batch._h, 1), "necg_batch_pin")
        self._pinned = True

    def replay(self):
        if self.graph is None:
            raise RuntimeError("the graph was released")
        self.graph.replay()
        return self.f, self.g

    def release(self):
        """Drop the graph and its pin on the batch's work lists."""
        self.graph = None
        if self._pinned:
            self._pinned = False
            self.batch.lib.necg_batch_pin(self.batch._h, 0)

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


def autotune(batch: Batch, x, item_flops=(1.5e7, 3e7, 6e7, 1.2e8, 2.4e8), grid_per_cu=(0,),
             reps: int = 3):
    """Time the batch evaluation for each (pair-tile cost, persistent grid)
    candidate and keep the fastest.  Returns (best, [(candidate, ms)])."""
    torch = _torch()
    xd = torch.as_tensor(np.asarray(x, dtype=np.float64), device=batch.device)
    f = torch.empty(len(batch), dtype=torch.float64, device=batch.device)
    g = torch.empty(batch.nx, dtype=torch.float64, device=batch.device)
    results = []
    for it in item_flops:
        for gp in grid_per_cu:
            batch.set_tiling(it, gp)
            batch.outprod_t(xd, f, g)
            batch.check("autotune warm-up")
            torch.cuda.synchronize(batch.device)
            t0 = time.perf_counter()
            for _ in range(reps):
                batch.outprod_t(xd, f, g)
            torch.cuda.synchronize(batch.device)
            results.append(((it, gp), (time.perf_counter() - t0) * 1e3 / reps))
    best = min(results, key=lambda r: r[1])[0]
    batch.set_tiling(*best)
    return best, results


SYSTEM_FACTORIES = {"He": he_system, "H2": h2_system, "LiH": lih_system}


def mixed_systems(count: int = 64, seed0: int = 100, kmin: int = 32, kmax: int = 1024):
    """Independent systems drawn uniformly from He (n=2), H2 (n=3) and LiH
    (n=5) with K log-uniform in [kmin, kmax]; system i uses seed seed0 + i
    for its type, its K and its parameters.  Returns (systems, Ks, xs)."""
    names = ("He", "H2", "LiH")
    cache = {}
    systems, Ks, xs = [], [], []
    for i in range(count):
        seed = seed0 + i
        rng = np.random.default_rng(seed)
        name = names[int(rng.integers(3))]
        K = int(round(math.exp(rng.uniform(math.log(kmin), math.log(kmax)))))
        if name not in cache:
            cache[name] = SYSTEM_FACTORIES[name]()
        s = cache[name]
        iv, C = generate_parameters(s, K, seed)
        systems.append(s)
        Ks.append(K)
        xs.append(pack_x(iv, C))
    return systems, Ks, xs
