"""Device engine: one system on one GPU behind the C ABI of ``libnecg_hip.so``.

PyTorch provides the device memory and the stream; every kernel is the
library's own.  Methods ending in ``_t`` take and return device tensors and
are asynchronous on the current torch stream; the others take NumPy arrays,
synchronise, and raise on any latched device error.
"""

from __future__ import annotations

import ctypes as ct
import threading

import numpy as np

from . import _lib
from .system import System


def _torch():
    import torch
    return torch


class Engine:
    """FP64 energy-and-gradient evaluations of one system on one GPU."""

    def __init__(self, system: System, device=None):
        torch = _torch()
        if not torch.cuda.is_available():
            raise RuntimeError("nbo_correlatedgaussians_amd needs a HIP GPU; none is visible")
        system.validate()
        self.system = system
        self.device = torch.device(device if device is not None else "cuda:0")
        if self.device.type != "cuda":
            raise ValueError("device must be a GPU device")
        self.index = self.device.index if self.device.index is not None else torch.cuda.current_device()
        self.device = torch.device("cuda", self.index)
        self.lib = _lib.load()
        self._host = [np.ascontiguousarray(a, dtype=np.float64) for a in
                      (system.mass, system.ncv, system.scvv, system.sml, system.sms)]
        m, q, c, l, s = self._host
        self._desc = _lib.NecgSystem(system.n, system.nst, float(system.efs), m.ctypes.data,
                                     q.ctypes.data, c.ctypes.data, l.ctypes.data, s.ctypes.data)
        ctx = ct.c_void_p()
        with torch.cuda.device(self.index):
            torch.cuda.init()
            rc = self.lib.necg_create(ct.byref(self._desc), self.index, ct.byref(ctx))
        _lib.raise_for(rc, "necg_create")
        self._ctx = ctx
        self.blk = system.blk
        # pinned host and device buffers of the host-array outprod, reused
        # while x keeps its length
        self._staging = None
        self._stage_lock = threading.Lock()

    # -- plumbing ----------------------------------------------------------
    def close(self):
        if getattr(self, "_ctx", None) and self._ctx.value:
            self.lib.necg_destroy(self._ctx)
            self._ctx = ct.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _stream(self):
        torch = _torch()
        return ct.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def _dev(self, a, dtype=None):
        torch = _torch()
        dtype = dtype or torch.float64
        if isinstance(a, torch.Tensor):
            t = a.to(device=self.device, dtype=dtype)
        else:
            t = torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=self.device)
        return t.contiguous()

    def _empty(self, *shape):
        torch = _torch()
        return torch.empty(*shape, dtype=torch.float64, device=self.device)

    def K_of(self, nx: int) -> int:
        if nx % (self.blk + 1):
            raise ValueError(f"x has length {nx}, not a multiple of blk+1 = {self.blk + 1}")
        return nx // (self.blk + 1)

    def check(self, where: str = "") -> None:
        """Synchronise the current stream and raise on a latched device error."""
        _lib.raise_for(self.lib.necg_status(self._ctx, self._stream()), where)

    def set_tiling(self, ket_tile: int = 0, term_split: int = 0, waves: int = 0) -> None:
        _lib.raise_for(self.lib.necg_set_tiling(self._ctx, ket_tile, term_split, waves),
                       "necg_set_tiling")

    def set_fold(self, fold: bool = True) -> None:
        """Folded diagonal blocks in the lower-triangle passes (default on)."""
        _lib.raise_for(self.lib.necg_set_fold(self._ctx, int(bool(fold))), "necg_set_fold")

    def set_launch_cus(self, ncu: int = 0) -> None:
        """Compute units the pair passes size their launches for (0: the device's)."""
        _lib.raise_for(self.lib.necg_set_launch_cus(self._ctx, ncu), "necg_set_launch_cus")

    def tiling(self):
        a, b, c = ct.c_int32(), ct.c_int32(), ct.c_int32()
        self.lib.necg_get_tiling(self._ctx, ct.byref(a), ct.byref(b), ct.byref(c))
        return a.value, b.value, c.value

    def timing(self, mode: int):
        """Pair-kernel HIP-event timing: 1 start, 0 stop, 2 read ->
        (milliseconds, launches)."""
        ms, n = ct.c_double(), ct.c_int64()
        _lib.raise_for(self.lib.necg_timing(self._ctx, mode, ct.byref(ms), ct.byref(n)),
                       "necg_timing")
        return ms.value, n.value

    @staticmethod
    def _ptr(t):
        return ct.c_void_p(t.data_ptr()) if t is not None else None

    # -- asynchronous tensor API -------------------------------------------
    def outprod_t(self, x, f=None, g=None):
        """(E, g) as device tensors for a device x (python/outprod.py:20)."""
        x = self._dev(x)
        K = self.K_of(x.numel())
        f = self._empty(1) if f is None else f
        g = self._empty(x.numel()) if g is None else g
        rc = self.lib.necg_outprod(self._ctx, K, self._ptr(x), self._ptr(f), self._ptr(g),
                                   self._stream())
        _lib.raise_for(rc, "necg_outprod")
        return f, g

    def energy_t(self, x, with_hs: bool = False):
        x = self._dev(x)
        K = self.K_of(x.numel())
        f = self._empty(1)
        H = self._empty(K, K) if with_hs else None
        S = self._empty(K, K) if with_hs else None
        rc = self.lib.necg_energy(self._ctx, K, self._ptr(x), self._ptr(f), self._ptr(H),
                                  self._ptr(S), self._stream())
        _lib.raise_for(rc, "necg_energy")
        return (f, H, S) if with_hs else f

    def energy_fwd_t(self, x):
        """(E, tape) for x: the gradient's pair partials stay on the context
        as the tape that ``energy_bwd_t`` consumes."""
        x = self._dev(x)
        K = self.K_of(x.numel())
        f = self._empty(1)
        tape = ct.c_int64(0)
        rc = self.lib.necg_energy_fwd(self._ctx, K, self._ptr(x), self._ptr(f), ct.byref(tape),
                                      self._stream())
        _lib.raise_for(rc, "necg_energy_fwd")
        return f, tape.value

    def energy_bwd_t(self, x, tape: int, gbar=None):
        """gbar * dE/dx from ``tape`` (of ``energy_fwd_t`` at x), or None if
        another evaluation ran on this engine since that forward."""
        x = self._dev(x)
        K = self.K_of(x.numel())
        gx = self._empty(x.numel())
        gb = None if gbar is None else self._dev(gbar).reshape(1)
        rc = self.lib.necg_energy_bwd(self._ctx, K, int(tape), self._ptr(gb), self._ptr(gx),
                                      self._stream())
        if rc == _lib.NECG_ERR_STALE:
            return None
        _lib.raise_for(rc, "necg_energy_bwd")
        return gx

    def vjp_t(self, x, gbar=None):
        x = self._dev(x)
        K = self.K_of(x.numel())
        gx = self._empty(x.numel())
        gb = None if gbar is None else self._dev(gbar).reshape(1)
        rc = self.lib.necg_energy_vjp(self._ctx, K, self._ptr(x), self._ptr(gb), self._ptr(gx),
                                      self._stream())
        _lib.raise_for(rc, "necg_energy_vjp")
        return gx

    def matrices_t(self, iv, K: int):
        iv = self._dev(iv)
        if iv.numel() != K * self.blk:
            raise ValueError("iv must hold K*blk numbers")
        H = self._empty(K, K)
        S = self._empty(K, K)
        rc = self.lib.necg_matrices(self._ctx, K, self._ptr(iv), self._ptr(H), self._ptr(S),
                                    self._stream())
        _lib.raise_for(rc, "necg_matrices")
        return H, S

    def kinetic_potential_t(self, x):
        """Kinetic and potential matrices (python/necg.py:36-58) for x."""
        x = self._dev(x)
        K = self.K_of(x.numel())
        T = self._empty(K, K)
        V = self._empty(K, K)
        rc = self.lib.necg_kinetic_potential(self._ctx, K, self._ptr(x), self._ptr(T),
                                             self._ptr(V), self._stream())
        _lib.raise_for(rc, "necg_kinetic_potential")
        return T, V

    def set_pair_mode(self, mode: int) -> None:
        """Form of the gradient pair pass: 0 two-sided where it applies
        (default), 1 two-sided (raises if the terms do not allow it), 2 one
        evaluation per ordered pair."""
        _lib.raise_for(self.lib.necg_set_pair_mode(self._ctx, int(mode)), "necg_set_pair_mode")

    def pair_form(self):
        """(two-sided form available for this system, last gradient pass used it)."""
        a, b = ct.c_int32(), ct.c_int32()
        _lib.raise_for(self.lib.necg_pair_form(self._ctx, ct.byref(a), ct.byref(b)),
                       "necg_pair_form")
        return bool(a.value), bool(b.value)

    def set_dipole_mode(self, mode: int) -> None:
        """0: symmetry-reduced dipole pass when the terms allow it (default);
        1: reduced form required; 2: the full nst^2 double sum."""
        _lib.raise_for(self.lib.necg_set_dipole_mode(self._ctx, int(mode)), "necg_set_dipole_mode")

    def dipole_t(self, x, out=None):
        """The driver's dipole pass (python/necg.py:200-232) for x: a device
        vector [C^T s C, <r_i,a> (3n, particle-major), dipole (3)]."""
        x = self._dev(x)
        K = self.K_of(x.numel())
        n = self.system.n
        out = self._empty(3 * n + 4) if out is None else out
        rc = self.lib.necg_dipole(self._ctx, K, self._ptr(x), self._ptr(out), self._stream())
        _lib.raise_for(rc, "necg_dipole")
        return out

    # -- synchronous NumPy API ---------------------------------------------
    @staticmethod
    def _host_x(x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        if x.ndim != 1:
            raise ValueError("x must be one-dimensional")
        if not np.all(np.isfinite(x)):
            raise FloatingPointError("x contains NaN or Inf")
        return x

    def _stage(self, n: int):
        """(pinned x, pinned g, pinned f, device x, device g, device f) for x
        of n doubles; made again only when n changes."""
        st = self._staging
        if st is None or st[0].numel() != n:
            torch = _torch()
            pin = [torch.empty(m, dtype=torch.float64, pin_memory=True) for m in (n, n, 1)]
            st = (*pin, self._empty(n), self._empty(n), self._empty(1))
            self._staging = st
        return st

    def outprod(self, x):
        """Drop-in for the reference ``outprod(x) -> (f, g)``
        (python/outprod.py:20-76): E = C^T H C / C^T S C and its gradient,
        g a new array.  x and g cross through pinned buffers with copies on
        the engine's stream and one synchronisation per call."""
        x = self._host_x(x)
        with self._stage_lock:
            px, pg, pf, dx, dg, df = self._stage(x.size)
            px.numpy()[:] = x
            dx.copy_(px, non_blocking=True)
            self.outprod_t(dx, df, dg)
            pg.copy_(dg, non_blocking=True)
            pf.copy_(df, non_blocking=True)
            self.check("outprod")            # synchronises the stream
            return float(pf[0]), pg.numpy().copy()

    def energy(self, x, with_hs: bool = False):
        x = self._host_x(x)
        out = self.energy_t(x, with_hs)
        self.check("energy")
        if with_hs:
            f, H, S = out
            return float(f.item()), H.cpu().numpy(), S.cpu().numpy()
        return float(out.item())

    def dipole(self, x):
        """(C^T s C, <r> as (n, 3), dipole (3,)) in atomic units; the
        reference prints the dipole also times 2.541765 (Debye)."""
        x = self._host_x(x)
        out = self.dipole_t(x)
        self.check("dipole")
        o = out.cpu().numpy()
        n = self.system.n
        return float(o[0]), o[1:1 + 3 * n].reshape(n, 3).copy(), o[1 + 3 * n:].copy()

    def matrices(self, iv, K: int):
        iv = self._host_x(iv)
        H, S = self.matrices_t(iv, K)
        self.check("matrices")
        return H.cpu().numpy(), S.cpu().numpy()

    def _queries(self, iv, K, queries):
        torch = _torch()
        q = np.ascontiguousarray(np.asarray(queries, dtype=np.int32).reshape(-1, 3))
        if q.size and (q[:, 0].min() < 0 or q[:, 0].max() >= self.system.nst or
                       q[:, 1:].min() < 0 or q[:, 1:].max() >= K):
            raise IndexError("query (term, bra, ket) out of range")
        iv = self._dev(self._host_x(iv))
        if iv.numel() != K * self.blk:
            raise ValueError("iv must hold K*blk numbers")
        return iv, torch.as_tensor(q, device=self.device), q.shape[0]

    def matel(self, iv, K: int, queries):
        """Rows (ov, tk, ert, zht) for 0-based (term, bra, ket) queries
        (python/matel.py:28)."""
        ivd, qd, nq = self._queries(iv, K, queries)
        out = self._empty(max(nq, 1), 4)
        rc = self.lib.necg_matel(self._ctx, K, self._ptr(ivd), nq, self._ptr(qd), self._ptr(out),
                                 self._stream())
        _lib.raise_for(rc, "necg_matel")
        self.check("matel")
        return out[:nq].cpu().numpy()

    def gradel(self, iv, K: int, queries):
        """(tgvl, tgvls) rows for 0-based (term, bra, ket) queries
        (python/gradel.py:27)."""
        ivd, qd, nq = self._queries(iv, K, queries)
        a = self._empty(max(nq, 1), self.blk)
        b = self._empty(max(nq, 1), self.blk)
        rc = self.lib.necg_gradel(self._ctx, K, self._ptr(ivd), nq, self._ptr(qd), self._ptr(a),
                                  self._ptr(b), self._stream())
        _lib.raise_for(rc, "necg_gradel")
        self.check("gradel")
        return a[:nq].cpu().numpy(), b[:nq].cpu().numpy()

    def dhds(self, iv, K: int):
        """DH, DS of shape (K*blk, K) (python/make_grad.py:34-55)."""
        ivd = self._dev(self._host_x(iv))
        DH = self._empty(K * self.blk, K)
        DS = self._empty(K * self.blk, K)
        rc = self.lib.necg_dhds(self._ctx, K, self._ptr(ivd), self._ptr(DH), self._ptr(DS),
                                self._stream())
        _lib.raise_for(rc, "necg_dhds")
        self.check("dhds")
        return DH.cpu().numpy(), DS.cpu().numpy()


class EngineGraph:
    """One system's evaluation (``outprod``: E and g) captured in a HIP graph:
    write new parameters into ``x`` in place, call ``replay()``, read ``f``
    and ``g``.  The capture follows one warm-up call on a side stream, which
    sizes the engine's workspace, so the captured sequence holds kernel
    launches only.  The graph keeps the addresses of that workspace and reads
    the tile list of its K, so it pins them (``necg_pin_workspace``): until
    ``release()`` (or garbage collection), an engine call at another K, one
    that needs more workspace, or a change of tiling, pair form or launch CU
    count raises instead of moving or rewriting them."""

    def __init__(self, engine: Engine, x0):
        self._pinned = False
        torch = _torch()
        self.engine = engine
        self.x = engine._dev(x0).clone()
        self.f = engine._empty(1)
        self.g = engine._empty(self.x.numel())
        dev = engine.device
        s = torch.cuda.Stream(device=dev)
        s.wait_stream(torch.cuda.current_stream(dev))
        with torch.cuda.stream(s):
            engine.outprod_t(self.x, self.f, self.g)
        torch.cuda.current_stream(dev).wait_stream(s)
        engine.check("graph warm-up")
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            engine.outprod_t(self.x, self.f, self.g)
        _lib.raise_for(engine.lib.necg_pin_workspace(engine._ctx, 1), "necg_pin_workspace")
        self._pinned = True

    def replay(self):
        if self.graph is None:
            raise RuntimeError("the graph was released")
        self.graph.replay()
        return self.f, self.g

    def release(self):
        """Drop the graph and its pin on the engine's workspace."""
        self.graph = None
        if self._pinned:
            self._pinned = False
            self.engine.lib.necg_pin_workspace(self.engine._ctx, 0)

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass
