"""ctypes front end of ``liboracle.so`` (built from ``necg_oracle.c``).

TEST INFRASTRUCTURE ONLY: the checker for the HIP path.  Index arguments are
0-based here; the reference's matel/gradel take 1-based indices
(python/matel.py:40-43).
"""

from __future__ import annotations

import ctypes as ct
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB):
        subprocess.run(["make", "-B" if force else "-s", "-C", HERE, "liboracle.so"],
                       check=True, stdout=subprocess.DEVNULL)
    return LIB


class _OSys(ct.Structure):
    _fields_ = [("n", ct.c_int), ("nst", ct.c_int), ("efs", ct.c_double),
                ("mass", ct.c_void_p), ("ncv", ct.c_void_p), ("scvv", ct.c_void_p),
                ("sml", ct.c_void_p), ("sms", ct.c_void_p)]


_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = ct.CDLL(LIB)
        P = ct.c_void_p
        _lib.necg_oracle_f0.restype = ct.c_double
        _lib.necg_oracle_f0.argtypes = [ct.c_double]
        _lib.necg_oracle_df0.restype = ct.c_double
        _lib.necg_oracle_df0.argtypes = [ct.c_double]
        _lib.necg_oracle_matel.argtypes = [P, P, ct.c_int, ct.c_int, ct.c_int, P]
        _lib.necg_oracle_gradel.argtypes = [P, P, ct.c_int, ct.c_int, ct.c_int, P, P]
        _lib.necg_oracle_matrices.argtypes = [P, ct.c_int, P, P, P, ct.c_int]
        _lib.necg_oracle_outprod.argtypes = [P, ct.c_int, P, P, P, P, P, ct.c_int]
        _lib.necg_oracle_outprod_ld.argtypes = [P, ct.c_int, P, P, P, P, ct.c_int]
        _lib.necg_oracle_grada_rows.argtypes = [P, ct.c_int, P, ct.c_double, ct.c_double,
                                                ct.c_int, ct.c_int, P, ct.c_int]
        _lib.necg_oracle_pair_sums.argtypes = [P, P, ct.c_int, P, P, P, P, P]
        _lib.necg_oracle_rrms.argtypes = [P, P, ct.c_int, ct.c_int, ct.c_int, ct.c_int, P, P]
        _lib.necg_oracle_dipole.argtypes = [P, ct.c_int, P, P, P, P, ct.c_int]
        _lib.necg_oracle_matel_rows.argtypes = [P, ct.c_int, P, ct.c_int, ct.c_int, P, P,
                                                ct.c_int]
    return _lib


def _p(a):
    return a.ctypes.data_as(ct.c_void_p)


class Oracle:
    """Reference-structured FP64 CPU evaluation for one system."""

    def __init__(self, sysd):
        self.sys = sysd
        self._keep = [np.ascontiguousarray(a, dtype=np.float64) for a in
                      (sysd.mass, sysd.ncv, sysd.scvv, sysd.sml, sysd.sms)]
        m, q, c, l, s = self._keep
        self._o = _OSys(sysd.n, sysd.nst, float(sysd.efs), m.ctypes.data, q.ctypes.data,
                        c.ctypes.data, l.ctypes.data, s.ctypes.data)
        self.lib = _load()

    @property
    def _po(self):
        return ct.byref(self._o)

    def f0(self, z):
        return self.lib.necg_oracle_f0(float(z))

    def df0(self, z):
        return self.lib.necg_oracle_df0(float(z))

    def matel(self, iv, stf, i, j):
        """(ov, tk, ert, zht) for term stf, bra i, ket j (0-based)."""
        iv = np.ascontiguousarray(iv, dtype=np.float64)
        out = np.zeros(4)
        if not self.lib.necg_oracle_matel(self._po, _p(iv), stf, i, j, _p(out)):
            raise np.linalg.LinAlgError("singular A_i + A_j")
        return out

    def gradel(self, iv, stf, i, j):
        """(tgvl, tgvls) for term stf, bra i, ket j (0-based)."""
        iv = np.ascontiguousarray(iv, dtype=np.float64)
        a = np.zeros(self.sys.blk)
        b = np.zeros(self.sys.blk)
        if not self.lib.necg_oracle_gradel(self._po, _p(iv), stf, i, j, _p(a), _p(b)):
            raise np.linalg.LinAlgError("singular matrix in gradel")
        return a, b

    def matrices(self, iv, K, nthreads=0):
        iv = np.ascontiguousarray(iv, dtype=np.float64)
        H = np.zeros((K, K))
        S = np.zeros((K, K))
        bad = self.lib.necg_oracle_matrices(self._po, K, _p(iv), _p(H), _p(S), nthreads)
        if bad:
            raise np.linalg.LinAlgError("singular pair matrix")
        return H, S

    def outprod(self, x, nthreads=0, return_hs=False):
        x = np.ascontiguousarray(x, dtype=np.float64)
        K = x.size // (self.sys.blk + 1)
        f = np.zeros(1)
        g = np.zeros(x.size)
        H = np.zeros((K, K))
        S = np.zeros((K, K))
        rc = self.lib.necg_oracle_outprod(self._po, K, _p(x), _p(f), _p(g), _p(H), _p(S),
                                          nthreads)
        if rc == 1:
            raise np.linalg.LinAlgError("singular pair matrix")
        if rc == 2:
            raise ZeroDivisionError("C^T S C == 0")
        if return_hs:
            return float(f[0]), g, H, S
        return float(f[0]), g

    def outprod_ld(self, x, nthreads=0):
        """(E, g) with double sums as ``outprod`` and (E, g) with every sum in
        long double, from one pass over the pair terms."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        K = x.size // (self.sys.blk + 1)
        f = np.zeros(2)
        g = np.zeros(x.size)
        gl = np.zeros(x.size)
        rc = self.lib.necg_oracle_outprod_ld(self._po, K, _p(x), _p(f), _p(g), _p(gl), nthreads)
        if rc == 1:
            raise np.linalg.LinAlgError("singular pair matrix")
        if rc == 2:
            raise ZeroDivisionError("C^T S C == 0")
        return (float(f[0]), g), (float(f[1]), gl)

    def grada_rows(self, x, E, norm, row0, nrows, nthreads=0):
        x = np.ascontiguousarray(x, dtype=np.float64)
        K = x.size // (self.sys.blk + 1)
        out = np.zeros(nrows * self.sys.blk)
        bad = self.lib.necg_oracle_grada_rows(self._po, K, _p(x), float(E), float(norm),
                                              row0, nrows, _p(out), nthreads)
        if bad:
            raise np.linalg.LinAlgError("singular pair matrix")
        return out

    def matel_rows(self, iv, K, row0, nrows, nthreads=0):
        """Rows [row0, row0+nrows) of the lower triangle of H and S."""
        iv = np.ascontiguousarray(iv, dtype=np.float64)
        H = np.zeros((nrows, K))
        S = np.zeros((nrows, K))
        if self.lib.necg_oracle_matel_rows(self._po, K, _p(iv), row0, nrows, _p(H), _p(S),
                                           nthreads):
            raise np.linalg.LinAlgError("singular pair matrix")
        return H, S

    def pair_sums(self, iv, pairs):
        """H_ij, S_ij (lower-triangle element) and the DH/DS blocks of the
        ordered pairs (bra i, ket j), 0-based."""
        iv = np.ascontiguousarray(iv, dtype=np.float64)
        pr = np.ascontiguousarray(np.asarray(pairs).reshape(-1, 2), dtype=np.int32)
        m = pr.shape[0]
        H = np.zeros(m)
        S = np.zeros(m)
        dh = np.zeros((m, self.sys.blk))
        ds = np.zeros((m, self.sys.blk))
        if not self.lib.necg_oracle_pair_sums(self._po, _p(iv), m, _p(pr), _p(H), _p(S),
                                              _p(dh), _p(ds)):
            raise np.linalg.LinAlgError("singular pair matrix")
        return H, S, dh, ds

    def rrms(self, iv, stfe, stf, i, j):
        """(ov, ov * st) of bra i under term stfe and ket j under term stf,
        0-based (python/rrms.py:21-120, all 3n components of st)."""
        iv = np.ascontiguousarray(iv, dtype=np.float64)
        ov = np.zeros(1)
        tk = np.zeros(3 * self.sys.n)
        if not self.lib.necg_oracle_rrms(self._po, _p(iv), stfe, stf, i, j, _p(ov), _p(tk)):
            raise np.linalg.LinAlgError("singular pair matrix")
        return float(ov[0]), tk

    def dipole(self, x, nthreads=0):
        """(C^T s C, rbar (n, 3), dipole (3,)) of the driver's dipole pass
        (python/necg.py:200-232)."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        K = x.size // (self.sys.blk + 1)
        css = np.zeros(1)
        rbar = np.zeros(3 * self.sys.n)
        dip = np.zeros(3)
        rc = self.lib.necg_oracle_dipole(self._po, K, _p(x), _p(css), _p(rbar), _p(dip), nthreads)
        if rc == 1:
            raise np.linalg.LinAlgError("singular pair matrix")
        if rc == 2:
            raise ZeroDivisionError("C^T s C == 0")
        return float(css[0]), rbar.reshape(self.sys.n, 3), dip
