"""Module-level drop-in for the reference's Python surface.

The reference keeps its state in module globals filled by ``read_input``
(python/read_input.py:16, python/shared.py) and exposes ``outprod(x)``
(python/outprod.py:20), ``matel(stf, ii, jj)`` (python/matel.py:28) and
``gradel(stf, ii, jj)`` (python/gradel.py:27) with 1-based indices, and runs
the dipole pass of its driver (``dipole()``, python/necg.py:200-232).  This
module mirrors that: after ``read_input(path)`` (or ``use(system, iv, C)``),
``outprod`` can be handed to ``scipy.optimize.minimize(..., jac=True)``
unchanged.  The arithmetic runs on the GPU through ``libnecg_hip.so``.
"""

from __future__ import annotations

import numpy as np

from . import system as _system
from .engine import Engine

# module state, as the reference's ``shared`` module
sys_desc = None
M = 0
iv = None
C = None
H = None
S = None
grad = None
_engine = None


def use(sysd, iv0, C0, device=None) -> Engine:
    """Install a system and starting parameters as the module state."""
    global sys_desc, M, iv, C, H, S, grad, _engine
    sys_desc = sysd
    M = len(C0)
    iv = np.array(iv0, dtype=np.float64)
    C = np.array(C0, dtype=np.float64)
    H = S = grad = None
    if _engine is None or _engine.system is not sysd:
        if _engine is not None:
            _engine.close()
        _engine = Engine(sysd, device)
    return _engine


def read_input(filename: str = "oldquit.dat", device=None) -> Engine:
    """Read an input file in the reference layout and install it."""
    sysd, _, iv0, C0 = _system.read_input(filename)
    return use(sysd, iv0, C0, device)


def engine() -> Engine:
    if _engine is None:
        raise RuntimeError("no system installed: call read_input() or use() first")
    return _engine


def x0() -> np.ndarray:
    """The current parameter vector [iv | C]."""
    return _system.pack_x(iv, C)


def outprod(x):
    """E = C^T H C / C^T S C and dE/dx for x = [iv | C]
    (python/outprod.py:20-76).  Returns (float, fresh ndarray)."""
    global iv, C, grad
    eng = engine()
    x = np.ascontiguousarray(x, dtype=np.float64)
    siv = M * sys_desc.blk
    if x.size != siv + M:
        raise ValueError(f"x has length {x.size}, expected {siv + M}")
    iv = x[:siv].copy()
    C = x[siv:].copy()
    f, g = eng.outprod(x)
    grad = g
    return f, g.copy()


def build_matrices():
    """H and S of the current state (python/necg.py:36-58)."""
    global H, S
    H, S = engine().matrices(iv, M)
    return H, S


def matel(stf: int, ii: int, jj: int):
    """(ov, tk, ert, zht) of term stf between bra ii and ket jj, 1-based."""
    r = engine().matel(iv, M, [(stf - 1, ii - 1, jj - 1)])[0]
    return float(r[0]), float(r[1]), float(r[2]), float(r[3])


def gradel(stf: int, ii: int, jj: int):
    """(tgvl, tgvls) of term stf between bra ii and ket jj, 1-based."""
    a, b = engine().gradel(iv, M, [(stf - 1, ii - 1, jj - 1)])
    return a[0], b[0]


DEBYE = 2.541765   # a.u. of dipole moment in Debye (python/necg.py:234)


def dipole():
    """(d7, d8, d9) of the current state in a.u. and in Debye, with the
    centres <r_i> (n, 3) of the pseudo-particles (python/necg.py:200-244)."""
    css, rbar, d = engine().dipole(x0())
    return {"dipole_au": d, "dipole_debye": d * DEBYE, "centres": rbar, "cSc": css}
