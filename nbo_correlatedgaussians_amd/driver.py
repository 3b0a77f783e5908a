"""The steps around the hot path that the reference's driver runs
(python/necg.py:88-244), on the GPU engine:

* ``lowest_eigenpair``: the generalised eigenproblem H v = lambda S v that
  initialises C (python/necg.py:129-134, scipy.linalg.eigh; DSYGV at
  source/necg.f90:175), solved on the device in FP64 by a Cholesky
  factorisation of S and a symmetric eigensolver (torch.linalg on ROCm).
* ``energetics``: Rayleigh quotient, average kinetic and potential energy and
  the virial coefficient -2 <T>/<V> (python/necg.py:61-71).
* ``optimize``: scipy.optimize.minimize over the GPU ``outprod`` with the
  reference's L-BFGS-B options (python/necg.py:159-171), or a truncated-Newton
  run with the LMQN constants of source/necg.f90:274-280 through scipy's TNC
  (maxCGit = 20, eta = 0.21, stepmx = 7, accuracy = 2.22e-16,
  maxfun = 150 (siv + M)), stopped after a given number of outer iterations.
* ``properties``: the dipole pass (python/necg.py:200-244): centres of the
  pseudo-particles and the dipole moment in a.u. and Debye.
* ``run``: the whole sequence for one input (matrices, eigenvector start,
  energetics, optimisation, final energetics and dipole).
"""

from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from .engine import Engine
from .system import System, pack_x, unpack_x

EPS = 2.220446649250313e-16


def _torch():
    import torch
    return torch


def generalized_eigh(H, S):
    """All eigenpairs of H v = lambda S v (symmetric H, positive definite S;
    device tensors or arrays): (lam ascending, V) with V^T S V = I and
    V^T H V = diag(lam), by a Cholesky factorisation S = L L^T and the
    symmetric eigensolver on L^-1 H L^-T, in FP64 on the device."""
    torch = _torch()
    dev = H.device if isinstance(H, torch.Tensor) else torch.device("cuda:0")
    H = torch.as_tensor(H, dtype=torch.float64, device=dev)
    S = torch.as_tensor(S, dtype=torch.float64, device=dev)
    L = torch.linalg.cholesky(S)
    Y = torch.linalg.solve_triangular(L, H, upper=False)              # L^-1 H
    A = torch.linalg.solve_triangular(L, Y.T.contiguous(), upper=False)  # L^-1 H L^-T
    A = 0.5 * (A + A.T)
    w, U = torch.linalg.eigh(A)
    V = torch.linalg.solve_triangular(L.T.contiguous(), U.contiguous(), upper=True)
    return w.contiguous(), V.contiguous()


def lowest_eigenpair(H, S):
    """(lambda_min, v) of H v = lambda S v for symmetric H and positive
    definite S (device tensors or arrays), v normalised so v^T S v = 1."""
    w, V = generalized_eigh(H, S)
    return float(w[0].item()), V[:, 0].contiguous()


def energetics(engine: Engine, x):
    """(E, <T>, <V>, virial) for x (python/necg.py:61-71)."""
    torch = _torch()
    xd = torch.as_tensor(np.asarray(x, dtype=np.float64), device=engine.device)
    K = engine.K_of(xd.numel())
    C = xd[K * engine.blk:]
    f, H, S = engine.energy_t(xd, with_hs=True)
    T, V = engine.kinetic_potential_t(xd)
    engine.check("energetics")
    cSc = float(C @ (S @ C))
    E = float(C @ (H @ C)) / cSc
    kin = float(C @ (T @ C)) / cSc
    pot = float(C @ (V @ C)) / cSc
    return {"energy": E, "kinetic": kin, "potential": pot, "virial": -2.0 * kin / pot}


DEBYE = 2.541765   # python/necg.py:234


def properties(engine: Engine, x):
    """Dipole pass for x: {'dipole_au', 'dipole_debye', 'centres', 'cSc'}."""
    css, rbar, d = engine.dipole(x)
    return {"dipole_au": d, "dipole_debye": d * DEBYE, "centres": rbar, "cSc": css}


class _Stop(Exception):
    pass


@dataclass
class OptResult:
    x: np.ndarray
    fun: float
    nit: int
    nfev: int
    message: str
    energies: list = field(default_factory=list)   # energy after each outer iteration
    seconds: float = 0.0


def write_punch(path: str, x) -> None:
    """The restart punch file of the truncated-Newton driver: x, one value
    per line, rewritten after every outer iteration (source/tn.f:338-347);
    read_punch reads it back."""
    with open(path + ".tmp", "w", encoding="utf-8") as fh:
        for v in np.asarray(x, dtype=np.float64):
            fh.write(f"{float(v)!r}\n")
    import os
    os.replace(path + ".tmp", path)


def read_punch(path: str) -> np.ndarray:
    with open(path, "r", encoding="utf-8") as fh:
        return np.array([float(t.replace("D", "E").replace("d", "e")) for t in fh.read().split()])


def optimize(fun, x0, method: str = "L-BFGS-B", maxiter: int = 20, siv: int = None,
             M: int = None, options: dict = None, punch: str = None):
    """Minimise E(x) with the analytic gradient of ``fun(x) -> (E, g)``.

    method "L-BFGS-B": the reference's call (python/necg.py:159-171).
    method "TN": truncated Newton with the LMQN constants, through scipy's TNC,
    stopped after ``maxiter`` outer iterations.
    punch: a path to write x to after every outer iteration (TNRESTRT.PT).
    """
    from scipy.optimize import minimize
    x0 = np.asarray(x0, dtype=np.float64)
    siv = siv if siv is not None else x0.size
    M = M if M is not None else 0
    last = {"x": x0.copy(), "f": None, "nfev": 0}

    def f(x):
        E, g = fun(x)
        last["nfev"] += 1
        return E, g

    energies = []
    t0 = time.perf_counter()
    if method.upper() == "L-BFGS-B":
        opts = {"maxiter": maxiter, "maxfun": 150 * (siv + M), "ftol": EPS, "gtol": np.sqrt(EPS)}
        opts.update(options or {})
        def cb_l(xk):
            energies.append(fun(xk)[0])
            if punch:
                write_punch(punch, xk)

        res = minimize(f, x0, method="L-BFGS-B", jac=True, options=opts, callback=cb_l)
        return OptResult(res.x, float(res.fun), int(res.nit), int(res.nfev), str(res.message),
                         energies, time.perf_counter() - t0)
    if method.upper() in ("TN", "TNC"):
        opts = {"maxCGit": 20, "eta": 0.21, "stepmx": 7.0, "accuracy": EPS,
                "maxfun": 150 * (siv + M), "ftol": 0.0, "xtol": np.sqrt(EPS), "gtol": 0.0}
        opts.update(options or {})
        state = {"nit": 0}

        def cb(xk):
            state["nit"] += 1
            E = fun(xk)[0]
            energies.append(E)
            last["x"], last["f"] = xk.copy(), E
            if punch:
                write_punch(punch, xk)
            if state["nit"] >= maxiter:
                raise _Stop

        try:
            res = minimize(f, x0, method="TNC", jac=True, options=opts, callback=cb)
            return OptResult(res.x, float(res.fun), state["nit"], int(res.nfev),
                             str(res.message), energies, time.perf_counter() - t0)
        except _Stop:
            return OptResult(last["x"], float(last["f"]), state["nit"], last["nfev"],
                             f"stopped after {maxiter} outer iterations", energies,
                             time.perf_counter() - t0)
    raise ValueError(f"unknown method {method!r}")


def run(system: System, iv, C=None, device=None, method: str = "L-BFGS-B", maxiter: int = 20,
        init_from_eigenvector: bool = True):
    """The reference driver's sequence for one system (python/necg.py:88-195):
    matrices, C from the lowest generalised eigenvector, energetics,
    optimisation of x = [iv | C], final energetics.  Returns a dict."""
    torch = _torch()
    eng = Engine(system, device)
    iv = np.asarray(iv, dtype=np.float64)
    K = iv.size // system.blk
    C0 = np.ones(K) / K if C is None else np.asarray(C, dtype=np.float64)
    out = {}
    if init_from_eigenvector:
        H, S = eng.matrices_t(iv, K)
        lam, v = lowest_eigenpair(H, S)
        out["eigenvalue"] = lam
        C0 = v.cpu().numpy()
    x0 = pack_x(iv, C0)
    out["initial"] = energetics(eng, x0)
    res = optimize(eng.outprod, x0, method=method, maxiter=maxiter, siv=K * system.blk, M=K)
    out["optimisation"] = {"energy": res.fun, "nit": res.nit, "nfev": res.nfev,
                           "message": res.message, "energies": res.energies,
                           "seconds": res.seconds}
    out["final"] = energetics(eng, res.x)
    out["final"].update(properties(eng, res.x))
    out["x"] = res.x
    return out
