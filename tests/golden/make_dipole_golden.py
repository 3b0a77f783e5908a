"""Write golden vectors of the reference's dipole pass (run in a checkout that
has the reference mounted; never on a machine without it).

    python tests/golden/make_dipole_golden.py [/path/to/reference]

For LiH K = 12 (seed 0) and H2 K = 8 (seed 4), the input file is written by
this package's writer and loaded by the reference's own ``read_input``
(python/read_input.py:16).  Stored per case:

* ``rrms`` (python/rrms.py:21-120) for a seeded sample of (bra term, ket
  term, bra, ket, component) queries, every term pair drawn from all nst^2
  (most of them with a non-identity bra term);
* the driver's dipole loop (python/necg.py:200-232) evaluated with the
  reference's ``rrms``: for every pseudo-particle and Cartesian component the
  centre d3 = C^T h C / C^T s C, the C^T s C of the loop, and the dipole
  components d7, d8, d9.

Only inputs and outputs are stored (``golden_dipole_<case>.npz``).
"""

from __future__ import annotations

import importlib
import os
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", ".."))

import nbo_correlatedgaussians_amd as necg  # noqa: E402


def _load_reference(ref_root):
    sys.path.insert(0, os.path.join(ref_root, "python"))
    return {name: importlib.import_module(name) for name in ("shared", "read_input", "rrms")}


def dipole_loop(ref):
    """The loop of python/necg.py:200-232 over the reference's rrms: the
    centres d3 (particle-major, n x 3), C^T s C and (d7, d8, d9)."""
    sh, rrms = ref["shared"], ref["rrms"].rrms
    M, ne, nst = sh.M, sh.ne, sh.nst
    C = np.asarray(sh.C, dtype=np.float64)
    d3 = np.zeros((ne, 3))
    dip = np.zeros(3)
    css = None
    kk = 1
    for ii in range(1, ne + 1):
        for comp in range(3):
            h = np.zeros((M, M))
            s = np.zeros((M, M))
            for stfe in range(1, nst + 1):
                for stf in range(1, nst + 1):
                    w = sh.scvv[stf - 1] * sh.scvv[stfe - 1]
                    for i in range(1, M + 1):
                        for j in range(i, M + 1):
                            ov, tk = rrms(stfe, stf, j, i, kk)
                            h[j - 1, i - 1] += w * tk
                            s[j - 1, i - 1] += w * ov
            h = np.tril(h) + np.tril(h, -1).T
            s = np.tril(s) + np.tril(s, -1).T
            chc, csc = float(C @ h @ C), float(C @ s @ C)
            d3[ii - 1, comp] = chc / csc
            dip[comp] += sh.ncv[ii] * d3[ii - 1, comp]
            css = csc
            kk += 1
    return d3, css, dip


def case(ref, name, sysd, K, seed, workdir, n_queries=48):
    iv, C = necg.generate_parameters(sysd, K, seed)
    path = os.path.join(workdir, "input.dat")
    necg.write_input(path, sysd, iv, C)
    ref["read_input"].read_input(path)
    rng = np.random.default_rng(77)
    q = []
    for _ in range(n_queries):
        stfe = int(rng.integers(1, sysd.nst + 1))
        stf = int(rng.integers(1, sysd.nst + 1))
        i = int(rng.integers(1, K + 1))
        j = int(rng.integers(1, K + 1))
        n1 = int(rng.integers(1, 3 * sysd.n + 1))
        q.append((stfe, stf, i, j, n1))
    q = np.array(q, dtype=np.int64)
    vals = np.array([ref["rrms"].rrms(*map(int, t)) for t in q])
    t0 = time.time()
    d3, css, dip = dipole_loop(ref)
    dt = time.time() - t0
    out = os.path.join(HERE, f"golden_dipole_{name}.npz")
    np.savez(out, system=np.array(sysd.name), x=necg.pack_x(iv, C), queries=q, rrms=vals,
             centres=d3, cSc=np.float64(css), dipole=dip, seconds=np.float64(dt))
    print(f"{name}: K={K} dipole {dip} ({dt:.1f} s)", flush=True)


def main():
    ref_root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    ref = _load_reference(ref_root)
    with tempfile.TemporaryDirectory() as wd:
        case(ref, "lih_k12_s0", necg.lih_system(), 12, 0, wd)
        case(ref, "h2_k8_s4", necg.h2_system(), 8, 4, wd)


if __name__ == "__main__":
    main()
