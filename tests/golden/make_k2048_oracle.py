"""Write the full-size fixture of the LiH K = 2048 evaluation (seed 1), or of
another K and seed (K = 4096 and K = 256 are stored too).

    python tests/golden/make_k2048_oracle.py [--K 2048] [--seed 1] [--threads 0]

The parameters are ``generate_parameters(lih_system(), K, seed)``: L
diagonals log-uniform in [0.1, 10], off-diagonals N(0, 0.3), shifts
N(0, 0.1), C ~ N(0, 1).  The CPU oracle (a term-by-term restatement of the
reference's outprod, python/outprod.py:20-76, with make_grad,
python/make_grad.py:18-90) evaluates it once.  Every sum is carried twice:
in double in the reference's order (the numbers the reference's own loops
would form) and in long double, whose result carries only the rounding of the
per-term matel/gradel values.  The file stores the SHA-256 of x (so a test
can prove it regenerated the same input), both energies and the gradient of
the long-double sums.

The oracle is pinned to the reference at small sizes by the golden vectors
of make_golden.py (E within 2e-14, g within 1e-12) and by the sampled K = 2048
pair sums of golden_lih_k2048_s1_pairs.npz; this file extends that pin to the
whole evaluation at the headline size.  About 7 minutes on 8 cores at
K = 2048, 30 minutes at K = 4096.
"""

from __future__ import annotations

import argparse
import hashlib
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", ".."))

import nbo_correlatedgaussians_amd as necg  # noqa: E402
from oracle.oracle import Oracle  # noqa: E402


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--K", type=int, default=2048)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--threads", type=int, default=0)
    a = ap.parse_args()
    s = necg.lih_system()
    iv, C = necg.generate_parameters(s, a.K, a.seed)
    x = necg.pack_x(iv, C)
    sha = hashlib.sha256(x.tobytes()).hexdigest()
    print(f"LiH K={a.K} seed={a.seed} x sha256 {sha}", flush=True)
    t0 = time.time()
    (E, g), (El, gl) = Oracle(s).outprod_ld(x, nthreads=a.threads)
    dt = time.time() - t0
    print(f"oracle {dt:.1f} s: E (double sums) {E:.17g}, E (long double sums) {El:.17g}, "
          f"|dE|/|E| {abs(E - El) / abs(El):.3e}, max|dg|/max|g| "
          f"{np.max(np.abs(g - gl)) / np.max(np.abs(gl)):.3e}", flush=True)
    out = os.path.join(HERE, f"golden_lih_k{a.K}_s{a.seed}_oracle.npz")
    # the double-summed gradient is left out (it differs from g_ld by the
    # printed amount only); E of both summations is kept
    np.savez_compressed(out, K=np.int64(a.K), seed=np.int64(a.seed),
                        x_sha256=np.array(sha), E=np.float64(E),
                        E_ld=np.float64(El), g_ld=gl, seconds=np.float64(dt))
    print("wrote", out, flush=True)


if __name__ == "__main__":
    main()
