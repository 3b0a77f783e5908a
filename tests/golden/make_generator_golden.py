"""Write golden vectors for the input generator (run in a checkout that has
the reference mounted; never on a machine without it).

    python tests/golden/make_generator_golden.py [/path/to/reference]

For each (M, seed) the reference's ``generate_lih_input``
(input_file_generation/generate_lih_input.py:310) writes its input file; the
file's SHA-256 and length are stored, and for the small cases the (iv, C) the
reference's ``read_input`` (python/read_input.py:16) reads from it.  For
M = 12, seed 0 (the LiH K = 12 case made by the generator from seed 0) the
reference's ``outprod`` (python/outprod.py:20) on that file is stored as
well.  Output: ``golden_generator.npz`` (plain arrays, no pickles).
"""

from __future__ import annotations

import hashlib
import importlib.util
import os
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

CASES = [(12, 0), (4, 7), (1, 3), (256, 0), (2048, 1)]
SMALL = 64          # arrays are stored for M up to this


def _module(path, name):
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def main():
    ref_root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    gen = _module(os.path.join(ref_root, "input_file_generation", "generate_lih_input.py"),
                  "generate_lih_input")
    sys.path.insert(0, os.path.join(ref_root, "python"))
    import read_input as ref_read   # noqa: E402
    import shared as ref_sh         # noqa: E402
    import outprod as ref_outprod   # noqa: E402

    out = {"cases": np.array(CASES, dtype=np.int64)}
    with tempfile.TemporaryDirectory() as wd:
        cwd = os.getcwd()
        os.chdir(wd)
        try:
            for M, seed in CASES:
                text = gen.generate_lih_input(M, reference=os.path.join(ref_root, "oldquit.dat"),
                                              seed=seed)
                tag = f"m{M}_s{seed}"
                out[f"{tag}_sha256"] = np.array(hashlib.sha256(text.encode()).hexdigest())
                out[f"{tag}_chars"] = np.int64(len(text))
                if M <= SMALL:
                    path = os.path.join(wd, f"{tag}.dat")
                    with open(path, "w", encoding="utf-8") as fh:
                        fh.write(text)
                    ref_read.read_input(path)
                    out[f"{tag}_iv"] = np.array(ref_sh.iv, dtype=np.float64).copy()
                    out[f"{tag}_C"] = np.array(ref_sh.C, dtype=np.float64).copy()
                    if (M, seed) == (12, 0):
                        x = np.concatenate([out[f"{tag}_iv"], out[f"{tag}_C"]])
                        t0 = time.time()
                        E, g = ref_outprod.outprod(x.copy())
                        out[f"{tag}_E"] = np.float64(E)
                        out[f"{tag}_g"] = np.array(g, dtype=np.float64)
                        print(f"{tag}: outprod E={E:.15e} ({time.time() - t0:.1f} s)")
                print(f"{tag}: {len(text)} chars")
        finally:
            os.chdir(cwd)
    np.savez(os.path.join(HERE, "golden_generator.npz"), **out)
    print("wrote", os.path.join(HERE, "golden_generator.npz"))


if __name__ == "__main__":
    main()
