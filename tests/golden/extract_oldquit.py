"""Extract the LiH non-BO system of the reference input ``oldquit.dat`` into
the package's ``data/lih_oldquit.json`` (run once, in a checkout that has the
reference).

The reference input is MIT-licensed (c) 2026 Mauricio Cafiero; the numbers are
data (mass tensor, charges, symmetry terms, and the 12-function basis).  The
token layout is the one ``python/read_input.py:16-157`` reads.

    python tests/golden/extract_oldquit.py /path/to/reference/oldquit.dat
"""

import json
import os
import sys

ROOT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "..")
sys.path.insert(0, ROOT)

from nbo_correlatedgaussians_amd.system import parse_tokens  # noqa: E402


def main(path):
    with open(path, "r", encoding="utf-8") as fh:
        sysd, K, iv, C = parse_tokens(fh.read().split())
    out = {
        "source": "oldquit.dat of MauricioCafiero/NBO_CorrelatedGaussians (MIT)",
        "name": "LiH",
        "n": sysd.n,
        "nst": sysd.nst,
        "efs": sysd.efs,
        "mass": sysd.mass.tolist(),
        "ncv": sysd.ncv.tolist(),
        "scvv_raw": sysd.scvv_raw.tolist(),
        "sml": sysd.sml.tolist(),
        # sms is kron(sml[p].T, I3) for every term of this file; it is stored
        # in full anyway so that the fixture is the file's data, not a rule.
        "sms": sysd.sms.tolist(),
        "K": K,
        "iv": iv.tolist(),
        "C": C.tolist(),
    }
    dst = os.path.join(ROOT, "nbo_correlatedgaussians_amd", "data", "lih_oldquit.json")
    with open(dst, "w", encoding="utf-8") as fh:
        json.dump(out, fh, separators=(",", ":"))
    print("wrote", dst)


if __name__ == "__main__":
    if len(sys.argv) != 2:
        sys.exit("usage: extract_oldquit.py /path/to/reference/oldquit.dat")
    main(sys.argv[1])
