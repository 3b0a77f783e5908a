"""Full-size parity: the whole energy and gradient of LiH K = 4096 and
K = 2048 (seed 1) and K = 256 (seed 0) against the oracle's evaluation of the
same x, stored by tests/golden/make_k2048_oracle.py.

The stored x is regenerated here and must hash to the stored SHA-256.  The
reference value is the oracle's evaluation with every sum in long double
(its double-summed evaluation is stored beside it, as the rounding spread of
a plain summation).  Tolerances: |dE| <= 1e-12 max(|E|, 1) and
max|dg| <= 1e-9 max(max|g|, 1), through both routes: ``outprod`` (the scipy
surface) and ``ecg_energy(x).backward()`` (the autograd Function)."""

import hashlib
import os

import numpy as np
import pytest

import nbo_correlatedgaussians_amd as necg

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
RTOL_E = 1e-12
RTOL_G = 1e-9


def _fixture(K, seed):
    path = os.path.join(HERE, "golden", f"golden_lih_k{K}_s{seed}_oracle.npz")
    d = np.load(path)
    s = necg.lih_system()
    x = necg.pack_x(*necg.generate_parameters(s, K, seed))
    assert hashlib.sha256(x.tobytes()).hexdigest() == str(d["x_sha256"]), "input drifted"
    return s, x, d


@pytest.mark.parametrize("K,seed", [(256, 0), (2048, 1), (4096, 1)])
def test_full_evaluation_vs_oracle_fixture(K, seed):
    import torch
    from nbo_correlatedgaussians_amd.autograd import ecg_energy
    from nbo_correlatedgaussians_amd.engine import Engine
    s, x, d = _fixture(K, seed)
    Eref, gref = float(d["E_ld"]), d["g_ld"]
    gscale = max(1.0, float(np.max(np.abs(gref))))
    print(f"LiH K={K} seed={seed}: outprod and autograd vs the oracle fixture", flush=True)
    eng = Engine(s, "cuda:0")
    E, g = eng.outprod(x)
    dE = abs(E - Eref) / max(abs(Eref), 1.0)
    dg = float(np.max(np.abs(g - gref))) / gscale
    print(f"  outprod  dE {dE:.2e}  dg {dg:.2e}", flush=True)
    assert dE <= RTOL_E, (E, Eref)
    assert dg <= RTOL_G, dg
    xr = torch.tensor(x, device="cuda:0").requires_grad_(True)
    Ea = ecg_energy(xr, eng)
    Ea.backward()
    ga = xr.grad.cpu().numpy()
    dEa = abs(Ea.item() - Eref) / max(abs(Eref), 1.0)
    dga = float(np.max(np.abs(ga - gref))) / gscale
    print(f"  autograd dE {dEa:.2e}  dg {dga:.2e}", flush=True)
    assert dEa <= RTOL_E, (Ea.item(), Eref)
    assert dga <= RTOL_G, dga
    # the double-summed oracle agrees with its long-double sums far inside the
    # tolerance: the fixture's own summation is not what is being tested
    assert abs(float(d["E"]) - Eref) <= 1e-3 * RTOL_E * max(abs(Eref), 1.0)
