"""The dipole pass (python/necg.py:200-232 over python/rrms.py:21-120) on the
GPU against the oracle, and its size-independent properties at K = 2048.

Tolerances: C^T s C within 1e-12 relative; centres and dipole within
1e-11 of max(max |<r>|, 1)."""

import numpy as np
import pytest

import nbo_correlatedgaussians_amd as necg
from conftest import golden, system_by_name

pytestmark = pytest.mark.gpu


def _close(gpu, ref):
    css, r, d = gpu
    css_o, r_o, d_o = ref
    scale = max(float(np.max(np.abs(r_o))), 1.0)
    assert abs(css - css_o) <= 1e-12 * abs(css_o), (css, css_o)
    assert np.max(np.abs(r - r_o)) <= 1e-11 * scale, np.max(np.abs(r - r_o))
    assert np.max(np.abs(d - d_o)) <= 1e-11 * scale * max(1.0, np.abs(d_o).max())


@pytest.mark.parametrize("name,K,seed", [("He", 24, 51), ("H2", 30, 52), ("LiH", 40, 53),
                                         ("LiH", 1, 54), ("He", 150, 55), ("Be", 20, 56),
                                         ("C", 14, 57)])
def test_dipole_vs_oracle(oracle_lib, name, K, seed):
    from nbo_correlatedgaussians_amd.engine import Engine
    s = system_by_name(name)
    iv, C = necg.generate_parameters(s, K, seed)
    x = necg.pack_x(iv, C)
    print(f"dipole {name} K={K} seed={seed}", flush=True)
    _close(Engine(s, "cuda:0").dipole(x), oracle_lib.Oracle(s).dipole(x, nthreads=16))


def test_dipole_oldquit_vs_oracle(oracle_lib):
    """The LiH input file's own basis (oldquit.dat, K = 12)."""
    from nbo_correlatedgaussians_amd.engine import Engine
    d = golden("lih_oldquit")
    s = system_by_name(d["system"])
    _close(Engine(s, "cuda:0").dipole(d["x"]), oracle_lib.Oracle(s).dipole(d["x"], nthreads=16))


@pytest.mark.parametrize("case", ["lih_k12_s0", "h2_k8_s4"])
def test_dipole_vs_reference_golden(case):
    """The driver's dipole loop as the reference computes it
    (python/necg.py:200-232 over python/rrms.py), both forms of the pass."""
    from nbo_correlatedgaussians_amd.engine import Engine
    d = golden(f"dipole_{case}")
    s = system_by_name(d["system"])
    eng = Engine(s, "cuda:0")
    ref = (float(d["cSc"]), d["centres"], d["dipole"])
    for mode in (0, 2):                # symmetry-reduced (where it applies) and full double sum
        eng.set_dipole_mode(mode)
        print(f"dipole {case} mode {mode} vs the reference", flush=True)
        _close(eng.dipole(d["x"]), ref)


def test_dipole_hydrogen_closed_form():
    """One particle, two Gaussians: the C-weighted mean of the centres A, B
    and P = (aA + bB)/(a + b) with the product-theorem overlap."""
    import math
    from nbo_correlatedgaussians_amd.engine import Engine
    s = necg.hydrogen_atom()
    a, b = 0.8, 2.3
    A, B = np.array([0.1, -0.2, 0.3]), np.array([-0.4, 0.25, 0.05])
    C = np.array([0.7, -1.3])
    x = np.concatenate([[math.sqrt(a)], A, [math.sqrt(b)], B, C])
    Sab = (4 * a * b / (a + b) ** 2) ** 0.75 * math.exp(-a * b / (a + b) * np.sum((A - B) ** 2))
    P = (a * A + b * B) / (a + b)
    den = C[0] ** 2 + C[1] ** 2 + 2 * C[0] * C[1] * Sab
    num = C[0] ** 2 * A + C[1] ** 2 * B + 2 * C[0] * C[1] * Sab * P
    css, r, d = Engine(s, "cuda:0").dipole(x)
    assert abs(css - den) <= 1e-14 * den
    assert np.max(np.abs(r[0] - num / den)) <= 1e-14


def test_dipole_k2048_properties():
    """LiH K = 2048, seed 1: repeated calls are bitwise equal, and moving
    every shift by one vector t moves every centre by t with C^T s C fixed."""
    from nbo_correlatedgaussians_amd.engine import Engine
    s = necg.lih_system()
    K = 2048
    iv, C = necg.generate_parameters(s, K, 1)
    eng = Engine(s, "cuda:0")
    r1 = eng.dipole(necg.pack_x(iv, C))
    r2 = eng.dipole(necg.pack_x(iv, C))
    assert r1[0] == r2[0] and np.array_equal(r1[1], r2[1]) and np.array_equal(r1[2], r2[2])
    t = np.array([0.03, -0.02, 0.05])
    iv2 = iv.reshape(K, s.blk).copy()
    iv2[:, s.svh:] += np.tile(t, s.n)
    r3 = eng.dipole(necg.pack_x(iv2.reshape(-1), C))
    assert abs(r3[0] - r1[0]) <= 1e-11 * abs(r1[0])
    np.testing.assert_allclose(r3[1] - r1[1], np.tile(t, (s.n, 1)), rtol=0, atol=1e-10)
    # C scaling leaves the centres unchanged
    r4 = eng.dipole(necg.pack_x(iv, 3.0 * C))
    assert abs(r4[0] - 9.0 * r1[0]) <= 1e-12 * 9.0 * abs(r1[0])
    np.testing.assert_allclose(r4[1], r1[1], rtol=0, atol=1e-12 * max(1.0, np.abs(r1[1]).max()))


def test_dipole_error_paths():
    from nbo_correlatedgaussians_amd import _lib
    from nbo_correlatedgaussians_amd.engine import Engine
    s = necg.he_system()
    eng = Engine(s, "cuda:0")
    iv, C = necg.generate_parameters(s, 4, 0)
    bad = iv.copy()
    bad[0] = 0.0
    with pytest.raises(_lib.NotPositiveDefinite):
        eng.dipole(necg.pack_x(bad, C))
    with pytest.raises(ZeroDivisionError):
        eng.dipole(necg.pack_x(iv, np.zeros(4)))
    css, r, d = eng.dipole(necg.pack_x(iv, C))
    assert np.isfinite(css) and np.all(np.isfinite(r))


def test_dipole_reduced_and_full_forms_agree(oracle_lib):
    """The symmetry-reduced pass (nst term pairs per basis pair) and the full
    nst^2 double sum agree; a term set that does not close under Q P^T (LiH
    without one term) falls back to the double sum and still matches the
    oracle, and refuses the reduced form when it is forced."""
    from nbo_correlatedgaussians_amd import _lib
    from nbo_correlatedgaussians_amd.engine import Engine
    s = necg.lih_system()
    iv, C = necg.generate_parameters(s, 40, 58)
    x = necg.pack_x(iv, C)
    eng = Engine(s, "cuda:0")
    red = eng.dipole(x)
    eng.set_dipole_mode(2)
    full = eng.dipole(x)
    _close(red, full)
    _close(red, oracle_lib.Oracle(s).dipole(x, nthreads=16))
    keep = [p for p in range(s.nst) if p != 7]
    s23 = necg.System(n=s.n, nst=len(keep), efs=s.efs, mass=s.mass, ncv=s.ncv,
                      scvv_raw=s.scvv_raw[keep], sml=s.sml[keep], sms=s.sms[keep], name="LiH-23")
    e23 = Engine(s23, "cuda:0")
    _close(e23.dipole(x), oracle_lib.Oracle(s23).dipole(x, nthreads=16))
    e23.set_dipole_mode(1)
    with pytest.raises(_lib.NecgError):
        e23.dipole(x)
