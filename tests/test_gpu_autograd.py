"""The autograd form: forward = element kernel + quotient, backward = the
hand-written adjoint kernel.  Checked against the fused outprod gradient,
central finite differences and torch.autograd.gradcheck."""

import numpy as np
import pytest

import nbo_correlatedgaussians_amd as necg

pytestmark = pytest.mark.gpu


def _setup(name, K, seed):
    import torch
    from nbo_correlatedgaussians_amd.engine import Engine
    s = {"LiH": necg.lih_system, "He": necg.he_system, "H2": necg.h2_system}[name]()
    iv, C = necg.generate_parameters(s, K, seed)
    x = torch.tensor(necg.pack_x(iv, C), dtype=torch.float64, device="cuda:0")
    return s, Engine(s, "cuda:0"), x


@pytest.mark.parametrize("name,K,seed", [("LiH", 12, 0), ("LiH", 300, 1), ("H2", 50, 2)])
def test_autograd_matches_outprod(name, K, seed):
    import torch
    from nbo_correlatedgaussians_amd.autograd import ecg_energy
    s, eng, x = _setup(name, K, seed)
    xr = x.clone().requires_grad_(True)
    E = ecg_energy(xr, eng)
    E.backward()
    Ef, gf = eng.outprod(x.cpu().numpy())
    assert abs(E.item() - Ef) <= 1e-13 * max(1.0, abs(Ef))
    g = xr.grad.cpu().numpy()
    assert np.max(np.abs(g - gf)) <= 1e-10 * max(1.0, np.max(np.abs(gf)))
    # chain rule through a scalar function of E
    xr.grad = None
    (3.0 * ecg_energy(xr, eng) ** 2).backward()
    np.testing.assert_allclose(xr.grad.cpu().numpy(), 6.0 * Ef * gf, rtol=1e-9,
                               atol=1e-9 * np.max(np.abs(6.0 * Ef * gf)))


def test_autograd_interleaved_forwards():
    """Two forwards on one engine before the backwards: each backward uses its
    own x."""
    from nbo_correlatedgaussians_amd.autograd import ecg_energy
    s, eng, x = _setup("He", 20, 3)
    x2 = x.clone()
    x2[0] *= 1.1
    a = x.clone().requires_grad_(True)
    b = x2.clone().requires_grad_(True)
    Ea = ecg_energy(a, eng)
    Eb = ecg_energy(b, eng)
    Ea.backward()
    Eb.backward()
    _, ga = eng.outprod(x.cpu().numpy())
    _, gb = eng.outprod(x2.cpu().numpy())
    assert np.max(np.abs(a.grad.cpu().numpy() - ga)) <= 1e-10 * np.max(np.abs(ga))
    assert np.max(np.abs(b.grad.cpu().numpy() - gb)) <= 1e-10 * np.max(np.abs(gb))


def test_gradcheck_small():
    import torch
    from nbo_correlatedgaussians_amd.autograd import ecg_energy
    s, eng, x = _setup("He", 5, 4)
    xr = x.clone().requires_grad_(True)
    assert torch.autograd.gradcheck(lambda t: ecg_energy(t, eng), (xr,), eps=1e-6, atol=1e-6,
                                    rtol=1e-4)


def test_tape_backward_is_outprod_gradient():
    """With gE = 1 the taped backward runs the same contraction as outprod:
    the gradient is bitwise equal; a forward without grad gives the same E
    (element-only kernel) to 1e-13; another evaluation drops the tape."""
    import torch
    from nbo_correlatedgaussians_amd.autograd import ecg_energy
    s, eng, x = _setup("LiH", 200, 7)
    xr = x.clone().requires_grad_(True)
    E = ecg_energy(xr, eng)
    E.backward()
    Ef, gf = eng.outprod(x.cpu().numpy())
    assert E.item() == Ef
    np.testing.assert_array_equal(xr.grad.cpu().numpy(), gf)
    with torch.no_grad():
        E0 = ecg_energy(x, eng)
    assert abs(E0.item() - Ef) <= 1e-13 * max(1.0, abs(Ef))
    _, t1 = eng.energy_fwd_t(x)
    _, t2 = eng.energy_fwd_t(x)
    assert t2 != t1 and eng.energy_bwd_t(x, t1) is None     # a later forward replaced it
    assert eng.energy_bwd_t(x, t2) is not None
    eng.outprod_t(x)
    assert eng.energy_bwd_t(x, t2) is None                  # another evaluation dropped it
    eng.check()
