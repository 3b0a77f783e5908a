"""The module-level drop-in (the reference's ``read_input`` / ``outprod`` /
``matel`` / ``gradel`` surface with 1-based indices, and its driver's dipole
pass) on the GPU, against the reference's golden vectors and the oracle."""

import numpy as np
import pytest

import nbo_correlatedgaussians_amd as necg
from conftest import golden, system_by_name

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def installed(tmp_path_factory):
    from nbo_correlatedgaussians_amd import api
    d = golden("lih_oldquit")
    s = system_by_name(d["system"])
    K = d["x"].size // (s.blk + 1)
    path = tmp_path_factory.mktemp("input") / "oldquit.dat"
    necg.write_input(str(path), s, d["x"][:K * s.blk], d["x"][K * s.blk:])
    api.read_input(str(path), device="cuda:0")
    return api, d, s


def test_api_outprod_and_terms(installed):
    api, d, s = installed
    np.testing.assert_array_equal(api.x0(), d["x"])
    E, g = api.outprod(api.x0())
    assert abs(E - float(d["E"])) <= 1e-12 * max(1.0, abs(float(d["E"])))
    assert np.max(np.abs(g - d["g"])) <= 1e-9 * max(1.0, np.max(np.abs(d["g"])))
    for t, (stf, ii, jj) in enumerate(d["queries"][:6]):
        m = api.matel(int(stf), int(ii), int(jj))
        scale = np.maximum(np.max(np.abs(d["matel"]), axis=0), 1e-300)
        assert np.all(np.abs(np.array(m) - d["matel"][t]) <= 1e-12 * scale)
        a, b = api.gradel(int(stf), int(ii), int(jj))
        assert np.max(np.abs(a - d["tgvl"][t])) <= 1e-10 * np.max(np.abs(d["tgvl"]))
        assert np.max(np.abs(b - d["tgvls"][t])) <= 1e-10 * np.max(np.abs(d["tgvls"]))
    H, S = api.build_matrices()
    assert np.max(np.abs(H - d["H"])) <= 1e-12 * np.max(np.abs(d["H"]))


def test_api_with_scipy_minimize(installed):
    """outprod is handed to scipy.optimize.minimize(jac=True) unchanged."""
    from scipy.optimize import minimize
    api, d, s = installed
    x0 = d["x"].copy()
    r = minimize(api.outprod, x0, method="L-BFGS-B", jac=True, options={"maxiter": 2})
    assert r.fun < float(d["E"])


def test_api_dipole_vs_oracle(installed, oracle_lib):
    api, d, s = installed
    api.use(s, d["x"][:-12], d["x"][-12:])
    out = api.dipole()
    css, rbar, dip = oracle_lib.Oracle(s).dipole(d["x"])
    scale = max(1.0, np.max(np.abs(rbar)))
    assert abs(out["cSc"] - css) <= 1e-12 * abs(css)
    assert np.max(np.abs(out["centres"] - rbar)) <= 1e-11 * scale
    assert np.max(np.abs(out["dipole_au"] - dip)) <= 1e-11 * scale * max(1.0, np.abs(dip).max())
    np.testing.assert_allclose(out["dipole_debye"], out["dipole_au"] * 2.541765, rtol=1e-15)
