"""Input format, generator and symmetry data (CPU)."""

import os

import numpy as np
import pytest

import nbo_correlatedgaussians_amd as necg
from nbo_correlatedgaussians_amd.system import (inverse_mass_tensor, lower_to_vech,
                                                vech_to_lower)


def test_lih_fixture_matches_reference_layout():
    s = necg.lih_system()
    assert (s.n, s.nst, s.svh, s.skp, s.blk) == (5, 24, 15, 15, 30)
    assert s.efs == -0.0002
    np.testing.assert_array_equal(s.ncv, [3, 1, -1, -1, -1, -1])
    assert s.scvv_raw.sum() == 0.0
    rep = s.symmetry_report()
    assert rep == {"sms_is_kron": True, "closed": True, "inverse_weights": True}
    # mass tensor of the file vs the Li-7 / proton masses (agree to ~3e-7)
    M = inverse_mass_tensor(necg.system.M_LI7_NUCLEUS, [necg.system.M_PROTON, 1, 1, 1, 1])
    np.testing.assert_allclose(s.mass, M, rtol=1e-6)


@pytest.mark.parametrize("factory", [necg.he_system, necg.h2_system, necg.hydrogen_atom])
def test_constructed_systems(factory):
    s = factory()
    s.validate()
    assert s.symmetry_report() == {"sms_is_kron": True, "closed": True, "inverse_weights": True}


def test_input_roundtrip(tmp_path):
    s = necg.lih_system()
    iv, C = necg.generate_parameters(s, 7, 3)
    p = tmp_path / "in.dat"
    necg.write_input(str(p), s, iv, C)
    s2, K, iv2, C2 = necg.read_input(str(p))
    assert K == 7
    np.testing.assert_array_equal(iv, iv2)
    np.testing.assert_array_equal(C, C2)
    for a in ("mass", "ncv", "scvv_raw", "sml", "sms"):
        np.testing.assert_array_equal(getattr(s, a), getattr(s2, a))
    assert s2.efs == s.efs


def test_fortran_d_exponents():
    toks = "1 1 1 -0.2d+00 0.5 1.0 -1.0 1.5d0 0.0 0.0 0.0 1.0 1.0 1.0".split()
    toks += ["1.0"] * 9
    s, K, iv, C = necg.parse_tokens(toks)
    assert s.efs == -0.2 and iv[0] == 1.5 and K == 1


def test_generator_distributions():
    s = necg.lih_system()
    iv, C = necg.generate_parameters(s, 4000, 11)
    blocks = iv.reshape(4000, s.blk)
    diag_idx = [0, 5, 9, 12, 14]  # vech positions of L_ii for n = 5
    d = blocks[:, diag_idx]
    assert d.min() >= 0.1 and d.max() <= 10.0
    assert abs(np.mean(np.log(d)) - 0.0) < 0.05            # log-uniform on [ln .1, ln 10]
    off = np.delete(blocks[:, :15], diag_idx, axis=1)
    assert abs(off.std() - 0.3) < 0.01
    assert abs(blocks[:, 15:].std() - 0.1) < 0.005
    assert abs(C.std() - 1.0) < 0.05
    iv2, C2 = necg.generate_parameters(s, 4000, 11)
    np.testing.assert_array_equal(iv, iv2)


def test_vech_roundtrip():
    rng = np.random.default_rng(0)
    L = np.tril(rng.normal(size=(5, 5)))
    v = lower_to_vech(L)
    assert v[1] == L[1, 0] and v[5] == L[1, 1]     # column-major lower triangle
    np.testing.assert_array_equal(vech_to_lower(v, 5), L)


def test_unpack_x():
    s = necg.he_system()
    iv, C = necg.generate_parameters(s, 5, 0)
    K, iv2, C2 = necg.unpack_x(s, necg.pack_x(iv, C))
    assert K == 5
    with pytest.raises(ValueError):
        necg.unpack_x(s, np.zeros(7))


def test_generator_matches_reference_generator_bytes():
    """The generator restatement writes the reference generator's file
    character for character (SHA-256 of the text from
    input_file_generation/generate_lih_input.py, tests/golden/
    make_generator_golden.py), and reading it back gives the (iv, C) the
    reference's read_input gave."""
    import hashlib
    gd = np.load(os.path.join(os.path.dirname(__file__), "golden", "golden_generator.npz"))
    for M, seed in gd["cases"]:
        tag = f"m{M}_s{seed}"
        text, sysd, iv, C = necg.generated_lih_input(int(M), int(seed))
        assert len(text) == int(gd[f"{tag}_chars"]), tag
        assert hashlib.sha256(text.encode()).hexdigest() == str(gd[f"{tag}_sha256"]), tag
        if f"{tag}_iv" in gd:
            np.testing.assert_array_equal(iv, gd[f"{tag}_iv"])
            np.testing.assert_array_equal(C, gd[f"{tag}_C"])
        assert sysd.n == 5 and sysd.nst == 24
        assert len(C) == M and np.all(C == 1.0 / M)     # C_k = 1/M


def test_generator_template_statistics():
    """Diagonals positive, other columns centred on the template's column
    means (a plausible mix-up of columns or of mean and spread fails)."""
    s = necg.lih_system()
    tiv, _ = necg.lih_oldquit_basis()
    mean = tiv.reshape(-1, s.blk).mean(axis=0)
    sd = tiv.reshape(-1, s.blk).std(axis=0)
    iv, C = necg.generate_like_template(s, 20000, 5)
    b = iv.reshape(-1, s.blk)
    diag_idx = [0, 5, 9, 12, 14]
    assert np.all(b[:, diag_idx] > 0)
    assert np.allclose(b.mean(axis=0), mean, atol=4 * sd.max() / np.sqrt(20000) + 1e-12)
    assert np.allclose(b.std(axis=0), np.where(sd < 1e-10, 1e-6, sd), rtol=0.05)


def test_mass_rank1_detection_and_op_count():
    """The op count follows the instance the kernels pick: LiH's, He's and
    H2's mass matrices are diagonal plus a constant; a matrix with unequal
    off-diagonal entries is not."""
    import numpy as np
    import nbo_correlatedgaussians_amd as necg
    from nbo_correlatedgaussians_amd.flops import (mass_is_rank1, two_sided_term,
                                                   two_sided_term_flops)
    for s in (necg.lih_system(), necg.he_system(), necg.h2_system()):
        assert mass_is_rank1(s.mass.tolist()), s.name
    m = necg.lih_system().mass.copy()
    m[0, 1:] *= 1.7
    m[1:, 0] *= 1.7
    assert not mass_is_rank1(m.tolist())
    full, r1 = two_sided_term(5), two_sided_term(5, True)
    # only the kinetic phase differs: M a1 as 75 FMAs against 12 adds, 3 products, 15 FMAs
    assert {k for k in full if full[k] != r1[k]} == {"kinetic"}
    assert full["kinetic"] - r1["kinetic"] == 150 - 45
    assert two_sided_term_flops(5, rank1=True) == 4031
    assert two_sided_term_flops(5, True, True) == 4031 - r1["ket"]
