"""The two-sided gradient pair pass (each unordered pair once per term, bra
and ket derivatives together) against the per-ordering pass and the oracle.

Tolerances: the two forms sum the same terms in different orders, so they
agree to |dE| <= 1e-13 max(|E|, 1) and max|dg| <= 1e-11 max(max|g|, 1); each
matches the oracle to the parity bar (1e-12 energy, 1e-9 gradient)."""

import numpy as np
import pytest

import nbo_correlatedgaussians_amd as necg
from nbo_correlatedgaussians_amd import _lib
from conftest import system_by_name

pytestmark = pytest.mark.gpu


def _engine(s):
    from nbo_correlatedgaussians_amd.engine import Engine
    return Engine(s, "cuda:0")


def _close(E, g, Eref, gref, te, tg):
    return (abs(E - Eref) <= te * max(abs(Eref), 1.0)
            and np.max(np.abs(g - gref)) <= tg * max(np.max(np.abs(gref)), 1.0))


@pytest.mark.parametrize("name,K,seed", [("LiH", 37, 5), ("He", 70, 6), ("H2", 130, 7),
                                         ("Be", 65, 8), ("LiH", 1, 9)])
def test_two_sided_matches_per_ordering_and_oracle(oracle_lib, name, K, seed):
    s = system_by_name(name)
    iv, C = necg.generate_parameters(s, K, seed)
    x = necg.pack_x(iv, C)
    print(f"two-sided {name} K={K} seed={seed}", flush=True)
    eng = _engine(s)
    assert eng.pair_form()[0]
    eng.set_pair_mode(1)
    E2, g2 = eng.outprod(x)
    assert eng.pair_form() == (True, True)
    eng.set_pair_mode(2)
    E1, g1 = eng.outprod(x)
    assert eng.pair_form() == (True, False)
    assert _close(E2, g2, E1, g1, 1e-13, 1e-11)
    Eo, go = oracle_lib.Oracle(s).outprod(x)
    assert _close(E2, g2, Eo, go, 1e-12, 1e-9)
    # the no-gradient forward runs the lower triangle only; same energy
    assert abs(eng.energy(x) - E2) <= 1e-13 * max(abs(E2), 1.0)


def test_two_sided_tilings_agree_and_repeat(oracle_lib):
    """Every ket tile / term-group shape of the two-sided pass gives the
    oracle's result, and repeated calls with one shape are bitwise equal."""
    s = necg.lih_system()
    K = 150
    iv, C = necg.generate_parameters(s, K, 3)
    x = necg.pack_x(iv, C)
    Eo, go = oracle_lib.Oracle(s).outprod(x)
    eng = _engine(s)
    eng.set_pair_mode(1)
    for kt, ts in ((0, 0), (1, 1), (5, 3), (16, 24), (64, 2), (150, 1), (7, 4)):
        print(f"two-sided tiling kt={kt} ts={ts}", flush=True)
        eng.set_tiling(kt, ts, 0)
        runs = [eng.outprod(x) for _ in range(3)]
        for E, g in runs:
            assert _close(E, g, Eo, go, 1e-12, 1e-9), (kt, ts)
        assert all(E == runs[0][0] and np.array_equal(g, runs[0][1]) for E, g in runs), (kt, ts)
        assert eng.pair_form()[1]


def test_two_sided_split_last_round(oracle_lib):
    """Lists whose last round is split into half tiles (a smaller launch CU
    count makes small K reach it) give the oracle's result, in the gradient
    pass, its tape's backward and the element-only forward."""
    import torch
    s = necg.lih_system()
    eng = _engine(s)
    eng.set_pair_mode(1)
    # (K, kt, CUs): 192 = 3 whole bra tiles, 6 (1 + 2 + 3) x 64 / kt tiles:
    # kt 8 -> 48 tiles, 8 left over on 20 CUs and split; kt 4 -> 96, 16 on
    # 40; kt 2 -> 192, 32 on 80.  K 150 adds the partial bra tile's tiles and
    # the upper pairs: 33, 66 and 129 tiles, of which 3, 3 and 9 are split.
    cases = ((192, 8, 20), (192, 4, 40), (192, 2, 80), (150, 8, 10), (150, 4, 21), (150, 2, 60))
    eng.set_fold(False)     # the tile counts above are those of the unfolded lists
    for K in (192, 150):
        iv, C = necg.generate_parameters(s, K, 5)
        x = necg.pack_x(iv, C)
        Eo, go = oracle_lib.Oracle(s).outprod(x)
        xd = torch.as_tensor(x, device="cuda:0")
        for KK, kt, ncu in cases:
            if KK != K:
                continue
            print(f"split tiles K={K} kt={kt} cus={ncu}", flush=True)
            eng.set_launch_cus(ncu)
            eng.set_tiling(kt, 1, 0)
            runs = [eng.outprod(x) for _ in range(2)]
            for E, g in runs:
                assert _close(E, g, Eo, go, 1e-12, 1e-9), (K, kt, ncu)
            assert runs[0][0] == runs[1][0] and np.array_equal(runs[0][1], runs[1][1])
            assert eng.pair_form()[1]
            Ef, tape = eng.energy_fwd_t(xd)
            gt = eng.energy_bwd_t(xd, tape).cpu().numpy()
            assert float(Ef) == runs[0][0] and np.array_equal(gt, runs[0][1])
            assert abs(eng.energy(x) - Eo) <= 1e-12 * abs(Eo)
    eng.set_launch_cus(0)
    eng.set_tiling(0, 0, 0)
    eng.set_fold(True)
    eng.check()


@pytest.mark.parametrize("name,K,seed", [("LiH", 64, 21), ("LiH", 150, 22), ("LiH", 192, 23),
                                         ("He", 130, 24), ("H2", 100, 25), ("Be", 65, 26)])
def test_folded_diagonal_blocks(oracle_lib, name, K, seed):
    """Folded diagonal blocks (two kets per sweep, the diagonal pass for the
    upper halves' diagonals) for every ket tile that divides 32, against the
    unfolded pass and the oracle: the gradient pass, repeated bitwise, its
    tape's backward, and the element-only forward that walks the same tile
    list.  K = 150 adds a partial bra tile with upper pairs (unfolded), 192
    a list whose last round is split on 16 CUs (fold tiles are not split)."""
    import torch
    s = system_by_name(name)
    iv, C = necg.generate_parameters(s, K, seed)
    x = necg.pack_x(iv, C)
    Eo, go = oracle_lib.Oracle(s).outprod(x)
    xd = torch.as_tensor(x, device="cuda:0")
    eng = _engine(s)
    eng.set_pair_mode(1)
    for kt in (2, 4, 8, 16, 32):
        for ncu in ((0, 16) if K == 192 else (0,)):
            eng.set_launch_cus(ncu)
            eng.set_tiling(kt, 1, 0)
            eng.set_fold(False)
            Eu, gu = eng.outprod(x)
            eng.set_fold(True)
            print(f"fold {name} K={K} kt={kt} cus={ncu}", flush=True)
            runs = [eng.outprod(x) for _ in range(2)]
            for E, g in runs:
                assert _close(E, g, Eo, go, 1e-12, 1e-9), (kt, ncu)
                assert _close(E, g, Eu, gu, 1e-13, 1e-11), (kt, ncu)
            assert runs[0][0] == runs[1][0] and np.array_equal(runs[0][1], runs[1][1])
            Ef, tape = eng.energy_fwd_t(xd)
            gt = eng.energy_bwd_t(xd, tape).cpu().numpy()
            assert float(Ef) == runs[0][0] and np.array_equal(gt, runs[0][1])
            assert abs(eng.energy(x) - Eo) <= 1e-12 * max(abs(Eo), 1.0)
    eng.set_launch_cus(0)
    eng.set_tiling(0, 0, 0)
    eng.check()


def test_two_sided_tape_and_vjp():
    """energy_fwd takes the two-sided pass and its tape gives the outprod
    gradient; a change of form drops the tape (the backward recomputes)."""
    import torch
    s = necg.lih_system()
    iv, C = necg.generate_parameters(s, 90, 11)
    x = necg.pack_x(iv, C)
    eng = _engine(s)
    E, g = eng.outprod(x)
    xd = torch.as_tensor(x, device="cuda:0")
    Ef, tape = eng.energy_fwd_t(xd)
    assert eng.pair_form()[1]
    gt = eng.energy_bwd_t(xd, tape).cpu().numpy()
    assert float(Ef) == E
    assert np.array_equal(gt, g)
    Ef, tape = eng.energy_fwd_t(xd)
    eng.set_pair_mode(2)
    assert eng.energy_bwd_t(xd, tape) is None       # stale: the form changed
    gv = eng.vjp_t(xd).cpu().numpy()
    eng.check()
    assert _close(E, gv, E, g, 0.0, 1e-11)


def _drop_term(s, p):
    keep = [q for q in range(s.nst) if q != p]
    t = necg.System(n=s.n, nst=len(keep), efs=s.efs, mass=s.mass.copy(), ncv=s.ncv.copy(),
                    scvv_raw=s.scvv_raw[keep].copy(), sml=s.sml[keep].copy(),
                    sms=s.sms[keep].copy(), name=s.name + "-drop")
    t.validate()
    return t


def test_two_sided_refused_where_it_does_not_apply(oracle_lib):
    """Without the inverse of a term (LiH less one 3-cycle), or with masses
    that a term does not preserve, the two-sided form is unavailable: forcing
    it is refused and the default runs the per-ordering pass (oracle parity)."""
    s = necg.lih_system()
    cyc = next(p for p in range(s.nst) if not np.array_equal(s.sml[p], s.sml[p].T))
    cases = [_drop_term(s, cyc)]
    m = s.mass.copy()
    m[1, 1] *= 1.5                      # electron 1 heavier: P(1 2) no longer a symmetry
    cases.append(necg.System(n=s.n, nst=s.nst, efs=s.efs, mass=m, ncv=s.ncv.copy(),
                             scvv_raw=s.scvv_raw.copy(), sml=s.sml.copy(), sms=s.sms.copy(),
                             name="LiH-mass"))
    for t in cases:
        print(f"fallback {t.name}", flush=True)
        iv, C = necg.generate_parameters(t, 20, 4)
        x = necg.pack_x(iv, C)
        eng = _engine(t)
        assert eng.pair_form()[0] is False
        with pytest.raises(_lib.NecgError) as ei:
            eng.set_pair_mode(1)
        assert ei.value.code == _lib.NECG_ERR_UNSUPPORTED
        E, g = eng.outprod(x)
        assert eng.pair_form() == (False, False)
        Eo, go = oracle_lib.Oracle(t).outprod(x)
        assert _close(E, g, Eo, go, 1e-12, 1e-9)


def test_two_sided_general_mass_matrix(oracle_lib):
    """LiH with the nucleus-electron mass couplings changed (still invariant
    under the electron permutations, so the two-sided form applies) but no
    longer equal to the electron-electron ones: the term takes the general
    M a1 product instead of the diagonal-plus-constant form."""
    s = necg.lih_system()
    m = s.mass.copy()
    m[0, 1:] *= 1.7
    m[1:, 0] *= 1.7
    t = necg.System(n=s.n, nst=s.nst, efs=s.efs, mass=m, ncv=s.ncv.copy(),
                    scvv_raw=s.scvv_raw.copy(), sml=s.sml.copy(), sms=s.sms.copy(),
                    name="LiH-coupled")
    t.validate()
    iv, C = necg.generate_parameters(t, 70, 12)
    x = necg.pack_x(iv, C)
    print("two-sided LiH-coupled K=70", flush=True)
    eng = _engine(t)
    E2, g2 = eng.outprod(x)
    assert eng.pair_form() == (True, True)
    eng.set_pair_mode(2)
    E1, g1 = eng.outprod(x)
    assert _close(E2, g2, E1, g1, 1e-13, 1e-11)
    Eo, go = oracle_lib.Oracle(t).outprod(x)
    assert _close(E2, g2, Eo, go, 1e-12, 1e-9)
    # the element pass takes the general product too
    assert abs(eng.energy(x) - E2) <= 1e-13 * max(abs(E2), 1.0)


@pytest.mark.parametrize("K", [1000, 4096])
def test_two_sided_large_k(K):
    """LiH at K = 1000 (the last bra tile's functions paired as upper kets)
    and K = 4096 (the largest served system): the two forms agree, and the
    gradient satisfies C . g_C = 0 (E is invariant under scaling C)."""
    s = necg.lih_system()
    iv, C = necg.generate_parameters(s, K, 31)
    x = necg.pack_x(iv, C)
    print(f"two-sided LiH K={K}", flush=True)
    eng = _engine(s)
    E2, g2 = eng.outprod(x)
    assert eng.pair_form() == (True, True)
    eng.set_pair_mode(2)
    E1, g1 = eng.outprod(x)
    assert _close(E2, g2, E1, g1, 1e-13, 1e-11)
    gC = g2[K * s.blk:]
    assert abs(np.dot(C, gC)) <= 1e-10 * np.linalg.norm(C) * max(np.linalg.norm(gC), 1.0)


def test_fp64_fma_rate_probe():
    """The FMA rate probe reports a rate below MI355X's published FP64 peak
    (78.6 TFLOP/s) and well above any rate the pair kernels reach when the
    device is healthy, and agrees with itself from run to run."""
    import ctypes as ct
    lib = _lib.load()
    rates = []
    for _ in range(2):
        r = ct.c_double(0.0)
        assert lib.necg_fp64_fma_rate(0, 50000, ct.byref(r)) == _lib.NECG_OK
        rates.append(r.value)
    print(f"fp64 fma rate {rates}", flush=True)
    assert all(20.0 < v < 78.6 * 1.02 for v in rates), rates
    assert abs(rates[0] - rates[1]) <= 0.15 * max(rates)   # clocks settle between runs
