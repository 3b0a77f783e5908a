"""Parity of the HIP path with the reference (golden vectors) and the oracle.

Tolerances (FP64 throughout): |dE| <= 1e-12 max(|E|, 1) and
max|dg| <= 1e-9 max(max|g|, 1); element matrices 1e-12 of their largest
entry; per-term gradients 1e-10 of their largest entry."""

import hashlib
import math

import numpy as np
import pytest

import nbo_correlatedgaussians_amd as necg
from conftest import FULL_CASES, golden, system_by_name

pytestmark = pytest.mark.gpu

RTOL_E = 1e-12
RTOL_G = 1e-9

_engines = {}


def engine_for(name):
    from nbo_correlatedgaussians_amd.engine import Engine
    if name not in _engines:
        _engines[name] = Engine(system_by_name(name), "cuda:0")
    return _engines[name]


def close_E(E, Eref):
    return abs(E - Eref) <= RTOL_E * max(abs(Eref), 1.0)


def close_g(g, gref):
    return np.max(np.abs(g - gref)) <= RTOL_G * max(np.max(np.abs(gref)), 1.0)


@pytest.mark.parametrize("case", FULL_CASES)
def test_outprod_vs_reference_golden(case):
    d = golden(case)
    eng = engine_for(str(d["system"]))
    print(f"case {case}", flush=True)
    E, g = eng.outprod(d["x"])
    assert close_E(E, float(d["E"])), (E, float(d["E"]))
    assert close_g(g, d["g"])


@pytest.mark.parametrize("case", FULL_CASES)
def test_energy_and_matrices_vs_reference_golden(case):
    d = golden(case)
    eng = engine_for(str(d["system"]))
    E, H, S = eng.energy(d["x"], with_hs=True)
    assert close_E(E, float(d["E"]))
    for a, b in ((H, d["H"]), (S, d["S"])):
        assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(b))
        np.testing.assert_array_equal(a, a.T)


@pytest.mark.parametrize("case", FULL_CASES)
def test_matel_gradel_terms_vs_reference_golden(case):
    d = golden(case)
    s = system_by_name(d["system"])
    eng = engine_for(str(d["system"]))
    K = len(d["x"]) // (s.blk + 1)
    iv = d["x"][:K * s.blk]
    q = d["queries"] - 1
    m = eng.matel(iv, K, q)
    scale = np.max(np.abs(d["matel"]), axis=0)
    scale[scale == 0] = 1.0
    assert np.all(np.max(np.abs(m - d["matel"]), axis=0) <= 1e-12 * scale)
    a, b = eng.gradel(iv, K, q)
    assert np.max(np.abs(a - d["tgvl"])) <= 1e-10 * np.max(np.abs(d["tgvl"]))
    assert np.max(np.abs(b - d["tgvls"])) <= 1e-10 * np.max(np.abs(d["tgvls"]))


def _sha(x):
    return hashlib.sha256(np.ascontiguousarray(x, dtype="<f8").tobytes()).hexdigest()


@pytest.mark.parametrize("case", ["lih_k256_s0_pairs", "lih_k2048_s1_pairs"])
def test_pair_sums_vs_reference_golden(case):
    """H_ij, S_ij and the DH/DS blocks of sampled pairs at K = 256 / 2048,
    summed over the 24 terms from the per-term entry points."""
    d = golden(case)
    s = system_by_name(d["system"])
    K = int(d["K"])
    iv, C = necg.generate_parameters(s, K, int(d["seed"]))
    assert _sha(necg.pack_x(iv, C)) == str(d["x_sha256"])
    eng = engine_for(str(d["system"]))
    pairs = d["pairs"] - 1
    P = len(pairs)
    q, qe = [], []
    for i, j in pairs:
        for p in range(s.nst):
            q.append((p, i, j))
            qe.append((p, max(i, j), min(i, j)))
    m = eng.matel(iv, K, qe).reshape(P, s.nst, 4)
    a, b = eng.gradel(iv, K, q)
    sc = s.scvv
    H = np.einsum("p,tp->t", sc, m[:, :, 1] + m[:, :, 2] - m[:, :, 3])
    S = np.einsum("p,tp->t", sc, m[:, :, 0])
    dh = np.einsum("p,tpr->tr", sc, a.reshape(P, s.nst, -1))
    ds = np.einsum("p,tpr->tr", sc, b.reshape(P, s.nst, -1))
    assert np.max(np.abs(H - d["H"])) <= 1e-12 * np.max(np.abs(d["H"]))
    assert np.max(np.abs(S - d["S"])) <= 1e-12 * np.max(np.abs(d["S"]))
    assert np.max(np.abs(dh - d["DH"])) <= 1e-10 * np.max(np.abs(d["DH"]))
    assert np.max(np.abs(ds - d["DS"])) <= 1e-10 * np.max(np.abs(d["DS"]))


@pytest.mark.parametrize("name,K,seed", [("LiH", 40, 3), ("He", 64, 4), ("H2", 64, 5),
                                         ("LiH", 130, 6), ("Be", 48, 7), ("C", 40, 8),
                                         ("C", 1, 9)])
def test_outprod_vs_oracle(oracle_lib, name, K, seed):
    s = system_by_name(name)
    iv, C = necg.generate_parameters(s, K, seed)
    x = necg.pack_x(iv, C)
    print(f"{name} K={K} seed={seed}", flush=True)
    E, g = engine_for(name).outprod(x)
    Eo, go = oracle_lib.Oracle(s).outprod(x)
    assert close_E(E, Eo), (E, Eo)
    assert close_g(g, go)


def test_dhds_vs_oracle(oracle_lib):
    s = necg.lih_system()
    K = 9
    iv, C = necg.generate_parameters(s, K, 8)
    DH, DS = engine_for("LiH").dhds(iv, K)
    o = oracle_lib.Oracle(s)
    pairs = [(l, k) for l in range(K) for k in range(K)]
    _, _, dh, ds = o.pair_sums(iv, pairs)
    DHo = np.zeros_like(DH)
    DSo = np.zeros_like(DS)
    for t, (l, k) in enumerate(pairs):
        DHo[l * s.blk:(l + 1) * s.blk, k] = dh[t]
        DSo[l * s.blk:(l + 1) * s.blk, k] = ds[t]
    assert np.max(np.abs(DH - DHo)) <= 1e-11 * np.max(np.abs(DHo))
    assert np.max(np.abs(DS - DSo)) <= 1e-11 * np.max(np.abs(DSo))


def test_k256_full_vs_oracle_and_fd(oracle_lib):
    """Config: LiH K = 256, seed 0; gradient against the oracle and central
    finite differences (h = 1e-6) on 64 seeded parameters."""
    s = necg.lih_system()
    K = 256
    iv, C = necg.generate_parameters(s, K, 0)
    x = necg.pack_x(iv, C)
    eng = engine_for("LiH")
    E, g = eng.outprod(x)
    Eo, go = oracle_lib.Oracle(s).outprod(x, nthreads=16)
    assert close_E(E, Eo), (E, Eo)
    assert close_g(g, go)
    rng = np.random.default_rng(0)
    idx = rng.choice(x.size, 64, replace=False)
    gmax = np.max(np.abs(g))
    worst = 0.0
    for i in idx:
        h = 1e-6 * max(1.0, abs(x[i]))
        xp, xm = x.copy(), x.copy()
        xp[i] += h
        xm[i] -= h
        fd = (eng.energy(xp) - eng.energy(xm)) / (2 * h)
        worst = max(worst, abs(fd - g[i]) / max(abs(g[i]), 1e-3 * gmax))
    assert worst < 1e-5, worst


def test_k2048_properties():
    """Config: LiH K = 2048, seed 1.  Size-independent checks: the split
    (outprod) and two-pass (energy + vjp) paths agree, C . g_C = 0, and
    scaling C leaves E and g_theta unchanged."""
    s = necg.lih_system()
    K = 2048
    iv, C = necg.generate_parameters(s, K, 1)
    x = necg.pack_x(iv, C)
    eng = engine_for("LiH")
    E, g = eng.outprod(x)
    assert math.isfinite(E) and np.all(np.isfinite(g))
    E2 = eng.energy(x)
    assert abs(E2 - E) <= 1e-13 * max(abs(E), 1.0)
    gx = eng.vjp_t(x).cpu().numpy()
    eng.check()
    assert close_g(gx, g)
    gC = g[K * s.blk:]
    assert abs(C @ gC) <= 1e-11 * (np.abs(C) @ np.abs(gC))
    E3, g3 = eng.outprod(necg.pack_x(iv, 2.0 * C))
    assert abs(E3 - E) <= 1e-12 * max(abs(E), 1.0)
    assert close_g(g3[:K * s.blk], g[:K * s.blk])


def test_error_paths():
    from nbo_correlatedgaussians_amd import _lib
    s = necg.he_system()
    eng = engine_for("He")
    iv, C = necg.generate_parameters(s, 4, 0)
    bad = iv.copy()
    bad[0] = 0.0          # L_11 of function 1 = 0: A_1 singular
    with pytest.raises(_lib.NotPositiveDefinite):
        eng.outprod(necg.pack_x(bad, C))
    with pytest.raises(ZeroDivisionError):
        eng.outprod(necg.pack_x(iv, np.zeros(4)))
    x = necg.pack_x(iv, C)
    x[3] = np.nan
    with pytest.raises(FloatingPointError):
        eng.outprod(x)
    # the latch is cleared: a good call afterwards succeeds
    E, g = eng.outprod(necg.pack_x(iv, C))
    assert math.isfinite(E)


@pytest.mark.parametrize("name,K,seed", [("Be", 30, 21), ("C", 26, 22)])
def test_n4_n6_energy_vjp_and_dhds(oracle_lib, name, K, seed):
    """n = 4 and n = 6: the element-only, adjoint and materialised-derivative
    paths against the oracle."""
    s = system_by_name(name)
    iv, C = necg.generate_parameters(s, K, seed)
    x = necg.pack_x(iv, C)
    eng = engine_for(name)
    o = oracle_lib.Oracle(s)
    Eo, go, Ho, So = o.outprod(x, return_hs=True)
    E, H, S = eng.energy(x, with_hs=True)
    assert close_E(E, Eo)
    assert np.max(np.abs(H - Ho)) <= 1e-12 * np.max(np.abs(Ho))
    gx = eng.vjp_t(x).cpu().numpy()
    eng.check()
    assert close_g(gx, go)
    DH, DS = eng.dhds(iv, K)
    _, _, dh, ds = o.pair_sums(iv, [(3, 5), (5, 3), (7, 7)])
    for t, (l, k) in enumerate([(3, 5), (5, 3), (7, 7)]):
        assert np.max(np.abs(DH[l * s.blk:(l + 1) * s.blk, k] - dh[t])) <= 1e-11 * np.max(np.abs(dh))
        assert np.max(np.abs(DS[l * s.blk:(l + 1) * s.blk, k] - ds[t])) <= 1e-11 * np.max(np.abs(ds))


@pytest.mark.parametrize("case", ["lih_oldquit", "lih_k32_s2"])
def test_tilings_agree_and_repeat(case):
    """Every launch shape (ket tile, term groups) gives the golden energy and
    gradient, and repeated calls are bitwise identical."""
    from nbo_correlatedgaussians_amd.engine import Engine
    d = golden(case)
    eng = Engine(system_by_name(d["system"]), "cuda:0")
    for kt, ts in ((0, 0), (1, 24), (4, 6), (16, 1), (1, 1), (3, 5)):
        eng.set_tiling(kt, ts, 0)
        runs = [eng.outprod(d["x"]) for _ in range(3)]
        for E, g in runs:
            assert close_E(E, float(d["E"])), (kt, ts, E)
            assert close_g(g, d["g"]), (kt, ts)
        assert all(E == runs[0][0] and np.array_equal(g, runs[0][1]) for E, g in runs), (kt, ts)
        Ee, H, S = eng.energy(d["x"], with_hs=True)
        assert close_E(Ee, float(d["E"]))
        gx = eng.vjp_t(d["x"]).cpu().numpy()
        eng.check()
        assert close_g(gx, d["g"]), (kt, ts)


def test_k2048_rotation_invariance():
    """Config: LiH K = 2048, seed 1.  Rotating every shift about z (the
    Stark field's axis) leaves E unchanged and rotates the shift gradients."""
    s = necg.lih_system()
    K = 2048
    iv, C = necg.generate_parameters(s, K, 1)
    x = necg.pack_x(iv, C)
    eng = engine_for("LiH")
    E, g = eng.outprod(x)
    c, sn = np.cos(1.1), np.sin(1.1)
    R = np.array([[c, -sn, 0.0], [sn, c, 0.0], [0.0, 0.0, 1.0]])
    y = x.copy()
    blocks = y[:K * s.blk].reshape(K, s.blk)
    blocks[:, s.svh:] = (blocks[:, s.svh:].reshape(K, s.n, 3) @ R.T).reshape(K, -1)
    E2, g2 = eng.outprod(y)
    assert abs(E2 - E) <= 1e-12 * max(1.0, abs(E)), (E, E2)
    gb, gb2 = g[:K * s.blk].reshape(K, s.blk), g2[:K * s.blk].reshape(K, s.blk)
    scale = max(1.0, np.max(np.abs(g)))
    rot = gb[:, s.svh:].reshape(K, s.n, 3) @ R.T
    assert np.max(np.abs(gb2[:, s.svh:].reshape(K, s.n, 3) - rot)) <= 1e-9 * scale
    assert np.max(np.abs(gb2[:, :s.svh] - gb[:, :s.svh])) <= 1e-9 * scale
    assert np.max(np.abs(g2[K * s.blk:] - g[K * s.blk:])) <= 1e-9 * scale


@pytest.mark.parametrize("cus,kt", [(0, 0), (14, 4)])
def test_engine_graph_replays(cus, kt):
    """The single-system evaluation captured in a HIP graph: every replay,
    with new parameters written into x in place and eager work between
    replays, is bitwise the eager outprod.  The second case launches for 14
    CUs with ket tile 4: K = 96 then has 32 tiles, 4 of them split in halves
    for the last round."""
    import torch
    from nbo_correlatedgaussians_amd.engine import Engine, EngineGraph
    s = necg.lih_system()
    K = 96
    eng = Engine(s, "cuda:0")
    eng.set_launch_cus(cus)
    eng.set_tiling(kt, 1 if kt else 0, 0)
    xs = [necg.pack_x(*necg.generate_parameters(s, K, seed)) for seed in (31, 32, 33)]
    G = EngineGraph(eng, xs[0])
    for rep in range(4):
        x = xs[rep % 3]
        G.x.copy_(torch.as_tensor(x, device="cuda:0"))
        f, g = G.replay()
        torch.cuda.synchronize()
        eng.check()
        Eg, gg = float(f.item()), g.cpu().numpy()
        Ee, ge = eng.outprod(x)          # eager work between replays
        assert Eg == Ee and np.array_equal(gg, ge), rep
    # while the graph lives, nothing may move or rewrite what it captured:
    # another K (smaller: a new tile list; larger: new buffers) and the
    # launch-shape setters are refused, and the replay still matches
    from nbo_correlatedgaussians_amd import _lib
    for Kx in (K - 32, K + 32):
        with pytest.raises(_lib.NecgError):
            eng.outprod(necg.pack_x(*necg.generate_parameters(s, Kx, 34)))
    for call in (lambda: eng.set_tiling(8, 1, 0), lambda: eng.set_launch_cus(0),
                 lambda: eng.set_pair_mode(2)):
        with pytest.raises(_lib.NecgError):
            call()
    G.x.copy_(torch.as_tensor(xs[1], device="cuda:0"))
    f, g = G.replay()
    Ee, ge = eng.outprod(xs[1])
    assert float(f.item()) == Ee and np.array_equal(g.cpu().numpy(), ge)
    G.release()
    E64, _ = eng.outprod(necg.pack_x(*necg.generate_parameters(s, K - 32, 34)))
    assert np.isfinite(E64)


def test_host_pointers_are_refused():
    """Pageable host memory is refused at the boundary (NECG_ERR_ARG), not
    handed to a kernel; device tensors pass the same check."""
    import ctypes as ct
    import torch
    from nbo_correlatedgaussians_amd import _lib
    lib = _lib.load()
    s = necg.he_system()
    iv, C = necg.generate_parameters(s, 4, 0)
    xh = necg.pack_x(iv, C)
    gh = np.zeros_like(xh)
    xd = torch.as_tensor(xh, device="cuda:0")
    assert lib.necg_device_visible(ct.c_void_p(xd.data_ptr())) == 1
    assert lib.necg_device_visible(ct.c_void_p(xh.ctypes.data)) == 0
    assert lib.necg_device_visible(ct.c_void_p(gh.ctypes.data)) == 0
    eng = engine_for("He")
    f = torch.empty(1, dtype=torch.float64, device="cuda:0")
    rc = lib.necg_outprod(eng._ctx, 4, ct.c_void_p(xh.ctypes.data), ct.c_void_p(f.data_ptr()),
                          ct.c_void_p(gh.ctypes.data), ct.c_void_p(0))
    assert rc == _lib.NECG_ERR_ARG
    E, g = eng.outprod(xh)          # the engine path stages host arrays on the device
    assert np.isfinite(E)


def test_outprod_on_generator_input_vs_reference_golden():
    """LiH K = 12 made by the reference's input generator from seed 0, read
    back from the generator's file: outprod against the reference's outprod
    on the same file."""
    gd = golden("generator")
    _, s, iv, C = necg.generated_lih_input(12, 0)
    from nbo_correlatedgaussians_amd.engine import Engine
    print("case generator m12_s0", flush=True)
    E, g = Engine(s, "cuda:0").outprod(necg.pack_x(iv, C))
    assert close_E(E, float(gd["m12_s0_E"])), (E, float(gd["m12_s0_E"]))
    assert close_g(g, gd["m12_s0_g"])


@pytest.mark.parametrize("K", [2, 63, 64, 65, 127, 129])
def test_block_boundaries_vs_oracle(oracle_lib, K):
    """LiH at K around the 64-function blocks of the record, mirror and
    gradient kernels (partial last blocks, one-block matrices): both pair
    forms, the element pass and H, S against the oracle."""
    s = necg.lih_system()
    iv, C = necg.generate_parameters(s, K, 40 + K)
    x = necg.pack_x(iv, C)
    print(f"LiH K={K} block boundaries", flush=True)
    o = oracle_lib.Oracle(s)
    Eo, go, Ho, So = o.outprod(x, return_hs=True)
    eng = engine_for("LiH")
    for mode in (1, 2):
        eng.set_pair_mode(mode)
        E, g = eng.outprod(x)
        assert close_E(E, Eo), (mode, E, Eo)
        assert close_g(g, go), mode
    eng.set_pair_mode(0)
    E, H, S = eng.energy(x, with_hs=True)
    assert close_E(E, Eo)
    assert np.max(np.abs(H - Ho)) <= 1e-12 * np.max(np.abs(Ho))
    assert np.max(np.abs(S - So)) <= 1e-12 * np.max(np.abs(So))
    assert np.array_equal(H, H.T) and np.array_equal(S, S.T)
