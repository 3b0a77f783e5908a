"""Serving: single-function replacement trials against resident systems and
full evaluations under continuous batching.  A trial's row and energy match a
full recomputation with the replaced function; every result is bitwise the
same alone and in the mix (full evaluations: a masked batch evaluation with
only that system active), and matches the per-system engine; a small memory budget still serves everything."""

import numpy as np
import pytest

import nbo_correlatedgaussians_amd as necg

pytestmark = pytest.mark.gpu


def _systems():
    out = []
    for name, K, seed in (("LiH", 40, 21), ("He", 60, 22), ("H2", 50, 23)):
        s = {"LiH": necg.lih_system, "He": necg.he_system, "H2": necg.h2_system}[name]()
        iv, C = necg.generate_parameters(s, K, seed)
        out.append((s, necg.pack_x(iv, C)))
    return out


def test_trial_matches_full_recomputation():
    from nbo_correlatedgaussians_amd.engine import Engine
    from nbo_correlatedgaussians_amd.serve import Request, Server
    sx = _systems()
    srv = Server([s for s, _ in sx], [x for _, x in sx], "cuda:0", trial_streams=1,
                 eval_streams=1)
    rng = np.random.default_rng(5)
    for si, (s, x) in enumerate(sx):
        K = srv.Ks[si]
        for k in (0, K // 2, K - 1):
            iv1, _ = necg.generate_parameters(s, 1, int(rng.integers(1 << 30)))
            E, (rh, rs) = srv.trial_alone(Request(0, "trial", si, 0.0, k, iv1))
            x2 = x.copy()
            x2[k * s.blk:(k + 1) * s.blk] = iv1
            eng = Engine(s, "cuda:0")
            Ef, H, S = eng.energy(x2, with_hs=True)
            assert abs(E - Ef) <= 1e-11 * max(1.0, abs(Ef)), (si, k, E, Ef)
            assert np.max(np.abs(rh - H[k])) <= 1e-12 * np.max(np.abs(H))
            assert np.max(np.abs(rs - S[k])) <= 1e-12 * np.max(np.abs(S))


def test_mix_equals_alone_and_budget():
    from nbo_correlatedgaussians_amd.engine import Engine
    from nbo_correlatedgaussians_amd.serve import Server, latency_stats, poisson_trace
    sx = _systems()
    systems = [s for s, _ in sx]
    xs = [x for _, x in sx]
    srv = Server(systems, xs, "cuda:0")
    reqs = poisson_trace(systems, srv.Ks, 300, rate=3000.0, seed=1)
    out = srv.run(reqs)
    assert all(r.done is not None for r in out)
    st = latency_stats(out)
    assert st["p99_ms"] >= st["p50_ms"] > 0
    full_alone, full_eng = {}, {}
    for si, (s, x) in enumerate(sx):
        full_alone[si] = srv.full_alone(si)
        full_eng[si] = Engine(s, "cuda:0").outprod(x)
    nfull = 0
    for r in out:
        if r.kind == "trial":
            E, _ = srv.trial_alone(r)
            assert r.E == E, r.rid
        else:
            nfull += 1
            Ef, gf = full_alone[r.sys]
            assert r.E == Ef and np.array_equal(r.g, gf), r.rid
    assert nfull > 0
    # the masked batch evaluation agrees with the per-system engine
    for si in full_alone:
        Ef, gf = full_alone[si]
        Ee, ge = full_eng[si]
        assert abs(Ef - Ee) <= 1e-12 * max(1.0, abs(Ee))
        assert np.max(np.abs(gf - ge)) <= 1e-9 * max(1.0, np.max(np.abs(ge)))
    # a budget of two trials' workspace still serves every request
    small = Server(systems, xs, "cuda:0", mem_budget=2 * srv.trial_bytes + max(srv.full_bytes))
    out2 = small.run(poisson_trace(systems, small.Ks, 60, rate=5000.0, seed=2))
    assert all(r.done is not None and np.isfinite(r.E) for r in out2)
