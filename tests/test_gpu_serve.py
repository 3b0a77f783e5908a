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
    """Row, Rayleigh quotient and lowest eigenvalue of (H', S') against a full
    recomputation with the replaced function (scipy.linalg.eigh for the
    eigenvalue: 1e-9 relative)."""
    from scipy.linalg import eigh
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
            E, (rh, rs), Ee = srv.trial_alone(Request(0, "trial", si, 0.0, k, iv1))
            x2 = x.copy()
            x2[k * s.blk:(k + 1) * s.blk] = iv1
            eng = Engine(s, "cuda:0")
            Ef, H, S = eng.energy(x2, with_hs=True)
            assert abs(E - Ef) <= 1e-11 * max(1.0, abs(Ef)), (si, k, E, Ef)
            assert np.max(np.abs(rh - H[k])) <= 1e-12 * np.max(np.abs(H))
            assert np.max(np.abs(rs - S[k])) <= 1e-12 * np.max(np.abs(S))
            w0 = eigh(H, S, eigvals_only=True, subset_by_index=[0, 0])[0]
            assert abs(Ee - w0) <= 1e-9 * max(1.0, abs(w0)), (si, k, Ee, w0)
            assert Ee <= E + 1e-12 * max(1.0, abs(E))


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
            E, _, Ee = srv.trial_alone(r)
            assert r.E == E and r.Eeig == Ee, r.rid
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


def test_trial_eigenvalue_edge_cases():
    """A replacement by a copy of another function (S' singular) has no
    eigenvalue; replacing a function by itself gives the resident lowest
    eigenvalue; malformed requests are refused on the host."""
    from scipy.linalg import eigh
    from nbo_correlatedgaussians_amd.serve import Request, Server
    sx = _systems()
    srv = Server([s for s, _ in sx], [x for _, x in sx], "cuda:0", trial_streams=1,
                 eval_streams=1)
    for si, (s, x) in enumerate(sx):
        K = srv.Ks[si]
        blocks = x[:K * s.blk].reshape(K, s.blk)
        E, _, Ee = srv.trial_alone(Request(0, "trial", si, 0.0, 3, blocks[7].copy()))
        assert np.isnan(Ee) and np.isfinite(E)
        H, S = srv.batch.matrices(si)
        w0 = eigh(H, S, eigvals_only=True, subset_by_index=[0, 0])[0]
        E, _, Ee = srv.trial_alone(Request(0, "trial", si, 0.0, 5, blocks[5].copy()))
        assert abs(Ee - w0) <= 1e-9 * max(1.0, abs(w0))
    with pytest.raises(ValueError):
        srv.trial_alone(Request(0, "trial", 0, 0.0, srv.Ks[0], blocks[0].copy()))
    with pytest.raises(ValueError):
        srv.run([Request(0, "trial", 7, 0.0, 0, blocks[0].copy())])


def test_served_config_at_full_size(oracle_lib):
    """Serving at its configured size: 32 resident He/H2/LiH systems with K
    log-uniform in [32, 4096] (seeds 200..231) and a trace of 1,000 Poisson
    requests (70 % trials).  Every full request is bitwise its masked
    evaluation alone and within 1e-12 (E) / 1e-9 (g) of the per-system
    engine; a seeded sample of trials matches a full recomputation with the
    replaced function (row of H and S, E) and scipy.linalg.eigh of (H', S')
    (Eeig, 1e-9 relative); the trials on systems with K <= 64 match the
    oracle's rows of H' and S'."""
    from scipy.linalg import eigh
    from nbo_correlatedgaussians_amd.batch import mixed_systems
    from nbo_correlatedgaussians_amd.engine import Engine
    from nbo_correlatedgaussians_amd.serve import Server, latency_stats, poisson_trace
    systems, Ks, xs = mixed_systems(count=32, seed0=200, kmax=4096)
    assert len(systems) == 32 and max(Ks) <= 4096
    print(f"serving 32 systems, K {min(Ks)}..{max(Ks)}", flush=True)
    srv = Server(systems, xs, "cuda:0")
    reqs = srv.run(poisson_trace(systems, srv.Ks, 1000, rate=2000.0, seed=11))
    assert all(r.done is not None for r in reqs)
    st = latency_stats(reqs)
    print(f"  p50 {st['p50_ms']:.3f} ms  p99 {st['p99_ms']:.3f} ms", flush=True)
    fulls = sorted({r.sys for r in reqs if r.kind == "full"})
    alone = {s: srv.full_alone(s) for s in fulls}
    for r in reqs:
        if r.kind == "full":
            Ef, gf = alone[r.sys]
            assert r.E == Ef and np.array_equal(r.g, gf), r.rid
    for s in fulls:
        Ee, ge = Engine(systems[s], "cuda:0").outprod(xs[s])
        Ef, gf = alone[s]
        assert abs(Ef - Ee) <= 1e-12 * max(1.0, abs(Ee)), s
        assert np.max(np.abs(gf - ge)) <= 1e-9 * max(1.0, np.max(np.abs(ge))), s
    print(f"  {len(fulls)} systems' full evaluations checked", flush=True)
    trials = [r for r in reqs if r.kind == "trial"]
    rng = np.random.default_rng(3)
    small = [r for r in trials if srv.Ks[r.sys] <= 64]
    mid = [r for r in trials if 64 < srv.Ks[r.sys] <= 1200]
    big = [r for r in trials if srv.Ks[r.sys] > 1200]
    sample = small[:6] + [mid[i] for i in rng.choice(len(mid), min(6, len(mid)), replace=False)]
    sample += [big[i] for i in rng.choice(len(big), min(2, len(big)), replace=False)]
    assert small and mid and big
    engines = {}
    for r in sample:
        K, sd = srv.Ks[r.sys], systems[r.sys]
        E, (rh, rs), Ee = srv.trial_alone(r)
        # the same in the mix as alone (NaN where S' is not positive definite)
        assert r.E == E and np.array_equal([r.Eeig], [Ee], equal_nan=True), r.rid
        x2 = xs[r.sys].copy()
        x2[r.func * sd.blk:(r.func + 1) * sd.blk] = r.newp
        eng = engines.setdefault(r.sys, Engine(sd, "cuda:0"))
        Ef, H, S = eng.energy(x2, with_hs=True)
        assert abs(E - Ef) <= 1e-11 * max(1.0, abs(Ef)), (r.rid, E, Ef)
        assert np.max(np.abs(rh - H[r.func])) <= 1e-12 * np.max(np.abs(H)), r.rid
        assert np.max(np.abs(rs - S[r.func])) <= 1e-12 * np.max(np.abs(S)), r.rid
        if np.isfinite(Ee):
            w0 = eigh(H, S, eigvals_only=True, subset_by_index=[0, 0])[0]
            assert abs(Ee - w0) <= 1e-9 * max(1.0, abs(w0)), (r.rid, K, Ee, w0)
        if K <= 64:
            Ho, So = oracle_lib.Oracle(sd).matrices(x2[:K * sd.blk], K)
            assert np.max(np.abs(rh - Ho[r.func])) <= 1e-12 * np.max(np.abs(Ho)), r.rid
            assert np.max(np.abs(rs - So[r.func])) <= 1e-12 * np.max(np.abs(So)), r.rid
        print(f"  trial {r.rid}: system {r.sys} K={K} function {r.func} ok", flush=True)


def test_accept_makes_replacement_resident():
    """An accept request puts its block into the resident x: its own result
    is outprod at the new x; a trial that arrives after it is computed on the
    new basis (row, E and, from the re-formed factors, the lowest eigenvalue
    against scipy.linalg.eigh); a later full request returns the accept's
    (E, g) bitwise; a trial that arrived before it saw the old basis."""
    from scipy.linalg import eigh
    from nbo_correlatedgaussians_amd.engine import Engine
    from nbo_correlatedgaussians_amd.serve import Request, Server
    sx = _systems()
    systems, xs = [s for s, _ in sx], [x.copy() for _, x in sx]
    srv = Server(systems, xs, "cuda:0")
    s = systems[0]
    x0 = xs[0].copy()
    p5, _ = necg.generate_parameters(s, 1, 71)
    p7, _ = necg.generate_parameters(s, 1, 72)
    E_old, _, _ = srv.trial_alone(Request(0, "trial", 0, 0.0, 5, p5))
    reqs = [Request(0, "trial", 0, 0.0, 5, p5), Request(1, "accept", 0, 0.0005, 5, p5),
            Request(2, "trial", 0, 0.001, 7, p7), Request(3, "full", 0, 0.0015),
            Request(4, "full", 1, 0.0)]
    out = {r.rid: r for r in srv.run(reqs)}
    assert all(r.done is not None for r in out.values())
    x1 = x0.copy()
    x1[5 * s.blk:6 * s.blk] = p5
    eng = Engine(s, "cuda:0")
    Ea, ga = eng.outprod(x1)
    assert out[0].E == E_old                                   # the trial before: old basis
    assert abs(out[1].E - Ea) <= 1e-12 * max(1.0, abs(Ea))
    assert np.max(np.abs(out[1].g - ga)) <= 1e-9 * max(1.0, np.max(np.abs(ga)))
    assert out[3].E == out[1].E and np.array_equal(out[3].g, out[1].g)
    x2 = x1.copy()
    x2[7 * s.blk:8 * s.blk] = p7
    E2, H2, S2 = eng.energy(x2, with_hs=True)
    assert abs(out[2].E - E2) <= 1e-11 * max(1.0, abs(E2))
    w0 = eigh(H2, S2, eigvals_only=True, subset_by_index=[0, 0])[0]
    assert abs(out[2].Eeig - w0) <= 1e-9 * max(1.0, abs(w0))
    assert np.array_equal(srv.xs[0], x1)


def test_trials_io_equals_separate_calls():
    """necg_pool_trials_io (copies in and out around the trial and eigenvalue
    kernels, one call) against necg_pool_trials + necg_pool_eig on device
    buffers: bitwise the same E and Eeig for a batch of trials."""
    import ctypes as ct
    import torch
    from nbo_correlatedgaussians_amd import _lib
    from nbo_correlatedgaussians_amd.serve import Request, Server
    sx = _systems()
    srv = Server([s for s, _ in sx], [x for _, x in sx], "cuda:0", trial_streams=1,
                 eval_streams=1)
    rng = np.random.default_rng(9)
    reqs = []
    for i in range(7):
        si = int(rng.integers(len(sx)))
        k = int(rng.integers(srv.Ks[si]))
        iv1, _ = necg.generate_parameters(sx[si][0], 1, int(rng.integers(1 << 30)))
        reqs.append(Request(i, "trial", si, 0.0, k, iv1))
    n, P, kmax = len(reqs), srv.pmax, srv.kmax
    L = srv._launch_trials(0, reqs)
    L.event.synchronize()
    out_io = srv._tbuf[0]["h_out"][:2 * n].clone().numpy()
    dev = dict(device="cuda:0")
    tsys = torch.tensor([r.sys for r in reqs], dtype=torch.int32, **dev)
    tfun = torch.tensor([r.func for r in reqs], dtype=torch.int32, **dev)
    newp = torch.zeros(n * P, dtype=torch.float64, **dev)
    for i, r in enumerate(reqs):
        newp[i * P:i * P + r.newp.size] = torch.as_tensor(r.newp)
    ws = torch.empty(n * srv.wstride, dtype=torch.float64, **dev)
    rowH = torch.empty(n * kmax, dtype=torch.float64, **dev)
    rowS = torch.empty_like(rowH)
    bg = torch.empty(n * 18 * kmax, dtype=torch.float64, **dev)
    E = torch.empty(n, dtype=torch.float64, **dev)
    Ee = torch.empty_like(E)
    st = ct.c_void_p(torch.cuda.current_stream().cuda_stream)
    p = lambda t: ct.c_void_p(t.data_ptr())   # noqa: E731
    lib = _lib.load()
    assert lib.necg_pool_trials(srv._pool, n, p(tsys), p(tfun), p(newp), P, p(ws), p(E), p(rowH),
                                p(rowS), kmax, st) == 0
    assert lib.necg_pool_eig(srv._pool, n, p(tsys), p(tfun), p(rowH), p(rowS), kmax, p(E), p(bg),
                             p(Ee), st) == 0
    torch.cuda.synchronize()
    assert lib.necg_pool_status(srv._pool, st) == 0
    assert np.array_equal(out_io[:n], E.cpu().numpy())
    assert np.array_equal(out_io[n:], Ee.cpu().numpy())
    # a misaligned device input buffer is refused before anything is enqueued
    assert lib.necg_pool_trials_io(srv._pool, 1, p(srv._tbuf[0]["h_in"]),
                                   ct.c_void_p(srv._tbuf[0]["d_in"].data_ptr() + 4), P, p(ws),
                                   p(rowH), p(rowS), kmax, p(bg), p(E),
                                   p(srv._tbuf[0]["h_out"]), st) == _lib.NECG_ERR_ARG
