"""Batched evaluation of uneven systems: every system's (E, g) equals its
single-system outprod (and the reference's golden values where they exist),
a captured HIP graph replays the same numbers, and autotuned tilings agree."""

import numpy as np
import pytest

import nbo_correlatedgaussians_amd as necg
from conftest import golden, system_by_name

pytestmark = pytest.mark.gpu


def _single(sysd, x):
    from nbo_correlatedgaussians_amd.engine import Engine
    return Engine(sysd, "cuda:0").outprod(x)


def _lih_without_an_inverse():
    """LiH less one term that is not its own inverse: the two-sided pair form
    does not apply, so a batch holding it runs both forms for n = 5."""
    s = necg.lih_system()
    p = next(q for q in range(s.nst) if not np.array_equal(s.sml[q], s.sml[q].T))
    keep = [q for q in range(s.nst) if q != p]
    t = necg.System(n=s.n, nst=len(keep), efs=s.efs, mass=s.mass.copy(), ncv=s.ncv.copy(),
                    scvv_raw=s.scvv_raw[keep].copy(), sml=s.sml[keep].copy(),
                    sms=s.sms[keep].copy(), name="LiH-23")
    t.validate()
    return t


def _mixed_small():
    cases = []
    for name in ("lih_k12_s0", "he_k8_s3", "h2_k8_s4"):
        d = golden(name)
        cases.append((system_by_name(d["system"]), d["x"], d))
    for name, K, seed in (("LiH", 70, 11), ("He", 130, 12), ("H2", 200, 13), ("LiH", 1, 14),
                          ("Be", 40, 15), ("C", 33, 16)):
        s = system_by_name(name)
        iv, C = necg.generate_parameters(s, K, seed)
        cases.append((s, necg.pack_x(iv, C), None))
    s = _lih_without_an_inverse()
    iv, C = necg.generate_parameters(s, 90, 17)
    cases.append((s, necg.pack_x(iv, C), None))
    return cases


def test_batch_matches_single_and_golden():
    from nbo_correlatedgaussians_amd.batch import Batch
    cases = _mixed_small()
    systems = [c[0] for c in cases]
    Ks = [c[1].size // (c[0].blk + 1) for c in cases]
    b = Batch(systems, Ks, "cuda:0")
    out = b.outprod([c[1] for c in cases])
    for (s, x, d), (E, g) in zip(cases, out):
        Es, gs = _single(s, x)
        assert abs(E - Es) <= 1e-13 * max(1.0, abs(Es))
        assert np.max(np.abs(g - gs)) <= 1e-11 * max(1.0, np.max(np.abs(gs)))
        if d is not None:
            assert abs(E - float(d["E"])) <= 1e-12 * max(1.0, abs(float(d["E"])))
            assert np.max(np.abs(g - d["g"])) <= 1e-9 * max(1.0, np.max(np.abs(d["g"])))
    H, S = b.matrices(0)
    d = golden("lih_k12_s0")
    assert np.max(np.abs(H - d["H"])) <= 1e-12 * np.max(np.abs(d["H"]))


def test_batch_graph_replays_and_tilings():
    import torch
    from nbo_correlatedgaussians_amd.batch import Batch, BatchGraph, autotune
    cases = _mixed_small()
    systems = [c[0] for c in cases]
    Ks = [c[1].size // (c[0].blk + 1) for c in cases]
    b = Batch(systems, Ks, "cuda:0")
    xa = b.pack([c[1] for c in cases])
    ref = b.outprod([c[1] for c in cases])
    G = BatchGraph(b, xa)
    for rep in range(4):
        f, g = G.replay()
        torch.cuda.synchronize()
        # eager work between replays
        b.outprod_t(torch.as_tensor(xa, device="cuda:0"))
        b.check()
        fh, gh = f.cpu().numpy(), g.cpu().numpy()
        for i, (E, gi) in enumerate(ref):
            assert fh[i] == E, (rep, i)
            np.testing.assert_array_equal(b.unpack(gh)[i], gi)
    # new parameters written in place are picked up by the replay
    x2 = xa.copy()
    x2[b.offsets[1] + 2] *= 1.05
    G.x.copy_(torch.as_tensor(x2, device="cuda:0"))
    f, g = G.replay()
    E1, _ = _single(systems[1], x2[b.offsets[1]:b.offsets[2]])
    assert abs(f[1].item() - E1) <= 1e-13 * max(1.0, abs(E1))
    from nbo_correlatedgaussians_amd import _lib
    with pytest.raises(_lib.NecgError):     # the graph pins the work lists
        b.set_tiling(1e6, 0)
    G.release()
    best, res = autotune(b, xa, item_flops=(1e6, 3e7, 3e8), reps=1)
    for (E, gi), (E2, g2) in zip(ref, b.outprod([c[1] for c in cases])):
        assert abs(E - E2) <= 1e-13 * max(1.0, abs(E))
        assert np.max(np.abs(gi - g2)) <= 1e-11 * max(1.0, np.max(np.abs(gi)))


def test_launch_cus_keeps_results():
    """Persistent pair launches sized for fewer CUs (as a serving loop keeps
    CUs for its trials): every partial keeps its slot, so E and g of whole and
    masked evaluations are bitwise those of the full-width launches; a pin
    refuses the re-plan."""
    import torch
    from nbo_correlatedgaussians_amd import _lib
    from nbo_correlatedgaussians_amd.batch import Batch, BatchGraph
    cases = _mixed_small()
    systems = [c[0] for c in cases]
    Ks = [c[1].size // (c[0].blk + 1) for c in cases]
    b = Batch(systems, Ks, "cuda:0")
    xs = [c[1] for c in cases]
    xd = torch.as_tensor(b.pack(xs), device="cuda:0")
    mask = torch.zeros(len(systems), dtype=torch.int32, device="cuda:0")
    mask[1] = mask[4] = 1
    ref = b.outprod(xs)
    fm, gm = b.outprod_t(xd, active=mask)
    ref_m = (fm.cpu().numpy().copy(), gm.cpu().numpy().copy())
    for ncu in (200, 17, 1):
        b.set_launch_cus(ncu)
        for (E, gi), (E2, g2) in zip(ref, b.outprod(xs)):
            assert E == E2 and np.array_equal(gi, g2), ncu
        fm, gm = b.outprod_t(xd, active=mask)
        b.check()
        for i in (1, 4):
            lo, hi = b.offsets[i], b.offsets[i + 1]
            assert fm[i].item() == ref_m[0][i], (ncu, i)
            assert np.array_equal(gm[lo:hi].cpu().numpy(), ref_m[1][lo:hi]), (ncu, i)
    b.set_launch_cus(0)
    G = BatchGraph(b, b.pack(xs))
    with pytest.raises(_lib.NecgError):
        b.set_launch_cus(100)
    G.release()


def test_mixed_64_systems_config():
    """64 systems (He/H2/LiH, K log-uniform in [32, 1024], seeds 100..163):
    a sample of systems against their single-system evaluation."""
    from nbo_correlatedgaussians_amd.batch import Batch, mixed_systems
    systems, Ks, xs = mixed_systems()
    assert len(systems) == 64 and min(Ks) >= 32 and max(Ks) <= 1024
    b = Batch(systems, Ks, "cuda:0")
    out = b.outprod(xs)
    assert all(np.isfinite(E) and np.all(np.isfinite(g)) for E, g in out)
    rng = np.random.default_rng(0)
    for i in rng.choice(64, 6, replace=False):
        Es, gs = _single(systems[i], xs[i])
        E, g = out[i]
        assert abs(E - Es) <= 1e-12 * max(1.0, abs(Es)), i
        assert np.max(np.abs(g - gs)) <= 1e-10 * max(1.0, np.max(np.abs(gs))), i
