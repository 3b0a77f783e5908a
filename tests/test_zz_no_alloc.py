"""necg_outprod_noalloc through the C ABI: the allocation-free variant of
necg_outprod (workspace carved from the library's static device scratch)
gives the same E and g, bitwise, as necg_outprod on the same context, on two
calls in a row with the second one smaller (sizes whose ordinary tiling fits
the scratch); an evaluation larger than the scratch is refused with
NECG_ERR_SCRATCH and leaves the context working, and sizes in between take a
tiling that fits and agree with necg_outprod to rounding."""

import ctypes as ct

import numpy as np
import pytest

import nbo_correlatedgaussians_amd as necg

pytestmark = pytest.mark.gpu


def _call(lib, name, eng, x):
    import torch
    f = torch.full((1,), float("nan"), dtype=torch.float64, device="cuda:0")
    g = torch.full((x.numel(),), float("nan"), dtype=torch.float64, device="cuda:0")
    K = x.numel() // (eng.blk + 1)
    st = ct.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = getattr(lib, name)(eng._ctx, K, ct.c_void_p(x.data_ptr()), ct.c_void_p(f.data_ptr()),
                            ct.c_void_p(g.data_ptr()), st)
    return rc, f, g


@pytest.mark.parametrize("name,Ks", [("LiH", (200, 70)), ("He", (300, 40)), ("C", (40, 9))])
def test_noalloc_equals_outprod(name, Ks):
    """LiH and He take the two-sided pair pass, the n = 6 system the
    per-ordering one."""
    import torch
    from conftest import system_by_name
    from nbo_correlatedgaussians_amd import _lib
    from nbo_correlatedgaussians_amd.engine import Engine
    lib = _lib.load()
    s = system_by_name(name)
    eng = Engine(s, "cuda:0")
    xs = [torch.as_tensor(necg.pack_x(*necg.generate_parameters(s, K, 60 + K)), device="cuda:0")
          for K in Ks]
    ref = []
    for x in xs:                      # the existing entry point first
        rc, f, g = _call(lib, "necg_outprod", eng, x)
        assert rc == 0
        eng.check()
        ref.append((f.item(), g.cpu().numpy()))
    for x, (Er, gr), K in zip(xs, ref, Ks):   # then two allocation-free calls, K falling
        print(f"{name} K={K}: necg_outprod_noalloc", flush=True)
        rc, f, g = _call(lib, "necg_outprod_noalloc", eng, x)
        assert rc == 0, _lib.NecgError(rc)
        eng.check()
        assert f.item() == Er, (K, f.item(), Er)
        np.testing.assert_array_equal(g.cpu().numpy(), gr)
    # the context's own workspace and results are untouched
    E0, g0 = eng.outprod(xs[0].cpu().numpy())
    assert E0 == ref[0][0] and np.array_equal(g0, ref[0][1])


def test_noalloc_refuses_what_does_not_fit():
    import torch
    from nbo_correlatedgaussians_amd import _lib
    from nbo_correlatedgaussians_amd.engine import Engine
    lib = _lib.load()
    s = necg.lih_system()
    eng = Engine(s, "cuda:0")
    big = torch.as_tensor(necg.pack_x(*necg.generate_parameters(s, 1500, 3)), device="cuda:0")
    print("LiH K=1500: refused", flush=True)
    rc, f, g = _call(lib, "necg_outprod_noalloc", eng, big)
    assert rc == _lib.NECG_ERR_SCRATCH
    eng.check()
    small = torch.as_tensor(necg.pack_x(*necg.generate_parameters(s, 64, 4)), device="cuda:0")
    rc, f, g = _call(lib, "necg_outprod_noalloc", eng, small)
    assert rc == 0
    eng.check()
    Er, gr = eng.outprod(small.cpu().numpy())
    assert f.item() == Er and np.array_equal(g.cpu().numpy(), gr)
    fits = []
    tried = (300, 400, 600, 900, 1200)
    for K in tried:    # where the 64 MiB scratch ends for LiH
        x = torch.as_tensor(necg.pack_x(*necg.generate_parameters(s, K, 5)), device="cuda:0")
        print(f"LiH K={K}: fits or refused", flush=True)
        rc, f, g = _call(lib, "necg_outprod_noalloc", eng, x)
        assert rc in (0, _lib.NECG_ERR_SCRATCH)
        eng.check()
        if rc == 0:
            fits.append(K)
            # where the ordinary call's tiling does not fit, the allocation-free
            # one takes a smaller one (other term groups or ket tiles): the same
            # sums in another grouping
            Er, gr = eng.outprod(x.cpu().numpy())
            assert abs(f.item() - Er) <= 1e-13 * abs(Er), (K, f.item(), Er)
            np.testing.assert_allclose(g.cpu().numpy(), gr, rtol=1e-9,
                                       atol=1e-11 * np.max(np.abs(gr)))
    print(f"LiH K that fit the scratch: {fits}", flush=True)
    assert fits == list(tried[:len(fits)])          # a prefix of the sizes tried
    assert 900 in fits and 1200 not in fits
