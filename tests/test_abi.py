"""The C-ABI library loads and exports every entry point declared in
include/necg.h (CPU: no compute calls)."""

import ctypes as ct
import os
import re

from nbo_correlatedgaussians_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    with open(os.path.join(ROOT, "include", "necg.h"), encoding="utf-8") as fh:
        src = fh.read()
    return sorted(set(re.findall(r"\b(necg_[a-z_0-9]+)\s*\(", src)))


def test_header_lists_the_entry_points():
    names = declared()
    assert set(names) == set(_lib.EXPORTS), names


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    for name in declared():
        assert hasattr(lib, name), name
    assert lib.necg_strerror(0) == b"ok"
    assert b"positive definite" in lib.necg_strerror(_lib.NECG_ERR_NOT_PD)


def test_create_rejects_bad_arguments_without_gpu_work():
    lib = _lib.load()
    out = ct.c_void_p()
    assert lib.necg_create(None, 0, ct.byref(out)) == _lib.NECG_ERR_ARG
    desc = _lib.NecgSystem(9, 1, 0.0, None, None, None, None, None)
    assert lib.necg_create(ct.byref(desc), 0, ct.byref(out)) == _lib.NECG_ERR_UNSUPPORTED
    assert lib.necg_destroy(None) == _lib.NECG_ERR_ARG
    assert lib.necg_set_tiling(None, 1, 1, 1) == _lib.NECG_ERR_ARG
    assert lib.necg_set_launch_cus(None, 8) == _lib.NECG_ERR_ARG
    r = ct.c_double(-1.0)
    assert lib.necg_fp64_fma_rate(0, 0, ct.byref(r)) == _lib.NECG_ERR_ARG
    assert lib.necg_fp64_fma_rate(0, 10, None) == _lib.NECG_ERR_ARG
    assert r.value == -1.0
