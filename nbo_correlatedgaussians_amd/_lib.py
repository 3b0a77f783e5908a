"""ctypes binding of ``libnecg_hip.so`` (the C ABI declared in include/necg.h).

The product path has no CPU fallback: if the library or a GPU is missing,
loading raises."""

from __future__ import annotations

import ctypes as ct
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# NECG_LIB_PATH selects another build of the same library (e.g. a tuning variant)
LIB_PATH = os.environ.get("NECG_LIB_PATH") or os.path.join(HERE, "libnecg_hip.so")

# status codes (include/necg.h)
NECG_OK = 0
NECG_ERR_ARG = 1
NECG_ERR_UNSUPPORTED = 2
NECG_ERR_DEVICE = 3
NECG_ERR_OOM = 4
NECG_ERR_NOT_PD = 5
NECG_ERR_CSC = 6
NECG_ERR_NONFINITE = 7
NECG_ERR_STALE = 8
NECG_ERR_SCRATCH = 9

EXPORTS = ("necg_create", "necg_destroy", "necg_outprod", "necg_outprod_noalloc", "necg_energy",
           "necg_kinetic_potential", "necg_dipole", "necg_set_dipole_mode", "necg_device_visible",
           "necg_set_pair_mode", "necg_pair_form", "necg_fp64_fma_rate",
           "necg_energy_fwd", "necg_energy_bwd",
           "necg_energy_vjp", "necg_matrices", "necg_matel", "necg_gradel",
           "necg_dhds", "necg_status", "necg_set_tiling", "necg_set_launch_cus", "necg_get_tiling",
           "necg_pin_workspace", "necg_batch_pin", "necg_set_fold",
           "necg_timing", "necg_strerror", "necg_batch_create", "necg_batch_destroy",
           "necg_batch_layout", "necg_batch_outprod", "necg_batch_matrices",
           "necg_batch_set_tiling", "necg_batch_status", "necg_batch_outprod_masked",
           "necg_batch_pool", "necg_pool_create",
           "necg_pool_destroy", "necg_pool_sizes", "necg_pool_trials", "necg_pool_status",
           "necg_pool_set_factors", "necg_pool_eig", "necg_pool_trials_io",
           "necg_batch_set_launch_cus")


class NecgError(RuntimeError):
    def __init__(self, code: int, where: str = ""):
        self.code = code
        msg = _LIB.necg_strerror(code).decode() if _LIB is not None else f"status {code}"
        super().__init__(f"{where}: {msg} (status {code})" if where else msg)


class NotPositiveDefinite(NecgError, ArithmeticError):
    pass


class ZeroNorm(NecgError, ZeroDivisionError):
    pass


class NonFinite(NecgError, FloatingPointError):
    pass


def raise_for(code: int, where: str = "") -> None:
    if code == NECG_OK:
        return
    cls = {NECG_ERR_NOT_PD: NotPositiveDefinite, NECG_ERR_CSC: ZeroNorm,
           NECG_ERR_NONFINITE: NonFinite}.get(code, NecgError)
    raise cls(code, where)


class NecgSystem(ct.Structure):
    _fields_ = [("n", ct.c_int32), ("nst", ct.c_int32), ("efs", ct.c_double),
                ("mass", ct.c_void_p), ("ncv", ct.c_void_p), ("scvv", ct.c_void_p),
                ("sml", ct.c_void_p), ("sms", ct.c_void_p)]


_LIB = None


def library_path() -> str:
    return LIB_PATH


def load() -> ct.CDLL:
    """Load the HIP library; raises if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: run `python -m nbo_correlatedgaussians_amd.build`")
    lib = ct.CDLL(LIB_PATH, mode=ct.RTLD_GLOBAL)
    P, I, D = ct.c_void_p, ct.c_int32, ct.c_double
    sig = {
        "necg_create": [ct.POINTER(NecgSystem), ct.c_int, ct.POINTER(P)],
        "necg_destroy": [P],
        "necg_outprod": [P, I, P, P, P, P],
        "necg_outprod_noalloc": [P, I, P, P, P, P],
        "necg_energy": [P, I, P, P, P, P, P],
        "necg_energy_vjp": [P, I, P, P, P, P],
        "necg_matrices": [P, I, P, P, P, P],
        "necg_kinetic_potential": [P, I, P, P, P, P],
        "necg_dipole": [P, I, P, P, P],
        "necg_set_dipole_mode": [P, I],
        "necg_set_pair_mode": [P, I],
        "necg_pair_form": [P, ct.POINTER(I), ct.POINTER(I)],
        "necg_fp64_fma_rate": [I, I, ct.POINTER(D)],
        "necg_device_visible": [P],
        "necg_energy_fwd": [P, I, P, P, ct.POINTER(ct.c_int64), P],
        "necg_energy_bwd": [P, I, ct.c_int64, P, P, P],
        "necg_matel": [P, I, P, I, P, P, P],
        "necg_gradel": [P, I, P, I, P, P, P, P],
        "necg_dhds": [P, I, P, P, P, P],
        "necg_status": [P, P],
        "necg_set_tiling": [P, I, I, I],
        "necg_set_launch_cus": [P, I],
        "necg_get_tiling": [P, ct.POINTER(I), ct.POINTER(I), ct.POINTER(I)],
        "necg_pin_workspace": [P, I],
        "necg_set_fold": [P, I],
        "necg_batch_pin": [P, I],
        "necg_timing": [P, I, ct.POINTER(D), ct.POINTER(ct.c_int64)],
        "necg_batch_create": [P, I, P, ct.c_int, ct.POINTER(P)],
        "necg_batch_destroy": [P],
        "necg_batch_layout": [P, P],
        "necg_batch_outprod": [P, P, P, P, P],
        "necg_batch_matrices": [P, I, P, P, P],
        "necg_batch_set_tiling": [P, D, I],
        "necg_batch_status": [P, P],
        "necg_batch_outprod_masked": [P, P, P, P, P, P],
        "necg_batch_pool": [P, ct.POINTER(P)],
        "necg_pool_create": [P, I, ct.POINTER(P)],
        "necg_pool_destroy": [P],
        "necg_pool_sizes": [P, ct.POINTER(I), ct.POINTER(I)],
        "necg_pool_trials": [P, I, P, P, P, I, P, P, P, P, I, P],
        "necg_pool_set_factors": [P, I, P, P],
        "necg_pool_eig": [P, I, P, P, P, P, I, P, P, P, P],
        "necg_pool_trials_io": [P, I, P, P, I, P, P, P, I, P, P, P, P],
        "necg_batch_set_launch_cus": [P, I],
        "necg_pool_status": [P, P],
    }
    for name, args in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = ct.c_int
    lib.necg_strerror.argtypes = [ct.c_int]
    lib.necg_strerror.restype = ct.c_char_p
    _LIB = lib
    return lib
