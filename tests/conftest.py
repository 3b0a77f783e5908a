import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU (MI355X)")


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import oracle
    oracle.build()
    return oracle


def golden(name):
    import numpy as np
    return np.load(os.path.join(ROOT, "tests", "golden", f"golden_{name}.npz"))


def system_by_name(name):
    import nbo_correlatedgaussians_amd as necg
    return {"LiH": necg.lih_system, "He": necg.he_system, "H2": necg.h2_system,
            "Be": necg.be_system, "C": necg.c_system}[str(name)]()


FULL_CASES = ["lih_oldquit", "lih_k12_s0", "he_k8_s3", "h2_k8_s4", "lih_k32_s2"]
