"""Host check of the lower-triangle tiling of the two-sided pair pass (CPU):
tests/native/tri_tiles_check.cpp, built with hipcc against the library's own
headers, checks for K = 1..300 and larger sizes and 12 ket-tile sizes that
every unordered pair of functions is evaluated exactly once and that each
function reduces exactly the tiles that wrote its partials; the same for the
kernels' lists whose last round is split into half tiles (tri_tile_list), over
a range of CU counts."""

import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def _hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    return None


def test_tri_tiles_cover_each_pair_once(tmp_path):
    hipcc = _hipcc()
    if hipcc is None:
        pytest.skip("hipcc not available")
    exe = str(tmp_path / "tri_tiles_check")
    src = os.path.join(HERE, "native", "tri_tiles_check.cpp")
    subprocess.run([hipcc, "-O2", "-std=c++17", "-o", exe, src], check=True,
                   stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stdout[-2000:]
