"""The exact fast path of the block / voxel index (tsdf_math.hpp div_index)
agrees with the reference arithmetic on >20M random and adversarial inputs
(built from tests/cpp/test_div_index.cpp as host code)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("hipcc") is None and not os.path.exists("/opt/rocm/bin/hipcc"),
                    reason="hipcc not available")
def test_div_index_matches_reference(tmp_path):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    exe = tmp_path / "test_div_index"
    subprocess.run([hipcc, "-O2", "-std=c++17", "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
                    "-o", str(exe), os.path.join(ROOT, "tests", "cpp", "test_div_index.cpp")],
                   check=True, capture_output=True)
    r = subprocess.run([str(exe), "20000000"], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "agree with the reference arithmetic" in r.stdout
