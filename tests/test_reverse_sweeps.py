"""Host check of the structured reverse product's hand-written sweeps (Riemann
fluxes, viscous terms, SST cell terms) against forward-mode duals: compiles
tests/reverse_sweeps_test.cpp for the host with hipcc and runs it."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not shutil.which(HIPCC) and not os.path.exists(HIPCC), reason="hipcc not found")
def test_reverse_sweeps_match_forward_duals(tmp_path):
    exe = str(tmp_path / "reverse_sweeps_test")
    subprocess.run([HIPCC, "-O1", "-std=c++17", "-x", "hip", "--offload-host-only",
                    os.path.join(HERE, "reverse_sweeps_test.cpp"), "-o", exe], check=True, capture_output=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
