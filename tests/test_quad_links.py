"""Host check of the tile stage's SWAR quad labelling (ss_quad.hpp) against
the per-run definition over all 2^20 nibble inputs.  Built with hipcc as
host code; needs no GPU."""
import os
import shutil
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_quad_links_exhaustive(tmp_path):
    exe = tmp_path / "quad_links_check"
    src = os.path.join(REPO, "tests", "quad_links_check.cpp")
    inc = os.path.join(REPO, "sparse-seam_amd", "csrc")
    subprocess.run([HIPCC, "-x", "hip", "--offload-host-only", "-O2", "-std=c++17", "-I", inc, "-o", str(exe), src],
                   check=True, capture_output=True, text=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "1048576 cases, 0 mismatches" in r.stdout
