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
    assert "1048576 cases (65536 one-run), 0 mismatches" in r.stdout


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_seam_unit_order(tmp_path):
    """Host check of the seam merge's unit order (ss_seam.hpp): a bijection onto
    the block-boundary tile edges, one block seam per aligned run of units."""
    exe = tmp_path / "seam_units_check"
    src = os.path.join(REPO, "tests", "seam_units_check.cpp")
    inc = os.path.join(REPO, "sparse-seam_amd", "csrc")
    subprocess.run([HIPCC, "-x", "hip", "--offload-host-only", "-O2", "-std=c++17", "-I", inc, "-o", str(exe), src],
                   check=True, capture_output=True, text=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "4096 grids, 0 failed" in r.stdout


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_autotune_admission_rule(tmp_path):
    """Host check of the autotuner's admission rule (ss_admit.hpp) over every
    ok / digest assignment of up to six candidates."""
    exe = tmp_path / "admit_check"
    src = os.path.join(REPO, "tests", "admit_check.cpp")
    inc = os.path.join(REPO, "sparse-seam_amd", "csrc")
    subprocess.run([HIPCC, "-x", "hip", "--offload-host-only", "-O2", "-std=c++17", "-I", inc, "-o", str(exe), src],
                   check=True, capture_output=True, text=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failed" in r.stdout


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_row_edges_partition(tmp_path):
    """Host check of the row tile stage's union edges (ss_rows.hpp): on random
    64 x 64 tiles the run union-find equals a breadth-first labelling of the
    8-connected zones; the run fills equal a per-bit walk."""
    exe = tmp_path / "rowtile_check"
    src = os.path.join(REPO, "tests", "rowtile_check.cpp")
    inc = os.path.join(REPO, "sparse-seam_amd", "csrc")
    subprocess.run([HIPCC, "-x", "hip", "--offload-host-only", "-O2", "-std=c++17", "-I", inc, "-o", str(exe), src],
                   check=True, capture_output=True, text=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "fills: 0 mismatches; 3000 tiles, 0 mismatches" in r.stdout
