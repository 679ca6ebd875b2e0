"""The C++ command-line gate (sparse-seam_amd/ss_harness): exit codes, CSV,
and the reference's oracle gate on the 206 golden triples."""
import csv
import json
import os
import subprocess
import sys

import pytest

from sscommon import GOLDEN, REPO, SUITES, oracle_quantize, oracle_seam_profile, read_pgm

HARNESS = os.path.join(REPO, "sparse-seam_amd", "ss_harness")
AGGREGATE = os.path.join(REPO, "sparse-seam_amd", "host", "ss_aggregate.py")


def run(args, timeout=300):
    return subprocess.run([HARNESS] + args, capture_output=True, text=True, timeout=timeout)


def test_usage_and_argument_errors():
    assert os.path.exists(HARNESS)
    assert run([]).returncode == 2
    assert run([GOLDEN, "--bogus"]).returncode == 2
    r = run([GOLDEN, "--skip-expected"])          # no --output
    assert r.returncode == 3 and "Missing --output" in r.stderr
    r = run([GOLDEN, "--output=/tmp/x", "--levels=1"])
    assert r.returncode == 3
    assert run(["--seam-profile"]).returncode == 2
    assert run(["--seam-profile", GOLDEN, "--bogus"]).returncode == 2
    r = run(["--seam-profile", GOLDEN])           # no --output
    assert r.returncode == 3 and "Missing --output" in r.stderr
    r = run(["--seam-profile", GOLDEN, "--output=/tmp/x.csv", "--tiles=96"])
    assert r.returncode == 3 and "Tile size" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("suite", SUITES)
def test_golden_gate(tmp_path, suite):
    root = os.path.join(GOLDEN, suite)
    out = tmp_path / "out"
    r = run([os.path.join(root, "inputs"), "--output=%s" % out,
             "--expected-max=%s" % os.path.join(root, "expected_max"),
             "--expected-min=%s" % os.path.join(root, "expected_min"), "--warmup=1", "--reps=2"])
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "[SS-HARNESS-VALIDATION-OK]" in r.stdout and "mismatches=0" in r.stdout
    rows = list(csv.DictReader(open(out / "ss_harness.csv")))
    assert len(rows) == len(os.listdir(os.path.join(root, "inputs")))
    assert all(row["oracle_max_mismatches"] == "0" and row["oracle_min_mismatches"] == "0" for row in rows)
    # the cross-route columns come from a real Global-UF run on every image
    assert all(row["cross_max_mismatches"] == "0" and row["cross_min_mismatches"] == "0" for row in rows)
    assert all(float(row["median_iterations"]) == 1.0 for row in rows)
    assert "cross_mismatch=0" in r.stdout and "cross_mismatches=0" in r.stdout
    assert "cross_route=Global-UF" in (out / "environment.txt").read_text()
    val = list(csv.DictReader(open(out / "validation.csv")))
    assert list(val[0].keys()) == ["image", "width", "height", "tiles", "fragment_capacity", "active_fragments",
                                   "auxiliary_bytes", "auxiliary_bytes_per_pixel", "elapsed_ms", "oracle_checked",
                                   "max_mismatches", "min_mismatches"]
    assert len(val) == len(rows) and all(v["oracle_checked"] == "1" and v["max_mismatches"] == "0" for v in val)
    agg = subprocess.run([sys.executable, AGGREGATE, str(out / "ss_harness.csv")], capture_output=True, text=True)
    assert agg.returncode == 0, agg.stdout + agg.stderr


@pytest.mark.gpu
def test_gate_detects_mismatch(tmp_path):
    root = os.path.join(GOLDEN, "oracle_v0_1")
    r = run([os.path.join(root, "inputs"), "--output=%s" % (tmp_path / "o"),
             "--expected-max=%s" % os.path.join(root, "expected_min"),   # swapped on purpose
             "--expected-min=%s" % os.path.join(root, "expected_max"), "--warmup=0", "--reps=1"])
    assert r.returncode == 1


@pytest.mark.gpu
def test_autotune_and_levels(tmp_path):
    root = os.path.join(GOLDEN, "oracle_v0_3_sparse_seam_adversarial")
    r = run([os.path.join(root, "inputs"), "--output=%s" % (tmp_path / "o"), "--skip-expected",
             "--autotune", "--levels=4", "--warmup=1", "--reps=2"])
    assert r.returncode == 0, r.stderr[-2000:]
    assert "[SS-AUTOTUNE]" in r.stdout and "[SS-HARNESS-MEASUREMENT-OK]" in r.stdout


@pytest.mark.gpu
def test_seam_profile_mode(tmp_path):
    """Seam-sparsity rows with the reference profiler's columns; counts equal
    the oracle restatement on the quantised inputs."""
    root = os.path.join(GOLDEN, "oracle_v0_3_sparse_seam_adversarial", "inputs")
    out = tmp_path / "profile.csv"
    r = run(["--seam-profile", root, "--output=%s" % out, "--tiles=16,128", "--levels=256,4"])
    assert r.returncode == 0, r.stderr[-2000:]
    assert "[SS-SEAM-PROFILE-OK]" in r.stdout
    rows = list(csv.DictReader(open(out)))
    names = sorted(os.listdir(root))
    assert len(rows) == len(names) * 4
    assert list(rows[0].keys()) == [
        "image", "width", "height", "megapixels", "levels", "tile", "total_tile_fragments",
        "seam_fragments", "interior_fragments", "seam_fragment_fraction", "seam_pixels",
        "seam_pixel_fraction", "equal_forward_edges", "equal_seam_edges", "seam_edge_fraction",
        "seam_fragments_per_mpix", "equal_seam_edges_per_mpix", "diagnostic_ms"]
    for row in rows[:8]:
        img = oracle_quantize(read_pgm(os.path.join(root, row["image"])), int(row["levels"]))
        want = oracle_seam_profile(img, int(row["tile"]))
        assert int(row["total_tile_fragments"]) == want["fragments"]
        assert int(row["seam_fragments"]) == want["seam_fragments"]
        assert int(row["seam_pixels"]) == want["seam_pixels"]
        assert int(row["equal_forward_edges"]) == want["equal_edges"]
        assert int(row["equal_seam_edges"]) == want["seam_edges"]
        assert int(row["interior_fragments"]) == want["fragments"] - want["seam_fragments"]


def test_aggregate_sum_px_over_sum_median(tmp_path):
    """Throughput = total pixels / sum of per-image median times (the
    reference aggregator's rule), mismatches summed, exit 1 on a mismatch."""
    cols = ["algorithm", "image", "width", "height", "median_compute_ms", "median_h2h_ms", "auxiliary_bytes",
            "oracle_max_mismatches", "oracle_min_mismatches", "cross_max_mismatches", "cross_min_mismatches",
            "tile"]
    path = tmp_path / "ss_harness.csv"
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(cols)
        w.writerow(["A", "a.pgm", 1000, 1000, 2.0, 8.0, 500000, 0, 0, 0, 0, 64])
        w.writerow(["A", "b.pgm", 2000, 1000, 6.0, 12.0, 1000000, 0, 0, 0, 0, 64])
    r = subprocess.run([sys.executable, AGGREGATE, str(path)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    a = json.loads(r.stdout.strip().splitlines()[-1])["A/T64"]
    assert a["pixels"] == 3_000_000 and abs(a["compute_mpix_s"] - 3.0 / 0.008) < 1e-9
    assert abs(a["h2h_mpix_s"] - 3.0 / 0.020) < 1e-9 and abs(a["aux_Bpp"] - 0.5) < 1e-12
    with open(path, "a", newline="") as f:
        csv.writer(f).writerow(["A", "c.pgm", 10, 10, 1.0, 1.0, 100, 1, 0, 0, 0, 64])
    assert subprocess.run([sys.executable, AGGREGATE, str(path)], capture_output=True).returncode == 1


def test_png_reader(tmp_path):
    """PNG decoding of the harness (--convert, no GPU): the reference's own PNG
    fixtures equal its PGM inputs; PIL-written gray, gray + alpha, palette, RGB
    and RGBA images decode to the ITU-R 601 gray of their pixels; 16-bit and
    interlaced files are refused."""
    Image = pytest.importorskip("PIL.Image")
    import numpy as np
    src = os.path.join(GOLDEN, "oracle_v0_1", "inputs_png")
    out = tmp_path / "conv"
    r = run(["--convert", src, "--output=%s" % out])
    assert r.returncode == 0, r.stderr
    for name in os.listdir(src):
        stem = name[:-4]
        assert np.array_equal(read_pgm(str(out / (stem + ".pgm"))),
                              read_pgm(os.path.join(GOLDEN, "oracle_v0_1", "inputs", stem + ".pgm"))), name
    rng = np.random.default_rng(3)
    d = tmp_path / "png"
    d.mkdir()
    rgb = rng.integers(0, 256, (37, 53, 3), dtype=np.uint8)
    rgb[5:20, 10:30] = 200                       # flat areas exercise every filter type
    gray601 = ((rgb[..., 0].astype(np.uint32) * 9798 + rgb[..., 1].astype(np.uint32) * 19235 +
                rgb[..., 2].astype(np.uint32) * 3735 + 16384) >> 15).astype(np.uint8)
    g = rng.integers(0, 256, (41, 29), dtype=np.uint8)
    want = {}
    Image.fromarray(g, "L").save(d / "gray.png")
    want["gray"] = g
    Image.fromarray(np.dstack([g, 255 - g]), "LA").save(d / "gray_alpha.png")
    want["gray_alpha"] = g
    Image.fromarray(rgb, "RGB").save(d / "rgb.png")
    want["rgb"] = gray601
    Image.fromarray(np.dstack([rgb, rng.integers(0, 256, rgb.shape[:2], dtype=np.uint8)]), "RGBA").save(d / "rgba.png")
    want["rgba"] = gray601
    pal = Image.fromarray(rgb, "RGB").quantize(colors=17)
    pal.save(d / "palette.png")
    prgb = np.array(pal.convert("RGB")).astype(np.uint32)
    want["palette"] = ((prgb[..., 0] * 9798 + prgb[..., 1] * 19235 + prgb[..., 2] * 3735 + 16384) >> 15).astype(np.uint8)
    out2 = tmp_path / "conv2"
    r = run(["--convert", str(d), "--output=%s" % out2])
    assert r.returncode == 0, r.stderr
    for name, arr in want.items():
        assert np.array_equal(read_pgm(str(out2 / (name + ".pgm"))), arr), name
    bad = tmp_path / "bad"
    bad.mkdir()
    Image.fromarray(g.astype(np.uint16) * 257).save(bad / "deep.png")
    r = run(["--convert", str(bad), "--output=%s" % (tmp_path / "conv3")])
    assert r.returncode == 3 and "Unsupported PNG" in r.stderr


@pytest.mark.gpu
def test_golden_gate_png_inputs(tmp_path):
    """The gate on the reference's PNG copies of its first golden suite."""
    root = os.path.join(GOLDEN, "oracle_v0_1")
    r = run([os.path.join(root, "inputs_png"), "--output=%s" % (tmp_path / "o"),
             "--expected-max=%s" % os.path.join(root, "expected_max"),
             "--expected-min=%s" % os.path.join(root, "expected_min"), "--warmup=0", "--reps=1"])
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "[SS-HARNESS-VALIDATION-OK] images=34 mismatches=0" in r.stdout
