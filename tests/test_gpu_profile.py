"""Seam-sparsity counters of the HIP path (ss_seam_profile_u8, through the C
ABI) against the oracle restatement of the reference profiler, exact."""
import sys

import numpy as np
import pytest
import torch

from sscommon import golden_cases, oracle_seam_profile
from ssx import sparse_seam as ss
from ssx import ssgen

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def profile(img_np, tile):
    img = torch.as_tensor(np.ascontiguousarray(img_np)).cuda()
    out = ss.seam_profile(img, tile)
    torch.cuda.synchronize()
    return out


@pytest.mark.parametrize("seed", range(8))
def test_profile_random(seed):
    rng = np.random.default_rng(500 + seed)
    h, w = int(rng.integers(1, 300)), int(rng.integers(1, 300))
    levels = int(rng.choice([2, 4, 16, 256]))
    img = ssgen.quantize(torch.as_tensor(rng.integers(0, 256, (h, w), dtype=np.uint8)), levels).numpy()
    if seed % 2:
        img = np.repeat(np.repeat(img[: -(-h // 3), : -(-w // 3)], 3, 0), 3, 1)[:h, :w].copy()
    for tile in (16, 32, 64, 128):
        say("profile random", seed, (h, w), levels, tile)
        assert profile(img, tile) == oracle_seam_profile(img, tile), (seed, tile)


def test_profile_golden_sample():
    cases = list(golden_cases())[::17]
    for suite, name, img, _, _, _ in cases:
        for tile in (16, 64, 128):
            say("profile golden", name, tile)
            assert profile(img, tile) == oracle_seam_profile(img, tile), (name, tile)


def test_profile_sweep_1024():
    for name, img in ssgen.sweep_images(1024, 0):
        if name not in ("iid", "smooth_s1_q8", "smooth_s4_q2", "smooth_s16_q256"):
            continue
        say("profile sweep", name)
        a = img.numpy()
        got = profile(a, 64)
        assert got == oracle_seam_profile(a, 64), name
        row = ss.seam_profile_row(got, 1024, 1024)
        assert 0.0 <= row["seam_fragment_fraction"] <= 1.0 and 0.0 <= row["seam_edge_fraction"] <= 1.0


def test_profile_constant_and_strided():
    # constant: one zone per tile, all of them on a seam
    img = torch.full((130, 260), 7, dtype=torch.uint8, device="cuda")
    got = ss.seam_profile(img, 32)
    assert got["fragments"] == 5 * 9 and got["seam_fragments"] == 45 and got["seam_pixels"] == 130 * 260
    # rows with a pitch larger than the width
    base = torch.as_tensor(np.random.default_rng(7).integers(0, 3, (90, 160), dtype=np.uint8)).cuda()
    view = base[:, 10:130]
    assert ss.seam_profile(view, 16) == oracle_seam_profile(view.cpu().numpy(), 16)
