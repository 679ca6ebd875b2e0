"""CPU checks of the synthetic generators used by the benchmark and the GPU
parity tests: determinism, the committed hash of the parity image, and the
structural properties each adversarial family is built to have."""
import hashlib

import numpy as np
import torch
from scipy import ndimage

from sscommon import oracle_bruteforce, oracle_quantize, oracle_sparse_seam
from ssx import ssgen

CONFIG1_SHA256 = "f74dafd397f7006f"  # first 16 hex digits, 512x512 sigma 2 q32 seed 0


def test_config1_hash_and_levels():
    img = ssgen.config1()
    assert img.shape == (512, 512) and img.dtype == torch.uint8
    assert hashlib.sha256(img.numpy().tobytes()).hexdigest()[:16] == CONFIG1_SHA256
    assert len(torch.unique(img)) == 32


def test_quantizer_matches_oracle():
    v = torch.arange(256, dtype=torch.uint8)
    for lv in (2, 3, 4, 8, 16, 32, 64, 256):
        assert np.array_equal(ssgen.quantize(v, lv).numpy(), oracle_quantize(v.numpy(), lv))


def test_deterministic_and_seeded():
    a = ssgen.smooth(97, 131, 4.0, 16, seed=3)
    b = ssgen.smooth(97, 131, 4.0, 16, seed=3)
    c = ssgen.smooth(97, 131, 4.0, 16, seed=4)
    assert torch.equal(a, b) and not torch.equal(a, c)
    assert torch.equal(ssgen.iid(50, 60, 1), ssgen.iid(50, 60, 1))


def test_iid_is_roughly_uniform():
    x = ssgen.iid(512, 512, 0).numpy()
    hist = np.bincount(x.ravel(), minlength=256)
    assert hist.min() > 0.8 * hist.mean() and hist.max() < 1.2 * hist.mean()


def test_smooth_stretch_covers_range():
    for s in (1, 4, 16):
        x = ssgen.smooth(256, 256, s, 256, seed=1).numpy()
        assert x.min() == 0 and x.max() == 255


def test_weights_symmetric_normalised():
    for s in (1.0, 2.0, 3.0, 4.0, 16.0):
        w = ssgen.gaussian_weights(s)
        assert len(w) == int(np.ceil(3 * s)) + 1
        assert abs(w[0] + 2 * sum(w[1:]) - 65536) <= len(w) * 2
        assert all(w[i] >= w[i + 1] for i in range(len(w) - 1))


def test_serpentine_is_one_non_maximal_zone():
    for vertical in (False, True):
        img = ssgen.serpentine(64, 96, 0, vertical).numpy()
        lab, n = ndimage.label(img == 254, structure=np.ones((3, 3)))
        assert n == 1
        mx, _ = oracle_bruteforce(img)
        assert not mx[img == 254].any()


def test_checker_colours_are_single_zones():
    img = ssgen.checker(100, 77, cell=7).numpy()
    for v in (64, 192):
        _, n = ndimage.label(img == v, structure=np.ones((3, 3)))
        assert n == 1
    mx, mn = oracle_bruteforce(img)
    assert np.array_equal(mx > 0, img == 192) and np.array_equal(mn > 0, img == 64)


def test_rects_boundaries_shifted():
    img = ssgen.rects(300, 300, 2, shift=16).numpy()
    # the first cell boundary is at 16 in both axes
    assert (img[:16, :16] == img[0, 0]).all()
    assert not (img[:16, 16:40] == img[0, 0]).all() or not (img[16:40, :16] == img[0, 0]).all()


def test_edt_disks():
    img = ssgen.edt_disks(400, 500, 3)
    assert img.dtype == np.uint8 and img.max() == 255
    assert (img == 0).mean() > 0.2
    # regional maxima lie inside disks
    mx, _ = oracle_bruteforce(img)
    assert (img[mx > 0] > 0).sum() > 0


def test_mixed_sizes_range():
    sizes = ssgen.mixed_sizes(200, 0)
    hs = [h for h, _ in sizes]
    ws = [w for _, w in sizes]
    assert min(hs) >= 256 and max(hs) <= 1024 and min(ws) >= 256 and max(ws) <= 1024


def test_seam_comb_fills_fragment_capacity():
    for tile in (16, 32):
        img = ssgen.seam_comb(256, 192, tile).numpy()
        _, _, counts = oracle_sparse_seam(img, tile)
        # interior tiles: every perimeter pixel but the two corners a row/column
        # pattern shares with its neighbour is a fragment of its own
        assert counts[1:-1, 1:-1].min() >= 4 * tile - 6
        assert counts.max() <= 4 * tile - 4
        assert np.array_equal(*oracle_bruteforce(img)) is False
