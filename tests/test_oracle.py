"""Pins the CPU oracle: the 206 reference golden triples, an independent scipy
closed form, tile-count invariants and special cases.  CPU only."""
import numpy as np
import pytest
from hypothesis import HealthCheck, given, settings
from hypothesis import strategies as hst
from scipy import ndimage

from sscommon import (golden_cases, mismatches, oracle_bruteforce, oracle_global_uf,
                      oracle_quantize, oracle_seam_profile, oracle_sparse_seam)

CASES = list(golden_cases())
EIGHT = np.ones((3, 3), dtype=bool)


def scipy_extrema(img, structure=EIGHT):
    """Independent closed form: label each grey level's components, then a zone
    is a maximum iff no pixel of the zone's dilation (inside the image) is
    strictly greater.  Absent (out-of-image) neighbours never count."""
    img = np.asarray(img, dtype=np.int16)
    mx = np.zeros(img.shape, np.uint8)
    mn = np.zeros(img.shape, np.uint8)
    for level in np.unique(img):
        lab, n = ndimage.label(img == level, structure=structure)
        if n == 0:
            continue
        nb_max = ndimage.maximum_filter(np.where(lab > 0, -1, img), footprint=structure,
                                        mode="constant", cval=-1)
        nb_min = ndimage.minimum_filter(np.where(lab > 0, 999, img), footprint=structure,
                                        mode="constant", cval=999)
        idx = np.arange(1, n + 1)
        zone_max = ndimage.maximum(nb_max, lab, idx)
        zone_min = ndimage.minimum(nb_min, lab, idx)
        is_max = np.concatenate([[False], np.asarray(zone_max) <= level])
        is_min = np.concatenate([[False], np.asarray(zone_min) >= level])
        mx[is_max[lab] & (lab > 0)] = 255
        mn[is_min[lab] & (lab > 0)] = 255
    return mx, mn


def test_fixture_count():
    assert len(CASES) == 206
    assert sum(1 for c in CASES if c[0] == "oracle_v0_1") == 34
    assert sum(1 for c in CASES if c[0] == "oracle_v0_2_fuzz") == 160
    assert sum(1 for c in CASES if c[0] == "oracle_v0_3_sparse_seam_adversarial") == 12


@pytest.mark.parametrize("case", CASES, ids=lambda c: c[1])
def test_manifest_matches_fixture(case):
    _, name, img, emax, emin, row = case
    assert row is not None
    assert int(row["height"]) == img.shape[0] and int(row["width"]) == img.shape[1]
    assert int(row["max_pixels"]) == int(np.count_nonzero(emax))
    assert int(row["min_pixels"]) == int(np.count_nonzero(emin))
    assert set(np.unique(emax)) <= {0, 255} and set(np.unique(emin)) <= {0, 255}


@pytest.mark.parametrize("case", CASES, ids=lambda c: c[1])
def test_routes_match_golden(case):
    _, name, img, emax, emin, _ = case
    for mx, mn in (oracle_bruteforce(img), oracle_global_uf(img, threads=2), scipy_extrema(img)):
        assert mismatches(mx, emax) == 0
        assert mismatches(mn, emin) == 0
    for tile in (1, 3, 16, 32, 64):
        mx, mn, _ = oracle_sparse_seam(img, tile)
        assert mismatches(mx, emax) == 0, tile
        assert mismatches(mn, emin) == 0, tile


def test_four_connectivity_is_caught():
    """A plausible mistake (4- instead of 8-connectivity) fails the fixtures."""
    four = ndimage.generate_binary_structure(2, 1)
    bad = 0
    for _, name, img, emax, emin, _ in CASES:
        mx, mn = scipy_extrema(img, structure=four)
        bad += mismatches(mx, emax) + mismatches(mn, emin) > 0
    assert bad > 10


def test_zero_padding_is_caught():
    """Treating out-of-image neighbours as value 0 (instead of absent) fails."""
    bad = 0
    for _, name, img, emax, emin, _ in CASES:
        padded = np.pad(img, 1, constant_values=0)
        _, mn = oracle_bruteforce(padded)
        bad += mismatches(mn[1:-1, 1:-1], emin) > 0
    assert bad > 10


def _random_image(rng, h, w, levels, kind):
    if kind == "iid":
        img = rng.integers(0, 256, (h, w), dtype=np.uint8)
    elif kind == "blocks":
        b = int(rng.integers(1, 9))
        small = rng.integers(0, 256, (-(-h // b), -(-w // b)), dtype=np.uint8)
        img = np.repeat(np.repeat(small, b, 0), b, 1)[:h, :w].copy()
    else:
        yy, xx = np.mgrid[0:h, 0:w]
        img = ((xx * 7 + yy * 3 + rng.integers(0, 4, (h, w))) % 256).astype(np.uint8)
    return oracle_quantize(img, levels)


@pytest.mark.parametrize("seed", range(40))
def test_routes_agree_on_random_images(seed):
    rng = np.random.default_rng(seed)
    h, w = int(rng.integers(1, 131)), int(rng.integers(1, 131))
    levels = int(rng.choice([2, 3, 4, 8, 16, 256]))
    kind = ["iid", "blocks", "ramp"][seed % 3]
    img = _random_image(rng, h, w, levels, kind)
    bmx, bmn = oracle_bruteforce(img)
    smx, smn = scipy_extrema(img)
    assert np.array_equal(bmx, smx) and np.array_equal(bmn, smn)
    gmx, gmn = oracle_global_uf(img, threads=4)
    assert np.array_equal(bmx, gmx) and np.array_equal(bmn, gmn)
    for tile in (16, 32, 64, int(rng.integers(1, 65))):
        tmx, tmn, _ = oracle_sparse_seam(img, tile)
        assert np.array_equal(bmx, tmx) and np.array_equal(bmn, tmn), tile


@settings(max_examples=60, deadline=None, derandomize=True, suppress_health_check=[HealthCheck.too_slow])
@given(h=hst.integers(1, 130), w=hst.integers(1, 130), levels=hst.sampled_from([2, 3, 4, 8, 16, 256]),
       kind=hst.sampled_from(["iid", "blocks", "ramp"]), seed=hst.integers(0, 2**32 - 1),
       tile=hst.integers(1, 64))
def test_routes_agree_property(h, w, levels, kind, seed, tile):
    """Property form of the cross-route check (sizes 1..130, any tile side):
    brute force, Global-UF and Sparse-Seam at T in {16, 32, 64} and one drawn
    T agree; hypothesis reaches the edge sizes (1, tile multiples +- 1) and
    shrinks a failure to its smallest image."""
    img = _random_image(np.random.default_rng(seed), h, w, levels, kind)
    bmx, bmn = oracle_bruteforce(img)
    gmx, gmn = oracle_global_uf(img, threads=2)
    assert np.array_equal(bmx, gmx) and np.array_equal(bmn, gmn)
    for t in (16, 32, 64, tile):
        tmx, tmn, _ = oracle_sparse_seam(img, t)
        assert np.array_equal(bmx, tmx) and np.array_equal(bmn, tmn), t


def tile_seam_counts(img, tile):
    """Independent count of tile-local flat zones that have an equal 8-neighbour
    in another tile (the reference profiler's seam-fragment definition)."""
    h, w = img.shape
    ty, tx = -(-h // tile), -(-w // tile)
    out = np.zeros((ty, tx), np.uint16)
    padded = np.pad(img.astype(np.int16), 1, constant_values=-1)
    for j in range(ty):
        for i in range(tx):
            y0, x0 = j * tile, i * tile
            sub = img[y0:y0 + tile, x0:x0 + tile]
            seam_zones = set()
            for level in np.unique(sub):
                lab, n = ndimage.label(sub == level, structure=EIGHT)
                for (ly, lx) in zip(*np.nonzero(lab)):
                    if lab[ly, lx] in seam_zones:
                        continue
                    if 0 < ly < sub.shape[0] - 1 and 0 < lx < sub.shape[1] - 1:
                        continue
                    win = padded[y0 + ly:y0 + ly + 3, x0 + lx:x0 + lx + 3]
                    for dy in range(3):
                        for dx in range(3):
                            gy, gx = y0 + ly + dy - 1, x0 + lx + dx - 1
                            inside = y0 <= gy < y0 + sub.shape[0] and x0 <= gx < x0 + sub.shape[1]
                            if not inside and win[dy, dx] == level:
                                seam_zones.add((level, lab[ly, lx]))
                seam_zones = {z for z in seam_zones}
            out[j, i] = len(seam_zones)
    return out


@pytest.mark.parametrize("seed", range(6))
def test_tile_fragment_counts(seed):
    rng = np.random.default_rng(100 + seed)
    h, w = int(rng.integers(20, 90)), int(rng.integers(20, 90))
    img = _random_image(rng, h, w, int(rng.choice([2, 4, 8])), ["iid", "blocks", "ramp"][seed % 3])
    for tile in (8, 16, 32):
        _, _, counts = oracle_sparse_seam(img, tile)
        assert np.array_equal(counts, tile_seam_counts(img, tile)), tile


def test_special_cases():
    for shape in ((1, 1), (1, 7), (9, 1), (64, 64), (129, 257)):
        img = np.full(shape, 77, np.uint8)
        for mx, mn in (oracle_bruteforce(img), oracle_global_uf(img),
                       oracle_sparse_seam(img, 32)[:2]):
            assert (mx == 255).all() and (mn == 255).all()
    img = np.zeros((5, 5), np.uint8)
    img[2, 2] = 9
    mx, mn = oracle_bruteforce(img)
    assert mx.sum() == 255 and mx[2, 2] == 255
    assert np.count_nonzero(mn) == 24 and mn[2, 2] == 0


def test_invariants_random():
    rng = np.random.default_rng(7)
    for _ in range(20):
        img = _random_image(rng, int(rng.integers(2, 60)), int(rng.integers(2, 60)), 4, "blocks")
        mx, mn = oracle_bruteforce(img)
        assert mx.any() and mn.any()          # at least one maximum and one minimum zone
        both = (mx > 0) & (mn > 0)
        assert both.any() == (img.min() == img.max())


def test_quantizer_closed_form():
    v = np.arange(256, dtype=np.uint8)
    assert np.array_equal(oracle_quantize(v, 256), v)
    assert np.array_equal(oracle_quantize(v, 2), np.where(v < 128, 0, 255).astype(np.uint8))
    q4 = oracle_quantize(v, 4)
    assert np.array_equal(q4, np.array([0, 85, 170, 255], np.uint8)[v // 64])
    q32 = oracle_quantize(v, 32)
    assert len(np.unique(q32)) == 32 and q32[0] == 0 and q32[255] == 255


def scipy_seam_profile(img, tile):
    """Independent closed form of the seam-sparsity counters: scipy labels of
    every grey level inside every tile; a zone is a seam zone iff one of its
    pixels has an equal 8-neighbour inside the image but outside the tile;
    equal pairs are counted by shifted array comparisons."""
    a = np.asarray(img, dtype=np.int16)
    h, w = a.shape
    zones = seam_zones = seam_px = 0
    for y0 in range(0, h, tile):
        for x0 in range(0, w, tile):
            sub = a[y0:y0 + tile, x0:x0 + tile]
            inside = np.zeros((h, w), bool)
            inside[y0:y0 + sub.shape[0], x0:x0 + sub.shape[1]] = True
            for level in np.unique(sub):
                lab, n = ndimage.label(sub == level, structure=EIGHT)
                zones += n
                full = np.zeros((h, w), np.int32)
                full[y0:y0 + sub.shape[0], x0:x0 + sub.shape[1]] = lab
                # equal pixels outside the tile, dilated back onto the tile
                outside_eq = (a == level) & ~inside
                touch = ndimage.binary_dilation(outside_eq, structure=EIGHT) & (full > 0)
                hit = np.unique(full[touch])
                seam_zones += len(hit)
                seam_px += int(np.isin(full, hit[hit > 0]).sum())
    equal = seam_e = 0
    ys, xs = np.indices((h, w))
    for dy, dx in ((0, 1), (1, -1), (1, 0), (1, 1)):
        y1, y2 = 0, h - dy
        x1, x2 = max(0, -dx), w - max(0, dx)
        p = a[y1:y2, x1:x2]
        q = a[y1 + dy:y2 + dy, x1 + dx:x2 + dx]
        eq = p == q
        cross = ((ys[y1:y2, x1:x2] // tile) != ((ys[y1:y2, x1:x2] + dy) // tile)) | \
                ((xs[y1:y2, x1:x2] // tile) != ((xs[y1:y2, x1:x2] + dx) // tile))
        equal += int(eq.sum())
        seam_e += int((eq & cross).sum())
    return {"fragments": zones, "seam_fragments": seam_zones, "seam_pixels": seam_px,
            "equal_edges": equal, "seam_edges": seam_e}


@pytest.mark.parametrize("seed", range(6))
def test_seam_profile_closed_form(seed):
    rng = np.random.default_rng(300 + seed)
    h, w = int(rng.integers(12, 80)), int(rng.integers(12, 80))
    img = _random_image(rng, h, w, int(rng.choice([2, 3, 8, 256])), ["iid", "blocks", "ramp"][seed % 3])
    for tile in (4, 16, 32):
        prof = oracle_seam_profile(img, tile)
        assert prof == scipy_seam_profile(img, tile), tile
        # the seam zones are the per-tile seam-fragment counts of the
        # Sparse-Seam restatement
        assert prof["seam_fragments"] == int(oracle_sparse_seam(img, tile)[2].sum())


def test_seam_profile_special_cases():
    # constant image: one zone per tile, every pair equal
    h, w, t = 37, 50, 16
    prof = oracle_seam_profile(np.full((h, w), 9, np.uint8), t)
    tiles = -(-h // t) * -(-w // t)
    pairs = h * (w - 1) + (h - 1) * w + 2 * (h - 1) * (w - 1)
    assert prof["fragments"] == tiles and prof["equal_edges"] == pairs
    assert prof["seam_fragments"] == tiles and prof["seam_pixels"] == h * w
    # one tile covering the image: no seams
    prof = oracle_seam_profile(np.full((8, 8), 1, np.uint8), 16)
    assert prof["seam_edges"] == 0 and prof["seam_fragments"] == 0
    # checkerboard: no equal horizontal or vertical pairs, all diagonal pairs equal
    cb = (np.indices((20, 20)).sum(0) % 2).astype(np.uint8)
    prof = oracle_seam_profile(cb, 8)
    assert prof["equal_edges"] == 2 * 19 * 19
