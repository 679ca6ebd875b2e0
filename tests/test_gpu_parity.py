"""Parity of the HIP path (through the C ABI) with the CPU oracle.

Bit-exact masks are required everywhere: the 206 golden triples, random
images at every (tile, threads) config, the seeded benchmark configurations
(oracle comparison where the oracle finishes in seconds, size-independent
properties beyond that), and the per-tile seam-fragment counts.
"""
import sys

import numpy as np
import pytest
import torch
from hypothesis import HealthCheck, given, settings
from hypothesis import strategies as hst

from sscommon import golden_cases, oracle_bruteforce, oracle_global_uf, oracle_quantize, oracle_sparse_seam
from ssx import sparse_seam as ss
from ssx import ssgen

pytestmark = pytest.mark.gpu

CASES = list(golden_cases())


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run(img_np, tile=None, threads=None, ws=None):
    img = torch.as_tensor(np.ascontiguousarray(img_np)).cuda()
    mx, mn, ws = ss.extrema(img, tile, threads, ws)
    torch.cuda.synchronize()
    st = ss.stats(ws)
    assert st["invariant"] == 0
    return mx.cpu().numpy(), mn.cpu().numpy(), ws, st


@pytest.fixture(scope="module")
def dev():
    assert torch.cuda.is_available()
    return torch.device("cuda")


@pytest.mark.parametrize("tile,threads", ss.CONFIGS)
def test_golden_fixtures(dev, tile, threads):
    say("golden fixtures tile=%d threads=%d" % (tile, threads))
    ws = None
    for suite, name, img, emax, emin, _ in CASES:
        mx, mn, ws, _ = run(img, tile, threads, ws)
        assert np.array_equal(mx, emax), (suite, name)
        assert np.array_equal(mn, emin), (suite, name)


@pytest.mark.parametrize("seed", range(12))
def test_random_images_all_configs(dev, seed):
    rng = np.random.default_rng(1000 + seed)
    h, w = int(rng.integers(1, 300)), int(rng.integers(1, 300))
    kind = seed % 4
    if kind == 0:
        img = rng.integers(0, 256, (h, w), dtype=np.uint8)
    elif kind == 1:
        b = int(rng.integers(2, 40))
        img = np.repeat(np.repeat(rng.integers(0, 4, (h // b + 1, w // b + 1), dtype=np.uint8), b, 0), b, 1)[:h, :w]
    elif kind == 2:
        img = ssgen.smooth(h, w, float(rng.choice([1.0, 3.0])), int(rng.choice([2, 4, 16])), seed).numpy()
    else:
        img = ssgen.checker(h, w, cell=int(rng.integers(1, 9))).numpy()
    say("random seed=%d %dx%d kind=%d" % (seed, h, w, kind))
    emax, emin = oracle_global_uf(img, threads=4)
    for tile, threads in ss.CONFIGS:
        mx, mn, _, _ = run(img, tile, threads)
        assert np.array_equal(mx, emax), (tile, threads)
        assert np.array_equal(mn, emin), (tile, threads)


@settings(max_examples=40, deadline=None, derandomize=True,
          suppress_health_check=[HealthCheck.too_slow, HealthCheck.function_scoped_fixture])
@given(h=hst.integers(1, 130), w=hst.integers(1, 130), levels=hst.sampled_from([2, 3, 4, 16, 256]),
       block=hst.integers(1, 12), seed=hst.integers(0, 2**32 - 1),
       config=hst.sampled_from(list(ss.CONFIGS)))
def test_property_small_images(dev, h, w, levels, block, seed, config):
    """Property form of the HIP-vs-oracle check (sizes 1..130 straddling the
    tile sides, blocky quantised content so that zones cross seams): the C-ABI
    path at a drawn (T, threads) equals the brute-force closed form."""
    rng = np.random.default_rng(seed)
    small = rng.integers(0, 256, (-(-h // block), -(-w // block)), dtype=np.uint8)
    img = np.ascontiguousarray(np.repeat(np.repeat(small, block, 0), block, 1)[:h, :w])
    img = oracle_quantize(img, levels)
    emax, emin = oracle_bruteforce(img)
    mx, mn, _, _ = run(img, *config)
    assert np.array_equal(mx, emax) and np.array_equal(mn, emin), (h, w, config)


@pytest.mark.parametrize("tile", ss.TILES)
def test_tile_counts_match_oracle(dev, tile):
    rng = np.random.default_rng(5)
    for kind in range(3):
        h, w = 197, 263
        if kind == 0:
            img = rng.integers(0, 256, (h, w), dtype=np.uint8)
        elif kind == 1:
            img = ssgen.smooth(h, w, 4.0, 8, 3).numpy()
        else:
            img = ssgen.rects(h, w, 1, tile // 2).numpy()
        say("tile counts tile=%d kind=%d" % (tile, kind))
        mx, mn, ws, st = run(img, tile, 256 if tile > 16 else 128)
        omx, omn, counts = oracle_sparse_seam(img, tile)
        assert np.array_equal(mx, omx) and np.array_equal(mn, omn)
        got = ss.tile_counts(ws, *counts.shape)
        assert np.array_equal(got, counts.astype(np.uint32))
        assert st["fragments"] == int(counts.sum())


# Every tile size with its wide quad configuration, plus the row tile stage
# (T = 64 with 128 threads, the default configuration).
SCALE_CONFIGS = ((16, 256), (32, 256), (64, 256), (64, 128))


def test_config1_image(dev):
    img = ssgen.config1()
    say("config1 512x512 sigma2 q32 seed0")
    emax, emin = oracle_global_uf(img.numpy())
    for tile, threads in SCALE_CONFIGS:
        mx, mn, _, _ = run(img.numpy(), tile, threads)
        assert np.array_equal(mx, emax) and np.array_equal(mn, emin)


def run_dev(img, tile, threads, ws=None):
    """The C-ABI path on a resident image; masks back on the host."""
    mx, mn, ws = ss.extrema(img, tile, threads, ws)
    torch.cuda.synchronize()
    assert ss.stats(ws)["invariant"] == 0
    return mx.cpu().numpy(), mn.cpu().numpy(), ws


@pytest.mark.parametrize("seed", range(5))
@pytest.mark.parametrize("side", [512, 1024, 2048])
def test_sweep_small_sides(dev, side, seed):
    """The resolution / gray-level sweep at seeds 0..4: all 19 images at every
    tile size, bit-exact against the CPU Global-UF route."""
    ws = {}
    for name, img in ssgen.sweep_images(side, seed, device="cuda"):
        say("sweep side=%d seed=%d %s" % (side, seed, name))
        emax, emin = oracle_global_uf(img.cpu().numpy())
        for cfg in SCALE_CONFIGS:
            mx, mn, ws[cfg] = run_dev(img, *cfg, ws.get(cfg))
            assert np.array_equal(mx, emax), (name, cfg)
            assert np.array_equal(mn, emin), (name, cfg)


def test_adversarial_4096(dev):
    for tile in ss.TILES:
        for name, img in ssgen.adversarial_images(4096, tile, 0, device="cuda"):
            h = img.cpu().numpy()
            emax, emin = oracle_global_uf(h)
            for threads in [n for t, n in SCALE_CONFIGS if t == tile]:
                say("adversarial 4096 tile=%d threads=%d %s" % (tile, threads, name))
                mx, mn, ws, st = run(h, tile, threads)
                assert np.array_equal(mx, emax), (name, tile, threads)
                assert np.array_equal(mn, emin), (name, tile, threads)
                if name == "constant":
                    assert (mx == 255).all() and (mn == 255).all()
                if name.startswith("serpentine"):
                    # the whole path is one zone with a greater neighbour at its start
                    assert not (mx[h == 254] == 255).any()


def test_seam_comb_fragment_capacity(dev):
    """Every perimeter pixel a fragment: tiles near 4T-4 fragments, and at
    T = 64 merge blocks with more fragments than the block merge keeps in LDS
    (its dense-array path)."""
    for tile, threads in SCALE_CONFIGS:
        img = ssgen.seam_comb(1024, 1024, tile).numpy()
        say("seam_comb 1024 tile=%d threads=%d" % (tile, threads))
        emax, emin, ecounts = oracle_sparse_seam(img, tile)
        mx, mn, ws, st = run(img, tile, threads)
        assert np.array_equal(mx, emax) and np.array_equal(mn, emin), tile
        counts = ss.tile_counts(ws, *ecounts.shape)
        assert np.array_equal(counts, ecounts), tile
        assert counts.max() >= 4 * tile - 6


@pytest.mark.parametrize("side", [4096, 8192])
def test_large_sweep_vs_global_uf(dev, side):
    """All 19 sweep images (seed 0) at 4096 and 8192, every tile size."""
    ws = {}
    for name, img in ssgen.sweep_images(side, 0, device="cuda"):
        say("large sweep side=%d %s" % (side, name))
        emax, emin = oracle_global_uf(img.cpu().numpy())
        for cfg in SCALE_CONFIGS:
            mx, mn, ws[cfg] = run_dev(img, *cfg, ws.get(cfg))
            assert np.array_equal(mx, emax), (name, cfg)
            assert np.array_equal(mn, emin), (name, cfg)


def test_full_size_sweep_vs_global_uf(dev):
    """The benchmark's workload at full size: every image of the
    16384x16384 sweep in every configuration above (T = 64 with 128 threads is
    the benchmark's default), bit-exact against the CPU Global-UF route (one
    image is a second or two of CPU work on the host's cores, done once per
    image)."""
    ws = {}
    for name, img in ssgen.sweep_images(16384, 0, device="cuda"):
        emax, emin = oracle_global_uf(img.cpu().numpy())
        for cfg in SCALE_CONFIGS:
            say("full-size sweep 16384 tile=%d threads=%d %s" % (cfg[0], cfg[1], name))
            mx, mn, ws[cfg] = run_dev(img, *cfg, ws.get(cfg))
            assert np.array_equal(mx, emax), (name, cfg)
            assert np.array_equal(mn, emin), (name, cfg)


def test_many_tiles_per_writeback_wave(dev):
    """T = 16 at 16384x16384: over a million tiles, so each write-back wave
    walks its tile range in several chunks of 64 (the counts and base words of
    a chunk are loaded at once); bit-exact against Global-UF."""
    name, img = "smooth_s4_q4", None
    for n, im in ssgen.sweep_images(16384, 0, device="cuda", sigmas=(4,), levels=(4,)):
        name, img = n, im
    say("writeback chunks 16384 T=16 %s" % name)
    mx, mn, ws = ss.extrema(img, 16, 256)
    torch.cuda.synchronize()
    st = ss.stats(ws)
    assert st["invariant"] == 0 and st["tiles"] == 1024 * 1024 and st["dirty_tiles"] > 0
    emax, emin = oracle_global_uf(img.cpu().numpy())
    assert np.array_equal(mx.cpu().numpy(), emax)
    assert np.array_equal(mn.cpu().numpy(), emin)


def test_strided_rows_and_outputs(dev):
    rng = np.random.default_rng(9)
    base = torch.as_tensor(rng.integers(0, 4, (130, 300), dtype=np.uint8)).cuda()
    view = base[:, 17:17 + 211]            # pitch 300, width 211
    omx = torch.full((130, 256), 7, dtype=torch.uint8, device="cuda")
    omn = torch.full((130, 256), 7, dtype=torch.uint8, device="cuda")
    emax, emin = oracle_global_uf(view.cpu().numpy())
    for tile, threads in ((32, 256), (64, 128)):
        omx.fill_(7)
        omn.fill_(7)
        mx, mn, ws = ss.extrema(view, tile, threads, out=(omx[:, :211], omn[:, :211]))
        torch.cuda.synchronize()
        assert np.array_equal(mx.cpu().numpy(), emax) and np.array_equal(mn.cpu().numpy(), emin)
        assert (omx[:, 211:] == 7).all() and (omn[:, 211:] == 7).all()


def test_repeat_runs_are_identical(dev):
    img = ssgen.smooth(1000, 1500, 4.0, 4, 2, device="cuda")
    ref = None
    ws = {}
    for k in range(6):
        cfg = (32, 256) if k % 2 else (64, 128)
        mx, mn, ws[cfg] = ss.extrema(img, *cfg, ws.get(cfg))
        cur = (mx.cpu().numpy(), mn.cpu().numpy())
        if ref is None:
            ref = cur
        assert np.array_equal(ref[0], cur[0]) and np.array_equal(ref[1], cur[1])


def test_global_uf_route_golden_and_sweep(dev):
    """The dense Global-UF route (ss_global_uf_u8, the harness's cross-route)
    equals the golden triples and the oracle on sweep images, including
    strided rows and outputs."""
    ws = None
    for suite, name, img, emax, emin, _ in CASES:
        mx, mn, ws = ss.global_uf(torch.as_tensor(img).cuda(), ws)
        torch.cuda.synchronize()
        assert np.array_equal(mx.cpu().numpy(), emax), (suite, name)
        assert np.array_equal(mn.cpu().numpy(), emin), (suite, name)
    for name, img in ssgen.sweep_images(2048, 3, device="cuda"):
        say("global-uf route 2048 %s" % name)
        mx, mn, ws = ss.global_uf(img, ws)
        emax, emin = oracle_global_uf(img.cpu().numpy())
        assert np.array_equal(mx.cpu().numpy(), emax), name
        assert np.array_equal(mn.cpu().numpy(), emin), name
    for name, img in ssgen.adversarial_images(2048, 32, 0, device="cuda"):
        say("global-uf route adversarial 2048 %s" % name)
        mx, mn, ws = ss.global_uf(img, ws)
        emax, emin = oracle_global_uf(img.cpu().numpy())
        assert np.array_equal(mx.cpu().numpy(), emax), name
        assert np.array_equal(mn.cpu().numpy(), emin), name
    base = ssgen.smooth(300, 500, 2.0, 4, 1, device="cuda")
    view = base[:, 13:13 + 401]
    omx = torch.full((300, 480), 7, dtype=torch.uint8, device="cuda")
    omn = torch.full((300, 480), 7, dtype=torch.uint8, device="cuda")
    mx, mn, _ = ss.global_uf(view, out=(omx[:, :401], omn[:, :401]))
    torch.cuda.synchronize()
    emax, emin = oracle_global_uf(view.cpu().numpy())
    assert np.array_equal(mx.cpu().numpy(), emax) and np.array_equal(mn.cpu().numpy(), emin)
    assert (omx[:, 401:] == 7).all() and (omn[:, 401:] == 7).all()
