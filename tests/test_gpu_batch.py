"""Batched form (one tile-stage launch over the tiles of many uneven images),
its hipGraph-captured plan, and the autotuner -- all through the C ABI, all
checked bit for bit against the single-image path and the oracle."""
import sys

import numpy as np
import pytest
import torch

from sscommon import oracle_global_uf
from ssx import sparse_seam as ss
from ssx import ssgen

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make_images(n, seed, lo=1, hi=300):
    rng = np.random.default_rng(seed)
    imgs = []
    for k in range(n):
        h, w = int(rng.integers(lo, hi)), int(rng.integers(lo, hi))
        kind = k % 4
        if kind == 0:
            a = rng.integers(0, 256, (h, w), dtype=np.uint8)
        elif kind == 1:
            a = ssgen.smooth(h, w, 3.0, 16, seed * 100 + k).numpy()
        elif kind == 2:
            a = ssgen.checker(h, w, cell=int(rng.integers(1, 6))).numpy()
        else:
            a = np.full((h, w), 9, np.uint8) if k % 8 == 3 else ssgen.rects(h, w, k, 8).numpy()
        imgs.append(torch.as_tensor(a).cuda())
    return imgs


def outputs_for(imgs, fill=0):
    return [(torch.full_like(i, fill), torch.full_like(i, fill)) for i in imgs]


@pytest.mark.parametrize("tile,threads", [(16, 256), (32, 256), (64, 256), (32, 64)])
def test_batch_matches_oracle(tile, threads):
    imgs = make_images(37, tile + threads)
    outs = outputs_for(imgs)
    d = ss.descs(imgs, outs)
    ws = ss.new_workspace(ss.batch_workspace_size(d, tile, threads))
    say("batch n=%d tile=%d threads=%d" % (len(imgs), tile, threads))
    ss.extrema_batch(d, ws, tile, threads)
    torch.cuda.synchronize()
    st = ss.stats(ws)
    assert st["invariant"] == 0
    for k, (img, (mx, mn)) in enumerate(zip(imgs, outs)):
        emax, emin = oracle_global_uf(img.cpu().numpy())
        assert np.array_equal(mx.cpu().numpy(), emax), k
        assert np.array_equal(mn.cpu().numpy(), emin), k


def test_batch_equals_single_image_path():
    imgs = make_images(20, 7, 100, 700)
    outs = outputs_for(imgs)
    d = ss.descs(imgs, outs)
    ws = ss.new_workspace(ss.batch_workspace_size(d, 32, 256))
    ss.extrema_batch(d, ws, 32, 256)
    torch.cuda.synchronize()
    for img, (mx, mn) in zip(imgs, outs):
        smx, smn, _ = ss.extrema(img, 32, 256)
        assert torch.equal(mx, smx) and torch.equal(mn, smn)


def test_batch_plan_graph_replays():
    imgs = make_images(24, 11, 50, 400)
    outs = outputs_for(imgs)
    d = ss.descs(imgs, outs)
    ws = ss.new_workspace(ss.batch_workspace_size(d, 32, 256))
    plan = ss.BatchPlan(d, ws, 32, 256)
    expected = [oracle_global_uf(i.cpu().numpy()) for i in imgs]
    for rep in range(4):
        for mx, mn in outs:          # eager work between replays
            mx.fill_(7)
            mn.fill_(7)
        torch.cuda.synchronize()
        say("graph replay", rep)
        plan.run()
        torch.cuda.synchronize()
        assert ss.stats(ws)["invariant"] == 0
        for (mx, mn), (emax, emin) in zip(outs, expected):
            assert np.array_equal(mx.cpu().numpy(), emax)
            assert np.array_equal(mn.cpu().numpy(), emin)
    plan.close()


def test_batch_adversarial_large():
    imgs = [ssgen.serpentine(1500, 1700, 1, False, "cuda"), ssgen.serpentine(1024, 777, 2, True, "cuda"),
            ssgen.checker(1200, 1100, 5, device="cuda"), ssgen.constant(513, 1025, device="cuda"),
            ssgen.rects(2048, 2048, 3, 16, "cuda")]
    outs = outputs_for(imgs)
    d = ss.descs(imgs, outs)
    for tile in (16, 32, 64):
        ws = ss.new_workspace(ss.batch_workspace_size(d, tile, 256))
        say("batch adversarial tile", tile)
        ss.extrema_batch(d, ws, tile, 256)
        torch.cuda.synchronize()
        assert ss.stats(ws)["invariant"] == 0
        for img, (mx, mn) in zip(imgs, outs):
            emax, emin = oracle_global_uf(img.cpu().numpy())
            assert np.array_equal(mx.cpu().numpy(), emax)
            assert np.array_equal(mn.cpu().numpy(), emin)


def test_autotune_picks_valid_config_and_keeps_results():
    imgs = make_images(6, 21, 200, 600)
    outs = outputs_for(imgs)
    d = ss.descs(imgs, outs)
    old = ss.default_config()
    (tile, threads), rate = ss.autotune(d, reps=2)
    assert (tile, threads) in ss.CONFIGS and rate > 0
    assert ss.default_config() == (tile, threads)
    for img, (mx, mn) in zip(imgs, outs):
        emax, emin = oracle_global_uf(img.cpu().numpy())
        assert np.array_equal(mx.cpu().numpy(), emax)
        assert np.array_equal(mn.cpu().numpy(), emin)
    ss.set_default_config(*old)
