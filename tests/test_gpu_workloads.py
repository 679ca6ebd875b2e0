"""Oracle parity at the size of the batched and serving workloads.

Batched: 4096 images, W and H iid U[256, 1024], content 1:1:1 iid noise /
Gaussian sigma 3 at 16 levels / EDT of 20..300 disks, run as one hipGraph
replay at the batch-autotuned (tile, threads) and as one eager batched launch
at another tile size; every image is checked against the CPU Global-UF route.

Serving: 51 requests (48 drawn, 3 at the corners of the distribution), side
log-uniform 64..8192, aspect 0.5..2, the same three generators, open-loop
Poisson arrivals under the 32 GB admission cap and under a cap that holds only
the largest request: each result equals the oracle, and a request run alone
equals the same request run in the mix.
"""
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


@pytest.mark.timeout(600)
def test_batched_4096_mixed_vs_oracle():
    n = 4096
    sizes = ssgen.mixed_sizes(n, seed=0)
    say("batched workload: generating %d images" % n)
    imgs = [ssgen.mixed_image(k, h, w, 0, "cuda") for k, (h, w) in enumerate(sizes)]
    torch.cuda.synchronize()
    px = sum(h * w for h, w in sizes)
    say("batched workload: %.3f Gpx generated" % (px / 1e9))
    old = ss.default_config()
    sample = list(range(0, n, n // 512))
    souts = [(torch.empty_like(imgs[i]), torch.empty_like(imgs[i])) for i in sample]
    (tile, threads), _ = ss.autotune(ss.descs([imgs[i] for i in sample], souts), reps=2, batch=True)
    ss.set_default_config(*old)
    other = 32 if tile != 32 else 16
    say("batched workload: autotuned tile=%d threads=%d, second run at tile=%d" % (tile, threads, other))
    outs_a = [(torch.full_like(i, 7), torch.full_like(i, 7)) for i in imgs]
    outs_b = [(torch.full_like(i, 7), torch.full_like(i, 7)) for i in imgs]
    da, db = ss.descs(imgs, outs_a), ss.descs(imgs, outs_b)
    wsa = ss.new_workspace(ss.batch_workspace_size(da, tile, threads))
    plan = ss.BatchPlan(da, wsa, tile, threads)
    plan.run()
    torch.cuda.synchronize()
    assert ss.stats(wsa)["invariant"] == 0
    plan.close()
    del wsa
    wsb = ss.new_workspace(ss.batch_workspace_size(db, other, 256))
    ss.extrema_batch(db, wsb, other, 256)
    torch.cuda.synchronize()
    assert ss.stats(wsb)["invariant"] == 0
    del wsb
    for k, img in enumerate(imgs):
        if k % 512 == 0:
            say("batched workload: checking image %d of %d" % (k, n))
        emax, emin = oracle_global_uf(img.cpu().numpy())
        for outs, t in ((outs_a, tile), (outs_b, other)):
            mx, mn = outs[k]
            assert np.array_equal(mx.cpu().numpy(), emax), (k, sizes[k], t)
            assert np.array_equal(mn.cpu().numpy(), emin), (k, sizes[k], t)


@pytest.fixture(scope="module")
def serving_set():
    # 48 drawn requests plus the corners of the distribution: the largest
    # side at aspect 1 and 2, and the smallest at aspect 0.5
    sizes = ssgen.serving_sizes(48, seed=1) + [(8192, 8192), (5793, 11585), (91, 45)]
    say("serving workload: generating %d requests, %.3f Gpx" % (len(sizes), sum(h * w for h, w in sizes) / 1e9))
    imgs = [np.ascontiguousarray(ssgen.mixed_image(k, h, w, 1, "cuda").cpu().numpy())
            for k, (h, w) in enumerate(sizes)]
    expected = [oracle_global_uf(i) for i in imgs]
    return imgs, expected


def open_loop(server, imgs, qps, seed):
    outs = [(np.full_like(i, 7), np.full_like(i, 7)) for i in imgs]
    reqs = (ss.Request * len(imgs))(*[ss.Server.request(i, mx, mn, k)
                                      for k, (i, (mx, mn)) in enumerate(zip(imgs, outs))])
    comps = server.run_open_loop(reqs, qps=qps, seed=seed)
    for k, c in enumerate(comps):
        assert c.status == 0 and c.tag == k and c.done_s >= c.launch_s >= c.submit_s, k
    return outs, comps


@pytest.mark.timeout(600)
def test_serving_32gb_cap_alone_equals_mixed_equals_oracle(serving_set):
    imgs, expected = serving_set
    server = ss.Server(hbm_cap_bytes=32 << 30, streams=3, max_batch_pixels=32 << 20)
    say("serving workload: open loop under the 32 GB cap")
    mixed, comps = open_loop(server, imgs, qps=400.0, seed=7)
    assert max(c.batch_size for c in comps) > 1       # arrivals were folded together
    for k, ((mx, mn), (emax, emin)) in enumerate(zip(mixed, expected)):
        assert np.array_equal(mx, emax) and np.array_equal(mn, emin), (k, imgs[k].shape)
    say("serving workload: each request alone")
    for k, img in enumerate(imgs):
        mx, mn = np.full_like(img, 7), np.full_like(img, 7)
        c = server.wait(server.submit(img, mx, mn, tag=k))
        assert c.status == 0 and c.batch_size == 1
        assert np.array_equal(mx, mixed[k][0]) and np.array_equal(mn, mixed[k][1]), (k, img.shape)
    st = server.stats()
    assert st["completed"] == 2 * len(imgs) and st["rejected"] == 0
    assert st["peak_inflight_bytes"] <= 32 << 30
    server.close()


@pytest.mark.timeout(600)
def test_serving_tight_cap_vs_oracle(serving_set):
    imgs, expected = serving_set
    big = max(imgs, key=lambda i: i.size)
    h, w = big.shape
    cap = 3 * big.size + ss.workspace_size(w, h) + (8 << 20)
    server = ss.Server(hbm_cap_bytes=cap, streams=2)
    say("serving workload: open loop under a %d-byte cap" % cap)
    outs, _ = open_loop(server, imgs, qps=400.0, seed=8)
    for k, ((mx, mn), (emax, emin)) in enumerate(zip(outs, expected)):
        assert np.array_equal(mx, emax) and np.array_equal(mn, emin), (k, imgs[k].shape)
    assert server.stats()["peak_inflight_bytes"] <= cap
    server.close()
