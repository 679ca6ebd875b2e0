"""Serving loop (ss_server_*): continuous batching under an HBM admission cap.
Every result must be bit-identical alone and in the mix, and equal to the
oracle."""
import sys

import numpy as np
import pytest

from sscommon import oracle_global_uf
from ssx import sparse_seam as ss
from ssx import ssgen

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def requests(n, seed, lo=16, hi=700):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        h, w = int(rng.integers(lo, hi)), int(rng.integers(lo, hi))
        kind = k % 3
        if kind == 0:
            img = rng.integers(0, 256, (h, w), dtype=np.uint8)
        elif kind == 1:
            img = ssgen.smooth(h, w, 3.0, 16, seed * 1000 + k).numpy()
        else:
            img = ssgen.edt_disks(max(h, 100), max(w, 100), seed * 1000 + k)
        out.append(np.ascontiguousarray(img))
    return out


def run_mixed(server, imgs):
    outs = [(np.zeros_like(i), np.zeros_like(i)) for i in imgs]
    ids = [server.submit(i, mx, mn, tag=k) for k, (i, (mx, mn)) in enumerate(zip(imgs, outs))]
    comps = [server.wait(r) for r in ids]
    return outs, comps


def test_alone_equals_mixed_equals_oracle():
    imgs = requests(30, 1)
    server = ss.Server(hbm_cap_bytes=1 << 30, streams=2)
    say("mixed submit of", len(imgs))
    mixed, comps = run_mixed(server, imgs)
    assert max(c.batch_size for c in comps) > 1          # requests were folded together
    alone = []
    for i in imgs:
        mx, mn = np.zeros_like(i), np.zeros_like(i)
        c = server.wait(server.submit(i, mx, mn))
        assert c.batch_size == 1
        alone.append((mx, mn))
    for img, (amx, amn), (mmx, mmn) in zip(imgs, alone, mixed):
        assert np.array_equal(amx, mmx) and np.array_equal(amn, mmn)
        emax, emin = oracle_global_uf(img)
        assert np.array_equal(amx, emax) and np.array_equal(amn, emin)
    st = server.stats()
    assert st["completed"] == 2 * len(imgs) and st["rejected"] == 0
    server.close()


def test_tight_cap_serialises_but_stays_exact():
    imgs = requests(12, 2, 200, 600)
    biggest = max(i.size for i in imgs)
    cap = 3 * biggest + ss.workspace_size(600, 600) + (1 << 20)
    server = ss.Server(hbm_cap_bytes=cap, streams=2)
    say("tight cap", cap)
    outs, comps = run_mixed(server, imgs)
    for img, (mx, mn) in zip(imgs, outs):
        emax, emin = oracle_global_uf(img)
        assert np.array_equal(mx, emax) and np.array_equal(mn, emin)
    assert server.stats()["peak_inflight_bytes"] <= cap
    server.close()


def test_admission_refuses_oversized_request():
    server = ss.Server(hbm_cap_bytes=1 << 20)
    big = np.zeros((2048, 2048), np.uint8)
    with pytest.raises(ss.SSError) as e:
        server.submit(big, np.zeros_like(big), np.zeros_like(big))
    assert e.value.status == ss.SS_ERR_CAPACITY
    assert server.stats()["rejected"] == 1
    server.close()


def test_open_loop_generator():
    imgs = requests(40, 3, 32, 400)
    outs = [(np.zeros_like(i), np.zeros_like(i)) for i in imgs]
    server = ss.Server(hbm_cap_bytes=1 << 30)
    reqs = (ss.Request * len(imgs))(*[ss.Server.request(i, mx, mn, k)
                                      for k, (i, (mx, mn)) in enumerate(zip(imgs, outs))])
    comps = server.run_open_loop(reqs, qps=2000.0, seed=5)
    for k, (img, (mx, mn)) in enumerate(zip(imgs, outs)):
        c = comps[k]
        assert c.status == 0 and c.tag == k and c.done_s >= c.launch_s >= c.submit_s
        emax, emin = oracle_global_uf(img)
        assert np.array_equal(mx, emax) and np.array_equal(mn, emin)
    server.close()
