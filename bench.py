#!/usr/bin/env python3
"""Benchmark of the exact dual regional-extrema operator on one MI355X.

Workload (one "step" = one pass of the plain single-image path over the set):
the resolution / gray-level sweep at 16384 x 16384, seed 0 -- iid uniform
noise plus Gaussian-smoothed noise (sigma 1, 4, 16) quantized to 2, 4, 8, 16,
64 and 256 levels: 19 images, 5.1 Gpixel per step, resident in HBM before the
timed region.  Each image is one ss_extrema_u8 call (both masks).

Printed: one JSON line with whole-job Gpixel/s, the roofline of the dominant
kernel (tile stage: 1 B image read + 2 B mask writes per pixel, timed in place
with HIP events on the pipeline's stream), per-image p50/p99 latency and the
CPU baseline (the Global-UF oracle route on the host cores, bounded sample).
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "sparse-seam_amd"))

METRIC = "dual-mask Gpixel/s vs HBM roofline; p50/p99 latency, batched mixed sizes"
HBM_PEAK_GBS = 8000.0
ALG_BYTES_PER_PX = 3.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def percentile(values, q):
    """Linear-interpolated percentile (the reference's rule, gpu_common_v0_6.cu:86-103)."""
    v = sorted(values)
    if not v:
        return float("nan")
    pos = (len(v) - 1) * q / 100.0
    lo = int(pos)
    hi = min(lo + 1, len(v) - 1)
    return v[lo] + (v[hi] - v[lo]) * (pos - lo)


def cpu_baseline(images, threads, runs=3):
    """Global-UF oracle route on the host cores: one warm-up run (page faults,
    thread pool), then `runs` timed runs of every image; sum(px) / sum(median)."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    from sscommon import oracle_global_uf
    px, med = 0, 0.0
    names = []
    oracle_global_uf(images[0][1], threads)
    for name, h in images:
        ts = []
        for _ in range(runs):
            t0 = time.perf_counter()
            oracle_global_uf(h, threads)
            ts.append(time.perf_counter() - t0)
        px += h.size
        med += sorted(ts)[len(ts) // 2]
        names.append(name)
    return px / med / 1e9, names


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def batched_leg(n, dev, ss, ssgen, torch, reps=3):
    """Mixed-size batch (widths and heights U[256, 1024], content 1:1:1 iid /
    smoothed sigma 3 at 16 levels / EDT of disks): autotuned (tile, threads),
    per-image launches of the plain path versus one hipGraph-replayed batched
    pipeline (one tile-stage launch over the tiles of all images)."""
    import numpy as np
    t0 = time.time()
    sizes = ssgen.mixed_sizes(n, seed=0)
    imgs = [ssgen.mixed_image(k, h, w, 0, dev) for k, (h, w) in enumerate(sizes)]
    outs = [(torch.empty_like(i), torch.empty_like(i)) for i in imgs]
    px = sum(h * w for h, w in sizes)
    torch.cuda.synchronize()
    log("batch: generated %d images (%.2f Gpx) in %.1fs" % (n, px / 1e9, time.time() - t0))
    step = max(1, n // 64)
    sample = list(range(0, n, step))
    sample_descs = ss.descs([imgs[i] for i in sample], [outs[i] for i in sample])
    # each path with the configuration its own autotuner picks on a sample:
    # per-image launches by the per-image rule on 64 images, the batch by
    # batched launches on 512 (a smaller batch is launch-bound, and its
    # ranking of the candidates says little about the full one)
    (ptile, pthreads), _ = ss.autotune(sample_descs, reps=2)
    bsample = list(range(0, n, max(1, n // 512)))
    bsample_descs = ss.descs([imgs[i] for i in bsample], [outs[i] for i in bsample])
    (tile, threads), _ = ss.autotune(bsample_descs, reps=3, batch=True)
    log("batch: autotuned per-image tile=%d threads=%d, batched tile=%d threads=%d" % (ptile, pthreads, tile, threads))
    stream = torch.cuda.current_stream()
    ev = lambda: torch.cuda.Event(enable_timing=True)
    # per-image launches of the plain path, timed from a C++ loop over
    # ss_extrema_u8 (one call per image, one stream)
    d = ss.descs(imgs, outs)
    ws = ss.new_workspace(max(ss.workspace_size(w, h, ptile, pthreads) for h, w in sizes), dev)
    plain_ms = ss.per_image_loop(d, ws, ptile, pthreads, reps=reps)
    ref = [(o[0].clone(), o[1].clone()) for o in outs[::step]]
    del ws
    # one graph-replayed batched pipeline
    wsb = ss.new_workspace(ss.batch_workspace_size(d, tile, threads), dev)
    plan = ss.BatchPlan(d, wsb, tile, threads)
    for o in outs:
        o[0].zero_()
        o[1].zero_()
    plan.run()
    torch.cuda.synchronize()
    st = ss.stats(wsb)
    same = all(torch.equal(o[0], r[0]) and torch.equal(o[1], r[1]) for o, r in zip(outs[::step], ref))
    batch_ms = []
    for _ in range(reps):
        a, b = ev(), ev()
        a.record(stream)
        plan.run()
        b.record(stream)
        torch.cuda.synchronize()
        batch_ms.append(a.elapsed_time(b))
    plan.close()
    p_ms, b_ms = percentile(plain_ms, 50), percentile(batch_ms, 50)
    return {
        "workload": "%d images, W,H iid U[256,1024], content iid / Gaussian sigma 3 q16 / EDT of 20..300 "
                    "disks (1:1:1), seed 0" % n,
        "pixels": px, "tile": tile, "threads": threads, "per_image_tile": ptile, "per_image_threads": pthreads,
        "batched_gpx": round(px / (b_ms * 1e-3) / 1e9, 3), "batched_ms": round(b_ms, 3),
        "per_image_launch_gpx": round(px / (p_ms * 1e-3) / 1e9, 3), "per_image_launch_ms": round(p_ms, 3),
        "per_image_launch_timing": "C++ loop over ss_extrema_u8 (libss_loop.so), hipEvents around each pass, "
                                   "median of %d passes after one untimed pass" % reps,
        "speedup": round(p_ms / b_ms, 3),
        "batched_equals_per_image": bool(same), "fragments": st["fragments"], "dirty_tiles": st["dirty_tiles"],
    }


def link_bandwidth(torch, dev, chunk=32 << 20, n=48, nstreams=3, reps=3):
    """The host link under the serving traffic pattern with no kernels: n
    "requests" of `chunk` pixels, each one H2D copy of 1 B/px then two D2H
    copies of 1 B/px (the two masks), issued round-robin on nstreams streams
    from pinned buffers (what the server does around its kernels).  Also the
    H2D-only and D2H-only rates of the same pattern.  Median of reps, GB/s and
    the pattern's Gpixel/s (its link bound)."""
    h_in = [torch.empty(chunk, dtype=torch.uint8, pin_memory=True) for _ in range(nstreams)]
    h_out = [torch.empty(2 * chunk, dtype=torch.uint8, pin_memory=True) for _ in range(nstreams)]
    d_in = [torch.empty(chunk, dtype=torch.uint8, device=dev) for _ in range(nstreams)]
    d_out = [torch.empty(2 * chunk, dtype=torch.uint8, device=dev) for _ in range(nstreams)]
    streams = [torch.cuda.Stream() for _ in range(nstreams)]

    def timed(h2d, d2h):
        ts = []
        for _ in range(reps + 1):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for k in range(n):
                j = k % nstreams
                with torch.cuda.stream(streams[j]):
                    if h2d:
                        d_in[j].copy_(h_in[j], non_blocking=True)
                    if d2h:
                        h_out[j][:chunk].copy_(d_out[j][:chunk], non_blocking=True)
                        h_out[j][chunk:].copy_(d_out[j][chunk:], non_blocking=True)
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        return percentile(ts[1:], 50)

    t_h2d, t_d2h, t_both = timed(True, False), timed(False, True), timed(True, True)
    del h_in, h_out, d_in, d_out
    return {"pattern": "%d requests of %d MiB px, 1 B/px H2D + 2 x 1 B/px D2H, %d streams, pinned" %
                       (n, chunk >> 20, nstreams),
            "h2d_gbs": round(n * chunk / t_h2d / 1e9, 2), "d2h_gbs": round(2 * n * chunk / t_d2h / 1e9, 2),
            "both_gbs": round(3 * n * chunk / t_both / 1e9, 2),
            "bound_gpx": round(n * chunk / t_both / 1e9, 3)}


def serving_leg(n, dev, ss, ssgen, torch, cap_bytes=32 << 30, plain_cfg=None):
    """Open-loop Poisson requests (side log-uniform 64..8192, aspect 0.5..2,
    the three generators), host buffers in pinned memory, HBM admission cap 32
    GB.  Plain per-request path (H2D, one launch, D2H, in order) versus the
    continuous-batching server at the plain path's own sustainable rate."""
    import numpy as np
    t0 = time.time()
    reqs_np = []
    keep = []
    for k, (h, w) in enumerate(ssgen.serving_sizes(n, seed=0)):
        img = ssgen.mixed_image(k, h, w, 1, dev) if k % 3 != 2 else torch.as_tensor(ssgen.edt_disks(h, w, k))
        host = torch.empty((h, w), dtype=torch.uint8, pin_memory=True)
        host.copy_(img)
        mx = torch.empty((h, w), dtype=torch.uint8, pin_memory=True)
        mn = torch.empty((h, w), dtype=torch.uint8, pin_memory=True)
        keep.append((host, mx, mn))
        reqs_np.append((host.numpy(), mx.numpy(), mn.numpy()))
    px = sum(r[0].size for r in reqs_np)
    log("serve: generated %d requests (%.2f Gpx) in %.1fs" % (n, px / 1e9, time.time() - t0))
    tile, threads = ss.default_config()
    reqs = (ss.Request * n)(*[ss.Server.request(i, mx, mn, k) for k, (i, mx, mn) in enumerate(reqs_np)])

    def run(server, qps, seed):
        comps = server.run_open_loop(reqs, qps=qps, seed=seed)
        lat = [(c.done_s - c.submit_s) * 1e3 for c in comps]
        span = max(c.done_s for c in comps) - min(c.submit_s for c in comps)
        return comps, lat, px / span / 1e9

    # plain per-request launches: one request per launch, one stream, same
    # arrival process (H2D, pipeline, D2H for each request in turn)
    # warm-up of each server: a saturated pass over all requests (first
    # launches of every kernel variant, first copies) before anything is timed
    ptile, pthreads = plain_cfg or (tile, threads)
    plain = ss.Server(hbm_cap_bytes=cap_bytes, max_batch=1, streams=1, tile=ptile, threads=pthreads)
    run(plain, 1e9, 3)
    _, _, plain_sat = run(plain, 1e9, 5)                 # capacity: all offered at once
    qps = 0.8 * plain_sat * 1e9 / (px / n)               # 80 % of the plain path's capacity
    _, plain_lat, plain_gpx = run(plain, qps, 11)
    plain_ref = [(m.clone(), n_.clone()) for _, m, n_ in keep]
    plain.close()
    # continuous batching: launches capped at 32 Mpixel, three in flight
    server = ss.Server(hbm_cap_bytes=cap_bytes, streams=3, max_batch_pixels=32 << 20)
    run(server, 1e9, 3)
    for _, m, n_ in keep:
        m.zero_()
        n_.zero_()
    comps, srv_lat, srv_gpx = run(server, qps, 11)
    same = all(torch.equal(m, r[0]) and torch.equal(n_, r[1]) for (_, m, n_), r in zip(keep, plain_ref))
    _, _, srv_sat = run(server, 1e9, 5)
    st = server.stats()
    server.close()
    # the link bound: every request moves 1 B/px host->device and 2 B/px back;
    # the same copies without kernels give the rate the link allows
    link = link_bandwidth(torch, dev)
    return {
        "workload": "%d requests, side log-uniform 64..8192, aspect 0.5..2, iid / sigma 3 q16 / EDT, "
                    "pinned host buffers (H2D + pipeline + D2H per request), HBM cap %d GB, open-loop "
                    "Poisson at 80%% of the plain path's capacity" % (n, cap_bytes >> 30),
        "pixels": px, "qps": round(qps, 2), "tile": tile, "threads": threads,
        "plain_p50_ms": round(percentile(plain_lat, 50), 3), "plain_p99_ms": round(percentile(plain_lat, 99), 3),
        "plain_gpx_at_qps": round(plain_gpx, 3), "plain_saturated_gpx": round(plain_sat, 3),
        "server_p50_ms": round(percentile(srv_lat, 50), 3), "server_p99_ms": round(percentile(srv_lat, 99), 3),
        "server_gpx_at_qps": round(srv_gpx, 3), "server_saturated_gpx": round(srv_sat, 3),
        "mean_batch_at_qps": round(sum(c.batch_size for c in comps) / n, 2),
        "launches": st["launches"], "peak_inflight_bytes": st["peak_inflight_bytes"],
        "server_equals_plain": bool(same),
        "plain_tile": ptile, "plain_threads": pthreads,
        "link_bytes_per_px": 3, "link": link,
        "server_saturated_frac_of_link_bound": round(srv_sat / link["bound_gpx"], 3),
    }


def load_traffic():
    path = os.path.join(REPO, "profiles", "pmc_tile_stage_16384.json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            return json.load(f).get("traffic_bytes_per_launch")
    except (OSError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--side", type=int, default=16384)
    ap.add_argument("--tile", type=int, default=None)
    ap.add_argument("--threads", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-threads", type=int, default=None)
    ap.add_argument("--images", default=None,
                    help="comma-separated subset of the sweep images (for profiling runs)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run one device-to-device copy of one image (known bytes), so "
                         "hardware byte counters can be scaled")
    ap.add_argument("--batch-images", type=int, default=4096, help="batched leg size (0 = skip)")
    ap.add_argument("--serve-requests", type=int, default=192, help="serving leg size (0 = skip)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py runs on one GPU only: images are independent and this library has no "
              "multi-GPU path (sharding would assign whole images, or tile-row bands with one "
              "seam-fragment all-gather, per GPU).", file=sys.stderr)
        sys.exit(2)

    import torch
    import sparse_seam as ss
    import ssgen

    assert torch.cuda.is_available(), "no GPU"
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    tile, threads = (args.tile or ss.default_config()[0]), (args.threads or ss.default_config()[1])

    t0 = time.time()
    images = list(ssgen.sweep_images(args.side, 0, device=dev))
    if args.images:
        keep = set(args.images.split(","))
        images = [(n, i) for n, i in images if n in keep]
        assert images, "no image matches --images"
    torch.cuda.synchronize()
    if args.copy_calibration:
        # one plain copy of side*side bytes: its counter readings give the
        # scale of the hardware byte counters (reads + writes = 2 * side^2 B)
        dst = torch.empty_like(images[0][1])
        dst.copy_(images[0][1])
        torch.cuda.synchronize()
        del dst
    log("generated %d images %dx%d in %.1fs" % (len(images), args.side, args.side, time.time() - t0))

    side = args.side
    npx = side * side
    mx = torch.empty((side, side), dtype=torch.uint8, device=dev)
    mn = torch.empty((side, side), dtype=torch.uint8, device=dev)
    ws = ss.new_workspace(ss.workspace_size(side, side, tile, threads), dev)
    stream = torch.cuda.current_stream()

    for w in range(args.warmup):
        for _, img in images:
            ss.extrema(img, tile, threads, ws, (mx, mn))
    torch.cuda.synchronize()
    if args.warmup > 0:
        log("warm-up done", ss.stats(ws))

    n = len(images)
    K = args.steps
    ev = lambda: torch.cuda.Event(enable_timing=True)
    img_ev = [[(ev(), ev()) for _ in range(n)] for _ in range(K)]
    k1_ev = [[(ev(), ev()) for _ in range(n)] for _ in range(K)]
    start, stop = ev(), ev()
    # torch creates the HIP event lazily at the first record; record the
    # tile-stage events once so their handles exist before the library uses them
    for row in k1_ev:
        for ka, kb in row:
            ka.record(stream)
            kb.record(stream)
    torch.cuda.synchronize()
    start.record(stream)
    for k in range(K):
        for i, (_, img) in enumerate(images):
            a, b = img_ev[k][i]
            ka, kb = k1_ev[k][i]
            ss.profile_events(ka, kb)
            a.record(stream)
            ss.extrema(img, tile, threads, ws, (mx, mn))
            b.record(stream)
    stop.record(stream)
    torch.cuda.synchronize()
    ss.profile_events(None, None)
    total_ms = start.elapsed_time(stop)
    st = ss.stats(ws)   # of the last timed image

    per_img = [[img_ev[k][i][0].elapsed_time(img_ev[k][i][1]) for k in range(K)] for i in range(n)]
    k1 = [k1_ev[k][i][0].elapsed_time(k1_ev[k][i][1]) for k in range(K) for i in range(n)]
    lat = [t for row in per_img for t in row]
    med = [percentile(row, 50) for row in per_img]
    value = K * n * npx / (total_ms * 1e-3) / 1e9
    k1_avg_ms = sum(k1) / len(k1)
    achieved = ALG_BYTES_PER_PX * npx / (k1_avg_ms * 1e-3) / 1e9
    pipeline_gbs = ALG_BYTES_PER_PX * npx * n * K / (total_ms * 1e-3) / 1e9

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Gpixel/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": round(total_ms / K, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": {
            "workload": "plain single-image sweep %dx%d seed 0: iid + Gaussian sigma {1,4,16} x "
                        "levels {2,4,8,16,64,256} (%d images, one ss_extrema_u8 call each)" % (side, side, n),
            "side": side, "images": n, "tile": tile, "threads": threads,
        },
        "roofline": {
            "bound": "hbm",
            # T = 64 with 128 threads runs the row tile stage; the other
            # configurations the quad tile stage
            "kernel": "k_row_tile" if (tile, threads) == (64, 128) else "k_tile_stage",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": load_traffic(),
            "alg_bytes_per_launch": int(ALG_BYTES_PER_PX * npx),
            "avg_launch_ms": round(k1_avg_ms, 4),
            "pipeline_achieved": round(pipeline_gbs, 1),
            "pipeline_frac": round(pipeline_gbs / HBM_PEAK_GBS, 4),
        },
        "p50_ms": round(percentile(lat, 50), 4),
        "p99_ms": round(percentile(lat, 99), 4),
        "sum_px_over_sum_median_gpx": round(n * npx / (sum(med) * 1e-3) / 1e9, 3),
        "per_image_gpx": {name: round(npx / (m * 1e-3) / 1e9, 2) for (name, _), m in zip(images, med)},
        "last_stats": {k: v for k, v in st.items() if k != "status"},
    }
    # seam sparsity of each image at the bench's tile size (the reference
    # profiler's ratios), outside the timed region
    sparsity = {}
    for name, img in images:
        row = ss.seam_profile_row(ss.seam_profile(img, tile, ws), side, side)
        sparsity[name] = {k: round(row[k], 5) for k in ("seam_fragment_fraction", "seam_pixel_fraction",
                                                        "seam_edge_fraction", "seam_fragments_per_mpix")}
    out["seam_sparsity"] = sparsity

    # CPU-baseline sample: every image of the step, kept on the host
    cpu_sample = [(name, img.cpu().numpy()) for name, img in images] if not args.no_cpu else []
    del images, mx, mn, ws
    torch.cuda.empty_cache()
    if args.batch_images > 0:
        t0 = time.time()
        out["batched"] = batched_leg(args.batch_images, dev, ss, ssgen, torch)
        log("batched leg %.1fs" % (time.time() - t0), out["batched"])
        torch.cuda.empty_cache()
    if args.serve_requests > 0:
        t0 = time.time()
        pcfg = None
        if "batched" in out:
            pcfg = (out["batched"]["per_image_tile"], out["batched"]["per_image_threads"])
        out["serving"] = serving_leg(args.serve_requests, dev, ss, ssgen, torch, plain_cfg=pcfg)
        log("serving leg %.1fs" % (time.time() - t0), out["serving"])
        torch.cuda.empty_cache()

    if not args.no_cpu:
        cores = args.cpu_threads or len(os.sched_getaffinity(0))
        t0 = time.time()
        rate, names = cpu_baseline(cpu_sample, cores)
        single = [s_ for s_ in cpu_sample if s_[0] in ("iid", "smooth_s4_q16")] or cpu_sample[:1]
        rate1, names1 = cpu_baseline(single, 1, runs=1)
        out["cpu_baseline"] = {
            "value": round(rate, 4), "unit": "Gpixel/s", "cores": cores, "kind": "port",
            "sample": "Global-UF oracle route on all %d images of the step (%s) at %dx%d, median of 3 "
                      "runs each after one warm-up run, sum(px)/sum(median)" % (len(names), ",".join(names),
                                                                               side, side),
            "single_thread_value": round(rate1, 4),
            "single_thread_sample": "%s, one run each" % ",".join(names1),
            "host_cpus": os.cpu_count(), "cpu_model": cpu_model(),
        }
        log("cpu baseline %.1fs" % (time.time() - t0))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
