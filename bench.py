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


def cpu_baseline(images, threads):
    """Global-UF oracle route on the host cores over a bounded sample."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    from sscommon import oracle_global_uf
    px, med = 0, 0.0
    names = []
    for name, img in images:
        h = img.cpu().numpy()
        oracle_global_uf(h, threads)  # warm-up (page faults, thread pool)
        ts = []
        for _ in range(2):
            t0 = time.perf_counter()
            oracle_global_uf(h, threads)
            ts.append(time.perf_counter() - t0)
        px += h.size
        med += sorted(ts)[len(ts) // 2]
        names.append(name)
    return px / med / 1e9, names


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
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py runs on one GPU only: images are independent and this build has no "
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
    st = ss.stats(ws)
    log("warm-up done", st)

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
            "kernel": "k_tile_stage",
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

    if not args.no_cpu:
        cores = args.cpu_threads or min(16, os.cpu_count() or 1)
        sample = [images[0], images[9]]  # iid and sigma-4 / 16 levels
        t0 = time.time()
        rate, names = cpu_baseline(sample, cores)
        out["cpu_baseline"] = {
            "value": round(rate, 4), "unit": "Gpixel/s", "cores": cores, "kind": "port",
            "sample": "Global-UF oracle route on %s at %dx%d (median of 2 runs each after a "
                      "warm-up, sum(px)/sum(median))" % ("+".join(names), side, side),
        }
        log("cpu baseline %.1fs" % (time.time() - t0))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
