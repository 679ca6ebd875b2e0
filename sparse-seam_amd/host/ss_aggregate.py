#!/usr/bin/env python3
"""Aggregates ss_harness.csv files the way the reference's campaign
aggregator does (tools/aggregate_campaign.py:33-50): throughput is the total
pixels divided by the sum of per-image median compute times (and, for the
host-to-host mode, of the per-image median H2H times); percentiles are
linear-interpolated; every figure carries its mismatch count.

Usage: python3 ss_aggregate.py <ss_harness.csv> [more.csv ...]
Prints one line per (algorithm, tile) group and a JSON summary line.
"""
import csv
import json
import sys
from collections import defaultdict


def pct(values, p):
    v = sorted(values)
    if not v:
        return float("nan")
    pos = p * (len(v) - 1)
    lo, hi = int(pos), min(int(pos) + 1, len(v) - 1)
    return v[lo] + (v[hi] - v[lo]) * (pos - lo)


def aggregate(rows):
    pix = sum(int(r["width"]) * int(r["height"]) for r in rows)
    sec = sum(float(r["median_compute_ms"]) for r in rows) / 1e3
    h2h = sum(float(r["median_h2h_ms"]) for r in rows) / 1e3
    aux = sum(float(r["auxiliary_bytes"]) for r in rows)
    mism = sum(int(r.get(k) or 0) for r in rows for k in ("oracle_max_mismatches", "oracle_min_mismatches",
                                                          "cross_max_mismatches", "cross_min_mismatches"))
    return {"images": len(rows), "pixels": pix, "compute_mpix_s": pix / 1e6 / sec if sec else float("nan"),
            "h2h_mpix_s": pix / 1e6 / h2h if h2h else float("nan"), "aux_Bpp": aux / pix if pix else 0.0,
            "p50_compute_ms": pct([float(r["median_compute_ms"]) for r in rows], 0.5),
            "p99_compute_ms": pct([float(r["median_compute_ms"]) for r in rows], 0.99), "mismatches": mism}


def main(paths):
    groups = defaultdict(list)
    for path in paths:
        with open(path, newline="", encoding="utf-8-sig") as f:
            for r in csv.DictReader(f):
                groups[(r["algorithm"], r.get("tile", ""))].append(r)
    summary = {}
    for (alg, tile), rows in sorted(groups.items()):
        a = aggregate(rows)
        summary["%s/T%s" % (alg, tile)] = a
        print("%-28s T=%-3s %10.3f MP/s compute  %10.3f MP/s h2h  aux %6.3f B/px  n=%d  mism=%d"
              % (alg, tile, a["compute_mpix_s"], a["h2h_mpix_s"], a["aux_Bpp"], a["images"], a["mismatches"]))
    print(json.dumps(summary))
    return 0 if all(a["mismatches"] == 0 for a in summary.values()) else 1


if __name__ == "__main__":
    if len(sys.argv) < 2:
        print(__doc__, file=sys.stderr)
        sys.exit(2)
    sys.exit(main(sys.argv[1:]))
