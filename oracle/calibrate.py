"""Times the oracle restatement against the reference's own CPU tree
(oracle/_ref) on identical inputs -- TEST INFRASTRUCTURE ONLY.

    python -m oracle.calibrate [--workload mixed_1m] [--threads N]

Prints one JSON line: build on 1 thread and search of every box on N threads
for both, and the ratios restatement / reference.  The restatement follows
the same algorithm (std::nth_element with the same comparator, the same push
rule), so the ratios should be close to 1; a bench "port" baseline is scaled
by them to stand in for the reference where oracle/_ref is absent.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from fkdtree_amd import gen  # noqa: E402


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="mixed_1m", choices=["mixed_1m", "uniform_100k"])
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    a = ap.parse_args(argv)
    if not oracle.ref_available():
        raise SystemExit("oracle/_ref is not built (reference sources absent)")
    pts, boxes = {"mixed_1m": gen.workload_mixed_1m, "uniform_100k": gen.workload_uniform_100k}[a.workload]()
    t0 = time.perf_counter()
    tc, ti = oracle.build(pts.coords, pts.ids)
    ob = time.perf_counter() - t0
    t0 = time.perf_counter()
    ooff, _ = oracle.search(tc, ti, boxes.mins, boxes.maxs, threads=a.threads)
    osr = time.perf_counter() - t0
    rt = oracle.RefTree(pts.coords, pts.ids)
    t0 = time.perf_counter()
    rt.build()
    rb = time.perf_counter() - t0
    t0 = time.perf_counter()
    roff, _ = rt.search(boxes.mins, boxes.maxs, variant=0, threads=a.threads)
    rs = time.perf_counter() - t0
    rc, rids = rt.tree()
    rt.close()
    assert np.array_equal(rc.view(np.uint32), tc.view(np.uint32)) and np.array_equal(rids, ti)
    assert np.array_equal(roff, ooff)
    print(json.dumps({"workload": a.workload, "threads": a.threads,
                      "build_ms": {"restatement": ob * 1e3, "reference": rb * 1e3, "ratio": ob / rb},
                      "search_s": {"restatement": osr, "reference": rs, "ratio": osr / rs}}))


if __name__ == "__main__":
    main()
