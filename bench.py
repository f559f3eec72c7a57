#!/usr/bin/env python3
"""Benchmark of the FKDTree hot paths on one MI355X.

Workload (default ``mixed_1m``): N = 1,048,576 points in 3-D, half uniform in
[0,1)^3 and half in 1,000 Gaussian blobs (sigma 0.005), quantised to a 2^-16
grid; 1,048,576 closed boxes of half-width 0.01 centred on random points.

A timed step is one box search of all queries against the resident tree:
spatial ordering of the boxes, count pass, exclusive scan, fill pass, writing
every hit id (CSR output), all enqueued without a host round trip.
``value`` is box queries per second over the timed steps.  Builds are timed
in their own loop (``builds_per_s``).  Inputs and outputs stay in HBM; timing
uses HIP events on the stream the kernels run on.

Prints one JSON line.  ``--gpus N`` with N > 1 is refused: the benchmark runs on
one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0


def progress(msg: str) -> None:
    """One line on stderr as each section ends (the JSON result goes to stdout)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="mixed_1m", choices=["mixed_1m", "uniform_100k", "uniform_4m"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--events", type=int, default=2048,
                    help="events of the batched workload (first N of the 2,048-event set; 0 skips)")
    ap.add_argument("--cpu-events", type=int, default=256,
                    help="events of the batched CPU baseline (a bounded sample: the first N events)")
    ap.add_argument("--event-steps", type=int, default=3)
    ap.add_argument("--serve-events", type=int, default=512,
                    help="events of the serving stream (0 skips)")
    ap.add_argument("--serve-rates", default="1000,5000,20000", help="Poisson arrival rates (events/s)")
    ap.add_argument("--per-box-threads", default="1,16,256",
                    help="host thread counts of the per-box search pattern (empty or 0 skips)")
    ap.add_argument("--events-only", action="store_true",
                    help="time only the batched events (build + search per batch) and exit; for profiling")
    ap.add_argument("--per-box-input", default="",
                    help="write the per-box driver's input (tests/cpp/fkdtree_perbox.cpp) for the workload "
                         "to this path and exit; for profiling the driver")
    ap.add_argument("--calibrate", action="store_true",
                    help="run a plain copy kernel (counter calibration) three times and exit")
    ap.add_argument("--calib-mib", type=int, default=2048,
                    help="MiB read and written by one calibration copy (well past the 256 MB MALL)")
    return ap.parse_args()


def workload(name):
    from fkdtree_amd import gen
    if name == "mixed_1m":
        return gen.workload_mixed_1m()
    if name == "uniform_100k":
        return gen.workload_uniform_100k()
    return gen.workload_uniform_4m()


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); "
              "it measures one GPU's share; trees shard by event across GPUs", file=sys.stderr)
        return 2
    import fkdtree_amd as F
    from fkdtree_amd import DeviceArray, Stream, Timer

    info = F.device_info()
    if args.calibrate:
        from fkdtree_amd._lib import check, lib
        n = args.calib_mib << 18
        a, b = DeviceArray(n, np.uint32), DeviceArray(n, np.uint32)
        s = Stream()
        for _ in range(3):
            check(lib().fkd_copy_u32(b.ptr, a.ptr, n, s.handle))
        s.sync()
        print(json.dumps({"calibration": "copy_u32_kernel", "bytes_read": 4 * n, "bytes_written": 4 * n}))
        return 0
    if args.events_only:
        print(json.dumps({"events": bench_events(args.events or 256, args.event_steps, batched_only=True)}))
        return 0
    pts, boxes = workload(args.workload)
    n, q, dim = pts.n, boxes.q, pts.dim
    stream = Stream()
    d_coords = DeviceArray.from_host(pts.coords, stream)
    d_ids = DeviceArray.from_host(pts.ids, stream)
    d_mins = DeviceArray.from_host(boxes.mins, stream)
    d_maxs = DeviceArray.from_host(boxes.maxs, stream)
    tree = F.FKDTree.from_device(dim, n, d_coords.ptr, d_ids.ptr, stream)

    # --- build -----------------------------------------------------------
    for _ in range(args.warmup):
        tree.build_async()
    stream.sync()
    tm = Timer()
    tm.start(stream)
    for _ in range(args.steps):
        tree.build_async()
    tm.stop(stream)
    build_ms = tm.ms() / args.steps
    progress(f"build {build_ms:.3f} ms")

    # --- search: size the output once ------------------------------------
    d_off = DeviceArray(q + 1, np.uint64)
    d_out = DeviceArray(1, np.uint32)
    total, ok = 0, False
    for _ in range(3):
        tree.search_single_device(q, d_mins.ptr, d_maxs.ptr, d_off.ptr, d_out.ptr, d_out.nbytes // 4, stream=stream)
        total, ok = tree.search_status(stream)
        if ok:
            break
        if total > d_out.nbytes // 4:
            d_out = DeviceArray(total, np.uint32)
    assert ok, "search output incomplete after regrowing"
    d_cnt = DeviceArray(q, np.uint32)
    d_vis = DeviceArray(q, np.uint32)
    from fkdtree_amd._lib import check, lib
    check(lib().fkd_search_stats(tree._h, q, d_mins.ptr, d_maxs.ptr, None, d_cnt.ptr, d_vis.ptr, stream.handle))
    visited = d_vis.download(stream).astype(np.int64)
    hits = d_cnt.download(stream).astype(np.int64)
    assert int(hits.sum()) == total

    def two_pass():
        tree.search_device(q, d_mins.ptr, d_maxs.ptr, d_off.ptr, d_out.ptr, total, stream=stream)

    def timed(fn, traced):
        for _ in range(args.warmup):
            fn()
        stream.sync()
        ktimers = [Timer() for _ in range(args.steps)]
        t_wall = time.perf_counter()
        tm.start(stream)
        for k in range(args.steps):
            if traced:
                check(lib().fkd_trace_events(tree._h, ktimers[k].a, ktimers[k].b))
            fn()
        tm.stop(stream)
        step_ms = tm.ms() / args.steps
        wall = (time.perf_counter() - t_wall) * 1e3 / args.steps
        check(lib().fkd_trace_events(tree._h, None, None))
        kms = sum(t.ms() for t in ktimers) / args.steps if traced else None
        return step_ms, wall, kms

    search_ms, wall_ms, fill_ms = timed(two_pass, True)
    progress(f"search {search_ms:.3f} ms (fill {fill_ms:.3f} ms)")
    if args.per_box_input:
        write_per_box_input(args.per_box_input, pts, boxes, d_off.download(stream), d_out)
        progress(f"per-box driver input written to {args.per_box_input}")
        return 0
    per_box = None
    threads = [int(x) for x in str(args.per_box_threads).split(",") if x.strip() and int(x) > 0]
    if threads:
        per_box = bench_per_box(pts, boxes, d_off.download(stream), d_out, threads)
        progress("per-box pattern done")

    V = int(visited.sum())
    H = total
    # Algorithmic bytes per query of the traversal kernel: 24 B box + 4 B
    # count + 12 B per visited node (3 coordinates) + 8 B per hit (id read +
    # id write).
    fill_bytes = 28 * q + 12 * V + 8 * H
    # Whole search (count + fill), the per-query model 28 + 12 V + 8 H.
    search_bytes = 28 * q + 12 * V + 8 * H
    # Build: per depth a 4 B key read plus a 16 B read and write per point.
    D = int(np.floor(np.log2(max(n, 1))))
    build_bytes = 36 * n * D
    achieved = fill_bytes / (fill_ms * 1e-3) / 1e9
    # compulsory bytes: points read once, boxes read and offsets written once,
    # every hit id written once (search); the records read and the tree written (build)
    search_compulsory = 16 * n + 28 * q + 4 * H
    build_compulsory = 32 * n
    fill_kernel = fill_kernel_name(tree, dim)
    traffic = measured_traffic(args.workload, fill_kernel)
    # HBM bytes the counters saw per fill launch (calibrated FETCH_SIZE +
    # WRITE_SIZE of the committed profile) over the launch time measured here
    counter_gbs = traffic / (fill_ms * 1e-3) / 1e9 if traffic else None
    search_gbs = search_bytes / (search_ms * 1e-3) / 1e9

    res = {
        "metric": "box queries/s + builds/s at 1M 3D pts; batched vs plain events/s; p99 ms",
        "value": q / (search_ms * 1e-3),
        "unit": "box queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": search_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (deterministic SplitMix64 generator, fkdtree_amd/gen.py)",
        "config": {"workload": args.workload, "points": n, "queries": q, "dim": dim,
                   "half_width": 0.01 if args.workload == "mixed_1m" else None},
        "builds_per_s": 1e3 / build_ms,
        "build_ms": build_ms,
        "search_ms": search_ms,
        "search_wall_ms": wall_ms,
        "search_mode": "two-pass (plan + count + scan + fill, one enqueue)",
        "fill_kernel_ms": fill_ms,
        "hits": H,
        "visited_nodes": V,
        "search_algorithmic_GBs": search_bytes / (search_ms * 1e-3) / 1e9,
        "build_algorithmic_GBs": build_bytes / (build_ms * 1e-3) / 1e9,
        "search_compulsory_bytes": search_compulsory,
        "search_compulsory_GBs": search_compulsory / (search_ms * 1e-3) / 1e9,
        "build_compulsory_bytes": build_compulsory,
        "build_compulsory_GBs": build_compulsory / (build_ms * 1e-3) / 1e9,
        "roofline": {"bound": "hbm", "kernel": f"{fill_kernel} (fill traversal)", "achieved": achieved,
                     "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": achieved / PEAK_HBM_GBS,
                     "traffic": traffic, "algorithmic_bytes": fill_bytes,
                     # the same launch by its counted HBM bytes: node re-reads hit L2 / MALL
                     "hbm_counter_GBs": counter_gbs,
                     "frac_counter": counter_gbs / PEAK_HBM_GBS if counter_gbs else None,
                     # the whole timed search (plan + count + scan + fill) by the same model
                     "search_model_GBs": search_gbs, "search_model_frac": search_gbs / PEAK_HBM_GBS,
                     # what limits the traversal (SQ counters, profiles/kernels_mixed_1m.md)
                     "limiter": "VALU issue of the dependent node walk (latency/issue bound, not HBM)"},
        "device": info["name"],
    }
    if per_box:
        res["per_box"] = per_box
    tree.close()
    del d_out
    if args.events:
        res["events"] = bench_events(args.events, args.event_steps)
        progress("events done")
    if args.serve_events:
        res["serving"] = bench_serving(args.serve_events, [float(r) for r in args.serve_rates.split(",")])
        progress("serving done")
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline(pts, boxes, args.cpu_threads, args.cpu_events if args.events else 0)
    print(json.dumps(res), flush=True)
    return 0


def write_per_box_input(path, pts, boxes, off, d_out, k: int = 1000) -> None:
    """Input of tests/cpp/fkdtree_perbox.cpp: points, boxes, every box's
    expected count and the first k boxes' sorted ids (from the batched
    search's offsets `off` and ids `d_out`)."""
    ids = d_out.download(count=int(off[k]))
    sorted_ids = np.concatenate([np.sort(ids[int(off[i]):int(off[i + 1])]) for i in range(k)])
    with open(path, "wb") as f:
        f.write(np.array([pts.n, boxes.q], np.uint32).tobytes())
        f.write(np.ascontiguousarray(pts.coords.T).tobytes())
        f.write(pts.ids.tobytes())
        f.write(np.ascontiguousarray(boxes.mins.T).tobytes())
        f.write(np.ascontiguousarray(boxes.maxs.T).tobytes())
        f.write(np.diff(off).astype(np.uint32).tobytes())
        f.write(np.array([k], np.uint32).tobytes())
        f.write(np.ascontiguousarray(off[:k + 1], dtype=np.uint64).tobytes())
        f.write(sorted_ids.astype(np.uint32).tobytes())


def bench_per_box(pts, boxes, off, d_out, thread_counts) -> list:
    """The reference driver's per-box pattern (main.cpp:351-359, :372): host
    threads call FKDTree_HIP::search_in_the_box once per box of the workload
    (tests/cpp/fkdtree_perbox.cpp; the calls go to the tree's resident search
    kernel, fkd_box_server_*).  Every result's size is checked against the
    batched search's offsets and the first 1,000 id sets against its ids."""
    import subprocess
    import tempfile

    from fkdtree_amd import build_lib
    if not os.path.exists(build_lib.CPP_PERBOX_BIN):
        return [{"skipped": "tests/cpp/fkdtree_perbox.bin not built"}]
    with tempfile.TemporaryDirectory() as d:
        fin = os.path.join(d, "in.bin")
        write_per_box_input(fin, pts, boxes, off, d_out)
        out = []
        for threads in thread_counts:
            r = subprocess.run([build_lib.CPP_PERBOX_BIN, fin, str(threads)], capture_output=True, text=True,
                               timeout=600)
            if r.returncode != 0:
                out.append({"threads": threads, "failed": r.returncode, "stderr": r.stderr[-500:],
                            "stdout": r.stdout[-500:]})
                continue
            res = json.loads(r.stdout.strip().splitlines()[-1])
            res["unit"] = "box queries/s"
            res["pattern"] = "search_in_the_box per box from host threads (main.cpp:372), FKDTree_HIP"
            out.append(res)
    return out


def bench_serving(n_events: int, rates) -> list:
    """p50 / p99 latency and events/s of the continuous-batching server for
    Poisson streams of n_events events (sizes log-uniform in [1,000, 500,000]
    points, mixed distribution; every point searched with half-width 0.01),
    received as FKDPoint records into a page-locked ingest ring."""
    from fkdtree_amd.serving import EventServer, Request, poisson_requests

    out = []
    # one stream of events, generated once: the sizes and points do not depend
    # on the rate, only the arrival times do (Poisson gaps scale as 1 / rate)
    base = poisson_requests(n_events, 1.0, records=True)
    ring = max(16_000_000, 2 * max(r.n for r in base))
    for rate in rates:
        # events arrive as FKDPoint records into the page-locked ingest ring
        reqs = [Request(r.event, None, None, r.arrival / rate, records=r.records) for r in base]
        srv = EventServer(n_streams=3, max_batch_points=4_000_000, max_batch_events=256, ring_points=ring)
        srv.run(reqs[: min(4, n_events)])          # warm-up (first allocations)
        srv.close()
        srv = EventServer(n_streams=3, max_batch_points=4_000_000, max_batch_events=256, ring_points=ring)
        rep = srv.run(reqs)
        srv.close()
        d = rep.summary()
        d.update({"rate_hz": rate, "points": int(sum(r.n for r in reqs))})
        progress(f"serving at {rate:.0f} events/s: {d.get('events_per_s', 0):.0f} events/s")
        out.append(d)
    return out


def fill_kernel_name(tree, dim: int) -> str:
    """Name of the fill-pass kernel the tree's traversal schedule launches
    (node records in a buffer resource: fewer than 2^28 nodes at dim <= 3)."""
    sched = tree.get_tuning("bucket") if dim <= 3 else 0
    if sched == 1:
        return f"fkd::bucket_search_kernel<{dim}, 1, true, 0>"
    if sched == 2:
        return f"fkd::bucket_search_kernel<{dim}, 1, true, 8>"
    return f"fkd::search_kernel<{dim}, 1, true, 0>"


def measured_traffic(workload: str, kernel: str):
    """HBM bytes per launch of `kernel` (FETCH_SIZE + WRITE_SIZE, calibrated on
    a copy kernel) from the committed rocprofv3 summary of this workload, or
    None when there is none."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    if not os.path.exists(path):
        return None
    k = json.load(open(path)).get("kernels", {}).get(kernel)
    if not k:
        return None
    return k.get("fetch_bytes", 0.0) + k.get("write_bytes", 0.0)


def events_batch(n_events: int):
    """The first n_events events of the 2,048-event workload and their boxes."""
    from fkdtree_amd import gen

    seed = gen.workload_seed("events_2048")
    sizes = gen.event_sizes(2048, seed, 1000, gen.LOG2_200)[:n_events]
    evs = [gen.mixed_event(int(sz), seed, e) for e, sz in enumerate(sizes)]
    batch = gen.concat_events(evs)
    return batch, gen.event_self_boxes(batch, 0.01)


def bench_events(n_events: int, steps: int, batched_only: bool = False) -> dict:
    """Events per second of build + self-search (every point of an event is
    the centre of one box of half-width 0.01 searched in its own tree), for
    the first n_events events of the 2,048-event workload (sizes log-uniform
    in [1,000, 200,000], the mixed distribution).  Three forms:
      batched - all events in shared launches (FKDEvents),
      graph   - the batched pipeline captured in one HIP graph,
      plain   - one tree per event, built and searched one after another."""
    import fkdtree_amd as F
    from fkdtree_amd import DeviceArray, Stream, Timer, gen
    from fkdtree_amd._lib import check, lib

    batch, boxes = events_batch(n_events)
    P = batch.points.n
    qev = np.repeat(np.arange(n_events, dtype=np.uint32), np.diff(batch.offsets))
    s = Stream()
    d_c = DeviceArray.from_host(batch.points.coords, s)
    d_i = DeviceArray.from_host(batch.points.ids, s)
    d_mi = DeviceArray.from_host(boxes.mins, s)
    d_ma = DeviceArray.from_host(boxes.maxs, s)
    d_qev = DeviceArray.from_host(qev, s)
    ev = F.FKDEvents.from_device(3, batch.offsets, d_c.ptr, d_i.ptr, s)
    d_off = DeviceArray(P + 1, np.uint64)
    cap = 1
    d_out = DeviceArray(cap, np.uint32)
    for _ in range(4):
        ev.build_async()
        ev.search_single_device(P, d_mi.ptr, d_ma.ptr, d_off.ptr, d_out.ptr, cap, d_qev=d_qev.ptr, stream=s)
        total, ok = ev.search_status(s)
        if ok:
            break
        if total > cap:
            cap = total
            d_out = DeviceArray(cap, np.uint32)
    assert ok
    tm = Timer()

    def batched():
        ev.build_async()
        ev.search_single_device(P, d_mi.ptr, d_ma.ptr, d_off.ptr, d_out.ptr, cap, d_qev=d_qev.ptr, stream=s)

    batched()
    s.sync()
    tm.start(s)
    for _ in range(steps):
        batched()
    tm.stop(s)
    batched_ms = tm.ms() / steps
    progress(f"events: {n_events} batched {batched_ms:.2f} ms")
    if batched_only:
        ev.close()
        return {"n_events": n_events, "points": P, "queries": P, "hits": int(total), "batched_ms": batched_ms,
                "batched_events_per_s": n_events / (batched_ms / 1e3)}
    g = F.Graph(ev, P, d_mi.ptr, d_ma.ptr, d_off.ptr, d_out.ptr, cap, s, d_qev=d_qev.ptr)
    g.launch()
    s.sync()
    tm.start(s)
    for _ in range(steps):
        g.launch()
    tm.stop(s)
    graph_ms = tm.ms() / steps
    progress(f"events: graph {graph_ms:.2f} ms")
    tot_g, ok_g = ev.search_status(s)
    assert ok_g and tot_g == total
    g.close()
    # algorithmic bytes of the batch: the per-event sums of the search model
    # (28 + 12 V + 8 H per box, V from the reference's push rule) and of the
    # build model (36 n D per event), and the compulsory bytes (16 N + 28 Q +
    # 4 H for the search, 32 N for the build)
    d_cnt = DeviceArray(P, np.uint32)
    d_vis = DeviceArray(P, np.uint32)
    check(lib().fkd_search_stats(ev._h, P, d_mi.ptr, d_ma.ptr, d_qev.ptr, d_cnt.ptr, d_vis.ptr, s.handle))
    V = int(d_vis.download(s).astype(np.int64).sum())
    H = int(d_cnt.download(s).astype(np.int64).sum())
    assert H == total
    del d_cnt, d_vis
    sizes = np.diff(batch.offsets).astype(np.int64)
    depths = np.floor(np.log2(np.maximum(sizes, 1))).astype(np.int64)
    search_bytes = 28 * P + 12 * V + 8 * H
    build_bytes = int((36 * sizes * depths).sum())
    compulsory = (16 * P + 28 * P + 4 * H) + 32 * P
    ev.close()
    del d_out

    # plain: one tree per event, launch after launch
    trees, bufs = [], []
    for e in range(n_events):
        a, b = int(batch.offsets[e]), int(batch.offsets[e + 1])
        n = b - a
        pe = batch.event(e)
        t = F.FKDTree(coords=pe.coords, ids=pe.ids, stream=s)
        bx = gen.boxes_around(pe.coords, 0.01)
        dmi, dma = DeviceArray.from_host(bx.mins, s), DeviceArray.from_host(bx.maxs, s)
        doff = DeviceArray(n + 1, np.uint64)
        t.build()
        t.search_single_device(n, dmi.ptr, dma.ptr, doff.ptr, None, 0, stream=s)
        tot_e, _ = t.search_status(s)
        dout = DeviceArray(max(tot_e, 1), np.uint32)
        for _ in range(3):
            t.search_single_device(n, dmi.ptr, dma.ptr, doff.ptr, dout.ptr, tot_e, stream=s)
            tot_e, ok_e = t.search_status(s)
            if ok_e:
                break
        trees.append(t)
        bufs.append((n, dmi, dma, doff, dout, tot_e))

    def plain():
        for t, (n, dmi, dma, doff, dout, tot_e) in zip(trees, bufs):
            t.build_async()
            t.search_single_device(n, dmi.ptr, dma.ptr, doff.ptr, dout.ptr, tot_e, stream=s)

    plain()
    s.sync()
    tm.start(s)
    for _ in range(max(1, steps // 2)):
        plain()
    tm.stop(s)
    plain_ms = tm.ms() / max(1, steps // 2)
    plain_hits = sum(b[5] for b in bufs)
    assert plain_hits == total
    for t in trees:
        t.close()
    return {"n_events": n_events, "points": P, "queries": P, "hits": total, "visited_nodes": V,
            "batched_ms": batched_ms, "graph_ms": graph_ms, "plain_ms": plain_ms,
            "algorithmic_bytes": {"search": search_bytes, "build": build_bytes},
            "algorithmic_GBs": (search_bytes + build_bytes) / (batched_ms * 1e-3) / 1e9,
            "algorithmic_frac": (search_bytes + build_bytes) / (batched_ms * 1e-3) / 1e9 / PEAK_HBM_GBS,
            "compulsory_bytes": compulsory,
            "compulsory_GBs": compulsory / (batched_ms * 1e-3) / 1e9,
            "batched_events_per_s": n_events / (batched_ms * 1e-3),
            "graph_events_per_s": n_events / (graph_ms * 1e-3),
            "plain_events_per_s": n_events / (plain_ms * 1e-3),
            "batched_over_plain": plain_ms / batched_ms}


def cpu_baseline(pts, boxes, threads, n_events=0):
    """CPU baselines, timed on this host.

    * Headline: the reference CPU tree (oracle/_ref, compiled from the
      reference sources) when present, else the oracle restatement: build on
      1 thread (main.cpp:201-205), search_in_the_box of every box on all host
      cores (OpenMP in place of tbb::parallel_for, main.cpp:372).
    * The same search on 1 thread over the first 131,072 boxes.
    * Side comparator: the reference's copy of the CMSSW KD-tree, built and
      searched box after box on 1 thread as main.cpp:560-629 does (first
      131,072 boxes).
    * Batched events: the oracle restatement, events built in parallel over
      the host cores and every point's box searched in parallel (the same
      events as the GPU events section)."""
    import oracle
    cores = threads or os.cpu_count() or 1
    sample = min(boxes.q, 131072)
    mins1, maxs1 = np.ascontiguousarray(boxes.mins[:, :sample]), np.ascontiguousarray(boxes.maxs[:, :sample])
    if oracle.ref_available():
        kind = "reference"
        rt = oracle.RefTree(pts.coords, pts.ids)
        t0 = time.perf_counter()
        rt.build()
        build_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        off, _ = rt.search(boxes.mins, boxes.maxs, variant=0, threads=cores)
        search_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        rt.search(mins1, maxs1, variant=0, threads=1)
        search1_s = time.perf_counter() - t0
        rt.close()
    else:
        kind = "port"
        t0 = time.perf_counter()
        tc, ti = oracle.build(pts.coords, pts.ids)
        build_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        off, _ = oracle.search(tc, ti, boxes.mins, boxes.maxs, threads=cores)
        search_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        oracle.search(tc, ti, mins1, maxs1, threads=1)
        search1_s = time.perf_counter() - t0
    res = {"value": boxes.q / search_s, "unit": "box queries/s", "cores": cores, "kind": kind,
           "sample": f"all {boxes.q} boxes of the workload, one build of all {pts.n} points on 1 thread",
           "build_ms_1thread": build_s * 1e3, "hits": int(off[-1]),
           "search_1thread": {"value": sample / search1_s, "unit": "box queries/s", "cores": 1,
                              "sample": f"first {sample} boxes"}}
    if oracle.cmssw_available() and pts.dim == 3:
        b_ms, s_ms, found = oracle.cmssw_build_search(pts.coords, mins1, maxs1)
        res["side_comparator"] = {"kind": "CMSSW KDTreeLinkerAlgo (reference copy, KDTreeLinkerAlgoT.h)",
                                  "value": sample / (s_ms * 1e-3), "unit": "box queries/s", "cores": 1,
                                  "build_ms_1thread": b_ms, "found": found,
                                  "sample": f"build of all {pts.n} points, first {sample} boxes"}
    if n_events:
        batch, eb = events_batch(n_events)
        qev = np.repeat(np.arange(n_events, dtype=np.uint32), np.diff(batch.offsets))
        t0 = time.perf_counter()
        tc, ti = oracle.build_events(batch.points.coords, batch.points.ids, batch.offsets, threads=cores)
        qoff, _ = oracle.search_events(tc, ti, batch.offsets, qev, eb.mins, eb.maxs, threads=cores)
        ev_s = time.perf_counter() - t0
        res["events"] = {"value": n_events / ev_s, "unit": "events/s", "cores": cores, "kind": "port",
                         "hits": int(qoff[-1]),
                         "sample": f"{n_events} events, {batch.points.n} points, every point's box"}
    return res


if __name__ == "__main__":
    sys.exit(main())
