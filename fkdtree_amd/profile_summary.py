"""Summaries of rocprofv3 result databases (rocpd SQLite, *_results.db).

    python -m fkdtree_amd.profile_summary trace  DIR            # kernel stats
    python -m fkdtree_amd.profile_summary counter DIR [DIR...]  # PMC per kernel
    python -m fkdtree_amd.profile_summary traffic --fetch D --write D \
        --calib-fetch D --calib-write D --calib-bytes B --out F.json
    python -m fkdtree_amd.profile_summary gaps DIR --split KERNEL  # idle time per batch

`traffic` scales FETCH_SIZE / WRITE_SIZE (KiB per dispatch, the median over
a kernel's dispatches) of every kernel
by the ratio true bytes / counted bytes of a plain copy kernel of known size
(copy_u32_kernel, `bench.py --calibrate`), and writes a JSON table that
bench.py reads for the `roofline.traffic` figure.
"""
from __future__ import annotations

import argparse
import collections
import glob
import json
import os
import sqlite3


def _db(path: str) -> sqlite3.Connection:
    if os.path.isdir(path):
        hits = glob.glob(os.path.join(path, "**", "*_results.db"), recursive=True)
        if not hits:
            raise FileNotFoundError(f"no *_results.db under {path}")
        path = hits[0]
    return sqlite3.connect(path)


def short(name: str) -> str:
    return name.split("(")[0].replace("void ", "").strip()


def kernel_stats(path: str):
    """[(kernel, calls, avg_us, total_ms, percent)] sorted by total time
    (top_kernels holds microseconds)."""
    rows = _db(path).execute("select name, total_calls, total_duration, average, percentage from top_kernels")
    return [(short(n), int(c), a, t / 1e3, p) for n, c, t, a, p in rows]


def counters(path: str, stat: str = "mean"):
    """{(kernel, counter): value per dispatch}: the mean over the kernel's
    dispatches, or their median (stat="median": a run's sizing calls, which
    write no output, do not pull it down)."""
    agg = collections.defaultdict(list)
    for k, c, v in _db(path).execute("select kernel_name, counter_name, value from counters_collection"):
        agg[(short(k), c)].append(v)
    if stat == "median":
        return {key: sorted(v)[len(v) // 2] for key, v in agg.items()}
    return {key: sum(v) / len(v) for key, v in agg.items()}


def dispatches(path: str):
    """[(kernel, start_ns, end_ns)] of every kernel dispatch, by start time."""
    rows = _db(path).execute("select name, start, end from kernels order by start")
    return [(short(n), int(a), int(b)) for n, a, b in rows]


def batch_gaps(path: str, split: str):
    """Per batch of dispatches (a batch starts at each dispatch of `split`):
    (span_us, busy_us, idle_us, dispatches), where busy is the union of the
    batch's kernel intervals and idle the rest of its span up to the next
    batch's first dispatch (or its own last end)."""
    ds = dispatches(path)
    starts = [i for i, d in enumerate(ds) if d[0] == split]
    out = []
    for j, i0 in enumerate(starts):
        i1 = starts[j + 1] if j + 1 < len(starts) else len(ds)
        seg = ds[i0:i1]
        t0 = seg[0][1]
        t1 = ds[i1][1] if i1 < len(ds) else max(d[2] for d in seg)
        busy, cur_a, cur_b = 0, None, None
        for _, a, b in seg:
            b = min(b, t1)
            if cur_b is None or a > cur_b:
                if cur_b is not None:
                    busy += cur_b - cur_a
                cur_a, cur_b = a, b
            else:
                cur_b = max(cur_b, b)
        if cur_b is not None:
            busy += cur_b - cur_a
        out.append(((t1 - t0) / 1e3, busy / 1e3, (t1 - t0 - busy) / 1e3, len(seg)))
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("trace")
    t.add_argument("dir")
    c = sub.add_parser("counter")
    c.add_argument("dirs", nargs="+")
    tr = sub.add_parser("traffic")
    tr.add_argument("--fetch", required=True)
    tr.add_argument("--write", required=True)
    tr.add_argument("--calib-fetch", required=True)
    tr.add_argument("--calib-write", required=True)
    tr.add_argument("--calib-bytes", type=float, required=True)
    tr.add_argument("--out", required=True)
    g = sub.add_parser("gaps")
    g.add_argument("dir")
    g.add_argument("--split", required=True, help="kernel whose dispatches start the batches")
    a = ap.parse_args(argv)
    if a.cmd == "trace":
        print(f"{'kernel':60s} {'calls':>6s} {'avg_us':>10s} {'total_ms':>9s} {'pct':>6s}")
        for k, n, avg, tot, pct in kernel_stats(a.dir):
            print(f"{k[:60]:60s} {n:6d} {avg:10.1f} {tot:9.3f} {pct:6.1f}")
    elif a.cmd == "gaps":
        print(f"{'batch':>5s} {'dispatches':>10s} {'span_us':>10s} {'busy_us':>10s} {'idle_us':>9s} {'idle%':>6s}")
        for j, (span, busy, idle, n) in enumerate(batch_gaps(a.dir, a.split)):
            print(f"{j:5d} {n:10d} {span:10.1f} {busy:10.1f} {idle:9.1f} {100 * idle / span:6.2f}")
    elif a.cmd == "counter":
        for d in a.dirs:
            for (k, cn), v in sorted(counters(d).items()):
                print(f"{k[:60]:60s} {cn:28s} {v:14.4g}")
    else:
        cf = a.calib_bytes / 1024.0 / counters(a.calib_fetch)[("fkd::copy_u32_kernel", "FETCH_SIZE")]
        cw = a.calib_bytes / 1024.0 / counters(a.calib_write)[("fkd::copy_u32_kernel", "WRITE_SIZE")]
        fetch, write = counters(a.fetch, "median"), counters(a.write, "median")
        table = {"fetch_scale": cf, "write_scale": cw, "kernels": {}}
        for (k, cn), v in fetch.items():
            if cn == "FETCH_SIZE":
                table["kernels"].setdefault(k, {})["fetch_bytes"] = v * 1024.0 * cf
        for (k, cn), v in write.items():
            if cn == "WRITE_SIZE":
                table["kernels"].setdefault(k, {})["write_bytes"] = v * 1024.0 * cw
        with open(a.out, "w") as f:
            json.dump(table, f, indent=1, sort_keys=True)
        for k, v in sorted(table["kernels"].items()):
            print(f"{k[:60]:60s} fetch {v.get('fetch_bytes', 0) / 1e6:10.1f} MB  write {v.get('write_bytes', 0) / 1e6:10.1f} MB")


if __name__ == "__main__":
    main()
