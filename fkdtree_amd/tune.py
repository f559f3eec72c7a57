"""Autotuner of the build and search launch parameters.

    python -m fkdtree_amd.tune [--workload mixed_1m] [--reps 3]

On a workload it times, with HIP events:

* the search, sweeping `query_key_bits` (Morton bits per dimension of the
  box order; 0 keeps the caller's order), `fill_order` (heaviest search
  blocks first in the fill pass), `fill_window` (boxes reordered by work
  inside windows of that many boxes for the fill pass), `count_window` and
  `key_volume_bits`, then, on top of the fastest of those, the traversal
  schedules `bucket` / `count_bucket` / `bucket_park` (bottom subtrees
  tested node by node);
* the build, sweeping `fin_log` (subtrees of fewer than 2^fin_log points
  are finished by one workgroup in LDS: 511, 1,023 or 2,047 points; fewer
  global level launches against bigger finisher workgroups) and
  `build_radix_items` (keys per thread of the presort's digit passes: tiles
  of 2,048 or 4,096 keys) and, at dim 3, `level_pairs` (one array moved
  through two depths per launch, or two arrays through one).

The fastest settings go to `tuning.json` next to this file, keyed by GPU
architecture, dimension, log2 of the (per-event) point count and kind of
handle.  Trees created by this package apply the nearest stored entry
(`apply`).  Results never depend on these parameters: only time does.
"""
from __future__ import annotations

import argparse
import itertools
import json
import math
import os
from typing import Dict, Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "tuning.json")

CANDIDATES = {"query_key_bits": (0, 5, 6, 7, 8, 10), "fill_order": (0, 1), "fill_window": (0, 256, 1024, 4096),
              "count_window": (0, 1024, 4096), "key_volume_bits": (0, 3, 4)}
BUILD_CANDIDATES = {"fin_log": (9, 10, 11), "build_radix_items": (8, 16), "level_pairs": (1, 0)}
# Traversal schedules, swept on top of the fastest CANDIDATES setting: the
# walk, or bucketed fill passes (queued, or parked with a threshold).
BUCKET_CANDIDATES = ({"bucket": 0, "count_bucket": 0, "bucket_park": 8}, {"bucket": 2, "count_bucket": 0, "bucket_park": 8},
                     {"bucket": 1, "bucket_park": 6, "count_bucket": 0}, {"bucket": 1, "bucket_park": 4, "count_bucket": 0},
                     {"bucket": 2, "count_bucket": 2, "bucket_park": 8})


_ARCH: Optional[str] = None
_TABLES: Dict[str, tuple] = {}  # path -> (modification time, table)


def _arch() -> str:
    """GPU architecture name (e.g. gfx950), asked once per process."""
    global _ARCH
    if _ARCH is None:
        from . import device_info

        name = device_info()["name"]
        _ARCH = name.split("(")[-1].split(":")[0].rstrip(")") or "unknown"
    return _ARCH


def _key(dim: int, n: int, events: bool = False) -> str:
    """Table key: dimension, log2 of the (mean per-event) tree size, and
    whether the handle holds several events (their searches tune apart)."""
    return f"dim{dim}_{'ev_' if events else ''}log2n{int(round(math.log2(max(n, 2))))}"


def load_table(path: str = TABLE) -> Dict:
    """The tuning table at `path` ({} if absent), parsed once per process and
    again only when the file changes (every new handle looks it up)."""
    if not os.path.exists(path):
        return {}
    mtime = os.path.getmtime(path)
    hit = _TABLES.get(path)
    if hit is not None and hit[0] == mtime:
        return hit[1]
    with open(path) as f:
        table = json.load(f)
    _TABLES[path] = (mtime, table)
    return table


def lookup(arch: str, dim: int, n: int, table: Optional[Dict] = None,
           events: bool = False) -> Optional[Dict[str, int]]:
    """Stored parameters of the nearest size bucket for (arch, dim, one tree
    or several events), if any."""
    table = load_table() if table is None else table
    entries = table.get(arch, {})
    best, dist = None, None
    want = math.log2(max(n, 2))
    prefix = f"dim{dim}_ev_" if events else f"dim{dim}_"
    for k, v in entries.items():
        if not k.startswith(prefix) or (not events and k.startswith(f"dim{dim}_ev_")):
            continue
        d = abs(int(k.split("log2n")[1]) - want)
        if dist is None or d < dist:
            best, dist = v["params"], d
    return best


def apply(tree, arch: Optional[str] = None) -> Optional[Dict[str, int]]:
    """Set the stored parameters on a tree handle; returns them (or None)."""
    try:
        params = lookup(arch or _arch(), tree.dim, max(tree.n // max(tree.n_events, 1), 1),
                        events=tree.n_events > 1)
    except Exception:
        return None
    applied = {}
    for k, v in (params or {}).items():
        try:            # keys a newer or older library does not know are skipped
            tree.set_tuning(**{k: v})
            applied[k] = v
        except Exception:
            pass
    return applied or None


def sweep(tree, nq: int, d_mins: int, d_maxs: int, reps: int = 3, candidates=None, d_qev: Optional[int] = None,
          settings=None):
    """Time the two-pass search for every candidate (the product of
    `candidates`, or the list `settings`); [(params, ms)] sorted."""
    from . import DeviceArray, Stream, Timer

    cands = candidates or CANDIDATES
    s = Stream()
    d_off = DeviceArray(nq + 1, np.uint64)
    tree.search_device(nq, d_mins, d_maxs, d_off.ptr, None, 0, d_qev=d_qev, stream=s, fill=False)
    total = int(d_off.download(s)[-1])
    d_out = DeviceArray(max(total, 1), np.uint32)
    tm = Timer()
    out = []
    keys = list(cands)
    if settings is None:
        settings = [dict(zip(keys, values)) for values in itertools.product(*(cands[k] for k in keys))]
    for params in settings:
        tree.set_tuning(**params)

        def run():
            tree.search_device(nq, d_mins, d_maxs, d_off.ptr, d_out.ptr, total, d_qev=d_qev, stream=s)

        run()
        s.sync()
        tm.start(s)
        for _ in range(reps):
            run()
        tm.stop(s)
        out.append((params, tm.ms() / reps))
    out.sort(key=lambda x: x[1])
    return out


def sweep_build(tree, reps: int = 5, candidates=None, settings=None):
    """Time the build for every candidate geometry (the product of
    `candidates`, or the list `settings`); [(params, ms)] sorted."""
    from . import Stream, Timer

    cands = candidates or BUILD_CANDIDATES
    s = tree.stream if tree.stream is not None else Stream()
    tm = Timer()
    out = []
    keys = list(cands)
    if settings is None:
        settings = [dict(zip(keys, values)) for values in itertools.product(*(cands[k] for k in keys))]
    for params in settings:
        tree.set_tuning(**params)
        tree.build()             # lays the tilings out for the new geometry
        tm.start(s)
        for _ in range(reps):
            tree.build_async()
        tm.stop(s)
        out.append((params, tm.ms() / reps))
    out.sort(key=lambda x: x[1])
    return out


def search_stage1():
    """The first search sweep: the product of CANDIDATES, each setting with the
    walk schedule named explicitly (set_tuning leaves unnamed keys as they
    were, so a setting must name every key the sweep ever changes)."""
    keys = list(CANDIDATES)
    base = dict(BUCKET_CANDIDATES[0])
    return [{**dict(zip(keys, values)), **base} for values in itertools.product(*(CANDIDATES[k] for k in keys))]


def search_stage2(best: Dict[str, int]):
    """The traversal schedules on top of the fastest first-stage setting."""
    return [{**best, **b} for b in BUCKET_CANDIDATES]


KEEP_SWEEP = 16  # sweep rows kept per table entry
FINAL_TOP = 8    # leading settings timed again ...
FINAL_REPS = 8   # ... with this many times the repetitions


def autotune(workload: str = "mixed_1m", reps: int = 3, write: bool = True, path: str = TABLE) -> Dict:
    """Sweep on a workload: one tree (mixed_1m, uniform_100k, uniform_4m) or
    events_256 (the first 256 events of the 2,048-event set, every point's
    box of half-width 0.01 searched in its own event's tree)."""
    from . import DeviceArray, FKDEvents, FKDTree, Stream, gen

    s = Stream()
    d_q = None
    if workload == "events_256":
        seed = gen.workload_seed("events_2048")
        sizes = gen.event_sizes(2048, seed, 1000, gen.LOG2_200)[:256]
        batch = gen.concat_events([gen.mixed_event(int(sz), seed, e) for e, sz in enumerate(sizes)])
        pts, boxes = batch.points, gen.event_self_boxes(batch, 0.01)
        qev = np.repeat(np.arange(256, dtype=np.uint32), np.diff(batch.offsets))
        d_q = DeviceArray.from_host(qev, s)
    else:
        pts, boxes = {"mixed_1m": gen.workload_mixed_1m, "uniform_100k": gen.workload_uniform_100k,
                      "uniform_4m": gen.workload_uniform_4m}[workload]()
    d_c = DeviceArray.from_host(pts.coords, s)
    d_i = DeviceArray.from_host(pts.ids, s)
    d_mi = DeviceArray.from_host(boxes.mins, s)
    d_ma = DeviceArray.from_host(boxes.maxs, s)
    if d_q is not None:
        tree = FKDEvents.from_device(pts.dim, batch.offsets, d_c.ptr, d_i.ptr, s)
        key = _key(pts.dim, pts.n // 256, events=True)
    else:
        tree = FKDTree.from_device(pts.dim, pts.n, d_c.ptr, d_i.ptr, s)
        key = _key(pts.dim, pts.n)
    tree.build()
    dq = d_q.ptr if d_q is not None else None
    res = sweep(tree, boxes.q, d_mi.ptr, d_ma.ptr, reps=reps, d_qev=dq, settings=search_stage1())
    # the traversal schedules on top of the fastest box order and windows
    sres = sweep(tree, boxes.q, d_mi.ptr, d_ma.ptr, reps=reps, d_qev=dq, settings=search_stage2(res[0][0]))
    res = sorted(sres + res[1:], key=lambda x: x[1])
    # single sweeps differ by noise near the top: the leaders are timed again, longer
    top = sweep(tree, boxes.q, d_mi.ptr, d_ma.ptr, reps=FINAL_REPS * reps, d_qev=dq,
                settings=[p for p, _ in res[:FINAL_TOP]])
    res = top + res[FINAL_TOP:]
    bres = sweep_build(tree, reps=max(reps, 5))
    btop = sweep_build(tree, reps=FINAL_REPS * max(reps, 5), settings=[p for p, _ in bres[:FINAL_TOP]])
    bres = btop + bres[FINAL_TOP:]
    tree.close()
    arch = _arch()
    # the table keeps the fastest settings of the sweeps; the whole sweeps are returned
    entry = {"params": {**res[0][0], **bres[0][0]}, "ms": res[0][1], "build_ms": bres[0][1], "workload": workload,
             "sweep": [{"params": p, "ms": ms} for p, ms in res[:KEEP_SWEEP]],
             "build_sweep": [{"params": p, "ms": ms} for p, ms in bres]}
    if write:
        table = load_table(path)
        table.setdefault(arch, {})[key] = entry
        with open(path, "w") as f:
            json.dump(table, f, indent=1, sort_keys=True)
    return {"arch": arch, "key": key, **entry, "sweep": [{"params": p, "ms": ms} for p, ms in res]}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="mixed_1m")
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--no-write", action="store_true")
    ap.add_argument("--out", default=TABLE, help="tuning table to update")
    a = ap.parse_args(argv)
    r = autotune(a.workload, a.reps, write=not a.no_write, path=a.out)
    print(json.dumps({k: r[k] for k in ("arch", "key", "params", "ms", "build_ms")}))
    for row in r["sweep"]:
        print(json.dumps(row))
    for row in r["build_sweep"]:
        print(json.dumps({"build": row}))


if __name__ == "__main__":
    main()
