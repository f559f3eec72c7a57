"""Continuous-batching event server over the GPU tree backend.

Events (independent point clouds) arrive over time.  Each must be built into
its own tree and have every one of its points searched with a box of
half-width `half_width`.

The server drives the GPU from one host thread:

* Requests that have arrived queue up in arrival order.  Whenever a stream
  slot is free, the longest prefix of the queue that fits the limits (points,
  events, free HBM) becomes one batch.  Requests that arrive while other
  batches run on the GPU join the next batch: this is continuous batching,
  not a single queue drained batch by batch.
* A batch is one FKDEvents handle: one build launch per depth for all its
  events, then one search of all its boxes (count, scan and fill enqueued
  together), on its own stream.
  Several batches overlap on the device.
* When a batch completes, its results are checked for completeness (the
  output is regrown and the search re-run if needed) and its latency is
  recorded.  Its build / search scratch is trimmed, and its trees stay
  resident for later queries (`query`), least recently used first out when
  HBM runs short.
* Admission is by free HBM: a batch is admitted only when its estimated
  footprint fits the budget, i.e. `hbm_fraction` of what hipMemGetInfo
  reports free.  Resident trees are evicted first, then the server waits for
  batches in flight to finish.

Latency of an event = completion time of its batch - arrival time (host
clock).  `ServeReport` holds p50 / p99 latency and events per second.
"""
from __future__ import annotations

import collections
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import gen


@dataclass
class Request:
    event: int
    coords: np.ndarray      # float32 [dim, n]
    ids: np.ndarray         # uint32 [n]
    arrival: float = 0.0    # seconds after the start of the stream

    @property
    def n(self) -> int:
        return int(self.ids.shape[0])


@dataclass
class EventResult:
    event: int
    hits: int
    arrival: float
    done: float
    batch: int
    offsets: Optional[np.ndarray] = None   # uint64 [n+1], when results are kept
    ids: Optional[np.ndarray] = None

    @property
    def latency(self) -> float:
        return self.done - self.arrival


@dataclass
class ServeReport:
    results: Dict[int, EventResult]
    n_batches: int
    wall_s: float
    peak_inflight: int
    evictions: int
    batch_sizes: List[int] = field(default_factory=list)

    def latencies_ms(self) -> np.ndarray:
        return np.array([r.latency * 1e3 for r in self.results.values()])

    def summary(self) -> dict:
        lat = self.latencies_ms()
        first = min(r.arrival for r in self.results.values())
        last = max(r.done for r in self.results.values())
        return {"events": len(self.results), "batches": self.n_batches,
                "events_per_s": len(self.results) / max(last - first, 1e-9),
                "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                "mean_batch_events": float(np.mean(self.batch_sizes)) if self.batch_sizes else 0.0,
                "peak_inflight_batches": self.peak_inflight, "evictions": self.evictions}


def batch_bytes(points: int, hits_per_point: float) -> int:
    """Estimated device footprint of one batch during build + search:
    build scratch and tree (about 300 B per point at dim 3), boxes and search
    plan (about 100 B per point), and the result ids."""
    return int(points * (400 + 8 * hits_per_point)) + (64 << 20)


def plan_batch(queue: Sequence[Request], budget_bytes: int, max_points: int, max_events: int,
               hits_per_point: float) -> int:
    """How many requests from the head of `queue` form the next batch: the
    longest prefix within the point, event and byte limits; at least one
    request when that one alone fits the byte budget, else zero."""
    n = 0
    pts = 0
    for r in queue:
        if n >= max_events:
            break
        if n and pts + r.n > max_points:
            break
        if batch_bytes(pts + r.n, hits_per_point) > budget_bytes:
            break
        pts += r.n
        n += 1
    return n


class _Batch:
    """One batch on its stream.  Ingestion is device side: each request's
    coordinates and ids are copied straight into one device array at the
    event's offset (no host concatenation or transposition), the handle is
    created from it, and the boxes around every point are made on the device
    (fkd_boxes_around).  All its device memory is allocated and freed in
    stream order on the batch's stream."""

    def __init__(self, idx: int, members: List[Request], server: "EventServer", stream):
        import fkdtree_amd as F

        self.idx = idx
        self.members = members
        self.stream = stream
        dim = int(members[0].coords.shape[0])
        offs = np.zeros(len(members) + 1, dtype=np.uint64)
        offs[1:] = np.cumsum([r.n for r in members])
        self.offsets = offs
        self.P = P = int(offs[-1])
        d_c = F.DeviceArray((dim, max(P, 1)), np.float32, stream=stream)
        d_i = F.DeviceArray(max(P, 1), np.uint32, stream=stream)
        hold = []  # host sources stay alive until the copies have run
        for r, a in zip(members, offs[:-1]):
            a = int(a)
            for m in range(dim):
                hold.append(np.ascontiguousarray(r.coords[m], dtype=np.float32))
                d_c.upload_at(hold[-1], 4 * (m * P + a), stream)
            hold.append(np.ascontiguousarray(r.ids, dtype=np.uint32))
            d_i.upload_at(hold[-1], 4 * a, stream)
        self.handle = F.FKDEvents.from_device(dim, offs, d_c.ptr, d_i.ptr, stream)  # synchronises the stream
        del hold
        d_c.free()
        d_i.free()
        self.d_mins = F.DeviceArray((dim, max(P, 1)), np.float32, stream=stream)
        self.d_maxs = F.DeviceArray((dim, max(P, 1)), np.float32, stream=stream)
        self.d_qev = F.DeviceArray(max(P, 1), np.uint32, stream=stream)
        self.handle.boxes_around(server.half_width, self.d_mins.ptr, self.d_maxs.ptr, self.d_qev.ptr, stream)
        self.d_off = F.DeviceArray(P + 1, np.uint64, stream=stream)
        self.cap = max(1, int(P * server.hits_per_point))
        self.d_out = F.DeviceArray(self.cap, np.uint32, stream=stream)
        self.done_ev = F.Event()
        self.launch()

    def launch(self, build: bool = True):
        if build:
            self.handle.build_async()
        self.handle.search_single_device(self.P, self.d_mins.ptr, self.d_maxs.ptr, self.d_off.ptr, self.d_out.ptr,
                                         self.cap, d_qev=self.d_qev.ptr, stream=self.stream)
        self.done_ev.record(self.stream)

    def free_search_buffers(self):
        self.d_mins = self.d_maxs = self.d_qev = self.d_off = self.d_out = None


class EventServer:
    def __init__(self, n_streams: int = 3, max_batch_points: int = 8_000_000, max_batch_events: int = 512,
                 hbm_fraction: float = 0.8, hbm_budget: Optional[int] = None, half_width: float = 0.01,
                 hits_per_point: float = 256.0, keep_results: bool = False, keep_resident: bool = True,
                 poll_sleep_s: float = 50e-6):
        import fkdtree_amd as F

        self.streams = [F.Stream() for _ in range(n_streams)]
        self.max_batch_points = max_batch_points
        self.max_batch_events = max_batch_events
        self.hbm_fraction = hbm_fraction
        self.hbm_budget = hbm_budget
        self.half_width = half_width
        self.hits_per_point = hits_per_point
        self.keep_results = keep_results
        self.keep_resident = keep_resident
        self.poll_sleep_s = poll_sleep_s
        self.resident: "collections.OrderedDict[int, _Batch]" = collections.OrderedDict()
        self.where: Dict[int, Tuple[int, int]] = {}   # event -> (batch, index in batch)
        self.evictions = 0
        self.resident_bytes = 0     # device bytes of the trimmed resident handles
        self.inflight_bytes = 0     # estimated footprint of the batches in flight

    # -- memory ---------------------------------------------------------------
    def _budget(self) -> int:
        """Bytes the next batch may use: `hbm_fraction` of the free HBM, and,
        with an explicit `hbm_budget`, what that budget leaves after the
        resident trees and the batches in flight."""
        import fkdtree_amd as F

        free = F.device_info()["free_bytes"]
        b = int(free * self.hbm_fraction)
        if self.hbm_budget is not None:
            b = min(b, self.hbm_budget - self.resident_bytes - self.inflight_bytes)
        return b

    def _evict_one(self) -> bool:
        if not self.resident:
            return False
        _, b = self.resident.popitem(last=False)
        for r in b.members:
            self.where.pop(r.event, None)
        self.resident_bytes -= b.resident_bytes
        b.handle.close()
        self.evictions += 1
        return True

    # -- serving loop -----------------------------------------------------------
    def run(self, requests: Sequence[Request]) -> ServeReport:
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        queue: "collections.deque[Request]" = collections.deque()
        inflight: List[_Batch] = []
        free_streams = list(self.streams)
        results: Dict[int, EventResult] = {}
        nb = 0
        sizes: List[int] = []
        peak = 0
        t0 = time.perf_counter()
        while pending or queue or inflight:
            now = time.perf_counter() - t0
            while pending and pending[0].arrival <= now:
                queue.append(pending.popleft())
            progressed = False
            # completions
            for b in list(inflight):
                if not b.done_ev.done():
                    continue
                total, ok = b.handle.search_status(b.stream)
                if not ok:
                    if total > b.cap:
                        import fkdtree_amd as F
                        b.cap = total
                        b.d_out = F.DeviceArray(total, np.uint32, stream=b.stream)
                    b.launch(build=False)
                    continue
                t_done = time.perf_counter() - t0
                self._finish(b, results, t_done)
                self.inflight_bytes -= b.est_bytes
                inflight.remove(b)
                free_streams.append(b.stream)
                progressed = True
            # admission
            while queue and free_streams:
                budget = self._budget()
                k = plan_batch(queue, budget, self.max_batch_points, self.max_batch_events, self.hits_per_point)
                if k == 0:
                    if self._evict_one():
                        continue
                    if inflight:
                        break           # wait for memory held by batches in flight
                    k = 1               # nothing else holds memory: try the head alone
                members = [queue.popleft() for _ in range(k)]
                s = free_streams.pop(0)
                b = _Batch(nb, members, self, s)
                b.est_bytes = batch_bytes(b.P, self.hits_per_point)
                self.inflight_bytes += b.est_bytes
                inflight.append(b)
                sizes.append(k)
                nb += 1
                peak = max(peak, len(inflight))
                progressed = True
            if not progressed:
                if pending and not queue and not inflight:
                    time.sleep(max(0.0, min(pending[0].arrival - (time.perf_counter() - t0), 0.01)))
                else:
                    time.sleep(self.poll_sleep_s)
        return ServeReport(results, nb, time.perf_counter() - t0, peak, self.evictions, sizes)

    def _finish(self, b: _Batch, results: Dict[int, EventResult], t_done: float):
        off = b.d_off.download(b.stream)
        ids = b.d_out.download(b.stream, count=int(off[-1])) if self.keep_results else None
        for i, r in enumerate(b.members):
            a, c = int(b.offsets[i]), int(b.offsets[i + 1])
            res = EventResult(r.event, int(off[c] - off[a]), r.arrival, t_done, b.idx)
            if self.keep_results:
                res.offsets = (off[a:c + 1] - off[a]).astype(np.uint64)
                res.ids = ids[int(off[a]):int(off[c])].copy()
            results[r.event] = res
        b.free_search_buffers()
        if self.keep_resident:
            b.handle.trim()
            b.resident_bytes = b.handle.device_bytes()
            self.resident_bytes += b.resident_bytes
            self.resident[b.idx] = b
            for i, r in enumerate(b.members):
                self.where[r.event] = (b.idx, i)
        else:
            b.handle.close()

    # -- queries on resident trees ------------------------------------------------
    def query(self, event: int, mins: np.ndarray, maxs: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
        """Boxes searched in the resident tree of `event`: (offsets, ids)."""
        bi, i = self.where[event]
        b = self.resident[bi]
        self.resident.move_to_end(bi)
        qev = np.full(mins.shape[1], i, dtype=np.uint32)
        return b.handle.search(mins, maxs, qev)

    def close(self):
        while self._evict_one():
            pass


def poisson_requests(n_events: int, rate_hz: float, size_scale: float = 1.0, seed_offset: int = 0) -> List[Request]:
    """The serving workload: Poisson arrivals at `rate_hz`, sizes log-uniform
    in [1,000, 500,000] points (scaled by size_scale), mixed distribution."""
    times, sizes = gen.serving_arrivals(n_events, rate_hz)
    out = []
    for e in range(n_events):
        n = max(1, int(sizes[e] * size_scale))
        p = gen.serving_event(n, e + seed_offset)
        out.append(Request(e, p.coords, p.ids, float(times[e])))
    return out
