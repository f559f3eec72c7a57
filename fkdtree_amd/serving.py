"""Continuous-batching event server over the GPU tree backend.

Events (independent point clouds) arrive over time.  Each must be built into
its own tree and have every one of its points searched with a box of
half-width `half_width`.

The server drives the GPU from one serving thread, with one worker thread
per stream that creates batches (creation synchronises the batch's stream):

* Requests that have arrived queue up in arrival order.  Whenever a stream
  slot is free, the longest prefix of the queue that fits the limits (points,
  events, free HBM) becomes one batch.  Requests that arrive while other
  batches run on the GPU join the next batch: this is continuous batching,
  not a single queue drained batch by batch.
* A batch is one FKDEvents handle: one build launch per depth for all its
  events, then one search of all its boxes (count, scan and fill enqueued
  together), on its own stream.  A worker thread creates it and enqueues its
  work; the serving thread meanwhile completes and admits other batches.
  Several batches overlap on the device.
* When a batch completes, its results are checked for completeness (the
  output is regrown and the search re-run if needed) and its latency is
  recorded.  Its trees stay resident for later queries (`query`).  When HBM
  runs short, resident trees first give up their build / search scratch
  (fkd_trim, oldest first), then the least recently used are evicted.
* Admission is by free HBM: a batch is admitted only when its estimated
  footprint fits the budget, i.e. `hbm_fraction` of what hipMemGetInfo
  reports free.  Resident trees are trimmed and then evicted first, then the server waits for
  batches in flight to finish.

* Ingestion: events that arrive as FKDPoint records (uint32 [n, dim+1],
  `Request.records`) are written into a page-locked ring (`IngestRing`) as
  they arrive: the serving thread reserves each event's slots in arrival
  order and receiver threads copy the records in, so the copying never
  holds up admission; a batch of consecutive events is then one contiguous range of
  the ring, handed to fkd_create_events as it lies (one asynchronous copy at
  full link rate, transposed on the device).  Events given as SoA arrays are
  copied per event into the batch's device arrays instead.

Latency of an event = completion time of its batch - arrival time (host
clock).  `ServeReport` holds p50 / p99 latency and events per second.
"""
from __future__ import annotations

import collections
import concurrent.futures
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import gen


@dataclass
class Request:
    event: int
    coords: Optional[np.ndarray]    # float32 [dim, n], or None when `records` is given
    ids: Optional[np.ndarray]       # uint32 [n]
    arrival: float = 0.0            # seconds after the start of the stream
    records: Optional[np.ndarray] = None  # FKDPoint records uint32 [n, dim+1] (coordinate bits, id)
    ring_start: int = -1            # slot of the records in the ingest ring, once received
    ingested: Optional[concurrent.futures.Future] = None  # the copy into the ring, once started

    @property
    def n(self) -> int:
        return int(self.records.shape[0] if self.records is not None else self.ids.shape[0])

    @property
    def dim(self) -> int:
        return int(self.records.shape[1] - 1 if self.records is not None else self.coords.shape[0])


# Page-locked ring buffers no longer used by a server, by (capacity, dim),
# handed to the next server instead of being freed: freeing page-locked
# memory empties the device's stream-ordered memory pool, and afterwards the
# pool no longer keeps what is freed (measured: 17 GB released at one
# hipHostFree, then every later synchronisation released the freed
# buffers), so every batch of every later server would map its buffers from
# the driver anew; runs then stalled for up to seconds.  The buffers live
# until the process ends.
_FREE_RINGS: Dict[Tuple[int, int], List[np.ndarray]] = {}


def _ring_buffer(capacity: int, dim: int) -> np.ndarray:
    free = _FREE_RINGS.get((capacity, dim))
    if free:
        return free.pop()
    import fkdtree_amd as F
    buf = F.pinned_empty((capacity, dim + 1), np.uint32)
    buf.fill(0)  # touch every page once, before events arrive
    return buf


class IngestRing:
    """Ring of FKDPoint records (uint32 [capacity, dim+1]: coordinate bits,
    then the id; FKDPoint.h:13-104) in page-locked host memory, into which
    arriving events are written in arrival order.  An event occupies one
    contiguous range; an event that does not fit before the end starts again
    at slot 0 (the tail end is skipped).  Space is released in the same
    order.  The reference's OpenCL backend instead unpacks every point on the
    host (FKDTree_opencl.h:22-33)."""

    def __init__(self, capacity: int, dim: int = 3, buffer: Optional[np.ndarray] = None):
        self.pooled = buffer is None
        if buffer is None:
            buffer = _ring_buffer(capacity, dim)
        assert buffer.shape == (capacity, dim + 1) and buffer.dtype == np.uint32
        self.buf = buffer
        self.capacity = capacity
        self.dim = dim
        self.tail = 0     # next free slot
        self.used = 0     # slots held, skipped tail ends included
        self.spans: "collections.deque[Tuple[int, int]]" = collections.deque()  # (slots held, n)

    def reserve(self, n: int) -> Optional[int]:
        """Claim the slots of an event of n points in arrival order; its
        first slot, or None when full.  write() fills them."""
        if n > self.capacity:
            raise ValueError(f"event of {n} points exceeds the ring ({self.capacity})")
        start, skip = self.tail, 0
        if start + n > self.capacity:
            start, skip = 0, self.capacity - self.tail
        if self.used + skip + n > self.capacity:
            return None
        self.tail = (start + n) % self.capacity
        self.used += skip + n
        self.spans.append((skip + n, n))
        return start

    def write(self, start: int, records: np.ndarray) -> None:
        """Copy an event's records into its reserved slots (any thread)."""
        self.buf[start:start + records.shape[0]] = records

    def put(self, records: np.ndarray) -> Optional[int]:
        """Reserve and copy in one step; the first slot, or None when full."""
        start = self.reserve(int(records.shape[0]))
        if start is not None:
            self.write(start, records)
        return start

    def release(self, count: int) -> None:
        """Free the oldest `count` events."""
        for _ in range(count):
            held, _ = self.spans.popleft()
            self.used -= held
        if not self.spans:
            self.tail = 0
            self.used = 0

    def view(self, start: int, n: int) -> np.ndarray:
        return self.buf[start:start + n]

    def close(self) -> None:
        """Hand a page-locked buffer back for the next ring (it is not freed)."""
        if self.pooled and self.buf is not None:
            _FREE_RINGS.setdefault((self.capacity, self.dim), []).append(self.buf)
        self.buf = None


def same_kind_prefix(queue: Sequence[Request], k: int) -> int:
    """How many of the first k queued requests arrive the way the head does
    (FKDPoint records through the ring, or SoA arrays): a batch is built from
    one kind only."""
    if k <= 1:
        return k
    head = queue[0].records is not None
    for i in range(1, k):
        if (queue[i].records is not None) != head:
            return i
    return k


def contiguous_prefix(queue: Sequence[Request], k: int) -> int:
    """How many of the first k queued events lie end to end in the ring."""
    if k <= 1:
        return k
    nxt = queue[0].ring_start + queue[0].n
    for i in range(1, k):
        if queue[i].ring_start < 0 or queue[i].ring_start != nxt:
            return i
        nxt += queue[i].n
    return k


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
        """Bookkeeping only (the serving thread); create() does the GPU work."""
        self.idx = idx
        self.members = members
        self.stream = stream
        offs = np.zeros(len(members) + 1, dtype=np.uint64)
        offs[1:] = np.cumsum([r.n for r in members])
        self.offsets = offs
        self.P = int(offs[-1])
        # one range of the page-locked ring: one copy, transposed on the device
        self.recs = server.ring.view(members[0].ring_start, self.P) if members[0].ring_start >= 0 else None

    def create(self, server: "EventServer") -> "_Batch":
        """Handle, device buffers and the launch of build + search.  Runs on a
        worker thread: creating a handle synchronises its stream (upload and
        input check), which waits behind the other batches' kernels; the
        library calls release the interpreter lock meanwhile."""
        import fkdtree_amd as F

        offs, P, stream = self.offsets, self.P, self.stream
        dim = self.members[0].dim
        for r in self.members:          # the members' copies into the ring
            if r.ingested is not None:
                r.ingested.result()
        if self.recs is not None:
            self.handle = F.FKDEvents.from_records(self.recs, offs, stream)  # synchronises the stream
        else:
            self.handle = self._from_soa(self.members, offs, dim, stream)
        self.d_mins = F.DeviceArray((dim, max(P, 1)), np.float32, stream=stream)
        self.d_maxs = F.DeviceArray((dim, max(P, 1)), np.float32, stream=stream)
        self.d_qev = F.DeviceArray(max(P, 1), np.uint32, stream=stream)
        self.handle.boxes_around(server.half_width, self.d_mins.ptr, self.d_maxs.ptr, self.d_qev.ptr, stream)
        self.d_off = F.DeviceArray(P + 1, np.uint64, stream=stream)
        # CSR offsets at the event boundaries are gathered on completion
        self.d_bidx = F.DeviceArray.from_host(self.offsets, stream)
        self.cap = max(1, int(P * server.hits_per_point))
        self.d_out = F.DeviceArray(self.cap, np.uint32, stream=stream)
        self.done_ev = F.Event()
        self.launch()
        return self

    @staticmethod
    def _from_soa(members, offs, dim, stream):
        import fkdtree_amd as F

        P = int(offs[-1])
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
        handle = F.FKDEvents.from_device(dim, offs, d_c.ptr, d_i.ptr, stream)  # synchronises the stream
        del hold
        d_c.free()
        d_i.free()
        return handle

    def collect(self, keep_results: bool, t0: float) -> bool:
        """After the batch's event: its status and results.  False when the
        output was too small and the search was launched again.  Done on the
        serving thread: handing it to a worker delays freeing the stream by a
        thread wake-up and measured slower (p50 2.2 -> 3.1 ms at 1,000 ev/s)."""
        import fkdtree_amd as F

        total, ok = self.handle.search_status(self.stream)
        if not ok:
            if total > self.cap:
                self.cap = total
                self.d_out = F.DeviceArray(total, np.uint32, stream=self.stream)
            self.launch(build=False)
            return False
        if keep_results:
            self.res_off = self.d_off.download(self.stream)
            self.res_ids = self.d_out.download(self.stream, count=int(self.res_off[-1]))
        else:
            self.res_off, self.res_ids = self.event_offsets(), None   # only the event boundaries
        self.t_done = time.perf_counter() - t0
        return True

    def launch(self, build: bool = True):
        if build:
            self.handle.build_async()
        self.handle.search_single_device(self.P, self.d_mins.ptr, self.d_maxs.ptr, self.d_off.ptr, self.d_out.ptr,
                                         self.cap, d_qev=self.d_qev.ptr, stream=self.stream)
        self.done_ev.record(self.stream)

    def event_offsets(self) -> np.ndarray:
        """CSR offsets at the event boundaries of the batch (device gather)."""
        import fkdtree_amd as F
        from fkdtree_amd._lib import check, lib

        out = F.DeviceArray(self.offsets.shape[0], np.uint64, stream=self.stream)
        check(lib().fkd_gather_u64(out.ptr, self.d_off.ptr, self.d_bidx.ptr, self.offsets.shape[0],
                                   self.stream.handle))
        return out.download(self.stream)

    def free_search_buffers(self):
        self.d_mins = self.d_maxs = self.d_qev = self.d_off = self.d_out = self.d_bidx = None


class EventServer:
    def __init__(self, n_streams: int = 3, max_batch_points: int = 8_000_000, max_batch_events: int = 512,
                 hbm_fraction: float = 0.8, hbm_budget: Optional[int] = None, half_width: float = 0.01,
                 hits_per_point: float = 256.0, keep_results: bool = False, keep_resident: bool = True,
                 poll_sleep_s: float = 50e-6, ring_points: int = 0, dim: int = 3, ingest_threads: int = 4):
        import fkdtree_amd as F

        self.streams = [F.Stream() for _ in range(n_streams)]
        # page-locked ingest ring for events that arrive as FKDPoint records
        self.ring = IngestRing(ring_points, dim) if ring_points else None
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
        # batches are created (upload, input check, launch) on worker threads,
        # one per stream, so the serving thread keeps admitting and completing
        self.pool = concurrent.futures.ThreadPoolExecutor(max_workers=n_streams)
        # arriving records are copied into the ring by receiver threads; the
        # serving thread only reserves their slots
        self.ingest = (concurrent.futures.ThreadPoolExecutor(max_workers=max(1, ingest_threads))
                       if ring_points else None)

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

    def _reclaim_one(self) -> bool:
        """Make room for a batch: release the build / search scratch of the
        oldest resident tree that still holds it (fkd_trim), else evict the
        least recently used tree."""
        for b in self.resident.values():
            if not b.trimmed:
                before = b.resident_bytes
                b.handle.trim()
                b.trimmed = True
                b.resident_bytes = b.handle.device_bytes()
                self.resident_bytes += b.resident_bytes - before
                return True
        return self._evict_one()

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
        creating: "collections.deque[Tuple[concurrent.futures.Future, _Batch]]" = collections.deque()
        inflight: List[_Batch] = []
        free_streams = list(self.streams)
        results: Dict[int, EventResult] = {}
        nb = 0
        sizes: List[int] = []
        peak = 0
        t0 = time.perf_counter()
        while pending or queue or creating or inflight:
            now = time.perf_counter() - t0
            while pending and pending[0].arrival <= now:
                r = pending[0]
                if r.records is not None:
                    if self.ring is None:
                        raise ValueError("events given as records need an ingest ring (ring_points)")
                    start = self.ring.reserve(r.n)
                    if start is None:
                        break       # ring full: later arrivals wait for batches to take their events
                    r.ring_start = start
                    r.ingested = self.ingest.submit(self.ring.write, start, r.records)
                queue.append(pending.popleft())
            progressed = False
            # created batches (launched by their worker); the ring frees
            # events oldest first, so ranges are released in admission order
            while creating and creating[0][0].done():
                fut, b = creating.popleft()
                fut.result()                    # a creation error surfaces here
                if b.recs is not None:
                    self.ring.release(len(b.members))
                    b.recs = None
                inflight.append(b)
                progressed = True
            # completions (a batch whose output was too small is searched again)
            for b in list(inflight):
                if not b.done_ev.done():
                    continue
                if not b.collect(self.keep_results, t0):
                    continue
                self._finish(b, results, b.t_done)
                self.inflight_bytes -= b.est_bytes
                inflight.remove(b)
                free_streams.append(b.stream)
                progressed = True
            # admission
            while queue and free_streams:
                budget = self._budget()
                k = plan_batch(queue, budget, self.max_batch_points, self.max_batch_events, self.hits_per_point)
                if k == 0:
                    if self._reclaim_one():
                        continue
                    if inflight or creating:
                        break           # wait for memory held by batches in flight
                    k = 1               # nothing else holds memory: try the head alone
                k = same_kind_prefix(queue, k)
                if queue[0].ring_start >= 0:
                    k = contiguous_prefix(queue, k)
                members = [queue.popleft() for _ in range(k)]
                s = free_streams.pop(0)
                b = _Batch(nb, members, self, s)
                b.est_bytes = batch_bytes(b.P, self.hits_per_point)
                self.inflight_bytes += b.est_bytes
                creating.append((self.pool.submit(b.create, self), b))
                sizes.append(k)
                nb += 1
                peak = max(peak, len(inflight) + len(creating))
                progressed = True
            if not progressed:
                if pending and not queue and not inflight and not creating:
                    time.sleep(max(0.0, min(pending[0].arrival - (time.perf_counter() - t0), 0.01)))
                else:
                    time.sleep(self.poll_sleep_s)
        return ServeReport(results, nb, time.perf_counter() - t0, peak, self.evictions, sizes)

    def _finish(self, b: _Batch, results: Dict[int, EventResult], t_done: float):
        off, ids = b.res_off, b.res_ids
        at = b.offsets.astype(np.int64) if self.keep_results else np.arange(b.offsets.shape[0])
        b.res_off = b.res_ids = None
        for i, r in enumerate(b.members):
            a, c = int(at[i]), int(at[i + 1])
            res = EventResult(r.event, int(off[c] - off[a]), r.arrival, t_done, b.idx)
            if self.keep_results:
                res.offsets = (off[a:c + 1] - off[a]).astype(np.uint64)
                res.ids = ids[int(off[a]):int(off[c])].copy()
            results[r.event] = res
        b.free_search_buffers()
        if self.keep_resident:
            # the scratch stays until memory is needed (_reclaim_one)
            b.trimmed = False
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
        """Evict every resident tree, stop the workers, destroy the streams
        and hand the page-locked ingest ring back for the next server (a
        server holds about 256 MB of page-locked memory and three streams
        until then)."""
        while self._evict_one():
            pass
        self.pool.shutdown(wait=True)
        if self.ingest is not None:
            self.ingest.shutdown(wait=True)
        for s in self.streams:
            s.sync()
            s.close()
        self.streams = []
        if self.ring is not None:
            self.ring.close()
        self.ring = None

    def __enter__(self) -> "EventServer":
        return self

    def __exit__(self, *exc) -> None:
        self.close()


def poisson_requests(n_events: int, rate_hz: float, size_scale: float = 1.0, seed_offset: int = 0,
                     records: bool = False) -> List[Request]:
    """The serving workload: Poisson arrivals at `rate_hz`, sizes log-uniform
    in [1,000, 500,000] points (scaled by size_scale), mixed distribution.
    records=True: every event arrives as FKDPoint records (uint32 [n, 4])."""
    times, sizes = gen.serving_arrivals(n_events, rate_hz)
    out = []
    for e in range(n_events):
        n = max(1, int(sizes[e] * size_scale))
        p = gen.serving_event(n, e + seed_offset)
        if records:
            rec = np.empty((n, p.coords.shape[0] + 1), dtype=np.uint32)
            rec[:, :-1] = p.coords.T.view(np.uint32)
            rec[:, -1] = p.ids
            out.append(Request(e, None, None, float(times[e]), records=rec))
        else:
            out.append(Request(e, p.coords, p.ids, float(times[e])))
    return out
