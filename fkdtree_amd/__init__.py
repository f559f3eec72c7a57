"""MI355X-native FKDTree: k-d tree build and box search in gfx950 HIP kernels.

Python mirror of the reference interface ``FKDTree<float, dim>``
(``FKDTree.h:6-52``) over the C ABI of ``include/fkdtree/fkd_capi.h``:

* ``FKDTree(points)``            - constructor (copies the points to the GPU)
* ``build()``                    - ``FKDTree_cpu.h:221-277``
* ``getDimensionVector(d)``      - heap-ordered coordinates (``:209-211``)
* ``getIdVector()``              - heap-ordered ids (``:212-213``)
* ``search_in_the_box(lo, hi)``  - ids in the closed box (``:123-169``)
* ``search_in_the_box_multiple`` - per-box counts, the OpenCL backend's result
                                   (``FKDTree_opencl.h:123-164``)
* ``search_in_the_box_multiple_csr`` - per-box ids as (offsets, ids)

``FKDEvents`` holds many independent trees (events) built and searched
together.  Device buffers (``DeviceArray``), streams and timers are thin
wrappers of the same library; nothing here falls back to the CPU.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Sequence, Tuple

import numpy as np

from . import gen  # noqa: F401
from ._lib import FKDError, check, lib

__all__ = ["FKDTree", "FKDEvents", "DeviceArray", "Stream", "Timer", "Event", "Graph", "FKDError",
           "device_info", "release_cached"]


def device_info() -> dict:
    name = C.create_string_buffer(256)
    cus = C.c_int()
    fr = C.c_size_t()
    tot = C.c_size_t()
    check(lib().fkd_device_info(name, 256, C.byref(cus), C.byref(fr), C.byref(tot)))
    res, used = C.c_uint64(), C.c_uint64()
    check(lib().fkd_pool_info(C.byref(res), C.byref(used)))
    cached = max(res.value - used.value, 0)
    # free_bytes: what the device reports free plus what the library's
    # device-memory cache holds unused (reusable by its allocations)
    return {"name": name.value.decode(), "compute_units": cus.value, "free_bytes": fr.value + cached,
            "device_free_bytes": fr.value, "pool_cached_bytes": cached, "total_bytes": tot.value}


def release_cached() -> None:
    """Give the device memory the library caches (freed buffers kept for
    reuse) back to the driver, where their last use has completed."""
    check(lib().fkd_release_cached())


class Stream:
    def __init__(self):
        h = C.c_void_p()
        check(lib().fkd_stream_create(C.byref(h)))
        self.handle = h.value

    def sync(self):
        check(lib().fkd_stream_sync(self.handle))

    def close(self):
        if self.handle:
            lib().fkd_stream_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Timer:
    """HIP events on a stream: elapsed milliseconds between two records."""

    def __init__(self):
        a, b = C.c_void_p(), C.c_void_p()
        check(lib().fkd_event_create(C.byref(a)))
        check(lib().fkd_event_create(C.byref(b)))
        self.a, self.b = a.value, b.value

    def start(self, stream=None):
        check(lib().fkd_event_record(self.a, _sh(stream)))

    def stop(self, stream=None):
        check(lib().fkd_event_record(self.b, _sh(stream)))

    def ms(self) -> float:
        v = C.c_float()
        check(lib().fkd_event_elapsed_ms(self.a, self.b, C.byref(v)))
        return float(v.value)

    def __del__(self):
        try:
            lib().fkd_event_destroy(self.a)
            lib().fkd_event_destroy(self.b)
        except Exception:
            pass


class Event:
    """A HIP event: record on a stream, poll without blocking."""

    def __init__(self):
        e = C.c_void_p()
        check(lib().fkd_event_create(C.byref(e)))
        self.h = e.value

    def record(self, stream=None):
        check(lib().fkd_event_record(self.h, _sh(stream)))

    def done(self) -> bool:
        r = lib().fkd_event_query(self.h)
        if r < 0:
            check(r)
        return r == 1

    def __del__(self):
        try:
            lib().fkd_event_destroy(self.h)
        except Exception:
            pass


def _sh(stream) -> Optional[int]:
    if stream is None:
        return None
    return stream.handle if isinstance(stream, Stream) else stream


class DeviceArray:
    """A device allocation holding a numpy-typed array."""

    def __init__(self, shape, dtype, ptr: Optional[int] = None, stream=None):
        """With `stream`, the memory is allocated and later freed in stream
        order on it (no device-wide synchronisation); the array keeps the
        stream alive."""
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        self._own = ptr is None
        self._stream = stream
        if ptr is None:
            p = C.c_void_p()
            if stream is None:
                check(lib().fkd_malloc(C.byref(p), max(self.nbytes, 1)))
            else:
                check(lib().fkd_malloc_async(C.byref(p), max(self.nbytes, 1), _sh(stream)))
            ptr = p.value
        self.ptr = ptr

    @classmethod
    def from_host(cls, a: np.ndarray, stream=None) -> "DeviceArray":
        """A copy of host array `a`; with `stream`, allocated (and later
        freed) in stream order on it."""
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype, stream=stream)
        d.upload(a, stream)
        return d

    def upload(self, a: np.ndarray, stream=None):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes, (a.nbytes, self.nbytes)
        check(lib().fkd_memcpy_htod(self.ptr, a.ctypes.data, a.nbytes, _sh(stream)))
        check(lib().fkd_stream_sync(_sh(stream)))

    def upload_at(self, a: np.ndarray, byte_offset: int, stream=None) -> None:
        """Enqueue a copy of host array `a` to byte_offset of this array; `a`
        must stay alive until the stream has passed the copy."""
        a = np.ascontiguousarray(a)
        assert byte_offset + a.nbytes <= self.nbytes, (byte_offset, a.nbytes, self.nbytes)
        if a.nbytes:
            check(lib().fkd_memcpy_htod(self.ptr + byte_offset, a.ctypes.data, a.nbytes, _sh(stream)))

    def download(self, stream=None, count: Optional[int] = None) -> np.ndarray:
        if count is None:
            out = np.empty(self.shape, dtype=self.dtype)
        else:
            out = np.empty(int(count), dtype=self.dtype)
        check(lib().fkd_memcpy_dtoh(out.ctypes.data, self.ptr, out.nbytes, _sh(stream)))
        return out

    def free(self):
        if self._own and self.ptr:
            if self._stream is None:
                lib().fkd_free(self.ptr)
            else:
                lib().fkd_free_async(self.ptr, _sh(self._stream))
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class _PinnedBlock:
    """Owner of one page-locked host allocation (fkd_host_alloc)."""

    def __init__(self, nbytes: int):
        p = C.c_void_p()
        check(lib().fkd_host_alloc(C.byref(p), max(int(nbytes), 1)))
        self.ptr = p.value
        self.nbytes = int(nbytes)

    def __del__(self):
        try:
            if self.ptr:
                lib().fkd_host_free(self.ptr)
        except Exception:
            pass
        self.ptr = None


def pinned_empty(shape, dtype) -> np.ndarray:
    """A numpy array in page-locked host memory; the memory lives as long as
    the array (or any view of it)."""
    dt = np.dtype(dtype)
    shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
    nbytes = int(np.prod(shape, dtype=np.int64)) * dt.itemsize
    blk = _PinnedBlock(nbytes)
    raw = (C.c_uint8 * max(nbytes, 1)).from_address(blk.ptr)
    raw._fkd_block = blk  # the ctypes buffer keeps the allocation alive
    return np.frombuffer(raw, dtype=np.uint8, count=nbytes).view(dt).reshape(shape)


def _as_soa(coords, dim=None) -> np.ndarray:
    c = np.ascontiguousarray(coords, dtype=np.float32)
    if c.ndim != 2:
        raise ValueError("coords must be 2-D [dim, n]")
    return c


def _points_aos(points=None, coords=None, ids=None) -> Tuple[int, int, np.ndarray]:
    """(dim, n, AoS uint32 records) from FKDPoint-like records or SoA coords+ids."""
    if points is not None:
        rec = np.ascontiguousarray(points)
        if rec.dtype != np.uint32:
            raise ValueError("points must be uint32 records [n, dim+1] (coordinate bits, id)")
        n, w = rec.shape
        return w - 1, n, rec
    c = _as_soa(coords)
    dim, n = c.shape
    i = np.arange(n, dtype=np.uint32) if ids is None else np.ascontiguousarray(ids, dtype=np.uint32)
    rec = np.empty((n, dim + 1), dtype=np.uint32)
    rec[:, :dim] = c.T.view(np.uint32)
    rec[:, dim] = i
    return dim, n, rec


class _TreeBase:
    _h = None

    def _init_handle(self, h: C.c_void_p, stream):
        self._h = h.value
        self.stream = stream
        d = C.c_int()
        e = C.c_uint32()
        n = C.c_uint64()
        check(lib().fkd_tree_info(self._h, C.byref(d), C.byref(e), C.byref(n)))
        self.dim, self.n_events, self.n = d.value, e.value, n.value
        self._mirror = None
        from . import tune

        if os.path.exists(tune.TABLE):
            tune.apply(self)

    _server = None

    def _drop_server(self) -> None:
        if self._server is not None:
            self._server.close()
            self._server = None

    def build(self) -> None:
        self._drop_server()
        check(lib().fkd_build(self._h, _sh(self.stream)))
        check(lib().fkd_stream_sync(_sh(self.stream)))
        self._mirror = None

    def build_async(self) -> None:
        """Enqueue the build on the tree's stream and return."""
        self._drop_server()
        check(lib().fkd_build(self._h, _sh(self.stream)))
        self._mirror = None

    def tree_arrays(self) -> Tuple[np.ndarray, np.ndarray]:
        """Heap-ordered SoA coordinates [dim, n] and ids [n] (host copies)."""
        if self._mirror is None:
            c = np.empty((self.dim, self.n), dtype=np.float32)
            i = np.empty(self.n, dtype=np.uint32)
            check(lib().fkd_get_tree(self._h, c.ctypes.data, i.ctypes.data, _sh(self.stream)))
            self._mirror = (c, i)
        return self._mirror

    def boxes_around(self, half_width: float, d_mins: int, d_maxs: int, d_qev: Optional[int] = None,
                     stream=None) -> None:
        """Enqueue fkd_boxes_around: boxes of half-width `half_width` around
        every point (input order, SoA [dim][n]) and each box's event."""
        check(lib().fkd_boxes_around(self._h, float(half_width), d_mins, d_maxs, d_qev,
                                     _sh(stream if stream is not None else self.stream)))

    def serialize(self) -> bytes:
        """The wire format of the built tree (fkd_serialize): header, event
        offsets, heap-ordered coordinates and ids."""
        size = C.c_uint64()
        check(lib().fkd_serialize(self._h, None, 0, C.byref(size), _sh(self.stream)))
        buf = np.empty(size.value, dtype=np.uint8)
        check(lib().fkd_serialize(self._h, buf.ctypes.data, size.value, C.byref(size), _sh(self.stream)))
        return buf.tobytes()

    def save(self, path: str) -> None:
        """Write the wire format to `path` (spill a resident tree)."""
        data = self.serialize()
        tmp = path + ".tmp"
        with open(tmp, "wb") as f:
            f.write(data)
        os.replace(tmp, path)

    def device_tree(self) -> Tuple[int, int]:
        c, i = C.c_void_p(), C.c_void_p()
        check(lib().fkd_tree_device(self._h, C.byref(c), C.byref(i)))
        return c.value, i.value

    def _search_host(self, mins, maxs, qev=None) -> Tuple[np.ndarray, np.ndarray]:
        mins = _as_soa(mins)
        maxs = _as_soa(maxs)
        if mins.shape != maxs.shape or mins.shape[0] != self.dim:
            raise ValueError("mins/maxs must be [dim, nq]")
        nq = mins.shape[1]
        off = np.empty(nq + 1, dtype=np.uint64)
        qp = None
        if qev is not None:
            qev = np.ascontiguousarray(qev, dtype=np.uint32)
            qp = qev.ctypes.data
        check(lib().fkd_search_host(self._h, nq, mins.ctypes.data, maxs.ctypes.data, qp, off.ctypes.data,
                                    _sh(self.stream)))
        ids = np.empty(int(off[-1]), dtype=np.uint32)
        check(lib().fkd_fetch_results(self._h, ids.ctypes.data, _sh(self.stream)))
        return off, ids

    def search_device(self, nq: int, d_mins: int, d_maxs: int, d_offsets: int, d_out: Optional[int],
                      capacity: int, d_qev: Optional[int] = None, stream=None, count: bool = True,
                      fill: bool = True) -> None:
        """Enqueue count+scan and/or fill on device buffers (raw pointers)."""
        s = _sh(stream if stream is not None else self.stream)
        if count:
            check(lib().fkd_search_count(self._h, nq, d_mins, d_maxs, d_qev, d_offsets, s))
        if fill:
            check(lib().fkd_search_fill(self._h, nq, d_mins, d_maxs, d_qev, d_offsets, d_out, capacity, s))

    def search_single_device(self, nq: int, d_mins: int, d_maxs: int, d_offsets: int, d_out: Optional[int],
                             capacity: int, d_qev: Optional[int] = None, stream=None) -> None:
        """Enqueue the whole search (fkd_search: plan, count, scan, fill) on
        device buffers; `search_status` tells whether `capacity` sufficed."""
        s = _sh(stream if stream is not None else self.stream)
        check(lib().fkd_search(self._h, nq, d_mins, d_maxs, d_qev, d_offsets, d_out, capacity, s))

    def search_status(self, stream=None) -> Tuple[int, bool]:
        """(total hits, complete) of the last one-call search (synchronises)."""
        tot = C.c_uint64()
        ok = C.c_int()
        s = _sh(stream if stream is not None else self.stream)
        check(lib().fkd_search_status(self._h, s, C.byref(tot), C.byref(ok)))
        return tot.value, bool(ok.value)

    def set_tuning(self, **params) -> None:
        """Launch parameters (fkd_set_tuning): query_key_bits, fill_order, fill_window, count_window,
        key_volume_bits."""
        for k, v in params.items():
            check(lib().fkd_set_tuning(self._h, k.encode(), int(v)))

    def get_tuning(self, key: str) -> int:
        v = C.c_uint32()
        check(lib().fkd_get_tuning(self._h, key.encode(), C.byref(v)))
        return v.value

    def trim(self) -> None:
        """Release build / search scratch; the tree stays resident."""
        check(lib().fkd_trim(self._h, _sh(self.stream)))

    def device_bytes(self) -> int:
        b = C.c_uint64()
        check(lib().fkd_device_bytes(self._h, C.byref(b)))
        return b.value

    def close(self):
        """Free the handle.  With a stream, its memory returns to the
        allocator in stream order (the handle's work must be on that stream
        or complete); without one, after a device-wide synchronisation."""
        self._drop_server()
        if self._h:
            if self.stream is None:
                lib().fkd_destroy(self._h)
            else:
                lib().fkd_destroy_async(self._h, _sh(self.stream))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class FKDTree(_TreeBase):
    """One k-d tree, ``FKDTree<float, dim>`` on the GPU."""

    def __init__(self, points=None, *, coords=None, ids=None, stream=None):
        dim, n, rec = _points_aos(points, coords, ids)
        h = C.c_void_p()
        check(lib().fkd_create(dim, n, rec.ctypes.data if n else None, _sh(stream), C.byref(h)))
        self._init_handle(h, stream)

    @classmethod
    def from_device(cls, dim: int, n: int, d_coords: int, d_ids: int, stream=None) -> "FKDTree":
        self = cls.__new__(cls)
        off = np.array([0, n], dtype=np.uint64)
        h = C.c_void_p()
        check(lib().fkd_create_device(dim, 1, off.ctypes.data, d_coords, d_ids, _sh(stream), C.byref(h)))
        self._init_handle(h, stream)
        return self

    # Reference-named accessors (FKDTree.h:20-22)
    def getDimensionVector(self, dimension: int) -> np.ndarray:
        c, _ = self.tree_arrays()
        return c[dimension] if 0 <= dimension < self.dim else c[0]

    def getIdVector(self) -> np.ndarray:
        return self.tree_arrays()[1]

    def box_server(self, slots: Optional[int] = None, slot_ids: int = 4096, idle_us: int = 20000) -> "BoxServer":
        """A resident per-box search kernel on this built tree
        (fkd_box_server_create); building again or closing the tree stops it.
        slots defaults to the host's hardware threads, 16 to 64 (each holds a
        compute unit while the kernel is live)."""
        if slots is None:
            slots = min(64, max(16, os.cpu_count() or 16))
        return BoxServer(self, slots, slot_ids, idle_us)

    def search_in_the_box(self, min_point: Sequence[float], max_point: Sequence[float]) -> np.ndarray:
        """Ids of the points in the closed box (FKDTree_cpu.h:123-169), through
        the tree's resident per-box search kernel (started at the first call
        after a build); a walk too wide for it or a result larger than its
        slot goes through the batched search."""
        if self._server is None:
            self._server = self.box_server()
        ids = self._server.search(min_point, max_point)
        if ids is not None:
            return ids
        lo = np.asarray(min_point, dtype=np.float32).reshape(self.dim, 1)
        hi = np.asarray(max_point, dtype=np.float32).reshape(self.dim, 1)
        _, ids = self._search_host(lo, hi)
        return ids

    # The reference's other search entry points (FKDTree_cpu.h:25-121,
    # :171-207): the same id set, so the same search.
    search_in_the_box_BFS = search_in_the_box
    search_in_the_box_recursive = search_in_the_box
    search_in_the_box_branchless = search_in_the_box

    def search_in_the_box_multiple(self, mins, maxs) -> np.ndarray:
        off, _ = self._search_host(mins, maxs)
        return np.diff(off).astype(np.uint32)

    def search_in_the_box_multiple_csr(self, mins, maxs) -> Tuple[np.ndarray, np.ndarray]:
        return self._search_host(mins, maxs)


class BoxServer:
    """Resident per-box search kernel of one built tree (fkd_box_server_*):
    search() may be called from several Python threads at once (ctypes
    releases the GIL), each call taking one request slot."""

    OVERFLOW = 0xFFFFFFFF

    def __init__(self, tree: "FKDTree", slots: int = 16, slot_ids: int = 4096, idle_us: int = 20000):
        h = C.c_void_p()
        check(lib().fkd_box_server_create(tree._h, slots, slot_ids, idle_us, C.byref(h)))
        self._h = h.value
        self.dim = tree.dim
        self.slot_ids = slot_ids
        self._tree = tree  # the tree outlives its server

    def search_raw(self, min_point, max_point, cap: Optional[int] = None) -> Tuple[int, np.ndarray]:
        """(count, ids): count is the box's hits or OVERFLOW; ids holds the
        first min(count, cap, slot_ids) of them."""
        box = np.concatenate([np.asarray(min_point, np.float32).ravel(),
                              np.asarray(max_point, np.float32).ravel()])
        assert box.size == 2 * self.dim
        cap = self.slot_ids if cap is None else cap
        out = np.empty(max(cap, 1), np.uint32)
        cnt = C.c_uint32()
        check(lib().fkd_box_server_search(self._h, box.ctypes.data, out.ctypes.data, cap, C.byref(cnt)))
        c = cnt.value
        return c, out[:min(c, cap, self.slot_ids)] if c != self.OVERFLOW else out[:0]

    def search(self, min_point, max_point) -> Optional[np.ndarray]:
        """The box's ids, or None when the walk was too wide for one wave or
        the result larger than a slot."""
        c, ids = self.search_raw(min_point, max_point)
        if c == self.OVERFLOW or c > self.slot_ids:
            return None
        return ids

    def stats(self) -> dict:
        la, rq, run = C.c_uint64(), C.c_uint64(), C.c_int()
        check(lib().fkd_box_server_stats(self._h, C.byref(la), C.byref(rq), C.byref(run)))
        rt, wk, rp = C.c_double(), C.c_double(), C.c_uint64()
        check(lib().fkd_box_server_timing(self._h, C.byref(rt), C.byref(wk), C.byref(rp)))
        return {"launches": la.value, "requests": rq.value, "running": bool(run.value),
                "round_trip_us": rt.value, "walk_us": wk.value, "reposts": rp.value}

    def close(self) -> None:
        if self._h:
            lib().fkd_box_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class FKDEvents(_TreeBase):
    """Independent trees of many events, built and searched in shared launches."""

    def __init__(self, coords, ids, offsets, stream=None):
        dim, n, rec = _points_aos(None, coords, ids)
        off = np.ascontiguousarray(offsets, dtype=np.uint64)
        h = C.c_void_p()
        check(lib().fkd_create_events(dim, off.shape[0] - 1, off.ctypes.data, rec.ctypes.data if n else None,
                                      _sh(stream), C.byref(h)))
        self.offsets = off
        self._init_handle(h, stream)

    @classmethod
    def from_records(cls, records: np.ndarray, offsets, stream=None) -> "FKDEvents":
        """From FKDPoint records (uint32 [n, dim+1]: coordinate bits, id),
        events end to end: one copy to the device (asynchronous at full rate
        from page-locked memory, see pinned_empty), transposed there."""
        self = cls.__new__(cls)
        rec = records if records.flags.c_contiguous else np.ascontiguousarray(records)
        if rec.dtype != np.uint32 or rec.ndim != 2:
            raise ValueError("records must be uint32 [n, dim+1]")
        off = np.ascontiguousarray(offsets, dtype=np.uint64)
        h = C.c_void_p()
        check(lib().fkd_create_events(rec.shape[1] - 1, off.shape[0] - 1, off.ctypes.data,
                                      rec.ctypes.data if rec.shape[0] else None, _sh(stream), C.byref(h)))
        self.offsets = off
        self._init_handle(h, stream)
        return self

    @classmethod
    def from_device(cls, dim: int, offsets, d_coords: int, d_ids: int, stream=None) -> "FKDEvents":
        self = cls.__new__(cls)
        off = np.ascontiguousarray(offsets, dtype=np.uint64)
        h = C.c_void_p()
        check(lib().fkd_create_device(dim, off.shape[0] - 1, off.ctypes.data, d_coords, d_ids, _sh(stream),
                                      C.byref(h)))
        self.offsets = off
        self._init_handle(h, stream)
        return self

    def search(self, mins, maxs, query_events) -> Tuple[np.ndarray, np.ndarray]:
        """Query q searches the tree of event query_events[q]: (offsets, ids)."""
        return self._search_host(mins, maxs, query_events)


def deserialize(data: bytes, verify: bool = False, stream=None):
    """A built tree from the wire format (fkd_deserialize): an FKDTree for one
    event, else an FKDEvents.  verify=True rebuilds from the points and checks
    that every array comes out identical."""
    buf = np.frombuffer(data, dtype=np.uint8)
    h = C.c_void_p()
    check(lib().fkd_deserialize(buf.ctypes.data, buf.size, 1 if verify else 0, _sh(stream), C.byref(h)))
    e = C.c_uint32()
    check(lib().fkd_tree_info(h.value, None, C.byref(e), None))
    cls = FKDTree if e.value == 1 else FKDEvents
    self = cls.__new__(cls)
    self._init_handle(h, stream)
    if cls is FKDEvents:
        hdr = np.frombuffer(data, dtype=np.uint8, count=24)
        E = int(hdr[12:16].view(np.uint32)[0])
        self.offsets = np.frombuffer(data, dtype=np.uint64, count=E + 1, offset=24).copy()
    return self


def load(path: str, verify: bool = False, stream=None):
    """A built tree from a file written by `save`."""
    with open(path, "rb") as f:
        return deserialize(f.read(), verify=verify, stream=stream)


class Graph:
    """Build + one-call search of fixed device boxes captured as one HIP
    graph (fkd_graph_create); launch() replays the whole pipeline."""

    def __init__(self, tree: _TreeBase, nq: int, d_mins: int, d_maxs: int, d_offsets: int, d_out: int,
                 capacity: int, stream: "Stream", d_qev: Optional[int] = None):
        h = C.c_void_p()
        check(lib().fkd_graph_create(tree._h, nq, d_mins, d_maxs, d_qev, d_offsets, d_out, capacity,
                                     _sh(stream), C.byref(h)))
        self.h = h.value
        self.tree = tree
        self.stream = stream

    def launch(self, stream=None) -> None:
        check(lib().fkd_graph_launch(self.h, _sh(stream if stream is not None else self.stream)))

    def close(self):
        if self.h:
            lib().fkd_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
