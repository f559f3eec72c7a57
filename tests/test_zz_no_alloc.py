"""fkd_search_noalloc: the search without any allocation (its scratch is a
fixed device array of the library).  Called through the C ABI, it returns
the same offsets and the same id sets as fkd_search and the oracle
(FKDTree_cpu.h:123-169), on calls in a row (the second and third with fewer
boxes, so stale scratch from the first would show), for one tree and for a
batch of events; more boxes than its limit are refused."""
import ctypes as C

import numpy as np
import pytest

import oracle
from fkdtree_amd import DeviceArray, FKDEvents, FKDTree, Stream, gen
from fkdtree_amd._lib import lib
from tests.helpers import sorted_sets

pytestmark = pytest.mark.gpu

MAX_BOXES = 1 << 22


def _dev(a, s):
    return DeviceArray.from_host(np.ascontiguousarray(a), s)


def _search(fn, tree, boxes, s, cap, qev=None):
    """(offsets, ids, status) of one call of fkd_search or fkd_search_noalloc."""
    nq = boxes.q
    d_mi, d_ma = _dev(boxes.mins, s), _dev(boxes.maxs, s)
    d_q = _dev(qev, s) if qev is not None else None
    d_off = DeviceArray(nq + 1, np.uint64)
    d_off.upload(np.full(nq + 1, 0xABABABABABABABAB, np.uint64), s)  # no stale zeros
    cap = int(cap)
    d_out = DeviceArray(max(cap, 1), np.uint32)
    rc = fn(tree._h, nq, d_mi.ptr, d_ma.ptr, d_q.ptr if d_q else None, d_off.ptr, d_out.ptr, cap, s.handle)
    assert rc == 0, lib().fkd_last_error()
    tot, ok = tree.search_status(s)
    off = d_off.download(s)
    ids = d_out.download(s, count=min(int(off[-1]), cap))
    return off, ids, (tot, ok)


def test_noalloc_matches_search_and_oracle(require_gpu):
    L = lib()
    p = gen.mixed_points(200_000, 31, n_blobs=200)
    s = Stream()
    t = FKDTree(coords=p.coords, ids=p.ids, stream=s)
    t.build()
    oc, oi = oracle.build(p.coords, p.ids)
    idx = (gen.u64(31, 40, 150_000) % np.uint64(p.n)).astype(np.int64)
    big = gen.boxes_around(p.coords[:, idx], 0.02)
    small = gen.boxes_around(p.coords[:, :30_000] + np.float32(0.001), 0.01)
    # every traversal schedule of the fill pass (walk, parked or queued buckets)
    for bucket, (name, boxes) in ((0, ("150000 boxes", big)), (0, ("30000 boxes", small)),
                                  (1, ("150000 boxes", big)), (2, ("30000 boxes", small))):
        print(f"[noalloc] one tree, {name}, bucket {bucket}", flush=True)
        t.set_tuning(bucket=bucket)
        o_off, o_ids = oracle.search(oc, oi, boxes.mins, boxes.maxs)
        total = int(o_off[-1])
        r_off, r_ids, _ = _search(L.fkd_search, t, boxes, s, cap=total)
        n_off, n_ids, st = _search(L.fkd_search_noalloc, t, boxes, s, cap=total)
        assert st == (int(o_off[-1]), True)
        assert np.array_equal(n_off, r_off) and np.array_equal(n_off, o_off)
        assert np.array_equal(sorted_sets(n_off, n_ids), sorted_sets(r_off, r_ids))
        assert np.array_equal(sorted_sets(n_off, n_ids), sorted_sets(o_off, o_ids))
    # too small an output: the total is still reported and nothing is written past capacity
    n_off, n_ids, st = _search(L.fkd_search_noalloc, t, small, s, cap=1000)
    assert st == (int(n_off[-1]), False) and n_ids.shape[0] == 1000
    t.close()


def test_noalloc_events_and_limits(require_gpu):
    L = lib()
    b = gen.workload_events_2048(n_events=40, scale=0.1)
    boxes = gen.event_self_boxes(b, 0.01)
    qev = np.repeat(np.arange(b.n_events, dtype=np.uint32), np.diff(b.offsets))
    print(f"[noalloc] events={b.n_events} points={b.points.n}", flush=True)
    s = Stream()
    ev = FKDEvents(b.points.coords, b.points.ids, b.offsets, stream=s)
    ev.build()
    oc, oi = oracle.build_events(b.points.coords, b.points.ids, b.offsets)
    o_off, o_ids = oracle.search_events(oc, oi, b.offsets, qev, boxes.mins, boxes.maxs)
    n_off, n_ids, st = _search(L.fkd_search_noalloc, ev, boxes, s, qev=qev, cap=int(o_off[-1]))
    assert st == (int(o_off[-1]), True)
    assert np.array_equal(n_off, o_off)
    assert np.array_equal(sorted_sets(n_off, n_ids), sorted_sets(o_off, o_ids))
    # zero boxes: offsets[0] = 0 and nothing else (a batch of trees names the
    # query events even then, as fkd_search requires)
    z_off = DeviceArray(1, np.uint64)
    z_off.upload(np.full(1, 7, np.uint64), s)
    assert L.fkd_search_noalloc(ev._h, 0, None, None, z_off.ptr, z_off.ptr, None, 0, s.handle) == 0, \
        L.fkd_last_error()
    assert z_off.download(s)[0] == 0
    # beyond the limit: refused before any device work
    rc = L.fkd_search_noalloc(ev._h, MAX_BOXES + 1, z_off.ptr, z_off.ptr, z_off.ptr, z_off.ptr, z_off.ptr, 1,
                              s.handle)
    assert rc == -1 and b"at most 4194304" in L.fkd_last_error()
    ev.close()
