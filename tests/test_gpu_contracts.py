"""Preconditions and call-order contracts of the C ABI on the GPU.

* Ids must be unique within a tree (an event): the reference comparator
  (FKDTree_cpu.h:246-252) is a strict order only then, so a tree of
  duplicated ids has no unique answer.  Creation rejects them with
  FKD_ERR_INVALID, whatever the id order, and accepts ids repeated across
  events of a batch.
* fkd_search_fill reuses the processing order fkd_search_count computed: a
  fill naming other boxes than the last count is refused (FKD_ERR_STATE)
  instead of pairing one box set's offsets with another's plan.
"""
import numpy as np
import pytest

import oracle
from fkdtree_amd import DeviceArray, FKDError, FKDEvents, FKDTree, Stream, gen
from tests.helpers import bits, sorted_sets

pytestmark = pytest.mark.gpu

FKD_ERR_INVALID, FKD_ERR_STATE = -1, -4


@pytest.mark.parametrize("order", ["increasing", "shuffled", "adjacent"])
def test_duplicate_ids_rejected(require_gpu, order):
    p = gen.mixed_points(30_000, 5, n_blobs=30)
    ids = np.arange(p.n, dtype=np.uint32) * np.uint32(3)
    if order == "increasing":
        ids[20_000] = ids[19_999]                 # equal neighbours
    elif order == "shuffled":
        perm = np.argsort(gen.u64(5, 9, p.n), kind="stable")
        ids = ids[perm]
        ids[17] = ids[29_000]                     # far apart, unordered ids
    else:
        ids = ids[::-1].copy()
        ids[100] = ids[101]
    print(f"[dup] order={order}", flush=True)
    with pytest.raises(FKDError) as e:
        FKDTree(coords=p.coords, ids=ids)
    assert e.value.code == FKD_ERR_INVALID and "duplicate" in str(e.value)


def test_unique_shuffled_ids_accepted_and_exact(require_gpu):
    p = gen.mixed_points(30_000, 5, n_blobs=30)
    perm = np.argsort(gen.u64(5, 9, p.n), kind="stable").astype(np.uint32)
    ids = perm * np.uint32(3) + np.uint32(7)
    t = FKDTree(coords=p.coords, ids=ids)
    t.build()
    c, i = t.tree_arrays()
    oc, oi = oracle.build(p.coords, ids)
    assert np.array_equal(bits(c), bits(oc)) and np.array_equal(i, oi)


def test_duplicate_ids_within_an_event_only(require_gpu):
    """Ids may repeat across events; a repeat inside one event is refused."""
    evs = [gen.mixed_event(n, 91, e) for e, n in enumerate([5000, 1, 0, 7000])]
    b = gen.concat_events(evs)
    ids = np.concatenate([np.arange(n, dtype=np.uint32)[::-1] for n in np.diff(b.offsets).astype(int)])
    ev = FKDEvents(b.points.coords, ids, b.offsets)   # same ids in every event: fine
    ev.build()
    c, i = ev.tree_arrays()
    oc, oi = oracle.build_events(b.points.coords, ids, b.offsets)
    assert np.array_equal(bits(c), bits(oc)) and np.array_equal(i, oi)
    ev.close()
    bad = ids.copy()
    a = int(b.offsets[3])
    bad[a + 10] = bad[a + 6000]
    with pytest.raises(FKDError) as e:
        FKDEvents(b.points.coords, bad, b.offsets)
    assert e.value.code == FKD_ERR_INVALID


def test_fill_refuses_other_boxes_than_the_count(require_gpu):
    p = gen.mixed_points(20_000, 8, n_blobs=20)
    a = gen.boxes_around(p.coords[:, :500], 0.02)
    b = gen.boxes_around(p.coords[:, 500:1000], 0.03)
    s = Stream()
    t = FKDTree(coords=p.coords, ids=p.ids, stream=s)
    t.build()
    da = (DeviceArray.from_host(a.mins, s), DeviceArray.from_host(a.maxs, s))
    db = (DeviceArray.from_host(b.mins, s), DeviceArray.from_host(b.maxs, s))
    off = DeviceArray(501, np.uint64)
    out = DeviceArray(1 << 20, np.uint32)
    t.search_device(500, da[0].ptr, da[1].ptr, off.ptr, None, 0, stream=s, fill=False)
    with pytest.raises(FKDError) as e:   # same count of boxes, other boxes
        t.search_device(500, db[0].ptr, db[1].ptr, off.ptr, out.ptr, 1 << 20, stream=s, count=False)
    assert e.value.code == FKD_ERR_STATE
    t.build_async()                      # a build between count and fill
    with pytest.raises(FKDError) as e:
        t.search_device(500, da[0].ptr, da[1].ptr, off.ptr, out.ptr, 1 << 20, stream=s, count=False)
    assert e.value.code == FKD_ERR_STATE
    # the proper order works and matches the oracle
    t.search_device(500, da[0].ptr, da[1].ptr, off.ptr, out.ptr, 1 << 20, stream=s)
    got_off = off.download(s)
    got = out.download(s, count=int(got_off[-1]))
    oc, oi = oracle.build(p.coords, p.ids)
    o_off, o_ids = oracle.search(oc, oi, a.mins, a.maxs)
    assert np.array_equal(got_off, o_off)
    assert np.array_equal(sorted_sets(got_off, got), sorted_sets(o_off, o_ids))
