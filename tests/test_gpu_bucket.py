"""The bucketed traversal (tuning key `bucket`): the walk stops three depths
above the last full depth and every subtree it reaches there (up to 15
nodes) is tested node by node.  The hits must stay the oracle's id sets
(FKDTree_cpu.h:123-169) whatever the tree's shape: every tree size around
the depths where the bucket root moves (n = 2^k - 2 .. 2^k + 1, trees of a
single bucket), dims 1-3, adversarial inputs, empty / degenerate / whole-space
/ NaN boxes, batches of events of uneven size, a too-small output capacity,
an unaligned output, and every park threshold (how many lanes wait before
the wave tests its buckets)."""
import numpy as np
import pytest

import oracle
from fkdtree_amd import DeviceArray, FKDEvents, FKDTree, gen
from tests.helpers import KINDS, adversarial_points, self_boxes, sorted_sets, special_boxes

pytestmark = pytest.mark.gpu

SIZES = (1, 2, 3, 6, 7, 8, 14, 15, 16, 17, 30, 31, 32, 62, 63, 64, 126, 127, 128, 129, 254, 255, 256,
         1000, 1022, 1023, 1024, 1025, 4094, 4095, 4096, 4097, 16383, 16385, 70_001)
# (bucket schedule, park threshold): 1 = lanes park at buckets, 2 = lanes queue them
SCHEDULES = ((1, 0), (1, 3), (1, 8), (2, 8))


def _boxes(p, hw):
    b = self_boxes(p, hw)
    s = special_boxes(p)
    return gen.Boxes(np.ascontiguousarray(np.concatenate([b.mins, s.mins], 1)),
                     np.ascontiguousarray(np.concatenate([b.maxs, s.maxs], 1)))


def _two_pass(t, dmi, dma, q, out_shift=0):
    """Count, then fill into a buffer whose first id sits `out_shift` ids past
    a 64-byte boundary."""
    off = DeviceArray(q + 1, np.uint64)
    t.search_device(q, dmi.ptr, dma.ptr, off.ptr, None, 0, fill=False)
    o = off.download()
    tot = int(o[-1])
    out = DeviceArray(tot + out_shift + 1, np.uint32)
    t.search_device(q, dmi.ptr, dma.ptr, off.ptr, out.ptr + 4 * out_shift, tot, count=False)
    return o, out.download(count=tot + out_shift)[out_shift:]


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_bucket_tree_sizes(require_gpu, dim):
    for n in SIZES:
        p = adversarial_points("mixed", n, dim=dim)
        boxes = _boxes(p, 0.05 if n < 5000 else 0.02)
        tc, ti = oracle.build(p.coords, p.ids)
        ooff, oids = oracle.search(tc, ti, boxes.mins, boxes.maxs)
        want = sorted_sets(ooff, oids)
        t = FKDTree(coords=p.coords, ids=p.ids)
        t.build()
        dmi = DeviceArray.from_host(boxes.mins)
        dma = DeviceArray.from_host(boxes.maxs)
        for mode, park in SCHEDULES:
            for kb in (0, 8):
                t.set_tuning(bucket=mode, count_bucket=mode, bucket_park=park, query_key_bits=min(kb, 32 // dim))
                assert t.get_tuning("bucket") == mode and t.get_tuning("bucket_park") == park
                o, ids = _two_pass(t, dmi, dma, boxes.q, out_shift=(n + park + mode) % 4)
                tag = f"dim {dim} n {n} schedule {mode} park {park} key bits {kb}"
                assert np.array_equal(o, ooff), f"{tag}: counts differ"
                assert np.array_equal(sorted_sets(o, ids), want), f"{tag}: ids differ"
        t.close()
    print(f"[bucket] dim {dim}: {len(SIZES)} tree sizes x {len(SCHEDULES)} schedules", flush=True)


@pytest.mark.parametrize("kind", KINDS)
def test_bucket_adversarial_inputs(require_gpu, kind):
    for dim, n in ((3, 60_000), (2, 9_000)):
        p = adversarial_points(kind, n, dim=dim)
        boxes = _boxes(p, 0.03)
        tc, ti = oracle.build(p.coords, p.ids)
        ooff, oids = oracle.search(tc, ti, boxes.mins, boxes.maxs)
        t = FKDTree(coords=p.coords, ids=p.ids)
        t.build()
        for mode in (1, 2):
            t.set_tuning(bucket=mode, count_bucket=mode, bucket_park=6)
            off, ids = t.search_in_the_box_multiple_csr(boxes.mins, boxes.maxs)
            assert np.array_equal(off, ooff), f"{kind} dim {dim} schedule {mode}: counts differ"
            assert np.array_equal(sorted_sets(off, ids), sorted_sets(ooff, oids)), f"{kind} dim {dim} schedule {mode}"
        t.close()


def test_bucket_events(require_gpu):
    """Events of every shape in one batch (empty, one point, single-bucket
    trees, sizes around bucket depth changes): each box searches its own
    event's tree."""
    sizes = [0, 1, 2, 7, 14, 15, 16, 31, 0, 127, 128, 1023, 1024, 1025, 3, 4097, 30_000, 16384, 5]
    evs = [gen.mixed_event(s, 91, e) for e, s in enumerate(sizes)]
    b = gen.concat_events(evs)
    boxes = gen.event_self_boxes(b, 0.02)
    qev = np.repeat(np.arange(b.n_events, dtype=np.uint32), np.diff(b.offsets))
    # boxes in reverse event order, and some naming an event that does not exist
    qev = qev[::-1].copy()
    mins, maxs = boxes.mins[:, ::-1].copy(), boxes.maxs[:, ::-1].copy()
    bad = np.zeros(boxes.q, bool)
    bad[::97] = True
    good = np.flatnonzero(~bad)
    tc, ti = oracle.build_events(b.points.coords, b.points.ids, b.offsets)
    ooff, oids = oracle.search_events(tc, ti, b.offsets, qev, mins, maxs)
    want_counts = np.diff(ooff).astype(np.int64)
    want_counts[bad] = 0
    want = sorted_sets(*_take(ooff, oids, good))
    qev_bad = qev.copy()
    qev_bad[bad] = b.n_events + 3   # a box naming no event matches nothing
    t = FKDEvents(b.points.coords, b.points.ids, b.offsets)
    t.build()
    for mode, park in ((1, 0), (1, 5), (1, 8), (2, 8)):
        for kb in (0, 10):
            t.set_tuning(bucket=mode, count_bucket=mode, bucket_park=park, query_key_bits=kb)
            off, ids = t.search(mins, maxs, qev_bad)
            assert np.array_equal(np.diff(off).astype(np.int64), want_counts), f"park {park} key bits {kb}: counts"
            assert np.array_equal(sorted_sets(*_take(off, ids, good)), want), f"park {park} key bits {kb}: ids"
    t.close()


def test_bucket_one_call_capacity(require_gpu):
    """The one-call search with too small an output: the status reports the
    true total, the kept ids are a prefix of the complete output, and a
    search into a large enough buffer returns the oracle's sets."""
    p = adversarial_points("mixed", 200_000)
    boxes = self_boxes(p, 0.015)
    tc, ti = oracle.build(p.coords, p.ids)
    ooff, oids = oracle.search(tc, ti, boxes.mins, boxes.maxs)
    t = FKDTree(coords=p.coords, ids=p.ids)
    t.build()
    t.set_tuning(bucket=2, count_bucket=2)
    q = boxes.q
    dmi, dma = DeviceArray.from_host(boxes.mins), DeviceArray.from_host(boxes.maxs)
    off = DeviceArray(q + 1, np.uint64)
    total = int(ooff[-1])
    small = total // 3 + 1
    out = DeviceArray(small, np.uint32)
    t.search_single_device(q, dmi.ptr, dma.ptr, off.ptr, out.ptr, small)
    tot, ok = t.search_status()
    assert tot == total and not ok
    o = off.download()
    assert np.array_equal(o, ooff)
    full = DeviceArray(total, np.uint32)
    t.search_single_device(q, dmi.ptr, dma.ptr, off.ptr, full.ptr, total)
    tot, ok = t.search_status()
    assert tot == total and ok
    ids = full.download()
    assert np.array_equal(out.download(), ids[:small])
    assert np.array_equal(sorted_sets(o, ids), sorted_sets(ooff, oids))
    t.close()


def test_bucket_matches_walk_on_mixed_1m(require_gpu):
    """The full 1,048,576-point workload tree (one point on its last depth):
    bucketed and walking traversals return the oracle's counts for every box
    and its id sets for every 16th box."""
    pts, boxes = gen.workload_mixed_1m()
    tc, ti = oracle.build(pts.coords, pts.ids)
    ooff, oids = oracle.search(tc, ti, boxes.mins, boxes.maxs)
    t = FKDTree(coords=pts.coords, ids=pts.ids)
    t.build()
    sel = np.arange(0, boxes.q, 16)
    want = oracle.sort_segments(*_take(ooff, oids, sel))
    for bucket, park in ((2, 8), (1, 4), (0, 8)):
        t.set_tuning(bucket=bucket, count_bucket=bucket, bucket_park=park)
        off, ids = t.search_in_the_box_multiple_csr(boxes.mins, boxes.maxs)
        assert np.array_equal(off, ooff), f"bucket {bucket} park {park}: counts differ"
        assert np.array_equal(oracle.sort_segments(*_take(off, ids, sel)), want), f"bucket {bucket} park {park}"
    t.close()


def _take(off, ids, sel):
    """CSR of the boxes `sel` only."""
    cnt = (off[sel + 1] - off[sel]).astype(np.int64)
    o = np.zeros(len(sel) + 1, np.uint64)
    o[1:] = np.cumsum(cnt)
    parts = [ids[int(off[i]):int(off[i + 1])] for i in sel]
    return o, (np.concatenate(parts) if parts else np.zeros(0, np.uint32))
