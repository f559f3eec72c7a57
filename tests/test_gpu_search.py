"""Box search parity on the GPU: every query's sorted id set bit-exact against the
oracle's level-BFS search (FKDTree_cpu.h:123-169) and a brute-force scan
(FKDTree.h:54-86), through the C ABI."""
import numpy as np
import pytest

import oracle
from fkdtree_amd import FKDEvents, FKDTree, gen
from tests.helpers import KINDS, adversarial_points, self_boxes, sorted_sets, special_boxes

pytestmark = pytest.mark.gpu


def check_search(p: gen.PointSet, boxes: gen.Boxes, label: str, brute: bool = True):
    print(f"[search] {label} n={p.n} q={boxes.q}", flush=True)
    t = FKDTree(coords=p.coords, ids=p.ids)
    t.build()
    goff, gids = t.search_in_the_box_multiple_csr(boxes.mins, boxes.maxs)
    tc, ti = oracle.build(p.coords, p.ids)
    ooff, oids = oracle.search(tc, ti, boxes.mins, boxes.maxs)
    assert np.array_equal(goff, ooff), f"{label}: counts differ"
    assert np.array_equal(sorted_sets(goff, gids), sorted_sets(ooff, oids)), f"{label}: ids differ"
    if brute:
        boff, bids = oracle.brute(p.coords, p.ids, boxes.mins, boxes.maxs)
        assert np.array_equal(goff, boff)
        assert np.array_equal(sorted_sets(goff, gids), sorted_sets(boff, bids))
    counts = t.search_in_the_box_multiple(boxes.mins, boxes.maxs)
    assert np.array_equal(counts, np.diff(ooff).astype(np.uint32))
    t.close()
    return goff, gids


@pytest.mark.parametrize("kind", KINDS)
def test_small_searches(require_gpu, kind):
    for n in (1, 2, 3, 4, 7, 8, 100, 1000, 4097):
        p = adversarial_points(kind, n)
        check_search(p, self_boxes(p, 0.05), f"{kind} hw0.05")
        check_search(p, special_boxes(p), f"{kind} special")


@pytest.mark.parametrize("dim", [1, 2, 4])
def test_other_dims(require_gpu, dim):
    for n in (1, 5, 1000, 20000):
        p = adversarial_points("mixed", n, dim=dim)
        check_search(p, self_boxes(p, 0.03), f"dim{dim}")


def test_single_box_api(require_gpu):
    p = adversarial_points("mixed", 5000)
    t = FKDTree(coords=p.coords, ids=p.ids)
    t.build()
    lo, hi = [0.2, 0.2, 0.2], [0.5, 0.6, 0.7]
    got = np.sort(t.search_in_the_box(lo, hi))
    inside = np.all((p.coords >= np.float32(lo)[:, None]) & (p.coords <= np.float32(hi)[:, None]), axis=0)
    assert np.array_equal(got, np.sort(p.ids[inside]))
    for variant in (t.search_in_the_box_BFS, t.search_in_the_box_recursive, t.search_in_the_box_branchless):
        assert np.array_equal(np.sort(variant(lo, hi)), got)


def test_uniform_100k_all_points(require_gpu):
    p, b = gen.workload_uniform_100k()
    check_search(p, b, "uniform_100k", brute=False)


def test_mixed_1m_sample(require_gpu):
    p, b = gen.workload_mixed_1m()
    sub = gen.Boxes(np.ascontiguousarray(b.mins[:, :20000]), np.ascontiguousarray(b.maxs[:, :20000]))
    check_search(p, sub, "mixed_1m[:20000]", brute=False)


def test_events_search(require_gpu):
    b = gen.workload_events_2048(n_events=64, scale=0.05)
    boxes = gen.event_self_boxes(b, 0.01)
    qev = np.repeat(np.arange(b.n_events, dtype=np.uint32), np.diff(b.offsets))
    print(f"[search] events n={b.points.n}", flush=True)
    t = FKDEvents(b.points.coords, b.points.ids, b.offsets)
    t.build()
    goff, gids = t.search(boxes.mins, boxes.maxs, qev)
    tc, ti = oracle.build_events(b.points.coords, b.points.ids, b.offsets)
    ooff, oids = oracle.search_events(tc, ti, b.offsets, qev, boxes.mins, boxes.maxs)
    assert np.array_equal(goff, ooff)
    assert np.array_equal(sorted_sets(goff, gids), sorted_sets(ooff, oids))


def test_events_query_naming_no_event(require_gpu):
    """A box whose event id is not an event of the handle matches nothing;
    the other boxes keep their results."""
    b = gen.workload_events_2048(n_events=8, scale=0.02)
    boxes = gen.event_self_boxes(b, 0.02)
    qev = np.repeat(np.arange(b.n_events, dtype=np.uint32), np.diff(b.offsets))
    bad = qev.copy()
    bad[::7] = np.uint32(b.n_events)
    bad[3::11] = np.uint32(0xFFFFFFFF)
    print(f"[search] events with bad ids n={b.points.n}", flush=True)
    t = FKDEvents(b.points.coords, b.points.ids, b.offsets)
    t.build()
    goff, gids = t.search(boxes.mins, boxes.maxs, bad)
    ooff, oids = t.search(boxes.mins, boxes.maxs, qev)
    t.close()
    cnt, ocnt = np.diff(goff), np.diff(ooff)
    named = bad < b.n_events
    assert np.all(cnt[~named] == 0)
    assert np.array_equal(cnt[named], ocnt[named])
    keep = np.repeat(named, ocnt.astype(np.int64))
    assert np.array_equal(sorted_sets(goff, gids), sorted_sets(np.concatenate([[0], np.cumsum(ocnt * named)]).astype(np.uint64), oids[keep]))


def _device_search(t, boxes, single: bool, capacity=None):
    from fkdtree_amd import DeviceArray
    q = boxes.q
    dmi = DeviceArray.from_host(boxes.mins)
    dma = DeviceArray.from_host(boxes.maxs)
    off = DeviceArray(q + 1, np.uint64)
    if single:
        out = DeviceArray(max(capacity or 1, 1), np.uint32)
        t.search_single_device(q, dmi.ptr, dma.ptr, off.ptr, out.ptr, capacity or 0)
        total, ok = t.search_status()
        return off, out, total, ok, (dmi, dma)
    t.search_device(q, dmi.ptr, dma.ptr, off.ptr, None, 0, fill=False)
    total = int(off.download()[-1])
    out = DeviceArray(max(total, 1), np.uint32)
    t.search_device(q, dmi.ptr, dma.ptr, off.ptr, out.ptr, total, count=False)
    return off, out, total, True, (dmi, dma)


def test_one_call_search_capacity_and_parity(require_gpu):
    """fkd_search (plan, count, scan and fill in one enqueue) reports
    incomplete output for too small a capacity, writes exactly the first
    `capacity` ids of the full result, and with enough capacity matches the
    two-pass calls and the oracle exactly."""
    p = adversarial_points("mixed", 60_000)
    boxes = self_boxes(p, 0.03)
    print(f"[search] one-call capacity n={p.n} q={boxes.q}", flush=True)
    t = FKDTree(coords=p.coords, ids=p.ids)
    t.build()
    off2, out2, total2, _, _ = _device_search(t, boxes, single=False)
    _, _, total, ok, _ = _device_search(t, boxes, single=True, capacity=0)
    assert total == total2 and not ok
    cut = total // 3
    offc, outc, total, ok, keep = _device_search(t, boxes, single=True, capacity=cut)
    assert total == total2 and not ok
    assert np.array_equal(outc.download(count=cut), out2.download(count=cut))
    off1, out1, total, ok, keep = _device_search(t, boxes, single=True, capacity=total)
    assert ok and total == total2
    o1, o2 = off1.download(), off2.download()
    assert np.array_equal(o1, o2)
    i1 = out1.download(count=total)
    i2 = out2.download(count=total)
    assert np.array_equal(sorted_sets(o1, i1), sorted_sets(o2, i2))
    tc, ti = oracle.build(p.coords, p.ids)
    ooff, oids = oracle.search(tc, ti, boxes.mins, boxes.maxs)
    assert np.array_equal(o1, ooff)
    assert np.array_equal(sorted_sets(o1, i1), sorted_sets(ooff, oids))

@pytest.mark.parametrize("kind", KINDS)
def test_visited_counts_match_oracle(require_gpu, kind):
    """fkd_search_stats' per-box hit and visited-node counts equal the
    oracle's level-BFS walk (FKDTree_cpu.h:150-157): the push rule visits the
    same nodes, including for inverted boxes and NaN bounds."""
    from fkdtree_amd import DeviceArray
    from fkdtree_amd._lib import check, lib
    for n in (1, 2, 7, 100, 4097):
        p = adversarial_points(kind, n)
        for label, boxes in (("hw0.05", self_boxes(p, 0.05)), ("special", special_boxes(p))):
            print(f"[visited] {kind} n={n} {label} q={boxes.q}", flush=True)
            t = FKDTree(coords=p.coords, ids=p.ids)
            t.build()
            dmi = DeviceArray.from_host(boxes.mins)
            dma = DeviceArray.from_host(boxes.maxs)
            cnt = DeviceArray(boxes.q, np.uint32)
            vis = DeviceArray(boxes.q, np.uint32)
            check(lib().fkd_search_stats(t._h, boxes.q, dmi.ptr, dma.ptr, None, cnt.ptr, vis.ptr, None))
            tc, ti = oracle.build(p.coords, p.ids)
            ocnt, ovis = oracle.search_counts(tc, ti, boxes.mins, boxes.maxs)
            assert np.array_equal(cnt.download(), ocnt), f"{kind} n={n} {label}: counts differ"
            assert np.array_equal(vis.download(), ovis), f"{kind} n={n} {label}: visited differ"
            t.close()


@pytest.mark.parametrize("dim", [1, 2, 4])
def test_visited_counts_other_dims(require_gpu, dim):
    """The same for dims 1, 2 and 4 (box bounds in registers for dim 1, 3 and
    7 bound pairs per lane in LDS for dims 2 and 4, rotated by depth % dim):
    hit and visited-node counts of every box equal the oracle's walk
    (FKDTree_cpu.h:150-157), the id sets of the searches its level-BFS search
    (FKDTree_cpu.h:123-169), special boxes included."""
    from fkdtree_amd import DeviceArray
    from fkdtree_amd._lib import check, lib
    for n in (1, 2, 7, 100, 4097, 20_000):
        p = adversarial_points("mixed", n, dim=dim)
        for label, boxes in (("hw0.04", self_boxes(p, 0.04)), ("special", special_boxes(p))):
            print(f"[visited] dim {dim} n={n} {label} q={boxes.q}", flush=True)
            t = FKDTree(coords=p.coords, ids=p.ids)
            t.build()
            dmi = DeviceArray.from_host(boxes.mins)
            dma = DeviceArray.from_host(boxes.maxs)
            cnt = DeviceArray(boxes.q, np.uint32)
            vis = DeviceArray(boxes.q, np.uint32)
            check(lib().fkd_search_stats(t._h, boxes.q, dmi.ptr, dma.ptr, None, cnt.ptr, vis.ptr, None))
            tc, ti = oracle.build(p.coords, p.ids)
            ocnt, ovis = oracle.search_counts(tc, ti, boxes.mins, boxes.maxs)
            assert np.array_equal(cnt.download(), ocnt), f"dim {dim} n={n} {label}: counts differ"
            assert np.array_equal(vis.download(), ovis), f"dim {dim} n={n} {label}: visited differ"
            goff, gids = t.search_in_the_box_multiple_csr(boxes.mins, boxes.maxs)
            ooff, oids = oracle.search(tc, ti, boxes.mins, boxes.maxs)
            assert np.array_equal(goff, ooff), f"dim {dim} n={n} {label}: search counts differ"
            assert np.array_equal(sorted_sets(goff, gids), sorted_sets(ooff, oids)), f"dim {dim} n={n} {label}"
            t.close()


def test_event_visited_counts_match_oracle(require_gpu):
    """The statistics pass over a batch of events (the visited-node totals the
    bench's per-event byte model uses): every box's hit and visited-node count
    equals the oracle's walk of its own event's tree (FKDTree_cpu.h:150-157),
    boxes in mixed event order, events of every size including empty ones."""
    from fkdtree_amd import DeviceArray, FKDEvents, gen
    from fkdtree_amd._lib import check, lib
    sizes = [5000, 0, 1, 2, 17, 1023, 1024, 30_000, 3, 0, 70_001]
    b = gen.concat_events([gen.mixed_event(s, 61, e) for e, s in enumerate(sizes)])
    boxes = gen.event_self_boxes(b, 0.015)
    qev = np.repeat(np.arange(b.n_events, dtype=np.uint32), np.diff(b.offsets))
    perm = np.argsort(gen.u64(61, 9, boxes.q), kind="stable")
    mins, maxs, qev = boxes.mins[:, perm].copy(), boxes.maxs[:, perm].copy(), qev[perm].copy()
    print(f"[visited] events n={b.points.n} q={boxes.q}", flush=True)
    tc, ti = oracle.build_events(b.points.coords, b.points.ids, b.offsets)
    ooff, _, ovis = oracle.search_events(tc, ti, b.offsets, qev, mins, maxs, with_visited=True)
    ev = FKDEvents(b.points.coords, b.points.ids, b.offsets)
    ev.build()
    dmi, dma, dq = DeviceArray.from_host(mins), DeviceArray.from_host(maxs), DeviceArray.from_host(qev)
    cnt = DeviceArray(boxes.q, np.uint32)
    vis = DeviceArray(boxes.q, np.uint32)
    check(lib().fkd_search_stats(ev._h, boxes.q, dmi.ptr, dma.ptr, dq.ptr, cnt.ptr, vis.ptr, None))
    assert np.array_equal(cnt.download().astype(np.int64), np.diff(ooff).astype(np.int64))
    assert np.array_equal(vis.download(), ovis)
    ev.close()
