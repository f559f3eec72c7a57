"""Launch parameters change time only: the two-pass search and the one-call
search return the oracle's id sets (FKDTree_cpu.h:123-169) for every query
order (`query_key_bits`, 0 = the caller's order), fill dispatch order, fill
window (boxes reordered by work) and count window (boxes reordered by
volume before counting); 30,000 boxes, so windows of 256, 1024 and 4096 with
a partial last window."""
import numpy as np
import pytest

import oracle
from fkdtree_amd import DeviceArray, FKDTree
from tests.helpers import adversarial_points, self_boxes, sorted_sets

pytestmark = pytest.mark.gpu


def _two_pass(t, dmi, dma, q):
    off = DeviceArray(q + 1, np.uint64)
    t.search_device(q, dmi.ptr, dma.ptr, off.ptr, None, 0, fill=False)
    o = off.download()
    out = DeviceArray(max(int(o[-1]), 1), np.uint32)
    t.search_device(q, dmi.ptr, dma.ptr, off.ptr, out.ptr, int(o[-1]), count=False)
    return o, out.download(count=int(o[-1]))


def _one_call(t, dmi, dma, q, cap):
    off = DeviceArray(q + 1, np.uint64)
    out = DeviceArray(max(cap, 1), np.uint32)
    for _ in range(3):
        t.search_single_device(q, dmi.ptr, dma.ptr, off.ptr, out.ptr, cap)
        total, ok = t.search_status()
        if ok:
            break
    assert ok and total <= cap
    o = off.download()
    return o, out.download(count=int(o[-1]))


def test_tuning_keeps_results(require_gpu):
    p = adversarial_points("mixed", 30_000)
    boxes = self_boxes(p, 0.04)
    print(f"[tuning] n={p.n} q={boxes.q}", flush=True)
    tc, ti = oracle.build(p.coords, p.ids)
    ooff, oids = oracle.search(tc, ti, boxes.mins, boxes.maxs)
    want = sorted_sets(ooff, oids)
    t = FKDTree(coords=p.coords, ids=p.ids)
    t.build()
    dmi = DeviceArray.from_host(boxes.mins)
    dma = DeviceArray.from_host(boxes.maxs)
    for kb, fo, fw, cw, vb in ((0, 1, 0, 0, 4), (1, 0, 256, 4096, 3), (5, 1, 4096, 1024, 4), (10, 0, 1024, 0, 6),
                               (8, 1, 1024, 256, 0), (8, 0, 0, 4096, 2), (2, 1, 4096, 0, 4)):
        t.set_tuning(query_key_bits=kb, fill_order=fo, fill_window=fw, count_window=cw, key_volume_bits=vb)
        assert t.get_tuning("query_key_bits") == kb and t.get_tuning("fill_order") == fo
        assert t.get_tuning("fill_window") == fw and t.get_tuning("count_window") == cw
        assert t.get_tuning("key_volume_bits") == vb
        tag = f"key bits {kb}, fill order {fo}, window {fw}, count window {cw}, volume bits {vb}"
        o, ids = _two_pass(t, dmi, dma, boxes.q)
        assert np.array_equal(o, ooff), f"{tag}: counts differ"
        assert np.array_equal(sorted_sets(o, ids), want), f"{tag}: ids differ"
    for fo, fw, cw in ((0, 0, 0), (1, 1024, 4096), (0, 4096, 1024)):
        t.set_tuning(fill_order=fo, fill_window=fw, count_window=cw)
        o, ids = _one_call(t, dmi, dma, boxes.q, int(ooff[-1]))
        assert np.array_equal(o, ooff), f"single call, fill order {fo}, window {fw}: counts differ"
        assert np.array_equal(sorted_sets(o, ids), want), f"single call, fill order {fo}, window {fw}: ids differ"
    t.close()


def test_tuning_changed_between_passes(require_gpu):
    """The count pass records each box's work only when the fill will order
    boxes or blocks by it; a fill whose tuning asks for ordering after a count
    that recorded none runs unordered, and the reverse ignores the recorded
    work.  Either way the id sets are the oracle's (FKDTree_cpu.h:123-169)."""
    p = adversarial_points("mixed", 30_000)
    boxes = self_boxes(p, 0.04)
    tc, ti = oracle.build(p.coords, p.ids)
    ooff, oids = oracle.search(tc, ti, boxes.mins, boxes.maxs)
    want = sorted_sets(ooff, oids)
    t = FKDTree(coords=p.coords, ids=p.ids)
    t.build()
    dmi = DeviceArray.from_host(boxes.mins)
    dma = DeviceArray.from_host(boxes.maxs)
    q = boxes.q
    for before, after in (((0, 0), (1, 1024)), ((1, 4096), (0, 0)), ((0, 0), (0, 256)), ((1, 0), (0, 4096))):
        off = DeviceArray(q + 1, np.uint64)
        t.set_tuning(fill_order=before[0], fill_window=before[1])
        t.search_device(q, dmi.ptr, dma.ptr, off.ptr, None, 0, fill=False)
        o = off.download()
        assert np.array_equal(o, ooff), f"{before} -> {after}: counts differ"
        out = DeviceArray(int(o[-1]), np.uint32)
        t.set_tuning(fill_order=after[0], fill_window=after[1])
        t.search_device(q, dmi.ptr, dma.ptr, off.ptr, out.ptr, int(o[-1]), count=False)
        assert np.array_equal(sorted_sets(o, out.download()), want), f"{before} -> {after}: ids differ"
    t.close()


def test_tuning_rejects_bad_values(require_gpu):
    p = adversarial_points("mixed", 100)
    t = FKDTree(coords=p.coords, ids=p.ids)
    for k, v in (("pool_batch", 32), ("query_key_bits", 11), ("fill_order", 2), ("fill_window", 512),
                 ("count_window", 100), ("key_volume_bits", 7), ("bucket", 3), ("count_bucket", 3),
                 ("bucket_park", 9), ("no_such_key", 1)):
        with pytest.raises(RuntimeError):
            t.set_tuning(**{k: v})
    t.close()


def test_event_batch_tuning_keeps_results(require_gpu):
    """A handle of several events: every launch setting returns the oracle's
    per-event id sets (FKDTree_cpu.h:123-169 on each event's own tree)."""
    from fkdtree_amd import FKDEvents, gen
    b = gen.workload_events_2048(n_events=24, scale=0.04)
    boxes = gen.event_self_boxes(b, 0.015)
    qev = np.repeat(np.arange(b.n_events, dtype=np.uint32), np.diff(b.offsets))
    print(f"[tuning] events n={b.points.n} q={boxes.q}", flush=True)
    tc, ti = oracle.build_events(b.points.coords, b.points.ids, b.offsets)
    ooff, oids = oracle.search_events(tc, ti, b.offsets, qev, boxes.mins, boxes.maxs)
    want = sorted_sets(ooff, oids)
    t = FKDEvents(b.points.coords, b.points.ids, b.offsets)
    t.build()
    assert t.get_tuning("fill_order") == 0   # the default for several events
    for kb, fo, fw, cw, vb in ((8, 0, 4096, 0, 4), (8, 1, 256, 1024, 0), (3, 0, 0, 4096, 3), (0, 1, 1024, 256, 4),
                               (10, 0, 4096, 0, 6)):
        t.set_tuning(query_key_bits=kb, fill_order=fo, fill_window=fw, count_window=cw, key_volume_bits=vb)
        off, ids = t.search(boxes.mins, boxes.maxs, qev)
        tag = f"key bits {kb}, fill order {fo}, window {fw}, count window {cw}, volume bits {vb}"
        assert np.array_equal(off, ooff), f"{tag}: counts differ"
        assert np.array_equal(sorted_sets(off, ids), want), f"{tag}: ids differ"
    t.close()


GEOMETRIES = [(fl, ri, lp) for fl in (9, 10, 11) for ri in (8, 16) for lp in (1, 0)]


@pytest.mark.parametrize("kind,n,dim", [("mixed", 300_000, 3), ("shuffled_ids", 70_000, 3), ("grid", 40_000, 2),
                                        ("uniform", 50_000, 4), ("zeros", 9_000, 1)])
def test_build_geometry_keeps_trees(require_gpu, kind, n, dim):
    """Finisher capacity (fin_log: subtrees of < 2^fin_log points built in
    LDS), digit-pass tile (build_radix_items) and the two-depth partition of
    dim 3 (level_pairs) change time only: every
    setting builds the oracle's tree bit for bit (FKDTree_cpu.h:221-277), on
    one handle re-laid out between builds, and searches it exactly."""
    from tests.helpers import bits
    p = adversarial_points(kind, n, dim=dim)
    boxes = self_boxes(p, 0.02)
    print(f"[geometry] {kind} n={n} dim={dim}", flush=True)
    oc, oi = oracle.build(p.coords, p.ids)
    q = min(n, 5000)
    ooff, oids = oracle.search(oc, oi, boxes.mins[:, :q], boxes.maxs[:, :q])
    t = FKDTree(coords=p.coords, ids=p.ids)
    for fl, ri, lp in GEOMETRIES:
        t.set_tuning(fin_log=fl, build_radix_items=ri, level_pairs=lp)
        assert t.get_tuning("fin_log") == fl and t.get_tuning("build_radix_items") == ri
        assert t.get_tuning("level_pairs") == lp
        t.build()
        c, i = t.tree_arrays()
        assert np.array_equal(bits(c), bits(oc)) and np.array_equal(i, oi), (fl, ri, lp)
        off, ids = t.search_in_the_box_multiple_csr(boxes.mins[:, :q], boxes.maxs[:, :q])
        assert np.array_equal(off, ooff), (fl, ri, lp)
        assert np.array_equal(sorted_sets(off, ids), sorted_sets(ooff, oids)), (fl, ri, lp)
    t.close()


def test_build_geometry_events(require_gpu):
    """The same for a batch of events of every size class: empty, one point,
    below and above each finisher capacity (511, 1,023, 2,047 points)."""
    from fkdtree_amd import FKDEvents, gen
    from tests.helpers import bits
    sizes = [0, 1, 2, 510, 511, 512, 1022, 1023, 1024, 2046, 2047, 2048, 4097, 30_000, 0, 70_001, 3]
    evs = [gen.mixed_event(s, 77, e) for e, s in enumerate(sizes)]
    b = gen.concat_events(evs)
    oc, oi = oracle.build_events(b.points.coords, b.points.ids, b.offsets)
    t = FKDEvents(b.points.coords, b.points.ids, b.offsets)
    for fl, ri, lp in GEOMETRIES[::-1]:
        t.set_tuning(fin_log=fl, build_radix_items=ri, level_pairs=lp)
        t.build()
        c, i = t.tree_arrays()
        assert np.array_equal(bits(c), bits(oc)) and np.array_equal(i, oi), (fl, ri, lp)
    t.close()


def test_build_geometry_rejects_bad_values(require_gpu):
    p = adversarial_points("mixed", 100)
    t = FKDTree(coords=p.coords, ids=p.ids)
    for k, v in (("fin_log", 8), ("fin_log", 12), ("build_radix_items", 4), ("build_radix_items", 32),
                 ("level_pairs", 2)):
        with pytest.raises(RuntimeError):
            t.set_tuning(**{k: v})
    t.close()
