"""Parity at the workloads' full sizes.

* 1,048,576-point mixed tree and 4,194,304-point uniform tree: the tree
  arrays hash to the digests the compiled reference build wrote
  (tests/golden/manifest.json, FKDTree_cpu.h:221-296), the first 10,000
  boxes' sorted id sets hash to the reference search's digest
  (FKDTree_cpu.h:123-169), and every one of the 1M / 4M boxes returns the
  oracle's exact id set.
* 2,048 events (77,201,253 points): every event's tree equals the oracle's,
  every one of the 77M self-centred boxes has the oracle's hit count (12.1e9
  hits in all), and the boxes of every 64th event return the oracle's exact
  id sets.
"""
import json
import os

import numpy as np
import pytest

import oracle
from fkdtree_amd import DeviceArray, FKDEvents, FKDTree, gen
from tests.helpers import bits, sorted_sets

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(HERE, "golden", "manifest.json")) as f:
    WORKLOADS = json.load(f)["workloads"]


def result_digest(off, ids):
    return gen.digest(np.diff(off).astype(np.uint32), sorted_sets(off, ids))


def check_single_tree(key, p, b):
    w = WORKLOADS[key]
    assert gen.digest(p.coords, p.ids) == w["input_sha"]
    assert gen.digest(b.mins, b.maxs) == w["boxes_sha"]
    print(f"[fullsize] {key} n={p.n} q={b.q}", flush=True)
    t = FKDTree(coords=p.coords, ids=p.ids)
    t.build()
    gc, gi = t.tree_arrays()
    assert gen.digest(gc, gi) == w["tree_sha"], f"{key}: tree differs from the reference build"
    goff, gids = t.search_in_the_box_multiple_csr(b.mins, b.maxs)
    t.close()
    s = w["search"]["sample"]
    assert int(goff[s]) == w["search"]["total"], f"{key}: sample hit total"
    assert result_digest(goff[: s + 1], gids[: int(goff[s])]) == w["search"]["result_sha"], \
        f"{key}: sample ids differ from the reference search"
    ooff, oids = oracle.search(gc, gi, b.mins, b.maxs)
    assert np.array_equal(goff, ooff), f"{key}: counts differ"
    assert np.array_equal(sorted_sets(goff, gids), sorted_sets(ooff, oids)), f"{key}: ids differ"
    print(f"[fullsize] {key} hits={int(goff[-1])}", flush=True)


@pytest.mark.timeout(600)
def test_mixed_1m_all_boxes(require_gpu):
    p, b = gen.workload_mixed_1m()
    check_single_tree("mixed_1m", p, b)


@pytest.mark.timeout(600)
def test_uniform_4m_all_boxes(require_gpu):
    p, b = gen.workload_uniform_4m()
    check_single_tree("uniform_4m", p, b)


@pytest.mark.timeout(900)
def test_events_2048_full(require_gpu):
    b = gen.workload_events_2048()
    P, E = b.points.n, b.n_events
    print(f"[fullsize] events E={E} P={P}", flush=True)
    t = FKDEvents(b.points.coords, b.points.ids, b.offsets)
    t.build()
    gc, gi = t.tree_arrays()
    oc, oi = oracle.build_events(b.points.coords, b.points.ids, b.offsets)
    assert np.array_equal(gi, oi), "event trees: ids differ"
    assert np.array_equal(bits(gc), bits(oc)), "event trees: coordinates differ"
    del oc, oi
    boxes = gen.event_self_boxes(b, 0.01)
    qev = np.repeat(np.arange(E, dtype=np.uint32), np.diff(b.offsets))
    dmi, dma, dq = DeviceArray.from_host(boxes.mins), DeviceArray.from_host(boxes.maxs), DeviceArray.from_host(qev)
    doff = DeviceArray(P + 1, np.uint64)
    t.search_device(P, dmi.ptr, dma.ptr, doff.ptr, None, 0, d_qev=dq.ptr, fill=False)
    goff = doff.download()
    del dmi, dma, dq, doff
    want = oracle.search_events_counts(gc, gi, b.offsets, qev, boxes.mins, boxes.maxs)
    assert np.array_equal(np.diff(goff), want.astype(np.uint64)), "event box counts differ"
    print(f"[fullsize] events hits={int(goff[-1])}", flush=True)
    sel = np.isin(qev, np.arange(0, E, 64, dtype=np.uint32))
    sub = gen.Boxes(np.ascontiguousarray(boxes.mins[:, sel]), np.ascontiguousarray(boxes.maxs[:, sel]))
    sq = np.ascontiguousarray(qev[sel])
    soff, sids = t.search(sub.mins, sub.maxs, sq)
    ooff, oids = oracle.search_events(gc, gi, b.offsets, sq, sub.mins, sub.maxs)
    assert np.array_equal(soff, ooff), "sampled events: counts differ"
    assert np.array_equal(sorted_sets(soff, sids), sorted_sets(ooff, oids)), "sampled events: ids differ"
    t.close()


@pytest.mark.timeout(900)
def test_output_beyond_2_32_ids(require_gpu):
    """More than 2^32 ids in one result (52,000 boxes of half-width 0.25 in a
    1,048,576-point uniform tree, ≈ 4.6e9 hits, 18 GB of output): 64-bit
    offsets, every box's count against the oracle, and the exact id sets of
    the boxes around id index 2^32, past it and at both ends."""
    from fkdtree_amd._lib import check, lib
    from fkdtree_amd import _sh
    p = gen.uniform_points(1 << 20, 77)
    nq = 52_000
    rng = np.random.default_rng(5)
    b = gen.boxes_around(rng.random((3, nq), dtype=np.float32), 0.25)
    print(f"[fullsize] beyond 2^32 n={p.n} q={nq}", flush=True)
    t = FKDTree(coords=p.coords, ids=p.ids)
    t.build()
    dmi, dma = DeviceArray.from_host(b.mins), DeviceArray.from_host(b.maxs)
    doff = DeviceArray(nq + 1, np.uint64)
    t.search_device(nq, dmi.ptr, dma.ptr, doff.ptr, None, 0, fill=False)
    o = doff.download()
    total = int(o[-1])
    assert total > 1 << 32
    tc, ti = t.tree_arrays()
    cnt, _ = oracle.search_counts(tc, ti, b.mins, b.maxs)
    assert np.array_equal(np.diff(o), cnt.astype(np.uint64)), "counts differ"
    out = DeviceArray(total, np.uint32)
    t.search_device(nq, dmi.ptr, dma.ptr, doff.ptr, out.ptr, total, count=False)
    at = int(np.searchsorted(o, 1 << 32, side="right")) - 1   # the box holding id index 2^32
    assert o[at] <= 1 << 32 < o[at + 1]
    sel = sorted({0, 1, at - 1, at, at + 1, nq - 2, nq - 1, *range(at + 500, at + 510)})
    for q in sel:
        n = int(o[q + 1] - o[q])
        got = np.empty(n, np.uint32)
        if n:
            check(lib().fkd_memcpy_dtoh(got.ctypes.data, out.ptr + 4 * int(o[q]), 4 * n, _sh(None)))
        bmin = np.ascontiguousarray(b.mins[:, q:q + 1])
        bmax = np.ascontiguousarray(b.maxs[:, q:q + 1])
        boff, bids = oracle.brute(p.coords, p.ids, bmin, bmax)
        assert np.array_equal(np.sort(got), np.sort(bids)), f"box {q} (offset {int(o[q])})"
    out.free()
    t.close()
    print(f"[fullsize] beyond 2^32 hits={total}", flush=True)


@pytest.mark.timeout(600)
@pytest.mark.parametrize("dim", [1, 2, 4])
def test_other_dims_1m(require_gpu, dim):
    """1,048,576 points of dimension 1, 2 or 4 (mixed distribution): the tree
    arrays bit for bit and the id sets of 200,000 self-centred boxes exactly
    against the oracle (FKDTree_cpu.h:123-169, 221-296)."""
    p = gen.mixed_points(1 << 20, 900 + dim, n_blobs=1000, dim=dim)
    hw = {1: 2e-6, 2: 2e-3, 4: 0.03}[dim]
    b = gen.boxes_around(np.ascontiguousarray(p.coords[:, ::5][:, :200_000]), hw)
    print(f"[fullsize] dim={dim} n={p.n} q={b.q}", flush=True)
    t = FKDTree(coords=p.coords, ids=p.ids)
    t.build()
    gc, gi = t.tree_arrays()
    oc, oi = oracle.build(p.coords, p.ids)
    assert np.array_equal(gi, oi) and np.array_equal(bits(gc), bits(oc)), f"dim {dim}: tree differs"
    goff, gids = t.search_in_the_box_multiple_csr(b.mins, b.maxs)
    t.close()
    ooff, oids = oracle.search(oc, oi, b.mins, b.maxs)
    assert np.array_equal(goff, ooff), f"dim {dim}: counts differ"
    assert np.array_equal(sorted_sets(goff, gids), sorted_sets(ooff, oids)), f"dim {dim}: ids differ"
    print(f"[fullsize] dim={dim} hits={int(goff[-1])}", flush=True)
