"""HIP-graph capture of build + search (fkd_graph_create): every replay,
interleaved with eager work and synchronisations, reproduces the oracle's
tree arrays (FKDTree_cpu.h:221-277) and every box's id set
(FKDTree_cpu.h:123-169) exactly.  While a graph exists the handle refuses
calls that would move the buffers the graph holds (a larger search, a trim),
and a graph whose tree is gone refuses to replay."""
import numpy as np
import pytest

import oracle
from fkdtree_amd import DeviceArray, FKDError, FKDEvents, FKDTree, Graph, Stream, gen
from tests.helpers import bits, sorted_sets

pytestmark = pytest.mark.gpu

FKD_ERR_STATE = -4


def test_graph_replays_match_oracle(require_gpu):
    b = gen.workload_events_2048(n_events=48, scale=0.05)
    boxes = gen.event_self_boxes(b, 0.01)
    P = b.points.n
    qev = np.repeat(np.arange(b.n_events, dtype=np.uint32), np.diff(b.offsets))
    print(f"[graph] events={b.n_events} points={P}", flush=True)
    oc, oi = oracle.build_events(b.points.coords, b.points.ids, b.offsets)
    o_off, o_ids = oracle.search_events(oc, oi, b.offsets, qev, boxes.mins, boxes.maxs)
    o_sets = sorted_sets(o_off, o_ids)
    s = Stream()
    ev = FKDEvents(b.points.coords, b.points.ids, b.offsets, stream=s)
    d_mi, d_ma = DeviceArray.from_host(boxes.mins, s), DeviceArray.from_host(boxes.maxs, s)
    d_q = DeviceArray.from_host(qev, s)
    total = int(o_off[-1])
    d_off = DeviceArray(P + 1, np.uint64)
    d_out = DeviceArray(total, np.uint32)
    g = Graph(ev, P, d_mi.ptr, d_ma.ptr, d_off.ptr, d_out.ptr, total, s, d_qev=d_q.ptr)
    for rep in range(4):
        d_out.upload(np.zeros(total, np.uint32), s)       # eager work between replays
        d_off.upload(np.zeros(P + 1, np.uint64), s)
        g.launch()
        s.sync()
        assert ev.search_status(s) == (total, True), rep
        off = d_off.download(s)
        ids = d_out.download(s)
        assert np.array_equal(off, o_off), rep
        assert np.array_equal(sorted_sets(off, ids), o_sets), rep
        ev._mirror = None
        c, i = ev.tree_arrays()
        assert np.array_equal(bits(c), bits(oc)) and np.array_equal(i, oi), rep
        # an eager search of a subset in between must not disturb the next replay
        k = 100
        e_off, e_ids = ev.search(boxes.mins[:, :k], boxes.maxs[:, :k], qev[:k])
        assert np.array_equal(sorted_sets(e_off, e_ids), sorted_sets(o_off[:k + 1], o_ids[:int(o_off[k])]))
    g.close()


def test_graph_pins_handle_buffers(require_gpu):
    """A larger eager search (its plan buffers would be reallocated) and a
    trim are refused while a graph holds the handle's scratch; after the
    graph is destroyed both work.  A graph whose tree was destroyed refuses
    to launch (nothing is replayed on freed memory)."""
    p = gen.mixed_points(20_000, 77, n_blobs=20)
    small = gen.boxes_around(p.coords[:, :1000], 0.02)
    large = gen.boxes_around(p.coords, 0.02)
    s = Stream()
    t = FKDTree(coords=p.coords, ids=p.ids, stream=s)
    t.build()
    oc, oi = oracle.build(p.coords, p.ids)
    o_off, o_ids = oracle.search(oc, oi, small.mins, small.maxs)
    nq = small.q
    d_mi, d_ma = DeviceArray.from_host(small.mins, s), DeviceArray.from_host(small.maxs, s)
    d_off = DeviceArray(nq + 1, np.uint64)
    d_out = DeviceArray(int(o_off[-1]), np.uint32)
    g = Graph(t, nq, d_mi.ptr, d_ma.ptr, d_off.ptr, d_out.ptr, int(o_off[-1]), s)
    # a smaller eager search fits the held buffers
    h_off, h_ids = t.search_in_the_box_multiple_csr(small.mins[:, :10], small.maxs[:, :10])
    assert np.array_equal(sorted_sets(h_off, h_ids), sorted_sets(o_off[:11], o_ids[:int(o_off[10])]))
    with pytest.raises(FKDError) as e:
        t.search_in_the_box_multiple_csr(large.mins, large.maxs)
    assert e.value.code == FKD_ERR_STATE and "graph" in str(e.value)
    with pytest.raises(FKDError) as e:
        t.trim()
    assert e.value.code == FKD_ERR_STATE
    g.launch()
    s.sync()
    assert np.array_equal(d_off.download(s), o_off)
    assert np.array_equal(sorted_sets(o_off, d_out.download(s)), sorted_sets(o_off, o_ids))
    g.close()
    l_off, l_ids = t.search_in_the_box_multiple_csr(large.mins, large.maxs)
    lo_off, lo_ids = oracle.search(oc, oi, large.mins, large.maxs)
    assert np.array_equal(sorted_sets(l_off, l_ids), sorted_sets(lo_off, lo_ids))
    t.trim()
    # a graph outliving its tree
    g2 = Graph(t, nq, d_mi.ptr, d_ma.ptr, d_off.ptr, d_out.ptr, int(o_off[-1]), s)
    t.close()
    with pytest.raises(FKDError) as e:
        g2.launch()
    assert e.value.code == FKD_ERR_STATE and "destroyed" in str(e.value)
    g2.close()
