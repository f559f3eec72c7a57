"""HIP-graph capture of build + search (fkd_graph_create): every replay,
interleaved with eager work and synchronisations, reproduces the eager
results exactly (tree arrays and every box's id set)."""
import numpy as np
import pytest

from fkdtree_amd import DeviceArray, FKDEvents, Graph, Stream, gen
from tests.helpers import bits, sorted_sets

pytestmark = pytest.mark.gpu


def test_graph_replays_match_eager(require_gpu):
    b = gen.workload_events_2048(n_events=48, scale=0.05)
    boxes = gen.event_self_boxes(b, 0.01)
    P = b.points.n
    qev = np.repeat(np.arange(b.n_events, dtype=np.uint32), np.diff(b.offsets))
    print(f"[graph] events={b.n_events} points={P}", flush=True)
    s = Stream()
    ev = FKDEvents(b.points.coords, b.points.ids, b.offsets, stream=s)
    d_mi, d_ma = DeviceArray.from_host(boxes.mins, s), DeviceArray.from_host(boxes.maxs, s)
    d_q = DeviceArray.from_host(qev, s)
    # eager reference
    ev.build()
    e_off, e_ids = ev.search(boxes.mins, boxes.maxs, qev)
    e_tree = [x.copy() for x in ev.tree_arrays()]
    total = int(e_off[-1])
    d_off = DeviceArray(P + 1, np.uint64)
    d_out = DeviceArray(total, np.uint32)
    g = Graph(ev, P, d_mi.ptr, d_ma.ptr, d_off.ptr, d_out.ptr, total, s, d_qev=d_q.ptr)
    for rep in range(4):
        d_out.upload(np.zeros(total, np.uint32), s)       # eager work between replays
        g.launch()
        s.sync()
        off = d_off.download(s)
        ids = d_out.download(s)
        assert np.array_equal(off, e_off), rep
        assert np.array_equal(sorted_sets(off, ids), sorted_sets(e_off, e_ids)), rep
        ev._mirror = None
        c, i = ev.tree_arrays()
        assert np.array_equal(bits(c), bits(e_tree[0])) and np.array_equal(i, e_tree[1]), rep
        # an eager search in between must not disturb the next replay
        ev.search(boxes.mins[:, :100], boxes.maxs[:, :100], qev[:100])
    g.close()
