"""Wire format (fkd_serialize / fkd_deserialize): a tree spilled and
reloaded without a build has the reference's arrays (FKDTree_cpu.h:209-213)
and the same search results; verification rebuilds and rejects arrays that
are not the k-d tree of their points."""
import numpy as np
import pytest

import fkdtree_amd as F
import oracle
from fkdtree_amd import gen
from fkdtree_amd._lib import FKDError
from tests.helpers import adversarial_points, self_boxes, sorted_sets

pytestmark = pytest.mark.gpu


def _same_search(a, b, boxes, qev=None):
    if qev is None:
        oa = a.search_in_the_box_multiple_csr(boxes.mins, boxes.maxs)
        ob = b.search_in_the_box_multiple_csr(boxes.mins, boxes.maxs)
    else:
        oa = a.search(boxes.mins, boxes.maxs, qev)
        ob = b.search(boxes.mins, boxes.maxs, qev)
    assert np.array_equal(oa[0], ob[0])
    assert np.array_equal(sorted_sets(*oa), sorted_sets(*ob))
    return ob


@pytest.mark.parametrize("dim", [1, 2, 3, 4])
def test_round_trip(require_gpu, dim):
    for n in ((1, 2, 7, 1000, 30000) if dim == 3 else (1, 7, 3000)):
        p = adversarial_points("mixed", n, dim=dim)
        print(f"[wire] dim={dim} n={n}", flush=True)
        t = F.FKDTree(coords=p.coords, ids=p.ids)
        t.build()
        data = t.serialize()
        assert len(data) == 24 + 16 + 4 * n * (dim + 1)
        u = F.deserialize(data)
        assert isinstance(u, F.FKDTree) and u.dim == dim and u.n == n
        tc, ti = t.tree_arrays()
        uc, ui = u.tree_arrays()
        assert np.array_equal(uc.view(np.uint32), tc.view(np.uint32)) and np.array_equal(ui, ti)
        boxes = self_boxes(p, 0.05)
        off, ids = _same_search(t, u, boxes)
        oc, oi = oracle.build(p.coords, p.ids)
        ooff, oids = oracle.search(oc, oi, boxes.mins, boxes.maxs)
        assert np.array_equal(off, ooff) and np.array_equal(sorted_sets(off, ids), sorted_sets(ooff, oids))
        u.build()  # the loaded points rebuild to the same tree
        uc, ui = u.tree_arrays()
        assert np.array_equal(uc.view(np.uint32), tc.view(np.uint32)) and np.array_equal(ui, ti)
        assert F.deserialize(data, verify=True).n == n
        t.close()
        u.close()


def test_events_save_load(require_gpu, tmp_path):
    sizes = [0, 1, 5, 0, 4097, 20000, 2]
    evs = [adversarial_points("shuffled_ids", s, seed=300 + i) if s else
           gen.PointSet(np.zeros((3, 0), np.float32), np.zeros(0, np.uint32)) for i, s in enumerate(sizes)]
    b = gen.concat_events(evs)
    print(f"[wire] events sizes={sizes}", flush=True)
    t = F.FKDEvents(b.points.coords, b.points.ids, b.offsets)
    t.build()
    path = str(tmp_path / "events.fkd")
    t.save(path)
    u = F.load(path, verify=True)
    assert isinstance(u, F.FKDEvents) and np.array_equal(u.offsets, b.offsets)
    tc, ti = t.tree_arrays()
    uc, ui = u.tree_arrays()
    assert np.array_equal(uc.view(np.uint32), tc.view(np.uint32)) and np.array_equal(ui, ti)
    boxes = gen.event_self_boxes(b, 0.03)
    qev = np.repeat(np.arange(len(sizes), dtype=np.uint32), np.diff(b.offsets).astype(np.int64))
    _same_search(t, F.load(path), boxes, qev)


def test_verify_rejects_a_tree_that_is_not_one(require_gpu):
    p = adversarial_points("mixed", 5000)
    t = F.FKDTree(coords=p.coords, ids=p.ids)
    t.build()
    data = bytearray(t.serialize())
    n = p.n
    # swap the coordinates and ids of heap nodes 1 and 2 (left and right son of the root)
    base = 24 + 16
    for k in range(4):
        a = base + k * 4 * n + 4 * 1
        c = base + k * 4 * n + 4 * 2
        data[a:a + 4], data[c:c + 4] = data[c:c + 4], data[a:a + 4]
    F.deserialize(bytes(data))  # accepted unverified: the arrays are taken as they are
    with pytest.raises(FKDError) as e:
        F.deserialize(bytes(data), verify=True)
    assert e.value.code == -1
