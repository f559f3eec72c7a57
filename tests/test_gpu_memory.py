"""The library's device-memory cache: buffers freed in stream order are
reused by later handles (on the same or other streams) only after the
freeing stream's earlier work has completed, so results stay exact under
heavy reuse; idle cached memory counts as free and goes back to the driver
on request."""
import numpy as np
import pytest

import oracle
import fkdtree_amd as F
from fkdtree_amd import FKDTree, Stream, gen
from tests.helpers import bits, sorted_sets

pytestmark = pytest.mark.gpu


def test_reuse_across_streams_stays_exact(require_gpu):
    streams = [Stream(), Stream()]
    sets = []
    for k in range(3):
        p = gen.mixed_points(60_000 + 1000 * k, 300 + k, n_blobs=60)
        b = gen.boxes_around(p.coords[:, :3000], 0.02)
        oc, oi = oracle.build(p.coords, p.ids)
        ooff, oids = oracle.search(oc, oi, b.mins, b.maxs)
        sets.append((p, b, oc, oi, ooff, sorted_sets(ooff, oids)))
    for rep in range(6):
        for k, (p, b, oc, oi, ooff, want) in enumerate(sets):
            s = streams[(rep + k) % 2]
            t = FKDTree(coords=p.coords, ids=p.ids, stream=s)
            t.build()
            c, i = t.tree_arrays()
            assert np.array_equal(bits(c), bits(oc)) and np.array_equal(i, oi), (rep, k)
            off, ids = t.search_in_the_box_multiple_csr(b.mins, b.maxs)
            assert np.array_equal(off, ooff) and np.array_equal(sorted_sets(off, ids), want), (rep, k)
            t.close()      # freed in stream order on s; the next handle may take the blocks
    for s in streams:
        s.sync()


def test_cached_memory_counts_as_free_and_is_released(require_gpu):
    s = Stream()
    p = gen.mixed_points(200_000, 9, n_blobs=100)
    t = FKDTree(coords=p.coords, ids=p.ids, stream=s)
    t.build()
    held = t.device_bytes()
    t.close()
    s.sync()
    info = F.device_info()
    assert info["pool_cached_bytes"] >= held          # kept for reuse, counted as free
    F.release_cached()
    info2 = F.device_info()
    assert info2["pool_cached_bytes"] == 0
    assert info2["device_free_bytes"] >= info["device_free_bytes"] + held
