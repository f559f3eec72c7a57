"""Host-side admission policy of the event server (no GPU)."""
import numpy as np
import pytest

from fkdtree_amd.serving import Request, batch_bytes, plan_batch


def reqs(sizes):
    return [Request(i, np.zeros((3, s), np.float32), np.arange(s, dtype=np.uint32), 0.0) for i, s in enumerate(sizes)]


def test_prefix_within_point_limit():
    q = reqs([1000, 2000, 3000, 4000])
    assert plan_batch(q, 1 << 40, max_points=6000, max_events=10, hits_per_point=10) == 3
    assert plan_batch(q, 1 << 40, max_points=10_000, max_events=2, hits_per_point=10) == 2


def test_memory_budget_limits_batch():
    q = reqs([100_000] * 5)
    one = batch_bytes(100_000, 100)
    assert plan_batch(q, int(one * 1.5), 10 ** 9, 100, 100) == 1
    assert plan_batch(q, int(batch_bytes(300_000, 100)), 10 ** 9, 100, 100) == 3
    assert plan_batch(q, one - 1, 10 ** 9, 100, 100) == 0


def test_oversized_head_still_admitted_alone_by_points():
    q = reqs([50_000, 10])
    assert plan_batch(q, 1 << 40, max_points=1000, max_events=10, hits_per_point=1) == 1


def _rec(n, tag):
    r = np.zeros((n, 4), dtype=np.uint32)
    r[:, 3] = tag
    return r


def test_ingest_ring_wraps_fills_and_releases_in_order():
    from fkdtree_amd.serving import IngestRing

    ring = IngestRing(10, 3, buffer=np.zeros((10, 4), dtype=np.uint32))
    assert ring.put(_rec(4, 1)) == 0
    assert ring.put(_rec(4, 2)) == 4
    assert ring.put(_rec(3, 3)) is None          # 2 slots left at the end, 0 at the start
    ring.release(1)                               # frees slots 0..3
    assert ring.put(_rec(3, 3)) == 0              # skips the 2-slot tail end, wraps
    assert ring.used == 4 + 2 + 3
    assert np.all(ring.view(0, 3)[:, 3] == 3) and np.all(ring.view(4, 4)[:, 3] == 2)
    assert ring.put(_rec(2, 4)) is None
    ring.release(2)
    assert ring.used == 0 and ring.tail == 0
    assert ring.put(_rec(10, 5)) == 0
    with pytest.raises(ValueError):
        ring.put(_rec(11, 6))


def test_contiguous_prefix_stops_at_the_wrap():
    from fkdtree_amd.serving import Request, contiguous_prefix

    def req(start, n):
        r = Request(0, None, None, 0.0, records=_rec(n, 0))
        r.ring_start = start
        return r

    q = [req(0, 3), req(3, 2), req(5, 4), req(0, 1), req(1, 1)]
    assert contiguous_prefix(q, 5) == 3
    assert contiguous_prefix(q, 2) == 2
    assert contiguous_prefix(q[3:], 2) == 2
    assert contiguous_prefix(q, 1) == 1


def test_batches_hold_one_request_kind():
    """A queue mixing SoA requests and FKDPoint-record requests is cut where
    the kind changes, so a batch never mixes them."""
    from fkdtree_amd.serving import Request, contiguous_prefix, same_kind_prefix

    def rec(start, n):
        r = Request(0, None, None, 0.0, records=_rec(n, 0))
        r.ring_start = start
        return r

    soa = reqs([5, 6])
    q = [soa[0], rec(0, 3), soa[1]]
    assert same_kind_prefix(q, 3) == 1
    q = [rec(0, 3), rec(3, 2), soa[0], rec(5, 1)]
    assert same_kind_prefix(q, 4) == 2
    assert contiguous_prefix(q, same_kind_prefix(q, 4)) == 2
    q = soa + reqs([7])
    assert same_kind_prefix(q, 3) == 3


def test_ingest_ring_reserve_then_write():
    """Slots are claimed in arrival order by reserve() and filled later by
    write() (from receiver threads), in any order."""
    from fkdtree_amd.serving import IngestRing

    ring = IngestRing(10, 3, buffer=np.zeros((10, 4), dtype=np.uint32))
    a, b = ring.reserve(4), ring.reserve(5)
    assert (a, b) == (0, 4) and ring.reserve(2) is None
    ring.write(b, _rec(5, 2))
    ring.write(a, _rec(4, 1))
    assert np.all(ring.view(0, 4)[:, 3] == 1) and np.all(ring.view(4, 5)[:, 3] == 2)
    ring.release(1)
    assert ring.reserve(3) == 0            # wraps past the 1-slot tail end
