"""Host-side admission policy of the event server (no GPU)."""
import numpy as np

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
