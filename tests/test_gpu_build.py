"""GPU build parity: heap-ordered tree arrays bit-exact against the oracle
(the restated FKDTree_CPU::build, FKDTree_cpu.h:221-277) through the C ABI."""
import numpy as np
import pytest

import oracle
from fkdtree_amd import FKDEvents, FKDTree, gen
from tests.helpers import KINDS, SMALL_N, adversarial_points, bits

pytestmark = pytest.mark.gpu


def check_tree(p: gen.PointSet, label: str):
    print(f"[build] {label} dim={p.dim} n={p.n}", flush=True)
    t = FKDTree(coords=p.coords, ids=p.ids)
    t.build()
    gc, gi = t.tree_arrays()
    oc, oi = oracle.build(p.coords, p.ids)
    assert np.array_equal(gi, oi), f"{label}: ids differ"
    assert np.array_equal(bits(gc), bits(oc)), f"{label}: coordinates differ"
    t.close()


@pytest.mark.parametrize("kind", KINDS)
def test_small_trees_dim3(require_gpu, kind):
    for n in SMALL_N:
        check_tree(adversarial_points(kind, n), f"{kind}")


@pytest.mark.parametrize("dim", [1, 2, 4])
def test_small_trees_other_dims(require_gpu, dim):
    for kind in ("uniform", "grid", "shuffled_ids", "zeros"):
        for n in (1, 2, 3, 17, 1000, 4097, 20000):
            check_tree(adversarial_points(kind, n, dim=dim), f"{kind}")


def test_empty_tree(require_gpu):
    t = FKDTree(coords=np.zeros((3, 0), np.float32))
    t.build()
    c, i = t.tree_arrays()
    assert c.shape == (3, 0) and i.shape == (0,)


def test_uniform_100k(require_gpu):
    p, _ = gen.workload_uniform_100k()
    check_tree(p, "uniform_100k")


def test_mixed_1m(require_gpu):
    p, _ = gen.workload_mixed_1m()
    check_tree(p, "mixed_1m")


def test_rebuild_is_idempotent(require_gpu):
    p = adversarial_points("mixed", 50_000)
    t = FKDTree(coords=p.coords, ids=p.ids)
    t.build()
    a = [x.copy() for x in t.tree_arrays()]
    t.build()
    b = t.tree_arrays()
    assert np.array_equal(bits(a[0]), bits(b[0])) and np.array_equal(a[1], b[1])


def test_nan_rejected(require_gpu):
    p = adversarial_points("uniform", 100)
    c = p.coords.copy()
    c[1, 17] = np.nan
    with pytest.raises(RuntimeError, match="NaN"):
        FKDTree(coords=c, ids=p.ids)


def test_events_small(require_gpu):
    sizes = [0, 1, 2, 3, 5, 0, 17, 1000, 4096, 4097, 9000, 1, 2]
    evs = [adversarial_points("mixed" if i % 2 else "shuffled_ids", s, seed=100 + i) if s else
           gen.PointSet(np.zeros((3, 0), np.float32), np.zeros(0, np.uint32)) for i, s in enumerate(sizes)]
    b = gen.concat_events(evs)
    print(f"[build] events sizes={sizes}", flush=True)
    t = FKDEvents(b.points.coords, b.points.ids, b.offsets)
    t.build()
    gc, gi = t.tree_arrays()
    oc, oi = oracle.build_events(b.points.coords, b.points.ids, b.offsets)
    assert np.array_equal(gi, oi)
    assert np.array_equal(bits(gc), bits(oc))


def test_events_scaled_batch(require_gpu):
    b = gen.workload_events_2048(n_events=256, scale=0.1)
    print(f"[build] events n_events=256 total={b.points.n}", flush=True)
    t = FKDEvents(b.points.coords, b.points.ids, b.offsets)
    t.build()
    gc, gi = t.tree_arrays()
    oc, oi = oracle.build_events(b.points.coords, b.points.ids, b.offsets)
    assert np.array_equal(gi, oi)
    assert np.array_equal(bits(gc), bits(oc))
