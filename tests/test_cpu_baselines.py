"""CPU baseline legs of bench.py (host only): the CMSSW side comparator
(main.cpp:560-629) finds the closed-box hits on continuous data, and the
baseline dictionary carries every line it promises."""
import numpy as np
import pytest

import oracle
from fkdtree_amd import gen


@pytest.mark.skipif(not oracle.cmssw_available(), reason="oracle/_ref not built (reference sources absent)")
def test_cmssw_side_comparator_counts():
    p = gen.PointSet(gen.uniform_f32(gen.workload_seed("uniform_100k"), 7, 3 * 5000).reshape(3, 5000),
                     np.arange(5000, dtype=np.uint32))
    b = gen.Boxes(p.coords[:, :500] - np.float32(0.05), p.coords[:, :500] + np.float32(0.05))
    _, _, found = oracle.cmssw_build_search(p.coords, b.mins, b.maxs)
    off, _ = oracle.brute(p.coords, p.ids, b.mins, b.maxs)
    assert found == int(off[-1])


def test_cpu_baseline_fields():
    import bench
    n = 20000
    p = gen.PointSet(gen.uniform_f32(gen.workload_seed("uniform_100k"), 8, 3 * n).reshape(3, n),
                     np.arange(n, dtype=np.uint32))
    b = gen.Boxes(p.coords - np.float32(0.02), p.coords + np.float32(0.02))
    r = bench.cpu_baseline(p, b, 2, 0)
    assert r["unit"] == "box queries/s" and r["value"] > 0 and r["cores"] == 2
    assert r["kind"] in ("reference", "port")
    assert r["search_1thread"]["cores"] == 1
    off, _ = oracle.brute(p.coords, p.ids, b.mins, b.maxs)
    assert r["hits"] == int(off[-1])
    if oracle.cmssw_available():
        assert r["side_comparator"]["found"] == r["hits"]
