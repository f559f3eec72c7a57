"""The per-box search path: fkd_search_small (one wave per box, breadth
first, results in page-locked host memory) and FKDTree_HIP's coalescing of
concurrent search_in_the_box calls (main.cpp:351-359's pattern).  Results
are the oracle's (FKDTree_cpu.h:123-169): exact id sets; counts above the
per-box slot are reported whole; walks wider than the frontier limit are
flagged and, in the C++ drop-in, searched again through fkd_search."""
import ctypes as C
import os
import subprocess
import tempfile

import numpy as np
import pytest

import oracle
from fkdtree_amd import FKDTree, gen
from fkdtree_amd._lib import lib
from tests.helpers import adversarial_points, special_boxes, sorted_sets

pytestmark = pytest.mark.gpu
PERBOX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "fkdtree_perbox.bin")
OVERFLOW = 0xFFFFFFFF


def _small(t, mins, maxs, slot):
    nb = mins.shape[1]
    boxes = np.ascontiguousarray(np.concatenate([mins.T, maxs.T], axis=1).astype(np.float32))
    counts = np.zeros(nb, np.uint32)
    ids = np.zeros(max(nb * slot, 1), np.uint32)
    rc = lib().fkd_search_small(t._h, nb, boxes.ctypes.data, slot, counts.ctypes.data, ids.ctypes.data, None)
    assert rc == 0, lib().fkd_last_error()
    return counts, ids.reshape(nb, slot) if nb else ids


@pytest.mark.parametrize("kind,n,dim", [("mixed", 100_000, 3), ("zeros", 3000, 3), ("grid", 5000, 2),
                                        ("uniform", 20_000, 4), ("equal", 700, 1), ("uniform", 1, 3)])
def test_small_search_matches_oracle(require_gpu, kind, n, dim):
    p = adversarial_points(kind, n, dim=dim)
    k = min(n, 300)
    b = gen.boxes_around(p.coords[:, :k], 0.02)
    sp = special_boxes(p)
    mins = np.ascontiguousarray(np.concatenate([b.mins, sp.mins], 1))
    maxs = np.ascontiguousarray(np.concatenate([b.maxs, sp.maxs], 1))
    print(f"[small] {kind} n={n} dim={dim} boxes={mins.shape[1]}", flush=True)
    t = FKDTree(coords=p.coords, ids=p.ids)
    t.build()
    oc, oi = oracle.build(p.coords, p.ids)
    ooff, oids, ovis = oracle.search(oc, oi, mins, maxs, with_visited=True)
    ocnt = np.diff(ooff)
    slot = 64
    counts, ids = _small(t, mins, maxs, slot)
    for q in range(mins.shape[1]):
        c = int(counts[q])
        if c == OVERFLOW:       # one depth of the walk held more than 2,048 nodes
            assert ovis[q] > 2048, (q, ovis[q])
            continue
        assert c == ocnt[q], (q, c, ocnt[q])
        if c <= slot:
            assert np.array_equal(np.sort(ids[q, :c]), np.sort(oids[int(ooff[q]):int(ooff[q + 1])])), q
    # a slot large enough for every box that is not flagged: all sets exact
    big = int(ocnt[counts != OVERFLOW].max(initial=1))
    counts2, ids2 = _small(t, mins, maxs, max(big, 1))
    for q in range(mins.shape[1]):
        if counts2[q] == OVERFLOW:
            continue
        c = int(counts2[q])
        assert np.array_equal(np.sort(ids2[q, :c]), np.sort(oids[int(ooff[q]):int(ooff[q + 1])])), q
    t.close()


def test_small_search_limits(require_gpu):
    p = adversarial_points("mixed", 1000)
    t = FKDTree(coords=p.coords, ids=p.ids)
    box = np.zeros(6, np.float32)
    cnt = np.zeros(1, np.uint32)
    ids = np.zeros(4, np.uint32)
    assert lib().fkd_search_small(t._h, 1, box.ctypes.data, 4, cnt.ctypes.data, ids.ctypes.data, None) == -4
    t.build()
    assert lib().fkd_search_small(t._h, 1, box.ctypes.data, 0, cnt.ctypes.data, ids.ctypes.data, None) == -1
    assert lib().fkd_search_small(t._h, 65537, box.ctypes.data, 1, cnt.ctypes.data, ids.ctypes.data, None) == -1
    assert lib().fkd_search_small(t._h, 0, None, 4, None, None, None) == 0
    t.close()


def write_perbox_input(path, pts, boxes, counts, k, off, sorted_ids):
    with open(path, "wb") as f:
        f.write(np.array([pts.n, boxes.q], np.uint32).tobytes())
        f.write(np.ascontiguousarray(pts.coords.T).tobytes())
        f.write(pts.ids.tobytes())
        f.write(np.ascontiguousarray(boxes.mins.T).tobytes())
        f.write(np.ascontiguousarray(boxes.maxs.T).tobytes())
        f.write(np.ascontiguousarray(counts, dtype=np.uint32).tobytes())
        f.write(np.array([k], np.uint32).tobytes())
        f.write(np.ascontiguousarray(off[:k + 1], dtype=np.uint64).tobytes())
        f.write(np.ascontiguousarray(sorted_ids[:int(off[k])], dtype=np.uint32).tobytes())


@pytest.mark.parametrize("threads", [1, 16])
def test_cpp_per_box_threads_match_oracle(require_gpu, threads):
    """16 host threads calling search_in_the_box box after box (coalesced into
    shared launches): every count and the first 5,000 id sets are the
    oracle's."""
    assert os.path.exists(PERBOX), "tests/cpp/fkdtree_perbox.bin missing: run __graft_entry__.build()"
    pts, boxes = gen.workload_mixed_1m(n=200_000, q=40_000)
    oc, oi = oracle.build(pts.coords, pts.ids)
    ooff, oids = oracle.search(oc, oi, boxes.mins, boxes.maxs)
    k = 5000
    with tempfile.TemporaryDirectory() as d:
        fin = os.path.join(d, "in.bin")
        write_perbox_input(fin, pts, boxes, np.diff(ooff), k, ooff, sorted_sets(ooff, oids))
        print(f"[perbox] threads={threads} n={pts.n} q={boxes.q}", flush=True)
        r = subprocess.run([PERBOX, fin, str(threads)], capture_output=True, text=True, timeout=240)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    import json
    res = json.loads(r.stdout.strip().splitlines()[-1])
    assert res["count_mismatches"] == 0 and res["set_mismatches"] == 0 and res["checked_sets"] == k
    assert res["hits"] == int(ooff[-1])
