"""The per-box search path: fkd_search_small (one wave per box, breadth
first, results in page-locked host memory), the resident per-box search
kernel (fkd_box_server_*) and FKDTree_HIP's search_in_the_box from many
host threads (main.cpp:351-359's pattern) through it.  Results
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
from fkdtree_amd import FKDEvents, FKDTree, gen
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


def _server_check(t, srv, mins, maxs, ooff, oids, ovis, cap=None):
    """Every box through the resident kernel against the oracle: counts whole
    (or flagged only where the oracle's walk visited more than 2,048 nodes),
    id sets exact where they fit."""
    cap = srv.slot_ids if cap is None else cap
    for q in range(mins.shape[1]):
        c, ids = srv.search_raw(mins[:, q], maxs[:, q], cap)
        if c == OVERFLOW:
            assert ovis[q] > 2048, (q, ovis[q])
            continue
        want = np.sort(oids[int(ooff[q]):int(ooff[q + 1])])
        assert c == want.size, (q, c, want.size)
        assert ids.size == min(c, cap, srv.slot_ids)
        if c <= min(cap, srv.slot_ids):
            assert np.array_equal(np.sort(ids), want), q
        else:  # cut off: the ids given are some of the box's
            assert np.isin(ids, want).all(), q


@pytest.mark.parametrize("kind,n,dim", [("mixed", 100_000, 3), ("zeros", 3000, 3), ("grid", 5000, 2),
                                        ("uniform", 20_000, 4), ("equal", 700, 1), ("uniform", 1, 3),
                                        ("shuffled_ids", 4097, 3)])
def test_box_server_matches_oracle(require_gpu, kind, n, dim):
    """fkd_box_server_search, box after box, against the oracle
    (FKDTree_cpu.h:123-169): ordinary, empty, degenerate, whole-space and NaN
    boxes; results cut off at a 64-id slot and at a smaller caller buffer."""
    p = adversarial_points(kind, n, dim=dim)
    k = min(n, 200)
    b = gen.boxes_around(p.coords[:, :k], 0.02)
    sp = special_boxes(p)
    mins = np.ascontiguousarray(np.concatenate([b.mins, sp.mins], 1))
    maxs = np.ascontiguousarray(np.concatenate([b.maxs, sp.maxs], 1))
    print(f"[server] {kind} n={n} dim={dim} boxes={mins.shape[1]}", flush=True)
    t = FKDTree(coords=p.coords, ids=p.ids)
    t.build()
    oc, oi = oracle.build(p.coords, p.ids)
    ooff, oids, ovis = oracle.search(oc, oi, mins, maxs, with_visited=True)
    srv = t.box_server(slots=8, slot_ids=4096)
    _server_check(t, srv, mins, maxs, ooff, oids, ovis)
    _server_check(t, srv, mins, maxs, ooff, oids, ovis, cap=5)
    small = t.box_server(slots=2, slot_ids=64)
    _server_check(t, small, mins, maxs, ooff, oids, ovis)
    st = srv.stats()
    assert st["requests"] == 2 * mins.shape[1] and st["launches"] >= 1
    small.close()
    srv.close()
    t.close()


def test_box_server_threads_share_few_slots(require_gpu):
    """Eight Python threads (ctypes drops the GIL) on a server of three slots:
    callers wait for a free slot, and every answer is the oracle's."""
    import threading

    pts, boxes = gen.workload_mixed_1m(n=100_000, q=4_000)
    oc, oi = oracle.build(pts.coords, pts.ids)
    ooff, oids = oracle.search(oc, oi, boxes.mins, boxes.maxs)
    t = FKDTree(coords=pts.coords, ids=pts.ids)
    t.build()
    srv = t.box_server(slots=3, slot_ids=4096)
    got = [None] * boxes.q
    errors = []

    def work(w):
        try:
            for q in range(w, boxes.q, 8):
                got[q] = srv.search(boxes.mins[:, q], boxes.maxs[:, q])
        except Exception as e:  # reported below
            errors.append(repr(e))

    print(f"[server] threads=8 slots=3 q={boxes.q}", flush=True)
    th = [threading.Thread(target=work, args=(w,)) for w in range(8)]
    for x in th:
        x.start()
    for x in th:
        x.join(120)
    assert not errors, errors[:3]
    for q in range(boxes.q):
        assert got[q] is not None, q
        assert np.array_equal(np.sort(got[q]), np.sort(oids[int(ooff[q]):int(ooff[q + 1])])), q
    assert srv.stats()["requests"] == boxes.q
    srv.close()
    t.close()


def test_box_server_leaves_when_idle_and_restarts(require_gpu):
    """The kernel leaves after its idle time; the next call starts it again
    and is answered; destroy waits for a live kernel to leave."""
    import time

    p = adversarial_points("mixed", 20_000)
    t = FKDTree(coords=p.coords, ids=p.ids)
    t.build()
    oc, oi = oracle.build(p.coords, p.ids)
    b = gen.boxes_around(p.coords[:, :4], 0.05)
    ooff, oids = oracle.search(oc, oi, b.mins, b.maxs)
    srv = t.box_server(slots=4, slot_ids=4096, idle_us=2000)
    print("[server] idle restart", flush=True)
    for rnd in range(3):
        for q in range(4):
            ids = srv.search(b.mins[:, q], b.maxs[:, q])
            assert np.array_equal(np.sort(ids), np.sort(oids[int(ooff[q]):int(ooff[q + 1])])), (rnd, q)
        deadline = time.time() + 5.0
        while srv.stats()["running"] and time.time() < deadline:
            time.sleep(0.002)
        assert not srv.stats()["running"]
    assert srv.stats()["launches"] == 3
    ids = srv.search(b.mins[:, 0], b.maxs[:, 0])  # live again
    assert srv.stats()["running"]
    srv.close()  # stops the live kernel
    t.close()


def test_box_server_lifecycle(require_gpu):
    """A server needs a built single tree; a build is refused while one
    exists; closing the tree stops its servers, whose calls then fail; the
    tree's own search_in_the_box keeps working across builds."""
    p = adversarial_points("uniform", 5000)
    t = FKDTree(coords=p.coords, ids=p.ids)
    h = C.c_void_p()
    assert lib().fkd_box_server_create(t._h, 4, 64, 1000, C.byref(h)) == -4  # not built
    t.build()
    assert lib().fkd_box_server_create(t._h, 0, 64, 1000, C.byref(h)) == -1
    assert lib().fkd_box_server_create(t._h, 65, 64, 1000, C.byref(h)) == -1
    assert lib().fkd_box_server_create(t._h, 4, 0, 1000, C.byref(h)) == -1
    assert lib().fkd_box_server_create(t._h, 4, 64, 0, C.byref(h)) == -1
    srv = t.box_server(slots=4, slot_ids=64)
    assert lib().fkd_build(t._h, None) == -4
    srv.close()
    oc, oi = oracle.build(p.coords, p.ids)
    lo, hi = np.full(3, 0.2, np.float32), np.full(3, 0.4, np.float32)
    ooff, oids = oracle.search(oc, oi, lo.reshape(3, 1), hi.reshape(3, 1))
    for _ in range(2):
        t.build()  # drops the tree's own server
        assert np.array_equal(np.sort(t.search_in_the_box(lo, hi)), np.sort(oids))
    srv = t.box_server(slots=4, slot_ids=64)
    c, _ = srv.search_raw(lo, hi)
    assert c == oids.size
    t.close()  # stops srv's kernel
    box = np.concatenate([lo, hi])
    cnt = C.c_uint32()
    assert lib().fkd_box_server_search(srv._h, box.ctypes.data, None, 0, C.byref(cnt)) == -4
    srv.close()
    ev = gen.concat_events([gen.mixed_event(500, 7, e) for e in range(3)])
    e = FKDEvents(ev.points.coords, ev.points.ids, ev.offsets)
    e.build()
    assert lib().fkd_box_server_create(e._h, 4, 64, 1000, C.byref(h)) == -1
    e.close()


def test_box_server_beside_batched_work(require_gpu):
    """Per-box calls from four threads on one tree while another tree is built
    and searched in batches on its own stream: the resident kernel's compute
    units and the build's look-back launches share the GPU, and every result
    of both is the oracle's."""
    import threading

    from fkdtree_amd import Stream

    pa, ba = gen.workload_mixed_1m(n=100_000, q=2_000)
    pb, bb = gen.workload_mixed_1m(n=300_000, q=50_000)
    oca, oia = oracle.build(pa.coords, pa.ids)
    offa, oidsa = oracle.search(oca, oia, ba.mins, ba.maxs)
    ocb, oib = oracle.build(pb.coords, pb.ids)
    offb, oidsb = oracle.search(ocb, oib, bb.mins, bb.maxs)
    ta = FKDTree(coords=pa.coords, ids=pa.ids)
    ta.build()
    srv = ta.box_server(slots=4)
    st = Stream()
    tb = FKDTree(coords=pb.coords, ids=pb.ids, stream=st)
    got = [None] * ba.q
    errors = []

    def work(w):
        try:
            for q in range(w, ba.q, 4):
                got[q] = srv.search(ba.mins[:, q], ba.maxs[:, q])
        except Exception as e:  # reported below
            errors.append(repr(e))

    print("[server] beside batched builds and searches", flush=True)
    th = [threading.Thread(target=work, args=(w,)) for w in range(4)]
    for x in th:
        x.start()
    rounds = 0
    while rounds < 2 or (any(x.is_alive() for x in th) and rounds < 20):
        tb.build()
        c, i = tb.tree_arrays()
        assert np.array_equal(i, oib) and np.array_equal(c.view(np.uint32), ocb.view(np.uint32)), rounds
        off, ids = tb.search_in_the_box_multiple_csr(bb.mins, bb.maxs)
        assert np.array_equal(off, offb), rounds
        rounds += 1
    for x in th:
        x.join(120)
    assert not errors, errors[:3]
    for q in range(ba.q):
        assert got[q] is not None, q
        assert np.array_equal(np.sort(got[q]), np.sort(oidsa[int(offa[q]):int(offa[q + 1])])), q
    assert np.array_equal(sorted_sets(off, ids), sorted_sets(offb, oidsb))
    print(f"[server] {rounds} batched rounds beside {ba.q} per-box calls", flush=True)
    srv.close()
    ta.close()
    tb.close()
    st.close()
