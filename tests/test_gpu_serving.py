"""Continuous-batching server: every event's result is the same alone and in
the mix (exact id sets), batches overlap on the device, HBM admission and
eviction work, and resident trees answer later queries."""
import numpy as np
import pytest

import oracle
from fkdtree_amd import FKDTree, gen
from fkdtree_amd.serving import EventServer, poisson_requests
from tests.helpers import sorted_sets

pytestmark = pytest.mark.gpu


def alone(req, hw=0.01):
    t = FKDTree(coords=req.coords, ids=req.ids)
    t.build()
    b = gen.boxes_around(req.coords, hw)
    off, ids = t.search_in_the_box_multiple_csr(b.mins, b.maxs)
    t.close()
    return off, ids


def test_mix_equals_alone(require_gpu):
    reqs = poisson_requests(60, rate_hz=20_000.0, size_scale=0.02)
    print(f"[serve] events=60 points={sum(r.n for r in reqs)}", flush=True)
    srv = EventServer(n_streams=3, max_batch_points=200_000, max_batch_events=16, keep_results=True)
    rep = srv.run(reqs)
    s = rep.summary()
    print(s, flush=True)
    assert len(rep.results) == 60
    assert rep.n_batches >= 4 and max(rep.batch_sizes) > 1
    for r in reqs:
        got = rep.results[r.event]
        off, ids = alone(r)
        assert np.array_equal(got.offsets, off), r.event
        assert np.array_equal(sorted_sets(got.offsets, got.ids), sorted_sets(off, ids)), r.event
    for r in reqs[:5]:
        tc, ti = oracle.build(r.coords, r.ids)
        b = gen.boxes_around(r.coords, 0.01)
        ooff, oids = oracle.search(tc, ti, b.mins, b.maxs)
        got = rep.results[r.event]
        assert np.array_equal(sorted_sets(got.offsets, got.ids), sorted_sets(ooff, oids))
    # resident trees answer new queries
    r = reqs[7]
    b = gen.boxes_around(r.coords[:, :50], 0.05)
    qoff, qids = srv.query(r.event, b.mins, b.maxs)
    t = FKDTree(coords=r.coords, ids=r.ids)
    t.build()
    aoff, aids = t.search_in_the_box_multiple_csr(b.mins, b.maxs)
    assert np.array_equal(sorted_sets(qoff, qids), sorted_sets(aoff, aids))
    srv.close()


def test_tiny_budget_forces_single_event_batches_and_eviction(require_gpu):
    reqs = poisson_requests(12, rate_hz=50_000.0, size_scale=0.01)
    biggest = max(r.n for r in reqs)
    from fkdtree_amd.serving import batch_bytes
    budget = batch_bytes(biggest, 64) + (1 << 20)
    srv = EventServer(n_streams=2, hbm_budget=budget, hits_per_point=64, keep_results=True)
    rep = srv.run(reqs)
    assert len(rep.results) == 12
    assert rep.n_batches >= 3          # the budget splits the stream
    assert rep.evictions > 0           # resident trees made room for new batches
    for r in reqs[:4]:
        off, ids = alone(r)
        got = rep.results[r.event]
        assert np.array_equal(sorted_sets(got.offsets, got.ids), sorted_sets(off, ids))
    srv.close()


def test_boxes_around_matches_host(require_gpu):
    """fkd_boxes_around: the same float32 boxes as the host formula
    (main.cpp:177-186 forms them around every point) and each box's event."""
    import fkdtree_amd as F
    sizes = [3, 0, 1000, 1, 4500]
    evs = [gen.serving_event(s, 40 + i) if s else gen.PointSet(np.zeros((3, 0), np.float32), np.zeros(0, np.uint32))
           for i, s in enumerate(sizes)]
    b = gen.concat_events(evs)
    s = F.Stream()
    t = F.FKDEvents(b.points.coords, b.points.ids, b.offsets, stream=s)
    P = b.points.n
    dmi = F.DeviceArray((3, P), np.float32, stream=s)
    dma = F.DeviceArray((3, P), np.float32, stream=s)
    dq = F.DeviceArray(P, np.uint32, stream=s)
    t.boxes_around(0.01, dmi.ptr, dma.ptr, dq.ptr)
    want = gen.boxes_around(b.points.coords, 0.01)
    assert np.array_equal(dmi.download(s).view(np.uint32), want.mins.view(np.uint32))
    assert np.array_equal(dma.download(s).view(np.uint32), want.maxs.view(np.uint32))
    assert np.array_equal(dq.download(s), np.repeat(np.arange(len(sizes), dtype=np.uint32),
                                                      np.diff(b.offsets).astype(np.int64)))
    t.close()


def test_ring_ingestion_equals_soa(require_gpu):
    """Events arriving as FKDPoint records through the page-locked ingest
    ring (small enough to wrap several times and to fill) give the same
    per-event results as events given as SoA arrays."""
    soa = poisson_requests(40, rate_hz=50_000.0, size_scale=0.02)
    rec = poisson_requests(40, rate_hz=50_000.0, size_scale=0.02, records=True)
    ring = 3 * max(r.n for r in rec)
    print(f"[serve] ring events=40 points={sum(r.n for r in rec)} ring={ring}", flush=True)
    a = EventServer(n_streams=2, max_batch_points=150_000, max_batch_events=16, keep_results=True)
    ra = a.run(soa)
    a.close()
    b = EventServer(n_streams=2, max_batch_points=150_000, max_batch_events=16, keep_results=True,
                    ring_points=ring)
    rb = b.run(rec)
    assert rb.n_batches >= 4
    for r in soa:
        x, y = ra.results[r.event], rb.results[r.event]
        assert np.array_equal(x.offsets, y.offsets), r.event
        assert np.array_equal(sorted_sets(x.offsets, x.ids), sorted_sets(y.offsets, y.ids)), r.event
    # ring-ingested results against the oracle (FKDTree_cpu.h:123-169) for a
    # sample of events, the largest included
    sample = sorted(soa, key=lambda r: -r.n)[:3] + soa[::9]
    for r in sample:
        tc, ti = oracle.build(r.coords, r.ids)
        bx = gen.boxes_around(r.coords, 0.01)
        ooff, oids = oracle.search(tc, ti, bx.mins, bx.maxs)
        y = rb.results[r.event]
        assert np.array_equal(y.offsets, ooff), r.event
        assert np.array_equal(sorted_sets(y.offsets, y.ids), sorted_sets(ooff, oids)), r.event
    assert b.ring.used == 0
    b.close()
    # without keeping results only the event boundaries come back; totals agree
    c = EventServer(n_streams=2, max_batch_points=150_000, max_batch_events=16, ring_points=ring)
    rc = c.run(rec)
    for r in soa:
        assert rc.results[r.event].hits == int(ra.results[r.event].offsets[-1])
    c.close()


def test_serving_full_size_events(require_gpu):
    """The serving workload at its stated sizes: 48 events of 1,000-500,000
    points (log-uniform, the uniform-plus-blob mix) arriving at 5,000 events/s
    as FKDPoint records through the ingest ring.  Every event's result in the
    mix equals its result alone, and eight events (the four largest and four
    others) equal the oracle's exact id sets."""
    soa = poisson_requests(48, rate_hz=5000.0)
    rec = poisson_requests(48, rate_hz=5000.0, records=True)
    pts = sum(r.n for r in rec)
    print(f"[serve-full] events=48 points={pts} largest={max(r.n for r in rec)}", flush=True)
    assert min(r.n for r in rec) >= 1000 and max(r.n for r in rec) <= 500_000
    with EventServer(n_streams=3, max_batch_points=4_000_000, max_batch_events=256, keep_results=True,
                     ring_points=max(16_000_000, 2 * max(r.n for r in rec))) as srv:
        rep = srv.run(rec)
        print(rep.summary(), flush=True)
        assert len(rep.results) == 48
        assert max(rep.batch_sizes) > 1           # events shared batches
        for r in soa:
            got = rep.results[r.event]
            off, ids = alone(r)
            assert np.array_equal(got.offsets, off), r.event
            assert np.array_equal(sorted_sets(got.offsets, got.ids), sorted_sets(off, ids)), r.event
        by_size = sorted(soa, key=lambda r: -r.n)
        for r in by_size[:4] + by_size[10::10][:4]:
            tc, ti = oracle.build(r.coords, r.ids)
            b = gen.boxes_around(r.coords, 0.01)
            ooff, oids = oracle.search(tc, ti, b.mins, b.maxs)
            got = rep.results[r.event]
            assert np.array_equal(got.offsets, ooff), r.event
            assert np.array_equal(sorted_sets(got.offsets, got.ids), sorted_sets(ooff, oids)), r.event
        print(f"[serve-full] hits={sum(r.hits for r in rep.results.values())}", flush=True)


def test_servers_in_turn_release_resources(require_gpu):
    """close() destroys a server's streams and hands its page-locked ingest
    ring to the next server: forty servers created and closed in turn (each
    with the 256 MB ring the benchmark uses) share one ring buffer and serve
    at the same rate at the end as at the start."""
    from fkdtree_amd import serving
    import time

    rec = poisson_requests(24, rate_hz=20_000.0, size_scale=0.05, records=True)
    rates = []
    for k in range(40):
        with EventServer(n_streams=3, max_batch_points=400_000, max_batch_events=64,
                         ring_points=16_000_000) as srv:
            t0 = time.perf_counter()
            rep = srv.run(rec)
            dt = time.perf_counter() - t0
        assert len(rep.results) == 24
        rates.append(24 / dt)
        if k % 10 == 9:
            print(f"[servers] {k + 1} servers: last events/s {rates[-1]:.0f}", flush=True)
    assert len(serving._FREE_RINGS[(16_000_000, 3)]) == 1   # one buffer, reused by every server
    first = float(np.median(rates[1:6]))
    last = float(np.median(rates[-5:]))
    print(f"[servers] first {first:.0f} last {last:.0f} events/s", flush=True)
    assert last * 1.5 >= first, (first, last, rates)
