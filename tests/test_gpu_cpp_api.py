"""The C++ drop-in (FKDTree_HIP<float,3>, include/fkdtree/FKDTree_hip.h) run as
the reference's driver runs its trees, compared bit-exactly with the oracle."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import oracle
from fkdtree_amd import gen
from tests.helpers import adversarial_points, sorted_sets

pytestmark = pytest.mark.gpu
BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "fkdtree_hip_test.bin")


@pytest.mark.parametrize("kind,n,hw", [("mixed", 30_000, 0.01), ("shuffled_ids", 5_000, 0.05), ("zeros", 300, 0.0)])
def test_cpp_wrapper_matches_oracle(require_gpu, kind, n, hw):
    assert os.path.exists(BIN), "tests/cpp/fkdtree_hip_test.bin missing: run __graft_entry__.build()"
    p = adversarial_points(kind, n)
    b = gen.boxes_around(p.coords[:, : min(n, 2000)], hw)
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        with open(fin, "wb") as f:
            f.write(np.array([p.n, b.q], np.uint32).tobytes())
            f.write(np.ascontiguousarray(p.coords.T).tobytes())
            f.write(p.ids.tobytes())
            f.write(np.ascontiguousarray(b.mins.T).tobytes())
            f.write(np.ascontiguousarray(b.maxs.T).tobytes())
        print(f"[cpp] {kind} n={n} q={b.q}", flush=True)
        r = subprocess.run([BIN, fin, fout], capture_output=True, text=True, timeout=120)
        print(r.stdout, r.stderr)
        assert r.returncode == 0
        raw = open(fout, "rb").read()
    tc = np.frombuffer(raw, np.float32, 3 * n).reshape(3, n)
    ti = np.frombuffer(raw, np.uint32, n, 12 * n)
    off = np.frombuffer(raw, np.uint64, b.q + 1, 16 * n)
    hits = np.frombuffer(raw, np.uint32, int(off[-1]), 16 * n + 8 * (b.q + 1))
    oc, oi = oracle.build(p.coords, p.ids)
    assert np.array_equal(tc.view(np.uint32), oc.view(np.uint32)) and np.array_equal(ti, oi)
    ooff, oids = oracle.search(oc, oi, b.mins, b.maxs)
    assert np.array_equal(off, ooff)
    assert np.array_equal(sorted_sets(off, hits), sorted_sets(ooff, oids))
