"""Tuning table lookup (host logic only): nearest size bucket per
architecture and dimension."""
import json

from fkdtree_amd import tune


def test_lookup_nearest_bucket(tmp_path):
    table = {"gfx950": {"dim3_log2n10": {"params": {"query_key_bits": 5}},
                        "dim3_log2n20": {"params": {"query_key_bits": 8, "pool_batch": 32}},
                        "dim2_log2n20": {"params": {"query_key_bits": 10}}}}
    assert tune.lookup("gfx950", 3, 1 << 19, table) == {"query_key_bits": 8, "pool_batch": 32}
    assert tune.lookup("gfx950", 3, 900, table) == {"query_key_bits": 5}
    assert tune.lookup("gfx950", 2, 10, table) == {"query_key_bits": 10}
    assert tune.lookup("gfx950", 4, 1 << 20, table) is None
    assert tune.lookup("gfx942", 3, 1 << 20, table) is None
    p = tmp_path / "t.json"
    p.write_text(json.dumps(table))
    assert tune.load_table(str(p)) == table
    assert tune.load_table(str(tmp_path / "missing.json")) == {}


def test_keys():
    assert tune._key(3, 1_000_000) == "dim3_log2n20"
    assert tune._key(3, 1) == "dim3_log2n1"


def test_candidates_are_valid():
    # query_key_bits * dim must fit a 32-bit key at dim 3; pool batch 1..1024
    assert all(0 <= kb <= 10 for kb in tune.CANDIDATES["query_key_bits"])
    assert all(1 <= b <= 1024 for b in tune.POOL_CANDIDATES["pool_batch"])
