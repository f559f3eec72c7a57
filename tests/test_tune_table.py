"""Tuning table lookup (host logic only): nearest size bucket per
architecture and dimension."""
import json

from fkdtree_amd import tune


def test_lookup_nearest_bucket(tmp_path):
    table = {"gfx950": {"dim3_log2n10": {"params": {"query_key_bits": 5}},
                        "dim3_log2n20": {"params": {"query_key_bits": 8, "fill_order": 1}},
                        "dim2_log2n20": {"params": {"query_key_bits": 10}}}}
    assert tune.lookup("gfx950", 3, 1 << 19, table) == {"query_key_bits": 8, "fill_order": 1}
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
    assert tune._key(3, 41_000, events=True) == "dim3_ev_log2n15"


def test_event_entries_are_kept_apart():
    """Handles of several events read only the event entries, single trees
    only the others (their best launch parameters differ)."""
    table = {"gfx950": {"dim3_log2n20": {"params": {"fill_window": 4096}},
                        "dim3_ev_log2n15": {"params": {"fill_window": 0}}}}
    assert tune.lookup("gfx950", 3, 1 << 15, table) == {"fill_window": 4096}
    assert tune.lookup("gfx950", 3, 1 << 15, table, events=True) == {"fill_window": 0}
    assert tune.lookup("gfx950", 3, 1 << 22, table, events=True) == {"fill_window": 0}
    assert tune.lookup("gfx950", 2, 1 << 15, table, events=True) is None


def test_candidates_are_valid():
    # query_key_bits * dim must fit a 32-bit key at dim 3; fill_order is 0/1
    assert all(0 <= kb <= 10 for kb in tune.CANDIDATES["query_key_bits"])
    assert set(tune.CANDIDATES["fill_order"]) <= {0, 1}
    assert set(tune.CANDIDATES["fill_window"]) <= {0, 256, 1024, 4096}
    assert set(tune.CANDIDATES["count_window"]) <= {0, 256, 1024, 4096}
    assert all(0 <= v <= 6 for v in tune.CANDIDATES["key_volume_bits"])
    # build geometry: finisher capacities the library instantiates, digit-pass tiles
    assert set(tune.BUILD_CANDIDATES["fin_log"]) <= {9, 10, 11}
    assert set(tune.BUILD_CANDIDATES["build_radix_items"]) <= {8, 16}
    assert set(tune.BUILD_CANDIDATES["level_pairs"]) <= {0, 1}
    # traversal schedules: walk, parked (threshold in eighths) or queued buckets
    for b in tune.BUCKET_CANDIDATES:
        assert b["bucket"] in (0, 1, 2) and b.get("count_bucket", 0) in (0, 1, 2)
        assert 0 <= b.get("bucket_park", 8) <= 8


class _FakeTree:
    dim, n, n_events = 3, 1 << 20, 1

    def __init__(self):
        self.set = {}

    def set_tuning(self, **kw):
        for k, v in kw.items():
            if k != "query_key_bits":
                raise RuntimeError("unknown tuning key " + k)
            self.set[k] = v


def test_apply_skips_unknown_keys(tmp_path, monkeypatch):
    table = {"gfx950": {"dim3_log2n20": {"params": {"query_key_bits": 7, "retired_key": 3}}}}
    monkeypatch.setattr(tune, "load_table", lambda path=tune.TABLE: table)
    t = _FakeTree()
    assert tune.apply(t, arch="gfx950") == {"query_key_bits": 7}
    assert t.set == {"query_key_bits": 7}


def test_table_is_reread_when_the_file_changes(tmp_path):
    """The table is parsed once per process but a rewritten file (a new
    autotuning run) is picked up."""
    import os
    p = tmp_path / "t.json"
    p.write_text(json.dumps({"gfx950": {"dim3_log2n20": {"params": {"query_key_bits": 5}}}}))
    assert tune.lookup("gfx950", 3, 1 << 20, tune.load_table(str(p))) == {"query_key_bits": 5}
    assert tune.load_table(str(p)) is tune.load_table(str(p))
    p.write_text(json.dumps({"gfx950": {"dim3_log2n20": {"params": {"query_key_bits": 7}}}}))
    st = os.stat(p)
    os.utime(p, ns=(st.st_atime_ns, st.st_mtime_ns + 1_000_000))
    assert tune.lookup("gfx950", 3, 1 << 20, tune.load_table(str(p))) == {"query_key_bits": 7}


def test_library_accepts_every_candidate_key():
    """Every key the tuner sweeps is one the library's fkd_set_tuning knows:
    its error message for an unknown key differs from a bad value's (no GPU
    call is needed to reject a NULL handle, so the names are checked by
    the header instead)."""
    import os
    hdr = open(os.path.join(os.path.dirname(tune.HERE), "include", "fkdtree", "fkd_capi.h")).read()
    for k in list(tune.CANDIDATES) + list(tune.BUILD_CANDIDATES):
        assert f'"{k}"' in hdr, k


def test_every_search_setting_names_every_swept_key():
    """set_tuning leaves keys a setting does not name as they were, so each
    setting of both search stages must name every key either stage sweeps:
    otherwise a setting is timed with its predecessor's traversal schedule
    and stored without it."""
    stage1 = tune.search_stage1()
    keys = set(tune.CANDIDATES) | {k for b in tune.BUCKET_CANDIDATES for k in b}
    assert len(stage1) == int(__import__("math").prod(len(v) for v in tune.CANDIDATES.values()))
    assert all(set(s) == keys for s in stage1)
    stage2 = tune.search_stage2(stage1[0])
    assert len(stage2) == len(tune.BUCKET_CANDIDATES)
    assert all(set(s) == keys for s in stage2)
    assert {s["bucket"] for s in stage2} == {b["bucket"] for b in tune.BUCKET_CANDIDATES}
