"""CPU checks of the C ABI: the library loads, exports every function that
include/fkdtree/fkd_capi.h declares, and its host-only entry points behave."""
import ctypes as C
import os
import re

import pytest

from fkdtree_amd import _lib

HDR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "fkdtree", "fkd_capi.h")


def declared():
    txt = open(HDR).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\*?\s+\**(fkd_\w+)\(", txt, re.M)))


def test_header_declares_entry_points():
    names = declared()
    assert "fkd_create" in names and "fkd_search" in names and "fkd_build" in names
    assert len(names) >= 25


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    missing = [n for n in declared() if not hasattr(L, n)]
    assert not missing, missing


def test_binding_covers_header():
    assert set(declared()) <= set(_lib.EXPORTS)


def test_error_paths_without_gpu():
    L = _lib.lib()
    h = C.c_void_p()
    assert L.fkd_create(7, 0, None, None, C.byref(h)) == -1
    assert b"dim" in L.fkd_last_error()
    assert L.fkd_build(None, None) == -1
    offs = (C.c_uint64 * 3)(0, 5, 3)
    assert L.fkd_create_events(3, 2, offs, None, None, C.byref(h)) == -1
    assert b"monotone" in L.fkd_last_error()
    big = (C.c_uint64 * 2)(0, (1 << 30) + 1)
    assert L.fkd_create_device(3, 1, big, None, None, None, C.byref(h)) == -1
    assert b"2^30" in L.fkd_last_error()
    t, st, sz = C.c_uint32(), C.c_uint32(), C.c_uint32()
    assert L.fkd_shape_locate(10, 5, 0, C.byref(t)) == -1
    srv = C.c_void_p()
    assert L.fkd_box_server_create(None, 64, 4096, 20000, C.byref(srv)) == -1
    box = (C.c_float * 6)()
    cnt = C.c_uint32()
    assert L.fkd_box_server_search(None, box, None, 0, C.byref(cnt)) == -1
    assert L.fkd_box_server_stats(None, None, None, None) == -1
    L.fkd_box_server_destroy(None)


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "nope.so"))
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(RuntimeError, match="missing"):
        _lib.lib()


def test_deserialize_rejects_bad_buffers():
    """Header checks run on the host before any device call."""
    import ctypes as C

    import numpy as np

    from fkdtree_amd._lib import lib

    h = C.c_void_p()
    for data in (b"", b"NOTATREE" + bytes(40), b"FKDTREE1" + np.array([3, 1], np.uint32).tobytes() +
                 np.array([10], np.uint64).tobytes() + np.array([0, 10], np.uint64).tobytes()):
        buf = np.frombuffer(data + b"\0", dtype=np.uint8)
        rc = lib().fkd_deserialize(buf.ctypes.data, len(data), 0, None, C.byref(h))
        assert rc == -1 and h.value is None


def _header(dim, n_events, n_points):
    import numpy as np

    return (b"FKDTREE1" + np.array([dim, n_events], np.uint32).tobytes() +
            np.array([n_points], np.uint64).tobytes())


@pytest.mark.parametrize("n_events,n_points,tail", [
    (0xFFFFFFFF, 0, 0),            # n_events + 1 wraps to 0 in 32 bits
    (0xFFFFFFFF, 0, 16),
    ((1 << 30) + 1, 0, 0),         # above the event limit
    (0xFFFFFFFE, 0, 8),            # a size check that wraps would accept 32 bytes
    (2, 5, 24),                    # truncated: offsets present, arrays missing
    (1, 1 << 31, 16),              # more points than a handle holds
])
def test_deserialize_rejects_wrapped_and_truncated_headers(n_events, n_points, tail):
    """Malformed wire headers fail with FKD_ERR_INVALID before any
    allocation, indexing or device call (no GPU needed)."""
    import numpy as np

    from fkdtree_amd._lib import lib

    data = _header(3, n_events, n_points) + bytes(tail)
    buf = np.frombuffer(data + b"\0" * 8, dtype=np.uint8)
    h = C.c_void_p()
    rc = lib().fkd_deserialize(buf.ctypes.data, len(data), 0, None, C.byref(h))
    assert rc == -1 and h.value is None, lib().fkd_last_error()


def test_create_rejects_too_many_events():
    L = _lib.lib()
    h = C.c_void_p()
    offs = (C.c_uint64 * 1)(0)
    assert L.fkd_create_device(3, (1 << 30) + 1, offs, None, None, None, C.byref(h)) == -1
    assert b"events" in L.fkd_last_error()
