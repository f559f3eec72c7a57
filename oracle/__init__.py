"""CPU oracle for the FKDTree hot paths -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only to check results or as the
timed CPU baseline.  The product (``fkdtree_amd``) never calls it.

Two native libraries are wrapped:

* ``liboracle.so`` (``fkd_oracle.cpp``): a restatement of the reference CPU
  tree, ``FKDTree_cpu.h:221-296`` (build) and ``:123-169, :302-321`` (search),
  with an unbounded BFS frontier, plus a brute-force scan (``FKDTree.h:54-86``).
* ``_ref/libfkdref.so`` (``ref_shim.cpp``): the reference's own
  ``FKDTree_cpu.h`` compiled from its sources.  Present only where it was built
  (it travels with the tree as a built library).
* ``_ref/libcmsswref.so`` (``cmssw_shim.cpp``): the reference's copy of the
  CMSSW KD-tree (``KDTreeLinkerAlgoT.h``), a side comparator for the CPU
  timing table (``main.cpp:560-629``).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libfkdref.so")
CMSSW_SO = os.path.join(HERE, "_ref", "libcmsswref.so")

_P = C.c_void_p
_U64 = C.c_uint64
_I = C.c_int


def build_libs(quiet: bool = True) -> None:
    """Compile liboracle.so (and _ref/ where the reference sources exist)."""
    out = subprocess.run(["make", "-C", HERE, "-B"], capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + out.stdout + out.stderr)
    if not quiet:
        print(out.stdout)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"{ORACLE_SO} is missing: run __graft_entry__.build()")
        L = C.CDLL(ORACLE_SO)
        L.orc_build.argtypes = [_I, _U64, _P, _P, _P, _P]
        L.orc_build_events.argtypes = [_I, _U64, _P, _P, _P, _P, _P, _I]
        L.orc_search_count.argtypes = [_I, _U64, _P, _P, _U64, _P, _P, _P, _P, _I]
        L.orc_search_fill.argtypes = [_I, _U64, _P, _P, _U64, _P, _P, _P, _P, _I]
        L.orc_search_events_count.argtypes = [_I, _U64, _P, _P, _P, _U64, _P, _P, _P, _P, _P, _I]
        L.orc_search_events_fill.argtypes = [_I, _U64, _P, _P, _P, _U64, _P, _P, _P, _P, _P, _I]
        L.orc_brute.argtypes = [_I, _U64, _P, _P, _U64, _P, _P, _P, _P, _P, _I]
        L.orc_sort_segments.argtypes = [_U64, _P, _P, _P, _I]
        L.orc_sort_segments.restype = None
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        if not ref_available():
            raise RuntimeError(f"{REF_SO} is missing (reference sources absent when built)")
        L = C.CDLL(REF_SO)
        L.ref_create.restype = _P
        L.ref_create.argtypes = [_I, _U64, _P, _P]
        L.ref_build.argtypes = [_P]
        L.ref_get_tree.argtypes = [_P, _P, _P]
        L.ref_search.argtypes = [_P, _U64, _P, _P, _I, _I, _P]
        L.ref_fetch.argtypes = [_P, _P, _P]
        L.ref_destroy.argtypes = [_P]
        _ref = L
    return _ref


def cmssw_available() -> bool:
    return os.path.exists(CMSSW_SO)


def cmssw_build_search(coords: np.ndarray, mins: np.ndarray, maxs: np.ndarray):
    """CMSSW KD-tree (reference copy) on one thread: build, then every box in
    turn, as main.cpp:560-629.  Returns (build_ms, search_ms, found)."""
    L = C.CDLL(CMSSW_SO)
    L.cms_build_search.argtypes = [_U64, _P, _U64, _P, _P, _P, _P, _P]
    c = np.ascontiguousarray(coords, dtype=np.float32)
    lo = np.ascontiguousarray(mins, dtype=np.float32)
    hi = np.ascontiguousarray(maxs, dtype=np.float32)
    assert c.shape[0] == 3 and lo.shape[0] == 3, "the CMSSW tree is instantiated for 3 dimensions"
    b, s, f = C.c_double(), C.c_double(), C.c_uint64()
    L.cms_build_search(c.shape[1], _ptr(c), lo.shape[1], _ptr(lo), _ptr(hi), C.byref(b), C.byref(s), C.byref(f))
    return b.value, s.value, f.value


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _soa(coords: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(coords, dtype=np.float32)


def max_threads() -> int:
    return int(lib().orc_max_threads())


def build(coords: np.ndarray, ids: np.ndarray):
    """Heap-ordered SoA tree arrays (coords[dim, n], ids[n])."""
    coords = _soa(coords)
    ids = np.ascontiguousarray(ids, dtype=np.uint32)
    dim, n = coords.shape
    out = np.empty_like(coords)
    out_ids = np.empty_like(ids)
    if lib().orc_build(dim, n, _ptr(coords), _ptr(ids), _ptr(out), _ptr(out_ids)) != 0:
        raise ValueError("bad dimension")
    return out, out_ids


def build_events(coords: np.ndarray, ids: np.ndarray, offsets: np.ndarray, threads: int = 0):
    coords = _soa(coords)
    ids = np.ascontiguousarray(ids, dtype=np.uint32)
    off = np.ascontiguousarray(offsets, dtype=np.uint64)
    out = np.empty_like(coords)
    out_ids = np.empty_like(ids)
    lib().orc_build_events(coords.shape[0], off.shape[0] - 1, _ptr(off), _ptr(coords), _ptr(ids),
                           _ptr(out), _ptr(out_ids), threads)
    return out, out_ids


def _csr(counts: np.ndarray) -> np.ndarray:
    off = np.zeros(counts.shape[0] + 1, dtype=np.uint64)
    np.cumsum(counts, out=off[1:])
    return off


def search(tree: np.ndarray, tids: np.ndarray, mins: np.ndarray, maxs: np.ndarray,
           threads: int = 0, with_visited: bool = False):
    """Level-BFS search of every box: (offsets[nq+1], ids) in BFS order [, visited[nq]]."""
    tree = _soa(tree)
    tids = np.ascontiguousarray(tids, dtype=np.uint32)
    mins, maxs = _soa(mins), _soa(maxs)
    dim, n = tree.shape
    nq = mins.shape[1]
    counts = np.empty(nq, dtype=np.uint32)
    vis = np.empty(nq, dtype=np.uint32) if with_visited else None
    L = lib()
    L.orc_search_count(dim, n, _ptr(tree), _ptr(tids), nq, _ptr(mins), _ptr(maxs), _ptr(counts),
                       _ptr(vis) if vis is not None else None, threads)
    off = _csr(counts)
    out = np.empty(int(off[-1]), dtype=np.uint32)
    L.orc_search_fill(dim, n, _ptr(tree), _ptr(tids), nq, _ptr(mins), _ptr(maxs), _ptr(off),
                      _ptr(out), threads)
    return (off, out, vis) if with_visited else (off, out)


def search_counts(tree, tids, mins, maxs, threads: int = 0):
    """(counts[nq], visited[nq]) of the level-BFS search."""
    tree = _soa(tree)
    tids = np.ascontiguousarray(tids, dtype=np.uint32)
    mins, maxs = _soa(mins), _soa(maxs)
    nq = mins.shape[1]
    counts = np.empty(nq, dtype=np.uint32)
    vis = np.empty(nq, dtype=np.uint32)
    lib().orc_search_count(tree.shape[0], tree.shape[1], _ptr(tree), _ptr(tids), nq, _ptr(mins),
                           _ptr(maxs), _ptr(counts), _ptr(vis), threads)
    return counts, vis


def search_events(tree, tids, offsets, qev, mins, maxs, threads: int = 0, with_visited=False):
    """Per-event search: query q searches the tree of event qev[q]."""
    tree = _soa(tree)
    tids = np.ascontiguousarray(tids, dtype=np.uint32)
    off = np.ascontiguousarray(offsets, dtype=np.uint64)
    qev = np.ascontiguousarray(qev, dtype=np.uint32)
    mins, maxs = _soa(mins), _soa(maxs)
    dim = tree.shape[0]
    nq = mins.shape[1]
    counts = np.empty(nq, dtype=np.uint32)
    vis = np.empty(nq, dtype=np.uint32) if with_visited else None
    L = lib()
    L.orc_search_events_count(dim, off.shape[0] - 1, _ptr(off), _ptr(tree), _ptr(tids), nq,
                              _ptr(qev), _ptr(mins), _ptr(maxs), _ptr(counts),
                              _ptr(vis) if vis is not None else None, threads)
    qoff = _csr(counts)
    out = np.empty(int(qoff[-1]), dtype=np.uint32)
    L.orc_search_events_fill(dim, off.shape[0] - 1, _ptr(off), _ptr(tree), _ptr(tids), nq,
                             _ptr(qev), _ptr(mins), _ptr(maxs), _ptr(qoff), _ptr(out), threads)
    return (qoff, out, vis) if with_visited else (qoff, out)


def search_events_counts(tree, tids, offsets, qev, mins, maxs, threads: int = 0):
    """counts[nq] of the per-event search (no ids): query q searches event qev[q]."""
    tree = _soa(tree)
    tids = np.ascontiguousarray(tids, dtype=np.uint32)
    off = np.ascontiguousarray(offsets, dtype=np.uint64)
    qev = np.ascontiguousarray(qev, dtype=np.uint32)
    mins, maxs = _soa(mins), _soa(maxs)
    nq = mins.shape[1]
    counts = np.empty(nq, dtype=np.uint32)
    lib().orc_search_events_count(tree.shape[0], off.shape[0] - 1, _ptr(off), _ptr(tree), _ptr(tids), nq,
                                  _ptr(qev), _ptr(mins), _ptr(maxs), _ptr(counts), None, threads)
    return counts


def brute(coords, ids, mins, maxs, threads: int = 0):
    """Brute-force closed-box membership: (offsets[nq+1], ids in point order)."""
    coords = _soa(coords)
    ids = np.ascontiguousarray(ids, dtype=np.uint32)
    mins, maxs = _soa(mins), _soa(maxs)
    dim, n = coords.shape
    nq = mins.shape[1]
    counts = np.empty(nq, dtype=np.uint32)
    L = lib()
    L.orc_brute(dim, n, _ptr(coords), _ptr(ids), nq, _ptr(mins), _ptr(maxs), _ptr(counts), None,
                None, threads)
    off = _csr(counts)
    out = np.empty(int(off[-1]), dtype=np.uint32)
    L.orc_brute(dim, n, _ptr(coords), _ptr(ids), nq, _ptr(mins), _ptr(maxs), _ptr(counts),
                _ptr(off), _ptr(out), threads)
    return off, out


def sort_segments(off: np.ndarray, ids: np.ndarray) -> np.ndarray:
    """Sort every CSR segment in place order-independently (returns a copy)."""
    off = np.ascontiguousarray(off, dtype=np.uint64)
    ids = np.ascontiguousarray(ids, dtype=np.uint32)
    nq = off.shape[0] - 1
    if nq < 0 or int(off[-1]) != ids.shape[0] or np.any(off[1:] < off[:-1]):
        raise ValueError("offsets do not describe ids")
    out = np.empty_like(ids)
    lib().orc_sort_segments(nq, _ptr(off), _ptr(ids), _ptr(out), 0)
    return out


class RefTree:
    """The reference's FKDTree_CPU<float, dim>, compiled from its sources."""

    def __init__(self, coords: np.ndarray, ids: np.ndarray):
        self.coords = _soa(coords)
        self.ids = np.ascontiguousarray(ids, dtype=np.uint32)
        self.dim, self.n = self.coords.shape
        self.h = ref().ref_create(self.dim, self.n, _ptr(self.coords), _ptr(self.ids))
        if not self.h:
            raise ValueError("bad dimension")

    def build(self):
        ref().ref_build(self.h)

    def tree(self):
        out = np.empty_like(self.coords)
        out_ids = np.empty_like(self.ids)
        ref().ref_get_tree(self.h, _ptr(out), _ptr(out_ids))
        return out, out_ids

    def search(self, mins, maxs, variant: int = 0, threads: int = 0):
        """variant 0 = search_in_the_box, 1 = _BFS, 2 = _recursive."""
        mins, maxs = _soa(mins), _soa(maxs)
        nq = mins.shape[1]
        counts = np.empty(nq, dtype=np.uint32)
        if ref().ref_search(self.h, nq, _ptr(mins), _ptr(maxs), variant, threads, _ptr(counts)) != 0:
            raise ValueError("bad variant")
        off = _csr(counts)
        out = np.empty(int(off[-1]), dtype=np.uint32)
        ref().ref_fetch(self.h, _ptr(off), _ptr(out))
        return off, out

    def close(self):
        if self.h:
            ref().ref_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
