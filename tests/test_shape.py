"""Closed-form heap shape used by the GPU build vs the reference's recursive
interval bookkeeping (FKDTree_cpu.h:224-267, partition_complete_kdtree
:287-296), evaluated on the host through the library's C ABI."""
import ctypes as C

import numpy as np
import pytest

from fkdtree_amd._lib import lib


def left_size(length):
    # FKDTree_cpu.h:287-296
    if length == 1:
        return 0
    p = 1 << (length.bit_length() - 1)
    return p - 1 if p // 2 - 1 <= length - p else length - p // 2


def recursive_intervals(n):
    """{(depth, t): (start, size)} following the reference's child interval rules."""
    out = {}
    D = n.bit_length() - 1
    imin = {0: 0}
    ilen = {0: n}
    for depth in range(D + 1):
        for t in range(1 << depth):
            node = (1 << depth) - 1 + t
            if node >= n:
                continue
            out[(depth, t)] = (imin[node], ilen[node])
            k = left_size(ilen[node]) if ilen[node] else 0
            l, r = 2 * node + 1, 2 * node + 2
            if l < n:
                imin[l], ilen[l] = imin[node], k
            if r < n:
                imin[r], ilen[r] = imin[node] + k + 1, ilen[node] - 1 - k
    return out


def c_subtree(n, L, t):
    s, z = C.c_uint32(), C.c_uint32()
    assert lib().fkd_shape_subtree(n, L, t, C.byref(s), C.byref(z)) == 0
    return s.value, z.value


def c_locate(n, L, j):
    t = C.c_uint32()
    assert lib().fkd_shape_locate(n, L, j, C.byref(t)) == 0
    return t.value


@pytest.mark.parametrize("n", list(range(1, 130)) + [255, 256, 257, 1000, 4097, 100_000])
def test_closed_form_matches_recursion(n):
    ref = recursive_intervals(n)
    D = n.bit_length() - 1
    for (L, t), (start, size) in ref.items():
        assert c_subtree(n, L, t) == (start, size), (n, L, t)
    if n > 5000:
        return
    for L in range(D + 1):
        for j in range(n):
            t = c_locate(n, L, j)
            assert 0 <= t < (1 << L)
            s, z = c_subtree(n, L, t)
            inside = s <= j < s + z
            # j is in subtree t, or it is the separator right after it
            assert inside or j == s + z, (n, L, j, t, s, z)
            if inside and (L, t) in ref:
                assert ref[(L, t)][0] <= j < sum(ref[(L, t)])


def test_left_size_table():
    for length in range(1, 70):
        s, z = c_subtree(length, 0, 0)
        assert (s, z) == (0, length)
    assert [left_size(x) for x in range(1, 10)] == [0, 1, 1, 2, 3, 3, 3, 4, 5]
