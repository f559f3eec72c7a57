"""Shared inputs and comparisons for the parity tests."""
import numpy as np

from fkdtree_amd import gen


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def adversarial_points(kind: str, n: int, dim: int = 3, seed: int = 11) -> gen.PointSet:
    """Inputs that stress the (coordinate, id) order: ties, signed zeros, negatives."""
    if kind == "uniform":
        return gen.uniform_points(n, seed, dim)
    if kind == "mixed":
        return gen.mixed_points(n, seed, n_blobs=max(1, n // 64), dim=dim)
    if kind == "equal":
        return gen.PointSet(np.full((dim, n), 0.5, np.float32), np.arange(n, dtype=np.uint32))
    if kind == "zeros":
        u = gen.u64(seed, 1, dim * n).reshape(dim, n)
        c = np.where((u & np.uint64(1)) == 1, np.float32(-0.0), np.float32(0.0)).astype(np.float32)
        c = np.where((u & np.uint64(6)) == 6, np.float32(1e-45), c)
        c = np.where((u & np.uint64(24)) == 24, np.float32(-1e-45), c).astype(np.float32)
        return gen.PointSet(np.ascontiguousarray(c), np.arange(n, dtype=np.uint32))
    if kind == "negative":
        p = gen.mixed_points(n, seed, n_blobs=max(1, n // 16), dim=dim)
        c = gen.quantise(p.coords.astype(np.float64) * 8.0 - 4.0, bits=4).astype(np.float32)
        return gen.PointSet(np.ascontiguousarray(c), p.ids)
    if kind == "grid":
        u = gen.u64(seed, 2, dim * n).reshape(dim, n)
        c = ((u % np.uint64(5)).astype(np.float32) - 2.0).astype(np.float32)
        return gen.PointSet(np.ascontiguousarray(c), np.arange(n, dtype=np.uint32))
    if kind == "shuffled_ids":
        p = gen.mixed_points(n, seed, n_blobs=max(1, n // 32), dim=dim)
        perm = np.argsort(gen.u64(seed, 3, n), kind="stable").astype(np.uint32)
        ids = (perm * np.uint32(7) + np.uint32(1000)).astype(np.uint32)
        return gen.PointSet(p.coords, ids)
    if kind == "huge_values":
        u = gen.u64(seed, 4, dim * n).reshape(dim, n)
        vals = np.array([np.inf, -np.inf, 3.4e38, -3.4e38, 0.0, -0.0, 1.0], np.float32)
        c = vals[(u % np.uint64(len(vals))).astype(np.int64)]
        return gen.PointSet(np.ascontiguousarray(c), np.arange(n, dtype=np.uint32))
    raise ValueError(kind)


KINDS = ("uniform", "mixed", "equal", "zeros", "negative", "grid", "shuffled_ids", "huge_values")
SMALL_N = (1, 2, 3, 4, 5, 7, 8, 9, 15, 16, 17, 31, 32, 33, 100, 1000, 2047, 2048, 2049, 4095, 4096, 4097)


def self_boxes(p: gen.PointSet, hw: float) -> gen.Boxes:
    return gen.boxes_around(p.coords, hw)


def special_boxes(p: gen.PointSet) -> gen.Boxes:
    """Empty (min > max), degenerate (min == max on points), whole space, NaN bounds."""
    dim, n = p.coords.shape
    k = min(n, 8)
    mins, maxs = [], []
    c = p.coords[:, :k]
    mins.append(c); maxs.append(c)                         # degenerate boxes on points
    mins.append(c + 1.0); maxs.append(c - 1.0)             # inverted: empty
    mins.append(np.full((dim, 1), -np.inf, np.float32)); maxs.append(np.full((dim, 1), np.inf, np.float32))
    mins.append(np.full((dim, 1), np.nan, np.float32)); maxs.append(np.full((dim, 1), 1.0, np.float32))
    mins.append(np.full((dim, 1), 0.0, np.float32)); maxs.append(np.full((dim, 1), np.nan, np.float32))
    mins.append(np.full((dim, 1), -0.0, np.float32)); maxs.append(np.full((dim, 1), 0.0, np.float32))
    return gen.Boxes(np.ascontiguousarray(np.concatenate(mins, 1).astype(np.float32)),
                     np.ascontiguousarray(np.concatenate(maxs, 1).astype(np.float32)))


def sorted_sets(off, ids):
    import oracle
    return oracle.sort_segments(off, ids)
