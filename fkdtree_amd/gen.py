"""Deterministic synthetic point clouds, query boxes and event streams.

Every value is produced with integer arithmetic and IEEE-754 basic operations
(+ - * /, floor, round-half-even) evaluated by numpy in a fixed order, so a seed
yields the same bits on any x86-64 host: no libm transcendental (whose last
bit may depend on the CPU's SIMD dispatch) is in the data path.  ``exp2`` and
``log`` below are fixed-order polynomial evaluations.

Random numbers come from counter-based SplitMix64: output ``i`` of stream ``s``
is ``mix64(base(seed, s) + (i + 1) * 0x9E3779B97F4A7C15)``.

The workloads mirror the reference's benchmark set-up (``main.cpp:167-187``
draws points uniformly and centres one box on every point) at the sizes used
for this hot path:

* ``uniform_100k``  - N=100,000 uniform in [0,1)^3, one box per point, hw 0.02
* ``mixed_1m``      - N=1,048,576, half uniform, half in 1,000 Gaussian blobs
                      (sigma 0.005), quantised to a 2^-16 grid; 1,048,576 boxes
                      centred on random points, hw 0.01
* ``uniform_4m``    - N=4,194,304 uniform; 4,194,304 boxes with uniform centres
                      and per-axis half-widths log-uniform in [1e-4, 3e-2]
* ``events_2048``   - 2,048 independent events, sizes log-uniform in
                      [1,000, 200,000], the ``mixed`` point mix; every point of
                      an event is the centre of one box, hw 0.01
* ``serving``       - Poisson arrivals, sizes log-uniform in [1,000, 500,000]
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

_U64 = np.uint64
_GAMMA = _U64(0x9E3779B97F4A7C15)
_M1 = _U64(0xBF58476D1CE4E5B9)
_M2 = _U64(0x94D049BB133111EB)
_MASK64 = (1 << 64) - 1

LN2 = 0.6931471805599453
# log2 of the ratio hi/lo of the log-uniform ranges used by the workloads.
LOG2_300 = 8.228818690495881
LOG2_200 = 7.643856189774724
LOG2_500 = 8.965784284662087

SEED_BASE = 0xF1D70000
WORKLOADS = ("uniform_100k", "mixed_1m", "uniform_4m", "events_2048", "serving")


def workload_seed(name: str) -> int:
    return SEED_BASE + WORKLOADS.index(name)


def _mix64(z: np.ndarray) -> np.ndarray:
    z = (z ^ (z >> _U64(30))) * _M1
    z = (z ^ (z >> _U64(27))) * _M2
    return z ^ (z >> _U64(31))


def _stream_base(seed: int, stream: int) -> np.uint64:
    b = (seed * 0xD1B54A32D192ED03 + (stream + 1) * 0x8CB92BA72F3D8DD7) & _MASK64
    return _U64(b)


def u64(seed: int, stream: int, n: int, start: int = 0) -> np.ndarray:
    """n SplitMix64 outputs of (seed, stream), counters start..start+n-1."""
    ctr = np.arange(start + 1, start + n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = ctr * _GAMMA + _stream_base(seed, stream)
        return _mix64(z)


def uniform_f32(seed: int, stream: int, n: int) -> np.ndarray:
    """float32 uniform on the 2^-24 grid of [0,1): exact k * 2^-24."""
    return (u64(seed, stream, n) >> _U64(40)).astype(np.float32) * np.float32(2.0 ** -24)


def uniform_f64(seed: int, stream: int, n: int) -> np.ndarray:
    """float64 uniform on the 2^-53 grid of [0,1)."""
    return (u64(seed, stream, n) >> _U64(11)).astype(np.float64) * (2.0 ** -53)


def irwin_hall12(seed: int, stream: int, n: int) -> np.ndarray:
    """Approximate standard normal: sum of 12 uniforms minus 6, summed in order."""
    acc = np.zeros(n, dtype=np.float64)
    for j in range(12):
        acc = acc + uniform_f64(seed, stream * 16 + j, n)
    return acc - 6.0


_EXP_COEF = [1.0 / math.factorial(k) for k in range(18)]


def exp2(f: np.ndarray) -> np.ndarray:
    """2**f for f >= 0 by range reduction and a fixed-order Taylor polynomial."""
    f = np.asarray(f, dtype=np.float64)
    n = np.floor(f)
    t = (f - n) * LN2
    acc = np.full_like(t, _EXP_COEF[-1])
    for c in reversed(_EXP_COEF[:-1]):
        acc = acc * t + c
    return np.ldexp(acc, n.astype(np.int64))


def log(x: np.ndarray) -> np.ndarray:
    """Natural log for x > 0: frexp reduction and the atanh series, fixed order."""
    x = np.asarray(x, dtype=np.float64)
    m, e = np.frexp(x)
    small = m < 0.7071067811865476
    m = np.where(small, m * 2.0, m)
    e = np.where(small, e - 1, e)
    z = (m - 1.0) / (m + 1.0)
    z2 = z * z
    acc = np.full_like(z, 1.0 / 25.0)
    for k in range(23, 0, -2):
        acc = acc * z2 + 1.0 / k
    return e.astype(np.float64) * LN2 + 2.0 * z * acc


def log_uniform(seed: int, stream: int, n: int, lo: float, log2_ratio: float) -> np.ndarray:
    """lo * 2**(u * log2_ratio), u uniform in [0,1): log-uniform in [lo, lo*2**log2_ratio)."""
    return lo * exp2(uniform_f64(seed, stream, n) * log2_ratio)


def quantise(x: np.ndarray, bits: int = 16) -> np.ndarray:
    """Round to the 2^-bits grid (round-half-even); keeps the sign of zero."""
    s = float(1 << bits)
    return np.round(np.asarray(x, dtype=np.float64) * s) / s


@dataclass
class PointSet:
    """SoA float32 coordinates ``coords[dim, n]`` and uint32 ids ``ids[n]``."""

    coords: np.ndarray
    ids: np.ndarray

    @property
    def n(self) -> int:
        return int(self.ids.shape[0])

    @property
    def dim(self) -> int:
        return int(self.coords.shape[0])

    def aos(self) -> np.ndarray:
        """FKDPoint<float,dim> records: dim float32 coordinates then the uint32 id."""
        rec = np.empty((self.n, self.dim + 1), dtype=np.uint32)
        rec[:, : self.dim] = self.coords.T.view(np.uint32)
        rec[:, self.dim] = self.ids
        return rec


@dataclass
class Boxes:
    """SoA float32 ``mins[dim, q]`` / ``maxs[dim, q]`` of closed query boxes."""

    mins: np.ndarray
    maxs: np.ndarray

    @property
    def q(self) -> int:
        return int(self.mins.shape[1])


def uniform_points(n: int, seed: int, dim: int = 3, stream0: int = 0) -> PointSet:
    coords = np.stack([uniform_f32(seed, stream0 + d, n) for d in range(dim)])
    return PointSet(np.ascontiguousarray(coords), np.arange(n, dtype=np.uint32))


def mixed_points(n: int, seed: int, n_blobs: int = 1000, sigma: float = 0.005,
                 dim: int = 3, stream0: int = 100) -> PointSet:
    """Even ids uniform in [0,1)^dim, odd ids in Gaussian blobs; 2^-16 grid.

    The grid forces equal coordinates (ties broken by id) and rounding of small
    negative blob coordinates produces -0.0, which must tie with +0.0.
    """
    uni = np.stack([uniform_f64(seed, stream0 + d, n) for d in range(dim)])
    centres = np.stack([uniform_f64(seed, stream0 + 10 + d, n_blobs) for d in range(dim)])
    which = (u64(seed, stream0 + 20, n) % _U64(n_blobs)).astype(np.int64)
    blob = np.stack([centres[d][which] + sigma * irwin_hall12(seed, stream0 + 30 + d, n)
                     for d in range(dim)])
    is_blob = (np.arange(n) & 1) == 1
    x = np.where(is_blob[None, :], blob, uni)
    coords = quantise(x).astype(np.float32)
    return PointSet(np.ascontiguousarray(coords), np.arange(n, dtype=np.uint32))


def boxes_around(centres: np.ndarray, hw) -> Boxes:
    """Closed boxes [c - hw, c + hw] computed in float32 (one rounding each)."""
    c = np.asarray(centres, dtype=np.float32)
    h = np.asarray(hw, dtype=np.float32)
    return Boxes(np.ascontiguousarray(c - h), np.ascontiguousarray(c + h))


def workload_uniform_100k(n: int = 100_000):
    seed = workload_seed("uniform_100k")
    pts = uniform_points(n, seed)
    return pts, boxes_around(pts.coords, 0.02)


def workload_mixed_1m(n: int = 1 << 20, q: int | None = None):
    seed = workload_seed("mixed_1m")
    pts = mixed_points(n, seed)
    q = n if q is None else q
    idx = (u64(seed, 500, q) % _U64(n)).astype(np.int64)
    return pts, boxes_around(pts.coords[:, idx], 0.01)


def workload_uniform_4m(n: int = 1 << 22, q: int | None = None):
    seed = workload_seed("uniform_4m")
    pts = uniform_points(n, seed)
    q = n if q is None else q
    centres = np.stack([uniform_f32(seed, 600 + d, q) for d in range(3)])
    hw = np.stack([log_uniform(seed, 610 + d, q, 1e-4, LOG2_300) for d in range(3)]).astype(np.float32)
    return pts, boxes_around(centres, hw)


@dataclass
class EventBatch:
    """Independent per-event point sets laid end to end.

    ``offsets[e]..offsets[e+1]`` are the points of event ``e`` in ``points``;
    ids restart at 0 in every event.
    """

    points: PointSet
    offsets: np.ndarray

    @property
    def n_events(self) -> int:
        return int(self.offsets.shape[0] - 1)

    def event(self, e: int) -> PointSet:
        a, b = int(self.offsets[e]), int(self.offsets[e + 1])
        return PointSet(np.ascontiguousarray(self.points.coords[:, a:b]), self.points.ids[a:b].copy())


def event_sizes(n_events: int, seed: int, lo: int, log2_ratio: float, stream: int = 700) -> np.ndarray:
    return np.floor(log_uniform(seed, stream, n_events, float(lo), log2_ratio)).astype(np.int64)


def mixed_event(n: int, seed: int, e: int) -> PointSet:
    """One event of the mixed distribution; one blob per 1,024 points."""
    return mixed_points(n, seed, n_blobs=max(1, n // 1024), stream0=1000 + 64 * e)


def workload_events_2048(n_events: int = 2048, scale: float = 1.0):
    """Batched events; ``scale`` shrinks every event (for quick tests)."""
    seed = workload_seed("events_2048")
    sizes = event_sizes(n_events, seed, 1000, LOG2_200)
    sizes = np.maximum(1, np.floor(sizes * scale).astype(np.int64))
    evs = [mixed_event(int(s), seed, e) for e, s in enumerate(sizes)]
    return concat_events(evs)


def concat_events(evs) -> EventBatch:
    offs = np.zeros(len(evs) + 1, dtype=np.int64)
    offs[1:] = np.cumsum([p.n for p in evs])
    dim = evs[0].dim if evs else 3
    coords = np.concatenate([p.coords for p in evs], axis=1) if evs else np.zeros((dim, 0), np.float32)
    ids = np.concatenate([p.ids for p in evs]) if evs else np.zeros(0, np.uint32)
    return EventBatch(PointSet(np.ascontiguousarray(coords), ids), offs)


def event_self_boxes(batch: EventBatch, hw: float = 0.01) -> Boxes:
    """One box per point of every event, centred on it (query order = point order)."""
    return boxes_around(batch.points.coords, hw)


def serving_arrivals(n_events: int, rate_hz: float, seed: int | None = None, lo: int = 1000,
                     log2_ratio: float = LOG2_500):
    """(arrival times in s, sizes) of a Poisson stream of events."""
    seed = workload_seed("serving") if seed is None else seed
    u = uniform_f64(seed, 800, n_events)
    gaps = -log(1.0 - u) / rate_hz
    times = np.cumsum(gaps)
    sizes = event_sizes(n_events, seed, lo, log2_ratio, stream=801)
    return times, sizes


def serving_event(n: int, e: int) -> PointSet:
    return mixed_event(n, workload_seed("serving"), e)


def digest(*arrays) -> str:
    """SHA-256 over the raw little-endian bytes of the arrays, in order."""
    import hashlib

    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()
