// Device helpers shared by the build and search kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fkd {

constexpr int kThreads = 256;  // every kernel runs 256-thread (4-wave) blocks
constexpr int kWave = 64;

__host__ __device__ inline uint32_t floor_log2(uint32_t x) { return 31u - (uint32_t)__builtin_clz(x | 1u); }

// Number of nodes in the left subtree of a left-complete binary tree with
// `len` nodes: the rank of the median the reference selects for an interval
// of that length (partition_complete_kdtree, FKDTree_cpu.h:287-296).
__host__ __device__ inline uint32_t left_size(uint32_t len) {
  // selects only (no branch in the build's per-item loops); for len <= 1 the
  // formula's value is discarded
  const uint32_t p = 1u << floor_log2(len), h = p >> 1;
  const uint32_t v = (h - 1u <= len - p) ? p - 1u : len - h;
  return len <= 1u ? 0u : v;
}

// Shape of the heap-ordered left-complete tree of n nodes: deepest depth D and
// the number R of nodes on depth D.
struct Shape {
  uint32_t n, D, R;
};
__host__ __device__ inline Shape shape_of(uint32_t n) {
  Shape s;
  s.n = n;
  s.D = floor_log2(n);
  s.R = n - ((1u << s.D) - 1u);
  return s;
}

// The build keeps every point array in in-order (symmetric) order of the final
// tree: the interval of the reference's node t on depth L (its theIntervalMin /
// theIntervalLength, FKDTree_cpu.h:224-227, 257-267) is the in-order range of
// that subtree.  Closed forms, with h = D - L and H = 2^h:
//   start(t) = t*H + min(R, t*H)
//   size(t)  = H - 1 + clamp(R - t*H, 0, H)
// and consecutive subtrees are separated by one position holding an ancestor.
__host__ __device__ inline void subtree_range(const Shape& s, uint32_t L, uint32_t t,
                                              uint32_t& start, uint32_t& size) {
  const uint32_t h = s.D - L;
  const uint32_t tH = t << h;
  const uint32_t H = 1u << h;
  start = tH + (s.R < tH ? s.R : tH);
  uint32_t m = s.R > tH ? s.R - tH : 0u;
  m = m < H ? m : H;
  size = H - 1u + m;
}

// Index t (0 <= t < 2^L) of the depth-L subtree whose in-order range holds
// position j, or of the subtree just before the separator j.
__host__ __device__ inline uint32_t subtree_at(const Shape& s, uint32_t L, uint32_t j) {
  const uint32_t h = s.D - L;
  const uint32_t B = (((s.R >> h) + 1u) << h) + s.R;
  return j < B ? (j >> (h + 1u)) : ((j - s.R) >> h);
}

// Total order of float keys under IEEE '<' with -0.0 == +0.0 (the reference's
// comparator, FKDTree_cpu.h:246-252, compares coordinates with '<' and '==').
// NaN is rejected before keys are formed.
__device__ inline uint32_t fkey(float f) {
  uint32_t u = __float_as_uint(f);
  u = (u << 1) == 0u ? 0u : u;
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__device__ inline float fkey_inv(uint32_t k) {
  return (k & 0x80000000u) ? __uint_as_float(k & 0x7fffffffu) : __uint_as_float(~k);
}

__device__ inline uint32_t lane_id() { return __lane_id(); }

template <class T>
__device__ inline T wave_incl_sum(T v) {
  const int lane = (int)lane_id();
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    T t = __shfl_up(v, o);
    if (lane >= o) v += t;
  }
  return v;
}

// Wave-wide inclusive scan of a 64-bit value under an associative op (first
// argument: the earlier lanes), by DPP moves: row_shr:1/2/4/8 inside rows of
// 16 lanes (lanes without a source read 0, which op must treat as its
// identity), then row_bcast:15 (lane 15 of rows 0 and 2 into rows 1 and 3)
// and row_bcast:31 (lane 31 into rows 2 and 3).  VALU moves instead of the
// LDS round trip of a shuffle.
template <int CTRL, int ROW_MASK>
__device__ inline uint64_t dpp_mov64(uint64_t v) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)v, CTRL, ROW_MASK, 0xf, true);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(v >> 32), CTRL, ROW_MASK, 0xf, true);
  return ((uint64_t)hi << 32) | lo;
}
template <class Op>
__device__ inline uint64_t wave_scan64(uint64_t v, Op op) {
  v = op(dpp_mov64<0x111, 0xf>(v), v);  // row_shr:1
  v = op(dpp_mov64<0x112, 0xf>(v), v);  // row_shr:2
  v = op(dpp_mov64<0x114, 0xf>(v), v);  // row_shr:4
  v = op(dpp_mov64<0x118, 0xf>(v), v);  // row_shr:8
  v = op(dpp_mov64<0x142, 0xa>(v), v);  // row_bcast:15
  v = op(dpp_mov64<0x143, 0xc>(v), v);  // row_bcast:31
  return v;
}

// Publication of decoupled look-back descriptors: status and value share one
// 64-bit word (status in the top two bits), so relaxed device-scope atomics
// suffice -- no other data is published through them.  Release/acquire
// fences would write back / invalidate the caches of an XCD on every tile.
constexpr uint64_t kStInvalid = 0, kStAggregate = 1ull << 62, kStPrefix = 2ull << 62;
constexpr uint64_t kStMask = 3ull << 62;

__device__ inline void publish(uint64_t* p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// One window of a look-back (wave-wide): lane l reads descriptor p[l] (null:
// a virtual prefix 0 before the first tile) until the nearest stopper (a
// descriptor for which is_stop holds) and every tile nearer than it have
// published; without a stopper, until all 64 have.  Waiting for all 64
// would hold every tile until the slowest of its 64 predecessors publishes.
// Returns the lane's descriptor and in *first the lane of the nearest
// stopper (kWave if none); lanes beyond it may be unpublished.  Only lanes
// still unpublished read again (the descriptors live past the caches, and
// re-reading all 64 per spin slows every waiting tile).
template <class Stop>
__device__ inline uint64_t look_window(uint64_t* p, Stop is_stop, uint32_t* first) {
  uint64_t w = p ? __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : kStPrefix;
  while (true) {
    const bool valid = (w & kStMask) != kStInvalid;
    const uint64_t sm = __ballot(valid && is_stop(w));
    const uint32_t f = sm ? (uint32_t)__builtin_ctzll(sm) : (uint32_t)kWave;
    const uint64_t need = f >= (uint32_t)kWave - 1u ? ~0ull : (2ull << f) - 1ull;
    if ((__ballot(!valid) & need) == 0ull) {
      *first = f;
      return w;
    }
    __builtin_amdgcn_s_sleep(1);
    if (!valid) w = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Dynamic block index, broadcast to the block.  A block of dispatch class
// c = blockIdx.x % 32 (four classes per XCD) takes the next rank r of its
// class and becomes block 32 r + c, so the indices are exactly
// 0 .. gridDim.x - 1.  One counter per class, kCtrStride words apart: a
// single counter for the whole grid serialises every workgroup of a launch
// on one address (about 17 ns each, 10 us of a 1,024-block launch's start;
// 8 classes still leave a 9 us tail, 32 none).  Look-back waits only
// on lower indices and stays deadlock-free under in-order dispatch, per
// queue or per XCD: every block dispatched before a class's next block holds
// an index below the one that block will take.  ctr needs kCtrWords zero
// 64-bit words.
constexpr uint32_t kCtrClasses = 32;
constexpr uint32_t kCtrStride = 512;  // uint32 words between two class counters
constexpr uint32_t kCtrWords = kCtrClasses * kCtrStride / 2;
__device__ inline uint32_t dynamic_block(uint32_t* ctr, uint32_t* lds) {
  if (threadIdx.x == 0) {
    const uint32_t c = blockIdx.x % kCtrClasses;
    *lds = atomicAdd(ctr + c * kCtrStride, 1u) * kCtrClasses + c;
  }
  __syncthreads();
  const uint32_t b = *lds;
  __syncthreads();
  return b;
}

// Workgroups are handed to the 8 XCDs in turn (linear workgroup id modulo
// 8).  XCD x's share of nb work items is the contiguous range starting at
// xcd_chunk_start(x, nb); xcd_item maps a linear workgroup id to its item, so
// that every XCD works on one contiguous range and its L2 holds that range's
// data (for instance whole events).
__host__ __device__ inline uint32_t xcd_chunk_start(uint32_t x, uint32_t nb) {
  return x * (nb >> 3) + (x < (nb & 7u) ? x : (nb & 7u));
}
__device__ inline uint32_t xcd_item(uint32_t lid, uint32_t nb) { return xcd_chunk_start(lid & 7u, nb) + (lid >> 3); }

__global__ void fill_u32_kernel(uint32_t* p, uint32_t v, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// Plain streaming copy (used to calibrate traffic counters).
__global__ void copy_u32_kernel(uint32_t* __restrict__ dst, const uint32_t* __restrict__ src, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// dst[i] = src[idx[i]]
__global__ void gather_u64_kernel(uint64_t* __restrict__ dst, const uint64_t* __restrict__ src,
                                  const uint64_t* __restrict__ idx, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = src[idx[i]];
}

__global__ void fill_u64_kernel(uint64_t* p, uint64_t v, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

}  // namespace fkd
