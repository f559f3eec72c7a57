// Single-pass device-wide exclusive prefix sum (decoupled look-back).
#pragma once
#include "common.hpp"

namespace fkd {

constexpr int kScanItems = 8;
constexpr int kScanTile = kThreads * kScanItems;

__host__ inline uint64_t scan_tiles(uint64_t n) { return (n + 1 + kScanTile - 1) / kScanTile; }

// out[j] = sum(in[0..j)) for j <= n, so out[n] is the total.  Launch
// scan_tiles(n) blocks; desc[0..scan_tiles(n)) and *ctr must be zero.
template <class OutT>
__global__ __launch_bounds__(kThreads) void scan_exclusive_kernel(const uint32_t* __restrict__ in,
                                                                  OutT* __restrict__ out, uint64_t n,
                                                                  uint64_t* desc, uint32_t* ctr) {
  __shared__ uint32_t s_blk;
  __shared__ uint64_t s_wave[kThreads / kWave];
  __shared__ uint64_t s_carry;
  const uint32_t tile = dynamic_block(ctr, &s_blk);
  const uint64_t base = (uint64_t)tile * kScanTile + (uint64_t)threadIdx.x * kScanItems;
  uint32_t v[kScanItems];
  uint64_t sum = 0;
  if (base + kScanItems <= n) {
    const uint4* p = reinterpret_cast<const uint4*>(in + base);
    uint4 a = p[0], b = p[1];
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
  } else {
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) v[i] = (base + i < n) ? in[base + i] : 0u;
  }
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) sum += v[i];
  const uint64_t incl = wave_incl_sum(sum);
  const int wave = threadIdx.x / kWave;
  if (lane_id() == kWave - 1) s_wave[wave] = incl;
  __syncthreads();
  uint64_t wave_off = 0, total = 0;
#pragma unroll
  for (int w = 0; w < kThreads / kWave; ++w) {
    if (w < wave) wave_off += s_wave[w];
    total += s_wave[w];
  }
  if (wave == 0) {
    // Look back over a window of 64 predecessors per step: lane l reads tile
    // tile-1-base-l; the sum stops at the nearest tile that has its prefix.
    uint64_t carry = 0;
    if (tile == 0) {
      if (lane_id() == 0) publish(&desc[0], kStPrefix | total);
    } else {
      if (lane_id() == 0) publish(&desc[tile], kStAggregate | total);
      for (int64_t base = 0;; base += kWave) {
        const int64_t j = (int64_t)tile - 1 - base - (int64_t)lane_id();
        uint32_t first;
        const uint64_t d = look_window(j >= 0 ? &desc[j] : nullptr,
                                       [](uint64_t x) { return (x & kStMask) == kStPrefix; }, &first);
        uint64_t v = lane_id() <= first ? (d & ~kStMask) : 0ull;
#pragma unroll
        for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o);
        carry += v;
        if (first < (uint32_t)kWave) break;
      }
      if (lane_id() == 0) publish(&desc[tile], kStPrefix | (carry + total));
    }
    if (lane_id() == 0) s_carry = carry;
  }
  __syncthreads();
  uint64_t run = s_carry + wave_off + incl - sum;
  OutT o[kScanItems];
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    o[i] = (OutT)run;
    run += v[i];
  }
  // the thread's outputs as 16-byte stores where all are in range and aligned
  constexpr int kVec = (int)(kScanItems * sizeof(OutT) / 16u);
  static_assert(kScanItems * sizeof(OutT) % 16u == 0, "whole 16-byte stores");
  if (base + kScanItems <= n + 1 && (reinterpret_cast<uintptr_t>(out + base) & 15u) == 0) {
    uint4 w[kVec];
    __builtin_memcpy(w, o, sizeof(o));
#pragma unroll
    for (int j = 0; j < kVec; ++j) reinterpret_cast<uint4*>(out + base)[j] = w[j];
  } else {
#pragma unroll
    for (int i = 0; i < kScanItems; ++i)
      if (base + i <= n) out[base + i] = o[i];
  }
}

}  // namespace fkd
