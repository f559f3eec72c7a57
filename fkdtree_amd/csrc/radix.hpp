// Per-event segmented LSD radix sort passes (8-bit digits), several
// independent key arrays at once (one per blockIdx.y).
//
// Tiles never straddle events: event e owns tiles [tb[e], tb[e+1]).  Digit
// counts are laid out per event as [digit][tile], so one global exclusive
// scan over all of them yields, for every (event, digit, tile), the first
// output slot inside that event's own range: the sort never mixes events.
#pragma once
#include "common.hpp"

namespace fkd {

// Keys per thread of the digit passes: a launch parameter of the build
// (8 or 16: tiles of 2,048 or 4,096 keys); the query plan uses kRadixItems.
constexpr int kRadixItems = 16;
constexpr int kRadixTile = kThreads * kRadixItems;

struct TileMap {
  const uint32_t* tile_ev;  // tile -> event; null: one event of single_n keys
  const uint32_t* tb;       // event -> first tile (n_events + 1 entries)
  const uint32_t* off;      // event -> first point (n_events + 1 entries)
  uint32_t T;               // number of tiles
  uint32_t single_n;
};

struct TileInfo {
  uint32_t e, lt, nt, base, end;
};

template <int TILE>
__device__ inline TileInfo tile_info(const TileMap& tm, uint32_t t) {
  TileInfo ti;
  if (!tm.tile_ev) {
    ti.e = 0;
    ti.lt = t;
    ti.nt = tm.T;
    ti.base = t * (uint32_t)TILE;
    ti.end = min(ti.base + (uint32_t)TILE, tm.single_n);
    return ti;
  }
  ti.e = tm.tile_ev[t];
  const uint32_t tb0 = tm.tb[ti.e];
  ti.lt = t - tb0;
  ti.nt = tm.tb[ti.e + 1] - tb0;
  const uint32_t o = tm.off[ti.e], o1 = tm.off[ti.e + 1];
  ti.base = o + ti.lt * (uint32_t)TILE;
  ti.end = min(ti.base + (uint32_t)TILE, o1);
  return ti;
}

// The first digit pass of a build whose points are in tie order reads the
// coordinates themselves (XDIM = dimension): the key of position i in array m
// is fkey(X[m][i]) and its value is i.  Its count kernel also writes the point
// records (dim <= 3: coordinates and id in 16 bytes, array 0's blocks) and the
// starting values of the key-space bounding box that the last pass narrows,
// and clears the build's scratch arena (zero words that later kernels need).
struct XSource {
  const float* X;        // [DIM][P] coordinates
  uint32_t P;
  const uint32_t* ids;   // point ids (records)
  float4* rec;           // point records, or null
  uint32_t* bbox;        // [2 DIM]: min keys, then max keys
  uint64_t* zero;        // words to clear
  uint64_t zero_words;
};

// counts[m][256*tb[e] + digit*nt + lt] = number of keys of tile t in array m
// whose digit (key >> shift) & 255 equals digit.
template <int ITEMS, int XDIM = 0>
__global__ __launch_bounds__(kThreads) void radix_count_kernel(const uint32_t* __restrict__ keys,
                                                               uint64_t kstride, uint32_t shift,
                                                               TileMap tm, uint32_t* counts, XSource xs) {
  constexpr int kItems = ITEMS;
  constexpr int kTile = kThreads * ITEMS;
  __shared__ uint32_t hist[4][256];
  const uint32_t t = blockIdx.x, m = blockIdx.y;
  if (XDIM && t == 0 && m == 0 && threadIdx.x < 2u * XDIM)
    xs.bbox[threadIdx.x] = threadIdx.x < (uint32_t)XDIM ? 0xffffffffu : 0u;
  if (XDIM) {
    const uint64_t nthreads = (uint64_t)gridDim.x * gridDim.y * kThreads;
    for (uint64_t i = ((uint64_t)m * gridDim.x + t) * kThreads + threadIdx.x; i < xs.zero_words; i += nthreads)
      xs.zero[i] = 0ull;
  }
  const int wave = threadIdx.x / kWave;
#pragma unroll
  for (int w = 0; w < 4; ++w) hist[w][threadIdx.x] = 0;
  __syncthreads();
  const TileInfo ti = tile_info<kTile>(tm, t);
  const uint32_t* K = keys + m * kstride;
  // all loads of the thread first, so they are in flight together
  uint32_t key[kItems];
  if (XDIM) {
    const float* Xm = xs.X + (uint64_t)m * xs.P;
    float x[kItems];
#pragma unroll
    for (int r = 0; r < kItems; ++r) {
      const uint32_t i = ti.base + threadIdx.x + (uint32_t)r * kThreads;
      x[r] = i < ti.end ? Xm[i] : 0.f;
      key[r] = fkey(x[r]);
    }
    if (XDIM <= 3 && m == 0 && xs.rec) {
#pragma unroll
      for (int r = 0; r < kItems; ++r) {
        const uint32_t i = ti.base + threadIdx.x + (uint32_t)r * kThreads;
        if (i < ti.end)
          xs.rec[i] = make_float4(x[r], XDIM > 1 ? xs.X[(uint64_t)xs.P + i] : 0.f,
                                  XDIM > 2 ? xs.X[2ull * xs.P + i] : 0.f, __uint_as_float(xs.ids[i]));
      }
    }
  } else {
#pragma unroll
    for (int r = 0; r < kItems; ++r) {
      const uint32_t i = ti.base + threadIdx.x + (uint32_t)r * kThreads;
      key[r] = i < ti.end ? K[i] : 0u;
    }
  }
  // Lanes that share a digit would serialise on one LDS address (a wave of
  // coordinates that differ only in their upper bits has a handful of
  // low digits).  Up to 4 times, the digit of the first lane not yet counted
  // is counted for all its lanes with one add when at least 8 lanes have it;
  // the remaining lanes add one each.
  const uint32_t lane = lane_id();
#pragma unroll
  for (int r = 0; r < kItems; ++r) {
    const uint32_t dig = (key[r] >> shift) & 255u;
    uint64_t act = __ballot(ti.base + threadIdx.x + (uint32_t)r * kThreads < ti.end);
    for (int it = 0; it < 4 && act != 0ull; ++it) {
      const uint32_t leader = (uint32_t)__builtin_ctzll(act);
      const uint32_t dl = (uint32_t)__builtin_amdgcn_readlane((int)dig, (int)leader);
      const uint64_t same = __ballot(dig == dl) & act;
      const uint32_t c = (uint32_t)__popcll(same);
      if (c < 8u) break;
      if (lane == leader) atomicAdd(&hist[wave][dl], c);
      act &= ~same;
    }
    if ((act >> lane) & 1ull) atomicAdd(&hist[wave][dig], 1u);
  }
  __syncthreads();
  const uint32_t c = hist[0][threadIdx.x] + hist[1][threadIdx.x] + hist[2][threadIdx.x] +
                     hist[3][threadIdx.x];
  counts[(uint64_t)m * 256u * tm.T + 256ull * (t - ti.lt) + (uint64_t)threadIdx.x * ti.nt + ti.lt] = c;
}

// Stable scatter of one digit pass.  offsets = exclusive scan of the counts of
// all arrays (array m's entries start at m * P).  Keys are written only when
// WRITE_KEYS; when WRITE_RANKS, ranks[4 val + m] = final slot of val (the
// ranks of a point in all dimensions form one 16-byte record), and, when
// bbox is given, the keys that land first and last in their event narrow
// bbox[m] (min) and bbox[M + m] (max), M = gridDim.y.
// The tile is first ordered by digit in LDS (stable: wave by wave, item by
// item, lane by lane), then written out in that order, so consecutive threads
// write consecutive slots of one digit's run instead of up to 64 buckets per
// store.
// FROM_X: the first pass of XSource: keys from the coordinates xs, values the
// positions (keys_in and vals_in unused).
template <bool WRITE_KEYS, bool WRITE_RANKS, int ITEMS, bool FROM_X = false>
__global__ __launch_bounds__(kThreads) void radix_scatter_kernel(
    const uint32_t* __restrict__ keys_in, uint32_t* __restrict__ keys_out,
    const uint32_t* __restrict__ vals_in, uint32_t* __restrict__ vals_out, uint64_t stride,
    uint32_t* __restrict__ ranks, uint32_t P, uint32_t shift, TileMap tm,
    const uint32_t* __restrict__ offsets, uint32_t* __restrict__ bbox, const float* __restrict__ xs) {
  constexpr int kItems = ITEMS;
  constexpr int kTile = kThreads * ITEMS;
  constexpr int kPerWave = kTile / (kThreads / kWave);
  __shared__ uint32_t cnt[4][256];
  __shared__ uint32_t goff[256];
  __shared__ uint32_t tstart[256];
  __shared__ uint32_t s_wsum[kThreads / kWave];
  __shared__ uint32_t s_key[kTile];
  __shared__ uint32_t s_val[kTile];
  // (tile, array) from the linear workgroup id, tile-major and XCD-chunked:
  // an XCD takes consecutive tiles, so the ranks of an event's points (its
  // 16-byte records, written at random inside the event's range) are
  // written by one XCD and merge in its L2
  const uint32_t w = xcd_item(blockIdx.x + blockIdx.y * gridDim.x, gridDim.x * gridDim.y);
  const uint32_t t = w / gridDim.y, m = w % gridDim.y;
  const int wave = threadIdx.x / kWave;
  const uint32_t lane = lane_id();
#pragma unroll
  for (int w = 0; w < 4; ++w) cnt[w][threadIdx.x] = 0;
  const TileInfo ti = tile_info<kTile>(tm, t);
  // the tile's first slot of digit threadIdx.x, needed only for the write-out:
  // loaded now, stored to LDS with the staging
  const uint32_t my_goff =
      offsets[(uint64_t)m * 256u * tm.T + 256ull * (t - ti.lt) + (uint64_t)threadIdx.x * ti.nt + ti.lt] - m * P;
  __syncthreads();
  const uint32_t* Ki = keys_in + m * stride;
  const uint32_t* Vi = vals_in + m * stride;
  uint32_t key[kItems], val[kItems], rank[kItems];
  const uint64_t lt_mask = (1ull << lane) - 1ull;
  const uint32_t wbase = ti.base + (uint32_t)wave * kPerWave + lane;
#pragma unroll
  for (int r = 0; r < kItems; ++r) {
    const uint32_t pos = wbase + (uint32_t)r * kWave;
    const bool valid = pos < ti.end;
    if (FROM_X) {
      key[r] = valid ? fkey(xs[(uint64_t)m * P + pos]) : 0u;
      val[r] = pos;
    } else {
      key[r] = valid ? Ki[pos] : 0u;
      val[r] = valid ? Vi[pos] : 0u;
    }
  }
#pragma unroll
  for (int r = 0; r < kItems; ++r) {
    const uint32_t pos = wbase + (uint32_t)r * kWave;
    const bool valid = pos < ti.end;
    const uint32_t dig = (key[r] >> shift) & 255u;
    // lanes with the same digit: for every bit, the ballot of the lanes that
    // have it set, complemented where this lane's bit is clear (s is all ones
    // for a set bit, so ~(ballot ^ s) selects without a per-lane branch)
    uint64_t match = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const uint32_t sb = (uint32_t)__builtin_amdgcn_sbfe((int)dig, (uint32_t)b, 1u);
      const uint64_t bal = __ballot(sb != 0u);
      match &= ~(bal ^ (((uint64_t)sb << 32) | sb));
    }
    uint32_t pre = 0;
    if (valid) pre = cnt[wave][dig];
    const uint64_t below = match & lt_mask;
    rank[r] = pre + (uint32_t)__popcll(below);
    if (valid && below == 0ull) cnt[wave][dig] = pre + (uint32_t)__popcll(match);
  }
  __syncthreads();
  {
    // tile-local start of every (wave, digit): digits in order, waves in order
    const uint32_t d = threadIdx.x;
    const uint32_t c0 = cnt[0][d], c1 = cnt[1][d], c2 = cnt[2][d], c3 = cnt[3][d];
    const uint32_t tot = c0 + c1 + c2 + c3;
    const uint32_t incl = wave_incl_sum(tot);
    if (lane == kWave - 1) s_wsum[wave] = incl;
    __syncthreads();
    uint32_t ts = incl - tot;
#pragma unroll
    for (int w = 0; w < kThreads / kWave; ++w) ts += (w < wave) ? s_wsum[w] : 0u;
    tstart[d] = ts;
    goff[d] = my_goff;
    cnt[0][d] = ts;
    cnt[1][d] = ts + c0;
    cnt[2][d] = ts + c0 + c1;
    cnt[3][d] = ts + c0 + c1 + c2;
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kItems; ++r) {
    const uint32_t pos = wbase + (uint32_t)r * kWave;
    if (pos < ti.end) {
      const uint32_t dig = (key[r] >> shift) & 255u;
      const uint32_t local = cnt[wave][dig] + rank[r];
      s_key[local] = key[r];
      s_val[local] = val[r];
    }
  }
  __syncthreads();
  uint32_t* Ko = keys_out + m * stride;
  uint32_t* Vo = vals_out + m * stride;
  const uint32_t tn = ti.end - ti.base;
  // event range of the tile (bounding box of the sorted keys)
  const uint32_t eo = tm.tile_ev ? tm.off[ti.e] : 0u, eo1 = tm.tile_ev ? tm.off[ti.e + 1] : tm.single_n;
#pragma unroll
  for (int r = 0; r < kItems; ++r) {
    const uint32_t i = threadIdx.x + (uint32_t)r * kThreads;
    if (i < tn) {
      const uint32_t k = s_key[i], v = s_val[i];
      const uint32_t dig = (k >> shift) & 255u;
      const uint32_t dest = goff[dig] + (i - tstart[dig]);
      Vo[dest] = v;
      if (WRITE_KEYS) Ko[dest] = k;
      if (WRITE_RANKS) {
        ranks[4 * (uint64_t)v + m] = dest;
        if (bbox && dest == eo) atomicMin(&bbox[m], k);
        if (bbox && dest + 1u == eo1) atomicMax(&bbox[gridDim.y + m], k);
      }
    }
  }
}

}  // namespace fkd
