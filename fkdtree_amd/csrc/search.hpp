// Box search over heap-ordered trees (one query per lane).
//
// The visited node set is the reference's: from a node with split value v on
// dimension d, the left son is visited unless v < min[d], the right son when
// v < min[d] or min[d] <= v <= max[d], and a son only if it exists
// (FKDTree_cpu.h:150-163).  A node is reported when v lies in [min[d], max[d]]
// and the point lies in the closed box (FKDTree_cpu.h:145-148, 302-321).
// Instead of the level-synchronous frontier (FQueue, FKDTree_cpu.h:126-167)
// every lane walks its query's visited set depth first without a stack: a
// 32-bit mask remembers on which depths a right son is still pending, and the
// heap index of that son follows from the current index by a shift.
//
// Queries are processed in spatial (Morton) order of their box centres so
// that the lanes of a wave walk nearly the same nodes: loads coalesce, the
// upper tree stays in L2, and lanes finish together.  Results stay in query
// order (CSR offsets come from a scan of the per-query counts).
#pragma once
#include "build.hpp"
#include "common.hpp"

namespace fkd {

struct SearchArgs {
  const float* nodes;     // packed records, NodeRec<DIM>::kFloats floats each
  const uint32_t* off;    // event -> first node (n_events + 1), or null
  uint32_t n_single;      // node count when off == null
  const uint32_t* qev;    // query -> event, or null (event 0)
  const uint32_t* perm;   // processing slot -> query (null: identity)
  uint32_t nq;
  const float* mins;      // [DIM][nq], indexed by processing slot
  const float* maxs;
  uint32_t* counts;       // count pass output (query order)
  uint32_t* counts_slot;  // pool mode: counts in processing order
  uint32_t* visited;      // stats pass: nodes visited per query
  const uint64_t* qoff;   // fill pass: first output slot of every query
  uint32_t* out;          // fill pass output
  uint64_t cap;           // capacity of out
  uint32_t out_aligned;   // out is 64-byte aligned
  uint32_t node_bytes;    // size of the node records (buffer resource range)
  // single-pass mode: hits go to 16-id chunks of a pool split into
  // kPoolRegions regions of region_chunks chunks; block b allocates in
  // region b % kPoolRegions, a wave reserves pool_batch chunks at a time.
  uint32_t* pool;         // kPoolRegions * region_chunks * 16 ids
  uint2* chunk_meta;      // per chunk: (processing slot, index of the chunk in its query)
  uint32_t* pool_ctr;     // per region: chunks reserved (may exceed region_chunks)
  uint32_t region_chunks;
  uint32_t pool_batch;    // chunks a wave reserves per atomic
};

constexpr uint32_t kPoolRegions = 256;
constexpr uint32_t kNoQuery = 0xffffffffu;

// depth % DIM for depth < 32, two bits per depth.
template <int DIM>
__device__ inline uint32_t depth_dim(uint32_t depth) {
  constexpr uint64_t tab = [] {
    uint64_t t = 0;
    for (int b = 0; b < 32; ++b) t |= (uint64_t)(b % DIM) << (2 * b);
    return t;
  }();
  return (uint32_t)(tab >> (2u * depth)) & 3u;
}

// Node records are read through a buffer resource (32-bit byte offsets, reads
// past num_records return zero) when the records fit in 4 GiB, else through a
// flat pointer.
struct NodeSource {
  const float* flat;
  __amdgpu_buffer_rsrc_t rsrc;
};

// Record i of the tree that starts at record `base` of the batch.  The
// buffer resource spans the whole batch (uniform across the wave), the
// per-lane tree offset goes into the 32-bit byte offset.
template <int DIM, bool BUF>
__device__ inline void load_rec(const NodeSource& src, uint32_t base, uint32_t i, float (&c)[DIM],
                                uint32_t& id) {
  constexpr int RF = NodeRec<DIM>::kFloats;
  float4 v;
  if (BUF)
    v = __builtin_bit_cast(float4,
                           __builtin_amdgcn_raw_buffer_load_b128(src.rsrc, (base + i) * (uint32_t)(RF * 4), 0, 0));
  else
    v = *reinterpret_cast<const float4*>(src.flat + ((uint64_t)base + i) * RF);
  c[0] = v.x;
  if (DIM > 1) c[DIM > 1 ? 1 : 0] = v.y;
  if (DIM > 2) c[DIM > 2 ? 2 : 0] = v.z;
  if (RF == 4) {
    id = __float_as_uint(v.w);
  } else {
    c[DIM > 3 ? 3 : 0] = v.w;
    if (BUF)
      id = __builtin_amdgcn_raw_buffer_load_b32(src.rsrc, (base + i) * (uint32_t)(RF * 4) + 16u, 0, 0);
    else
      id = __float_as_uint(src.flat[((uint64_t)base + i) * RF + 4]);
  }
}

enum SearchMode { kCount = 0, kFill = 1, kStats = 2, kPool = 3 };

constexpr int kStage = 16;  // ids staged per lane: one 64-byte line

// The loop body is branch free except for the emission of a hit in the fill
// pass: the next node is selected among (left son, right son, pending right
// son of the deepest pending depth) with conditional moves.
template <int DIM, int MODE, bool BUF>
__global__ __launch_bounds__(kThreads) void search_kernel(SearchArgs a) {
  constexpr bool FILL = MODE == kFill;
  constexpr bool POOL = MODE == kPool;
  constexpr bool STAGE = FILL || POOL;
  __shared__ uint32_t stage[STAGE ? kThreads : 1][kStage + 1];
  const uint32_t slot = blockIdx.x * kThreads + threadIdx.x;
  if (slot >= a.nq) return;
  const uint32_t q = a.perm ? a.perm[slot] : slot;
  uint32_t base = 0, n = a.n_single;
  if (a.off) {
    const uint32_t e = a.qev ? a.qev[q] : 0u;
    base = a.off[e];
    n = a.off[e + 1] - base;
  }
  float mn[DIM], mx[DIM];
#pragma unroll
  for (int m = 0; m < DIM; ++m) {
    mn[m] = a.mins[(uint64_t)m * a.nq + slot];
    mx[m] = a.maxs[(uint64_t)m * a.nq + slot];
  }
  NodeSource src;
  src.flat = a.nodes;
  if (BUF) src.rsrc = __builtin_amdgcn_make_buffer_rsrc((void*)a.nodes, (short)0, (int)a.node_bytes, 0x00020000);
  uint64_t w = FILL ? a.qoff[q] : 0;  // next output slot not yet written
  uint32_t nb = 0;                    // ids staged
  uint32_t c = 0;
  uint32_t* st = stage[STAGE ? threadIdx.x : 0];
  auto flush = [&]() {
    if (nb == kStage && a.out_aligned && (w & (kStage - 1)) == 0 && w + kStage <= a.cap) {
      uint4* dst = reinterpret_cast<uint4*>(a.out + w);
#pragma unroll
      for (int i = 0; i < kStage / 4; ++i)
        dst[i] = make_uint4(st[4 * i], st[4 * i + 1], st[4 * i + 2], st[4 * i + 3]);
    } else {
      for (uint32_t i = 0; i < nb; ++i)
        if (w + i < a.cap) a.out[w + i] = st[i];
    }
    w += nb;
    nb = 0;
  };
  uint32_t nchunk = 0;  // chunks of this query handed out so far (pool mode)
  const uint32_t lane = lane_id();
  const uint32_t wid = threadIdx.x / kWave;
  __shared__ uint32_t s_wbase[POOL ? kThreads / kWave : 1], s_wleft[POOL ? kThreads / kWave : 1];
  const uint32_t region = blockIdx.x % kPoolRegions;
  if (POOL && lane == 0) {
    s_wbase[wid] = 0;
    s_wleft[wid] = 0;
  }
  // Chunk allocation for the lanes with `need` set; called by every active
  // lane of the wave.  The wave's current reservation lives in LDS so that
  // lanes that already left the traversal loop see it too.
  auto pool_flush = [&](bool need) {
    const uint64_t m = __ballot(need);
    if (m == 0ull) return;
    const uint32_t cnt = (uint32_t)__popcll(m);
    const uint32_t leader = (uint32_t)__builtin_ctzll(m);
    uint32_t wbase = s_wbase[wid], wleft = s_wleft[wid];
    if (cnt > wleft) {
      const uint32_t grab = cnt > a.pool_batch ? cnt : a.pool_batch;
      uint32_t base = 0;
      if (lane == leader) {
        base = atomicAdd(&a.pool_ctr[region], grab);
        // mark the unused rest of the old reservation
        for (uint32_t i = 0; i < wleft; ++i) {
          const uint32_t k = wbase + i;
          if (k < a.region_chunks)
            a.chunk_meta[(uint64_t)region * a.region_chunks + k] = make_uint2(kNoQuery, 0u);
        }
      }
      wbase = __shfl(base, (int)leader);
      wleft = grab;
    }
    if (lane == leader) {
      s_wbase[wid] = wbase + cnt;
      s_wleft[wid] = wleft - cnt;
    }
    if (need) {
      const uint32_t k = wbase + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
      if (k < a.region_chunks) {
        const uint64_t ch = (uint64_t)region * a.region_chunks + k;
        uint4* dst = reinterpret_cast<uint4*>(a.pool + ch * kStage);
#pragma unroll
        for (int i = 0; i < kStage / 4; ++i)
          dst[i] = make_uint4(st[4 * i], st[4 * i + 1], st[4 * i + 2], st[4 * i + 3]);
        a.chunk_meta[ch] = make_uint2(slot, nchunk);
      }
      ++nchunk;
      nb = 0;
    }
  };
  uint32_t node = 0, depth = 0, dd = 0, pending = 0, vis = 0;
  bool alive = n != 0;
  while (alive) {
    if (MODE == kStats) ++vis;
    float x[DIM];
    uint32_t id;
    load_rec<DIM, BUF>(src, base, node, x, id);
    float v = x[0], lo = mn[0], hi = mx[0];
#pragma unroll
    for (int m = 1; m < DIM; ++m) {
      const bool sel = dd == (uint32_t)m;
      v = sel ? x[m] : v;
      lo = sel ? mn[m] : lo;
      hi = sel ? mx[m] : hi;
    }
    bool in = true;
#pragma unroll
    for (int m = 0; m < DIM; ++m) in = in & (x[m] <= mx[m]) & (x[m] >= mn[m]);
    const bool lower = v < lo;
    const bool inter = (v <= hi) & (v >= lo);
    if (FILL) {
      if (in) {
        st[nb++] = id;
        if (((w + nb) & (kStage - 1)) == 0) flush();
      }
    } else if (POOL) {
      if (in) st[nb++] = id;
      c += in ? 1u : 0u;
      pool_flush(nb == (uint32_t)kStage);
    } else {
      c += in ? 1u : 0u;
    }
    const uint32_t lc = 2u * node + 1u;
    const bool go_l = (!lower) & (lc < n);
    const bool go_r = (lower | inter) & (lc + 1u < n);
    const bool desc = go_l | go_r;
    pending |= (go_l & go_r) ? (2u << depth) : 0u;
    const uint32_t b = 31u - (uint32_t)__builtin_clz(pending | 1u);
    const uint32_t up = (node + 1u) >> (depth - b);
    alive = desc | (pending != 0u);
    node = go_l ? lc : (go_r ? lc + 1u : up);
    pending = desc ? pending : (pending & ~(1u << b));
    const uint32_t dn = (dd + 1u == (uint32_t)DIM) ? 0u : dd + 1u;
    dd = desc ? dn : depth_dim<DIM>(b);
    depth = desc ? depth + 1u : b;
  }
  if (FILL && nb) flush();
  if (POOL) {
    pool_flush(nb != 0u);
    // mark the unused rest of this wave's reservation (by one lane: the
    // wave may be partly empty)
    const uint32_t first = (uint32_t)__builtin_ctzll(__ballot(true));
    if (lane == first) {
      const uint32_t wb = s_wbase[wid], wl = s_wleft[wid];
      for (uint32_t i = 0; i < wl; ++i) {
        const uint32_t k = wb + i;
        if (k < a.region_chunks) a.chunk_meta[(uint64_t)region * a.region_chunks + k] = make_uint2(kNoQuery, 0u);
      }
    }
  }
  if (!FILL) a.counts[q] = c;
  if (POOL) a.counts_slot[slot] = c;
  if (MODE == kStats) a.visited[q] = vis;
}

// Copies the pooled chunks into the query-ordered CSR: slot t of chunk k
// (processing slot s, k-th chunk j of its query) goes to
// out[soff[s] + 16 j + t] when 16 j + t < counts[s]; soff / counts are in
// processing order, so neighbouring chunks read neighbouring entries.  Region r's used chunks are its first
// min(pool_ctr[r], region_chunks); reserved but unused ones carry kNoQuery.
// Every block forms the prefix sum of the used chunks per region in LDS and
// walks only used chunks, 16 threads per chunk.
__global__ __launch_bounds__(kThreads) void compact_kernel(const uint32_t* __restrict__ pool,
                                                           const uint2* __restrict__ meta,
                                                           const uint32_t* __restrict__ pool_ctr,
                                                           uint32_t region_chunks, uint32_t nq,
                                                           const uint32_t* __restrict__ counts,
                                                           const uint64_t* __restrict__ soff,
                                                           uint32_t* __restrict__ out, uint64_t cap) {
  static_assert(kPoolRegions == kThreads, "one thread per region");
  __shared__ uint32_t pre[kPoolRegions + 1];
  __shared__ uint32_t s_wave[kThreads / kWave];
  const uint32_t used = min(pool_ctr[threadIdx.x], region_chunks);
  const uint32_t incl = wave_incl_sum(used);
  const int wave = threadIdx.x / kWave;
  if (lane_id() == kWave - 1) s_wave[wave] = incl;
  __syncthreads();
  uint32_t woff = 0;
#pragma unroll
  for (int w = 0; w < kThreads / kWave; ++w) woff += (w < wave) ? s_wave[w] : 0u;
  pre[threadIdx.x + 1] = woff + incl;
  if (threadIdx.x == 0) pre[0] = 0;
  __syncthreads();
  const uint32_t chunks = pre[kPoolRegions];
  for (uint32_t kc = blockIdx.x * kThreads + threadIdx.x; kc < chunks; kc += gridDim.x * kThreads) {
    uint32_t lo = 0, hi = kPoolRegions;  // largest r with pre[r] <= kc
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (pre[mid] <= kc) lo = mid; else hi = mid;
    }
    const uint64_t k = (uint64_t)lo * region_chunks + (kc - pre[lo]);
    const uint2 mq = meta[k];
    const uint4* src = reinterpret_cast<const uint4*>(pool + k * kStage);
    uint4 v[kStage / 4];
#pragma unroll
    for (int i = 0; i < kStage / 4; ++i) v[i] = src[i];
    if (mq.x >= nq) continue;  // kNoQuery: reserved but unused
    const uint32_t pos = mq.y * kStage;
    const uint32_t cnt = counts[mq.x];
    const uint64_t dst = soff[mq.x] + pos;
    const uint32_t m = cnt > pos ? min(cnt - pos, (uint32_t)kStage) : 0u;
    const uint32_t* ids = reinterpret_cast<const uint32_t*>(v);
#pragma unroll
    for (int i = 0; i < kStage; ++i)
      if ((uint32_t)i < m && dst + i < cap) out[dst + i] = ids[i];
  }
}

// Offsets in processing order: soff[i] = qoff[perm[i]].
__global__ void gather_offsets_kernel(const uint64_t* __restrict__ qoff, const uint32_t* __restrict__ perm,
                                      uint32_t nq, uint64_t* __restrict__ soff) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nq; i += gridDim.x * blockDim.x)
    soff[i] = qoff[perm[i]];
}

// Spatial sort key of every query: Morton interleave of its box centre,
// quantised to kb bits per dimension inside the tree's bounding box.
// Any key order gives the same results; this one only groups similar work.
template <int DIM>
__global__ void query_keys_kernel(const float* __restrict__ mins, const float* __restrict__ maxs,
                                  uint32_t nq, const uint32_t* __restrict__ bbox, uint32_t kb,
                                  uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  float lo[DIM], sc[DIM];
  const float cells = (float)(1u << kb);
#pragma unroll
  for (int m = 0; m < DIM; ++m) {
    const float a = fkey_inv(bbox[m]), b = fkey_inv(bbox[DIM + m]);
    const float ext = b - a;
    lo[m] = a;
    sc[m] = (ext > 0.f && ext < 3.0e38f) ? cells / ext : 0.f;
    if (!(lo[m] > -3.0e38f && lo[m] < 3.0e38f)) sc[m] = 0.f;
  }
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nq; i += gridDim.x * blockDim.x) {
    uint32_t key = 0;
    uint32_t u[DIM];
#pragma unroll
    for (int m = 0; m < DIM; ++m) {
      const float c = 0.5f * mins[(uint64_t)m * nq + i] + 0.5f * maxs[(uint64_t)m * nq + i];
      float f = (c - lo[m]) * sc[m];
      f = (f >= 0.f) ? f : 0.f;  // also maps NaN to 0
      f = f < cells - 1.f ? f : cells - 1.f;
      u[m] = (uint32_t)f;
    }
    for (uint32_t b = 0; b < kb; ++b)
#pragma unroll
      for (int m = 0; m < DIM; ++m) key |= ((u[m] >> b) & 1u) << (b * DIM + m);
    keys[i] = key;
    vals[i] = i;
  }
}

// Boxes in processing order: smin[d][i] = mins[d][perm[i]].
template <int DIM>
__global__ void permute_boxes_kernel(const float* __restrict__ mins, const float* __restrict__ maxs,
                                     const uint32_t* __restrict__ perm, uint32_t nq,
                                     float* __restrict__ smins, float* __restrict__ smaxs) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nq; i += gridDim.x * blockDim.x) {
    const uint32_t q = perm[i];
#pragma unroll
    for (int m = 0; m < DIM; ++m) {
      smins[(uint64_t)m * nq + i] = mins[(uint64_t)m * nq + q];
      smaxs[(uint64_t)m * nq + i] = maxs[(uint64_t)m * nq + q];
    }
  }
}

}  // namespace fkd
