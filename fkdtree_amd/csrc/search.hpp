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
// heap index of that son follows from the current index by a shift (Walk).
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
  const float* recs;      // node records, NodeRec<DIM>::kFloats floats each
  const uint32_t* off;    // event -> first node (n_events + 1), or null
  uint32_t n_events;      // events of off (a query naming another event matches nothing)
  uint32_t n_single;      // node count when off == null
  const uint32_t* qev;    // query -> event, or null (event 0)
  const uint32_t* perm;   // processing slot -> query (null: identity)
  const uint32_t* block_order;  // dispatch position -> search block (null: identity)
  const uint32_t* slot_order;   // dispatch slot -> processing slot (null: identity)
  uint32_t nq;
  const float* mins;      // [DIM][nq], indexed by processing slot
  const float* maxs;
  uint32_t* counts;       // count pass output (query order)
  uint32_t* work_slot;    // count pass: nodes visited per query in processing order, or null
  uint32_t* visited;      // stats pass: nodes visited per query
  const uint64_t* qoff;   // fill pass: first output slot of every query
  uint32_t* out;          // fill pass output
  uint64_t cap;           // capacity of out
  uint32_t out_aligned;   // out is 64-byte aligned
  uint32_t node_bytes;    // size of the node records (buffer resource range)
  uint32_t bucket_park;   // bucketed passes: parked eighths of the live lanes that start the bucket tests
};

// Workgroups are handed to the 8 XCDs in turn.  Workgroup b of a search
// launch takes position start(b % 8) + b / 8, so that every XCD walks one
// contiguous range of the spatially ordered queries and its L2 holds that
// range's part of the tree; `order` (if any) maps positions to blocks.
__device__ inline uint32_t search_block(uint32_t b, uint32_t nb, const uint32_t* order) {
  const uint32_t j = xcd_chunk_start(b & 7u, nb) + (b >> 3);
  return order ? order[j] : j;
}

// Node records are read through a buffer resource (32-bit byte offsets,
// reads past num_records return zero) when the records fit in 4 GiB, else
// through a flat pointer.  The resource spans the whole batch (uniform across
// the wave); the per-lane record index goes into the byte offset.
struct NodeSource {
  const float* flat;
  __amdgpu_buffer_rsrc_t rsrc;
};

// Rotated coordinates and the id bits of record `rec`.
template <int DIM, bool BUF>
__device__ inline void load_rec(const NodeSource& src, uint32_t rec, float (&c)[DIM], uint32_t& id) {
  constexpr int RF = NodeRec<DIM>::kFloats;
  float4 q;
  if (BUF)
    q = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(src.rsrc, rec * (uint32_t)(RF * 4), 0, 0));
  else
    q = *reinterpret_cast<const float4*>(src.flat + (uint64_t)rec * RF);
  c[0] = q.x;
  if (DIM > 1) c[DIM > 1 ? 1 : 0] = q.y;
  if (DIM > 2) c[DIM > 2 ? 2 : 0] = q.z;
  if (RF == 4) {
    id = __float_as_uint(q.w);
  } else {
    c[DIM > 3 ? 3 : 0] = q.w;
    if (BUF)
      id = __builtin_amdgcn_raw_buffer_load_b32(src.rsrc, rec * (uint32_t)(RF * 4) + 16u, 0, 0);
    else
      id = __float_as_uint(src.flat[(uint64_t)rec * RF + NodeRec<DIM>::kId]);
  }
}

enum SearchMode { kCount = 0, kFill = 1, kStats = 2 };

// Depth-first walk state.  k is the current node's heap index counted from 1
// (record k - 1; sons 2k and 2k + 1; depth 31 - clz(k)).  For j below the
// current depth, bit j of p is set when the right sibling of the current
// node's ancestor j depths up (j = 0: the node itself) is still to be
// visited; that ancestor is then a left son.  Bit `depth` of p is always set
// (a sentinel standing for the root, which has no sibling).  Descending
// shifts k and p left by one; returning to the deepest pending right son
// shifts both right by the lowest set bit t of p: that son is (k >> t) + 1
// and its own bit is bit 0 of p >> t.  When only the sentinel is left, the
// same shift leads back to the root (k = 1): the walk is over.
struct Walk {
  uint32_t k = 1, p = 1;
  // After testing node k: on to its left son if go_l (its right son pending
  // too if go_r), else to its right son if go_r, else to the deepest pending
  // right son.  False when none is left.  go_l and go_r come as the wave's
  // lane masks (ballots), so that they combine on the scalar unit, and each
  // of the son's k and p is one add with carry in (2k + carry), which the
  // compiler would otherwise form from a select and a shift.
  __device__ inline bool next(uint64_t go_l, uint64_t go_r) {
    const bool desc = __builtin_amdgcn_inverse_ballot_w64(go_l | go_r);
    const uint32_t t = (uint32_t)__builtin_ctz(p);  // p != 0 (sentinel)
    uint32_t dk, dp;
    uint64_t co;
    asm("v_addc_co_u32 %0, %1, %2, %2, %3" : "=v"(dk), "=s"(co) : "v"(k), "s"(~go_l));
    asm("v_addc_co_u32 %0, %1, %2, %2, %3" : "=v"(dp), "=s"(co) : "v"(p), "s"(go_l & go_r));
    const uint32_t uk = (k >> t) | 1u;
    const uint32_t up = (p >> t) ^ 1u;
    k = desc ? dk : uk;
    p = desc ? dp : up;
    return desc || uk != 1u;
  }
  // After testing node k, whose split value x gave ge = (x >= lo) and
  // le = (x <= hi) against the bounds of its dimension: the reference's rule
  // (left son unless x < lo, right son if x < lo or x <= hi, a son only if it
  // exists: k < tl, k < tr).  Neither x nor lo is NaN, so x < lo is !ge.
  __device__ inline bool next(bool ge, bool le, uint32_t tl, uint32_t tr) {
    const uint64_t g = __builtin_amdgcn_ballot_w64(ge), l = __builtin_amdgcn_ballot_w64(le);
    const uint64_t el = __builtin_amdgcn_ballot_w64(k < tl), er = __builtin_amdgcn_ballot_w64(k < tr);
    return next(g & el, (~g | l) & er);
  }
  // Leaves node k without visiting its sons.
  __device__ inline bool pop() { return next(0ull, 0ull); }
};

// Byte offset, in a lane's box (the pairs of dimensions 0 .. DIM-1, then
// 0 .. DIM-2), of the bounds rotated like the records of depth 31 - c (split
// dimension first): the rotation tables below are indexed by c = clz(k), so
// the walk needs no depth variable.
template <int DIM>
__device__ inline uint8_t rot_of_clz(uint32_t c) {
  return (uint8_t)(((31u - c) % (uint32_t)DIM) * (uint32_t)sizeof(float2));
}

// Fill staging: hits go to a ring of kRing ids per lane in LDS, ring index
// i holding the id for output slot wb + i (wb = the query's first output slot
// rounded down to 4 ids) at ring slot i % kRing.  Whenever some lane's ring
// is full, every lane writes its complete 4-id groups to `out` as 16-byte
// stores, so that stores are issued a few times per wave instead of once per
// lane and line.
constexpr int kRing = 16;
constexpr int kRingRow = 16;  // words per lane (16-byte aligned rows)
constexpr int kBucketQueue = 8;       // bucket roots queued per lane in the fill pass (bucketed traversal)
constexpr int kBucketQueueCount = 4;  // in the count pass (its LDS allows 8 waves per SIMD)

// The loop body is branch free except for writing out staged ids: the next
// node is selected among (left son, right son, pending right son of the
// deepest pending depth) with conditional moves.  Every lane keeps its box
// in LDS as the (min, max) pairs of dimensions 0 .. DIM-1, 0 .. DIM-2 (2 DIM - 1
// pairs, an odd stride, so the lanes' reads avoid bank conflicts), so the box
// rotated like the node records of depth d (split dimension first) is the DIM
// pairs starting at pair d % DIM, found through a 32-entry table: no
// selection by dimension in the loop.
// WK (count pass): the work figure each box reports in work_slot for the
// fill's window sort and block order: 0 none (the fill sorts nothing, and the
// loop keeps no counter), 1 nodes visited, 2 the work of the bucketed fill (a
// step per node down to the bucket depth B, none inside buckets; B as in
// bucket_search_kernel), so that the window sort balances the bucketed walk.
// The fill pass asks for the waves per SIMD its LDS allows: 6 blocks a CU up
// to dim 3, 5 above.
template <int DIM, int MODE, bool BUF, int WK>
__global__ __launch_bounds__(kThreads)
__attribute__((amdgpu_waves_per_eu(MODE == kFill ? (DIM <= 3 ? 6 : 5) : 1)))
void search_kernel(SearchArgs a) {
  constexpr bool FILL = MODE == kFill;
  constexpr bool LDS_BOX = DIM >= 2;
  constexpr int BS = 2 * DIM - 1;  // pairs per lane
  __shared__ __align__(16) uint32_t ring[FILL ? kThreads * kRingRow : 1];
  __shared__ float2 s_box[LDS_BOX ? kThreads * BS : 1];
  __shared__ uint8_t s_rot[LDS_BOX ? 32 : 1];  // clz(k) -> byte offset of its rotation
  if (LDS_BOX) {
    if (threadIdx.x < 32u) s_rot[threadIdx.x] = rot_of_clz<DIM>(threadIdx.x);
    __syncthreads();
  }
  uint32_t slot = search_block(blockIdx.x, gridDim.x, a.block_order) * kThreads + threadIdx.x;
  if (slot >= a.nq) return;
  if (a.slot_order) slot = a.slot_order[slot];
  const uint32_t q = a.perm ? a.perm[slot] : slot;
  uint32_t base = 0, n = a.n_single;
  if (a.off) {
    const uint32_t e = a.qev ? a.qev[q] : 0u;
    if (e < a.n_events) {
      base = a.off[e];
      n = a.off[e + 1] - base;
    } else {
      n = 0;
    }
  }
  float mn[DIM], mx[DIM];
  float2* const my_box = &s_box[LDS_BOX ? threadIdx.x * BS : 0];
#pragma unroll
  for (int m = 0; m < DIM; ++m) {
    mn[m] = a.mins[(uint64_t)m * a.nq + slot];
    mx[m] = a.maxs[(uint64_t)m * a.nq + slot];
    // A NaN lower bound matches nothing and sends the walk left only; with
    // (-inf, NaN) the two compares below give exactly that.
    if (mn[m] != mn[m]) {
      mn[m] = -__builtin_huge_valf();
      mx[m] = __builtin_nanf("");
    }
    if (LDS_BOX) {
      my_box[m] = make_float2(mn[m], mx[m]);
      if (m + 1 < DIM) my_box[m + DIM] = make_float2(mn[m], mx[m]);
    }
  }
  NodeSource src;
  src.flat = a.recs;
  if (BUF) src.rsrc = __builtin_amdgcn_make_buffer_rsrc((void*)a.recs, (short)0, (int)a.node_bytes, 0x00020000);
  uint32_t c = 0;
  uint32_t* rw = &ring[FILL ? threadIdx.x * kRingRow : 0];
  const uint64_t w0 = FILL ? a.qoff[q] : 0;
  const uint64_t wb = w0 & ~(uint64_t)3;
  const uint32_t i0 = (uint32_t)w0 & 3u;  // first ring index of this query
  uint32_t wr = i0;                       // next ring index
  uint32_t fq = 0;                        // first ring index not written out (multiple of 4)
  uint32_t wlim = kRing;                  // ring index at which the ring is full
  // Writes the complete 4-id groups; the query's first group is partial when
  // its first slot is not 16-byte aligned.
  auto write_groups = [&]() {
    while (fq + 4u <= wr) {
      const uint4 g = *reinterpret_cast<const uint4*>(rw + (fq & (kRing - 1)));
      const uint64_t p = wb + fq;
      if (fq >= i0 && a.out_aligned && p + 4u <= a.cap) {
        *reinterpret_cast<uint4*>(a.out + p) = g;
      } else {
        const uint32_t e[4] = {g.x, g.y, g.z, g.w};
#pragma unroll
        for (uint32_t j = 0; j < 4u; ++j)
          if (fq + j >= i0 && p + j < a.cap) a.out[p + j] = e[j];
      }
      fq += 4u;
    }
  };
  // node k has a left (right) son iff 2k (2k + 1) <= n, i.e. k < tl (tr)
  const uint32_t tl = (n >> 1) + 1u, tr = ((n - 1u) >> 1) + 1u;
  const uint32_t rb = base - 1u;  // record of node k: rb + k
  Walk w;
  uint32_t vis = 0;
  bool alive = n != 0;
  // WK 2: nodes on depths 0 .. B count, i.e. clz(k) >= 31 - B
  uint32_t wb_clz = 0;
  if (MODE == kCount && WK == 2) {
    const uint32_t F = 31u - (uint32_t)__builtin_clz(n + 1u);
    wb_clz = F > 3u ? 34u - F : (F == 3u ? 31u : 0u);
  }
  while (alive) {
    const uint32_t cz = (uint32_t)__builtin_clz(w.k);
    if (MODE == kStats || (MODE == kCount && WK == 1))
      ++vis;
    else if (MODE == kCount && WK == 2)
      vis += cz >= wb_clz ? 1u : 0u;
    float x[DIM], lo[DIM], hi[DIM];
    uint32_t id;
    load_rec<DIM, BUF>(src, rb + w.k, x, id);
    if (LDS_BOX) {
      const float2* bx = reinterpret_cast<const float2*>(reinterpret_cast<const char*>(my_box) + s_rot[cz]);
#pragma unroll
      for (int j = 0; j < DIM; ++j) {
        const float2 t = bx[j];
        lo[j] = t.x;
        hi[j] = t.y;
      }
    } else {
      lo[0] = mn[0];
      hi[0] = mx[0];
    }
    // x[0] is the split value and lo[0], hi[0] the bounds on the split
    // dimension.  The reference's rule (left son unless the split value is
    // below the box, right son if it is below or inside) from two compares:
    // neither x (points hold no NaN) nor lo[0] (a NaN lower bound became
    // -inf) is NaN, so "below" (x < lo) is !(x >= lo), and below || inside
    // is !(x >= lo) || (x <= hi).
    const bool ge = x[0] >= lo[0], le = x[0] <= hi[0];
    bool in = ge & le;
#pragma unroll
    for (int j = 1; j < DIM; ++j) in = in & (x[j] <= hi[j]) & (x[j] >= lo[j]);
    if (FILL) {
      rw[wr & (kRing - 1)] = id;  // the slot is free; it is kept only for a hit
      wr += in ? 1u : 0u;
      if (__ballot(wr == wlim) != 0ull) {
        write_groups();
        wlim = fq + (uint32_t)kRing;
      }
    } else {
      c += in ? 1u : 0u;
    }
    alive = w.next(ge, le, tl, tr);
  }
  if (FILL) {
    write_groups();
    for (uint32_t j = fq > i0 ? fq : i0; j < wr; ++j)
      if (wb + j < a.cap) a.out[wb + j] = rw[j & (kRing - 1)];
  }
  if (!FILL) a.counts[q] = c;
  if (MODE == kCount && WK != 0) a.work_slot[slot] = vis;
  if (MODE == kStats) a.visited[q] = vis;
}

// Bucketed traversal for the count and fill passes (dim <= 3).
//
// The hits of a box are the points inside it, whichever nodes are visited to
// find them: the push rule above only skips subtrees that cannot hold one
// (a left subtree has x[d] <= split, a right one x[d] >= split).  Most of a
// box's visits are on the bottom depths, where its hits are and where pruning
// saves little.  So the walk stops at depth B = F - 3 (F = number of full
// depths of the tree) and every subtree it reaches there, a bucket of the 7
// nodes on depths B..B+2 and up to 8 on the partial depth B+3, is tested node
// by node with independent loads (about 9 VALU instructions a node against
// about 30 for a walk step).  The visited-count statistics (kStats) keep the
// reference walk.
//
// Two schedules (BQ):
//  * BQ = 0, parking: a lane that reaches a bucket waits until `bucket_park`
//    eighths of the wave's live lanes wait or none can walk; then the waiting
//    lanes test their buckets together.
//  * BQ > 0, queueing: a lane puts the bucket roots it reaches into its own
//    LDS queue of BQ entries and walks on; every iteration issues the load of
//    its walk step and the loads of its oldest queued bucket together, so one
//    memory latency covers both (a lane whose queue is full waits).
// Occupancy: the parked count pass through a buffer resource fits 8 waves
// per SIMD without spilling (the flat-pointer one would spill); the parked
// fill pass 6, which its LDS (ring and box: 26.6 KB a block) also allows; the
// queued passes 4.
template <int DIM, int MODE, bool BUF, int BQ>
__global__ __launch_bounds__(kThreads)
__attribute__((amdgpu_waves_per_eu(BQ > 0 ? 4 : (MODE == kCount ? (BUF ? 8 : 4) : 6))))
void bucket_search_kernel(SearchArgs a) {
  static_assert(DIM >= 1 && DIM <= 3, "one 16-byte record per node");
  static_assert(MODE == kCount || MODE == kFill, "count and fill passes only");
  constexpr bool FILL = MODE == kFill;
  constexpr int BS = 2 * DIM - 1;  // box pairs per lane, as in search_kernel
  __shared__ __align__(16) uint32_t ring[FILL ? kThreads * kRingRow : 1];
  __shared__ float2 s_box[kThreads * BS];
  __shared__ uint8_t s_rot[32];  // clz(k) -> byte offset of its rotation
  __shared__ uint32_t s_bq[BQ > 0 ? BQ * kThreads : 1];  // queue entry k of thread t at [k * kThreads + t]
  if (threadIdx.x < 32u) s_rot[threadIdx.x] = rot_of_clz<DIM>(threadIdx.x);
  __syncthreads();
  uint32_t slot = search_block(blockIdx.x, gridDim.x, a.block_order) * kThreads + threadIdx.x;
  if (slot >= a.nq) return;
  if (a.slot_order) slot = a.slot_order[slot];
  const uint32_t q = a.perm ? a.perm[slot] : slot;
  uint32_t base = 0, n = a.n_single;
  if (a.off) {
    const uint32_t e = a.qev ? a.qev[q] : 0u;
    if (e < a.n_events) {
      base = a.off[e];
      n = a.off[e + 1] - base;
    } else {
      n = 0;
    }
  }
  float2* const my_box = &s_box[threadIdx.x * BS];
#pragma unroll
  for (int m = 0; m < DIM; ++m) {
    float lo = a.mins[(uint64_t)m * a.nq + slot];
    float hi = a.maxs[(uint64_t)m * a.nq + slot];
    if (lo != lo) {  // a NaN lower bound: as in search_kernel
      lo = -__builtin_huge_valf();
      hi = __builtin_nanf("");
    }
    my_box[m] = make_float2(lo, hi);
    if (m + 1 < DIM) my_box[m + DIM] = make_float2(lo, hi);
  }
  // F full depths (n >= 2^F - 1); buckets are rooted on depth B.  A tree of
  // fewer than 7 nodes (F < 3) has no full bucket and is walked (B = 32 is
  // never reached).
  const uint32_t F = 31u - (uint32_t)__builtin_clz(n + 1u);
  const uint32_t B = F > 3u ? F - 3u : (F == 3u ? 0u : 32u);
  const uint32_t cB = 31u - B;  // clz(k) of the nodes on depth B
  // The bounds compared with slot j of a record on depth d are the pair at
  // byte s_rot[31 - d] + 8 j of my_box (dimension (d + j) % DIM).  The fill pass
  // keeps those of depths B .. B + DIM - 1 in registers (bk); the count pass
  // reads them per use, so they take no registers across its loop.
  constexpr bool kRegBounds = FILL;
  float2 bk[kRegBounds ? DIM : 1][DIM];
  if (kRegBounds) {
    const uint32_t r0 = B % (uint32_t)DIM;
#pragma unroll
    for (int k = 0; k < (kRegBounds ? DIM : 1); ++k) {
      uint32_t r = r0 + (uint32_t)k;
      r = r >= (uint32_t)DIM ? r - (uint32_t)DIM : r;
#pragma unroll
      for (int j = 0; j < DIM; ++j) bk[k][j] = my_box[r + (uint32_t)j];
    }
  }
  // bounds of depth B + k
  auto bounds = [&](uint32_t k, float2 (&bb)[DIM]) {
    if (kRegBounds) {
#pragma unroll
      for (int j = 0; j < DIM; ++j) bb[j] = bk[kRegBounds ? k % (uint32_t)DIM : 0][j];
    } else {
      const float2* bx =
          reinterpret_cast<const float2*>(reinterpret_cast<const char*>(my_box) + s_rot[(cB - k) & 31u]);
#pragma unroll
      for (int j = 0; j < DIM; ++j) bb[j] = bx[j];
    }
  };
  NodeSource src;
  src.flat = a.recs;
  if (BUF) src.rsrc = __builtin_amdgcn_make_buffer_rsrc((void*)a.recs, (short)0, (int)a.node_bytes, 0x00020000);
  uint32_t c = 0;
  uint32_t* rw = &ring[FILL ? threadIdx.x * kRingRow : 0];
  const uint64_t w0 = FILL ? a.qoff[q] : 0;
  const uint64_t wb = w0 & ~(uint64_t)3;
  const uint32_t i0 = (uint32_t)w0 & 3u;
  uint32_t wr = i0;  // next ring index
  uint32_t fq = 0;   // first ring index not written out (multiple of 4)
  auto write_groups = [&]() {
    while (fq + 4u <= wr) {
      const uint4 g = *reinterpret_cast<const uint4*>(rw + (fq & (kRing - 1)));
      const uint64_t p = wb + fq;
      if (fq >= i0 && a.out_aligned && p + 4u <= a.cap) {
        *reinterpret_cast<uint4*>(a.out + p) = g;
      } else {
        const uint32_t e[4] = {g.x, g.y, g.z, g.w};
#pragma unroll
        for (uint32_t j = 0; j < 4u; ++j)
          if (fq + j >= i0 && p + j < a.cap) a.out[p + j] = e[j];
      }
      fq += 4u;
    }
  };
  // Keeps room for 8 more ids in every lane's ring: when some lane has more
  // than 8 pending, the lanes holding at least 8 write their complete groups
  // (32 bytes or more each; lanes with fewer keep theirs for a longer run).
  auto make_room = [&]() {
    if (FILL && __ballot(wr - fq > 8u) != 0ull) {
      if (wr - fq >= 8u) write_groups();
    }
  };
  // a hit candidate: kept (fill) or counted when `in`
  auto report = [&](uint32_t id, bool in) {
    if (FILL) {
      rw[wr & (kRing - 1)] = id;  // the slot is free; it is kept only for a hit
      wr += in ? 1u : 0u;
    } else {
      c += in ? 1u : 0u;
    }
  };
  // one record: the closed-box test against the bounds of its depth
  auto inside = [&](const float4 r, const float2 (&bb)[DIM]) -> bool {
    bool in = (r.x >= bb[0].x) & (r.x <= bb[0].y);
    if (DIM > 1) in = in & (r.y >= bb[DIM > 1 ? 1 : 0].x) & (r.y <= bb[DIM > 1 ? 1 : 0].y);
    if (DIM > 2) in = in & (r.z >= bb[DIM > 2 ? 2 : 0].x) & (r.z <= bb[DIM > 2 ? 2 : 0].y);
    return in;
  };
  // a node record; the count pass needs only the coordinates (12 bytes)
  auto rec16 = [&](uint32_t rec) -> float4 {
    if (FILL) {
      if (BUF) return __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(src.rsrc, rec * 16u, 0, 0));
      return reinterpret_cast<const float4*>(a.recs)[rec];
    }
    if (BUF) {
      const auto v = __builtin_amdgcn_raw_buffer_load_b96(src.rsrc, rec * 16u, 0, 0);
      return make_float4(__uint_as_float(v[0]), __uint_as_float(v[1]), __uint_as_float(v[2]), 0.f);
    }
    const float* f = a.recs + (uint64_t)rec * 4u;
    return make_float4(f[0], f[1], f[2], 0.f);
  };
  const uint32_t rb = base - 1u;  // record of node k (heap index from 1): rb + k
  // depths B, B+1, B+2 of the bucket rooted at node k: k; 2k, 2k+1;
  // 4k .. 4k+3 (all exist)
  auto load7 = [&](uint32_t k, float4 (&v)[7]) {
    const uint32_t r1 = rb + 2u * k, r2 = rb + 4u * k;
    v[0] = rec16(rb + k);
#pragma unroll
    for (int i = 0; i < 2; ++i) v[1 + i] = rec16(r1 + (uint32_t)i);
#pragma unroll
    for (int i = 0; i < 4; ++i) v[3 + i] = rec16(r2 + (uint32_t)i);
  };
  auto test7 = [&](const float4 (&v)[7]) {
    float2 bb[DIM];
    bounds(0u, bb);
    report(__float_as_uint(v[0].w), inside(v[0], bb));
    bounds(1u, bb);
    report(__float_as_uint(v[1].w), inside(v[1], bb));
    report(__float_as_uint(v[2].w), inside(v[2], bb));
    bounds(2u, bb);
#pragma unroll
    for (int i = 3; i < 7; ++i) report(__float_as_uint(v[i].w), inside(v[i], bb));
  };
  // depth B + 3 = F under node k: nodes 8k .. 8k+7 that exist (a prefix of the depth)
  auto last_depth = [&](uint32_t k) {
    const uint32_t l3 = 8u * k;
    if (l3 <= n) {
      make_room();
      float4 u[8];
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const uint32_t j = l3 + (uint32_t)i;
        u[i] = rec16(rb + (j < n ? j : n));
      }
      float2 bb[DIM];
      bounds(3u, bb);
#pragma unroll
      for (int i = 0; i < 8; ++i) report(__float_as_uint(u[i].w), (l3 + (uint32_t)i <= n) & inside(u[i], bb));
    }
  };
  // node k has a left (right) son iff 2k (2k + 1) <= n, i.e. k < tl (tr)
  const uint32_t tl = (n >> 1) + 1u, tr = ((n - 1u) >> 1) + 1u;
  Walk w;
  uint32_t work = 0;
  bool alive = n != 0;
  auto at_bucket = [&]() { return (uint32_t)__builtin_clz(w.k) == cB; };
  // one step of the reference walk from node w.k (not a bucket root)
  auto walk_step = [&]() {
    ++work;
    float x[DIM];
    uint32_t id;
    load_rec<DIM, BUF>(src, rb + w.k, x, id);
    const float2* bx =
        reinterpret_cast<const float2*>(reinterpret_cast<const char*>(my_box) + s_rot[__builtin_clz(w.k)]);
    float lo[DIM], hi[DIM];
#pragma unroll
    for (int j = 0; j < DIM; ++j) {
      const float2 t = bx[j];
      lo[j] = t.x;
      hi[j] = t.y;
    }
    // the walk kernel's rule (neither x nor lo[0] is NaN)
    const bool ge = x[0] >= lo[0], le = x[0] <= hi[0];
    bool in = ge & le;
#pragma unroll
    for (int j = 1; j < DIM; ++j) in = in & (x[j] <= hi[j]) & (x[j] >= lo[j]);
    report(id, in);
    alive = w.next(ge, le, tl, tr);
  };
  // leave the bucket root w.k: on to the deepest pending right son
  auto pop = [&]() { alive = w.pop(); };
  if (BQ == 0) {
    bool parked = alive && B == 0u;
    const uint32_t park8 = a.bucket_park;
    for (;;) {
      if (__ballot(alive) == 0ull) break;
      // room for one walk step and a bucket's first 7 nodes
      make_room();
      if (alive && !parked) {
        walk_step();
        parked = alive && at_bucket();
      }
      const uint64_t pm = __ballot(parked);
      if (pm == 0ull) continue;
      const uint64_t wm = __ballot(alive && !parked);
      if (wm != 0ull && (uint32_t)__popcll(pm) * 8u < park8 * (uint32_t)__popcll(pm | wm)) continue;
      if (parked) {
        work += 2u;
        float4 v[7];
        load7(w.k, v);
        test7(v);
        last_depth(w.k);
        pop();
        parked = alive && at_bucket();
      }
    }
  } else {
    uint32_t* const my_q = &s_bq[BQ > 0 ? threadIdx.x : 0];
    uint32_t qh = 0, qt = 0;  // queue head and tail (entries qh .. qt-1 are pending)
    for (;;) {
      const bool has_b = qh != qt;
      if (__ballot(alive || has_b) == 0ull) break;
      // room for one walk step and a bucket's first 7 nodes
      make_room();
      // queue the bucket roots at the head of the walk while there is room
      while (alive && at_bucket() && qt - qh < (uint32_t)BQ) {
        my_q[(qt % (uint32_t)BQ) * kThreads] = w.k;
        ++qt;
        pop();
      }
      uint32_t r = 0;
      float4 v[7];
      if (has_b) {
        r = my_q[(qh % (uint32_t)BQ) * kThreads];
        load7(r, v);
      }
      if (alive && !at_bucket()) walk_step();
      if (has_b) {
        work += 2u;
        ++qh;
        test7(v);
        last_depth(r);
      }
    }
  }
  if (FILL) {
    write_groups();
    for (uint32_t j = fq > i0 ? fq : i0; j < wr; ++j)
      if (wb + j < a.cap) a.out[wb + j] = rw[j & (kRing - 1)];
  }
  if (!FILL) a.counts[q] = c;
  if (MODE == kCount && a.work_slot) a.work_slot[slot] = work;
}

// Low-latency search of a few boxes: one wave per box, breadth first.
//
// A lane-per-box walk is a chain of ~ visited-nodes dependent loads (280 on
// 1M mixed points), hundreds of microseconds on a GPU, so a caller that
// issues one box at a time (the reference's per-box search_in_the_box,
// main.cpp:351-359) needs another shape: the 64 lanes of a wave take the
// nodes of one depth of the box's frontier together, as the reference's
// level-synchronous FQueue walk does (FKDTree_cpu.h:126-167), so the chain
// is one load per depth.  The visited set and the reported ids are the
// reference's (push rule FKDTree_cpu.h:150-163, closed box :145-148).
constexpr uint32_t kWaveFrontier = 2048;
constexpr uint32_t kSmallOverflow = 0xffffffffu;
constexpr uint32_t kTopLevels = 10;
constexpr uint32_t kTopNodes = (1u << kTopLevels) - 1u;

// Copies the records of the top min(n, 2^TOP - 1) nodes to LDS with
// independent, coalesced loads (one wave).
template <int DIM, uint32_t TOP>
__device__ inline void wave_load_top(const float* recs, uint32_t n, float* top) {
  constexpr int RF = NodeRec<DIM>::kFloats;
  constexpr uint32_t kNodes = (1u << TOP) - 1u;
  const uint32_t ntop = n < kNodes ? n : kNodes;
  for (uint32_t i = lane_id(); i < ntop * (uint32_t)(RF / 4); i += kWave)
    reinterpret_cast<float4*>(top)[i] = reinterpret_cast<const float4*>(recs)[i];
}

// Set bits of a wave mask in the lanes below the calling lane.
__device__ inline uint32_t below_count(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Window ancestors of every lane, for each window size lw = 1..6 (64 >> lw
// frontier nodes of 2^lw lanes each): the lanes holding the ancestors from
// which the lane's node is a left son (x, y: low and high words) and a right
// son (z, w).  Filled once per workgroup; lw = 0 has no ancestors.
constexpr uint32_t kAncSizes = 7;
__device__ inline void wave_fill_ancestors(uint4* anc) {
  const uint32_t lane = lane_id();
  for (uint32_t lw = 0; lw < kAncSizes; ++lw) {
    uint64_t lm = 0, rm = 0;
    if (lw) {
      const uint32_t e = lane >> lw, j = lane & ((1u << lw) - 1u);
      const uint32_t lvl = j ? 31u - (uint32_t)__builtin_clz(j) : 0u;
      for (uint32_t a = 1; a <= lvl; ++a) {
        const uint64_t bit = 1ull << ((e << lw) | (j >> a));
        if ((j >> (a - 1u)) & 1u)
          rm |= bit;
        else
          lm |= bit;
      }
    }
    anc[lw * kWave + lane] = make_uint4((uint32_t)lm, (uint32_t)(lm >> 32), (uint32_t)rm, (uint32_t)(rm >> 32));
  }
}

// One box walked by the calling wave (a workgroup of exactly one wave: the
// rounds synchronise with __syncthreads).  The records of depths < TOP come
// from `top` (wave_load_top), the rest from `src`.  Hits go to out[i] for
// i < slot; returns the true number of hits, or kSmallOverflow when a
// stored frontier exceeds kWaveFrontier nodes.  A NaN lower bound is read as
// (-inf, NaN), as in search_kernel.
//
// Breadth first, several depths per round while the frontier is small: with
// f <= 16 frontier nodes each takes 2^K lanes (K = 6 - ceil(log2 f)) for the
// 2^K - 1 nodes of its subtree's top K depths, all loaded at once.  A lane's
// node is visited when every ancestor inside the window pushed towards it
// (the reference's push rule, tested on the ancestors' lanes through the
// wave's ballots of left and right pushes), so the visited set and the hits
// are exactly the depth-by-depth walk's; only the window's bottom depth
// writes the next frontier.  Wider frontiers take one depth per round, one
// node per lane.  A box around one point (a frontier of 1-4 nodes down to
// the last depths) needs about a third of the dependent rounds.  `anc` is
// the wave_fill_ancestors table: a lane's ancestor test is two masked
// compares of the push ballots.
template <int DIM, bool BUF, uint32_t TOP>
__device__ inline uint32_t wave_box_walk(const NodeSource& src, const float* top, uint32_t n, float (&mn)[DIM],
                                         float (&mx)[DIM], uint32_t (*fr)[kWaveFrontier], const uint4* anc,
                                         float2* rbox, uint32_t* out, uint32_t slot) {
  constexpr int RF = NodeRec<DIM>::kFloats;
  const uint32_t lane = lane_id();
#pragma unroll
  for (int m = 0; m < DIM; ++m)
    if (mn[m] != mn[m]) {
      mn[m] = -__builtin_huge_valf();
      mx[m] = __builtin_nanf("");
    }
  // the box rotated like the records of a depth d with d % DIM = r:
  // rbox[r * DIM + i] = (min, max) of dimension (r + i) % DIM
  // (compile-time indices only: a per-lane index into mn / mx would put
  // them in private memory)
  if (lane == 0) {
#pragma unroll
    for (int r = 0; r < DIM; ++r)
#pragma unroll
      for (int i = 0; i < DIM; ++i) rbox[r * DIM + i] = make_float2(mn[(r + i) % DIM], mx[(r + i) % DIM]);
  }
  const uint32_t tl = n >> 1, tr = (n - 1u) >> 1;
  const bool narrow = n <= (1u << 26);
  uint32_t nf = n ? 1u : 0u, cur = 0, hits = 0, depth = 0;
  bool overflow = false;
  if (lane == 0) fr[0][0] = 0;
  __syncthreads();
  while (nf) {
    // lw: log2 of the lanes per frontier node; K: depths this round
    const uint32_t lw = nf > 16u ? 0u : 6u - (nf > 1u ? 32u - (uint32_t)__builtin_clz(nf - 1u) : 0u);
    const uint32_t K = lw ? lw : 1u;
    const uint32_t per = kWave >> lw;  // frontier nodes per pass
    const uint32_t e = lane >> lw;
    const uint32_t j = lw ? (lane & ((1u << lw) - 1u)) : 1u;  // 1-based heap index inside the window
    const uint32_t lvl = j ? 31u - (uint32_t)__builtin_clz(j) : 0u;
    const uint32_t ndepth = depth + lvl;
    const uint4 am = anc[lw * kWave + lane];
    const uint64_t anc_l = (uint64_t)am.x | ((uint64_t)am.y << 32), anc_r = (uint64_t)am.z | ((uint64_t)am.w << 32);
    // the box rotated like this lane's records: slot i is dimension (ndepth + i) % DIM
    const float2* rb = rbox + (ndepth % (uint32_t)DIM) * (uint32_t)DIM;
    float lo[DIM], hi[DIM];
#pragma unroll
    for (int i = 0; i < DIM; ++i) {
      const float2 t = rb[i];
      lo[i] = t.x;
      hi[i] = t.y;
    }
    uint32_t nn = 0;
    for (uint32_t base = 0; base < nf; base += per) {
      const bool slot_ok = j != 0u && base + e < nf;
      const uint32_t root = slot_ok ? fr[cur][base + e] : 0u;
      // a window node below a deep root can pass 2^32 (such nodes are past
      // n); with n <= 2^26 none can (roots < 2^26, windows of 5 depths below)
      bool valid;
      uint32_t node;
      if (narrow) {
        node = ((root + 1u) << lvl) + j - (1u << lvl) - 1u;
        valid = slot_ok && node < n;
      } else {
        const uint64_t node64 = (((uint64_t)root + 1ull) << lvl) + (uint64_t)(j - (1u << lvl)) - 1ull;
        valid = slot_ok && node64 < (uint64_t)n;
        node = (uint32_t)node64;
      }
      node = valid ? node : 0u;
      float x[DIM];
      uint32_t id = 0;
#pragma unroll
      for (int i = 0; i < DIM; ++i) x[i] = 0.f;
      if (valid && ndepth < TOP) {
        const float* r = top + (uint64_t)node * RF;
#pragma unroll
        for (int i = 0; i < DIM; ++i) x[i] = r[i];
        id = __float_as_uint(r[NodeRec<DIM>::kId]);
      } else if (valid) {
        load_rec<DIM, BUF>(src, node, x, id);
      }
      const bool ge = x[0] >= lo[0], le = x[0] <= hi[0], lower = x[0] < lo[0];
      bool in = valid & ge & le;
#pragma unroll
      for (int i = 1; i < DIM; ++i) in = in & (x[i] <= hi[i]) & (x[i] >= lo[i]);
      const bool go_l = valid && ge && node < tl;
      const bool go_r = valid && (lower || le) && node < tr;
      const uint64_t gl = __ballot(go_l), gr = __ballot(go_r);
      // visited: every window ancestor pushed towards this lane's node
      const bool reach = valid && (gl & anc_l) == anc_l && (gr & anc_r) == anc_r;
      const bool hit = reach && in;
      const uint64_t hm = __ballot(hit);
      if (hit) {
        const uint32_t pos = hits + below_count(hm);
        if (pos < slot) out[pos] = id;
      }
      hits += (uint32_t)__popcll(hm);
      // the window's bottom depth writes the next frontier
      const bool bottom = reach && lvl + 1u == K;
      const bool pl = bottom && go_l, pr = bottom && go_r;
      const uint64_t lm = __ballot(pl), rm = __ballot(pr);
      const uint32_t at = nn + below_count(lm) + below_count(rm);
      const uint32_t nxt = cur ^ 1u;
      if (pl && at < kWaveFrontier) fr[nxt][at] = 2u * node + 1u;
      const uint32_t atr = at + (pl ? 1u : 0u);
      if (pr && atr < kWaveFrontier) fr[nxt][atr] = 2u * node + 2u;
      nn += (uint32_t)(__popcll(lm) + __popcll(rm));
    }
    __syncthreads();
    if (nn > kWaveFrontier) {
      overflow = true;
      break;
    }
    cur ^= 1u;
    nf = nn;
    depth += K;
  }
  return overflow ? kSmallOverflow : hits;
}

// fkd_search_small: box b walked by workgroup b.  Boxes are read as
// [nb][2 * DIM] (mins, then maxs); hits go to ids[b * slot + i] for i < slot
// and counts[b] is wave_box_walk's result.  Boxes and outputs live in
// page-locked host memory: no copies before or after the launch.
template <int DIM, bool BUF>
__global__ __launch_bounds__(kWave) void box_wave_kernel(const float* recs, uint32_t node_bytes, uint32_t n,
                                                         const float* boxes, uint32_t nb, uint32_t slot,
                                                         uint32_t* counts, uint32_t* ids) {
  constexpr int RF = NodeRec<DIM>::kFloats;
  __shared__ uint32_t fr[2][kWaveFrontier];
  __shared__ uint4 anc[kAncSizes * kWave];
  __shared__ float2 rbox[DIM * DIM];
  __shared__ __align__(16) float top[kTopNodes * RF];
  const uint32_t b = blockIdx.x;
  if (b >= nb) return;
  wave_load_top<DIM, kTopLevels>(recs, n, top);
  wave_fill_ancestors(anc);
  float mn[DIM], mx[DIM];
#pragma unroll
  for (int m = 0; m < DIM; ++m) {
    mn[m] = boxes[(uint64_t)b * 2 * DIM + m];
    mx[m] = boxes[(uint64_t)b * 2 * DIM + DIM + m];
  }
  NodeSource src;
  src.flat = recs;
  if (BUF) src.rsrc = __builtin_amdgcn_make_buffer_rsrc((void*)recs, (short)0, (int)node_bytes, 0x00020000);
  const uint32_t c =
      wave_box_walk<DIM, BUF, kTopLevels>(src, top, n, mn, mx, fr, anc, rbox, ids + (uint64_t)b * slot, slot);
  if (lane_id() == 0) counts[b] = c;
}

// Resident per-box search (fkd_box_server_*).  Each workgroup (one wave)
// owns one request slot in page-locked host memory and polls it: the host
// writes a box and then a new sequence number into the slot's 64-byte line;
// the wave reads the whole line with one load, walks the box
// (wave_box_walk), writes the ids, then the count and the box it searched
// (the host re-posts a request whose echo differs: the line's read need not
// be one snapshot), and publishes the sequence number after a system-scope
// release fence.  A call therefore costs no launch and no stream
// synchronisation, only PCIe round trips and the walk.  The top
// ServerTop::kLevels depths stay in LDS for the kernel's lifetime.
//
// Every wave ends: on a stop sequence number in its slot, when any wave of
// the launch has closed it (ctl->closing == gen), or when the whole launch
// has seen no request for idle_ticks of the device wall clock
// (ctl->last_act, raised by every request) or for max_idle_polls checks in
// a row with last_act unchanged, whichever comes first; an idle wave then
// closes the launch.  A wave whose slot has been idle for a while polls
// less often (s_sleep between polls), so idle slots do not crowd the PCIe
// reads of busy ones.  The last wave to leave writes gen to *exit_word
// (host memory), which tells the host that requests posted since must be
// served by a new launch.
struct alignas(64) BoxReq {
  uint32_t seq;
  uint32_t pad0;
  float box[8];  // mins then maxs, 2 * dim floats
  uint32_t pad1[6];
};
struct alignas(64) BoxRep {
  uint32_t seq;
  uint32_t count;  // true hits, or kSmallOverflow
  float echo[8];   // the box the wave searched
  uint32_t walk_ticks;  // wall-clock ticks from reading the request to the end of the walk
  uint32_t pad[5];
};
struct BoxCtl {
  unsigned long long last_act;  // wall clock of the last request (or of a wave's start)
  uint32_t closing;             // generation of a launch that is closing
  uint32_t exited;              // waves of the current launch that have left
};
constexpr uint32_t kServerStop = 0xffffffffu;
template <int DIM>
struct ServerTop {
  static constexpr uint32_t kLevels = DIM <= 3 ? 13u : 12u;  // 128 KB of records in LDS
};

__device__ inline uint32_t sys_load(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

template <int DIM, bool BUF>
__global__ __launch_bounds__(kWave) void box_server_kernel(const float* recs, uint32_t node_bytes, uint32_t n,
                                                           const BoxReq* req, BoxRep* rep, uint32_t* ids,
                                                           uint32_t slot_ids, uint32_t* done, BoxCtl* ctl,
                                                           uint32_t* exit_word, uint32_t gen,
                                                           unsigned long long idle_ticks,
                                                           uint32_t max_idle_polls) {
  constexpr int RF = NodeRec<DIM>::kFloats;
  constexpr uint32_t TOP = ServerTop<DIM>::kLevels;
  __shared__ uint32_t fr[2][kWaveFrontier];
  __shared__ uint4 anc[kAncSizes * kWave];
  __shared__ float2 rbox[DIM * DIM];
  __shared__ __align__(16) float top[((1u << TOP) - 1u) * RF];
  const uint32_t s = blockIdx.x;
  const uint32_t lane = lane_id();
  wave_load_top<DIM, TOP>(recs, n, top);
  wave_fill_ancestors(anc);
  NodeSource src;
  src.flat = recs;
  if (BUF) src.rsrc = __builtin_amdgcn_make_buffer_rsrc((void*)recs, (short)0, (int)node_bytes, 0x00020000);
  uint32_t handled = done[s];
  if (lane == 0) atomicMax(&ctl->last_act, (unsigned long long)wall_clock64());
  const uint32_t* line = reinterpret_cast<const uint32_t*>(req + s);
  uint32_t idle = 0, checks = 0;
  unsigned long long seen_act = 0;
  while (true) {
    // the slot's line: word 0 the sequence number, words 2.. the box
    const uint32_t w = sys_load(line + (lane & 15u));
    const uint32_t q = __builtin_amdgcn_readfirstlane(w);
    if (q == kServerStop) break;
    if (q != handled) {
      const unsigned long long t0 = wall_clock64();
      float mn[DIM], mx[DIM];
#pragma unroll
      for (int m = 0; m < DIM; ++m) {
        mn[m] = __uint_as_float((uint32_t)__shfl((int)w, 2 + m));
        mx[m] = __uint_as_float((uint32_t)__shfl((int)w, 2 + DIM + m));
      }
      const uint32_t c = wave_box_walk<DIM, BUF, TOP>(src, top, n, mn, mx, fr, anc, rbox,
                                                      ids + (uint64_t)s * slot_ids, slot_ids);
      const unsigned long long t1 = wall_clock64();
      BoxRep* r = rep + s;
      if (lane == 0) {
        r->count = c;
        r->walk_ticks = (uint32_t)(t1 - t0);
      }
      if (lane >= 2u && lane < 2u + 2u * (uint32_t)DIM) r->echo[lane - 2u] = __uint_as_float(w);
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
      if (lane == 0) {
        __hip_atomic_store(&r->seq, q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        done[s] = q;
        atomicMax(&ctl->last_act, (unsigned long long)t1);
      }
      handled = q;
      idle = 0;
      continue;
    }
    ++idle;
    if ((idle & 15u) == 0u) {
      const uint32_t cl = __hip_atomic_load(&ctl->closing, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (__builtin_amdgcn_readfirstlane(cl) == gen) break;
      const unsigned long long la = __hip_atomic_load(&ctl->last_act, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const unsigned long long now = wall_clock64();
      // the poll-count bound counts checks since the launch last saw a request
      checks = la != seen_act ? 0u : checks + 1u;
      seen_act = la;
      if ((now > la && now - la > idle_ticks) || checks > max_idle_polls) {
        if (lane == 0) __hip_atomic_store(&ctl->closing, gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
    }
    // a slot idle for a while polls about every 2 us instead of back to back
    if (idle > 256u)
      __builtin_amdgcn_s_sleep(127);
    else
      __builtin_amdgcn_s_sleep(2);
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
  if (lane == 0 && atomicAdd(&ctl->exited, 1u) == gridDim.x - 1u)
    __hip_atomic_store(exit_word, gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Work estimate of every search block for the fill pass: the most nodes any
// of its queries visited in the count pass (visits in processing order).
__global__ __launch_bounds__(kThreads) void block_work_kernel(const uint32_t* __restrict__ work_slot, uint32_t nq,
                                                              uint32_t* __restrict__ work) {
  __shared__ uint32_t s_max[kThreads / kWave];
  const uint32_t slot = blockIdx.x * kThreads + threadIdx.x;
  uint32_t h = slot < nq ? work_slot[slot] : 0u;
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) h = max(h, (uint32_t)__shfl_xor((int)h, o));
  if (lane_id() == 0) s_max[threadIdx.x / kWave] = h;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t m = 0;
#pragma unroll
    for (int w = 0; w < kThreads / kWave; ++w) m = max(m, s_max[w]);
    work[blockIdx.x] = m;
  }
}

// Fill-pass lane balance: a wave runs as long as its slowest lane, and
// neighbouring boxes in the spatial order can differ in work by orders of
// magnitude (box sizes vary, blobs sit next to sparse regions).  Every window
// of 256 ITEMS consecutive processing slots is reordered by decreasing work
// (the count pass's visits, bucketed on log2 with three fraction bits), so
// the lanes of a wave get similar work while staying inside one spatial
// window.  Writes slot_order (dispatch slot -> processing slot) and the work
// of every search block (its first, heaviest, slot).  The order inside a
// bucket follows LDS atomics and only affects time.
constexpr uint32_t kSortBuckets = 256;
__device__ inline uint32_t sort_bucket(uint32_t w) {
  const uint32_t lg = floor_log2(w | 1u);
  const uint32_t frac = lg >= 3u ? (w >> (lg - 3u)) & 7u : (w << (3u - lg)) & 7u;
  return (31u - lg) * 8u + (7u - frac);  // 0 = heaviest
}
template <int ITEMS>
__global__ __launch_bounds__(kThreads) void window_sort_kernel(const uint32_t* __restrict__ work_slot, uint32_t nq,
                                                               uint32_t* __restrict__ slot_order,
                                                               uint32_t* __restrict__ block_work) {
  static_assert(kSortBuckets == kThreads, "one bucket per thread in the scan");
  __shared__ uint32_t hist[kSortBuckets];
  __shared__ uint32_t wsum[kThreads / kWave];
  constexpr uint32_t W = kThreads * ITEMS;
  const uint32_t base = blockIdx.x * W;
  hist[threadIdx.x] = 0;
  __syncthreads();
  uint32_t w[ITEMS], bk[ITEMS], rk[ITEMS];
#pragma unroll
  for (int j = 0; j < ITEMS; ++j) {
    const uint32_t s = base + (uint32_t)j * kThreads + threadIdx.x;
    w[j] = s < nq ? work_slot[s] : 0u;
    bk[j] = sort_bucket(w[j]);
    rk[j] = s < nq ? atomicAdd(&hist[bk[j]], 1u) : 0u;
  }
  __syncthreads();
  // exclusive scan of the bucket sizes (bucket = thread)
  const uint32_t h = hist[threadIdx.x];
  uint32_t x = h;
  const uint32_t lane = lane_id();
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    const uint32_t y = (uint32_t)__shfl_up((int)x, o);
    if (lane >= (uint32_t)o) x += y;
  }
  if (lane == kWave - 1) wsum[threadIdx.x / kWave] = x;
  __syncthreads();
  uint32_t pre = 0;
#pragma unroll
  for (int v = 0; v < kThreads / kWave; ++v)
    if ((uint32_t)v < threadIdx.x / kWave) pre += wsum[v];
  const uint32_t start = pre + x - h;
  __syncthreads();
  hist[threadIdx.x] = start;
  __syncthreads();
#pragma unroll
  for (int j = 0; j < ITEMS; ++j) {
    const uint32_t s = base + (uint32_t)j * kThreads + threadIdx.x;
    if (s < nq) {
      const uint32_t d = base + hist[bk[j]] + rk[j];
      slot_order[d] = s;
      if ((d & (kThreads - 1u)) == 0u) block_work[d / kThreads] = w[j];
    }
  }
}

// Work estimate of every box before counting (processing order): the
// expected number of hits, box volume times the mean point density of the
// trees (at least 1, saturating).  Only its order matters.
template <int DIM>
__global__ void volume_work_kernel(const float* __restrict__ mins, const float* __restrict__ maxs, uint32_t nq,
                                   float density, uint32_t* __restrict__ work) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nq; i += gridDim.x * blockDim.x) {
    float v = density;
#pragma unroll
    for (int m = 0; m < DIM; ++m) {
      const float e = maxs[(uint64_t)m * nq + i] - mins[(uint64_t)m * nq + i];
      v *= e > 0.f ? e : 0.f;  // also maps NaN to 0
    }
    v = v * 16.f + 1.f;
    work[i] = v < 2.0e9f ? (uint32_t)v : 2000000000u;
  }
}

// Dispatch order of the fill pass: within every XCD chunk, blocks by
// decreasing work, bucketed on log2 with two fraction bits, so that the
// heaviest blocks start first and the last to finish are light ones.  The
// order inside a bucket is arbitrary; it only affects time.  One workgroup.
constexpr uint32_t kWorkBuckets = 128;
__device__ inline uint32_t work_bucket(uint32_t w) {
  const uint32_t lg = floor_log2(w);
  const uint32_t frac = lg >= 2u ? (w >> (lg - 2u)) & 3u : (lg == 1u ? (w & 1u) << 1 : 0u);
  return (31u - lg) * 4u + (3u - frac);
}
__global__ __launch_bounds__(1024) void block_order_kernel(const uint32_t* __restrict__ work, uint32_t nb,
                                                           uint32_t* __restrict__ order) {
  __shared__ uint32_t cur[8][kWorkBuckets];
  for (uint32_t i = threadIdx.x; i < 8u * kWorkBuckets; i += blockDim.x) (&cur[0][0])[i] = 0;
  __syncthreads();
  auto chunk_of = [&](uint32_t b) {
    uint32_t x = 0;
    while (x < 7u && b >= xcd_chunk_start(x + 1u, nb)) ++x;
    return x;
  };
  for (uint32_t b = threadIdx.x; b < nb; b += blockDim.x) atomicAdd(&cur[chunk_of(b)][work_bucket(work[b])], 1u);
  __syncthreads();
  if (threadIdx.x < 8u) {
    uint32_t run = xcd_chunk_start(threadIdx.x, nb);
    for (uint32_t k = 0; k < kWorkBuckets; ++k) {
      const uint32_t c = cur[threadIdx.x][k];
      cur[threadIdx.x][k] = run;
      run += c;
    }
  }
  __syncthreads();
  for (uint32_t b = threadIdx.x; b < nb; b += blockDim.x)
    order[atomicAdd(&cur[chunk_of(b)][work_bucket(work[b])], 1u)] = b;
}

// Spatial sort key of every query: Morton interleave of its box centre,
// quantised to kb bits per dimension inside the tree's bounding box, below
// the query's event (qev, at bit ev_shift) when the handle holds several
// trees, so that a wave walks one event's tree.  With vb > 0 a class of the
// box volume (relative to the bounding box, 2.5 octaves per class) is
// inserted below the top two Morton bits per dimension: inside each of those
// regions boxes of similar size, and so similar work, are neighbours.  Also
// writes the boxes as 32-byte records (mins, then maxs, each padded to 4), so
// that the permutation gathers one record per box instead of 2 DIM floats.
// Any key order gives the same results; this one only groups similar work.
template <int DIM>
__global__ void query_keys_kernel(const float* __restrict__ mins, const float* __restrict__ maxs,
                                  uint32_t nq, const uint32_t* __restrict__ bbox, uint32_t kb, uint32_t vb,
                                  const uint32_t* __restrict__ qev, uint32_t ev_shift,
                                  uint32_t* __restrict__ keys, uint32_t* __restrict__ vals,
                                  float4* __restrict__ aos) {
  static_assert(DIM <= 4, "box records hold 4 floats per bound");
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
    float b[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int m = 0; m < DIM; ++m) {
      b[m] = mins[(uint64_t)m * nq + i];
      b[4 + m] = maxs[(uint64_t)m * nq + i];
      const float c = 0.5f * b[m] + 0.5f * b[4 + m];
      float f = (c - lo[m]) * sc[m];
      f = (f >= 0.f) ? f : 0.f;  // also maps NaN to 0
      f = f < cells - 1.f ? f : cells - 1.f;
      u[m] = (uint32_t)f;
    }
    for (uint32_t b = 0; b < kb; ++b)
#pragma unroll
      for (int m = 0; m < DIM; ++m) key |= ((u[m] >> b) & 1u) << (b * DIM + m);
    if (vb) {
      // -log2 of the box volume relative to the bounding box
      float nl = 0.f;
#pragma unroll
      for (int m = 0; m < DIM; ++m) {
        const float r = (b[4 + m] - b[m]) * sc[m] / cells;
        nl -= r > 0.f ? __log2f(r) : -1000.f;  // empty, flat or NaN boxes: the last class
      }
      const float cf = nl * 0.4f;
      const uint32_t cmax = (1u << vb) - 1u;
      const uint32_t cls = cf > 0.f ? (cf < (float)cmax ? (uint32_t)cf : cmax) : 0u;
      const uint32_t mb = kb * (uint32_t)DIM;
      const uint32_t lowb = mb > 2u * DIM ? mb - 2u * DIM : 0u;  // bits below the top regions
      const uint32_t low = key & ((1u << lowb) - 1u);
      key = ((key >> lowb) << (lowb + vb)) | (cls << lowb) | low;
    }
    if (qev && ev_shift < 32u) key |= qev[i] << ev_shift;
    keys[i] = key;
    vals[i] = i;
    aos[2 * (uint64_t)i] = make_float4(b[0], b[1], b[2], b[3]);
    aos[2 * (uint64_t)i + 1] = make_float4(b[4], b[5], b[6], b[7]);
  }
}

// Boxes in processing order: smin[d][i] = mins[d][perm[i]].
template <int DIM>
__global__ void permute_boxes_kernel(const float4* __restrict__ aos, const uint32_t* __restrict__ perm,
                                     uint32_t nq, float* __restrict__ smins, float* __restrict__ smaxs) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nq; i += gridDim.x * blockDim.x) {
    const uint32_t q = perm[i];
    const float4 lo = aos[2 * (uint64_t)q], hi = aos[2 * (uint64_t)q + 1];
    const float l[4] = {lo.x, lo.y, lo.z, lo.w}, h[4] = {hi.x, hi.y, hi.z, hi.w};
#pragma unroll
    for (int m = 0; m < DIM; ++m) {
      smins[(uint64_t)m * nq + i] = l[m];
      smaxs[(uint64_t)m * nq + i] = h[m];
    }
  }
}

}  // namespace fkd
