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
#pragma once
#include "build.hpp"
#include "common.hpp"

namespace fkd {

struct SearchArgs {
  const float* nodes;     // packed records, NodeRec<DIM>::kFloats floats each
  const uint32_t* off;    // event -> first node (n_events + 1), or null
  uint32_t n_single;      // node count when off == null
  const uint32_t* qev;    // query -> event, or null (event 0)
  uint32_t nq;
  const float* mins;      // [DIM][nq]
  const float* maxs;
  uint32_t* counts;       // count pass output
  uint32_t* visited;      // stats pass: nodes visited per query
  const uint64_t* qoff;   // fill pass: first output slot of every query
  uint32_t* out;          // fill pass output
  uint64_t cap;           // capacity of out
};

template <int DIM>
__device__ inline void load_rec(const float* nodes, uint32_t i, float (&c)[DIM], uint32_t& id) {
  constexpr int RF = NodeRec<DIM>::kFloats;
  const float4 v = *reinterpret_cast<const float4*>(nodes + (uint64_t)i * RF);
  c[0] = v.x;
  if (DIM > 1) c[DIM > 1 ? 1 : 0] = v.y;
  if (DIM > 2) c[DIM > 2 ? 2 : 0] = v.z;
  if (RF == 4) {
    id = __float_as_uint(v.w);
  } else {
    c[DIM > 3 ? 3 : 0] = v.w;
    id = __float_as_uint(*(nodes + (uint64_t)i * RF + 4));
  }
}

enum SearchMode { kCount = 0, kFill = 1, kStats = 2 };

template <int DIM, int MODE>
__global__ __launch_bounds__(kThreads) void search_kernel(SearchArgs a) {
  constexpr bool FILL = MODE == kFill;
  const uint32_t q = blockIdx.x * kThreads + threadIdx.x;
  if (q >= a.nq) return;
  uint32_t base = 0, n = a.n_single;
  if (a.off) {
    const uint32_t e = a.qev ? a.qev[q] : 0u;
    base = a.off[e];
    n = a.off[e + 1] - base;
  }
  float mn[DIM], mx[DIM];
#pragma unroll
  for (int m = 0; m < DIM; ++m) {
    mn[m] = a.mins[(uint64_t)m * a.nq + q];
    mx[m] = a.maxs[(uint64_t)m * a.nq + q];
  }
  const float* nodes = a.nodes + (uint64_t)base * NodeRec<DIM>::kFloats;
  uint64_t w = FILL ? a.qoff[q] : 0;
  uint32_t c = 0;
  uint32_t node = 0, depth = 0, dd = 0, pending = 0, vis = 0;
  while (n != 0) {
    if (MODE == kStats) ++vis;
    float x[DIM];
    uint32_t id;
    load_rec<DIM>(nodes, node, x, id);
    float v = x[0], lo = mn[0], hi = mx[0];
#pragma unroll
    for (int m = 1; m < DIM; ++m) {
      if (dd == (uint32_t)m) {
        v = x[m];
        lo = mn[m];
        hi = mx[m];
      }
    }
    const bool lower = v < lo;
    const bool inter = v <= hi && v >= lo;
    if (inter) {
      bool in = true;
#pragma unroll
      for (int m = 0; m < DIM; ++m) in = in && (x[m] <= mx[m] && x[m] >= mn[m]);
      if (in) {
        if (FILL && w < a.cap) a.out[w] = id;
        ++w;
        ++c;
      }
    }
    const uint32_t lc = 2u * node + 1u;
    const bool go_l = !lower && lc < n;
    const bool go_r = (lower || inter) && lc + 1u < n;
    if (go_l || go_r) {
      if (go_l && go_r) pending |= 2u << depth;
      node = go_l ? lc : lc + 1u;
      ++depth;
      dd = (dd + 1u == (uint32_t)DIM) ? 0u : dd + 1u;
    } else {
      if (pending == 0u) break;
      const uint32_t b = 31u - (uint32_t)__builtin_clz(pending);
      pending &= ~(1u << b);
      node = (node + 1u) >> (depth - b);
      depth = b;
      dd = b % (uint32_t)DIM;
    }
  }
  if (!FILL) a.counts[q] = c;
  if (MODE == kStats) a.visited[q] = vis;
}

}  // namespace fkd
