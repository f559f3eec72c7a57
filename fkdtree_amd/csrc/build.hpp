// Build kernels: key preparation and the per-depth stable partition.
//
// Data layout (P = points of the batch, all events end to end):
//   X[dim][P] float   input coordinates (SoA),  ids[P] uint32
//   A[m][P]  uint32   point indices; for every dimension m one array kept in
//                     in-order order of the tree under construction, each
//                     interval sorted by (coordinate m, id)
//   R[m][P]  uint32   rank of every point in the (event, coordinate m, id)
//                     order: comparing ranks equals the reference comparator
//                     (FKDTree_cpu.h:246-252)
//   tree outputs: Tc[dim][P] float, Tid[P] uint32 (heap order per event,
//                 the arrays getDimensionVector/getIdVector return,
//                 FKDTree_cpu.h:209-213) and packed node records for search.
#pragma once
#include "common.hpp"
#include "radix.hpp"

namespace fkd {

// Search record of a node (kFloats floats): its coordinates rotated so that
// the split dimension comes first -- slot j holds coordinate (dd + j) % DIM,
// dd = depth % DIM -- then at index kId the id bits.  The traversal reads the
// split value from slot 0 and compares slot j with the box rotated the same
// way, so it never selects by dimension; hits are emitted without a second
// load.
template <int DIM>
struct NodeRec {
  static constexpr int kFloats = DIM <= 3 ? 4 : 8;
  static constexpr int kId = DIM <= 3 ? 3 : 4;
};

// AoS FKDPoint<float, DIM> records (DIM floats then a uint32 id) -> SoA.
template <int DIM>
__global__ void aos_to_soa_kernel(const uint32_t* __restrict__ aos, float* __restrict__ X,
                                  uint32_t* __restrict__ ids, uint32_t P) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < P; i += gridDim.x * blockDim.x) {
#pragma unroll
    for (int m = 0; m < DIM; ++m) X[(uint64_t)m * P + i] = __uint_as_float(aos[(uint64_t)i * (DIM + 1) + m]);
    ids[i] = aos[(uint64_t)i * (DIM + 1) + DIM];
  }
}

// Keys and values for the coordinate sorts.  order == nullptr means point
// order already is (event, id) order.  Sets bit 0 of *err on a NaN.  Also
// sets the bounding box that the last digit pass narrows to its starting
// values (min keys all ones, max keys zero), and, when rec is given (dim <= 3,
// point order), the point records of records_kernel from the same loads.
template <int DIM>
__global__ void keys_kernel(const float* __restrict__ X, const uint32_t* __restrict__ order,
                            uint32_t P, uint32_t* __restrict__ K, uint32_t* __restrict__ V,
                            uint64_t stride, uint32_t* err, uint32_t* bbox, const uint32_t* __restrict__ ids,
                            float4* __restrict__ rec) {
  if (blockIdx.x == 0 && threadIdx.x < 2u * DIM) bbox[threadIdx.x] = threadIdx.x < (uint32_t)DIM ? 0xffffffffu : 0u;
  bool nan = false;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < P; i += gridDim.x * blockDim.x) {
    const uint32_t p = order ? order[i] : i;
    float c[DIM];
#pragma unroll
    for (int m = 0; m < DIM; ++m) {
      const float x = X[(uint64_t)m * P + p];
      c[m] = x;
      nan |= (x != x);
      K[m * stride + i] = fkey(x);
      V[m * stride + i] = p;
    }
    if (DIM <= 3 && rec)
      rec[i] = make_float4(c[0], DIM > 1 ? c[DIM > 1 ? 1 : 0] : 0.f, DIM > 2 ? c[DIM > 2 ? 2 : 0] : 0.f,
                           __uint_as_float(ids[i]));
  }
  if (nan) atomicOr(err, 1u);
}

// Rank planes of the level arrays: array m at position j holds the ranks of
// its point in every dimension k: (k == m) ? j : R[A_m[j]][k], where the
// ranks of a point form one 16-byte record (one gather per array element).
// Plane k of an array starts at k * ps; ps = plane_stride(P).  Work items
// are (radix tile, array) pairs, XCD-chunked as in the last digit pass that
// wrote R, so that an XCD gathers the records of the events it wrote.
template <int DIM>
__global__ __launch_bounds__(kThreads) void planes_kernel(const uint32_t* __restrict__ A, uint64_t astride,
                                                          const uint4* __restrict__ R, TileMap tm, uint32_t tile,
                                                          uint32_t* __restrict__ planes, uint64_t ps,
                                                          uint64_t pstride) {
  static_assert(DIM <= 4, "rank records hold 4 dimensions");
  constexpr int U = 16;  // elements per thread, their gathers in flight together
  const uint32_t w = xcd_item(blockIdx.x, gridDim.x);
  const uint32_t t = w / (uint32_t)DIM, m = w % (uint32_t)DIM;
  uint32_t base, end;
  if (!tm.tile_ev) {
    base = t * tile;
    end = min(base + tile, tm.single_n);
  } else {
    const uint32_t e = tm.tile_ev[t];
    base = tm.off[e] + (t - tm.tb[e]) * tile;
    end = min(base + tile, tm.off[e + 1]);
  }
  uint32_t* dst = planes + m * pstride;
  for (uint32_t j0 = base + threadIdx.x; j0 < end; j0 += U * kThreads) {
    uint32_t pt[U];
    uint4 rr[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t j = j0 + (uint32_t)u * kThreads;
      pt[u] = j < end ? A[m * astride + j] : 0u;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t j = j0 + (uint32_t)u * kThreads;
      rr[u] = j < end ? R[pt[u]] : make_uint4(0u, 0u, 0u, 0u);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t j = j0 + (uint32_t)u * kThreads;
      if (j >= end) continue;
      const uint32_t rk[4] = {rr[u].x, rr[u].y, rr[u].z, rr[u].w};
#pragma unroll
      for (int k = 0; k < DIM; ++k) dst[(uint64_t)k * ps + j] = (k == (int)m) ? j : rk[k];
    }
  }
}

// Keys and values for the id sort (tie order inside every event).
__global__ void id_keys_kernel(const uint32_t* __restrict__ ids, uint32_t P, uint32_t* __restrict__ K,
                               uint32_t* __restrict__ V) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < P; i += gridDim.x * blockDim.x) {
    K[i] = ids[i];
    V[i] = i;
  }
}

// Stride of the rank planes: a multiple of 64 (256-byte aligned planes, so a
// thread's 8 ranks never straddle a 32-byte sector when the event starts on
// a multiple of 8) with at least 4 slack elements, so the level kernel's
// aligned 16-byte windows around an event's positions stay inside their
// plane.
__host__ __device__ inline uint64_t plane_stride(uint64_t P) { return (P + 4u + 63u) & ~(uint64_t)63u; }

constexpr int kLevelItems = 8;
constexpr int kLevelTile = kThreads * kLevelItems;

// Depths whose subtrees have at most 2^fin_log - 1 points are built by one
// workgroup per subtree in LDS (finish_kernel); the global level loop stops
// above them.  In an event of depth D every subtree of depth L has at most
// 2^(D-L+1) - 1 points, so the finisher starts at depth fin_depth(D, fin_log).
// fin_log is a launch parameter (kFinLogMin..kFinLogMax); the finisher's
// threads take 2^(fin_log - 8) consecutive positions each.
constexpr uint32_t kFinLogDefault = 10;
constexpr uint32_t kFinLogMin = 9;
constexpr uint32_t kFinLogMax = 11;
__host__ __device__ inline uint32_t fin_depth(uint32_t D, uint32_t fin_log) {
  return D + 1u > fin_log ? D + 1u - fin_log : 0u;
}
// Largest finisher subtree of an event of n >= 1 points: H - 1 + min(R, H)
// with H = 2^(D - fin_depth(D)) and R the nodes on the last depth D.
__host__ __device__ inline uint32_t fin_subtree_max(uint32_t n, uint32_t fin_log) {
  const uint32_t D = floor_log2(n);
  const uint32_t H = 1u << (D - fin_depth(D, fin_log));
  const uint32_t R = n - ((1u << D) - 1u);
  return H - 1u + (R < H ? R : H);
}
// Dynamic LDS of finish_kernel for subtrees of at most cap points: the rank
// table (dim x cap uint32) and the names (dim x 2 x (cap + 1) uint16).  Sized
// per launch so that trees whose subtrees are half full (2^k points) keep
// twice as many workgroups resident.
__host__ __device__ inline uint32_t fin_lds_bytes(uint32_t dim, uint32_t cap) {
  return (dim * cap * 4u + dim * 2u * (cap + 1u) * 2u + 15u) & ~15u;
}
// Positions per finisher thread for subtrees of at most cap points: the
// smallest of 1, 2, 4, 8 with 256 * items >= cap (cap < 2^kFinLogMax).
__host__ __device__ inline uint32_t fin_items(uint32_t cap) {
  return cap <= 256u ? 1u : cap <= 512u ? 2u : cap <= 1024u ? 4u : 8u;
}

template <int DIM>
struct LevelArgs {
  uint32_t L;                 // depth whose intervals are split
  uint32_t d;                 // split dimension L % DIM
  uint32_t fin_log;           // depths from fin_depth(D, fin_log) on are finish_kernel's
  const uint32_t* Ad;         // rank planes of array d (not moved at this depth)
  const uint32_t* src[3];     // rank planes of the DIM-1 other arrays ...
  uint32_t* dst[3];           // ... and where their partitioned copies go
  const uint32_t* S0;         // rank in dimension 0 -> point index
  uint32_t* F;                // heap index -> rank in dimension 0 of the node's point
  TileMap tm;                 // tiles of kLevelTile points, event aligned
  const float* X;             // input points
  const uint32_t* ids;
  uint32_t P;                 // points of the batch (stride of X and the tree arrays)
  uint64_t PS;                // plane stride, plane_stride(P)
  float* Tc;                  // tree outputs
  uint32_t* Tid;
  float* recs;                // search records (rotated coordinates, id)
  uint64_t* desc;             // look-back descriptors, zero at launch
  uint32_t* ctr;              // dynamic block counter, zero at launch
};

// Event, event-local tile, first point and size of level tile t; a map
// without a tile table is one tree of single_n points.
struct LevelTile {
  uint32_t e, lt, o, n;
};
__device__ inline LevelTile level_tile(const TileMap& tm, uint32_t t) {
  LevelTile r;
  if (!tm.tile_ev) {
    r.e = 0;
    r.lt = t;
    r.o = 0;
    r.n = tm.single_n;
    return r;
  }
  r.e = tm.tile_ev[t];
  r.lt = t - tm.tb[r.e];
  r.o = tm.off[r.e];
  r.n = tm.off[r.e + 1] - r.o;
  return r;
}

// The node record of point p: its coordinates and id (one 16-byte gather of
// the point record for dim <= 3).
template <int DIM>
__device__ inline void node_point(const LevelArgs<DIM>& a, uint32_t p, const float4* prec, float (&c)[DIM],
                                  uint32_t& id) {
  if (DIM <= 3) {
    const float4 r = prec[p];
    const float rc[3] = {r.x, r.y, r.z};
#pragma unroll
    for (int m = 0; m < DIM; ++m) c[m] = rc[m < 3 ? m : 0];
    id = __float_as_uint(r.w);
  } else {
#pragma unroll
    for (int m = 0; m < DIM; ++m) c[m] = a.X[(uint64_t)m * a.P + p];
    id = a.ids[p];
  }
}

// Heap-ordered tree arrays and the rotated search record of one node.
template <int DIM>
__device__ inline void write_node(const LevelArgs<DIM>& a, uint32_t o, uint32_t node, uint32_t depth,
                                  const float (&c)[DIM], uint32_t id) {
  constexpr int RF = NodeRec<DIM>::kFloats;
  const uint32_t dd = depth % (uint32_t)DIM;
  float r[4] = {0.f, 0.f, 0.f, 0.f};  // rotated coordinates
#pragma unroll
  for (int m = 0; m < DIM; ++m) {
    a.Tc[(uint64_t)m * a.P + o + node] = c[m];
#pragma unroll
    for (int j = 0; j < DIM; ++j)  // slot j holds coordinate (dd + j) % DIM
      if ((dd + (uint32_t)j) % (uint32_t)DIM == (uint32_t)m) r[j < 4 ? j : 0] = c[m];
  }
  a.Tid[o + node] = id;
  float* rec = a.recs + (uint64_t)(o + node) * RF;
  if (RF == 4) {
    *reinterpret_cast<float4*>(rec) = make_float4(r[0], r[1], r[2], __uint_as_float(id));
  } else {
    reinterpret_cast<float4*>(rec)[0] = make_float4(r[0], r[1], r[2], r[3]);
    reinterpret_cast<float4*>(rec)[1] = make_float4(__uint_as_float(id), 0.f, 0.f, 0.f);
  }
}

// Segmented (reset-at-interval-start) counts of left / right elements.
struct SegCount {
  uint32_t f, l, r;
};
__device__ inline SegCount seg_op(SegCount a, SegCount b) {
  SegCount c;
  c.f = a.f | b.f;
  c.l = b.f ? b.l : a.l + b.l;
  c.r = b.f ? b.r : a.r + b.r;
  return c;
}
// Descriptor word: status (2) | f (1) | l (30) | r (30).
__device__ inline uint64_t seg_pack(SegCount s) {
  return ((uint64_t)s.f << 60) | ((uint64_t)s.l << 30) | (uint64_t)s.r;
}
__device__ inline SegCount seg_unpack(uint64_t w) {
  SegCount s;
  s.f = (uint32_t)(w >> 60) & 1u;
  s.l = (uint32_t)(w >> 30) & 0x3fffffffu;
  s.r = (uint32_t)w & 0x3fffffffu;
  return s;
}

enum : uint32_t { kNone = 0, kLeft = 1, kMed = 2, kRight = 3 };

// One depth of the build for every interval of every event at once.
// Block b handles moving array b % (DIM-1) over tile b / (DIM-1) (b is the
// dynamic block index, so look-back waits only on blocks already running).
// For each interval [start, start+size) of depth L with median rank
// k = left_size(size) (FKDTree_cpu.h:239): the median is the element at
// start+k of array d; every other array is stably partitioned into
// [start, start+k) (rank in dimension d below the median's), start+k (the
// median) and (start+k, start+size) (rank above).  Elements carry their
// ranks in every dimension, so the partition reads and writes only
// coalesced planes.  Blocks of array 0 also store the depth-L nodes
// (add_at_position, FKDTree_cpu.h:253-255).  Depths from fin_depth(D) on,
// the last split and the leaves included, are finish_kernel's.
template <int DIM>
__global__ __launch_bounds__(kThreads) void level_kernel(LevelArgs<DIM> a) {
  constexpr uint32_t M = DIM - 1;
  __shared__ uint32_t s_blk;
  __shared__ SegCount s_wave[kThreads / kWave];
  __shared__ SegCount s_carry;
  __shared__ uint32_t s_cls[kThreads / kWave];
  // moved elements of the tile in write-out order, with their destinations
  __shared__ uint32_t s_out[M ? DIM + 1 : 1][M ? kLevelTile : 1];
  const uint32_t bid = dynamic_block(a.ctr, &s_blk);
  const uint32_t arr = M ? bid % M : 0u;
  const uint32_t t = M ? bid / M : bid;
  const LevelTile lti = level_tile(a.tm, t);
  const uint32_t lt = lti.lt, o = lti.o, n = lti.n;
  const Shape sh = shape_of(n);
  const bool split = a.L < fin_depth(sh.D, a.fin_log);  // deeper levels: finish_kernel
  const bool root_only = (n == 1 && a.L == 0);
  const uint64_t P = a.PS;  // plane stride
  if (!split) {
    if (M && threadIdx.x == 0) publish(&a.desc[bid], kStPrefix | seg_pack(SegCount{1u, 0u, 0u}));
    if (root_only && arr == 0 && threadIdx.x == 0) a.F[o] = a.Ad[o];
    return;
  }
  const uint32_t first = lt * (uint32_t)kLevelTile + threadIdx.x * kLevelItems;
  const uint32_t* Ad_d = a.Ad + (uint64_t)a.d * P;  // plane d of array d
  const uint32_t* S = M ? a.src[arr] : nullptr;

  // Phase 1: interval of every item and all loads (no stores in between so
  // the loads of all items can be in flight together).
  uint32_t cls[kLevelItems], dest0[kLevelItems], rk[kLevelItems][DIM], rm[kLevelItems], fr[kLevelItems];
  uint32_t st0[kLevelItems], mp[kLevelItems], hn[kLevelItems];
  // The thread's 8 ranks of a plane from aligned 16-byte loads where the
  // tile holds all 8 items: two when the event starts on a multiple of 4,
  // else three around them, shifted by sh4 = o % 4 (the same for the whole
  // block; planes are 16-byte aligned and padded, see plane_stride).  One
  // 4-byte load per item otherwise.  The windows also read separator
  // positions and up to 3 positions past the items, whose ranks go unused.
  static_assert(kLevelItems == 8, "two or three 16-byte loads per plane");
  const bool vec = M != 0 && first + (uint32_t)kLevelItems <= n;
  const uint32_t sh4 = o & 3u;
  if (vec && sh4 == 0u) {
#pragma unroll
    for (int k = 0; k < DIM; ++k) {
      const uint4* q = reinterpret_cast<const uint4*>(S + (uint64_t)k * P + o + first);
      const uint4 x0 = q[0], x1 = q[1];
      const uint32_t e[8] = {x0.x, x0.y, x0.z, x0.w, x1.x, x1.y, x1.z, x1.w};
#pragma unroll
      for (int i = 0; i < kLevelItems; ++i) rk[i][k] = e[i];
    }
  } else if (vec) {
#pragma unroll
    for (int k = 0; k < DIM; ++k) {
      const uint4* q = reinterpret_cast<const uint4*>(S + (uint64_t)k * P + o + first - sh4);
      const uint4 x0 = q[0], x1 = q[1], x2 = q[2];
      const uint32_t e[12] = {x0.x, x0.y, x0.z, x0.w, x1.x, x1.y, x1.z, x1.w, x2.x, x2.y, x2.z, x2.w};
#pragma unroll
      for (int i = 0; i < kLevelItems; ++i) rk[i][k] = sh4 == 1u ? e[i + 1] : sh4 == 2u ? e[i + 2] : e[i + 3];
    }
  }
#pragma unroll
  for (int i = 0; i < kLevelItems; ++i) {
    const uint32_t jl = first + i;
    cls[i] = kNone;
    st0[i] = 0;
    mp[i] = 0;
    hn[i] = 0;
    rm[i] = 0;
    fr[i] = 0;
    if (!vec) {
#pragma unroll
      for (int k = 0; k < DIM; ++k) rk[i][k] = 0;
    }
    if (jl < n) {
      const uint32_t ts = subtree_at(sh, a.L, jl);
      uint32_t start, size;
      subtree_range(sh, a.L, ts, start, size);
      if (jl - start < size) {
        const uint32_t mpos = start + left_size(size);
        st0[i] = start;
        mp[i] = mpos;
        hn[i] = (1u << a.L) - 1u + ts;  // heap index of the interval's node
        cls[i] = 8u;  // in an interval
        if (M) {
          rm[i] = Ad_d[o + mpos];
          if (!vec) {
#pragma unroll
            for (int k = 0; k < DIM; ++k) rk[i][k] = S[(uint64_t)k * P + o + jl];
          }
        }
        if (arr == 0 && jl == mpos) fr[i] = a.Ad[o + jl];  // plane 0 of array d
      }
    }
  }
  // Phase 2: classification and the thread's segmented counts.  cnt packs
  // the thread's unsegmented counts of left (bits 0-11), right (12-23) and
  // median (24-31) elements: a tile of 2,048 holds at most six intervals
  // (every interval here has at least 2^fin_log - 1 >= 511 points), so no
  // field overflows.
  SegCount agg{0u, 0u, 0u};
  uint32_t cnt = 0;
#pragma unroll
  for (int i = 0; i < kLevelItems; ++i) {
    const uint32_t jl = first + i;
    if (!(cls[i] & 8u)) continue;
    if (arr == 0 && jl == mp[i]) a.F[o + hn[i]] = fr[i];  // node of depth L
    if (!M) continue;
    uint32_t r = rk[i][0];
#pragma unroll
    for (int k = 1; k < DIM; ++k) r = (a.d == (uint32_t)k) ? rk[i][k] : r;
    uint32_t c;
    if (jl == st0[i]) agg = SegCount{1u, 0u, 0u};
    if (r < rm[i]) {
      c = kLeft;
      dest0[i] = o + st0[i];
      agg.l++;
      cnt += 1u;
    } else if (r > rm[i]) {
      c = kRight;
      dest0[i] = o + mp[i] + 1u;
      agg.r++;
      cnt += 1u << 12;
    } else {
      c = kMed;
      dest0[i] = o + mp[i];
      cnt += 1u << 24;
    }
    if (jl == st0[i]) c |= 4u;  // interval starts here
    cls[i] = c;
  }
  if (!M) return;

  // Block-wide segmented exclusive scan of the per-thread aggregates.
  SegCount incl = agg;
  const uint32_t lane = lane_id();
#pragma unroll
  for (int ofs = 1; ofs < kWave; ofs <<= 1) {
    SegCount u;
    u.f = __shfl_up(incl.f, ofs);
    u.l = __shfl_up(incl.l, ofs);
    u.r = __shfl_up(incl.r, ofs);
    if ((int)lane >= ofs) incl = seg_op(u, incl);
  }
  SegCount excl;
  excl.f = __shfl_up(incl.f, 1);
  excl.l = __shfl_up(incl.l, 1);
  excl.r = __shfl_up(incl.r, 1);
  if (lane == 0) excl = SegCount{0u, 0u, 0u};
  const int wave = threadIdx.x / kWave;
  const uint32_t cnt_incl = wave_incl_sum(cnt);
  if (lane == kWave - 1) {
    s_wave[wave] = incl;
    s_cls[wave] = cnt_incl;
  }
  __syncthreads();
  SegCount wcarry{0u, 0u, 0u}, tile_agg{0u, 0u, 0u};
  uint32_t cnt_excl = cnt_incl - cnt, cnt_tot = 0;
#pragma unroll
  for (int w = 0; w < kThreads / kWave; ++w) {
    if (w < wave) {
      wcarry = seg_op(wcarry, s_wave[w]);
      cnt_excl += s_cls[w];
    }
    tile_agg = seg_op(tile_agg, s_wave[w]);
    cnt_tot += s_cls[w];
  }
  excl = seg_op(wcarry, excl);

  if (wave == 0) {
    // Windowed look-back: lane l reads the descriptor of tile lt-1-base-l of
    // the same array; the nearest tile that has its prefix or an interval
    // start ends the walk, and only it and the tiles nearer than it are
    // waited for.  Values combine far-to-near with seg_op.
    SegCount carry{0u, 0u, 0u};
    if (lt == 0) {
      if (lane == 0) publish(&a.desc[bid], kStPrefix | seg_pack(tile_agg));
    } else {
      if (lane == 0) publish(&a.desc[bid], kStAggregate | seg_pack(tile_agg));
      for (uint32_t base = 0;; base += kWave) {
        const int64_t j = (int64_t)lt - 1 - (int64_t)base - (int64_t)lane;
        uint32_t first;
        const uint64_t w = look_window(
            j >= 0 ? &a.desc[bid - (base + lane + 1u) * M] : nullptr,
            [](uint64_t d) { return (d & kStMask) == kStPrefix || seg_unpack(d).f; }, &first);
        SegCount v = seg_unpack(w);
        if (lane > first) v = SegCount{0u, 0u, 0u};
#pragma unroll
        for (int o = 1; o < kWave; o <<= 1) {
          SegCount u;
          u.f = __shfl_down(v.f, o);
          u.l = __shfl_down(v.l, o);
          u.r = __shfl_down(v.r, o);
          if (lane + o < (uint32_t)kWave) v = seg_op(u, v);
        }
        carry = seg_op(v, carry);
        if (first < (uint32_t)kWave) break;
      }
      if (lane == 0) publish(&a.desc[bid], kStPrefix | seg_pack(seg_op(carry, tile_agg)));
    }
    if (lane == 0) s_carry = carry;
  }
  __syncthreads();
  const SegCount run0 = seg_op(s_carry, excl);
  uint32_t nl = run0.l, nr = run0.r;
  // Staging: the tile's left elements in tile order, then its right ones,
  // then its medians.  In tile order the destinations of each class rise,
  // jumping only between intervals, so the write-out below stores
  // consecutive slots of one run with consecutive threads.
  const uint32_t tot_l = cnt_tot & 0xfffu, tot_r = (cnt_tot >> 12) & 0xfffu;
  uint32_t sl = cnt_excl & 0xfffu;
  uint32_t sr = tot_l + ((cnt_excl >> 12) & 0xfffu);
  uint32_t sm = tot_l + tot_r + (cnt_excl >> 24);
#pragma unroll
  for (int i = 0; i < kLevelItems; ++i) {
    const uint32_t c = (cls[i] & 8u) ? 0u : cls[i];
    if (c & 4u) {
      nl = 0;
      nr = 0;
    }
    uint32_t dest = dest0[i], slot = sm;
    switch (c & 3u) {
      case kLeft: dest += nl; ++nl; slot = sl++; break;
      case kRight: dest += nr; ++nr; slot = sr++; break;
      case kMed: ++sm; break;
      default: break;
    }
    if ((c & 3u) != kNone) {
      s_out[DIM][slot] = dest;
#pragma unroll
      for (int k = 0; k < DIM; ++k) s_out[k][slot] = rk[i][k];
    }
  }
  __syncthreads();
  uint32_t* D = a.dst[arr];
  const uint32_t tn = tot_l + tot_r + (cnt_tot >> 24);
  for (uint32_t q = threadIdx.x; q < tn; q += kThreads) {
    const uint32_t dest = s_out[DIM][q];
#pragma unroll
    for (int k = 0; k < DIM; ++k) D[(uint64_t)k * P + dest] = s_out[k][q];
  }
}

// Two depths for one array at a time (dim 3).  An array of dim 3 moves on two
// consecutive depths and then splits the next one, so its two moves can be
// one: array m = (L + 2) % 3 goes from depth-L intervals straight to its
// place under depth L + 1, a stable partition of every depth-L interval into
// four runs (left-left, left-right, right-left, right-right) and three medians
// (of the interval on dimension dA = L % 3, of its two children on dB =
// (L + 1) % 3).  The medians' ranks come from the arrays that already hold
// depth L (array dA, moved through depth L - 1, is split on dA) and depth
// L + 1 (array dB, moved through depth L by the launch before).  One launch per
// depth then moves one array instead of two, half the bytes of level_kernel;
// the first launch of a build moves array 1 through depth 0 alone (single),
// and an event whose depth L + 1 is the finisher's gets depth L alone.
struct PairArgs {
  uint32_t L;            // first depth of the pair
  uint32_t dA, dB;       // split dimensions of depths L and L + 1
  uint32_t single;       // 1: depth L only, for every event
  uint32_t write_f;      // store the depth-L nodes in F
  uint32_t fin_log;
  const uint32_t* Aa;    // rank planes of array dA (holds depth L)
  const uint32_t* Ab;    // rank planes of array dB (holds depth L + 1)
  const uint32_t* src;   // rank planes of the moved array ...
  uint32_t* dst;         // ... and where its partitioned copy goes
  uint32_t* F;
  TileMap tm;            // tiles of kThreads * ITEMS positions, event aligned
  uint64_t PS;           // plane stride
  uint64_t* desc;        // two look-back words per block, zero at launch
  uint32_t* ctr;         // dynamic block counter, zero at launch
};

// Segmented counts of the four runs (reset at depth-L interval starts).
struct SegCount4 {
  uint32_t f, c[4];
};
__device__ inline SegCount4 seg4_op(const SegCount4& a, const SegCount4& b) {
  SegCount4 r;
  r.f = a.f | b.f;
#pragma unroll
  for (int k = 0; k < 4; ++k) r.c[k] = b.f ? b.c[k] : a.c[k] + b.c[k];
  return r;
}
__device__ inline SegCount4 seg4_shfl_up(const SegCount4& v, int ofs) {
  SegCount4 u;
  u.f = __shfl_up(v.f, ofs);
#pragma unroll
  for (int k = 0; k < 4; ++k) u.c[k] = __shfl_up(v.c[k], ofs);
  return u;
}
__device__ inline SegCount4 seg4_shfl_down(const SegCount4& v, int ofs) {
  SegCount4 u;
  u.f = __shfl_down(v.f, ofs);
#pragma unroll
  for (int k = 0; k < 4; ++k) u.c[k] = __shfl_down(v.c[k], ofs);
  return u;
}
// Look-back words: word 0 = status | f << 60 | c0 << 30 | c1, word 1 =
// status | c2 << 30 | c3 (counts below 2^30: events hold fewer points).  A
// reader takes a pair only when both words carry the same status.
__device__ inline void seg4_pack(const SegCount4& s, uint64_t st, uint64_t& w0, uint64_t& w1) {
  w0 = st | ((uint64_t)s.f << 60) | ((uint64_t)s.c[0] << 30) | (uint64_t)s.c[1];
  w1 = st | ((uint64_t)s.c[2] << 30) | (uint64_t)s.c[3];
}
__device__ inline SegCount4 seg4_unpack(uint64_t w0, uint64_t w1) {
  SegCount4 s;
  s.f = (uint32_t)(w0 >> 60) & 1u;
  s.c[0] = (uint32_t)(w0 >> 30) & 0x3fffffffu;
  s.c[1] = (uint32_t)w0 & 0x3fffffffu;
  s.c[2] = (uint32_t)(w1 >> 30) & 0x3fffffffu;
  s.c[3] = (uint32_t)w1 & 0x3fffffffu;
  return s;
}

enum : uint32_t { kLL = 0, kLR = 1, kRL = 2, kRR = 3, kMedian = 4, kOutside = 5 };

// The depth-L interval that holds position j (or the one just before the
// separator j), its median positions, the medians' ranks and the F entry of
// its node.
struct PairIv {
  uint32_t st, sz, mp, mpl, mpr, hn, rma, rml, rmr, fv;
};
__device__ inline PairIv pair_interval(const Shape& sh, uint32_t L, uint32_t j, bool two, const uint32_t* Aa0,
                                       const uint32_t* Aa_d, const uint32_t* Ab_d, uint32_t o) {
  PairIv v;
  const uint32_t ts = subtree_at(sh, L, j);
  subtree_range(sh, L, ts, v.st, v.sz);
  const uint32_t k = left_size(v.sz);
  v.mp = v.st + k;
  v.hn = (1u << L) - 1u + ts;
  v.rma = Aa_d[o + v.mp];
  v.fv = Aa0[o + v.mp];  // rank in dimension 0 of the median: its entry in F
  v.mpl = v.mpr = v.rml = v.rmr = 0u;
  if (two) {  // children of at least 2^fin_log - 1 points: both exist
    v.mpl = v.st + left_size(k);
    v.mpr = v.mp + 1u + left_size(v.sz - k - 1u);
    v.rml = Ab_d[o + v.mpl];
    v.rmr = Ab_d[o + v.mpr];
  }
  return v;
}

// DA = a.dA, the split dimension of depth L (a template argument, so the
// ranks of the item on dA and dB = (DA + 1) % 3 are registers, not selects).
template <int ITEMS, int DA>
__global__ __launch_bounds__(kThreads) void level_pair_kernel(PairArgs a) {
  constexpr int DIM = 3;
  constexpr uint32_t kTile = kThreads * ITEMS;
  static_assert(ITEMS == 4 || ITEMS == 8, "one or two 16-byte loads per plane");
  __shared__ uint32_t s_blk;
  __shared__ uint64_t s_wave[2][kThreads / kWave];
  __shared__ SegCount4 s_carry;
  // staging slot q at q + q / 8: a thread's items of one run are 8 slots
  // apart, and the padding spreads a wave's writes over all banks
  __shared__ uint32_t s_out[DIM + 1][kTile + kTile / 8 + 1];  // + a slot for positions that stay
  const uint32_t bid = dynamic_block(a.ctr, &s_blk);
  const LevelTile lti = level_tile(a.tm, bid);
  const uint32_t lt = lti.lt, o = lti.o, n = lti.n;
  const Shape sh = shape_of(n);
  const uint32_t fd = fin_depth(sh.D, a.fin_log);
  const uint64_t P = a.PS;
  if (a.L >= fd) {  // depth L is the finisher's for this event
    if (threadIdx.x == 0) {
      uint64_t w0, w1;
      seg4_pack(SegCount4{1u, {0u, 0u, 0u, 0u}}, kStPrefix, w0, w1);
      publish(&a.desc[2 * bid + 1], w1);
      publish(&a.desc[2 * bid], w0);
      if (n == 1 && a.L == 0 && a.write_f) a.F[o] = a.Aa[o];
    }
    return;
  }
  const bool two = !a.single && a.L + 1u < fd;
  const uint32_t first = lt * kTile + threadIdx.x * (uint32_t)ITEMS;
  const uint32_t* Aa_d = a.Aa + (uint64_t)a.dA * P;  // plane dA of array dA
  const uint32_t* Ab_d = a.Ab + (uint64_t)a.dB * P;  // plane dB of array dB

  // Phase 1: the ranks of the thread's items and the medians of their
  // intervals, all loads before any use.
  uint32_t rk[ITEMS][DIM];
  bool in[ITEMS], second[ITEMS];  // inside an interval; that interval is v1
  const bool vec = first + (uint32_t)ITEMS <= n;
  const uint32_t sh4 = o & 3u;
  if (vec) {
#pragma unroll
    for (int k = 0; k < DIM; ++k) {
      const uint4* q = reinterpret_cast<const uint4*>(a.src + (uint64_t)k * P + o + first - sh4);
      uint32_t e[ITEMS + 4];
#pragma unroll
      for (int v = 0; v < ITEMS / 4; ++v) {
        const uint4 x = q[v];
        e[4 * v] = x.x;
        e[4 * v + 1] = x.y;
        e[4 * v + 2] = x.z;
        e[4 * v + 3] = x.w;
      }
      if (sh4) {
        const uint4 x = q[ITEMS / 4];
        e[ITEMS] = x.x;
        e[ITEMS + 1] = x.y;
        e[ITEMS + 2] = x.z;
        e[ITEMS + 3] = x.w;
      } else {
        e[ITEMS] = e[ITEMS + 1] = e[ITEMS + 2] = e[ITEMS + 3] = 0u;
      }
#pragma unroll
      for (int i = 0; i < ITEMS; ++i)
        rk[i][k] = sh4 == 0u ? e[i] : sh4 == 1u ? e[i + 1] : sh4 == 2u ? e[i + 2] : e[i + 3];
    }
  }
  // The thread's ITEMS positions lie in at most two depth-L intervals (every
  // interval has at least 2^fin_log - 1 > ITEMS points) and a separator:
  // the intervals and medians are found for the first and the last position
  // only.
  PairIv v0{0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u}, v1 = v0;
  if (first < n) {
    const uint32_t jlast = (first + (uint32_t)ITEMS <= n ? first + (uint32_t)ITEMS : n) - 1u;
    v0 = pair_interval(sh, a.L, first, two, a.Aa, Aa_d, Ab_d, o);
    v1 = pair_interval(sh, a.L, jlast, two, a.Aa, Aa_d, Ab_d, o);
  }
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    const uint32_t jl = first + i;
    const bool in0 = jl < n && jl - v0.st < v0.sz;
    const bool in1 = !in0 && jl < n && jl - v1.st < v1.sz;
    in[i] = in0 || in1;
    second[i] = !in0;
    if (!vec) {
#pragma unroll
      for (int k = 0; k < DIM; ++k) rk[i][k] = in[i] ? a.src[(uint64_t)k * P + o + jl] : 0u;
    }
  }
  // Phase 2: classes, destinations relative to the run starts, and the
  // thread's segmented run counts.  cnt0 packs the thread's unsegmented
  // counts of left-left (bits 0-11), left-right (12-23) and median (24-31)
  // elements, cnt1 right-left and right-right: a tile holds at most six
  // depth-L intervals (each of at least 2^fin_log - 1 >= 511 points), so
  // none overflows.
  // Selects and arithmetic only, no branches: the lanes of a wave take every
  // class.  The thread's counts are packed 12 bits a class (a tile holds at
  // most 2,048 elements): seg = the segmented run counts (LL, LR, RL, RR in
  // bits 0-47, bit 48: an interval starts here), cnt = the unsegmented counts
  // of the four runs and the medians (bits 48-59).
  constexpr int DB = (DA + 1) % 3;
  constexpr uint64_t kSegF = 1ull << 48;
  uint32_t cls[ITEMS], base[ITEMS];
  bool starts[ITEMS];
  uint64_t seg = 0, cnt = 0;
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    const uint32_t jl = first + i;
    const uint32_t ra = rk[i][DA], rb = rk[i][DB];
    const bool s1 = second[i];
    const uint32_t st = s1 ? v1.st : v0.st, mp = s1 ? v1.mp : v0.mp, rma = s1 ? v1.rma : v0.rma;
    const bool start = in[i] && jl == st;
    const bool med = ra == rma, left = ra < rma;
    const uint32_t rm2 = left ? (s1 ? v1.rml : v0.rml) : (s1 ? v1.rmr : v0.rmr);
    const uint32_t mpl = s1 ? v1.mpl : v0.mpl, mpr = s1 ? v1.mpr : v0.mpr, mp2 = left ? mpl : mpr;
    const bool med2 = two & !med & (rb == rm2), hi2 = two & (rb > rm2);
    uint32_t c = (left ? kLL : kRL) + (hi2 ? 1u : 0u);  // kLR = kLL + 1, kRR = kRL + 1
    c = (med | med2) ? kMedian : c;
    c = in[i] ? c : kOutside;
    const uint32_t run = left ? (hi2 ? mpl + 1u : st) : (hi2 ? mpr + 1u : mp + 1u);
    base[i] = o + (med ? mp : med2 ? mp2 : run);
    cls[i] = c;
    starts[i] = start;
    const uint64_t one = c < kOutside ? 1ull << (12u * c) : 0ull;  // the class's field
    seg = (start ? kSegF : seg) + (c < kMedian ? one : 0ull);
    cnt += one;
  }
  if (a.write_f) {
#pragma unroll
    for (int i = 0; i < ITEMS; ++i)
      if (in[i] && first + i == (second[i] ? v1.mp : v0.mp))
        a.F[o + (second[i] ? v1.hn : v0.hn)] = second[i] ? v1.fv : v0.fv;  // the depth-L node
  }

  // Block-wide scans of the packed counts: segmented (a flagged element
  // restarts the sum) for the destinations, plain for the staging slots.
  const auto seg_add = [](uint64_t x, uint64_t y) { return (y & kSegF) ? y : x + y; };
  const auto add = [](uint64_t x, uint64_t y) { return x + y; };
  const uint32_t lane = lane_id();
  const int wave = threadIdx.x / kWave;
  const uint64_t seg_incl = wave_scan64(seg, seg_add), cnt_incl = wave_scan64(cnt, add);
  uint64_t seg_excl = ((uint64_t)(uint32_t)__shfl_up((int)(uint32_t)(seg_incl >> 32), 1) << 32) |
                      (uint32_t)__shfl_up((int)(uint32_t)seg_incl, 1);
  if (lane == 0) seg_excl = 0;
  if (lane == kWave - 1) {
    s_wave[0][wave] = seg_incl;
    s_wave[1][wave] = cnt_incl;
  }
  __syncthreads();
  uint64_t wcarry = 0, tile_seg = 0, cnt_excl = cnt_incl - cnt, cnt_tot = 0;
#pragma unroll
  for (int w = 0; w < kThreads / kWave; ++w) {
    if (w < wave) {
      wcarry = seg_add(wcarry, s_wave[0][w]);
      cnt_excl += s_wave[1][w];
    }
    tile_seg = seg_add(tile_seg, s_wave[0][w]);
    cnt_tot += s_wave[1][w];
  }
  seg_excl = seg_add(wcarry, seg_excl);
  const auto unpack = [](uint64_t x) {
    return SegCount4{(uint32_t)(x >> 48) & 1u,
                     {(uint32_t)x & 0xfffu, (uint32_t)(x >> 12) & 0xfffu, (uint32_t)(x >> 24) & 0xfffu,
                      (uint32_t)(x >> 36) & 0xfffu}};
  };
  const SegCount4 tile_agg = unpack(tile_seg), excl = unpack(seg_excl);

  if (wave == 0) {
    // Windowed look-back over the tiles of the same event (as level_kernel),
    // two words per tile.
    SegCount4 carry{0u, {0u, 0u, 0u, 0u}};
    uint64_t w0, w1;
    if (lt == 0) {
      if (lane == 0) {
        seg4_pack(tile_agg, kStPrefix, w0, w1);
        publish(&a.desc[2 * bid + 1], w1);
        publish(&a.desc[2 * bid], w0);
      }
    } else {
      if (lane == 0) {
        seg4_pack(tile_agg, kStAggregate, w0, w1);
        publish(&a.desc[2 * bid + 1], w1);
        publish(&a.desc[2 * bid], w0);
      }
      for (uint32_t b0 = 0;; b0 += kWave) {
        const int64_t j = (int64_t)lt - 1 - (int64_t)b0 - (int64_t)lane;
        uint64_t* p = j >= 0 ? &a.desc[2 * (bid - (b0 + lane + 1u))] : nullptr;
        uint64_t x0 = kStPrefix, x1 = kStPrefix;
        if (p) {
          x0 = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          x1 = __hip_atomic_load(p + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        uint32_t fst;
        while (true) {
          const bool valid = (x0 & kStMask) != kStInvalid && (x0 & kStMask) == (x1 & kStMask);
          const bool stop = valid && ((x0 & kStMask) == kStPrefix || ((x0 >> 60) & 1ull));
          const uint64_t sm = __ballot(stop);
          fst = sm ? (uint32_t)__builtin_ctzll(sm) : (uint32_t)kWave;
          const uint64_t need = fst >= (uint32_t)kWave - 1u ? ~0ull : (2ull << fst) - 1ull;
          if ((__ballot(!valid) & need) == 0ull) break;
          __builtin_amdgcn_s_sleep(1);
          if (!valid) {
            x0 = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            x1 = __hip_atomic_load(p + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          }
        }
        SegCount4 v = seg4_unpack(x0, x1);
        if (lane > fst) v = SegCount4{0u, {0u, 0u, 0u, 0u}};
#pragma unroll
        for (int ofs = 1; ofs < kWave; ofs <<= 1) {
          const SegCount4 u = seg4_shfl_down(v, ofs);
          if (lane + ofs < (uint32_t)kWave) v = seg4_op(u, v);
        }
        carry = seg4_op(v, carry);
        if (fst < (uint32_t)kWave) break;
      }
      if (lane == 0) {
        seg4_pack(seg4_op(carry, tile_agg), kStPrefix, w0, w1);
        publish(&a.desc[2 * bid + 1], w1);
        publish(&a.desc[2 * bid], w0);
      }
    }
    if (lane == 0) s_carry = carry;
  }
  __syncthreads();
  const SegCount4 run0 = seg4_op(s_carry, excl);
  uint32_t nLL = run0.c[0], nLR = run0.c[1], nRL = run0.c[2], nRR = run0.c[3];
  // Staging in write-out order: the tile's four runs, then its medians; in
  // tile order the destinations of each run rise, jumping only between
  // intervals.
  const uint32_t tLL = (uint32_t)cnt_tot & 0xfffu, tLR = (uint32_t)(cnt_tot >> 12) & 0xfffu;
  const uint32_t tRL = (uint32_t)(cnt_tot >> 24) & 0xfffu, tRR = (uint32_t)(cnt_tot >> 36) & 0xfffu;
  const uint32_t tMD = (uint32_t)(cnt_tot >> 48) & 0xfffu;
  uint32_t sLL = (uint32_t)cnt_excl & 0xfffu, sLR = tLL + ((uint32_t)(cnt_excl >> 12) & 0xfffu);
  uint32_t sRL = tLL + tLR + ((uint32_t)(cnt_excl >> 24) & 0xfffu);
  uint32_t sRR = tLL + tLR + tRL + ((uint32_t)(cnt_excl >> 36) & 0xfffu);
  uint32_t sMD = tLL + tLR + tRL + tRR + ((uint32_t)(cnt_excl >> 48) & 0xfffu);
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    const uint32_t c = cls[i];
    if (starts[i]) nLL = nLR = nRL = nRR = 0;
    const bool ll = c == kLL, lr = c == kLR, rl = c == kRL, rr = c == kRR, md = c == kMedian;
    const uint32_t dest = base[i] + (ll ? nLL : lr ? nLR : rl ? nRL : rr ? nRR : 0u);
    const uint32_t sl = ll ? sLL : lr ? sLR : rl ? sRL : rr ? sRR : sMD;
    nLL += ll ? 1u : 0u;
    nLR += lr ? 1u : 0u;
    nRL += rl ? 1u : 0u;
    nRR += rr ? 1u : 0u;
    sLL += ll ? 1u : 0u;
    sLR += lr ? 1u : 0u;
    sRL += rl ? 1u : 0u;
    sRR += rr ? 1u : 0u;
    sMD += md ? 1u : 0u;
    const uint32_t sp = c == kOutside ? kTile + kTile / 8 : sl + (sl >> 3);
    s_out[DIM][sp] = dest;
#pragma unroll
    for (int k = 0; k < DIM; ++k) s_out[k][sp] = rk[i][k];
  }
  __syncthreads();
  const uint32_t tn = tLL + tLR + tRL + tRR + tMD;
  for (uint32_t q = threadIdx.x; q < tn; q += kThreads) {
    const uint32_t qp = q + (q >> 3);
    const uint32_t dest = s_out[DIM][qp];
#pragma unroll
    for (int k = 0; k < DIM; ++k) a.dst[(uint64_t)k * P + dest] = s_out[k][qp];
  }
}

// Segmented counts of N arrays at once: one interval-start flag, and the
// left / right counts of every array.
template <int N>
struct SegCountN {
  uint32_t f, l[N], r[N];
};
template <int N>
__device__ inline SegCountN<N> seg_zero() {
  SegCountN<N> z;
  z.f = 0;
#pragma unroll
  for (int k = 0; k < N; ++k) z.l[k] = z.r[k] = 0;
  return z;
}
template <int N>
__device__ inline SegCountN<N> seg_op(const SegCountN<N>& a, const SegCountN<N>& b) {
  SegCountN<N> c;
  c.f = a.f | b.f;
#pragma unroll
  for (int k = 0; k < N; ++k) {
    c.l[k] = b.f ? b.l[k] : a.l[k] + b.l[k];
    c.r[k] = b.f ? b.r[k] : a.r[k] + b.r[k];
  }
  return c;
}
template <int N>
__device__ inline SegCountN<N> seg_shfl_up(const SegCountN<N>& v, int ofs) {
  SegCountN<N> u;
  u.f = __shfl_up(v.f, ofs);
#pragma unroll
  for (int k = 0; k < N; ++k) {
    u.l[k] = __shfl_up(v.l[k], ofs);
    u.r[k] = __shfl_up(v.r[k], ofs);
  }
  return u;
}

// Block-wide exclusive segmented scan of the threads' aggregates (every
// thread of an NT-thread block calls it); returns the thread's exclusive
// prefix.  s_wave holds NT / kWave entries; the caller syncs before reusing it.
template <int N, int NT = kThreads>
__device__ inline SegCountN<N> block_seg_excl(SegCountN<N> agg, SegCountN<N>* s_wave) {
  SegCountN<N> incl = agg;
  const uint32_t lane = lane_id();
#pragma unroll
  for (int ofs = 1; ofs < kWave; ofs <<= 1) {
    const SegCountN<N> u = seg_shfl_up(incl, ofs);
    if ((int)lane >= ofs) incl = seg_op(u, incl);
  }
  SegCountN<N> excl = seg_shfl_up(incl, 1);
  if (lane == 0) excl = seg_zero<N>();
  const int wave = threadIdx.x / kWave;
  if (lane == kWave - 1) s_wave[wave] = incl;
  __syncthreads();
  SegCountN<N> wcarry = seg_zero<N>();
#pragma unroll
  for (int w = 0; w < NT / kWave; ++w)
    if (w < wave) wcarry = seg_op(wcarry, s_wave[w]);
  return seg_op(wcarry, excl);
}

// Level launches that moved array m of an event before depth L0: every depth
// L < L0 whose split dimension L % dim is not m.  Array m of that event sits
// in buffer (moves & 1) when the finisher starts; events with fewer global
// depths than the launch count were skipped by the later launches.
__host__ __device__ inline uint32_t moves_before(uint32_t L0, uint32_t m, uint32_t dim) {
  const uint32_t same = L0 > m ? (L0 - 1u - m) / dim + 1u : 0u;
  return L0 - same;
}

// The same count under level_pair_kernel (dim 3): array 1 in the first
// launch (when depth 0 is built globally), and array m in launch L for every
// L < L0 with (L + 2) % 3 == m, i.e. L = m + 1 (mod 3).
__host__ __device__ inline uint32_t pair_moves_before(uint32_t L0, uint32_t m) {
  const uint32_t r = (m + 1u) % 3u;
  return ((m == 1u && L0 >= 1u) ? 1u : 0u) + (L0 > r ? (L0 - 1u - r) / 3u + 1u : 0u);
}

template <int DIM>
struct FinArgs {
  const uint32_t* planes[4][2];  // array m, buffer c: plane k (rank in dimension k) at planes[m][c] + k * P
  uint32_t* F;                // heap index -> rank in dimension 0 of the node's point
  TileMap fm;                 // workgroup -> event; local index = subtree of depth fin_depth(D)
  uint64_t P;                 // plane stride, plane_stride(points of the batch)
  uint32_t cap;               // most points of one subtree (fin_subtree_max over the events)
  uint32_t fin_log;           // the subtrees start at depth fin_depth(D, fin_log)
  uint32_t pairs;             // the depths above were built by level_pair_kernel
  // With out.S0, the tree arrays and search records of the subtree's nodes are
  // written here (out's S0, X, ids, P, Tc, Tid, recs; prec the point records)
  // instead of F, and workgroup t of an event also writes the event's top
  // node t (above the subtrees' depth) from the F the level launches wrote.
  LevelArgs<DIM> out;
  const float4* prec;
};

// Tree arrays and search record of heap node v of the event at o, whose point
// has dimension-0 rank `rank` (S0: rank -> point).
template <int DIM>
__device__ inline void out_node(const LevelArgs<DIM>& a, const float4* prec, uint32_t o, uint32_t v, uint32_t rank) {
  float c[DIM];
  uint32_t id;
  node_point<DIM>(a, a.S0[rank], prec, c, id);
  write_node<DIM>(a, o, v, floor_log2(v + 1u), c, id);
}

// The levels fin_depth(D) .. D-1 of one subtree in LDS: the same medians and
// stable partitions as level_kernel (FKDTree_cpu.h:229-268), with the
// subtree's points named by their index in array 0 (sorted by dimension 0)
// and their ranks kept in a table, so only 16-bit names move.  Writes F
// (heap index -> rank) for the subtree's nodes, leaves included, or the nodes'
// tree arrays and search records themselves (FinArgs::out).
// ITEMS consecutive positions per thread: subtrees of up to NT * ITEMS
// points.
// NT threads per workgroup (one to four waves).
template <int DIM, int ITEMS, int NT>
// Eight waves per SIMD (at most 64 registers) keep a whole launch of
// half-full subtrees resident (2,048 workgroups of 2 items per thread on 1M
// points); more items per thread would spill at that bound and keep the
// default (1,024 workgroups of 4 items need four waves per SIMD).
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(DIM <= 3 && ITEMS <= 2 ? 8 : 1))) void finish_kernel(FinArgs<DIM> a) {
  constexpr int kFinItems = ITEMS;
  constexpr int MA = DIM > 1 ? DIM - 1 : 1;  // arrays moved per depth
  __shared__ SegCountN<MA> s_wave[NT / kWave];
  extern __shared__ __align__(16) uint32_t fin_lds[];
  const uint32_t cap = a.cap;
  // s_rank(k, j): rank in dimension k of name j; s_arr(m, c, j): buffer c of array m
  uint32_t* const s_rank_base = fin_lds;
  uint16_t* const s_arr_base = reinterpret_cast<uint16_t*>(fin_lds + (uint64_t)DIM * cap);
  auto s_rank = [&](uint32_t k) { return s_rank_base + (uint64_t)k * cap; };
  auto s_arr = [&](uint32_t m, uint32_t c) { return s_arr_base + ((uint64_t)m * 2u + c) * (cap + 1u); };
  const uint32_t sub = xcd_item(blockIdx.x, gridDim.x);  // an XCD takes consecutive subtrees (whole events)
  const uint32_t e = a.fm.tile_ev[sub];
  const uint32_t t0 = sub - a.fm.tb[e];
  const uint32_t o = a.fm.off[e];
  const uint32_t n = a.fm.off[e + 1] - o;
  const Shape sh = shape_of(n);
  const uint32_t L0 = fin_depth(sh.D, a.fin_log);
  uint32_t start0, size0;
  subtree_range(sh, L0, t0, start0, size0);
  const uint64_t P = a.P;
  const uint32_t* pl[DIM];
#pragma unroll
  for (int m = 0; m < DIM; ++m)
    pl[m] = a.planes[m][(a.pairs ? pair_moves_before(L0, (uint32_t)m) : moves_before(L0, (uint32_t)m, (uint32_t)DIM)) & 1u];
  // All global loads of the subtree at once: array 0's ranks in every
  // dimension, and the other arrays' ranks in dimension 0 (for their names).
  constexpr int kS = ITEMS * (DIM - 1);
  uint32_t r0[kS > 0 ? kS : 1], lo[kS > 0 ? kS : 1];
  {
    uint32_t rr[ITEMS][DIM];
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
      const uint32_t j = threadIdx.x + (uint32_t)i * NT;
#pragma unroll
      for (int k = 0; k < DIM; ++k) rr[i][k] = j < size0 ? pl[0][(uint64_t)k * P + o + start0 + j] : 0u;
    }
#pragma unroll
    for (int q = 0; q < kS; ++q) {
      const uint32_t j = threadIdx.x + (uint32_t)(q % ITEMS) * NT;
      r0[q] = j < size0 ? pl[1 + q / ITEMS][o + start0 + j] : 0u;
      lo[q] = 0;
    }
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
      const uint32_t j = threadIdx.x + (uint32_t)i * NT;
      if (j < size0) {
#pragma unroll
        for (int k = 0; k < DIM; ++k) s_rank(k)[j] = rr[i][k];
        s_arr(0, 0)[j] = (uint16_t)j;
      }
    }
  }
  if (sh.D == 0) {  // a single point
    if (threadIdx.x == 0 && n == 1) {
      if (a.out.S0)
        out_node<DIM>(a.out, a.prec, o, 0u, pl[0][o]);
      else
        a.F[o] = pl[0][o];
    }
    return;
  }
  __syncthreads();
  // names of the other arrays' points: index of their dimension-0 rank in
  // array 0, by a binary search of fixed length (the largest power of two
  // not above size0 halved each step), all of a thread's searches in step
  // together so that their LDS reads are in flight at once
  {
    for (uint32_t step = 1u << floor_log2(size0); step > 0u; step >>= 1) {
      uint32_t v[kS > 0 ? kS : 1];
#pragma unroll
      for (int q = 0; q < kS; ++q) {  // every read issued before any is used
        const uint32_t idx = lo[q] + step - 1u;
        v[q] = s_rank(0)[idx < size0 ? idx : size0 - 1u];
      }
#pragma unroll
      for (int q = 0; q < kS; ++q)  // lo = number of entries below r0 seen so far
        lo[q] += (lo[q] + step <= size0 && v[q] < r0[q]) ? step : 0u;
    }
#pragma unroll
    for (int q = 0; q < kS; ++q) {
      const uint32_t j = threadIdx.x + (uint32_t)(q % ITEMS) * NT;
      if (j < size0) s_arr(1 + q / ITEMS, 0)[j] = (uint16_t)(lo[q] < size0 ? lo[q] : size0 - 1u);  // always found
    }
  }
  __syncthreads();
  uint32_t curbits = 0;  // bit m: the buffer of s_arr(m, .) that holds array m
  const uint32_t first = threadIdx.x * (uint32_t)kFinItems;
  // The interval of every item, followed down the depths instead of found
  // from its position each time: at depth L0 it is the whole subtree; below,
  // the left (right) son's range when the item lies before (after) the
  // median, none once the item is a node.  ts is the interval's index among
  // the nodes of its depth.
  uint32_t st[kFinItems], sz[kFinItems], ts[kFinItems], mp[kFinItems];
  // Every position becomes a node (or, at the last split, a leaf) at one
  // depth: its heap index is kept, and F is written once at the end from the
  // array that held the node at that depth.
  uint32_t fh[kFinItems];
  bool inside[kFinItems];
#pragma unroll
  for (int i = 0; i < kFinItems; ++i) {
    st[i] = start0;
    sz[i] = size0;
    ts[i] = t0;
    inside[i] = first + i < size0;
    fh[i] = 0;
  }
  for (uint32_t L = L0; L < sh.D; ++L) {
    const uint32_t d = L % (uint32_t)DIM;
    const bool last = L + 1u == sh.D;
    const uint16_t* med = s_arr(d, (curbits >> d) & 1u);
#pragma unroll
    for (int i = 0; i < kFinItems; ++i) {
      const uint32_t jl = start0 + first + i;
      mp[i] = st[i] + left_size(sz[i]);
      // node of depth L, or (last split) a leaf of depth D: the son on the
      // item's side of the median; array d holds it at this depth
      const bool placed = inside[i] && (jl == mp[i] || last);
      const uint32_t h = jl == mp[i] ? (1u << L) - 1u + ts[i] : (2u << L) - 1u + 2u * ts[i] + (jl > mp[i] ? 1u : 0u);
      fh[i] = placed ? h : fh[i];
    }
    const auto advance = [&]() {
#pragma unroll
      for (int i = 0; i < kFinItems; ++i) {
        const uint32_t jl = start0 + first + i;
        const bool right = jl > mp[i];
        inside[i] = inside[i] && jl != mp[i];
        sz[i] = right ? st[i] + sz[i] - mp[i] - 1u : mp[i] - st[i];
        st[i] = right ? mp[i] + 1u : st[i];
        ts[i] = 2u * ts[i] + (right ? 1u : 0u);
      }
    };
    if (last) continue;
    if (DIM == 1) {  // one array, split by position only
      advance();
      continue;
    }
    // The DIM-1 arrays other than d (array (d + 1 + k) % DIM for slot k) are
    // partitioned together: one segmented scan of all their counts.
    uint32_t rmed[kFinItems];  // rank in dimension d of the item's interval median
#pragma unroll
    for (int i = 0; i < kFinItems; ++i) rmed[i] = inside[i] ? s_rank(d)[med[mp[i] - start0]] : 0u;
    uint32_t cls[MA][kFinItems], dest0[MA][kFinItems], who[MA][kFinItems];
    if constexpr (MA <= 2) {
      // Packed counts, 12 bits each (a subtree holds fewer than 4,096
      // points): left / right of array slot k in bits 24k and 24k + 12, bit
      // 48 an interval start; classes by selects, the scan by DPP moves.
      constexpr uint64_t kSegF = 1ull << 48;
      uint64_t seg = 0;
#pragma unroll
      for (int i = 0; i < kFinItems; ++i) {
        const uint32_t jl = start0 + first + i;
        const bool start = inside[i] && jl == st[i];
        seg = start ? kSegF : seg;
#pragma unroll
        for (int k = 0; k < MA; ++k) {
          uint32_t m = d + 1u + (uint32_t)k;
          m = m >= (uint32_t)DIM ? m - (uint32_t)DIM : m;
          const uint32_t w = inside[i] ? s_arr(m, (curbits >> m) & 1u)[first + i] : 0u;
          const uint32_t r = s_rank(d)[w];
          const bool lo = r < rmed[i], hi = r > rmed[i];
          who[k][i] = w;
          cls[k][i] = !inside[i] ? kNone : (lo ? kLeft : hi ? kRight : kMed) | (start ? 4u : 0u);
          dest0[k][i] = lo ? st[i] : hi ? mp[i] + 1u : mp[i];
          seg += inside[i] ? ((lo ? 1ull : 0ull) << (24 * k)) + ((hi ? 1ull : 0ull) << (24 * k + 12)) : 0ull;
        }
      }
      const auto seg_add = [](uint64_t x, uint64_t y) { return (y & kSegF) ? y : x + y; };
      __shared__ uint64_t s_w64[NT / kWave];
      const uint32_t lane = lane_id();
      const int wave = threadIdx.x / kWave;
      const uint64_t incl = wave_scan64(seg, seg_add);
      uint64_t excl = ((uint64_t)(uint32_t)__shfl_up((int)(uint32_t)(incl >> 32), 1) << 32) |
                      (uint32_t)__shfl_up((int)(uint32_t)incl, 1);
      if (lane == 0) excl = 0;
      if (lane == kWave - 1) s_w64[wave] = incl;
      __syncthreads();
      uint64_t wcarry = 0;
#pragma unroll
      for (int w = 0; w < NT / kWave; ++w)
        if (w < wave) wcarry = seg_add(wcarry, s_w64[w]);
      excl = seg_add(wcarry, excl);
#pragma unroll
      for (int k = 0; k < MA; ++k) {
        uint32_t m = d + 1u + (uint32_t)k;
        m = m >= (uint32_t)DIM ? m - (uint32_t)DIM : m;
        uint16_t* dst = s_arr(m, ((curbits >> m) & 1u) ^ 1u);
        uint32_t nl = (uint32_t)(excl >> (24 * k)) & 0xfffu, nr = (uint32_t)(excl >> (24 * k + 12)) & 0xfffu;
#pragma unroll
        for (int i = 0; i < kFinItems; ++i) {
          const uint32_t c = cls[k][i];
          nl = (c & 4u) ? 0u : nl;
          nr = (c & 4u) ? 0u : nr;
          const bool l = (c & 3u) == kLeft, r = (c & 3u) == kRight;
          const uint32_t dest = dest0[k][i] + (l ? nl : r ? nr : 0u);
          nl += l ? 1u : 0u;
          nr += r ? 1u : 0u;
          // elements outside every interval write the spare slot at index cap
          dst[(c & 3u) != kNone ? dest - start0 : cap] = (uint16_t)who[k][i];
        }
      }
    } else {
    SegCountN<MA> agg = seg_zero<MA>();
#pragma unroll
    for (int i = 0; i < kFinItems; ++i) {
      const uint32_t j = first + i;
      const uint32_t jl = start0 + j;
      if (inside[i] && jl == st[i]) {
        agg = seg_zero<MA>();
        agg.f = 1u;
      }
#pragma unroll
      for (int k = 0; k < MA; ++k) {
        cls[k][i] = kNone;
        dest0[k][i] = 0;
        who[k][i] = 0;
        if (!inside[i]) continue;
        uint32_t m = d + 1u + (uint32_t)k;
        m = m >= (uint32_t)DIM ? m - (uint32_t)DIM : m;
        who[k][i] = s_arr(m, (curbits >> m) & 1u)[j];
        const uint32_t r = s_rank(d)[who[k][i]];
        uint32_t c;
        if (r < rmed[i]) {
          c = kLeft;
          dest0[k][i] = st[i];
          agg.l[k]++;
        } else if (r > rmed[i]) {
          c = kRight;
          dest0[k][i] = mp[i] + 1u;
          agg.r[k]++;
        } else {
          c = kMed;
          dest0[k][i] = mp[i];
        }
        if (jl == st[i]) c |= 4u;
        cls[k][i] = c;
      }
    }
    const SegCountN<MA> run0 = block_seg_excl<MA, NT>(agg, s_wave);
#pragma unroll
    for (int k = 0; k < MA; ++k) {
      uint32_t m = d + 1u + (uint32_t)k;
      m = m >= (uint32_t)DIM ? m - (uint32_t)DIM : m;
      uint16_t* dst = s_arr(m, ((curbits >> m) & 1u) ^ 1u);
      uint32_t nl = run0.l[k], nr = run0.r[k];
#pragma unroll
      for (int i = 0; i < kFinItems; ++i) {
        const uint32_t c = cls[k][i];
        if (c & 4u) {
          nl = 0;
          nr = 0;
        }
        uint32_t dest = dest0[k][i];
        switch (c & 3u) {
          case kLeft: dest += nl; ++nl; break;
          case kRight: dest += nr; ++nr; break;
          default: break;
        }
        if ((c & 3u) != kNone) dst[dest - start0] = (uint16_t)who[k][i];
      }
    }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < MA; ++k) {
      uint32_t m = d + 1u + (uint32_t)k;
      m = m >= (uint32_t)DIM ? m - (uint32_t)DIM : m;
      curbits ^= 1u << m;
    }
    advance();
  }
  // F from the array split at the depth where each position was placed (the
  // depth of its heap index; D - 1 for the leaves of depth D), in the buffer
  // it held then: array m moves on every depth of [L0, L) except those with
  // L % DIM == m, and starts in buffer 0.
  const auto same = [](uint32_t x, uint32_t m) { return (x + (uint32_t)DIM - 1u - m) / (uint32_t)DIM; };
  uint32_t rk[kFinItems];
#pragma unroll
  for (int i = 0; i < kFinItems; ++i) {
    const uint32_t j = first + i;
    if (j >= size0) continue;
    const uint32_t hd = floor_log2(fh[i] + 1u), L = hd < sh.D ? hd : sh.D - 1u, m = L % (uint32_t)DIM;
    const uint32_t moves = (L - L0) - (same(L, m) - same(L0, m));
    rk[i] = s_rank(0)[s_arr(m, moves & 1u)[j]];
    if (!a.out.S0) a.F[o + fh[i]] = rk[i];
  }
  if (a.out.S0) {
    // point, then its record, for all items before any is used; the writes
    // land in this subtree's heap ranges (one per depth) and merge in L2
    uint32_t pt[kFinItems];
#pragma unroll
    for (int i = 0; i < kFinItems; ++i) pt[i] = first + i < size0 ? a.out.S0[rk[i]] : 0u;
    float c[kFinItems][DIM];
    uint32_t id[kFinItems];
#pragma unroll
    for (int i = 0; i < kFinItems; ++i)
      if (first + i < size0) node_point<DIM>(a.out, pt[i], a.prec, c[i], id[i]);
#pragma unroll
    for (int i = 0; i < kFinItems; ++i)
      if (first + i < size0) write_node<DIM>(a.out, o, fh[i], floor_log2(fh[i] + 1u), c[i], id[i]);
    // one of the event's 2^L0 - 1 nodes above the subtrees per workgroup
    if (threadIdx.x == 0 && t0 + 1u < (1u << L0)) out_node<DIM>(a.out, a.prec, o, t0, a.F[o + t0]);
  }
}

// Bounding box of all points in fkey order (trees loaded without a build):
// per-wave reductions, one atomic per wave and bound.
template <int DIM>
__global__ void bbox_points_kernel(const float* __restrict__ X, uint32_t P, uint32_t* bbox) {
  uint32_t lo[DIM], hi[DIM];
#pragma unroll
  for (int m = 0; m < DIM; ++m) {
    lo[m] = 0xffffffffu;
    hi[m] = 0u;
  }
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < P; p += gridDim.x * blockDim.x) {
#pragma unroll
    for (int m = 0; m < DIM; ++m) {
      const uint32_t k = fkey(X[(uint64_t)m * P + p]);
      lo[m] = min(lo[m], k);
      hi[m] = max(hi[m], k);
    }
  }
#pragma unroll
  for (int m = 0; m < DIM; ++m) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) {
      lo[m] = min(lo[m], (uint32_t)__shfl_xor((int)lo[m], o));
      hi[m] = max(hi[m], (uint32_t)__shfl_xor((int)hi[m], o));
    }
    if (lane_id() == 0 && lo[m] <= hi[m]) {
      atomicMin(&bbox[m], lo[m]);
      atomicMax(&bbox[DIM + m], hi[m]);
    }
  }
}

// 1 in *err if the rebuilt tree arrays differ from the loaded ones.
__global__ void compare_kernel(const uint32_t* __restrict__ a, const uint32_t* __restrict__ b, uint64_t n,
                               uint32_t* err) {
  bool diff = false;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    diff |= a[i] != b[i];
  if (__ballot(diff) != 0ull && lane_id() == 0) atomicOr(err, 1u);
}

// Closed boxes [c - hw, c + hw] around every point of a batch, in input
// order (SoA [DIM][P], one float rounding per bound), and the event of every
// box: queries that search each point's neighbourhood in its own tree.
template <int DIM>
__global__ __launch_bounds__(kThreads) void boxes_around_kernel(const float* __restrict__ X, uint32_t P, float hw,
                                                                TileMap tm, float* __restrict__ mins,
                                                                float* __restrict__ maxs, uint32_t* __restrict__ qev) {
  const uint32_t t = blockIdx.x;
  const LevelTile lti = level_tile(tm, t);
  const uint32_t e = lti.e, lt = lti.lt, o = lti.o, n = lti.n;
  for (uint32_t j = lt * (uint32_t)kLevelTile + threadIdx.x; j < min(n, (lt + 1) * (uint32_t)kLevelTile);
       j += kThreads) {
    const uint64_t i = (uint64_t)o + j;
#pragma unroll
    for (int m = 0; m < DIM; ++m) {
      const float c = X[(uint64_t)m * P + i];
      mins[(uint64_t)m * P + i] = c - hw;
      maxs[(uint64_t)m * P + i] = c + hw;
    }
    if (qev) qev[i] = e;
  }
}

// Point records (coordinates, id) in input order, one 16-byte gather per
// node for the output (DIM <= 3).
template <int DIM>
__global__ void records_kernel(const float* __restrict__ X, const uint32_t* __restrict__ ids, uint32_t P,
                               float4* __restrict__ rec) {
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < P; p += gridDim.x * blockDim.x) {
    float4 r;
    r.x = X[p];
    r.y = DIM > 1 ? X[(uint64_t)(DIM > 1 ? 1 : 0) * P + p] : 0.f;
    r.z = DIM > 2 ? X[(uint64_t)(DIM > 2 ? 2 : 0) * P + p] : 0.f;
    r.w = __uint_as_float(ids[p]);
    rec[p] = r;
  }
}

// Tree arrays and search records: heap node v of every event holds the
// point S0[F[v]] (S0: rank in dimension 0 -> point), or point o + v when S0
// is null (a tree loaded from its arrays).  One thread per node, writes
// coalesced.
template <int DIM>
__global__ __launch_bounds__(kThreads) void output_kernel(LevelArgs<DIM> a, const float4* __restrict__ rec) {
  constexpr int IT = kLevelTile / kThreads;
  const uint32_t t = xcd_item(blockIdx.x, gridDim.x);  // an XCD takes consecutive tiles (whole events)
  const LevelTile lti = level_tile(a.tm, t);
  const uint32_t lt = lti.lt, o = lti.o, n = lti.n;
  const uint32_t v0 = lt * (uint32_t)kLevelTile + threadIdx.x;
  const uint32_t vend = min(n, (lt + 1u) * (uint32_t)kLevelTile);
  // each round of loads for all the thread's nodes before the next (F, then
  // the point index, then the point record), so they are in flight together
  uint32_t p[IT];
  bool ok[IT];
#pragma unroll
  for (int r = 0; r < IT; ++r) {
    const uint32_t v = v0 + (uint32_t)r * kThreads;
    ok[r] = v < vend;
    p[r] = o + v;  // S0 == null: the points already are the heap-ordered tree
  }
  if (a.S0) {
    uint32_t f[IT];
#pragma unroll
    for (int r = 0; r < IT; ++r) f[r] = ok[r] ? a.F[o + v0 + (uint32_t)r * kThreads] : a.P;
#pragma unroll
    for (int r = 0; r < IT; ++r) {
      ok[r] = ok[r] && f[r] < a.P;  // F is complete; the test keeps reads in range
      p[r] = ok[r] ? a.S0[f[r]] : 0u;
    }
  }
  float c[IT][DIM];
  uint32_t id[IT];
#pragma unroll
  for (int r = 0; r < IT; ++r) {
    id[r] = 0;
#pragma unroll
    for (int m = 0; m < DIM; ++m) c[r][m] = 0.f;
    if (ok[r]) node_point<DIM>(a, p[r], rec, c[r], id[r]);
  }
#pragma unroll
  for (int r = 0; r < IT; ++r) {
    const uint32_t v = v0 + (uint32_t)r * kThreads;
    if (ok[r]) write_node<DIM>(a, o, v, floor_log2(v + 1u), c[r], id[r]);
  }
}

}  // namespace fkd
