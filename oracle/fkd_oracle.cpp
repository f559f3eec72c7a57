// CPU oracle for the FKDTree hot paths.  TEST INFRASTRUCTURE ONLY: tests,
// the smoke check and the benchmark's CPU-baseline leg load this library to
// check results; the product path never calls it.
//
// It restates the reference's CPU tree (FKDTree_cpu.h) for float points of
// dimension 1..4:
//   * build                     FKDTree_cpu.h:221-277  (level loop, nth_element
//                               with the (coordinate, id) comparator :246-252,
//                               leaf copy :271-276)
//   * partition_complete_kdtree FKDTree_cpu.h:287-296
//   * add_at_position           FKDTree_cpu.h:280-285
//   * search_in_the_box         FKDTree_cpu.h:123-169  (level-synchronous BFS,
//                               push rule :150-163) with an unbounded frontier
//                               instead of the 1024-slot FQueue (FQueue.h:158)
//   * intersects / is_in_the_box FKDTree_cpu.h:302-321
//   * brute force               FKDTree.h:54-86 (closed-box membership)
//
// Pinning (each checked in tests/test_oracle.py): full tree arrays and search
// results written by the compiled reference (tests/golden), a brute-force
// builder (full sort of every interval) and the subtree-bounds invariant for
// the build, and a brute-force scan for the search.
// ORC_MUTATE (tests only, never set in a normal build) plants a known mistake
// so the tests can show that their pins catch it.
#ifndef ORC_MUTATE
#define ORC_MUTATE 0
#endif

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

inline unsigned floor_log2(uint64_t x) { return 63u - (unsigned)__builtin_clzll(x | 1); }

// Left-subtree size of a left-complete tree with `len` nodes
// (FKDTree_cpu.h:287-296).
inline uint64_t left_size(uint64_t len) {
  if (len == 1) return 0;
  uint64_t p = 1ull << floor_log2(len);
  return (p / 2 - 1 <= len - p) ? p - 1 : len - p / 2;
}

struct Pt {
  float x[4];
  uint32_t id;
};

// One tree: n points, SoA in, SoA heap-ordered out.
void build_one(int dim, uint64_t n, const float* coords, uint64_t cstride,
               const uint32_t* ids, float* out, uint64_t ostride, uint32_t* out_ids) {
  if (n == 0) return;
  std::vector<Pt> pts(n);
  for (uint64_t i = 0; i < n; ++i) {
    for (int d = 0; d < dim; ++d) pts[i].x[d] = coords[d * cstride + i];
    pts[i].id = ids[i];
  }
  const unsigned depth_max = floor_log2(n);
  std::vector<uint64_t> imin(n), ilen(n);
  imin[0] = 0;
  ilen[0] = n;
  auto store = [&](const Pt& p, uint64_t pos) {
    for (int d = 0; d < dim; ++d) out[d * ostride + pos] = p.x[d];
    out_ids[pos] = p.id;
  };
  for (unsigned depth = 0; depth < depth_max; ++depth) {
    const int d = (depth + (ORC_MUTATE == 2 ? 1 : 0)) % dim;
    const uint64_t first = (1ull << depth) - 1;
    for (uint64_t t = 0; t < (1ull << depth); ++t) {
      const uint64_t node = first + t;
      const uint64_t k = left_size(ilen[node]);
      auto b = pts.begin() + imin[node];
      std::nth_element(b, b + k, b + ilen[node], [d](const Pt& a, const Pt& c) {
        if (a.x[d] == c.x[d]) return ORC_MUTATE == 1 ? a.id > c.id : a.id < c.id;
        return a.x[d] < c.x[d];
      });
      store(pts[imin[node] + k], node);
      const uint64_t l = 2 * node + 1, r = l + 1;
      if (l < n) {
        imin[l] = imin[node];
        ilen[l] = k;
      }
      if (r < n) {
        imin[r] = imin[node] + k + 1;
        ilen[r] = ilen[node] - 1 - k;
      }
    }
  }
  for (uint64_t node = (1ull << depth_max) - 1; node < n; ++node) store(pts[imin[node]], node);
}

struct Box {
  float mn[4], mx[4];
};

// search_in_the_box (FKDTree_cpu.h:123-169): level-synchronous BFS.  `emit`
// receives every hit; returns the number of visited nodes.
template <class Emit>
uint64_t search_one(int dim, uint64_t n, const float* tree, uint64_t tstride,
                    const uint32_t* tids, const Box& b, std::vector<uint64_t>& cur,
                    std::vector<uint64_t>& nxt, Emit emit) {
  if (n == 0) return 0;
  uint64_t visited = 0;
  cur.clear();
  cur.push_back(0);
  const unsigned depth_max = floor_log2(n);
  for (unsigned depth = 0; depth < depth_max + 1 && !cur.empty(); ++depth) {
    const int d = depth % dim;
    nxt.clear();
    for (uint64_t idx : cur) {
      ++visited;
      const float v = tree[d * tstride + idx];
      const bool inter = v <= b.mx[d] && v >= b.mn[d];
      if (inter) {
        bool in = true;
        for (int e = 0; e < dim; ++e) {
          const float w = tree[e * tstride + idx];
          const bool below_max = ORC_MUTATE == 4 ? w < b.mx[e] : w <= b.mx[e];
          if (!(below_max && w >= b.mn[e])) {
            in = false;
            break;
          }
        }
        if (in) emit(tids[idx]);
      }
      const bool lower = v < b.mn[d];
      const int s0 = lower, s1 = ORC_MUTATE == 3 ? lower : (lower || inter);
      for (int s = s0; s <= s1; ++s) {
        const uint64_t c = 2 * idx + 1 + s;
        if (c < n) nxt.push_back(c);
      }
    }
    cur.swap(nxt);
  }
  return visited;
}

inline Box load_box(int dim, uint64_t nq, const float* mins, const float* maxs, uint64_t q) {
  Box b;
  for (int d = 0; d < dim; ++d) {
    b.mn[d] = mins[d * nq + q];
    b.mx[d] = maxs[d * nq + q];
  }
  return b;
}

inline void set_threads(int threads) {
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#else
  (void)threads;
#endif
}

}  // namespace

extern "C" {

int orc_version() { return 1; }

// Tree of n points; coords is SoA [dim][n]; outputs are heap-ordered SoA.
int orc_build(int dim, uint64_t n, const float* coords, const uint32_t* ids, float* out,
              uint32_t* out_ids) {
  if (dim < 1 || dim > 4) return -1;
  build_one(dim, n, coords, n, ids, out, n, out_ids);
  return 0;
}

// Independent trees, one per event; event e owns points [off[e], off[e+1]) of
// the SoA arrays (row stride = off[n_events]).  Events are built in parallel.
int orc_build_events(int dim, uint64_t n_events, const uint64_t* off, const float* coords,
                     const uint32_t* ids, float* out, uint32_t* out_ids, int threads) {
  if (dim < 1 || dim > 4) return -1;
  const uint64_t total = off[n_events];
  set_threads(threads);
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t e = 0; e < (int64_t)n_events; ++e) {
    const uint64_t a = off[e], n = off[e + 1] - a;
    build_one(dim, n, coords + a, total, ids + a, out + a, total, out_ids + a);
  }
  return 0;
}

// Level-BFS search of nq boxes (SoA mins/maxs [dim][nq]) in one tree.
// counts[q] = hits; visited[q] = nodes visited (either may be null).
int orc_search_count(int dim, uint64_t n, const float* tree, const uint32_t* tids, uint64_t nq,
                     const float* mins, const float* maxs, uint32_t* counts, uint32_t* visited,
                     int threads) {
  if (dim < 1 || dim > 4) return -1;
  set_threads(threads);
#pragma omp parallel
  {
    std::vector<uint64_t> cur, nxt;
#pragma omp for schedule(dynamic, 1024)
    for (int64_t q = 0; q < (int64_t)nq; ++q) {
      uint32_t c = 0;
      const uint64_t v = search_one(dim, n, tree, n, tids, load_box(dim, nq, mins, maxs, q), cur, nxt,
                                    [&](uint32_t) { ++c; });
      if (counts) counts[q] = c;
      if (visited) visited[q] = (uint32_t)v;
    }
  }
  return 0;
}

// Same traversal; writes the ids of query q (BFS order) at out[offsets[q]...].
int orc_search_fill(int dim, uint64_t n, const float* tree, const uint32_t* tids, uint64_t nq,
                    const float* mins, const float* maxs, const uint64_t* offsets, uint32_t* out,
                    int threads) {
  if (dim < 1 || dim > 4) return -1;
  set_threads(threads);
#pragma omp parallel
  {
    std::vector<uint64_t> cur, nxt;
#pragma omp for schedule(dynamic, 1024)
    for (int64_t q = 0; q < (int64_t)nq; ++q) {
      uint64_t w = offsets[q];
      search_one(dim, n, tree, n, tids, load_box(dim, nq, mins, maxs, q), cur, nxt,
                 [&](uint32_t id) { out[w++] = id; });
    }
  }
  return 0;
}

// Per-event search: query q searches the tree of event qev[q]; trees laid out
// as in orc_build_events.  counts/visited as in orc_search_count.
int orc_search_events_count(int dim, uint64_t n_events, const uint64_t* off, const float* tree,
                            const uint32_t* tids, uint64_t nq, const uint32_t* qev,
                            const float* mins, const float* maxs, uint32_t* counts,
                            uint32_t* visited, int threads) {
  if (dim < 1 || dim > 4) return -1;
  const uint64_t total = off[n_events];
  set_threads(threads);
#pragma omp parallel
  {
    std::vector<uint64_t> cur, nxt;
#pragma omp for schedule(dynamic, 1024)
    for (int64_t q = 0; q < (int64_t)nq; ++q) {
      const uint64_t a = off[qev[q]], n = off[qev[q] + 1] - a;
      uint32_t c = 0;
      const uint64_t v = search_one(dim, n, tree + a, total, tids + a,
                                    load_box(dim, nq, mins, maxs, q), cur, nxt,
                                    [&](uint32_t) { ++c; });
      if (counts) counts[q] = c;
      if (visited) visited[q] = (uint32_t)v;
    }
  }
  return 0;
}

int orc_search_events_fill(int dim, uint64_t n_events, const uint64_t* off, const float* tree,
                           const uint32_t* tids, uint64_t nq, const uint32_t* qev,
                           const float* mins, const float* maxs, const uint64_t* offsets,
                           uint32_t* out, int threads) {
  if (dim < 1 || dim > 4) return -1;
  const uint64_t total = off[n_events];
  set_threads(threads);
#pragma omp parallel
  {
    std::vector<uint64_t> cur, nxt;
#pragma omp for schedule(dynamic, 1024)
    for (int64_t q = 0; q < (int64_t)nq; ++q) {
      const uint64_t a = off[qev[q]], n = off[qev[q] + 1] - a;
      uint64_t w = offsets[q];
      search_one(dim, n, tree + a, total, tids + a, load_box(dim, nq, mins, maxs, q), cur, nxt,
                 [&](uint32_t id) { out[w++] = id; });
    }
  }
  return 0;
}

// Brute force (FKDTree.h:54-86): ids of all points inside each closed box, in
// point order.  Pass out == nullptr to get counts only.
int orc_brute(int dim, uint64_t n, const float* coords, const uint32_t* ids, uint64_t nq,
              const float* mins, const float* maxs, uint32_t* counts, const uint64_t* offsets,
              uint32_t* out, int threads) {
  if (dim < 1 || dim > 4) return -1;
  set_threads(threads);
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t q = 0; q < (int64_t)nq; ++q) {
    const Box b = load_box(dim, nq, mins, maxs, q);
    uint32_t c = 0;
    uint64_t w = out ? offsets[q] : 0;
    for (uint64_t i = 0; i < n; ++i) {
      bool in = true;
      for (int d = 0; d < dim; ++d) {
        const float v = coords[d * n + i];
        in &= (v <= b.mx[d] && v >= b.mn[d]);
      }
      if (in) {
        if (out) out[w++] = ids[i];
        ++c;
      }
    }
    if (counts) counts[q] = c;
  }
  return 0;
}

// Sorts every CSR segment [offsets[q], offsets[q+1]) of `ids` into `out`
// (ascending), so result sets can be compared independently of hit order.
void orc_sort_segments(uint64_t nq, const uint64_t* offsets, const uint32_t* ids, uint32_t* out,
                       int threads) {
  set_threads(threads);
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t q = 0; q < (int64_t)nq; ++q) {
    const uint64_t a = offsets[q], b = offsets[q + 1];
    std::copy(ids + a, ids + b, out + a);
    std::sort(out + a, out + b);
  }
}

int orc_max_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

}  // extern "C"
