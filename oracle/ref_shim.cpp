// C entry points around the reference's own CPU tree (FKDTree_cpu.h), built
// from the reference sources in place by oracle/Makefile into oracle/_ref/.
// TEST INFRASTRUCTURE ONLY: it writes the golden fixtures, cross-checks the
// oracle restatement, and serves as the "reference" CPU baseline.
//
// Wrapped reference interface:
//   ctor                 FKDTree_cpu.h:19-23
//   build()              FKDTree_cpu.h:221-277
//   getDimensionVector / getIdVector  FKDTree_cpu.h:209-213
//   search_in_the_box    FKDTree_cpu.h:123-169   (variant 0)
//   search_in_the_box_BFS FKDTree_cpu.h:72-121   (variant 1)
//   search_in_the_box_recursive FKDTree_cpu.h:171-207 (variant 2)
// search_in_the_box_branchless (FKDTree_cpu.h:25-70) is not wrapped: it reads
// past the arrays whenever N is even (it pushes firstSon+1 unchecked, :58-60).
#include <cstdint>
#include <memory>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "FKDTree_cpu.h"

namespace {

struct RefBase {
  virtual ~RefBase() {}
  virtual void build() = 0;
  virtual void get_tree(float* out, uint32_t* out_ids) = 0;
  virtual void search(uint64_t nq, const float* mins, const float* maxs, int variant,
                      uint32_t* counts) = 0;
  std::vector<std::vector<unsigned int>> results;
};

template <unsigned D>
struct RefTree : RefBase {
  std::unique_ptr<FKDTree_CPU<float, D>> tree;
  uint64_t n;

  RefTree(uint64_t n_, const float* coords, const uint32_t* ids) : n(n_) {
    std::vector<FKDPoint<float, D>> pts(n);
    for (uint64_t i = 0; i < n; ++i) {
      for (unsigned d = 0; d < D; ++d) pts[i].setDimension(d, coords[d * n + i]);
      pts[i].setId(ids[i]);
    }
    tree.reset(new FKDTree_CPU<float, D>(pts));
  }

  void build() override { tree->build(); }

  void get_tree(float* out, uint32_t* out_ids) override {
    for (unsigned d = 0; d < D; ++d) {
      const auto& v = tree->getDimensionVector(d);
      for (uint64_t i = 0; i < n; ++i) out[d * n + i] = v[i];
    }
    const auto& ids = tree->getIdVector();
    for (uint64_t i = 0; i < n; ++i) out_ids[i] = ids[i];
  }

  void search(uint64_t nq, const float* mins, const float* maxs, int variant,
              uint32_t* counts) override {
    results.assign(nq, {});
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t q = 0; q < (int64_t)nq; ++q) {
      FKDPoint<float, D> lo, hi;
      for (unsigned d = 0; d < D; ++d) {
        lo.setDimension(d, mins[d * nq + q]);
        hi.setDimension(d, maxs[d * nq + q]);
      }
      std::vector<unsigned int> r;
      if (variant == 0)
        r = tree->search_in_the_box(lo, hi);
      else if (variant == 1)
        r = tree->search_in_the_box_BFS(lo, hi);
      else
        tree->search_in_the_box_recursive(lo, hi, r);
      counts[q] = (uint32_t)r.size();
      results[q].swap(r);
    }
  }
};

}  // namespace

extern "C" {

void* ref_create(int dim, uint64_t n, const float* coords, const uint32_t* ids) {
  switch (dim) {
    case 1: return new RefTree<1>(n, coords, ids);
    case 2: return new RefTree<2>(n, coords, ids);
    case 3: return new RefTree<3>(n, coords, ids);
    case 4: return new RefTree<4>(n, coords, ids);
    default: return nullptr;
  }
}

void ref_build(void* h) { static_cast<RefBase*>(h)->build(); }

void ref_get_tree(void* h, float* out, uint32_t* out_ids) {
  static_cast<RefBase*>(h)->get_tree(out, out_ids);
}

// Runs variant (0 = search_in_the_box, 1 = _BFS, 2 = _recursive) for every
// box; counts[q] receives the hit count, the ids stay in the handle until
// ref_fetch copies them out (in the reference's traversal order).
int ref_search(void* h, uint64_t nq, const float* mins, const float* maxs, int variant,
               int threads, uint32_t* counts) {
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#endif
  if (variant < 0 || variant > 2) return -1;
  static_cast<RefBase*>(h)->search(nq, mins, maxs, variant, counts);
  return 0;
}

void ref_fetch(void* h, const uint64_t* offsets, uint32_t* out) {
  auto& res = static_cast<RefBase*>(h)->results;
  for (size_t q = 0; q < res.size(); ++q)
    for (size_t i = 0; i < res[q].size(); ++i) out[offsets[q] + i] = res[q][i];
}

void ref_destroy(void* h) { delete static_cast<RefBase*>(h); }

}  // extern "C"
