// C entry point around the reference's copy of the CMSSW KD-tree
// (KDTreeLinkerAlgoT.h, KDTreeLinkerToolsT.h), built from those headers in
// place by oracle/Makefile into oracle/_ref/libcmsswref.so.
// TEST INFRASTRUCTURE ONLY: a side comparator for the CPU timing table, run
// the way main.cpp:560-629 runs it: nodes (index, x, y, z), the points'
// bounding cube as region, one build, then every box searched in turn on one
// thread, found points counted.  Its boxes are closed for the leaf test but
// its region-intersection test is strict (KDTreeLinkerAlgoT.h:195-204), so
// its count can differ from the closed-box hit count at boundaries.
#include <algorithm>
#include <array>
#include <chrono>
#include <cstdint>
#include <vector>

#include "KDTreeLinkerAlgoT.h"

extern "C" int cms_build_search(uint64_t n, const float* coords, uint64_t nq, const float* mins,
                                const float* maxs, double* build_ms, double* search_ms, uint64_t* found) {
  using clk = std::chrono::steady_clock;
  const auto t0 = clk::now();
  KDTreeLinkerAlgo<unsigned, 3> tree;
  std::vector<KDTreeNodeInfoT<unsigned, 3>> nodes;
  std::vector<KDTreeNodeInfoT<unsigned, 3>> founds;
  nodes.reserve(n);
  std::array<float, 3> lo{{0.f, 0.f, 0.f}}, hi{{0.f, 0.f, 0.f}};
  for (uint64_t i = 0; i < n; ++i) {
    const float x = coords[i], y = coords[n + i], z = coords[2 * n + i];
    nodes.emplace_back((unsigned)i, x, y, z);
    const float p[3] = {x, y, z};
    for (int d = 0; d < 3; ++d) {
      lo[d] = i == 0 ? p[d] : std::min(p[d], lo[d]);
      hi[d] = i == 0 ? p[d] : std::max(p[d], hi[d]);
    }
  }
  KDTreeCube bounds(lo[0], hi[0], lo[1], hi[1], lo[2], hi[2]);
  tree.build(nodes, bounds);
  const auto t1 = clk::now();
  uint64_t total = 0;
  for (uint64_t q = 0; q < nq; ++q) {
    KDTreeCube box(mins[q], maxs[q], mins[nq + q], maxs[nq + q], mins[2 * nq + q], maxs[2 * nq + q]);
    tree.search(box, founds);
    total += founds.size();
    founds.clear();
  }
  const auto t2 = clk::now();
  *build_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  *search_ms = std::chrono::duration<double, std::milli>(t2 - t1).count();
  *found = total;
  return 0;
}
