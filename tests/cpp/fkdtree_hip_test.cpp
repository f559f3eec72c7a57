// Drives FKDTree_HIP<float, 3> the way the reference's benchmark driver
// drives its trees (main.cpp:167-218 build + test_correct_build,
// main.cpp:342-367 search_in_the_box per point from several host threads,
// test_correct_search on a sample).  Input and output are raw binary files so
// that the Python test can compare the results with the oracle:
//   in : uint32 n, uint32 q, float pts[n][3], uint32 ids[n], float lo[q][3], float hi[q][3]
//   out: float tree[3][n], uint32 tree_ids[n], uint64 offsets[q+1], uint32 hits[...]
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <thread>
#include <vector>

#include "fkdtree/FKDTree_hip.h"

int main(int argc, char** argv) {
  if (argc != 3) {
    std::cerr << "usage: fkdtree_hip_test in.bin out.bin\n";
    return 2;
  }
  std::ifstream in(argv[1], std::ios::binary);
  uint32_t n = 0, q = 0;
  in.read(reinterpret_cast<char*>(&n), 4);
  in.read(reinterpret_cast<char*>(&q), 4);
  std::vector<float> xyz(3 * (size_t)n), lo(3 * (size_t)q), hi(3 * (size_t)q);
  std::vector<uint32_t> ids(n);
  in.read(reinterpret_cast<char*>(xyz.data()), xyz.size() * 4);
  in.read(reinterpret_cast<char*>(ids.data()), ids.size() * 4);
  in.read(reinterpret_cast<char*>(lo.data()), lo.size() * 4);
  in.read(reinterpret_cast<char*>(hi.data()), hi.size() * 4);
  if (!in) {
    std::cerr << "short input\n";
    return 2;
  }
  std::vector<FKDPoint<float, 3>> points, mins, maxs;
  for (uint32_t i = 0; i < n; ++i) points.push_back(make_FKDPoint(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], ids[i]));
  for (uint32_t i = 0; i < q; ++i) {
    mins.push_back(FKDPoint<float, 3>(lo[3 * i], lo[3 * i + 1], lo[3 * i + 2], 0));
    maxs.push_back(FKDPoint<float, 3>(hi[3 * i], hi[3 * i + 1], hi[3 * i + 2], 0));
  }

  FKDTree_HIP<float, 3> tree(points);
  tree.build();
  FKDTree<float, 3>& base = tree;
  if (!base.test_correct_build()) {
    std::cerr << "test_correct_build failed\n";
    return 1;
  }

  // search_in_the_box from 4 host threads, as main.cpp's parallel_for does
  std::vector<std::vector<unsigned int>> res(q);
  std::vector<std::thread> pool;
  for (int t = 0; t < 4; ++t)
    pool.emplace_back([&, t]() {
      for (uint32_t i = t; i < q; i += 4) res[i] = base.search_in_the_box(mins[i], maxs[i]);
    });
  for (auto& th : pool) th.join();

  int bad = 0;
  for (uint32_t i = 0; i < q && i < 64; ++i)
    if (!test_correct_search(points, res[i], mins[i], maxs[i])) ++bad;
  if (bad) {
    std::cerr << bad << " searches failed test_correct_search\n";
    return 1;
  }

  // batched forms agree with the per-box calls
  std::vector<uint64_t> off;
  std::vector<unsigned int> all;
  tree.search_in_the_box_multiple_csr(mins, maxs, off, all);
  std::vector<unsigned int> counts = base.search_in_the_box_multiple(mins, maxs);
  for (uint32_t i = 0; i < q; ++i) {
    if (counts[i] != res[i].size() || off[i + 1] - off[i] != res[i].size()) {
      std::cerr << "count mismatch at box " << i << "\n";
      return 1;
    }
  }

  // the reference's other search names give the same ids
  for (uint32_t i = 0; i < q && i < 16; ++i) {
    auto a = tree.search_in_the_box_BFS(mins[i], maxs[i]);
    auto b = tree.search_in_the_box_recursive(mins[i], maxs[i]);
    auto c = tree.search_in_the_box_branchless(mins[i], maxs[i]);
    if (a.size() != res[i].size() || b.size() != res[i].size() || c.size() != res[i].size()) {
      std::cerr << "variant mismatch at box " << i << "\n";
      return 1;
    }
  }

  // a tree restored from its wire format, unverified and verified, is the same tree
  const std::vector<unsigned char> wire = tree.serialize();
  for (int verify = 0; verify < 2; ++verify) {
    FKDTree_HIP<float, 3> back(wire, verify != 0);
    for (int d = 0; d < 3; ++d)
      if (back.getDimensionVector(d) != tree.getDimensionVector(d)) {
        std::cerr << "restored tree differs in dimension " << d << "\n";
        return 1;
      }
    if (back.getIdVector() != tree.getIdVector()) {
      std::cerr << "restored ids differ\n";
      return 1;
    }
    std::vector<uint64_t> off2;
    std::vector<unsigned int> all2;
    back.search_in_the_box_multiple_csr(mins, maxs, off2, all2);
    if (off2 != off) {
      std::cerr << "restored tree searches differently\n";
      return 1;
    }
  }

  std::ofstream out(argv[2], std::ios::binary);
  for (int d = 0; d < 3; ++d) out.write(reinterpret_cast<const char*>(base.getDimensionVector(d).data()), n * 4);
  out.write(reinterpret_cast<const char*>(base.getIdVector().data()), n * 4);
  std::vector<uint64_t> o2(q + 1, 0);
  for (uint32_t i = 0; i < q; ++i) o2[i + 1] = o2[i] + res[i].size();
  out.write(reinterpret_cast<const char*>(o2.data()), o2.size() * 8);
  for (uint32_t i = 0; i < q; ++i) out.write(reinterpret_cast<const char*>(res[i].data()), res[i].size() * 4);
  std::cout << "fkdtree_hip_test ok: n=" << n << " boxes=" << q << " hits=" << o2[q] << std::endl;
  return 0;
}
