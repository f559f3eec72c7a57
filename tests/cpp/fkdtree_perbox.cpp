// The reference driver's per-box search pattern on FKDTree_HIP<float, 3>:
// T host threads take box indices in turn and call search_in_the_box once
// per box (main.cpp:351-359, :372, whose tbb::parallel_for(0, N, 1) hands
// out single indices), after one build (main.cpp:201-205).  Every result is
// checked: its size against the expected count of every box, and its sorted
// ids against the expected set for the first boxes.  Prints one JSON line
// with the rate.
//   in: uint32 n, uint32 q, float pts[n][3], uint32 ids[n], float lo[q][3],
//       float hi[q][3], uint32 counts[q], uint32 k, uint64 off[k+1],
//       uint32 sorted_ids[off[k]]
//   usage: fkdtree_perbox in.bin threads
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <string>
#include <thread>
#include <vector>

#include "fkdtree/FKDTree_hip.h"

template <class T>
static bool read_vec(std::ifstream& in, std::vector<T>& v, size_t n) {
  v.resize(n);
  in.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(n * sizeof(T)));
  return (bool)in;
}

int main(int argc, char** argv) {
  if (argc != 3) {
    std::cerr << "usage: fkdtree_perbox in.bin threads\n";
    return 2;
  }
  const int threads = std::max(1, std::stoi(argv[2]));
  std::ifstream in(argv[1], std::ios::binary);
  uint32_t n = 0, q = 0, k = 0;
  in.read(reinterpret_cast<char*>(&n), 4);
  in.read(reinterpret_cast<char*>(&q), 4);
  std::vector<float> xyz, lo, hi;
  std::vector<uint32_t> ids, want_counts, want_ids;
  std::vector<uint64_t> want_off;
  bool ok = read_vec(in, xyz, 3 * (size_t)n) && read_vec(in, ids, n) && read_vec(in, lo, 3 * (size_t)q) &&
            read_vec(in, hi, 3 * (size_t)q) && read_vec(in, want_counts, q);
  in.read(reinterpret_cast<char*>(&k), 4);
  ok = ok && in && k <= q && read_vec(in, want_off, (size_t)k + 1);
  ok = ok && read_vec(in, want_ids, k ? want_off[k] : 0);
  if (!ok) {
    std::cerr << "short input\n";
    return 2;
  }
  std::vector<FKDPoint<float, 3>> points, mins, maxs;
  points.reserve(n);
  for (uint32_t i = 0; i < n; ++i) points.push_back(make_FKDPoint(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], ids[i]));
  for (uint32_t i = 0; i < q; ++i) {
    mins.push_back(FKDPoint<float, 3>(lo[3 * i], lo[3 * i + 1], lo[3 * i + 2], 0));
    maxs.push_back(FKDPoint<float, 3>(hi[3 * i], hi[3 * i + 1], hi[3 * i + 2], 0));
  }
  FKDTree_HIP<float, 3> tree(points);
  tree.build();
  FKDTree<float, 3>& base = tree;
  // warm-up: the first call starts the resident search kernel
  for (uint32_t i = 0; i < std::min<uint32_t>(q, 64); ++i) base.search_in_the_box(mins[i], maxs[i]);

  const uint64_t l0 = tree.server_launches(), r0 = tree.server_requests();
  std::vector<uint32_t> got(q, 0);
  std::vector<std::vector<unsigned int>> sample(k);
  std::atomic<uint32_t> next(0);
  const auto t0 = std::chrono::steady_clock::now();
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; ++t)
    pool.emplace_back([&]() {
      for (uint32_t i = next++; i < q; i = next++) {
        std::vector<unsigned int> r = base.search_in_the_box(mins[i], maxs[i]);
        got[i] = (uint32_t)r.size();
        if (i < k) sample[i] = std::move(r);
      }
    });
  for (auto& th : pool) th.join();
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  const uint64_t launches = tree.server_launches() - l0, answered = tree.server_requests() - r0;
  double round_trip_us = 0, walk_us = 0;
  uint64_t reposts = 0;
  tree.server_timing(round_trip_us, walk_us, reposts);

  uint64_t count_bad = 0, set_bad = 0, hits = 0;
  for (uint32_t i = 0; i < q; ++i) {
    hits += got[i];
    count_bad += got[i] != want_counts[i];
  }
  for (uint32_t i = 0; i < k; ++i) {
    std::vector<unsigned int>& r = sample[i];
    std::sort(r.begin(), r.end());
    const bool same = r.size() == want_off[i + 1] - want_off[i] &&
                      std::equal(r.begin(), r.end(), want_ids.begin() + (std::ptrdiff_t)want_off[i]);
    set_bad += !same;
  }
  std::printf(
      "{\"threads\": %d, \"boxes\": %u, \"points\": %u, \"seconds\": %.6f, \"box_queries_per_s\": %.1f, "
      "\"hits\": %llu, \"count_mismatches\": %llu, \"checked_sets\": %u, \"set_mismatches\": %llu, "
      "\"server_launches\": %llu, \"server_answered\": %llu, \"mean_round_trip_us\": %.2f, "
      "\"mean_walk_us\": %.2f, \"reposts\": %llu}\n",
      threads, q, n, secs, q / secs, (unsigned long long)hits, (unsigned long long)count_bad, k,
      (unsigned long long)set_bad, (unsigned long long)launches, (unsigned long long)answered, round_trip_us, walk_us,
      (unsigned long long)reposts);
  return (count_bad || set_bad) ? 1 : 0;
}
