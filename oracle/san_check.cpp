// Sanitizer driver for the CPU oracle and the header-only C++ interface.
// TEST INFRASTRUCTURE ONLY.  `make -C oracle sanitize` compiles it together
// with fkd_oracle.cpp under -fsanitize=address,undefined (no OpenMP), and
// tests/test_sanitize.py runs it: any out-of-bounds access, leak or
// undefined behaviour in the oracle's build, search, brute force and segment
// sort, or in FKDPoint.h / FKDTree.h (the validators test_correct_build and
// test_correct_search, FKDTree.h:24-86), aborts the run.
//
// Inputs cover the edge cases of the parity tests: sizes 1..4097 around
// powers of two, empty events, all-equal and signed-zero coordinates,
// infinities, and empty / inverted / whole-space / NaN boxes.  Each search is
// also compared with the brute-force scan, so the run checks results as well
// as memory safety.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <limits>
#include <vector>

#include "fkdtree/FKDPoint.h"
#include "fkdtree/FKDTree.h"

extern "C" {
int orc_build(int dim, uint64_t n, const float* coords, const uint32_t* ids, float* out, uint32_t* out_ids);
int orc_build_events(int dim, uint64_t n_events, const uint64_t* off, const float* coords, const uint32_t* ids,
                     float* out, uint32_t* out_ids, int threads);
int orc_search_count(int dim, uint64_t n, const float* tree, const uint32_t* tids, uint64_t nq,
                     const float* mins, const float* maxs, uint32_t* counts, uint32_t* visited, int threads);
int orc_search_fill(int dim, uint64_t n, const float* tree, const uint32_t* tids, uint64_t nq,
                    const float* mins, const float* maxs, const uint64_t* offsets, uint32_t* out, int threads);
int orc_search_events_count(int dim, uint64_t n_events, const uint64_t* off, const float* tree,
                            const uint32_t* tids, uint64_t nq, const uint32_t* qev, const float* mins,
                            const float* maxs, uint32_t* counts, uint32_t* visited, int threads);
int orc_search_events_fill(int dim, uint64_t n_events, const uint64_t* off, const float* tree,
                           const uint32_t* tids, uint64_t nq, const uint32_t* qev, const float* mins,
                           const float* maxs, const uint64_t* offsets, uint32_t* out, int threads);
int orc_brute(int dim, uint64_t n, const float* coords, const uint32_t* ids, uint64_t nq, const float* mins,
              const float* maxs, uint32_t* counts, const uint64_t* offsets, uint32_t* out, int threads);
void orc_sort_segments(uint64_t nq, const uint64_t* offsets, const uint32_t* ids, uint32_t* out, int threads);
}

namespace {

int g_fail = 0;
#define CHECK(c)                                                          \
  do {                                                                    \
    if (!(c)) {                                                           \
      std::fprintf(stderr, "san_check: %s failed at line %d\n", #c, __LINE__); \
      ++g_fail;                                                           \
    }                                                                     \
  } while (0)

uint64_t g_state = 0x5eed1234abcdULL;
uint64_t next_u64() {  // SplitMix64
  uint64_t z = (g_state += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
float next_unit() { return (float)(next_u64() >> 40) * (1.0f / 16777216.0f); }

// SoA coordinates [dim][n] of one of the edge-case kinds.
std::vector<float> points(int dim, uint64_t n, int kind) {
  std::vector<float> c((size_t)dim * n);
  const float specials[] = {std::numeric_limits<float>::infinity(), -std::numeric_limits<float>::infinity(),
                            0.0f, -0.0f, 1.0f, -1.0f, 3.0e38f};
  for (auto& x : c) {
    switch (kind) {
      case 0: x = next_unit(); break;
      case 1: x = 0.5f; break;                                    // all equal
      case 2: x = (next_u64() & 1) ? -0.0f : 0.0f; break;         // signed zeros
      case 3: x = specials[next_u64() % 7]; break;                // infinities, huge
      default: x = (float)(next_u64() % 5) - 2.0f; break;         // grid duplicates
    }
  }
  return c;
}

// Boxes around the first points, plus empty, inverted, whole-space and NaN boxes.
void boxes(int dim, const std::vector<float>& c, uint64_t n, std::vector<float>& mins, std::vector<float>& maxs,
           uint64_t& nq) {
  const uint64_t k = std::min<uint64_t>(n, 16);
  nq = k + 4;
  mins.assign((size_t)dim * nq, 0.f);
  maxs.assign((size_t)dim * nq, 0.f);
  const float inf = std::numeric_limits<float>::infinity(), nan = std::numeric_limits<float>::quiet_NaN();
  for (int d = 0; d < dim; ++d) {
    for (uint64_t q = 0; q < k; ++q) {
      mins[d * nq + q] = c[d * n + q] - 0.1f;
      maxs[d * nq + q] = c[d * n + q] + 0.1f;
    }
    mins[d * nq + k] = 1.0f, maxs[d * nq + k] = -1.0f;        // inverted
    mins[d * nq + k + 1] = -inf, maxs[d * nq + k + 1] = inf;  // whole space
    mins[d * nq + k + 2] = nan, maxs[d * nq + k + 2] = 1.0f;  // NaN bounds
    mins[d * nq + k + 3] = 0.0f, maxs[d * nq + k + 3] = nan;
  }
}

std::vector<uint64_t> csr(const std::vector<uint32_t>& counts) {
  std::vector<uint64_t> off(counts.size() + 1, 0);
  for (size_t i = 0; i < counts.size(); ++i) off[i + 1] = off[i] + counts[i];
  return off;
}

// A CPU-held tree behind the reference interface, so that the validators of
// FKDTree.h run on oracle arrays.
template <unsigned D>
class OracleTree : public FKDTree<float, D> {
 public:
  explicit OracleTree(const std::vector<FKDPoint<float, D>>& pts) : FKDTree<float, D>(pts), pts_(pts) {}
  void build() override {
    const uint64_t n = pts_.size();
    std::vector<float> c((size_t)D * n);
    std::vector<uint32_t> ids(n);
    for (uint64_t i = 0; i < n; ++i) {
      for (unsigned d = 0; d < D; ++d) c[d * n + i] = pts_[i][d];
      ids[i] = pts_[i].getId();
    }
    std::vector<float> out(c.size());
    tids_.assign(n, 0);
    orc_build(D, n, c.data(), ids.data(), out.data(), tids_.data());
    for (unsigned d = 0; d < D; ++d) dims_[d].assign(out.begin() + d * n, out.begin() + (d + 1) * n);
  }
  std::vector<unsigned int> search_in_the_box(const FKDPoint<float, D>& lo, const FKDPoint<float, D>& hi) override {
    float mn[D], mx[D];
    for (unsigned d = 0; d < D; ++d) mn[d] = lo[d], mx[d] = hi[d];
    std::vector<float> tree((size_t)D * pts_.size());
    for (unsigned d = 0; d < D; ++d) std::copy(dims_[d].begin(), dims_[d].end(), tree.begin() + d * pts_.size());
    uint32_t cnt = 0;
    orc_search_count(D, pts_.size(), tree.data(), tids_.data(), 1, mn, mx, &cnt, nullptr, 1);
    std::vector<unsigned int> ids(cnt);
    const uint64_t off[2] = {0, cnt};
    orc_search_fill(D, pts_.size(), tree.data(), tids_.data(), 1, mn, mx, off, ids.data(), 1);
    return ids;
  }
  std::vector<unsigned int> search_in_the_box_multiple(const std::vector<FKDPoint<float, D>>&,
                                                       const std::vector<FKDPoint<float, D>>&) override {
    return {};
  }
  const std::vector<float>& getDimensionVector(const int d) override {
    return dims_[(d >= 0 && (unsigned)d < D) ? d : 0];
  }
  const std::vector<unsigned int>& getIdVector() override { return tids_; }

 private:
  std::vector<FKDPoint<float, D>> pts_;
  std::vector<float> dims_[D];
  std::vector<unsigned int> tids_;
};

void check_one(int dim, uint64_t n, int kind) {
  const std::vector<float> c = points(dim, n, kind);
  std::vector<uint32_t> ids(n);
  for (uint64_t i = 0; i < n; ++i) ids[i] = (uint32_t)(n - 1 - i) * 3u + 1u;  // unique, decreasing
  std::vector<float> tree(c.size());
  std::vector<uint32_t> tids(n);
  CHECK(orc_build(dim, n, c.data(), ids.data(), tree.data(), tids.data()) == 0);
  std::vector<uint32_t> sorted_ids(tids);
  std::sort(sorted_ids.begin(), sorted_ids.end());
  std::vector<uint32_t> want(ids);
  std::sort(want.begin(), want.end());
  CHECK(sorted_ids == want);  // a permutation of the input ids
  std::vector<float> mins, maxs;
  uint64_t nq = 0;
  boxes(dim, c, n, mins, maxs, nq);
  std::vector<uint32_t> cnt(nq), vis(nq), bcnt(nq);
  orc_search_count(dim, n, tree.data(), tids.data(), nq, mins.data(), maxs.data(), cnt.data(), vis.data(), 1);
  const std::vector<uint64_t> off = csr(cnt);
  std::vector<uint32_t> out(off.back() + 1), sorted(off.back() + 1);
  orc_search_fill(dim, n, tree.data(), tids.data(), nq, mins.data(), maxs.data(), off.data(), out.data(), 1);
  orc_brute(dim, n, c.data(), ids.data(), nq, mins.data(), maxs.data(), bcnt.data(), nullptr, nullptr, 1);
  CHECK(bcnt == cnt);
  std::vector<uint32_t> bout(off.back() + 1), bsorted(off.back() + 1);
  orc_brute(dim, n, c.data(), ids.data(), nq, mins.data(), maxs.data(), bcnt.data(), off.data(), bout.data(), 1);
  orc_sort_segments(nq, off.data(), out.data(), sorted.data(), 1);
  orc_sort_segments(nq, off.data(), bout.data(), bsorted.data(), 1);
  CHECK(std::equal(sorted.begin(), sorted.begin() + off.back(), bsorted.begin()));
  CHECK(cnt[nq - 4] == 0 && cnt[nq - 2] == 0 && cnt[nq - 1] == 0);  // inverted, NaN
  CHECK(cnt[nq - 3] == n);                                           // whole space, infinities included
}

void check_events(int dim) {
  const std::vector<uint64_t> sizes = {0, 1, 7, 0, 300, 2, 1025};
  std::vector<uint64_t> off(1, 0);
  for (uint64_t s : sizes) off.push_back(off.back() + s);
  const uint64_t P = off.back(), E = sizes.size();
  const std::vector<float> c = points(dim, P, 4);
  std::vector<uint32_t> ids(P);
  for (uint64_t e = 0; e < E; ++e)
    for (uint64_t i = off[e]; i < off[e + 1]; ++i) ids[i] = (uint32_t)(off[e + 1] - i);
  std::vector<float> tree(c.size());
  std::vector<uint32_t> tids(P);
  CHECK(orc_build_events(dim, E, off.data(), c.data(), ids.data(), tree.data(), tids.data(), 1) == 0);
  // every event's tree equals its own single-tree build
  for (uint64_t e = 0; e < E; ++e) {
    const uint64_t a = off[e], n = off[e + 1] - a;
    std::vector<float> ce((size_t)dim * n), te((size_t)dim * n);
    std::vector<uint32_t> tie(n);
    for (int d = 0; d < dim; ++d) std::copy(c.begin() + d * P + a, c.begin() + d * P + a + n, ce.begin() + d * n);
    orc_build(dim, n, ce.data(), ids.data() + a, te.data(), tie.data());
    for (int d = 0; d < dim; ++d)
      CHECK(std::equal(te.begin() + d * n, te.begin() + (d + 1) * n, tree.begin() + d * P + a));
    CHECK(std::equal(tie.begin(), tie.end(), tids.begin() + a));
  }
  // self boxes of every point, searched in their own event
  std::vector<float> mins((size_t)dim * P), maxs((size_t)dim * P);
  std::vector<uint32_t> qev(P);
  for (uint64_t e = 0; e < E; ++e)
    for (uint64_t i = off[e]; i < off[e + 1]; ++i) qev[i] = (uint32_t)e;
  for (size_t i = 0; i < mins.size(); ++i) mins[i] = c[i] - 1.0f, maxs[i] = c[i] + 1.0f;
  std::vector<uint32_t> cnt(P), vis(P);
  orc_search_events_count(dim, E, off.data(), tree.data(), tids.data(), P, qev.data(), mins.data(), maxs.data(),
                          cnt.data(), vis.data(), 1);
  const std::vector<uint64_t> qoff = csr(cnt);
  std::vector<uint32_t> out(qoff.back() + 1);
  orc_search_events_fill(dim, E, off.data(), tree.data(), tids.data(), P, qev.data(), mins.data(), maxs.data(),
                         qoff.data(), out.data(), 1);
  for (uint64_t q = 0; q < P; ++q) CHECK(cnt[q] >= 1 && vis[q] >= cnt[q]);  // at least the point itself
}

template <unsigned D>
void check_interface(uint64_t n) {
  std::vector<FKDPoint<float, D>> pts(n);
  for (uint64_t i = 0; i < n; ++i) {
    for (unsigned d = 0; d < D; ++d) pts[i].setDimension((int)d, (float)(next_u64() % 64) * 0.125f);
    pts[i].setId((unsigned)i);
  }
  OracleTree<D> t(pts);
  t.build();
  CHECK(t.test_correct_build());
  CHECK(t.getIdVector().size() == n);
  CHECK(&t.getDimensionVector(-1) == &t.getDimensionVector(0));  // out of range: dimension 0
  for (uint64_t q = 0; q < std::min<uint64_t>(n, 32); ++q) {
    FKDPoint<float, D> lo = pts[q], hi = pts[q];
    for (unsigned d = 0; d < D; ++d) lo[(int)d] -= 1.0f, hi[(int)d] += 1.0f;
    const std::vector<unsigned> found = t.search_in_the_box(lo, hi);
    const bool good = test_correct_search<float, (int)D>(pts, found, lo, hi);
    CHECK(good);
  }
  const FKDPoint<float, 3> p = make_FKDPoint<float>(1.f, 2.f, 3.f, 9u);
  CHECK(p[0] == 1.f && p[2] == 3.f && p.getId() == 9u);
}

}  // namespace

int main() {
  const uint64_t sizes[] = {1, 2, 3, 4, 5, 7, 8, 9, 15, 16, 17, 31, 32, 33, 100, 1000, 4095, 4096, 4097};
  for (int dim = 1; dim <= 4; ++dim)
    for (uint64_t n : sizes)
      for (int kind = 0; kind < 5; ++kind) check_one(dim, n, kind);
  for (int dim = 1; dim <= 4; ++dim) check_events(dim);
  check_interface<1>(500);
  check_interface<2>(777);
  check_interface<3>(1024);
  check_interface<4>(333);
  if (g_fail) {
    std::fprintf(stderr, "san_check: %d checks failed\n", g_fail);
    return 1;
  }
  std::printf("san_check: all checks passed\n");
  return 0;
}
