// FKDTree_HIP<float, dim>: the MI355X backend of the FKDTree interface
// (FKDTree.h), a header-only C++ wrapper over the C ABI of fkd_capi.h
// (link with fkdtree_amd/libfkdtree_hip.so).  It takes the place of the
// reference's FKDTree_CPU / FKDTree_OpenCL classes (FKDTree_cpu.h:16-327,
// FKDTree_opencl.h:17-290) with the same constructor and virtual methods.
//
// Differences from the reference, by design:
//  * float only (static_assert); ids must be unique; NaN coordinates are
//    rejected at construction (std::runtime_error), as are HIP errors;
//  * search_in_the_box_multiple returns per-box counts, the only result the
//    reference implements (FKDTree_opencl.h:123-164);
//    search_in_the_box_multiple_csr returns every box's ids;
//  * the order of ids within one result is unspecified (the set is exact);
//  * calls on one object are thread safe.  Concurrent search_in_the_box calls
//    from many host threads (main.cpp:351-359) go to a resident search
//    kernel (fkd_box_server_*, started at the first call after a build): each
//    call writes its box into a free request slot in page-locked host memory
//    and spins until the kernel's wave has written the ids back, so a call
//    costs no launch and no stream synchronisation.  Boxes whose walk is too
//    wide for one wave or whose result exceeds a slot go through fkd_search.
//    Other calls are serialised by a mutex; the host copies behind
//    getDimensionVector / getIdVector are made once per build.
#ifndef FKDTREE_AMD_FKDTREE_HIP_H_
#define FKDTREE_AMD_FKDTREE_HIP_H_

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdint>
#include <exception>
#include <mutex>
#include <type_traits>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "FKDPoint.h"
#include "FKDTree.h"
#include "fkd_capi.h"

// Microseconds without a search_in_the_box call after which the resident
// search kernel leaves the GPU (the next call starts it again).
#ifndef FKDTREE_HIP_SERVER_IDLE_US
#define FKDTREE_HIP_SERVER_IDLE_US 20000
#endif

template <class TYPE, unsigned int numberOfDimensions>
class FKDTree_HIP : public FKDTree<TYPE, numberOfDimensions> {
  static_assert(std::is_same<TYPE, float>::value, "FKDTree_HIP supports float coordinates only");
  static_assert(numberOfDimensions >= 1 && numberOfDimensions <= 4, "dimension must be 1..4");
  typedef FKDPoint<TYPE, numberOfDimensions> Point;
  static_assert(sizeof(Point) == 4 * (numberOfDimensions + 1), "unexpected FKDPoint layout");

 public:
  explicit FKDTree_HIP(const std::vector<Point>& points, fkd_stream_t stream = nullptr)
      : FKDTree<TYPE, numberOfDimensions>(points), stream_(stream) {
    check(fkd_create(numberOfDimensions, static_cast<uint32_t>(points.size()),
                     points.empty() ? nullptr : points.data(), stream_, &tree_));
    n_ = points.size();
  }

  ~FKDTree_HIP() override {
    fkd_box_server_destroy(server_.exchange(nullptr));
    if (tree_) fkd_destroy(tree_);
  }

  FKDTree_HIP(const FKDTree_HIP&) = delete;
  FKDTree_HIP& operator=(const FKDTree_HIP&) = delete;

  void build() override {
    std::lock_guard<std::mutex> g(mu_);
    fkd_box_server_destroy(server_.exchange(nullptr));  // it reads the tree being rebuilt
    check(fkd_build(tree_, stream_));
    check(fkd_stream_sync(stream_));
    built_ = true;
    mirror_ = false;
  }

  std::vector<unsigned int> search_in_the_box(const Point& minPoint, const Point& maxPoint) override {
    float box[2 * numberOfDimensions];
    for (unsigned d = 0; d < numberOfDimensions; ++d) {
      box[d] = minPoint[d];
      box[numberOfDimensions + d] = maxPoint[d];
    }
    fkd_box_server_t srv = server();
    thread_local std::vector<unsigned int> buf;
    if (buf.size() < kSlotIds) buf.resize(kSlotIds);
    uint32_t c = 0;
    check(fkd_box_server_search(srv, box, buf.data(), kSlotIds, &c));
    if (c != FKD_SMALL_OVERFLOW && c <= kSlotIds) return std::vector<unsigned int>(buf.begin(), buf.begin() + c);
    // a walk too wide for one wave, or more ids than a slot holds
    std::vector<Point> lo(1, minPoint), hi(1, maxPoint);
    std::vector<uint64_t> off;
    std::vector<unsigned int> ids;
    search_csr(lo, hi, off, ids);
    return ids;
  }

  std::vector<unsigned int> search_in_the_box_multiple(const std::vector<Point>& minPoints,
                                                       const std::vector<Point>& maxPoints) override {
    std::vector<uint64_t> off;
    std::vector<unsigned int> ids;
    search_csr(minPoints, maxPoints, off, ids);
    std::vector<unsigned int> counts(minPoints.size());
    for (size_t q = 0; q < counts.size(); ++q) counts[q] = static_cast<unsigned int>(off[q + 1] - off[q]);
    return counts;
  }

  // The reference's other search entry points (FKDTree_cpu.h:25-121,
  // :171-207) return the same id set in other orders; here they are one
  // search.  (_branchless reads past the arrays on even N in the reference,
  // :58-60; this one does not.)
  std::vector<unsigned int> search_in_the_box_BFS(const Point& minPoint, const Point& maxPoint) {
    return search_in_the_box(minPoint, maxPoint);
  }
  std::vector<unsigned int> search_in_the_box_recursive(const Point& minPoint, const Point& maxPoint) {
    return search_in_the_box(minPoint, maxPoint);
  }
  std::vector<unsigned int> search_in_the_box_branchless(const Point& minPoint, const Point& maxPoint) {
    return search_in_the_box(minPoint, maxPoint);
  }

  // Wire format of the built tree (fkd_serialize) and a tree restored from
  // it without a build (fkd_deserialize; verify rebuilds and compares).
  std::vector<unsigned char> serialize() {
    std::lock_guard<std::mutex> g(mu_);
    uint64_t size = 0;
    check(fkd_serialize(tree_, nullptr, 0, &size, stream_));
    std::vector<unsigned char> buf(size);
    check(fkd_serialize(tree_, buf.data(), size, &size, stream_));
    return buf;
  }
  FKDTree_HIP(const std::vector<unsigned char>& wire, bool verify, fkd_stream_t stream = nullptr)
      : FKDTree<TYPE, numberOfDimensions>(std::vector<Point>()), stream_(stream) {
    check(fkd_deserialize(wire.data(), wire.size(), verify ? 1 : 0, stream_, &tree_));
    int dim = 0;
    uint32_t events = 0;
    uint64_t n = 0;
    check(fkd_tree_info(tree_, &dim, &events, &n));
    if (dim != (int)numberOfDimensions || events != 1) {
      fkd_destroy(tree_);
      tree_ = nullptr;
      throw std::invalid_argument("fkdtree: wire format holds another dimension or several events");
    }
    n_ = n;
    built_ = true;
  }

  // Ids of box q are ids[offsets[q] .. offsets[q+1]).
  void search_in_the_box_multiple_csr(const std::vector<Point>& minPoints, const std::vector<Point>& maxPoints,
                                      std::vector<uint64_t>& offsets, std::vector<unsigned int>& ids) {
    search_csr(minPoints, maxPoints, offsets, ids);
  }

  std::vector<TYPE> const& getDimensionVector(const int dimension) override {
    mirror();
    return dims_[(dimension >= 0 && static_cast<unsigned>(dimension) < numberOfDimensions) ? dimension : 0];
  }

  std::vector<unsigned int> const& getIdVector() override {
    mirror();
    return ids_;
  }

  fkd_tree_t handle() const { return tree_; }

  // Launches of the resident search kernel and the search_in_the_box calls
  // it answered, since the last build.
  uint64_t server_launches() const { return server_stat(0); }
  uint64_t server_requests() const { return server_stat(1); }
  // Mean search_in_the_box round trip and the walk inside it (microseconds),
  // and the requests posted again after a torn read of a request slot.
  void server_timing(double& round_trip_us, double& walk_us, uint64_t& reposts) const {
    round_trip_us = walk_us = 0.0;
    reposts = 0;
    fkd_box_server_t s = server_.load(std::memory_order_acquire);
    if (s) check(fkd_box_server_timing(s, &round_trip_us, &walk_us, &reposts));
  }

 private:
  static void check(int rc) {
    if (rc != FKD_OK) throw std::runtime_error(std::string("fkdtree: ") + fkd_last_error());
  }

  // The resident search kernel's handle, created at the first
  // search_in_the_box after a build.
  fkd_box_server_t server() {
    fkd_box_server_t s = server_.load(std::memory_order_acquire);
    if (s) return s;
    std::lock_guard<std::mutex> g(mu_);
    if (!built_) throw std::logic_error("fkdtree: search before build()");
    s = server_.load(std::memory_order_acquire);
    if (!s) {
      check(fkd_box_server_create(tree_, server_slots(), kSlotIds, FKDTREE_HIP_SERVER_IDLE_US, &s));
      server_.store(s, std::memory_order_release);
    }
    return s;
  }

  uint64_t server_stat(int which) const {
    fkd_box_server_t s = server_.load(std::memory_order_acquire);
    uint64_t launches = 0, requests = 0;
    if (s) check(fkd_box_server_stats(s, &launches, &requests, nullptr));
    return which == 0 ? launches : requests;
  }

  void search_csr(const std::vector<Point>& minPoints, const std::vector<Point>& maxPoints,
                  std::vector<uint64_t>& offsets, std::vector<unsigned int>& ids) {
    if (minPoints.size() != maxPoints.size()) throw std::invalid_argument("fkdtree: box count mismatch");
    const size_t nq = minPoints.size();
    std::vector<float> lo(nq * numberOfDimensions), hi(nq * numberOfDimensions);
    for (size_t q = 0; q < nq; ++q)
      for (unsigned d = 0; d < numberOfDimensions; ++d) {
        lo[d * nq + q] = minPoints[q][d];
        hi[d * nq + q] = maxPoints[q][d];
      }
    offsets.assign(nq + 1, 0);
    std::lock_guard<std::mutex> g(mu_);
    if (!built_) throw std::logic_error("fkdtree: search before build()");
    check(fkd_search_host(tree_, static_cast<uint32_t>(nq), lo.data(), hi.data(), nullptr, offsets.data(),
                          stream_));
    ids.resize(offsets[nq]);
    check(fkd_fetch_results(tree_, ids.empty() ? nullptr : ids.data(), stream_));
  }

  void mirror() {
    std::lock_guard<std::mutex> g(mu_);
    if (mirror_) return;
    std::vector<float> c(n_ * numberOfDimensions);
    ids_.assign(n_, 0);
    if (built_) check(fkd_get_tree(tree_, c.data(), ids_.data(), stream_));
    for (unsigned d = 0; d < numberOfDimensions; ++d)
      dims_[d].assign(c.begin() + d * n_, c.begin() + (d + 1) * n_);
    mirror_ = true;
  }

  // Concurrent search_in_the_box calls served: one per host hardware thread
  // (a waiting call spins on its core), 16 to 64; each slot holds a compute
  // unit while the resident kernel is live.
  static uint32_t server_slots() {
    const unsigned hw = std::thread::hardware_concurrency();
    return std::min<uint32_t>(FKD_SERVER_MAX_SLOTS, std::max<uint32_t>(16u, hw));
  }
  static constexpr uint32_t kSlotIds = 4096;               // ids per call answered by the resident kernel
  fkd_tree_t tree_ = nullptr;
  fkd_stream_t stream_ = nullptr;
  std::atomic<fkd_box_server_t> server_{nullptr};
  size_t n_ = 0;
  bool built_ = false;
  bool mirror_ = false;
  std::mutex mu_;
  std::vector<TYPE> dims_[numberOfDimensions];
  std::vector<unsigned int> ids_;
};

#endif  // FKDTREE_AMD_FKDTREE_HIP_H_
