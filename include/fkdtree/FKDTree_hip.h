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
//    from many host threads (main.cpp:351-359) are coalesced: one caller
//    leads and searches every box pending at that moment in one
//    fkd_search_small launch (one wave per box, breadth first), the others
//    wait for their result, spinning; boxes whose result does not fit the per-box
//    output slot (which then grows) or whose walk is too wide go through
//    fkd_search.  Other calls are serialised by a mutex; the host copies
//    behind getDimensionVector / getIdVector are made once per build.
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

// Longest wait, in microseconds, of a leading search_in_the_box call for the
// other callers to queue their boxes before it launches.
#ifndef FKDTREE_HIP_GATHER_US
#define FKDTREE_HIP_GATHER_US 8
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
    if (tree_) fkd_destroy(tree_);
  }

  FKDTree_HIP(const FKDTree_HIP&) = delete;
  FKDTree_HIP& operator=(const FKDTree_HIP&) = delete;

  void build() override {
    std::lock_guard<std::mutex> g(mu_);
    check(fkd_build(tree_, stream_));
    check(fkd_stream_sync(stream_));
    built_ = true;
    mirror_ = false;
  }

  std::vector<unsigned int> search_in_the_box(const Point& minPoint, const Point& maxPoint) override {
    Pending r;
    for (unsigned d = 0; d < numberOfDimensions; ++d) {
      r.box[d] = minPoint[d];
      r.box[numberOfDimensions + d] = maxPoint[d];
    }
    {
      std::lock_guard<std::mutex> g(q_mu_);
      queue_.push_back(&r);
    }
    // Wait by spinning (a sleeping waiter takes tens of microseconds to wake,
    // longer than a launch), leading a launch whenever none is running.
    while (!r.done.load(std::memory_order_acquire)) {
      if (leading_.exchange(true, std::memory_order_acquire)) {
        std::this_thread::yield();
        continue;
      }
      // Callers released by the last launch come back within microseconds:
      // wait (at most FKDTREE_HIP_GATHER_US) until as many boxes are pending as that
      // launch held, so launches stay full instead of alternating with
      // nearly empty ones.
      const size_t want = expect_;
      const auto t0 = std::chrono::steady_clock::now();
      while (true) {
        {
          std::lock_guard<std::mutex> g(q_mu_);
          if (queue_.size() >= want) break;
        }
        if (std::chrono::steady_clock::now() - t0 > std::chrono::microseconds(FKDTREE_HIP_GATHER_US)) break;
        std::this_thread::yield();
      }
      std::vector<Pending*> batch;
      {
        std::lock_guard<std::mutex> g(q_mu_);
        const size_t k = std::min<size_t>(queue_.size(), FKD_SMALL_MAX_BOXES);
        batch.assign(queue_.begin(), queue_.begin() + k);
        queue_.erase(queue_.begin(), queue_.begin() + k);
      }
      expect_ = std::max<size_t>(batch.size(), 1);
      if (!batch.empty()) {
        run_batch(batch);
        launches_.fetch_add(1);
        launched_boxes_.fetch_add(batch.size());
      }
      for (Pending* p : batch) p->done.store(true, std::memory_order_release);
      leading_.store(false, std::memory_order_release);
    }
    if (r.error) std::rethrow_exception(r.error);
    return std::move(r.ids);
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

  // Launches made for search_in_the_box calls, and the boxes they carried.
  uint64_t coalesced_launches() const { return launches_.load(); }
  uint64_t coalesced_boxes() const { return launched_boxes_.load(); }

 private:
  static void check(int rc) {
    if (rc != FKD_OK) throw std::runtime_error(std::string("fkdtree: ") + fkd_last_error());
  }

  struct Pending {
    float box[2 * numberOfDimensions];
    std::vector<unsigned int> ids;
    std::exception_ptr error;
    std::atomic<bool> done{false};
  };

  // One fkd_search_small launch for the pending boxes; a box whose result
  // exceeded the slot grows the slot for later batches and, like a box whose
  // walk was too wide, is searched again through fkd_search.
  void run_batch(const std::vector<Pending*>& batch) {
    const size_t nb = batch.size();
    std::vector<float> boxes(nb * 2 * numberOfDimensions);
    for (size_t b = 0; b < nb; ++b)
      std::copy(batch[b]->box, batch[b]->box + 2 * numberOfDimensions, boxes.begin() + b * 2 * numberOfDimensions);
    std::vector<uint32_t> counts(nb);
    uint32_t slot = 0;
    try {
      {
        std::lock_guard<std::mutex> g(mu_);
        if (!built_) throw std::logic_error("fkdtree: search before build()");
        slot = std::min<uint32_t>(slot_, static_cast<uint32_t>((1u << 28) / std::max<size_t>(nb, 1)));
        small_ids_.resize(nb * (size_t)slot);
        check(fkd_search_small(tree_, static_cast<uint32_t>(nb), boxes.data(), slot, counts.data(),
                               small_ids_.data(), stream_));
        for (size_t b = 0; b < nb; ++b) {
          const uint32_t c = counts[b];
          if (c == FKD_SMALL_OVERFLOW || c > slot) continue;
          batch[b]->ids.assign(small_ids_.begin() + b * slot, small_ids_.begin() + b * slot + c);
        }
      }
      for (size_t b = 0; b < nb; ++b) {
        const uint32_t c = counts[b];
        if (c != FKD_SMALL_OVERFLOW && c <= slot) continue;
        if (c != FKD_SMALL_OVERFLOW) {
          std::lock_guard<std::mutex> g(mu_);
          uint32_t grow = slot_;
          while (grow < c && grow < kMaxSlot) grow *= 2;
          slot_ = std::max(slot_, grow);
        }
        std::vector<Point> lo(1), hi(1);
        for (unsigned d = 0; d < numberOfDimensions; ++d) {
          lo[0][d] = batch[b]->box[d];
          hi[0][d] = batch[b]->box[numberOfDimensions + d];
        }
        std::vector<uint64_t> off;
        search_csr(lo, hi, off, batch[b]->ids);
      }
    } catch (...) {
      for (Pending* p : batch) p->error = std::current_exception();
    }
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

  static constexpr uint32_t kMaxSlot = 1u << 16;  // ids per box in one small launch
  fkd_tree_t tree_ = nullptr;
  fkd_stream_t stream_ = nullptr;
  // coalescing of concurrent search_in_the_box calls
  std::mutex q_mu_;
  std::vector<Pending*> queue_;
  std::atomic<bool> leading_{false};
  size_t expect_ = 1;  // boxes of the last launch (read and written by the leader only)
  std::atomic<uint64_t> launches_{0}, launched_boxes_{0};
  uint32_t slot_ = 256;
  std::vector<uint32_t> small_ids_;
  size_t n_ = 0;
  bool built_ = false;
  bool mirror_ = false;
  std::mutex mu_;
  std::vector<TYPE> dims_[numberOfDimensions];
  std::vector<unsigned int> ids_;
};

#endif  // FKDTREE_AMD_FKDTREE_HIP_H_
