// Host runtime of the MI355X FKDTree library: tree objects, device memory,
// launch orchestration and the C ABI declared in include/fkdtree/fkd_capi.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <unordered_map>
#include <type_traits>
#include <thread>
#include <vector>

#include "build.hpp"
#include "common.hpp"
#include "fkdtree/fkd_capi.h"
#include "radix.hpp"
#include "scan.hpp"
#include "search.hpp"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define FKD_HIP(x)                                                                         \
  do {                                                                                     \
    hipError_t e_ = (x);                                                                   \
    if (e_ != hipSuccess) return fail(FKD_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define FKD_TRY(x)         \
  do {                     \
    int r_ = (x);          \
    if (r_ != FKD_OK) return r_; \
  } while (0)

constexpr uint32_t kMaxPoints = 1u << 30;
// Events per handle: n_events + 1 offsets stay far from any 32-bit wrap.
constexpr uint32_t kMaxEvents = 1u << 30;

// Stream on which the current call allocates and frees its buffers
// (stream-ordered allocator), set by every entry point that enqueues work;
// outside such a call (fkd_destroy) buffers are freed synchronously.
thread_local hipStream_t tl_stream = nullptr;
thread_local bool tl_ordered = false;

struct StreamScope {
  hipStream_t prev;
  bool prev_ordered;
  explicit StreamScope(hipStream_t s) : prev(tl_stream), prev_ordered(tl_ordered) {
    tl_stream = s;
    tl_ordered = true;
  }
  ~StreamScope() {
    tl_stream = prev;
    tl_ordered = prev_ordered;
  }
};

// Device memory cache of the library's stream-ordered allocations.
//
// A freed block stays with the library, with an event recorded on the stream
// that freed it, and is handed out again (to any stream) once that event has
// completed; a new block comes from the driver only when no cached block of
// the size class is free.  Blocks go back to the driver only when an
// allocation finds the device full, or on fkd_release_cached.  The runtime's
// stream-ordered pool cannot serve here: any synchronous free in the process
// (hipFree, hipHostFree, the runtime's staging of pageable copies) empties
// it, and a long-running server then mapped multi-gigabyte buffers from the
// driver for every batch and stalled for up to seconds.  Sizes are rounded
// up to four classes per power of two (at most 25% slack).  Each device has
// its own cache; the caches live until the process ends.
// Which runtime call of the last failed cache operation on this thread.
thread_local const char* tl_cache_call = "";

class DeviceCache {
 public:
  hipError_t alloc(void** out, size_t bytes) {
    const size_t cls = size_class(bytes);
    {
      std::lock_guard<std::mutex> g(mu_);
      auto range = free_.equal_range(cls);
      for (auto it = range.first; it != range.second; ++it) {
        const Block b = it->second;
        if (b.done) {
          const hipError_t q = hipEventQuery(b.done);
          if (q == hipErrorNotReady) {
            (void)hipGetLastError();  // "not ready" is an answer, not an error of a later call
            continue;
          }
          if (q != hipSuccess) {
            // An event whose stream was destroyed outside fkd_stream_destroy
            // can no longer answer: once the whole device is idle the block
            // is free regardless.
            (void)hipGetLastError();
            const hipError_t e = hipDeviceSynchronize();
            if (e != hipSuccess) {
              tl_cache_call = "hipDeviceSynchronize after an unanswerable cached-block event";
              return e;
            }
          }
          events_.push_back(b.done);
        }
        free_.erase(it);
        free_bytes_ -= cls;
        live_[b.p] = cls;
        live_bytes_ += cls;
        *out = b.p;
        return hipSuccess;
      }
    }
    hipError_t e = hipMalloc(out, cls);
    if (e == hipErrorOutOfMemory) {  // give the completed cached blocks back and try once more
      (void)hipGetLastError();
      (void)release();
      e = hipMalloc(out, cls);
    }
    if (e != hipSuccess) {
      tl_cache_call = "hipMalloc";
      return e;
    }
    std::lock_guard<std::mutex> g(mu_);
    live_[*out] = cls;
    live_bytes_ += cls;
    return hipSuccess;
  }

  // Stream-ordered free on `s` (ordered), or after the whole device is idle.
  hipError_t free(void* p, hipStream_t s, bool ordered) {
    if (!p) return hipSuccess;
    size_t cls = 0;
    hipEvent_t ev = nullptr;
    {
      std::lock_guard<std::mutex> g(mu_);
      auto it = live_.find(p);
      if (it == live_.end()) return hipErrorInvalidValue;
      cls = it->second;
      live_.erase(it);
      live_bytes_ -= cls;
      if (ordered && !events_.empty()) {
        ev = events_.back();
        events_.pop_back();
      }
    }
    hipError_t e = hipSuccess;
    if (ordered) {
      if (!ev) e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
      if (e == hipSuccess) e = hipEventRecord(ev, s);
      if (e != hipSuccess) tl_cache_call = "hipEventRecord at a stream-ordered free";
    } else {
      e = hipDeviceSynchronize();  // as hipFree would
      if (e != hipSuccess) tl_cache_call = "hipDeviceSynchronize at a free";
    }
    std::lock_guard<std::mutex> g(mu_);
    if (e != hipSuccess && ev) {  // the block is kept out of the cache: nothing tells when it is free
      events_.push_back(ev);
      return e;
    }
    free_.emplace(cls, Block{p, ev, ordered ? s : nullptr});
    free_bytes_ += cls;
    return e;
  }

  // The blocks freed on stream s need no event once s has drained (its
  // caller synchronised it): they are free, and no cached event refers to s
  // after it is destroyed.
  void retire_stream(hipStream_t s) {
    std::lock_guard<std::mutex> g(mu_);
    for (auto& kv : free_) {
      Block& b = kv.second;
      if (b.done && b.stream == s) {
        events_.push_back(b.done);
        b.done = nullptr;
        b.stream = nullptr;
      }
    }
  }

  // Cached blocks whose last use has completed go back to the driver.
  hipError_t release() {
    std::vector<void*> give;
    {
      std::lock_guard<std::mutex> g(mu_);
      for (auto it = free_.begin(); it != free_.end();) {
        const Block b = it->second;
        if (b.done) {
          const hipError_t q = hipEventQuery(b.done);
          if (q == hipErrorNotReady) {
            (void)hipGetLastError();
            ++it;
            continue;
          }
          events_.push_back(b.done);
        }
        free_bytes_ -= it->first;
        give.push_back(b.p);
        it = free_.erase(it);
      }
    }
    hipError_t r = hipSuccess;
    for (void* p : give) {
      const hipError_t e = hipFree(p);
      if (e != hipSuccess) r = e;
    }
    return r;
  }

  void info(uint64_t* held, uint64_t* used) {
    std::lock_guard<std::mutex> g(mu_);
    if (held) *held = live_bytes_ + free_bytes_;
    if (used) *used = live_bytes_;
  }

 private:
  struct Block {
    void* p;
    hipEvent_t done;     // recorded on the freeing stream; null: free at once
    hipStream_t stream;  // the freeing stream (while `done` is pending)
  };
  static size_t size_class(size_t bytes) {
    bytes = std::max<size_t>(bytes, 512);
    const unsigned k = 63u - (unsigned)__builtin_clzll(bytes);
    const size_t step = (size_t)1 << (k - 2);
    return (bytes + step - 1) & ~(step - 1);
  }
  std::mutex mu_;
  std::unordered_map<void*, size_t> live_;  // block -> its size class
  std::multimap<size_t, Block> free_;       // size class -> free blocks
  std::vector<hipEvent_t> events_;          // spare events
  uint64_t live_bytes_ = 0, free_bytes_ = 0;
};

DeviceCache* device_cache() {
  static std::mutex mu;
  static DeviceCache* caches[64] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
  std::lock_guard<std::mutex> g(mu);
  if (!caches[dev]) caches[dev] = new DeviceCache();  // kept until the process ends
  return caches[dev];
}

hipError_t dev_alloc(void** p, size_t bytes) { return device_cache()->alloc(p, bytes); }

hipError_t dev_free(void* p) { return device_cache()->free(p, tl_stream, tl_ordered); }

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  // Live HIP graphs of the owning handle (null: not a handle buffer).  A
  // captured graph holds this buffer's address, so while one exists the
  // buffer neither moves nor goes away.
  const uint32_t* graphs = nullptr;
  ~DevBuf() {
    if (p) (void)dev_free(p);
  }
  int ensure(size_t want) {
    if (want <= bytes) return FKD_OK;
    if (graphs && *graphs)
      return fail(FKD_ERR_STATE, "a HIP graph of this handle holds its buffers: this call needs larger ones "
                                 "(destroy the graph first)");
    if (p) {
      FKD_HIP(dev_free(p));
      p = nullptr;
      bytes = 0;
    }
    want = std::max<size_t>(want, 256);
    const hipError_t e = dev_alloc(&p, want);
    if (e != hipSuccess)
      return fail(FKD_ERR_HIP, std::string("device allocation (") + tl_cache_call + "): " + hipGetErrorString(e));
    bytes = want;
    return FKD_OK;
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(p);
  }
  void release() {
    if (p) (void)dev_free(p);
    p = nullptr;
    bytes = 0;
  }
};

// Page-locked host memory the kernels read and write directly (coherent,
// mapped): the staging of fkd_search_small.
struct PinnedBuf {
  void* h = nullptr;  // host address
  void* d = nullptr;  // the same memory as the device addresses it
  size_t bytes = 0;
  ~PinnedBuf() {
    if (h) (void)hipHostFree(h);
  }
  int ensure(size_t want) {
    if (want <= bytes) return FKD_OK;
    if (h) FKD_HIP(hipHostFree(h));
    h = d = nullptr;
    bytes = 0;
    want = std::max<size_t>(want, 4096);
    FKD_HIP(hipHostMalloc(&h, want, hipHostMallocCoherent | hipHostMallocMapped));
    FKD_HIP(hipHostGetDevicePointer(&d, h, 0));
    bytes = want;
    return FKD_OK;
  }
};

int launch_check(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(FKD_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
  return FKD_OK;
}

int fill_u32(uint32_t* p, uint32_t v, uint64_t n, hipStream_t s) {
  if (n == 0) return FKD_OK;
  const uint32_t blocks = (uint32_t)std::min<uint64_t>((n + 255) / 256, 4096);
  fkd::fill_u32_kernel<<<blocks, 256, 0, s>>>(p, v, n);
  return launch_check("fill_u32");
}

int fill_u64(uint64_t* p, uint64_t v, uint64_t n, hipStream_t s) {
  if (n == 0) return FKD_OK;
  const uint32_t blocks = (uint32_t)std::min<uint64_t>((n + 255) / 256, 4096);
  fkd::fill_u64_kernel<<<blocks, 256, 0, s>>>(p, v, n);
  return launch_check("fill_u64");
}

uint32_t grid_for(uint64_t n) { return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((n + 255) / 256, 8192)); }

// Event-aligned tiling: event e owns tiles [tb[e], tb[e+1]).
struct Tiling {
  uint32_t T = 0;
  DevBuf tile_ev, tb;
  std::vector<uint32_t> htb, hev;  // host sources of the uploads (kept until the handle goes)
  int init(const std::vector<uint64_t>& off, uint32_t tile, hipStream_t s) {
    const size_t E = off.size() - 1;
    htb.assign(E + 1, 0);
    hev.clear();
    uint32_t t = 0;
    for (size_t e = 0; e < E; ++e) {
      htb[e] = t;
      const uint64_t n = off[e + 1] - off[e];
      const uint32_t k = (uint32_t)((n + tile - 1) / tile);
      for (uint32_t i = 0; i < k; ++i) hev.push_back((uint32_t)e);
      t += k;
    }
    htb[E] = t;
    T = t;
    FKD_TRY(tb.ensure(htb.size() * 4));
    FKD_TRY(tile_ev.ensure(std::max<size_t>(hev.size(), 1) * 4));
    FKD_HIP(hipMemcpyAsync(tb.p, htb.data(), htb.size() * 4, hipMemcpyHostToDevice, s));
    if (!hev.empty())
      FKD_HIP(hipMemcpyAsync(tile_ev.p, hev.data(), hev.size() * 4, hipMemcpyHostToDevice, s));
    return FKD_OK;
  }
  // Tiles given per event.
  int init_counts(const std::vector<uint32_t>& per_event, hipStream_t s) {
    const size_t E = per_event.size();
    htb.assign(E + 1, 0);
    hev.clear();
    uint32_t t = 0;
    for (size_t e = 0; e < E; ++e) {
      htb[e] = t;
      for (uint32_t i = 0; i < per_event[e]; ++i) hev.push_back((uint32_t)e);
      t += per_event[e];
    }
    htb[E] = t;
    T = t;
    FKD_TRY(tb.ensure(htb.size() * 4));
    FKD_TRY(tile_ev.ensure(std::max<size_t>(hev.size(), 1) * 4));
    FKD_HIP(hipMemcpyAsync(tb.p, htb.data(), htb.size() * 4, hipMemcpyHostToDevice, s));
    if (!hev.empty())
      FKD_HIP(hipMemcpyAsync(tile_ev.p, hev.data(), hev.size() * 4, hipMemcpyHostToDevice, s));
    return FKD_OK;
  }
  fkd::TileMap map(const uint32_t* off) const {
    fkd::TileMap m;
    m.tile_ev = tile_ev.as<uint32_t>();
    m.tb = tb.as<uint32_t>();
    m.off = off;
    m.T = T;
    return m;
  }
};

}  // namespace

struct fkd_tuning {
  uint32_t query_key_bits = 8;  // search: Morton bits per dimension of the box order
  uint32_t fill_order = 1;      // fill pass: heaviest search blocks first (0 for handles of several events)
  uint32_t fill_window = 4096;  // fill pass: boxes reordered by work inside windows of this many (0: off)
  uint32_t count_window = 0;    // count pass: boxes reordered by volume inside windows of this many (0: off)
  uint32_t key_volume_bits = 0; // box order: bits of a box-volume class below the top Morton regions
  uint32_t fin_log = fkd::kFinLogDefault;  // build: subtrees of < 2^fin_log points are finished in LDS
  uint32_t build_radix_items = fkd::kRadixItems;  // build: keys per thread of the digit passes (8 or 16)
  uint32_t level_pairs = 1;     // build, dim 3: one array moved through two depths per launch
  uint32_t bucket = 0;          // fill pass, dim <= 3: bottom subtrees tested node by node (bucket_search_kernel):
                                // 0 walk every node, 1 lanes park at buckets, 2 lanes queue them
  uint32_t count_bucket = 0;    // count pass: the same choice
  uint32_t bucket_park = 8;     // bucketed search: eighths of the live lanes parked before the bucket tests run
};

struct fkd_tree_s {
  fkd_tree_s();
  fkd_tuning tune;
  int dim = 3;
  uint32_t n_events = 0;
  uint32_t P = 0;
  uint32_t Dmax = 0;
  uint32_t global_levels = 0;  // depths built by level_kernel; the rest by finish_kernel
  uint32_t fin_cap = 1;        // most points of one finisher subtree over all events
  uint32_t lay_fin_log = 0;    // fin_log and radix items the tilings were laid out for
  uint32_t lay_radix_items = 0;
  bool ids_in_order = true;
  bool built = false;
  std::vector<uint64_t> h_off;
  DevBuf off;                  // uint32 event offsets (n_events + 1)
  Tiling radix_tiles, level_tiles, fin_tiles;
  DevBuf X, ids;               // input copy
  DevBuf Tc, Tid, recs;        // outputs and search records
  DevBuf A, K, R, planes, F, Xr;   // build scratch
  DevBuf counts, offsets, err, bbox;
  DevBuf q_counts, q_off, q_hits, q_boxes, q_ev, q_aos;  // search scratch
  DevBuf q_keys, q_perm, q_sbox, q_rcounts, q_roffsets;
  DevBuf q_bwork, q_border;    // fill pass: work per search block and dispatch order
  DevBuf q_wslot;              // count pass: work per box in processing order ...
  bool q_work_valid = false;   // ... recorded by the last count pass (else the fill sorts nothing)
  DevBuf q_sorder;             // fill pass: dispatch slot -> processing slot
  DevBuf zarena;               // look-back descriptors and block counters, zeroed per build / search
  uint64_t last_capacity = 0;
  uint64_t* last_offsets = nullptr;
  hipEvent_t trace_start = nullptr, trace_stop = nullptr;  // around the traversal kernel
  uint32_t q_plan_nq = 0;  // queries whose processing order q_perm holds
  uint32_t q_perm_half = 0;  // which half of q_perm holds it
  uint64_t* q_scan_desc = nullptr;  // zeroed descriptors for the scan of the count pass
  uint64_t last_nq = 0, last_total = 0;
  // boxes of the plan in q_perm / q_sbox / q_wslot (fkd_search_fill must name the same)
  const float* plan_mins = nullptr;
  const float* plan_maxs = nullptr;
  const uint32_t* plan_qev = nullptr;
  PinnedBuf small_in, small_out;         // staging of fkd_search_small
  uint32_t graph_count = 0;              // live graphs captured on this handle
  std::vector<fkd_graph_s*> graph_list;  // ... detached when the handle goes first
  std::vector<fkd_box_server_s*> server_list;  // live box servers (they read recs): stopped and
                                               // detached when the handle goes first
  ~fkd_tree_s();

  // Every device buffer of the handle.
  std::vector<DevBuf*> buffers() {
    return {&off, &radix_tiles.tile_ev, &radix_tiles.tb, &level_tiles.tile_ev, &level_tiles.tb,
            &fin_tiles.tile_ev, &fin_tiles.tb, &X, &ids, &Tc, &Tid, &recs, &A, &K, &R, &Xr, &planes, &F,
            &counts, &offsets, &err, &bbox, &q_counts, &q_off, &q_hits, &q_boxes, &q_ev, &q_keys, &q_perm,
            &q_sbox, &q_rcounts, &q_roffsets, &q_bwork, &q_border, &q_wslot, &q_sorder, &q_aos, &zarena};
  }
  // The buffers a captured build + fkd_search can reference: all but the
  // host-search staging (q_boxes, q_ev, q_off, q_hits of fkd_search_host).
  std::vector<DevBuf*> graph_held() {
    std::vector<DevBuf*> v;
    for (DevBuf* b : buffers())
      if (b != &q_boxes && b != &q_ev && b != &q_off && b != &q_hits) v.push_back(b);
    return v;
  }
  // The build and search scratch fkd_trim releases.
  std::vector<DevBuf*> scratch() {
    return {&A, &K, &R, &planes, &F, &Xr, &counts, &offsets, &q_counts, &q_off, &q_hits, &q_boxes, &q_ev,
            &q_keys, &q_perm, &q_sbox, &q_rcounts, &q_roffsets, &q_bwork, &q_border, &q_wslot, &q_sorder,
            &q_aos, &zarena};
  }
};

struct fkd_graph_s {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  fkd_tree_s* tree = nullptr;  // null once the handle is destroyed
  uint32_t nq = 0;
  uint64_t* offsets = nullptr;
  uint64_t capacity = 0;
  ~fkd_graph_s() {
    if (tree) {
      --tree->graph_count;
      auto& l = tree->graph_list;
      l.erase(std::remove(l.begin(), l.end(), this), l.end());
    }
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
  }
};

// A resident per-box search kernel and its request / reply slots
// (box_server_kernel).  Host memory the kernel polls and writes is
// page-locked, coherent and mapped; the sequence numbers in it are the only
// synchronisation between the callers and the kernel.
struct fkd_box_server_s {
  fkd_tree_s* tree = nullptr;  // null once the tree is destroyed
  int dim = 3;
  uint32_t nslots = 0, slot_ids = 0;
  hipStream_t stream = nullptr;
  PinnedBuf req, rep, ids, word;  // BoxReq[nslots], BoxRep[nslots], ids[nslots][slot_ids], exit word
  DevBuf ctl;                     // BoxCtl, then done[nslots] (last sequence number served per slot)
  unsigned long long idle_ticks = 0;
  uint32_t max_idle_polls = 0;
  std::mutex launch_mu;
  std::atomic<uint32_t> gen{0};         // generation of the last launch (0: none yet)
  std::atomic<uint64_t> free_mask{0};   // slots no caller holds
  std::vector<uint32_t> seq;            // last sequence number posted per slot (by its holder)
  std::atomic<uint64_t> launches{0}, requests{0};
  std::atomic<uint64_t> reposts{0};        // requests posted again after a torn read (echo differed)
  std::atomic<uint64_t> round_trip_ns{0};  // summed over answered requests: post to reply, host clock
  std::atomic<uint64_t> walk_ticks{0};     // ... and the walks, device wall-clock ticks
  int rate_khz = 0;                        // device wall-clock rate
  bool stopped = false;

  fkd::BoxReq* h_req() const { return static_cast<fkd::BoxReq*>(req.h); }
  fkd::BoxRep* h_rep() const { return static_cast<fkd::BoxRep*>(rep.h); }
  uint32_t* h_exit() const { return static_cast<uint32_t*>(word.h); }
  bool running() const {
    return gen.load(std::memory_order_acquire) != __atomic_load_n(h_exit(), __ATOMIC_ACQUIRE);
  }
  // Every slot's next request is the stop value: the kernel's waves leave at
  // their next poll; then the stream drains.
  void stop() {
    if (stopped) return;
    stopped = true;
    for (uint32_t s = 0; s < nslots; ++s) __atomic_store_n(&h_req()[s].seq, fkd::kServerStop, __ATOMIC_RELEASE);
    if (stream) (void)hipStreamSynchronize(stream);
  }
  ~fkd_box_server_s() {
    stop();
    if (stream) (void)hipStreamDestroy(stream);
  }
};

fkd_tree_s::fkd_tree_s() {
  for (DevBuf* b : graph_held()) b->graphs = &graph_count;
}

fkd_tree_s::~fkd_tree_s() {
  for (fkd_graph_s* g : graph_list) g->tree = nullptr;
  for (fkd_box_server_s* v : server_list) {
    v->stop();
    v->tree = nullptr;
  }
}

namespace {

int ensure_build_scratch(fkd_tree_s* t);
int layout(fkd_tree_s* t, hipStream_t s);

// Zeroed words for the decoupled look-back scans of one build or search:
// sized up front, cleared by a single launch, handed out in slices.
struct ZeroArena {
  uint64_t* base = nullptr;
  uint64_t used = 0;
  uint64_t cap = 0;  // words cleared by zero_arena
  // The next `words` words, or null when they are not all inside the
  // cleared arena (callers fail instead of launching on such memory).
  uint64_t* take(uint64_t words) {
    if (!base || used + words > cap) return nullptr;
    uint64_t* p = base + used;
    used += words;
    return p;
  }
};

int arena_exhausted() { return fail(FKD_ERR_STATE, "internal: zeroed scratch arena too small for this launch"); }

int zero_arena(fkd_tree_s* t, uint64_t words, hipStream_t s, ZeroArena& za, bool clear = true);

// Words one radix pass takes from the arena: scan descriptors + counter.
uint64_t radix_pass_words(uint32_t M, uint32_t T) { return fkd::scan_tiles((uint64_t)M * 256u * T) + fkd::kCtrWords; }

// One stable LSD digit pass over M key arrays (count, scan, scatter) with
// ITEMS keys per thread; desc (scan_tiles of the digit counts) and ctr must
// be zero.
template <int ITEMS>
int radix_pass(const fkd::TileMap& tm, uint32_t P, uint32_t M, uint32_t* counts, uint32_t* offsets,
               uint64_t* desc, uint32_t* ctr, const uint32_t* kin, uint32_t* kout, const uint32_t* vin,
               uint32_t* vout, uint64_t stride, uint32_t* ranks, uint32_t shift, bool write_keys,
               bool write_ranks, hipStream_t s, uint32_t* bbox = nullptr, const fkd::XSource* xs = nullptr) {
  const uint64_t nc = (uint64_t)M * 256u * tm.T;
  dim3 grid(tm.T, M);
  if (xs) {
    // first pass of a build in tie order: keys from the coordinates (M = dim)
    if (M == 1) fkd::radix_count_kernel<ITEMS, 1><<<grid, fkd::kThreads, 0, s>>>(kin, stride, shift, tm, counts, *xs);
    else if (M == 2) fkd::radix_count_kernel<ITEMS, 2><<<grid, fkd::kThreads, 0, s>>>(kin, stride, shift, tm, counts, *xs);
    else if (M == 3) fkd::radix_count_kernel<ITEMS, 3><<<grid, fkd::kThreads, 0, s>>>(kin, stride, shift, tm, counts, *xs);
    else fkd::radix_count_kernel<ITEMS, 4><<<grid, fkd::kThreads, 0, s>>>(kin, stride, shift, tm, counts, *xs);
  } else {
    fkd::radix_count_kernel<ITEMS><<<grid, fkd::kThreads, 0, s>>>(kin, stride, shift, tm, counts, fkd::XSource{});
  }
  FKD_TRY(launch_check("radix_count"));
  const uint64_t st = fkd::scan_tiles(nc);
  fkd::scan_exclusive_kernel<uint32_t><<<(uint32_t)st, fkd::kThreads, 0, s>>>(counts, offsets, nc, desc, ctr);
  FKD_TRY(launch_check("radix_scan"));
  if (xs && write_keys && !write_ranks)
    fkd::radix_scatter_kernel<true, false, ITEMS, true><<<grid, fkd::kThreads, 0, s>>>(
        kin, kout, vin, vout, stride, ranks, P, shift, tm, offsets, bbox, xs->X);
  else if (xs)
    return fail(FKD_ERR_STATE, "internal: a pass from coordinates writes keys only");
  else if (write_keys && !write_ranks)
    fkd::radix_scatter_kernel<true, false, ITEMS><<<grid, fkd::kThreads, 0, s>>>(kin, kout, vin, vout, stride, ranks,
                                                                                 P, shift, tm, offsets, bbox, nullptr);
  else if (!write_keys && write_ranks)
    fkd::radix_scatter_kernel<false, true, ITEMS><<<grid, fkd::kThreads, 0, s>>>(kin, kout, vin, vout, stride, ranks,
                                                                                 P, shift, tm, offsets, bbox, nullptr);
  else if (write_keys && write_ranks)
    fkd::radix_scatter_kernel<true, true, ITEMS><<<grid, fkd::kThreads, 0, s>>>(kin, kout, vin, vout, stride, ranks,
                                                                                P, shift, tm, offsets, bbox, nullptr);
  else
    fkd::radix_scatter_kernel<false, false, ITEMS><<<grid, fkd::kThreads, 0, s>>>(kin, kout, vin, vout, stride,
                                                                                  ranks, P, shift, tm, offsets, bbox,
                                                                                  nullptr);
  return launch_check("radix_scatter");
}

template <int DIM>
int run_radix_pass(fkd_tree_s* t, ZeroArena& za, uint32_t M, const uint32_t* kin, uint32_t* kout,
                   const uint32_t* vin, uint32_t* vout, uint64_t stride, uint32_t* ranks, uint32_t shift,
                   bool write_keys, bool write_ranks, hipStream_t s, uint32_t* bbox = nullptr,
                   const fkd::XSource* xs = nullptr) {
  const fkd::TileMap tm = t->radix_tiles.map(t->off.as<uint32_t>());
  uint64_t* desc = za.take(radix_pass_words(M, tm.T));
  if (!desc) return arena_exhausted();
  uint32_t* ctr = reinterpret_cast<uint32_t*>(desc + fkd::scan_tiles((uint64_t)M * 256u * tm.T));
  if (t->lay_radix_items == 8)
    return radix_pass<8>(tm, t->P, M, t->counts.as<uint32_t>(), t->offsets.as<uint32_t>(), desc, ctr, kin, kout, vin,
                         vout, stride, ranks, shift, write_keys, write_ranks, s, bbox, xs);
  return radix_pass<16>(tm, t->P, M, t->counts.as<uint32_t>(), t->offsets.as<uint32_t>(), desc, ctr, kin, kout, vin,
                        vout, stride, ranks, shift, write_keys, write_ranks, s, bbox, xs);
}

// Level tiles of the handle; one tree needs no tile table (its kernels take
// event 0, offset 0 and the point count from the map itself).
fkd::TileMap level_map(fkd_tree_s* t) {
  fkd::TileMap m = t->level_tiles.map(t->off.as<uint32_t>());
  if (t->n_events == 1) {
    m.tile_ev = nullptr;
    m.single_n = t->P;
  }
  return m;
}

template <int DIM>
int build_impl(fkd_tree_s* t, hipStream_t s) {
  if (!t->server_list.empty())
    return fail(FKD_ERR_STATE, "a box server reads this tree: destroy it before building again");
  const uint32_t P = t->P;
  if (P == 0) {
    t->built = true;
    return FKD_OK;
  }
  if (t->lay_fin_log != t->tune.fin_log || t->lay_radix_items != t->tune.build_radix_items) {
    if (t->graph_count) return fail(FKD_ERR_STATE, "the build's tilings cannot change while a HIP graph exists");
    FKD_TRY(layout(t, s));
  }
  FKD_TRY(ensure_build_scratch(t));
  const uint32_t levels = t->global_levels;
  const fkd::TileMap tm = level_map(t);
  constexpr uint32_t M = DIM - 1;
  const uint32_t RT = t->radix_tiles.T;
  // dim 3: one launch per depth moves one array through two depths, plus a
  // first launch for array 1's lone move at depth 0 (level_pair_kernel)
  const bool pairs = DIM == 3 && t->tune.level_pairs && levels > 0;
  const uint64_t level_words =
      pairs ? (uint64_t)(levels + 1) * (2ull * tm.T + fkd::kCtrWords) : (uint64_t)levels * ((uint64_t)M * tm.T + fkd::kCtrWords);
  // In point order the first digit pass's count kernel clears the arena (no
  // kernel before it uses zeroed words); otherwise a fill launch does.
  ZeroArena za;
  FKD_TRY(zero_arena(t,
                     (t->ids_in_order ? 0 : 4 * radix_pass_words(1, RT)) + 4 * radix_pass_words(DIM, RT) + level_words,
                     s, za, !t->ids_in_order));
  const uint64_t S = P;  // stride of one array
  uint32_t* A = t->A.as<uint32_t>();  // A[m][2][P]
  uint32_t* K = t->K.as<uint32_t>();  // K[m][2][P]
  uint32_t* R = t->R.as<uint32_t>();  // R[p][4]: ranks of point p (first the id order, see 1.)
  auto Abuf = [&](int m, int c) { return A + ((uint64_t)m * 2 + c) * S; };
  auto Kbuf = [&](int m, int c) { return K + ((uint64_t)m * 2 + c) * S; };
  const float* X = t->X.as<float>();
  const uint32_t grid = grid_for(P);

  // 1. Tie order: if ids are not increasing inside every event, sort point
  //    indices by (event, id) first.
  const uint32_t* order = nullptr;
  if (!t->ids_in_order) {
    fkd::id_keys_kernel<<<grid, 256, 0, s>>>(t->ids.as<uint32_t>(), P, Kbuf(0, 0), Abuf(0, 0));
    FKD_TRY(launch_check("id_keys"));
    int c = 0;
    for (int pass = 0; pass < 4; ++pass) {
      const bool lastp = pass == 3;
      FKD_TRY(run_radix_pass<DIM>(t, za, 1, Kbuf(0, c), Kbuf(0, 1 - c), Abuf(0, c), lastp ? R : Abuf(0, 1 - c),
                                  S, nullptr, 8u * pass, !lastp, false, s));
      c = 1 - c;
    }
    order = R;  // R holds the (event, id) order until the last key pass writes the ranks
  }
  // 2. Keys of every coordinate, in tie order.
  // point records (search nodes' payload) with the keys when point order is
  // the tie order; otherwise records_kernel below
  float4* fused_rec = (DIM <= 3 && !order) ? t->Xr.as<float4>() : nullptr;
  // In point order the first digit pass reads the coordinates itself (keys,
  // positions, point records and the bounding box's starting values).
  const fkd::XSource xsrc{X, P, t->ids.as<uint32_t>(), fused_rec, t->bbox.as<uint32_t>(), za.base, za.cap};
  const fkd::XSource* first_pass_x = order ? nullptr : &xsrc;
  if (order) {
    fkd::keys_kernel<DIM><<<grid, 256, 0, s>>>(X, order, P, Kbuf(0, 0), Abuf(0, 0), 2 * S, t->err.as<uint32_t>(),
                                               t->bbox.as<uint32_t>(), t->ids.as<uint32_t>(), nullptr);
    FKD_TRY(launch_check("keys"));
    if (DIM <= 3) {  // the point records the finisher gathers
      fkd::records_kernel<DIM><<<grid, 256, 0, s>>>(X, t->ids.as<uint32_t>(), P, t->Xr.as<float4>());
      FKD_TRY(launch_check("records"));
    }
  }
  // 3. Stable LSD sort of every coordinate inside every event; the last pass
  //    stores each point's rank R[p][m] (the values end in buffer 0) and
  //    narrows the key-space bounding box (used to order search queries)
  //    from the starting values keys_kernel set.
  {
    int c = 0;
    for (int pass = 0; pass < 4; ++pass) {
      const bool lastp = pass == 3;
      FKD_TRY(run_radix_pass<DIM>(t, za, DIM, Kbuf(0, c), Kbuf(0, 1 - c), Abuf(0, c), Abuf(0, 1 - c), 2 * S,
                                  lastp ? R : nullptr, 8u * pass, !lastp, lastp, s,
                                  lastp ? t->bbox.as<uint32_t>() : nullptr, pass == 0 ? first_pass_x : nullptr));
      c = 1 - c;
    }
  }
  // 4. Rank planes: every array element carries its ranks in all dimensions.
  uint32_t* PL = t->planes.as<uint32_t>();  // PL[m][c][k][PS]
  const uint64_t PS = fkd::plane_stride(P);
  auto Pbuf = [&](int m, int c) { return PL + (((uint64_t)m * 2 + c) * DIM) * PS; };
  {
    const fkd::TileMap rtm = t->radix_tiles.map(t->off.as<uint32_t>());
    fkd::planes_kernel<DIM><<<rtm.T * DIM, fkd::kThreads, 0, s>>>(Abuf(0, 0), 2 * S, reinterpret_cast<const uint4*>(R),
                                                                  rtm, fkd::kThreads * t->lay_radix_items, Pbuf(0, 0), PS,
                                                                  2 * (uint64_t)DIM * PS);
  }
  FKD_TRY(launch_check("planes"));
  // 5. One partition launch per depth.
  int cur[4] = {0, 0, 0, 0};
  if (pairs) {
    // launch -1 moves array 1 through depth 0; launch L moves array (L + 2) % 3
    // through depths L and L + 1 (L alone where L + 1 is the finisher's)
    for (int64_t L = -1; L < (int64_t)levels; ++L) {
      fkd::PairArgs a;
      const uint32_t m = L < 0 ? 1u : (uint32_t)(L + 2) % 3u;
      a.L = L < 0 ? 0u : (uint32_t)L;
      a.dA = a.L % 3u;
      a.dB = (a.L + 1u) % 3u;
      a.single = L < 0 ? 1u : 0u;
      a.write_f = L < 0 ? 0u : 1u;
      a.fin_log = t->lay_fin_log;
      a.Aa = Pbuf((int)a.dA, cur[a.dA]);
      a.Ab = L < 0 ? a.Aa : Pbuf((int)a.dB, cur[a.dB]);
      a.src = Pbuf((int)m, cur[m]);
      a.dst = Pbuf((int)m, 1 - cur[m]);
      a.F = t->F.as<uint32_t>();
      a.tm = tm;
      a.PS = PS;
      a.desc = za.take(2ull * tm.T + fkd::kCtrWords);
      if (!a.desc) return arena_exhausted();
      a.ctr = reinterpret_cast<uint32_t*>(a.desc + 2ull * tm.T);
      if (a.dA == 0)
        fkd::level_pair_kernel<fkd::kLevelItems, 0><<<tm.T, fkd::kThreads, 0, s>>>(a);
      else if (a.dA == 1)
        fkd::level_pair_kernel<fkd::kLevelItems, 1><<<tm.T, fkd::kThreads, 0, s>>>(a);
      else
        fkd::level_pair_kernel<fkd::kLevelItems, 2><<<tm.T, fkd::kThreads, 0, s>>>(a);
      FKD_TRY(launch_check("level_pair"));
      cur[m] ^= 1;
    }
  }
  for (uint32_t L = 0; L < (pairs ? 0u : levels); ++L) {
    fkd::LevelArgs<DIM> a;
    a.L = L;
    a.d = L % DIM;
    a.fin_log = t->lay_fin_log;
    a.Ad = Pbuf(a.d, cur[a.d]);
    int k = 0;
    int moved[3] = {0, 0, 0};
    for (int m = 0; m < DIM; ++m) {
      if ((uint32_t)m == a.d) continue;
      a.src[k] = Pbuf(m, cur[m]);
      a.dst[k] = Pbuf(m, 1 - cur[m]);
      moved[k] = m;
      ++k;
    }
    for (; k < 3; ++k) {
      a.src[k] = nullptr;
      a.dst[k] = nullptr;
    }
    a.S0 = Abuf(0, 0);
    a.F = t->F.as<uint32_t>();
    a.tm = tm;
    a.X = X;
    a.ids = t->ids.as<uint32_t>();
    a.P = P;
    a.PS = PS;
    a.Tc = t->Tc.as<float>();
    a.Tid = t->Tid.as<uint32_t>();
    a.recs = t->recs.as<float>();
    a.desc = za.take((uint64_t)M * tm.T + fkd::kCtrWords);
    if (!a.desc) return arena_exhausted();
    a.ctr = reinterpret_cast<uint32_t*>(a.desc + (uint64_t)M * tm.T);
    const uint32_t blocks = (M ? M : 1u) * tm.T;
    fkd::level_kernel<DIM><<<blocks, fkd::kThreads, 0, s>>>(a);
    FKD_TRY(launch_check("level"));
    for (uint32_t i = 0; i < M; ++i) cur[moved[i]] ^= 1;
  }
  // 6. The depths whose subtrees fit a workgroup, one subtree per workgroup.
  if (t->fin_tiles.T) {
    fkd::FinArgs<DIM> f{};
    for (int m = 0; m < DIM; ++m) {
      f.planes[m][0] = Pbuf(m, 0);
      f.planes[m][1] = Pbuf(m, 1);
    }
    f.F = t->F.as<uint32_t>();
    f.fm = t->fin_tiles.map(t->off.as<uint32_t>());
    f.P = PS;
    f.cap = t->fin_cap;
    f.fin_log = t->lay_fin_log;
    f.pairs = pairs ? 1u : 0u;
    // the finisher writes the tree arrays and search records (step 7)
    f.out.S0 = Abuf(0, 0);
    f.out.X = X;
    f.out.ids = t->ids.as<uint32_t>();
    f.out.P = P;
    f.out.Tc = t->Tc.as<float>();
    f.out.Tid = t->Tid.as<uint32_t>();
    f.out.recs = t->recs.as<float>();
    f.prec = t->Xr.as<float4>();
    const uint32_t lds = fkd::fin_lds_bytes(DIM, f.cap);
    const uint32_t items = fkd::fin_items(f.cap);
    if (items == 1)
      fkd::finish_kernel<DIM, 1, fkd::kThreads><<<t->fin_tiles.T, fkd::kThreads, lds, s>>>(f);
    else if (items == 2)
      fkd::finish_kernel<DIM, 2, fkd::kThreads><<<t->fin_tiles.T, fkd::kThreads, lds, s>>>(f);
    else if (items == 4)
      fkd::finish_kernel<DIM, 4, fkd::kThreads><<<t->fin_tiles.T, fkd::kThreads, lds, s>>>(f);
    else
      fkd::finish_kernel<DIM, 8, fkd::kThreads><<<t->fin_tiles.T, fkd::kThreads, lds, s>>>(f);
    FKD_TRY(launch_check("finish"));
  }
  // 7. Tree arrays and search records from F, where no finisher wrote them.
  if (!t->fin_tiles.T) {
    fkd::LevelArgs<DIM> a{};
    a.tm = tm;
    a.S0 = Abuf(0, 0);
    a.F = t->F.as<uint32_t>();
    a.X = X;
    a.ids = t->ids.as<uint32_t>();
    a.P = P;
    a.Tc = t->Tc.as<float>();
    a.Tid = t->Tid.as<uint32_t>();
    a.recs = t->recs.as<float>();
    fkd::output_kernel<DIM><<<tm.T, fkd::kThreads, 0, s>>>(a, t->Xr.as<float4>());
    FKD_TRY(launch_check("output"));
  }
  t->built = true;
  t->q_plan_nq = 0;
  return FKD_OK;
}


template <int DIM, int MODE>
int launch_search(fkd_tree_s* t, const fkd::SearchArgs& a, hipStream_t s, const char* what) {
  fkd::SearchArgs b = a;
  const uint32_t blocks = std::max<uint32_t>(1, (a.nq + fkd::kThreads - 1) / fkd::kThreads);
  const uint64_t bytes = (uint64_t)t->P * fkd::NodeRec<DIM>::kFloats * 4;
  b.node_bytes = (uint32_t)std::min<uint64_t>(bytes, 0xffffffffull);
  b.bucket_park = t->tune.bucket_park;
  constexpr bool kBucketable = DIM <= 3 && MODE != fkd::kStats;
  const uint32_t sched = MODE == fkd::kFill ? t->tune.bucket : t->tune.count_bucket;
  if (kBucketable && sched) {
    constexpr int BD = DIM <= 3 ? DIM : 3;
    constexpr int BM = MODE == fkd::kStats ? fkd::kCount : MODE;
    const bool buf = bytes < (1ull << 32);
    if (sched == 1) {
      if (buf) fkd::bucket_search_kernel<BD, BM, true, 0><<<blocks, fkd::kThreads, 0, s>>>(b);
      else fkd::bucket_search_kernel<BD, BM, false, 0><<<blocks, fkd::kThreads, 0, s>>>(b);
    } else {
      constexpr int BQ = BM == fkd::kFill ? fkd::kBucketQueue : fkd::kBucketQueueCount;
      if (buf) fkd::bucket_search_kernel<BD, BM, true, BQ><<<blocks, fkd::kThreads, 0, s>>>(b);
      else fkd::bucket_search_kernel<BD, BM, false, BQ><<<blocks, fkd::kThreads, 0, s>>>(b);
    }
  } else if (MODE == fkd::kCount) {
    // the work figure the fill will sort by: none, nodes visited, or the
    // bucketed fill's work when the fill pass is bucketed
    const int wk = !b.work_slot ? 0 : (DIM <= 3 && t->tune.bucket ? 2 : 1);
    const bool buf = bytes < (1ull << 32);
    if (wk == 0) {
      if (buf) fkd::search_kernel<DIM, MODE, true, 0><<<blocks, fkd::kThreads, 0, s>>>(b);
      else fkd::search_kernel<DIM, MODE, false, 0><<<blocks, fkd::kThreads, 0, s>>>(b);
    } else if (wk == 1) {
      if (buf) fkd::search_kernel<DIM, MODE, true, 1><<<blocks, fkd::kThreads, 0, s>>>(b);
      else fkd::search_kernel<DIM, MODE, false, 1><<<blocks, fkd::kThreads, 0, s>>>(b);
    } else {
      if (buf) fkd::search_kernel<DIM, MODE, true, 2><<<blocks, fkd::kThreads, 0, s>>>(b);
      else fkd::search_kernel<DIM, MODE, false, 2><<<blocks, fkd::kThreads, 0, s>>>(b);
    }
  } else if (bytes < (1ull << 32)) {
    fkd::search_kernel<DIM, MODE, true, 0><<<blocks, fkd::kThreads, 0, s>>>(b);
  } else {
    fkd::search_kernel<DIM, MODE, false, 0><<<blocks, fkd::kThreads, 0, s>>>(b);
  }
  return launch_check(what);
}

template <int DIM>
fkd::SearchArgs search_args(fkd_tree_s* t, uint32_t nq, const uint32_t* qev) {
  fkd::SearchArgs a{};
  a.recs = t->recs.as<float>();
  a.off = t->n_events > 1 || qev ? t->off.as<uint32_t>() : nullptr;
  a.n_events = t->n_events;
  a.n_single = t->P;
  a.qev = qev;
  a.nq = nq;
  return a;
}

// Processing order of the queries (spatial sort of the box centres) and the
// boxes permuted into it.  Kept in the tree for the fill pass that follows.
// Processing slot -> query of the current plan.
uint32_t* plan_perm(fkd_tree_s* t) {
  return t->q_perm.as<uint32_t>() + (uint64_t)t->q_perm_half * std::max<uint32_t>(t->q_plan_nq, 1);
}

template <int DIM>
int plan_queries(fkd_tree_s* t, uint32_t nq, const float* mins, const float* maxs, const uint32_t* qev,
                 hipStream_t s) {
  const size_t n = std::max<uint32_t>(nq, 1);
  FKD_TRY(t->q_keys.ensure(n * 2 * 4));
  FKD_TRY(t->q_perm.ensure(n * 2 * 4));
  FKD_TRY(t->q_sbox.ensure(n * 2 * DIM * 4));
  const uint32_t T = (uint32_t)((nq + fkd::kRadixTile - 1) / fkd::kRadixTile);
  const uint64_t nc = 256ull * std::max<uint32_t>(T, 1);
  FKD_TRY(t->q_rcounts.ensure(nc * 4));
  FKD_TRY(t->q_roffsets.ensure((nc + 1) * 4));
  uint32_t* keys = t->q_keys.as<uint32_t>();
  uint32_t* perm = t->q_perm.as<uint32_t>();
  const uint32_t g = grid_for(nq);
  // Key: the event (when queries name one of several trees) above the
  // Morton bits; Morton bits are dropped until both fit in 32 bits.
  uint32_t kb = t->tune.query_key_bits;
  uint32_t vb = kb > 0 ? t->tune.key_volume_bits : 0u;
  uint32_t ev_bits = 0;
  if (qev && t->n_events > 1 && kb > 0)
    while (ev_bits < 32u && (1ull << ev_bits) < (uint64_t)t->n_events) ++ev_bits;
  vb = std::min<uint32_t>(vb, 32u - std::min<uint32_t>(ev_bits, 32u));
  if (kb > 0) kb = std::min<uint32_t>(kb, (32u - ev_bits - vb) / (uint32_t)DIM);
  const uint32_t ev_shift = kb * (uint32_t)DIM + vb;
  FKD_TRY(t->q_aos.ensure(n * 32));
  fkd::query_keys_kernel<DIM><<<g, 256, 0, s>>>(mins, maxs, nq, t->bbox.as<uint32_t>(), kb, vb,
                                                ev_bits ? qev : nullptr, ev_shift, keys, perm,
                                                t->q_aos.as<float4>());
  FKD_TRY(launch_check("query_keys"));
  fkd::TileMap tm{};
  tm.tile_ev = nullptr;
  tm.T = T;
  tm.single_n = nq;
  const int passes = (int)(ev_shift + ev_bits + 7) / 8;
  // one clear for the digit passes and the scan of the count pass
  ZeroArena za;
  FKD_TRY(zero_arena(t, passes * radix_pass_words(1, T) + fkd::scan_tiles(nq) + fkd::kCtrWords, s, za));
  int c = 0;
  for (int pass = 0; pass < passes; ++pass) {
    const bool lastp = pass == passes - 1;
    uint64_t* desc = za.take(radix_pass_words(1, T));
    if (!desc) return arena_exhausted();
    uint32_t* ctr = reinterpret_cast<uint32_t*>(desc + fkd::scan_tiles(nc));
    FKD_TRY(radix_pass<fkd::kRadixItems>(tm, nq, 1, t->q_rcounts.as<uint32_t>(), t->q_roffsets.as<uint32_t>(), desc, ctr,
                       keys + (uint64_t)c * n,
                       keys + (uint64_t)(1 - c) * n, perm + (uint64_t)c * n, perm + (uint64_t)(1 - c) * n, n,
                       nullptr, 8u * pass, !lastp, false, s));
    c = 1 - c;
  }
  t->q_perm_half = (uint32_t)c;
  t->q_scan_desc = za.take(fkd::scan_tiles(nq) + fkd::kCtrWords);
  if (!t->q_scan_desc) return arena_exhausted();
  float* smins = t->q_sbox.as<float>();
  float* smaxs = smins + (uint64_t)DIM * n;
  fkd::permute_boxes_kernel<DIM><<<g, 256, 0, s>>>(t->q_aos.as<float4>(), perm + (uint64_t)c * n, nq, smins,
                                                   smaxs);
  FKD_TRY(launch_check("permute_boxes"));
  t->q_plan_nq = nq;
  return FKD_OK;
}

// Dispatch order (t->q_sorder) with every window of `win` processing slots
// ordered by decreasing work[slot], and the work of every search block
// (t->q_bwork).
int window_sort(fkd_tree_s* t, uint32_t win, const uint32_t* work, uint32_t nq, hipStream_t s) {
  const uint32_t nb = (uint32_t)((nq + fkd::kThreads - 1) / fkd::kThreads);
  FKD_TRY(t->q_sorder.ensure((size_t)std::max<uint32_t>(nq, 1) * 4));
  FKD_TRY(t->q_bwork.ensure((size_t)std::max<uint32_t>(nb, 1) * 4));
  const uint32_t nw = (uint32_t)((nq + win - 1) / win);
  uint32_t* so = t->q_sorder.as<uint32_t>();
  uint32_t* bw = t->q_bwork.as<uint32_t>();
  if (win == 256) fkd::window_sort_kernel<1><<<nw, fkd::kThreads, 0, s>>>(work, nq, so, bw);
  else if (win == 1024) fkd::window_sort_kernel<4><<<nw, fkd::kThreads, 0, s>>>(work, nq, so, bw);
  else fkd::window_sort_kernel<16><<<nw, fkd::kThreads, 0, s>>>(work, nq, so, bw);
  return launch_check("window_sort");
}

template <int DIM>
int search_count_impl(fkd_tree_s* t, uint32_t nq, const float* mins, const float* maxs, const uint32_t* qev,
                      uint64_t* qoff, hipStream_t s) {
  FKD_TRY(t->q_counts.ensure((size_t)std::max<uint32_t>(nq, 1) * 4));
  if (nq > 0) {
    FKD_TRY(plan_queries<DIM>(t, nq, mins, maxs, qev, s));
    t->plan_mins = mins;
    t->plan_maxs = maxs;
    t->plan_qev = qev;
    fkd::SearchArgs a = search_args<DIM>(t, nq, qev);
    const size_t n = std::max<uint32_t>(nq, 1);
    a.perm = plan_perm(t);
    a.mins = t->q_sbox.as<float>();
    a.maxs = a.mins + (uint64_t)DIM * n;
    a.counts = t->q_counts.as<uint32_t>();
    FKD_TRY(t->q_wslot.ensure((size_t)std::max<uint32_t>(nq, 1) * 4));
    // The count pass records each box's work only for a fill that orders
    // boxes or blocks by it (fill_window, fill_order as set now).
    const uint32_t nb = (uint32_t)((nq + fkd::kThreads - 1) / fkd::kThreads);
    t->q_work_valid = (t->tune.fill_window != 0 && nq > t->tune.fill_window) || (t->tune.fill_order && nb >= 64);
    a.work_slot = t->q_work_valid ? t->q_wslot.as<uint32_t>() : nullptr;
    const uint32_t cw = t->tune.count_window;
    if (cw != 0 && nq > cw) {
      // Work is unknown before counting; the box volume (expected hits for
      // the handle's mean point density) stands in for it.  The count pass
      // overwrites the estimate with its own work figure.
      const float density = (float)t->P / (float)std::max<uint32_t>(t->n_events, 1u);
      uint32_t* est = t->q_wslot.as<uint32_t>();
      fkd::volume_work_kernel<DIM><<<grid_for(nq), 256, 0, s>>>(a.mins, a.maxs, nq, density, est);
      FKD_TRY(launch_check("volume_work"));
      FKD_TRY(window_sort(t, cw, est, nq, s));
      a.slot_order = t->q_sorder.as<uint32_t>();
    }
    FKD_TRY((launch_search<DIM, fkd::kCount>(t, a, s, "search_count")));
  }
  const uint64_t st = fkd::scan_tiles(nq);
  uint64_t* desc = t->q_scan_desc;
  if (nq == 0) {
    ZeroArena za;
    FKD_TRY(zero_arena(t, st + fkd::kCtrWords, s, za));
    desc = za.take(st + fkd::kCtrWords);
    if (!desc) return arena_exhausted();
  }
  fkd::scan_exclusive_kernel<uint64_t><<<(uint32_t)st, fkd::kThreads, 0, s>>>(
      t->q_counts.as<uint32_t>(), qoff, nq, desc, reinterpret_cast<uint32_t*>(desc + st));
  return launch_check("search_scan");
}

template <int DIM>
int search_stats_impl(fkd_tree_s* t, uint32_t nq, const float* mins, const float* maxs, const uint32_t* qev,
                      uint32_t* counts, uint32_t* visited, hipStream_t s) {
  if (nq == 0) return FKD_OK;
  fkd::SearchArgs a = search_args<DIM>(t, nq, qev);
  a.mins = mins;
  a.maxs = maxs;
  a.counts = counts;
  a.visited = visited;
  return launch_search<DIM, fkd::kStats>(t, a, s, "search_stats");
}

template <int DIM>
int search_fill_impl(fkd_tree_s* t, uint32_t nq, const float* mins, const float* maxs, const uint32_t* qev,
                     const uint64_t* qoff, uint32_t* out, uint64_t cap, hipStream_t s) {
  if (nq == 0) return FKD_OK;
  if (t->q_plan_nq != nq || t->plan_mins != mins || t->plan_maxs != maxs || t->plan_qev != qev)
    return fail(FKD_ERR_STATE, "fkd_search_fill must follow fkd_search_count with the same queries");
  fkd::SearchArgs a = search_args<DIM>(t, nq, qev);
  const size_t n = std::max<uint32_t>(nq, 1);
  a.perm = plan_perm(t);
  a.mins = t->q_sbox.as<float>();
  a.maxs = a.mins + (uint64_t)DIM * n;
  a.qoff = qoff;
  a.out = out;
  a.cap = cap;
  a.out_aligned = (reinterpret_cast<uintptr_t>(out) & 63u) == 0;
  const uint32_t nb = (uint32_t)((nq + fkd::kThreads - 1) / fkd::kThreads);
  const uint32_t win = t->tune.fill_window;
  // by the work the count pass recorded (none if the tuning asked for no
  // ordering when it ran: any order gives the same results)
  const bool sorted = win != 0 && nq > win && t->q_work_valid;
  const bool ordered = t->tune.fill_order && nb >= 64 && t->q_work_valid;
  if (sorted || ordered) FKD_TRY(t->q_bwork.ensure((size_t)nb * 4));
  if (sorted) {
    FKD_TRY(window_sort(t, win, t->q_wslot.as<uint32_t>(), nq, s));
    a.slot_order = t->q_sorder.as<uint32_t>();
  }
  if (ordered) {
    FKD_TRY(t->q_border.ensure((size_t)nb * 4));
    if (!sorted) {
      fkd::block_work_kernel<<<nb, fkd::kThreads, 0, s>>>(t->q_wslot.as<uint32_t>(), nq, t->q_bwork.as<uint32_t>());
      FKD_TRY(launch_check("block_work"));
    }
    fkd::block_order_kernel<<<1, 1024, 0, s>>>(t->q_bwork.as<uint32_t>(), nb, t->q_border.as<uint32_t>());
    FKD_TRY(launch_check("block_order"));
    a.block_order = t->q_border.as<uint32_t>();
  }
  if (t->trace_start) FKD_HIP(hipEventRecord(t->trace_start, s));
  FKD_TRY((launch_search<DIM, fkd::kFill>(t, a, s, "search_fill")));
  if (t->trace_stop) FKD_HIP(hipEventRecord(t->trace_stop, s));
  return FKD_OK;
}

// fkd_search: count pass, scan and fill pass back to back on the stream (the
// fill pass reads the offsets on the device, no host round trip); ids beyond
// `cap` are dropped and reported by fkd_search_status.
template <int DIM>
int search_single_impl(fkd_tree_s* t, uint32_t nq, const float* mins, const float* maxs, const uint32_t* qev,
                       uint64_t* qoff, uint32_t* out, uint64_t cap, hipStream_t s) {
  FKD_TRY(search_count_impl<DIM>(t, nq, mins, maxs, qev, qoff, s));
  FKD_TRY(search_fill_impl<DIM>(t, nq, mins, maxs, qev, qoff, out, cap, s));
  t->last_nq = nq;
  t->last_capacity = cap;
  t->last_offsets = qoff;
  return FKD_OK;
}

__global__ void check_input_kernel(const float* X, const uint32_t* ids, const uint32_t* off,
                                   uint32_t n_events, uint32_t P, int dim, uint32_t* err) {
  bool nan = false, unordered = false;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < P; i += gridDim.x * blockDim.x) {
    for (int m = 0; m < dim; ++m) {
      const float x = X[(uint64_t)m * P + i];
      nan |= (x != x);
    }
    if (i + 1 < P && ids[i] >= ids[i + 1]) {
      // Only a problem when i and i+1 are in the same event.
      uint32_t lo = 0, hi = n_events;  // find e with off[e] <= i < off[e+1]
      while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (off[mid] <= i) lo = mid; else hi = mid;
      }
      if (i + 1 < off[lo + 1]) unordered = true;
    }
  }
  if (nan) atomicOr(err, 1u);
  if (unordered) atomicOr(err, 2u);
}

// Sets *err when two neighbours of the (event, id)-sorted ids are equal
// inside one event.
__global__ void duplicate_ids_kernel(const uint32_t* __restrict__ sorted, const uint32_t* __restrict__ off,
                                     uint32_t n_events, uint32_t P, uint32_t* err) {
  bool dup = false;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i + 1 < P; i += gridDim.x * blockDim.x) {
    if (sorted[i] != sorted[i + 1]) continue;
    uint32_t lo = 0, hi = n_events;  // event e with off[e] <= i < off[e+1]
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) / 2;
      if (off[mid] <= i) lo = mid; else hi = mid;
    }
    dup |= i + 1 < off[lo + 1];
  }
  if (dup) atomicOr(err, 1u);
}

int validate_offsets(uint32_t n_events, const uint64_t* event_offsets, uint64_t* P) {
  if (n_events == 0) return fail(FKD_ERR_INVALID, "n_events must be >= 1");
  if (n_events > kMaxEvents) return fail(FKD_ERR_INVALID, "more than 2^30 events in one handle");
  if (!event_offsets) return fail(FKD_ERR_INVALID, "event_offsets is NULL");
  if (event_offsets[0] != 0) return fail(FKD_ERR_INVALID, "event_offsets[0] must be 0");
  for (uint32_t e = 0; e < n_events; ++e) {
    if (event_offsets[e + 1] < event_offsets[e]) return fail(FKD_ERR_INVALID, "event_offsets not monotone");
    if (event_offsets[e + 1] - event_offsets[e] > kMaxPoints)
      return fail(FKD_ERR_INVALID, "more than 2^30 points in one event");
  }
  *P = event_offsets[n_events];
  if (*P > kMaxPoints) return fail(FKD_ERR_INVALID, "more than 2^30 points in one handle");
  return FKD_OK;
}

// Allocates everything, given the input already in t->X / t->ids; validates.
// Scratch of the build (released by fkd_trim, re-created on the next build).
// clear = false: the caller's first kernel clears the arena (it must run
// before any kernel that relies on zeroed words).
int zero_arena(fkd_tree_s* t, uint64_t words, hipStream_t s, ZeroArena& za, bool clear) {
  FKD_TRY(t->zarena.ensure(std::max<uint64_t>(words, 1) * 8));
  za.base = t->zarena.as<uint64_t>();
  za.used = 0;
  za.cap = words;
  return clear ? fill_u64(za.base, 0, words, s) : FKD_OK;
}

int ensure_build_scratch(fkd_tree_s* t) {
  const size_t S = std::max<uint32_t>(t->P, 1);
  const int dim = t->dim;
  FKD_TRY(t->A.ensure(S * dim * 2 * 4));
  FKD_TRY(t->K.ensure(S * dim * 2 * 4));
  FKD_TRY(t->R.ensure(S * 4 * 4));
  FKD_TRY(t->planes.ensure(fkd::plane_stride(S) * dim * dim * 2 * 4));
  FKD_TRY(t->F.ensure(S * 4));
  FKD_TRY(t->Xr.ensure(S * 16));
  const uint64_t nc = (uint64_t)dim * 256u * std::max<uint32_t>(t->radix_tiles.T, 1);
  FKD_TRY(t->counts.ensure(nc * 4));
  FKD_TRY(t->offsets.ensure((nc + 1) * 4));
  return FKD_OK;
}

// Ids must be unique within an event: the reference comparator
// (FKDTree_cpu.h:246-252) is a strict order only then, and the tree is a
// function of the points only then.  Increasing ids are unique; otherwise the
// ids of every event are sorted (the build's own id-sort passes) and
// neighbours compared.  Synchronous.
int check_unique_ids(fkd_tree_s* t, hipStream_t s) {
  const uint32_t P = t->P;
  if (P < 2) return FKD_OK;
  const uint64_t S = P;
  uint32_t* K = t->K.as<uint32_t>();
  uint32_t* A = t->A.as<uint32_t>();
  const fkd::TileMap tm = t->radix_tiles.map(t->off.as<uint32_t>());
  ZeroArena za;
  FKD_TRY(zero_arena(t, 4 * radix_pass_words(1, tm.T), s, za));
  fkd::id_keys_kernel<<<grid_for(P), 256, 0, s>>>(t->ids.as<uint32_t>(), P, K, A);
  FKD_TRY(launch_check("id_keys"));
  int c = 0;
  for (int pass = 0; pass < 4; ++pass) {
    FKD_TRY(run_radix_pass<1>(t, za, 1, K + c * S, K + (1 - c) * S, A + c * S, A + (1 - c) * S, S, nullptr, 8u * pass,
                              true, false, s));
    c = 1 - c;
  }
  uint32_t* flag = t->err.as<uint32_t>() + 1;
  FKD_TRY(fill_u32(flag, 0, 1, s));
  duplicate_ids_kernel<<<grid_for(P), 256, 0, s>>>(K + c * S, t->off.as<uint32_t>(), t->n_events, P, flag);
  FKD_TRY(launch_check("duplicate_ids"));
  uint32_t dup = 0;
  FKD_HIP(hipMemcpyAsync(&dup, flag, 4, hipMemcpyDeviceToHost, s));
  FKD_HIP(hipStreamSynchronize(s));
  if (dup) return fail(FKD_ERR_INVALID, "duplicate point id within an event (ids must be unique)");
  return FKD_OK;
}

// Tilings that depend on the build's launch parameters (finisher capacity,
// digit-pass tile): the global depths, the finisher's subtrees and the radix
// tiles.  Laid out at creation and again by a build after those change.
int layout(fkd_tree_s* t, hipStream_t s) {
  const uint32_t E = t->n_events, fl = t->tune.fin_log;
  t->Dmax = 0;
  t->global_levels = 0;
  t->fin_cap = 1;
  std::vector<uint32_t> fin_count(E, 0);
  for (uint32_t e = 0; e < E; ++e) {
    const uint64_t n = t->h_off[e + 1] - t->h_off[e];
    if (!n) continue;
    const uint32_t D = fkd::floor_log2((uint32_t)n);
    t->Dmax = std::max(t->Dmax, D);
    t->global_levels = std::max(t->global_levels, fkd::fin_depth(D, fl));
    fin_count[e] = 1u << fkd::fin_depth(D, fl);  // subtrees the finisher builds
    t->fin_cap = std::max(t->fin_cap, fkd::fin_subtree_max((uint32_t)n, fl));
  }
  FKD_TRY(t->radix_tiles.init(t->h_off, fkd::kThreads * t->tune.build_radix_items, s));
  FKD_TRY(t->fin_tiles.init_counts(fin_count, s));
  t->lay_fin_log = fl;
  t->lay_radix_items = t->tune.build_radix_items;
  return FKD_OK;
}

int finish_create(fkd_tree_s* t, hipStream_t s) {
  const uint32_t P = t->P, E = t->n_events;
  const int dim = t->dim;
  std::vector<uint32_t> off32((size_t)E + 1);
  for (size_t e = 0; e <= E; ++e) off32[e] = (uint32_t)t->h_off[e];
  FKD_TRY(t->off.ensure(((size_t)E + 1) * 4));
  FKD_HIP(hipMemcpyAsync(t->off.p, off32.data(), ((size_t)E + 1) * 4, hipMemcpyHostToDevice, s));
  FKD_TRY(t->level_tiles.init(t->h_off, fkd::kLevelTile, s));
  FKD_TRY(layout(t, s));
  const size_t S = std::max<uint32_t>(P, 1);
  const int rf = dim <= 3 ? 4 : 8;
  FKD_TRY(t->Tc.ensure(S * dim * 4));
  FKD_TRY(t->Tid.ensure(S * 4));
  FKD_TRY(t->recs.ensure(S * rf * 4));
  FKD_TRY(ensure_build_scratch(t));
  FKD_TRY(t->err.ensure(64));
  FKD_TRY(t->bbox.ensure(64));
  FKD_TRY(fill_u32(t->err.as<uint32_t>(), 0, 1, s));
  if (P) {
    check_input_kernel<<<grid_for(P), 256, 0, s>>>(t->X.as<float>(), t->ids.as<uint32_t>(), t->off.as<uint32_t>(),
                                                    E, P, dim, t->err.as<uint32_t>());
    FKD_TRY(launch_check("check_input"));
  }
  uint32_t err = 0;
  FKD_HIP(hipMemcpyAsync(&err, t->err.p, 4, hipMemcpyDeviceToHost, s));
  FKD_HIP(hipStreamSynchronize(s));
  if (err & 1u) return fail(FKD_ERR_NAN, "NaN coordinate in the input points");
  t->ids_in_order = (err & 2u) == 0;
  if (!t->ids_in_order) FKD_TRY(check_unique_ids(t, s));
  return FKD_OK;
}

template <int DIM>
int aos_upload(fkd_tree_s* t, const void* host_points, hipStream_t s) {
  const uint32_t P = t->P;
  if (P == 0) return FKD_OK;
  DevBuf staging;
  FKD_TRY(staging.ensure((size_t)P * (DIM + 1) * 4));
  FKD_HIP(hipMemcpyAsync(staging.p, host_points, (size_t)P * (DIM + 1) * 4, hipMemcpyHostToDevice, s));
  fkd::aos_to_soa_kernel<DIM><<<grid_for(P), 256, 0, s>>>(staging.as<uint32_t>(), t->X.as<float>(),
                                                           t->ids.as<uint32_t>(), P);
  FKD_TRY(launch_check("aos_to_soa"));
  FKD_HIP(hipStreamSynchronize(s));
  return FKD_OK;
}

template <class F>
int dispatch_dim(int dim, F&& f) {
  switch (dim) {
    case 1: return f(std::integral_constant<int, 1>());
    case 2: return f(std::integral_constant<int, 2>());
    case 3: return f(std::integral_constant<int, 3>());
    case 4: return f(std::integral_constant<int, 4>());
    default: return fail(FKD_ERR_INVALID, "dim must be 1..4");
  }
}

int create_common(int dim, uint32_t n_events, const uint64_t* event_offsets, fkd_tree_s** out) {
  if (!out) return fail(FKD_ERR_INVALID, "out is NULL");
  *out = nullptr;
  if (dim < 1 || dim > 4) return fail(FKD_ERR_INVALID, "dim must be 1..4");
  uint64_t P = 0;
  FKD_TRY(validate_offsets(n_events, event_offsets, &P));
  fkd_tree_s* t = new fkd_tree_s();
  t->dim = dim;
  t->n_events = n_events;
  t->P = (uint32_t)P;
  // Several trees: starting the heaviest blocks first scatters the writes of
  // an event's boxes over the whole output (measured slower for batches).
  if (n_events > 1) t->tune.fill_order = 0;
  t->h_off.assign(event_offsets, event_offsets + (size_t)n_events + 1);
  const size_t S = std::max<uint64_t>(P, 1);
  int r = t->X.ensure(S * dim * 4);
  if (r == FKD_OK) r = t->ids.ensure(S * 4);
  if (r != FKD_OK) {
    delete t;
    return r;
  }
  *out = t;
  return FKD_OK;
}

// Tree arrays and search records of a handle whose points already are the
// heap-ordered tree (loaded from the wire format), without a build.
template <int DIM>
int adopt_tree_impl(fkd_tree_s* t, hipStream_t s) {
  const uint32_t P = t->P;
  if (P) {
    const uint32_t grid = grid_for(P);
    if (DIM <= 3) {
      fkd::records_kernel<DIM><<<grid, 256, 0, s>>>(t->X.as<float>(), t->ids.as<uint32_t>(), P, t->Xr.as<float4>());
      FKD_TRY(launch_check("records"));
    }
    fkd::LevelArgs<DIM> a{};
    a.tm = level_map(t);
    a.S0 = nullptr;
    a.F = nullptr;
    a.X = t->X.as<float>();
    a.ids = t->ids.as<uint32_t>();
    a.P = P;
    a.Tc = t->Tc.as<float>();
    a.Tid = t->Tid.as<uint32_t>();
    a.recs = t->recs.as<float>();
    fkd::output_kernel<DIM><<<a.tm.T, fkd::kThreads, 0, s>>>(a, t->Xr.as<float4>());
    FKD_TRY(launch_check("output"));
  }
  FKD_TRY(fill_u32(t->bbox.as<uint32_t>(), 0xffffffffu, DIM, s));
  FKD_TRY(fill_u32(t->bbox.as<uint32_t>() + DIM, 0u, DIM, s));
  if (P) {
    fkd::bbox_points_kernel<DIM><<<grid_for(P), 256, 0, s>>>(t->X.as<float>(), P, t->bbox.as<uint32_t>());
    FKD_TRY(launch_check("bbox_points"));
  }
  t->built = true;
  t->q_plan_nq = 0;
  return FKD_OK;
}

// Scratch of fkd_search_noalloc: a file-scope device array, one copy per GPU,
// held until the process ends.  It holds the zeroed look-back descriptors
// of the offsets scan, then the per-box hit counts.  Host code never uses
// the symbol itself as an address: noalloc_scratch() looks up each device's
// copy once (hipGetSymbolAddress) and every kernel argument, fill and copy
// uses that pointer.
constexpr uint32_t kNoAllocMaxBoxes = 1u << 22;
// descriptor words, rounded up to 256 bytes: the scan reads the counts that
// follow as 16-byte vectors
constexpr uint64_t kNoAllocDescWords =
    ((kNoAllocMaxBoxes + 1 + fkd::kScanTile - 1) / fkd::kScanTile + fkd::kCtrWords + 31) & ~uint64_t(31);
static_assert(kNoAllocDescWords * 8 % 256 == 0, "counts start 256-byte aligned");
constexpr uint64_t kNoAllocBytes = kNoAllocDescWords * 8 + (uint64_t)kNoAllocMaxBoxes * 4;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The current device's copy of `scratch` (looked up on the first call on
// each device, with that device current).
int noalloc_scratch(void** out) {
  int dev = 0;
  FKD_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(FKD_ERR_INVALID, "device index beyond the scratch table");
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    FKD_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(FKD_ERR_HIP, "hipGetSymbolAddress returned no address for the search scratch");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return FKD_OK;
}

// fkd_search without the spatial box plan: count pass in the caller's box
// order, scan, fill pass.  Scratch comes from `scr` (kNoAllocBytes); nothing
// is allocated.
template <int DIM>
int search_noalloc_impl(fkd_tree_s* t, uint32_t nq, const float* mins, const float* maxs, const uint32_t* qev,
                        uint64_t* qoff, uint32_t* out, uint64_t cap, void* scr, hipStream_t s) {
  const uint64_t st = fkd::scan_tiles(nq);
  uint64_t* desc = static_cast<uint64_t*>(scr);
  uint32_t* counts = reinterpret_cast<uint32_t*>(desc + kNoAllocDescWords);
  FKD_TRY(fill_u64(desc, 0, st + fkd::kCtrWords, s));  // a kernel: no memset node under capture
  fkd::SearchArgs a = search_args<DIM>(t, nq, qev);
  a.mins = mins;
  a.maxs = maxs;
  if (nq > 0) {
    a.counts = counts;
    FKD_TRY((launch_search<DIM, fkd::kCount>(t, a, s, "search_count_noalloc")));
  }
  fkd::scan_exclusive_kernel<uint64_t><<<(uint32_t)st, fkd::kThreads, 0, s>>>(
      counts, qoff, nq, desc, reinterpret_cast<uint32_t*>(desc + st));
  FKD_TRY(launch_check("search_scan_noalloc"));
  if (nq > 0) {
    a.counts = nullptr;
    a.qoff = qoff;
    a.out = out;
    a.cap = cap;
    a.out_aligned = (reinterpret_cast<uintptr_t>(out) & 63u) == 0;
    FKD_TRY((launch_search<DIM, fkd::kFill>(t, a, s, "search_fill_noalloc")));
  }
  t->last_nq = nq;
  t->last_capacity = cap;
  t->last_offsets = qoff;
  return FKD_OK;
}

constexpr char kWireMagic[8] = {'F', 'K', 'D', 'T', 'R', 'E', 'E', '1'};
constexpr uint64_t kWireHeader = 24;  // magic, uint32 dim, uint32 n_events, uint64 n_points

uint64_t wire_bytes(int dim, uint32_t n_events, uint64_t P) {
  return kWireHeader + 8ull * ((uint64_t)n_events + 1) + 4ull * P * (uint64_t)(dim + 1);
}

}  // namespace

extern "C" {

int fkd_serialize(fkd_tree_t t, void* buf, uint64_t capacity, uint64_t* size, fkd_stream_t stream) {
  if (!t || !size) return fail(FKD_ERR_INVALID, "tree / size is NULL");
  if (!t->built) return fail(FKD_ERR_STATE, "tree not built");
  const uint64_t bytes = wire_bytes(t->dim, t->n_events, t->P);
  *size = bytes;
  if (!buf) return FKD_OK;
  if (capacity < bytes) return fail(FKD_ERR_CAPACITY, "buffer holds " + std::to_string(capacity) + " of " +
                                                          std::to_string(bytes) + " bytes");
  char* b = static_cast<char*>(buf);
  const uint32_t dim = (uint32_t)t->dim, E = t->n_events;
  const uint64_t P = t->P;
  std::memcpy(b, kWireMagic, 8);
  std::memcpy(b + 8, &dim, 4);
  std::memcpy(b + 12, &E, 4);
  std::memcpy(b + 16, &P, 8);
  std::memcpy(b + kWireHeader, t->h_off.data(), 8ull * ((uint64_t)E + 1));
  char* data = b + kWireHeader + 8ull * ((uint64_t)E + 1);
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (P) {
    FKD_HIP(hipMemcpyAsync(data, t->Tc.p, P * dim * 4, hipMemcpyDeviceToHost, s));
    FKD_HIP(hipMemcpyAsync(data + P * dim * 4, t->Tid.p, P * 4, hipMemcpyDeviceToHost, s));
  }
  FKD_HIP(hipStreamSynchronize(s));
  return FKD_OK;
}

int fkd_deserialize(const void* buf, uint64_t size, int verify, fkd_stream_t stream, fkd_tree_t* out) {
  StreamScope scope(static_cast<hipStream_t>(stream));
  if (!buf || !out) return fail(FKD_ERR_INVALID, "buf / out is NULL");
  *out = nullptr;
  const char* b = static_cast<const char*>(buf);
  if (size < kWireHeader || std::memcmp(b, kWireMagic, 8) != 0)
    return fail(FKD_ERR_INVALID, "not an FKDTREE1 buffer");
  uint32_t dim = 0, E = 0;
  uint64_t P = 0;
  std::memcpy(&dim, b + 8, 4);
  std::memcpy(&E, b + 12, 4);
  std::memcpy(&P, b + 16, 8);
  if (dim < 1 || dim > 4 || E < 1 || E > kMaxEvents || P > kMaxPoints) return fail(FKD_ERR_INVALID, "bad FKDTREE1 header");
  if (size != wire_bytes((int)dim, E, P))
    return fail(FKD_ERR_INVALID, "FKDTREE1 size " + std::to_string(size) + " != " +
                                     std::to_string(wire_bytes((int)dim, E, P)));
  std::vector<uint64_t> off((size_t)E + 1);
  std::memcpy(off.data(), b + kWireHeader, 8ull * ((uint64_t)E + 1));
  if (off[E] != P) return fail(FKD_ERR_INVALID, "FKDTREE1 offsets do not end at n_points");
  hipStream_t s = static_cast<hipStream_t>(stream);
  fkd_tree_s* t = nullptr;
  FKD_TRY(create_common((int)dim, E, off.data(), &t));
  const char* data = b + kWireHeader + 8ull * ((uint64_t)E + 1);
  int r = FKD_OK;
  if (P) {
    hipError_t e = hipMemcpyAsync(t->X.p, data, P * dim * 4, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(t->ids.p, data + P * dim * 4, P * 4, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) r = fail(FKD_ERR_HIP, std::string("hipMemcpyAsync: ") + hipGetErrorString(e));
  }
  if (r == FKD_OK) r = finish_create(t, s);
  if (r == FKD_OK && verify) {
    // The tree is a unique function of its points: a valid buffer rebuilds to itself.
    r = dispatch_dim((int)dim, [&](auto D) { return build_impl<decltype(D)::value>(t, s); });
    if (r == FKD_OK && P) r = fill_u32(t->err.as<uint32_t>(), 0, 1, s);
    if (r == FKD_OK && P) {
      fkd::compare_kernel<<<grid_for(P * dim), 256, 0, s>>>(t->Tc.as<uint32_t>(), t->X.as<uint32_t>(), P * dim,
                                                             t->err.as<uint32_t>());
      fkd::compare_kernel<<<grid_for(P), 256, 0, s>>>(t->Tid.as<uint32_t>(), t->ids.as<uint32_t>(), P,
                                                       t->err.as<uint32_t>());
      r = launch_check("compare");
    }
    uint32_t err = 0;
    if (r == FKD_OK && P) {
      hipError_t e = hipMemcpyAsync(&err, t->err.p, 4, hipMemcpyDeviceToHost, s);
      if (e == hipSuccess) e = hipStreamSynchronize(s);
      if (e != hipSuccess) r = fail(FKD_ERR_HIP, std::string("verify: ") + hipGetErrorString(e));
    }
    if (r == FKD_OK && err) r = fail(FKD_ERR_INVALID, "FKDTREE1 arrays are not the k-d tree of their points");
  } else if (r == FKD_OK) {
    r = dispatch_dim((int)dim, [&](auto D) { return adopt_tree_impl<decltype(D)::value>(t, s); });
    if (r == FKD_OK) {
      hipError_t e = hipStreamSynchronize(s);
      if (e != hipSuccess) r = fail(FKD_ERR_HIP, std::string("load: ") + hipGetErrorString(e));
    }
  }
  if (r != FKD_OK) {
    delete t;
    return r;
  }
  *out = t;
  return FKD_OK;
}

const char* fkd_last_error(void) { return g_err.c_str(); }

int fkd_create_events(int dim, uint32_t n_events, const uint64_t* event_offsets, const void* host_points,
                      fkd_stream_t stream, fkd_tree_t* out) {
  StreamScope scope(static_cast<hipStream_t>(stream));
  hipStream_t s = static_cast<hipStream_t>(stream);
  fkd_tree_s* t = nullptr;
  FKD_TRY(create_common(dim, n_events, event_offsets, &t));
  if (t->P && !host_points) {
    delete t;
    return fail(FKD_ERR_INVALID, "host_points is NULL");
  }
  int r = dispatch_dim(dim, [&](auto D) { return aos_upload<decltype(D)::value>(t, host_points, s); });
  if (r == FKD_OK) r = finish_create(t, s);
  if (r != FKD_OK) {
    delete t;
    return r;
  }
  *out = t;
  return FKD_OK;
}

int fkd_create(int dim, uint32_t n, const void* host_points, fkd_stream_t stream, fkd_tree_t* out) {
  const uint64_t off[2] = {0, n};
  return fkd_create_events(dim, 1, off, host_points, stream, out);
}

int fkd_create_device(int dim, uint32_t n_events, const uint64_t* event_offsets, const float* dev_coords,
                      const uint32_t* dev_ids, fkd_stream_t stream, fkd_tree_t* out) {
  StreamScope scope(static_cast<hipStream_t>(stream));
  hipStream_t s = static_cast<hipStream_t>(stream);
  fkd_tree_s* t = nullptr;
  FKD_TRY(create_common(dim, n_events, event_offsets, &t));
  int r = FKD_OK;
  if (t->P) {
    if (!dev_coords || !dev_ids) {
      delete t;
      return fail(FKD_ERR_INVALID, "dev_coords / dev_ids is NULL");
    }
    hipError_t e = hipMemcpyAsync(t->X.p, dev_coords, (size_t)t->P * dim * 4, hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess)
      e = hipMemcpyAsync(t->ids.p, dev_ids, (size_t)t->P * 4, hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) r = fail(FKD_ERR_HIP, std::string("hipMemcpyAsync: ") + hipGetErrorString(e));
  }
  if (r == FKD_OK) r = finish_create(t, s);
  if (r != FKD_OK) {
    delete t;
    return r;
  }
  *out = t;
  return FKD_OK;
}

void fkd_destroy(fkd_tree_t tree) { delete tree; }

int fkd_destroy_async(fkd_tree_t tree, fkd_stream_t stream) {
  StreamScope scope(static_cast<hipStream_t>(stream));
  delete tree;
  return launch_check("destroy");
}

int fkd_build(fkd_tree_t t, fkd_stream_t stream) {
  StreamScope scope(static_cast<hipStream_t>(stream));
  if (!t) return fail(FKD_ERR_INVALID, "tree is NULL");
  hipStream_t s = static_cast<hipStream_t>(stream);
  return dispatch_dim(t->dim, [&](auto D) { return build_impl<decltype(D)::value>(t, s); });
}

int fkd_get_tree(fkd_tree_t t, float* coords_out, uint32_t* ids_out, fkd_stream_t stream) {
  if (!t) return fail(FKD_ERR_INVALID, "tree is NULL");
  if (!t->built) return fail(FKD_ERR_STATE, "tree not built");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (t->P) {
    if (coords_out)
      FKD_HIP(hipMemcpyAsync(coords_out, t->Tc.p, (size_t)t->P * t->dim * 4, hipMemcpyDeviceToHost, s));
    if (ids_out) FKD_HIP(hipMemcpyAsync(ids_out, t->Tid.p, (size_t)t->P * 4, hipMemcpyDeviceToHost, s));
  }
  FKD_HIP(hipStreamSynchronize(s));
  return FKD_OK;
}

int fkd_tree_device(fkd_tree_t t, const float** dev_coords, const uint32_t** dev_ids) {
  if (!t) return fail(FKD_ERR_INVALID, "tree is NULL");
  if (dev_coords) *dev_coords = t->Tc.as<float>();
  if (dev_ids) *dev_ids = t->Tid.as<uint32_t>();
  return FKD_OK;
}

int fkd_tree_info(fkd_tree_t t, int* dim, uint32_t* n_events, uint64_t* n_points) {
  if (!t) return fail(FKD_ERR_INVALID, "tree is NULL");
  if (dim) *dim = t->dim;
  if (n_events) *n_events = t->n_events;
  if (n_points) *n_points = t->P;
  return FKD_OK;
}

int fkd_search_count(fkd_tree_t t, uint32_t nq, const float* mins, const float* maxs, const uint32_t* qev,
                     uint64_t* dev_offsets, fkd_stream_t stream) {
  StreamScope scope(static_cast<hipStream_t>(stream));
  if (!t) return fail(FKD_ERR_INVALID, "tree is NULL");
  if (!t->built) return fail(FKD_ERR_STATE, "tree not built");
  if (!dev_offsets || (nq && (!mins || !maxs))) return fail(FKD_ERR_INVALID, "NULL search buffer");
  if (!qev && t->n_events > 1) return fail(FKD_ERR_INVALID, "query events required for a batch of trees");
  hipStream_t s = static_cast<hipStream_t>(stream);
  return dispatch_dim(t->dim, [&](auto D) {
    return search_count_impl<decltype(D)::value>(t, nq, mins, maxs, qev, dev_offsets, s);
  });
}

int fkd_search_fill(fkd_tree_t t, uint32_t nq, const float* mins, const float* maxs, const uint32_t* qev,
                    const uint64_t* dev_offsets, uint32_t* out, uint64_t capacity, fkd_stream_t stream) {
  StreamScope scope(static_cast<hipStream_t>(stream));
  if (!t) return fail(FKD_ERR_INVALID, "tree is NULL");
  if (!t->built) return fail(FKD_ERR_STATE, "tree not built");
  if (!dev_offsets || (nq && (!mins || !maxs))) return fail(FKD_ERR_INVALID, "NULL search buffer");
  if (!qev && t->n_events > 1) return fail(FKD_ERR_INVALID, "query events required for a batch of trees");
  if (capacity && !out) return fail(FKD_ERR_INVALID, "out is NULL");
  hipStream_t s = static_cast<hipStream_t>(stream);
  return dispatch_dim(t->dim, [&](auto D) {
    return search_fill_impl<decltype(D)::value>(t, nq, mins, maxs, qev, dev_offsets, out, capacity, s);
  });
}

int fkd_search_stats(fkd_tree_t t, uint32_t nq, const float* mins, const float* maxs, const uint32_t* qev,
                     uint32_t* dev_counts, uint32_t* dev_visited, fkd_stream_t stream) {
  StreamScope scope(static_cast<hipStream_t>(stream));
  if (!t) return fail(FKD_ERR_INVALID, "tree is NULL");
  if (!t->built) return fail(FKD_ERR_STATE, "tree not built");
  if (nq && (!mins || !maxs || !dev_counts || !dev_visited)) return fail(FKD_ERR_INVALID, "NULL search buffer");
  if (!qev && t->n_events > 1) return fail(FKD_ERR_INVALID, "query events required for a batch of trees");
  hipStream_t s = static_cast<hipStream_t>(stream);
  return dispatch_dim(t->dim, [&](auto D) {
    return search_stats_impl<decltype(D)::value>(t, nq, mins, maxs, qev, dev_counts, dev_visited, s);
  });
}

int fkd_search(fkd_tree_t t, uint32_t nq, const float* mins, const float* maxs, const uint32_t* qev,
               uint64_t* dev_offsets, uint32_t* dev_ids_out, uint64_t capacity, fkd_stream_t stream) {
  StreamScope scope(static_cast<hipStream_t>(stream));
  if (!t) return fail(FKD_ERR_INVALID, "tree is NULL");
  if (!t->built) return fail(FKD_ERR_STATE, "tree not built");
  if (!dev_offsets || (nq && (!mins || !maxs))) return fail(FKD_ERR_INVALID, "NULL search buffer");
  if (!qev && t->n_events > 1) return fail(FKD_ERR_INVALID, "query events required for a batch of trees");
  if (capacity && !dev_ids_out) return fail(FKD_ERR_INVALID, "dev_ids_out is NULL");
  hipStream_t s = static_cast<hipStream_t>(stream);
  return dispatch_dim(t->dim, [&](auto D) {
    return search_single_impl<decltype(D)::value>(t, nq, mins, maxs, qev, dev_offsets, dev_ids_out, capacity, s);
  });
}

int fkd_boxes_around(fkd_tree_t t, float half_width, float* dev_mins, float* dev_maxs, uint32_t* dev_query_events,
                     fkd_stream_t stream) {
  if (!t) return fail(FKD_ERR_INVALID, "tree is NULL");
  if (t->P && (!dev_mins || !dev_maxs)) return fail(FKD_ERR_INVALID, "NULL box buffer");
  if (!(half_width >= 0.f)) return fail(FKD_ERR_INVALID, "half_width must be >= 0");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (!t->P) return FKD_OK;
  const fkd::TileMap tm = level_map(t);
  return dispatch_dim(t->dim, [&](auto D) {
    fkd::boxes_around_kernel<decltype(D)::value><<<tm.T, fkd::kThreads, 0, s>>>(
        t->X.as<float>(), t->P, half_width, tm, dev_mins, dev_maxs, dev_query_events);
    return launch_check("boxes_around");
  });
}

int fkd_trace_events(fkd_tree_t t, fkd_event_t start, fkd_event_t stop) {
  if (!t) return fail(FKD_ERR_INVALID, "tree is NULL");
  t->trace_start = static_cast<hipEvent_t>(start);
  t->trace_stop = static_cast<hipEvent_t>(stop);
  return FKD_OK;
}

int fkd_search_noalloc(fkd_tree_t t, uint32_t nq, const float* mins, const float* maxs, const uint32_t* qev,
                       uint64_t* dev_offsets, uint32_t* dev_ids_out, uint64_t capacity, fkd_stream_t stream) {
  if (!t) return fail(FKD_ERR_INVALID, "tree is NULL");
  if (!t->built) return fail(FKD_ERR_STATE, "tree not built");
  if (!dev_offsets || (nq && (!mins || !maxs))) return fail(FKD_ERR_INVALID, "NULL search buffer");
  if (!qev && t->n_events > 1) return fail(FKD_ERR_INVALID, "query events required for a batch of trees");
  if (capacity && !dev_ids_out) return fail(FKD_ERR_INVALID, "dev_ids_out is NULL");
  if (nq > kNoAllocMaxBoxes)
    return fail(FKD_ERR_INVALID, "fkd_search_noalloc takes at most " + std::to_string(kNoAllocMaxBoxes) +
                                     " boxes per call (got " + std::to_string(nq) + ")");
  void* scr = nullptr;
  FKD_TRY(noalloc_scratch(&scr));
  hipStream_t s = static_cast<hipStream_t>(stream);
  return dispatch_dim(t->dim, [&](auto D) {
    return search_noalloc_impl<decltype(D)::value>(t, nq, mins, maxs, qev, dev_offsets, dev_ids_out, capacity,
                                                   scr, s);
  });
}

int fkd_search_small(fkd_tree_t t, uint32_t nb, const float* host_boxes, uint32_t slot, uint32_t* host_counts,
                     uint32_t* host_ids, fkd_stream_t stream) {
  if (!t) return fail(FKD_ERR_INVALID, "tree is NULL");
  if (!t->built) return fail(FKD_ERR_STATE, "tree not built");
  if (t->n_events != 1) return fail(FKD_ERR_INVALID, "fkd_search_small searches a single tree (use fkd_search)");
  if (nb > FKD_SMALL_MAX_BOXES)
    return fail(FKD_ERR_INVALID, "fkd_search_small takes at most " + std::to_string(FKD_SMALL_MAX_BOXES) + " boxes");
  if (slot == 0 || slot > FKD_SMALL_MAX_SLOT) return fail(FKD_ERR_INVALID, "slot must be 1..2^20 ids");
  if ((uint64_t)nb * slot > (1ull << 28)) return fail(FKD_ERR_INVALID, "nb * slot above 2^28 ids");
  if (nb && (!host_boxes || !host_counts || !host_ids)) return fail(FKD_ERR_INVALID, "NULL box or result buffer");
  if (nb == 0) return FKD_OK;
  const int dim = t->dim;
  const size_t in_bytes = (size_t)nb * 2 * dim * 4;
  FKD_TRY(t->small_in.ensure(in_bytes));
  FKD_TRY(t->small_out.ensure(((size_t)nb + (size_t)nb * slot) * 4));
  std::memcpy(t->small_in.h, host_boxes, in_bytes);
  uint32_t* d_counts = static_cast<uint32_t*>(t->small_out.d);
  uint32_t* d_ids = d_counts + nb;
  hipStream_t s = static_cast<hipStream_t>(stream);
  FKD_TRY(dispatch_dim(dim, [&](auto D) {
    constexpr int DIM = decltype(D)::value;
    const uint64_t bytes = (uint64_t)t->P * fkd::NodeRec<DIM>::kFloats * 4;
    const float* boxes = static_cast<const float*>(t->small_in.d);
    if (bytes < (1ull << 32))
      fkd::box_wave_kernel<DIM, true><<<nb, fkd::kWave, 0, s>>>(t->recs.as<float>(), (uint32_t)bytes, t->P, boxes,
                                                                nb, slot, d_counts, d_ids);
    else
      fkd::box_wave_kernel<DIM, false><<<nb, fkd::kWave, 0, s>>>(t->recs.as<float>(), 0u, t->P, boxes, nb, slot,
                                                                 d_counts, d_ids);
    return launch_check("box_wave");
  }));
  FKD_HIP(hipStreamSynchronize(s));
  const uint32_t* h_counts = static_cast<const uint32_t*>(t->small_out.h);
  const uint32_t* h_ids = h_counts + nb;
  std::memcpy(host_counts, h_counts, (size_t)nb * 4);
  for (uint32_t b = 0; b < nb; ++b) {
    const uint32_t c = h_counts[b];
    if (c != fkd::kSmallOverflow && c) std::memcpy(host_ids + (size_t)b * slot, h_ids + (size_t)b * slot,
                                                   (size_t)std::min(c, slot) * 4);
  }
  return FKD_OK;
}

namespace {

// One launch of box_server_kernel as generation gen + 1 (caller holds
// launch_mu and has seen no live launch).
int server_launch(fkd_box_server_s* v) {
  fkd_tree_s* t = v->tree;
  const uint32_t g = v->gen.load(std::memory_order_relaxed) + 1u;
  fkd::BoxCtl* ctl = v->ctl.as<fkd::BoxCtl>();
  uint32_t* done = reinterpret_cast<uint32_t*>(ctl + 1);
  FKD_HIP(hipMemsetAsync(&ctl->exited, 0, sizeof(uint32_t), v->stream));
  FKD_TRY(dispatch_dim(v->dim, [&](auto D) {
    constexpr int DIM = decltype(D)::value;
    const uint64_t bytes = (uint64_t)t->P * fkd::NodeRec<DIM>::kFloats * 4;
    const fkd::BoxReq* rq = static_cast<const fkd::BoxReq*>(v->req.d);
    fkd::BoxRep* rp = static_cast<fkd::BoxRep*>(v->rep.d);
    uint32_t* ids = static_cast<uint32_t*>(v->ids.d);
    uint32_t* ew = static_cast<uint32_t*>(v->word.d);
    if (bytes < (1ull << 32))
      fkd::box_server_kernel<DIM, true><<<v->nslots, fkd::kWave, 0, v->stream>>>(
          t->recs.as<float>(), (uint32_t)bytes, t->P, rq, rp, ids, v->slot_ids, done, ctl, ew, g, v->idle_ticks,
          v->max_idle_polls);
    else
      fkd::box_server_kernel<DIM, false><<<v->nslots, fkd::kWave, 0, v->stream>>>(
          t->recs.as<float>(), 0u, t->P, rq, rp, ids, v->slot_ids, done, ctl, ew, g, v->idle_ticks,
          v->max_idle_polls);
    return launch_check("box_server");
  }));
  v->gen.store(g, std::memory_order_release);
  v->launches.fetch_add(1, std::memory_order_relaxed);
  return FKD_OK;
}

// Starts a launch unless one is live.  A launch that is closing still counts
// as live until its last wave has left; a request it did not take is served
// by the next one, which the waiting caller starts.
int server_ensure(fkd_box_server_s* v) {
  if (v->running()) return FKD_OK;
  std::lock_guard<std::mutex> g(v->launch_mu);
  if (v->running()) return FKD_OK;
  return server_launch(v);
}

}  // namespace

int fkd_box_server_create(fkd_tree_t t, uint32_t slots, uint32_t slot_ids, uint32_t idle_us,
                          fkd_box_server_t* out) {
  if (!t || !out) return fail(FKD_ERR_INVALID, "tree / out is NULL");
  *out = nullptr;
  if (!t->built) return fail(FKD_ERR_STATE, "tree not built");
  if (t->n_events != 1) return fail(FKD_ERR_INVALID, "a box server searches a single tree");
  if (slots == 0 || slots > FKD_SERVER_MAX_SLOTS)
    return fail(FKD_ERR_INVALID, "slots must be 1.." + std::to_string(FKD_SERVER_MAX_SLOTS));
  if (slot_ids == 0 || slot_ids > FKD_SMALL_MAX_SLOT) return fail(FKD_ERR_INVALID, "slot_ids must be 1..2^20");
  if (idle_us == 0 || idle_us > 10u * 1000u * 1000u) return fail(FKD_ERR_INVALID, "idle_us must be 1..10^7");
  int dev = 0, rate_khz = 0;
  FKD_HIP(hipGetDevice(&dev));
  FKD_HIP(hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, dev));
  if (rate_khz <= 0) return fail(FKD_ERR_HIP, "the device reports no wall-clock rate");
  auto* v = new fkd_box_server_s();
  auto bail = [&](int rc) {
    delete v;
    return rc;
  };
  v->tree = t;
  v->dim = t->dim;
  v->nslots = slots;
  v->slot_ids = slot_ids;
  v->rate_khz = rate_khz;
  v->idle_ticks = (unsigned long long)rate_khz * idle_us / 1000ull;
  // Checks of an idle launch come at least 16 PCIe round trips apart (about
  // 16 us): this bound only matters if the clock were to stall.
  v->max_idle_polls = std::max<uint32_t>(idle_us, 1000u);
  v->seq.assign(slots, 0);
  int rc = v->req.ensure((size_t)slots * sizeof(fkd::BoxReq));
  if (rc == FKD_OK) rc = v->rep.ensure((size_t)slots * sizeof(fkd::BoxRep));
  if (rc == FKD_OK) rc = v->ids.ensure((size_t)slots * slot_ids * 4);
  if (rc == FKD_OK) rc = v->word.ensure(64);
  if (rc == FKD_OK) rc = v->ctl.ensure(sizeof(fkd::BoxCtl) + (size_t)slots * 4);
  if (rc != FKD_OK) return bail(rc);
  std::memset(v->req.h, 0, v->req.bytes);
  std::memset(v->rep.h, 0, v->rep.bytes);
  std::memset(v->word.h, 0, v->word.bytes);
  hipError_t e = hipStreamCreateWithFlags(&v->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMemsetAsync(v->ctl.p, 0, v->ctl.bytes, v->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(v->stream);
  if (e != hipSuccess) return bail(fail(FKD_ERR_HIP, std::string("box server set-up: ") + hipGetErrorString(e)));
  v->free_mask.store(slots == 64 ? ~0ull : ((1ull << slots) - 1ull));
  t->server_list.push_back(v);
  *out = v;
  return FKD_OK;
}

int fkd_box_server_search(fkd_box_server_t v, const float* box, uint32_t* ids_out, uint32_t cap,
                          uint32_t* count) {
  if (!v || !box || !count) return fail(FKD_ERR_INVALID, "server / box / count is NULL");
  if (cap && !ids_out) return fail(FKD_ERR_INVALID, "ids_out is NULL");
  if (!v->tree || v->stopped) return fail(FKD_ERR_STATE, "the box server's tree was destroyed");
  // take a free slot (yielding: its holder may need this core)
  uint32_t s = 0;
  uint64_t m = v->free_mask.load(std::memory_order_acquire);
  while (true) {
    if (!m) {
      std::this_thread::yield();
      m = v->free_mask.load(std::memory_order_acquire);
      continue;
    }
    const uint64_t bit = m & (~m + 1ull);
    if (v->free_mask.compare_exchange_weak(m, m & ~bit, std::memory_order_acquire)) {
      s = (uint32_t)__builtin_ctzll(bit);
      break;
    }
  }
  auto release = [&]() { v->free_mask.fetch_or(1ull << s, std::memory_order_release); };
  const int nf = 2 * v->dim;
  fkd::BoxReq* rq = v->h_req() + s;
  const fkd::BoxRep* rp = v->h_rep() + s;
  for (int j = 0; j < nf; ++j) rq->box[j] = box[j];
  const auto t0 = std::chrono::steady_clock::now();
  while (true) {
    uint32_t q = v->seq[s] + 1u;
    if (q == fkd::kServerStop) q = 0;
    v->seq[s] = q;
    __atomic_store_n(&rq->seq, q, __ATOMIC_RELEASE);
    int rc = server_ensure(v);
    if (rc != FKD_OK) {
      release();  // a later launch may still answer it; the next holder waits for its own number
      return rc;
    }
    // Spin (a sleeping caller takes tens of microseconds to wake, longer than
    // an answer); yield once the wait is long, in case callers outnumber cores.
    uint32_t spins = 0;
    while (__atomic_load_n(&rp->seq, __ATOMIC_ACQUIRE) != q) {
      ++spins;
      if (spins < 2048u) {
        __builtin_ia32_pause();
        continue;
      }
      std::this_thread::yield();
      if ((spins & 255u) == 0u) {
        rc = server_ensure(v);
        if (rc != FKD_OK) {
          release();
          return rc;
        }
        if (std::chrono::steady_clock::now() - t0 > std::chrono::seconds(10))
          return fail(FKD_ERR_HIP, "box server: no reply within 10 s (the slot stays taken)");
      }
    }
    // The wave read the slot's line with one load, which need not be one
    // snapshot of it: a box other than the one posted is searched again.
    if (std::memcmp(rp->echo, rq->box, (size_t)nf * 4) == 0) break;
    v->reposts.fetch_add(1, std::memory_order_relaxed);
  }
  const uint32_t c = rp->count;
  *count = c;
  if (c != fkd::kSmallOverflow && cap) {
    const uint32_t k = std::min(c, std::min(cap, v->slot_ids));
    std::memcpy(ids_out, static_cast<const uint32_t*>(v->ids.h) + (size_t)s * v->slot_ids, (size_t)k * 4);
  }
  v->round_trip_ns.fetch_add(
      (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now() - t0).count(),
      std::memory_order_relaxed);
  v->walk_ticks.fetch_add(rp->walk_ticks, std::memory_order_relaxed);
  v->requests.fetch_add(1, std::memory_order_relaxed);
  release();
  return FKD_OK;
}

int fkd_box_server_stats(fkd_box_server_t v, uint64_t* launches, uint64_t* requests, int* running) {
  if (!v) return fail(FKD_ERR_INVALID, "server is NULL");
  if (launches) *launches = v->launches.load();
  if (requests) *requests = v->requests.load();
  if (running) *running = v->running() ? 1 : 0;
  return FKD_OK;
}

int fkd_box_server_timing(fkd_box_server_t v, double* round_trip_us, double* walk_us, uint64_t* reposts) {
  if (!v) return fail(FKD_ERR_INVALID, "server is NULL");
  const uint64_t n = v->requests.load();
  if (round_trip_us) *round_trip_us = n ? (double)v->round_trip_ns.load() / n / 1000.0 : 0.0;
  if (walk_us) *walk_us = n ? (double)v->walk_ticks.load() * 1000.0 / v->rate_khz / n : 0.0;
  if (reposts) *reposts = v->reposts.load();
  return FKD_OK;
}

void fkd_box_server_destroy(fkd_box_server_t v) {
  if (!v) return;
  v->stop();
  if (v->tree) {
    auto& l = v->tree->server_list;
    l.erase(std::remove(l.begin(), l.end(), v), l.end());
  }
  delete v;
}

int fkd_search_status(fkd_tree_t t, fkd_stream_t stream, uint64_t* total, int* complete) {
  if (!t) return fail(FKD_ERR_INVALID, "tree is NULL");
  if (!t->last_offsets) return fail(FKD_ERR_STATE, "no fkd_search to report on");
  hipStream_t s = static_cast<hipStream_t>(stream);
  uint64_t tot = 0;
  FKD_HIP(hipMemcpyAsync(&tot, t->last_offsets + t->last_nq, 8, hipMemcpyDeviceToHost, s));
  FKD_HIP(hipStreamSynchronize(s));
  if (total) *total = tot;
  if (complete) *complete = tot <= t->last_capacity ? 1 : 0;
  return FKD_OK;
}

int fkd_graph_create(fkd_tree_t t, uint32_t nq, const float* mins, const float* maxs, const uint32_t* qev,
                     uint64_t* dev_offsets, uint32_t* dev_ids_out, uint64_t capacity, fkd_stream_t stream,
                     fkd_graph_t* out) {
  StreamScope scope(static_cast<hipStream_t>(stream));
  if (!t || !out) return fail(FKD_ERR_INVALID, "tree / out is NULL");
  *out = nullptr;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (!s) return fail(FKD_ERR_INVALID, "graph capture needs a created stream, not the null stream");
  hipEvent_t ts = t->trace_start, te = t->trace_stop;
  t->trace_start = t->trace_stop = nullptr;
  // Eager runs until every buffer has its final size and the results are
  // complete; then the same sequence is captured.
  int complete = 0;
  uint64_t total = 0;
  for (int attempt = 0; attempt < 3 && !complete; ++attempt) {
    FKD_TRY(fkd_build(t, stream));
    FKD_TRY(fkd_search(t, nq, mins, maxs, qev, dev_offsets, dev_ids_out, capacity, stream));
    FKD_TRY(fkd_search_status(t, stream, &total, &complete));
    if (!complete && total > capacity) {
      t->trace_start = ts;
      t->trace_stop = te;
      return fail(FKD_ERR_CAPACITY, "capacity " + std::to_string(capacity) + " < " + std::to_string(total) + " hits");
    }
  }
  if (!complete) return fail(FKD_ERR_CAPACITY, "search did not complete");
  fkd_graph_s* g = new fkd_graph_s();
  int rc = FKD_OK;
  hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    rc = fkd_build(t, stream);
    if (rc == FKD_OK) rc = fkd_search(t, nq, mins, maxs, qev, dev_offsets, dev_ids_out, capacity, stream);
    hipError_t e2 = hipStreamEndCapture(s, &g->graph);
    if (rc == FKD_OK && e2 != hipSuccess) e = e2;
  }
  if (rc == FKD_OK && e == hipSuccess) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
  t->trace_start = ts;
  t->trace_stop = te;
  if (rc != FKD_OK || e != hipSuccess) {
    delete g;
    return rc != FKD_OK ? rc : fail(FKD_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
  }
  g->tree = t;
  g->nq = nq;
  g->offsets = dev_offsets;
  g->capacity = capacity;
  ++t->graph_count;
  t->graph_list.push_back(g);
  *out = g;
  return FKD_OK;
}

int fkd_graph_launch(fkd_graph_t g, fkd_stream_t stream) {
  if (!g) return fail(FKD_ERR_INVALID, "graph is NULL");
  if (!g->tree) return fail(FKD_ERR_STATE, "the graph's tree was destroyed");
  FKD_HIP(hipGraphLaunch(g->exec, static_cast<hipStream_t>(stream)));
  // fkd_search_status reports on the replayed search
  g->tree->last_nq = g->nq;
  g->tree->last_offsets = g->offsets;
  g->tree->last_capacity = g->capacity;
  return FKD_OK;
}

void fkd_graph_destroy(fkd_graph_t g) { delete g; }

int fkd_trim(fkd_tree_t t, fkd_stream_t stream) {
  StreamScope scope(static_cast<hipStream_t>(stream));
  if (!t) return fail(FKD_ERR_INVALID, "tree is NULL");
  if (t->graph_count)
    return fail(FKD_ERR_STATE, "a HIP graph of this handle holds its scratch (destroy the graph first)");
  FKD_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
  for (DevBuf* b : t->scratch()) b->release();
  t->q_plan_nq = 0;
  t->last_offsets = nullptr;
  t->last_total = 0;
  return FKD_OK;
}

int fkd_device_bytes(fkd_tree_t t, uint64_t* bytes) {
  if (!t || !bytes) return fail(FKD_ERR_INVALID, "tree / bytes is NULL");
  uint64_t b = 0;
  for (const DevBuf* d : t->buffers()) b += d->bytes;
  *bytes = b;
  return FKD_OK;
}

int fkd_set_tuning(fkd_tree_t t, const char* key, uint32_t value) {
  if (!t || !key) return fail(FKD_ERR_INVALID, "tree / key is NULL");
  const std::string k(key);
  if (k == "query_key_bits") {
    if (value > 32u / (uint32_t)t->dim) return fail(FKD_ERR_INVALID, "query_key_bits * dim must be <= 32");
    t->tune.query_key_bits = value;
  } else if (k == "fill_order") {
    if (value > 1) return fail(FKD_ERR_INVALID, "fill_order must be 0 or 1");
    t->tune.fill_order = value;
  } else if (k == "key_volume_bits") {
    if (value > 6) return fail(FKD_ERR_INVALID, "key_volume_bits must be <= 6");
    t->tune.key_volume_bits = value;
  } else if (k == "fin_log" || k == "build_radix_items") {
    const bool fl = k == "fin_log";
    if (fl && (value < fkd::kFinLogMin || value > fkd::kFinLogMax)) return fail(FKD_ERR_INVALID, "fin_log must be 9..11");
    if (!fl && value != 8 && value != 16) return fail(FKD_ERR_INVALID, "build_radix_items must be 8 or 16");
    // a captured graph launches with the tilings these lay out
    if (t->graph_count && value != (fl ? t->tune.fin_log : t->tune.build_radix_items))
      return fail(FKD_ERR_STATE, k + " cannot change while a HIP graph of this handle exists");
    (fl ? t->tune.fin_log : t->tune.build_radix_items) = value;
  } else if (k == "level_pairs") {
    if (value > 1) return fail(FKD_ERR_INVALID, "level_pairs must be 0 or 1");
    t->tune.level_pairs = value;
  } else if (k == "bucket" || k == "count_bucket") {
    if (value > 2) return fail(FKD_ERR_INVALID, k + " must be 0, 1 or 2");
    (k == "bucket" ? t->tune.bucket : t->tune.count_bucket) = value;
  } else if (k == "bucket_park") {
    if (value > 8) return fail(FKD_ERR_INVALID, "bucket_park must be 0..8");
    t->tune.bucket_park = value;
  } else if (k == "fill_window" || k == "count_window") {
    if (value != 0 && value != 256 && value != 1024 && value != 4096)
      return fail(FKD_ERR_INVALID, k + " must be 0, 256, 1024 or 4096");
    (k == "fill_window" ? t->tune.fill_window : t->tune.count_window) = value;
  } else {
    return fail(FKD_ERR_INVALID, "unknown tuning key " + k);
  }
  return FKD_OK;
}

int fkd_get_tuning(fkd_tree_t t, const char* key, uint32_t* value) {
  if (!t || !key || !value) return fail(FKD_ERR_INVALID, "NULL argument");
  const std::string k(key);
  if (k == "query_key_bits") *value = t->tune.query_key_bits;
  else if (k == "fill_order") *value = t->tune.fill_order;
  else if (k == "fill_window") *value = t->tune.fill_window;
  else if (k == "count_window") *value = t->tune.count_window;
  else if (k == "key_volume_bits") *value = t->tune.key_volume_bits;
  else if (k == "fin_log") *value = t->tune.fin_log;
  else if (k == "build_radix_items") *value = t->tune.build_radix_items;
  else if (k == "level_pairs") *value = t->tune.level_pairs;
  else if (k == "bucket") *value = t->tune.bucket;
  else if (k == "count_bucket") *value = t->tune.count_bucket;
  else if (k == "bucket_park") *value = t->tune.bucket_park;
  else return fail(FKD_ERR_INVALID, "unknown tuning key " + k);
  return FKD_OK;
}

int fkd_search_host(fkd_tree_t t, uint32_t nq, const float* host_mins, const float* host_maxs,
                    const uint32_t* host_qev, uint64_t* host_offsets, fkd_stream_t stream) {
  StreamScope scope(static_cast<hipStream_t>(stream));
  if (!t) return fail(FKD_ERR_INVALID, "tree is NULL");
  if (!host_offsets || (nq && (!host_mins || !host_maxs))) return fail(FKD_ERR_INVALID, "NULL search buffer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const size_t bb = (size_t)nq * t->dim * 4;
  const size_t bb_al = (bb + 255) & ~size_t(255);
  FKD_TRY(t->q_boxes.ensure(bb_al + bb + 256));
  float* dmins = t->q_boxes.as<float>();
  float* dmaxs = reinterpret_cast<float*>(reinterpret_cast<char*>(t->q_boxes.p) + bb_al);
  const uint32_t* dqev = nullptr;
  if (nq) {
    FKD_HIP(hipMemcpyAsync(dmins, host_mins, bb, hipMemcpyHostToDevice, s));
    FKD_HIP(hipMemcpyAsync(dmaxs, host_maxs, bb, hipMemcpyHostToDevice, s));
    if (host_qev) {
      FKD_TRY(t->q_ev.ensure((size_t)nq * 4));
      FKD_HIP(hipMemcpyAsync(t->q_ev.p, host_qev, (size_t)nq * 4, hipMemcpyHostToDevice, s));
      dqev = t->q_ev.as<uint32_t>();
    }
  }
  FKD_TRY(t->q_off.ensure(((size_t)nq + 1) * 8));
  uint64_t total = 0;
  int complete = 0;
  for (int attempt = 0; attempt < 3 && !complete; ++attempt) {
    const uint64_t cap = t->q_hits.bytes / 4;
    FKD_TRY(fkd_search(t, nq, dmins, dmaxs, dqev, t->q_off.as<uint64_t>(), t->q_hits.as<uint32_t>(), cap, stream));
    FKD_TRY(fkd_search_status(t, stream, &total, &complete));
    if (!complete) FKD_TRY(t->q_hits.ensure(std::max<uint64_t>(total, 1) * 4));
  }
  if (!complete) return fail(FKD_ERR_CAPACITY, "search output did not fit after regrowing");
  FKD_HIP(hipMemcpyAsync(host_offsets, t->q_off.p, ((size_t)nq + 1) * 8, hipMemcpyDeviceToHost, s));
  FKD_HIP(hipStreamSynchronize(s));
  t->last_nq = nq;
  t->last_total = total;
  return FKD_OK;
}

int fkd_fetch_results(fkd_tree_t t, uint32_t* host_ids, fkd_stream_t stream) {
  StreamScope scope(static_cast<hipStream_t>(stream));
  if (!t) return fail(FKD_ERR_INVALID, "tree is NULL");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (t->last_total) {
    if (!host_ids) return fail(FKD_ERR_INVALID, "host_ids is NULL");
    FKD_HIP(hipMemcpyAsync(host_ids, t->q_hits.p, t->last_total * 4, hipMemcpyDeviceToHost, s));
  }
  FKD_HIP(hipStreamSynchronize(s));
  return FKD_OK;
}

int fkd_device_info(char* name, size_t name_len, int* compute_units, size_t* free_bytes, size_t* total_bytes) {
  int dev = 0;
  FKD_HIP(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  FKD_HIP(hipGetDeviceProperties(&prop, dev));
  if (name && name_len) {
    std::snprintf(name, name_len, "%s%s(%s)", prop.name, prop.name[0] ? " " : "", prop.gcnArchName);
  }
  if (compute_units) *compute_units = prop.multiProcessorCount;
  size_t fr = 0, tot = 0;
  FKD_HIP(hipMemGetInfo(&fr, &tot));
  if (free_bytes) *free_bytes = fr;
  if (total_bytes) *total_bytes = tot;
  return FKD_OK;
}

int fkd_host_alloc(void** ptr, size_t bytes) {
  if (!ptr) return fail(FKD_ERR_INVALID, "ptr is NULL");
  // non-coherent (host-cached) pages: host writes run at memcpy speed and
  // copies to the device at full link rate (measured: 57 GB/s against 3 GB/s
  // with the default flags on this platform)
  FKD_HIP(hipHostMalloc(ptr, std::max<size_t>(bytes, 64), hipHostMallocNonCoherent));
  return FKD_OK;
}
int fkd_host_free(void* ptr) {
  if (ptr) FKD_HIP(hipHostFree(ptr));
  return FKD_OK;
}

int fkd_pool_info(uint64_t* reserved, uint64_t* used) {
  device_cache()->info(reserved, used);
  return FKD_OK;
}

int fkd_release_cached(void) {
  FKD_HIP(device_cache()->release());
  return FKD_OK;
}

int fkd_malloc(void** ptr, size_t bytes) {
  if (!ptr) return fail(FKD_ERR_INVALID, "ptr is NULL");
  FKD_HIP(hipMalloc(ptr, std::max<size_t>(bytes, 256)));
  return FKD_OK;
}
int fkd_free(void* ptr) {
  if (ptr) FKD_HIP(hipFree(ptr));
  return FKD_OK;
}
int fkd_malloc_async(void** ptr, size_t bytes, fkd_stream_t stream) {
  if (!ptr) return fail(FKD_ERR_INVALID, "ptr is NULL");
  StreamScope scope(static_cast<hipStream_t>(stream));
  FKD_HIP(dev_alloc(ptr, std::max<size_t>(bytes, 256)));
  return FKD_OK;
}
int fkd_free_async(void* ptr, fkd_stream_t stream) {
  StreamScope scope(static_cast<hipStream_t>(stream));
  if (ptr) FKD_HIP(dev_free(ptr));
  return FKD_OK;
}
int fkd_memcpy_htod(void* dst, const void* src, size_t bytes, fkd_stream_t stream) {
  if (bytes) FKD_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, static_cast<hipStream_t>(stream)));
  return FKD_OK;
}
int fkd_memcpy_dtoh(void* dst, const void* src, size_t bytes, fkd_stream_t stream) {
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (bytes) FKD_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
  FKD_HIP(hipStreamSynchronize(s));
  return FKD_OK;
}
int fkd_copy_u32(uint32_t* dst, const uint32_t* src, uint64_t n, fkd_stream_t stream) {
  if (n == 0) return FKD_OK;
  fkd::copy_u32_kernel<<<2048, 256, 0, static_cast<hipStream_t>(stream)>>>(dst, src, n);
  return launch_check("copy_u32");
}
int fkd_gather_u64(uint64_t* dst, const uint64_t* src, const uint64_t* idx, uint64_t n, fkd_stream_t stream) {
  if (n == 0) return FKD_OK;
  if (!dst || !src || !idx) return fail(FKD_ERR_INVALID, "NULL gather buffer");
  fkd::gather_u64_kernel<<<grid_for(n), 256, 0, static_cast<hipStream_t>(stream)>>>(dst, src, idx, n);
  return launch_check("gather_u64");
}
int fkd_stream_create(fkd_stream_t* stream) {
  hipStream_t s;
  FKD_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *stream = s;
  return FKD_OK;
}
int fkd_stream_destroy(fkd_stream_t stream) {
  hipStream_t s = static_cast<hipStream_t>(stream);
  FKD_HIP(hipStreamSynchronize(s));
  device_cache()->retire_stream(s);
  FKD_HIP(hipStreamDestroy(s));
  return FKD_OK;
}
int fkd_stream_sync(fkd_stream_t stream) {
  FKD_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
  return FKD_OK;
}
int fkd_event_create(fkd_event_t* ev) {
  hipEvent_t e;
  FKD_HIP(hipEventCreate(&e));
  *ev = e;
  return FKD_OK;
}
int fkd_event_destroy(fkd_event_t ev) {
  FKD_HIP(hipEventDestroy(static_cast<hipEvent_t>(ev)));
  return FKD_OK;
}
int fkd_event_record(fkd_event_t ev, fkd_stream_t stream) {
  FKD_HIP(hipEventRecord(static_cast<hipEvent_t>(ev), static_cast<hipStream_t>(stream)));
  return FKD_OK;
}
int fkd_event_query(fkd_event_t ev) {
  const hipError_t e = hipEventQuery(static_cast<hipEvent_t>(ev));
  if (e == hipSuccess) return 1;
  if (e == hipErrorNotReady) {
    (void)hipGetLastError();  // an answer, not an error for the next launch check
    return 0;
  }
  return fail(FKD_ERR_HIP, std::string("hipEventQuery: ") + hipGetErrorString(e));
}
int fkd_event_elapsed_ms(fkd_event_t a, fkd_event_t b, float* ms) {
  FKD_HIP(hipEventSynchronize(static_cast<hipEvent_t>(b)));
  FKD_HIP(hipEventElapsedTime(ms, static_cast<hipEvent_t>(a), static_cast<hipEvent_t>(b)));
  return FKD_OK;
}

int fkd_shape_subtree(uint32_t n, uint32_t L, uint32_t t, uint32_t* start, uint32_t* size) {
  if (n == 0) return fail(FKD_ERR_INVALID, "n == 0");
  const fkd::Shape sh = fkd::shape_of(n);
  if (L > sh.D || t >= (1u << L)) return fail(FKD_ERR_INVALID, "L / t out of range");
  fkd::subtree_range(sh, L, t, *start, *size);
  return FKD_OK;
}

int fkd_shape_locate(uint32_t n, uint32_t L, uint32_t j, uint32_t* t) {
  if (n == 0) return fail(FKD_ERR_INVALID, "n == 0");
  const fkd::Shape sh = fkd::shape_of(n);
  if (L > sh.D || j >= n) return fail(FKD_ERR_INVALID, "L / j out of range");
  *t = fkd::subtree_at(sh, L, j);
  return FKD_OK;
}

}  // extern "C"
