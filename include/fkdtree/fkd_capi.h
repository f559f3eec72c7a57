/*
 * C ABI of the MI355X (gfx950) FKDTree hot paths: the level-by-level k-d tree
 * build and the closed-box search, for float points of dimension 1..4.
 *
 * It replaces the work behind the reference's tree interface
 * (FKDTree.h:6-52) and its GPU backends (FKDTree_opencl.h:17-290,
 * searchInTheBoxCuda.cu:174-185).  Plain pointers and sizes only.
 *
 * Conventions
 *  - Every call returns FKD_OK (0) or a negative FKD_ERR_* code; the message
 *    of the last failure on the calling thread is fkd_last_error().
 *  - `stream` is a hipStream_t (NULL = the default stream).  Calls that only
 *    enqueue GPU work return before it finishes; calls documented as
 *    synchronous wait for `stream`.
 *  - Calls on one tree need external synchronisation; different trees are
 *    independent.
 *  - Device coordinates are SoA: coords[d * P + i] is coordinate d of point i
 *    (P = number of points).  Query boxes are SoA too: mins[d * nq + q].
 *  - Limits: at most 2^30 points per tree / batch and per event.
 *  - Preconditions (as for the reference comparator, FKDTree_cpu.h:246-252):
 *    ids unique within a tree (an event of a batch); coordinates not NaN.
 *    Creation checks both and fails with FKD_ERR_NAN / FKD_ERR_INVALID.
 *  - At most 2^30 events per handle.
 */
#ifndef FKDTREE_FKD_CAPI_H_
#define FKDTREE_FKD_CAPI_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FKD_OK 0
#define FKD_ERR_INVALID (-1)  /* bad argument */
#define FKD_ERR_HIP (-2)      /* HIP runtime error */
#define FKD_ERR_NAN (-3)      /* NaN coordinate */
#define FKD_ERR_STATE (-4)    /* e.g. search before build */
#define FKD_ERR_CAPACITY (-5) /* output buffer too small */

typedef struct fkd_tree_s* fkd_tree_t;
typedef struct fkd_graph_s* fkd_graph_t;
typedef struct fkd_box_server_s* fkd_box_server_t;
typedef void* fkd_stream_t; /* hipStream_t */

const char* fkd_last_error(void);

/* ---- tree life cycle ------------------------------------------------- */

/* Tree over n FKDPoint<float, dim> records in host memory (dim floats then a
 * uint32 id, 4*(dim+1) bytes each; FKDPoint.h:13-104).  Copies the points to
 * the device (the reference constructors copy them too: FKDTree_cpu.h:19-23,
 * FKDTree_opencl.h:20-56).  Synchronous.
 * Replaces: FKDTree_CPU / FKDTree_OpenCL constructors (FKDTree.h:9). */
int fkd_create(int dim, uint32_t n, const void* host_points, fkd_stream_t stream,
               fkd_tree_t* out);

/* Batch of independent trees ("events"): event e owns records
 * [event_offsets[e], event_offsets[e+1]) of host_points.  Ids need only be
 * unique within an event.  Synchronous.  No reference counterpart: the
 * reference builds one tree per object. */
int fkd_create_events(int dim, uint32_t n_events, const uint64_t* event_offsets,
                      const void* host_points, fkd_stream_t stream, fkd_tree_t* out);

/* Same from device-resident SoA coordinates and ids (copied on the device).
 * event_offsets is a host array; n_events = 1 with {0, P} is a single tree. */
int fkd_create_device(int dim, uint32_t n_events, const uint64_t* event_offsets,
                      const float* dev_coords, const uint32_t* dev_ids, fkd_stream_t stream,
                      fkd_tree_t* out);

void fkd_destroy(fkd_tree_t tree);

/* As fkd_destroy, but the device memory returns to the library's cache in
 * stream order on `stream` instead of after a device-wide synchronisation:
 * every call on the handle must be ordered before it on `stream` (or
 * complete). */
int fkd_destroy_async(fkd_tree_t tree, fkd_stream_t stream);

/* Build every tree of the handle.  Enqueues only.
 * Replaces: FKDTree_CPU::build (FKDTree_cpu.h:221-277),
 *           FKDTree_OpenCL::build (FKDTree_opencl.h:208-246). */
int fkd_build(fkd_tree_t tree, fkd_stream_t stream);

/* Heap-ordered tree arrays of all events end to end, copied to the host:
 * coords_out[d * P + i], ids_out[i].  Synchronous.
 * Replaces: getDimensionVector / getIdVector (FKDTree_cpu.h:209-213,
 *           FKDTree_opencl.h:251-262). */
int fkd_get_tree(fkd_tree_t tree, float* coords_out, uint32_t* ids_out, fkd_stream_t stream);

/* Device pointers of the same arrays (valid until fkd_destroy). */
int fkd_tree_device(fkd_tree_t tree, const float** dev_coords, const uint32_t** dev_ids);

/* Shape queries. */
int fkd_tree_info(fkd_tree_t tree, int* dim, uint32_t* n_events, uint64_t* n_points);

/* Release the build and search scratch of a built tree (synchronises
 * `stream`): the tree arrays, search records and input points stay, about
 * 64 B per point at dim 3.  Later builds and searches re-create what they need.  For keeping
 * many trees resident.  No reference counterpart. */
int fkd_trim(fkd_tree_t tree, fkd_stream_t stream);

/* Wire format of a built handle (little-endian), for spilling resident
 * trees and reloading them without a build:
 *   "FKDTREE1" (8 bytes), uint32 dim, uint32 n_events, uint64 n_points,
 *   uint64 event_offsets[n_events + 1],
 *   float coords[dim][n_points], uint32 ids[n_points]: the heap-ordered tree
 *   arrays of every event, events end to end.
 * fkd_serialize writes it into `buf` (NULL: only *size is set; synchronises
 * `stream`).  Serialises what getDimensionVector / getIdVector expose
 * (FKDTree_cpu.h:209-213); the reference has no file format. */
int fkd_serialize(fkd_tree_t tree, void* buf, uint64_t capacity, uint64_t* size, fkd_stream_t stream);

/* A built handle from the wire format: the arrays are uploaded as they are
 * and the search records derived from them (no build); the handle's points
 * are the tree's own, so fkd_build reproduces the same tree.  verify = 1
 * rebuilds from the points instead and fails with FKD_ERR_INVALID unless
 * every array comes out identical.  A buffer whose header, size or offsets
 * disagree fails with FKD_ERR_INVALID before anything is allocated.
 * Synchronises `stream`. */
int fkd_deserialize(const void* buf, uint64_t size, int verify, fkd_stream_t stream, fkd_tree_t* out);

/* Launch parameters of the handle (chosen by fkdtree_amd/tune.py; results
 * never depend on them):
 *   "query_key_bits"  Morton bits per dimension of the box order (0 keeps
 *                     the caller's order), key bits * dim <= 32;
 *   "fill_order"      1: the fill pass starts the search blocks with the
 *                     most work (nodes visited) first, 0: in box order
 *                     (default 1 for one tree, 0 for several events);
 *   "fill_window"     the fill pass reorders the boxes by work inside
 *                     windows of this many consecutive boxes of the box
 *                     order, so the lanes of a wave get similar work
 *                     (0, 256, 1024 or 4096; 0 = off; default 4096);
 *   "count_window"    the count pass reorders the boxes by volume (their
 *                     expected work) inside windows of this many boxes
 *                     (0, 256, 1024 or 4096; default 0 = off: it pays when
 *                     box sizes vary widely);
 *   "key_volume_bits" bits of a box-volume class inserted in the box order
 *                     below the top two Morton bits per dimension (0..6,
 *                     default 0), so boxes of similar size are neighbours;
 *                     it pays when box sizes vary widely;
 *   "fin_log"         build: subtrees of fewer than 2^fin_log points are
 *                     finished by one workgroup in LDS (9, 10 or 11:
 *                     511 / 1,023 / 2,047 points; default 10), the depths
 *                     above them by one launch per depth;
 *   "build_radix_items" build: keys per thread of the presort's digit
 *                     passes (8 or 16: tiles of 2,048 or 4,096 keys;
 *                     default 16);
 *   "level_pairs"     build, dim 3: 1 (default) moves one point array
 *                     through two depths per launch (half the bytes of
 *                     moving two arrays through one), 0 two arrays through
 *                     one depth;
 *   "bucket"          fill pass, dim <= 3: 0 (default) walks every node;
 *                     1 and 2 stop the walk three depths above the last
 *                     full depth and test each subtree reached there (7
 *                     nodes, and up to 8 on the partial depth) node by
 *                     node: 1 lets a lane that reaches one wait until
 *                     "bucket_park" eighths of its wave's live lanes wait
 *                     (0..8, default 8), 2 queues up to 8 of them per lane
 *                     and tests one beside every walk step; it pays for
 *                     boxes with many hits;
 *   "count_bucket"    the same choice for the count pass (default 0).
 * The count pass records each box's work only when fill_order or
 * fill_window, as set when it runs, will order the fill by it; a fill
 * whose settings change after its count pass orders only by work that
 * count recorded.
 * fin_log and build_radix_items cannot change while a graph of the handle
 * exists (FKD_ERR_STATE); the next build lays the tilings out for new values.
 * They replace the reference's fixed launch geometry
 * (FKDTree_opencl.h:13-15, 288-289). */
int fkd_set_tuning(fkd_tree_t tree, const char* key, uint32_t value);
int fkd_get_tuning(fkd_tree_t tree, const char* key, uint32_t* value);

/* Device memory held by the handle, in bytes. */
int fkd_device_bytes(fkd_tree_t tree, uint64_t* bytes);

/* ---- search ---------------------------------------------------------- */

/* Count pass + scan: dev_offsets[q] = first output slot of query q,
 * dev_offsets[nq] = total hits (uint64).  dev_query_events (may be NULL:
 * event 0) names the event whose tree query q searches; a query naming no
 * event of the handle (>= n_events) matches nothing.  Enqueues only.
 * Replaces, with search_fill: search_in_the_box (FKDTree_cpu.h:123-169) for
 * nq boxes at once; the counts alone are the result of
 * FKDTree_OpenCL::search_in_the_box_multiple (FKDTree_opencl.h:123-164). */
int fkd_search_count(fkd_tree_t tree, uint32_t nq, const float* dev_mins, const float* dev_maxs,
                     const uint32_t* dev_query_events, uint64_t* dev_offsets,
                     fkd_stream_t stream);

/* Whole search in one call: ids of the points inside box q go to
 * dev_ids_out[dev_offsets[q] .. dev_offsets[q+1]), dev_offsets[nq] = total.
 * Enqueues the count pass, the scan and the fill pass back to back (the fill
 * pass reads the offsets on the device: no host round trip, so the call can
 * be captured in a HIP graph).  Ids beyond `capacity` are dropped:
 * fkd_search_status reports whether the results are complete; when not, the
 * call must be repeated with a capacity of at least the reported total.
 * Replaces: search_in_the_box (FKDTree_cpu.h:123-169) for nq boxes at once. */
int fkd_search(fkd_tree_t tree, uint32_t nq, const float* dev_mins, const float* dev_maxs,
               const uint32_t* dev_query_events, uint64_t* dev_offsets, uint32_t* dev_ids_out,
               uint64_t capacity, fkd_stream_t stream);

/* The search of fkd_search (same arguments, same offsets and id sets) for
 * callers that cannot allocate device memory: code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc.  No reference
 * counterpart.  It calls no hipMalloc / hipFree: its scratch (the per-box
 * counts and the zeroed descriptors of the offsets scan) is a fixed device
 * array of the library, 16.1 MiB on each GPU it runs on, held until the
 * process ends; so it takes at most 4,194,304 boxes per call (more fail with
 * FKD_ERR_INVALID) and walks the boxes in the caller's order (no spatial
 * plan, which needs per-call buffers; sort boxes spatially for speed).  The
 * tree must be built.  Every call on one GPU shares that GPU's scratch:
 * calls that overlap on one GPU (other streams, other threads, other
 * handles) race and must be ordered by the caller.  Ids beyond `capacity`
 * are dropped; fkd_search_status reports the total.  Enqueues only. */
int fkd_search_noalloc(fkd_tree_t tree, uint32_t nq, const float* dev_mins, const float* dev_maxs,
                       const uint32_t* dev_query_events, uint64_t* dev_offsets, uint32_t* dev_ids_out,
                       uint64_t capacity, fkd_stream_t stream);

/* Low-latency search of a few boxes given in host memory, for callers that
 * issue one box (or a handful) at a time and wait for it: the reference's
 * per-box search_in_the_box called from many threads (main.cpp:351-359,
 * FKDTree_cpu.h:123-169).  One wave walks each box breadth first (one
 * dependent load per depth instead of one per visited node), reading the
 * boxes from and writing the results to page-locked host memory, so a call
 * is one launch and one synchronisation.  Single trees only (n_events 1).
 *   host_boxes   [nb][2 * dim] floats: the box's mins, then its maxs;
 *   host_counts  [nb]: hits of box b, or FKD_SMALL_OVERFLOW when one depth
 *                of its walk holds more than 2,048 nodes (search it with
 *                fkd_search instead);
 *   host_ids     [nb][slot]: the first min(count, slot) ids of box b in
 *                unspecified order; a count above `slot` means the ids were
 *                cut off (search the box again with a larger slot).
 * At most FKD_SMALL_MAX_BOXES boxes, FKD_SMALL_MAX_SLOT ids per box and
 * 2^28 ids in all (nb * slot).
 * Synchronous; calls on one tree must not overlap. */
#define FKD_SMALL_MAX_BOXES 65536u
#define FKD_SMALL_MAX_SLOT (1u << 20)
#define FKD_SMALL_OVERFLOW 0xffffffffu
int fkd_search_small(fkd_tree_t tree, uint32_t nb, const float* host_boxes, uint32_t slot,
                     uint32_t* host_counts, uint32_t* host_ids, fkd_stream_t stream);

/* Resident per-box search: the reference's search_in_the_box
 * (FKDTree_cpu.h:123-169) as the reference driver calls it, one box per call
 * from many host threads that each wait for their result (main.cpp:351-359,
 * :372).  A kernel stays resident on its own stream and polls `slots`
 * request slots in page-locked host memory; fkd_box_server_search writes its
 * box into a free slot and spins until the reply is there, so a call costs
 * no launch and no stream synchronisation, only PCIe round trips and a
 * breadth-first walk of one box by one wave (the top 12 depths, 11 at dim
 * 4, read from LDS).  The kernel leaves after `idle_us` microseconds without
 * a request (1..10^7); the next call starts it again.  While it is live it
 * holds one compute unit per slot (a one-wave workgroup with 144 KB of LDS),
 * which other kernels on the GPU cannot use meanwhile.
 *   fkd_box_server_create   single built trees only; 1..FKD_SERVER_MAX_SLOTS
 *                           slots of `slot_ids` ids (1..2^20) each.  While a
 *                           server exists fkd_build on its tree fails
 *                           (FKD_ERR_STATE); destroying the tree stops the
 *                           server, whose searches then fail.
 *   fkd_box_server_search   box = 2 * dim floats, mins then maxs.  *count is
 *                           the box's number of hits, or FKD_SMALL_OVERFLOW
 *                           when one depth of its walk holds more than 2,048
 *                           nodes (search it with fkd_search); the first
 *                           min(count, cap, slot_ids) ids, in unspecified
 *                           order, go to ids_out: a count above min(cap,
 *                           slot_ids) means the ids were cut off.  Thread
 *                           safe: calls from several threads run together,
 *                           one slot each (more callers than slots wait for
 *                           one).  Must not overlap fkd_box_server_destroy.
 *   fkd_box_server_stats    launches made, requests answered, and whether a
 *                           launch is live.
 *   fkd_box_server_timing   mean time of an answered call from posting its
 *                           box to reading the reply (host clock), the mean
 *                           walk inside it (device clock), and the requests
 *                           posted again because the wave read a box other
 *                           than the one posted (its one load of the request
 *                           line need not be a snapshot; the reply echoes
 *                           the box it searched).
 *   fkd_box_server_destroy  stops the kernel (waits for it to leave) and
 *                           frees the slots. */
#define FKD_SERVER_MAX_SLOTS 64u
int fkd_box_server_create(fkd_tree_t tree, uint32_t slots, uint32_t slot_ids, uint32_t idle_us,
                          fkd_box_server_t* out);
int fkd_box_server_search(fkd_box_server_t server, const float* box, uint32_t* ids_out, uint32_t cap,
                          uint32_t* count);
int fkd_box_server_stats(fkd_box_server_t server, uint64_t* launches, uint64_t* requests, int* running);
int fkd_box_server_timing(fkd_box_server_t server, double* round_trip_us, double* walk_us, uint64_t* reposts);
void fkd_box_server_destroy(fkd_box_server_t server);

/* Boxes [c - half_width, c + half_width] around every point of the handle,
 * in the order the points were given (SoA [dim][n_points], one float
 * rounding per bound), and the event of every box (dev_query_events may be
 * NULL): the boxes of a self-search of every point in its own tree, made on
 * the device.  Enqueues only.  No reference counterpart (the drivers form
 * such boxes on the host around every point, main.cpp:174-187). */
int fkd_boxes_around(fkd_tree_t tree, float half_width, float* dev_mins, float* dev_maxs,
                     uint32_t* dev_query_events, fkd_stream_t stream);

/* Synchronises `stream`; *total = hits of the last fkd_search on this tree,
 * *complete = 1 when every id was written. */
int fkd_search_status(fkd_tree_t tree, fkd_stream_t stream, uint64_t* total, int* complete);

/* Build + one-call search (fkd_search) of the same boxes captured as one HIP graph.
 * Runs the sequence eagerly first (so every internal buffer has its final
 * size; FKD_ERR_CAPACITY when `capacity` is below the hits), then captures
 * it on `stream` (a created stream).  fkd_graph_launch replays it: a full
 * rebuild of every tree of the handle and the search, after which
 * fkd_search_status reports on the replayed search.  The boxes and outputs
 * must stay alive while the graph exists.  The graph holds the handle's
 * build and search buffers: while it exists, fkd_trim and any call that
 * would need larger buffers (a build or search of more boxes) fail with
 * FKD_ERR_STATE; smaller eager searches work.  Destroying the tree first
 * detaches its graphs: their fkd_graph_launch fails with FKD_ERR_STATE.
 * No reference counterpart (the reference launches kernel by kernel). */
int fkd_graph_create(fkd_tree_t tree, uint32_t nq, const float* dev_mins, const float* dev_maxs,
                     const uint32_t* dev_query_events, uint64_t* dev_offsets, uint32_t* dev_ids_out,
                     uint64_t capacity, fkd_stream_t stream, fkd_graph_t* out);
int fkd_graph_launch(fkd_graph_t graph, fkd_stream_t stream);
void fkd_graph_destroy(fkd_graph_t graph);

/* Fill pass: ids of the points inside box q go to
 * dev_ids_out[dev_offsets[q] .. dev_offsets[q+1]) (slots >= capacity are
 * dropped; check dev_offsets[nq] <= capacity).  Within a query the order is
 * unspecified.  Must follow fkd_search_count on the same tree with the same
 * boxes (it reuses the processing order computed there): a fill whose box
 * pointers, event pointer or count differ from the last count's, or that
 * follows a build or another search of the handle, fails with
 * FKD_ERR_STATE.  Enqueues only. */
int fkd_search_fill(fkd_tree_t tree, uint32_t nq, const float* dev_mins, const float* dev_maxs,
                    const uint32_t* dev_query_events, const uint64_t* dev_offsets,
                    uint32_t* dev_ids_out, uint64_t capacity, fkd_stream_t stream);

/* Traversal statistics: hits and visited nodes per query (the visited set is
 * the reference's, FKDTree_cpu.h:150-163), for traffic accounting.
 * Enqueues only.  No reference counterpart. */
int fkd_search_stats(fkd_tree_t tree, uint32_t nq, const float* dev_mins, const float* dev_maxs,
                     const uint32_t* dev_query_events, uint32_t* dev_counts, uint32_t* dev_visited,
                     fkd_stream_t stream);

/* Host convenience: boxes in host memory (SoA), host_offsets[nq+1] filled,
 * result ids kept on the device until fkd_fetch_results.  Synchronous. */
int fkd_search_host(fkd_tree_t tree, uint32_t nq, const float* host_mins, const float* host_maxs,
                    const uint32_t* host_query_events, uint64_t* host_offsets,
                    fkd_stream_t stream);
int fkd_fetch_results(fkd_tree_t tree, uint32_t* host_ids, fkd_stream_t stream);

/* ---- device plumbing (memory, streams, timing) ------------------------- */

int fkd_device_info(char* name, size_t name_len, int* compute_units, size_t* free_bytes,
                    size_t* total_bytes);
int fkd_malloc(void** ptr, size_t bytes);
int fkd_free(void* ptr);
/* Stream-ordered device allocation from the library's device-memory cache,
 * which also holds every handle's buffers: a block freed on a stream is
 * reused once the work queued on that stream before the free has completed;
 * blocks return to the driver only when the device is full or on
 * fkd_release_cached (the runtime's own stream-ordered pool is emptied by
 * any synchronous free in the process, so it cannot be relied on). */
int fkd_malloc_async(void** ptr, size_t bytes, fkd_stream_t stream);
/* Page-locked host memory (hipHostMalloc): copies from it to the device run
 * asynchronously at full link rate.  For receiving FKDPoint records
 * (FKDPoint.h:13-104) that go to fkd_create_events without a staging copy. */
int fkd_host_alloc(void** ptr, size_t bytes);
int fkd_host_free(void* ptr);
/* Bytes the library's device-memory cache holds from the device, and of
 * those in use: the difference is free for the library's allocations though
 * hipMemGetInfo counts it as used. */
int fkd_pool_info(uint64_t* reserved, uint64_t* used);
/* Give the cached blocks whose last use has completed back to the driver. */
int fkd_release_cached(void);
int fkd_free_async(void* ptr, fkd_stream_t stream);
int fkd_memcpy_htod(void* dst, const void* src, size_t bytes, fkd_stream_t stream);
int fkd_memcpy_dtoh(void* dst, const void* src, size_t bytes, fkd_stream_t stream);
/* Streaming device copy of n uint32 (a plain kernel, for counter calibration). */
int fkd_copy_u32(uint32_t* dst, const uint32_t* src, uint64_t n, fkd_stream_t stream);
/* dst[i] = src[idx[i]] for i < n (device arrays; e.g. the CSR offsets at the
 * event boundaries of a batch).  Enqueues only. */
int fkd_gather_u64(uint64_t* dst, const uint64_t* src, const uint64_t* idx, uint64_t n, fkd_stream_t stream);
int fkd_stream_create(fkd_stream_t* stream);
/* Synchronises the stream, marks the cached memory freed on it as free, and
 * destroys it.  Streams of the caller's own may be destroyed with
 * hipStreamDestroy; cached blocks freed on them then become reusable after
 * a device-wide synchronisation at the next allocation that needs them. */
int fkd_stream_destroy(fkd_stream_t stream);
int fkd_stream_sync(fkd_stream_t stream);
typedef void* fkd_event_t;
int fkd_event_create(fkd_event_t* ev);
int fkd_event_destroy(fkd_event_t ev);
int fkd_event_record(fkd_event_t ev, fkd_stream_t stream);
int fkd_event_elapsed_ms(fkd_event_t start, fkd_event_t stop, float* ms);
/* 1 when all work before the event has completed, 0 when not yet, <0 on error. */
int fkd_event_query(fkd_event_t ev);

/* Record `start` / `stop` (NULL: none) on the stream right before and after
 * the traversal kernel of every following fkd_search / fkd_search_fill on
 * this tree. */
int fkd_trace_events(fkd_tree_t tree, fkd_event_t start, fkd_event_t stop);

/* ---- host-side helpers (no GPU needed) --------------------------------- */

/* Closed-form shape of the heap-ordered tree of n nodes, for tests:
 * start/size of subtree t on depth L, and the subtree holding position j. */
int fkd_shape_subtree(uint32_t n, uint32_t L, uint32_t t, uint32_t* start, uint32_t* size);
int fkd_shape_locate(uint32_t n, uint32_t L, uint32_t j, uint32_t* t);

#ifdef __cplusplus
}
#endif

#endif /* FKDTREE_FKD_CAPI_H_ */
