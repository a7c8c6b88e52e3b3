/*
 * kdtree_hip.h -- C ABI of libkdtree_hip.so, the MI355X (gfx950) KD-tree build
 * and exact k-NN query engine.
 *
 * The ABI is what the reference's binding layer would call in place of its Rust
 * core: plain pointers and sizes, no framework types.  Each entry point names the
 * reference interface it replaces (paths are into FujishigeTemma/kdtree-rs).
 *
 * Conventions
 *   - Points and queries are row-major float64 matrices, n x ndim.
 *   - Returned indices are int64 original row numbers; distances are float64.
 *   - Missing neighbours are padded with (+inf, n_points), src/query.rs:425-428
 *     and src/query.rs:488-491.
 *   - Every function returns 0 on success or a KDT_ERR_* code; the message of
 *     the last failure on the calling thread is kdt_last_error().  Codes 1-11
 *     are the reference's KDTreeError variants (src/error.rs:3-16) and their
 *     messages are the reference's Display strings (src/error.rs:18-51).
 *   - A built tree is immutable; concurrent kdt_query calls on one tree from
 *     several host threads are safe (the reference's frozen pyclass,
 *     src/lib.rs:100 and src/lib.rs:185).  Host-pointer entry points run on the
 *     calling thread's per-thread HIP stream.
 */
#ifndef KDTREE_HIP_H
#define KDTREE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct kdt_tree kdt_tree;

enum kdt_error {
  KDT_OK = 0,
  KDT_ERR_EMPTY_DATA = 1,           /* KDTreeError::EmptyData */
  KDT_ERR_INVALID_LEAFSIZE = 2,     /* KDTreeError::InvalidLeafsize */
  KDT_ERR_INVALID_K = 3,            /* KDTreeError::InvalidK */
  KDT_ERR_INVALID_METRIC = 4,       /* KDTreeError::InvalidMetric(p) */
  KDT_ERR_INVALID_EPS = 5,          /* KDTreeError::InvalidEps(eps) */
  KDT_ERR_INVALID_MAX_DISTANCE = 6, /* KDTreeError::InvalidMaxDistance(d) */
  KDT_ERR_NON_FINITE = 7,           /* KDTreeError::NonFiniteData */
  KDT_ERR_TOO_MANY_POINTS = 8,      /* KDTreeError::TooManyPoints(n) */
  KDT_ERR_TOO_MANY_NODES = 9,       /* KDTreeError::TooManyNodes(n) */
  KDT_ERR_INVALID_SHAPE = 10,       /* KDTreeError::InvalidShape(msg) */
  KDT_ERR_DIMENSION_MISMATCH = 11,  /* KDTreeError::DimensionMismatch */
  KDT_ERR_DEVICE = 100,             /* HIP runtime failure: no device, out of memory */
  KDT_ERR_ARGUMENT = 101            /* null pointer or size overflow at the ABI */
};

/* Query flags. */
#define KDT_QUERY_FP32_SCREEN 0x1u /* FP32 screen with a proven error margin, FP64
                                      recheck: identical results (p = 2, eps = 0 only;
                                      ignored for eps > 0).  ndim <= 16: batches of
                                      8192 queries or more on rows of 8 or 16
                                      coordinates (k <= 32) take the tile walk (see
                                      KDT_QUERY_TILE_WALK), others the traversal's
                                      own leaf screen.
                                      For 32 <= ndim <= 64, p = 2, eps = 0, k <= 32:
                                      matrix-core screen over every point (rows
                                      rounded to fp16 in a per-tree frame, with a
                                      proven error bound), FP64 recheck and a visit
                                      certificate; queries the
                                      certificate does not cover run the traversal */
#define KDT_QUERY_ASSUME_FINITE 0x2u /* kdt_query_device only: the caller has already
                                        checked that every query coordinate is finite
                                        (e.g. by an earlier call on the same buffer), so
                                        the call skips the check and returns without
                                        synchronising the stream.  Batches of 8192 or
                                        more queries still take stream-ordered scratch
                                        (hipMallocAsync) for the leaf-order sort. */

#define KDT_QUERY_MATRIX_SCREEN 0x8u /* with KDT_QUERY_FP32_SCREEN: the matrix-core
                                        screen also for 8 <= ndim < 32 (rows padded
                                        to 16 or 32 columns), where the traversal
                                        still prunes part of the tree (a per-tree
                                        choice, made by timing both) */
#define KDT_QUERY_GROUP_SHIFT 4u /* bits 4-6: lanes per query in the traversal, 0 = chosen
                                    from k (16 for k <= 16, 32 for k <= 32, else 64),
                                    1 = 8, 2 = 16, 3 = 32, 4 = 64 lanes (ignored when fewer
                                    than k); a per-tree choice, made by timing (kdtree.forest) */
#define KDT_QUERY_GROUP_MASK 0x70u
#define KDT_QUERY_SCREEN_ONLY 0x4u /* diagnostic, with KDT_QUERY_FP32_SCREEN on the
                                      matrix-core path: skip the traversal of the
                                      queries the certificate does not cover, whose
                                      output rows are then left as they were */
#define KDT_QUERY_TILE_WALK 0x400u /* with KDT_QUERY_FP32_SCREEN, rows of 1-4, 8 or 16
                                       coordinates, p = 2, eps = 0, k <= 32: the tile walk
                                       (one tree walk per 64 queries, FP32 screen of the
                                       points it reaches with a proven margin, FP64
                                       recheck and visit certificate) for any batch size
                                       and these widths; without this bit it is taken from
                                       8192 queries on, for rows of 8 or 16 */
#define KDT_QUERY_NO_TILE_WALK 0x800u /* with KDT_QUERY_FP32_SCREEN: never the tile walk (the
                                          traversal's own FP32 leaf screen instead); both
                                          give the same results */
#define KDT_QUERY_SLICES_SHIFT 12u /* bits 12-13, with the matrix-core screen: its point
                                      slices (the range of point tiles one workgroup
                                      sweeps for its block of queries, i.e. the work
                                      tile's shape) relative to its occupancy rule:
                                      0 = the rule, 1 = half as many, 2 = twice, 3 = four
                                      times; a per-tree choice, made by timing
                                      (kdtree.forest); results are the same */
#define KDT_QUERY_SLICES_MASK 0x3000u
#define KDT_QUERY_NO_ORDER 0x200u /* keep the given query order: batches of 8192 or more
                                     queries otherwise run grouped by leaf (a leaf-order
                                     sort with stream-ordered scratch); results are the
                                     same either way.  Same bit as KDT_BATCH_NO_ORDER. */

/* Batch flags (kdt_query_batch_device). */
#define KDT_BATCH_PER_TREE_LAUNCHES 0x100u /* one launch per item instead of one persistent launch */
#define KDT_BATCH_NO_ORDER 0x200u          /* keep the given query order: every item runs its own
                                              launch sequence on the caller's stream, without
                                              leaf-order sort; in kdt_query_batch_device also
                                              without the matrix-core screen (items it would take
                                              run the traversal, same results), so that the call
                                              allocates nothing, copies nothing from the host and
                                              only enqueues kernels (capturable in a graph).  A
                                              batch plan keeps the screen (its scratch is the
                                              plan's) */

/* One item of a multi-tree batch: queries (device, n_queries x ndim, ndim =
 * the tree's) against one tree, k neighbours, outputs n_queries x k (device). */
typedef struct kdt_batch_item {
  const kdt_tree* tree;
  const double* queries;
  uint64_t n_queries;
  uint64_t ndim;
  uint64_t k;
  double* out_dist;
  int64_t* out_idx;
  uint32_t group;  /* lanes per query for this item's tree (8, 16, 32, 64; 0 = from the
                      flags or from k), e.g. the width an autotuner picked */
  uint32_t flags;  /* KDT_QUERY_FP32_SCREEN, KDT_QUERY_MATRIX_SCREEN, KDT_QUERY_TILE_WALK,
                      KDT_QUERY_NO_TILE_WALK, KDT_QUERY_SLICES_* for this item on top of
                      the call's flags (other bits ignored), e.g. an autotuner's per-tree
                      choice of a screen and its work-tile shape; 0 = the call's */
} kdt_batch_item;

/* Node record exported for inspection (24 bytes, the reference's Node budget,
 * src/tree.rs:97-100).  Leaf: a = first row, b = end row.  Inner: a = left child
 * id, b = right child id. */
typedef struct kdt_node_record {
  double split_value;
  uint32_t a;
  uint32_t b;
  uint32_t split_dim;
  uint8_t is_leaf;
  uint8_t order_by_box;
  uint16_t pad;
} kdt_node_record;

/* Build a tree from host data, copied.  Replaces Tree::new / build::build,
 * src/tree.rs:39-46 and src/build.rs:19-82 (validation order :25-56).
 * `parallel` is accepted for parity with src/lib.rs:108; the device build is
 * always parallel and its result does not depend on it (src/build.rs:345-359). */
int kdt_build(const double* data, uint64_t n_points, uint64_t ndim, uint64_t leafsize,
              int parallel, kdt_tree** out);

/* Same, with `data` already in device memory (row-major, n_points x ndim).  The
 * build runs on `stream` (a hipStream_t; 0 = the null stream).  The tree keeps
 * its own copy. */
int kdt_build_device(const double* d_data, uint64_t n_points, uint64_t ndim, uint64_t leafsize,
                     void* stream, kdt_tree** out);

/* Exact k-NN for host queries.  Replaces Tree::query, src/query.rs:29-78, with
 * the dimension check of src/lib.rs:173-174 first.  max_distance = +inf means
 * None.  out_dist and out_idx hold n_queries * k values, row-major. */
int kdt_query(const kdt_tree* tree, const double* queries, uint64_t n_queries, uint64_t ndim,
              uint64_t k, double p, double max_distance, double eps, uint32_t flags,
              double* out_dist, int64_t* out_idx);

/* Same, with queries and outputs in device memory, on `stream`.  The finiteness
 * check of the queries runs on the device first and this call synchronises
 * `stream` to read it; only finite queries are then queried (the outputs are
 * untouched on an error) and the call returns once the query is enqueued.  With
 * KDT_QUERY_ASSUME_FINITE the call skips the check and only enqueues. */
int kdt_query_device(const kdt_tree* tree, const double* d_queries, uint64_t n_queries,
                     uint64_t ndim, uint64_t k, double p, double max_distance, double eps,
                     uint32_t flags, double* d_out_dist, int64_t* d_out_idx, void* stream);

/* kdt_query_device without any allocation, for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  Same arguments, checks and
 * results as kdt_query_device; the call's scratch (finiteness flag, leaf-order
 * sort, matrix-core screen lists) comes from one fixed device array of
 * KDT_NOALLOC_SCRATCH_BYTES per GPU, held from the library's load on that GPU
 * until the process ends, and the call makes no hipMalloc, hipFree or other
 * allocating call.  A query whose scratch would exceed that size is rejected
 * with KDT_ERR_ARGUMENT before any launch (e.g. more than about 80,000 queries
 * on the 64-d matrix-core screen, or about 5 million through the leaf-order
 * sort).  Every call on one GPU shares the array and overwrites it: calls
 * whose work overlaps on one GPU (several streams or threads) race, so the
 * caller orders them (e.g. one stream). */
#define KDT_NOALLOC_SCRATCH_BYTES (64u << 20)
int kdt_query_device_noalloc(const kdt_tree* tree, const double* d_queries, uint64_t n_queries,
                             uint64_t ndim, uint64_t k, double p, double max_distance, double eps,
                             uint32_t flags, double* d_out_dist, int64_t* d_out_idx, void* stream);

/* Many trees and query sets in one persistent launch.  No reference
 * counterpart (the reference answers one matrix against one tree per call,
 * src/query.rs:29-78); each item's result is identical to kdt_query_device on
 * that item.  The caller guarantees finite queries (as KDT_QUERY_ASSUME_FINITE);
 * p, max_distance and eps apply to every item and are validated like
 * kdt_query, and each item like kdt_query_device (dimension, k, buffers,
 * n_queries * k).  Runs on `stream` without synchronising: it uploads the
 * item descriptors from host memory and takes stream-ordered scratch, so it is
 * not capturable in a graph unless flags hold KDT_BATCH_NO_ORDER (use a batch
 * plan to capture the full path).  Items that run their own launch sequence
 * may run on side streams of `stream`, forked from it and joined back into it
 * before the call returns; each caller stream has its own side streams. */
int kdt_query_batch_device(const kdt_batch_item* items, uint64_t n_items, double p,
                           double max_distance, double eps, uint32_t flags, void* stream);

/* A prepared batch: the items are validated and classified once, one device
 * region holds the item descriptors and every piece of scratch the batch
 * needs (leaf-order sort, screen lists), and the plan makes its own side
 * streams (default priority).  kdt_batch_plan_run then allocates nothing,
 * copies nothing from the host and only enqueues kernels and event waits on
 * `stream` (and the plan's side streams, joined back into `stream`), so it
 * can be captured in a HIP graph and replayed; every run gives the same
 * results as kdt_query_batch_device on the same items.  The items' trees,
 * queries and outputs must outlive the plan; runs of one plan must not
 * overlap (they share its region).  No reference counterpart. */
typedef struct kdt_batch_plan kdt_batch_plan;
int kdt_batch_plan_create(const kdt_batch_item* items, uint64_t n_items, double p, double max_distance,
                          double eps, uint32_t flags, kdt_batch_plan** out);
int kdt_batch_plan_run(kdt_batch_plan* plan, void* stream);
uint64_t kdt_batch_plan_bytes(const kdt_batch_plan* plan); /* its device region */
void kdt_batch_plan_free(kdt_batch_plan* plan);            /* waits for the device first */

/* Kernel timing, a measurement aid with no reference counterpart.  While it is
 * enabled (kdt_timing_enable(1); returns the previous state), every launch of
 * the main kernels -- traversal (knn_reg, knn_lds, knn_batch) and matrix-core
 * screen (k_screen, k_screen_select, k_screen_recheck) -- is bracketed by HIP
 * events on the stream it runs on.  kdt_timing_read waits for those events and
 * returns per kernel name the launches, their summed duration, their summed
 * algorithmic work (FLOPs of the screen's contraction, 2 ndim per (query,
 * point) pair; queries for the others) and the kernel's busy time, the length
 * of the union of its launches' intervals (launches overlap when several
 * streams run them; -1 when unknown), since the last read, then one entry
 * named "*" for all of them together (its busy time: the device ran at least
 * one timed kernel); *count = entries (those beyond `capacity` are dropped).
 * Not for use under graph capture. */
typedef struct kdt_kernel_time {
  char name[48];
  uint64_t launches;
  double ms;
  double work;
  double busy_ms;
} kdt_kernel_time;
int kdt_timing_enable(int on);
int kdt_timing_read(kdt_kernel_time* out, uint32_t capacity, uint32_t* count);

/* Rows in the caller's original order (n_points x ndim).  Replaces
 * Tree::original_data, src/tree.rs:60-68 (the `data` getter, src/lib.rs:138-146). */
int kdt_original_data(const kdt_tree* tree, double* out);

/* Accessors, src/tree.rs:48-58 and src/lib.rs:123-136. */
uint64_t kdt_ndim(const kdt_tree* tree);
uint64_t kdt_n_points(const kdt_tree* tree);
uint64_t kdt_leafsize(const kdt_tree* tree);
uint64_t kdt_n_nodes(const kdt_tree* tree);

/* Copy the tree's internal arrays to host memory (any pointer may be null):
 * rows in tree order (n_points x ndim), original indices (n_points), nodes
 * (n_nodes) and boxes (n_nodes x 2 x ndim, [lo | hi]).  These are the reference
 * Tree fields data/indices/nodes/boxes (src/tree.rs:29-36); used by the build
 * parity tests. */
int kdt_export(const kdt_tree* tree, double* data, uint32_t* indices, kdt_node_record* nodes,
               double* boxes);

/* Release a tree and its device memory.  Replaces Drop for Tree. */
void kdt_free(kdt_tree* tree);

/* Message of the last error on this thread ("" if none).  Replaces
 * <KDTreeError as Display>::fmt, src/error.rs:18-51. */
const char* kdt_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* KDTREE_HIP_H */
