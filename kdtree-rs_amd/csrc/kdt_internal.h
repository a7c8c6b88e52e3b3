// kdt_internal.h -- layouts shared by the build, the query and the C ABI.
//
// Device tree image (all in HBM, owned by kdt_tree):
//   nodes  DevNode[n_nodes]            16 B per node, preorder ids (left = id+1)
//   boxes  f64[n_nodes][2][ndim]       tight box per node, [lo | hi]
//   pts    f64 leaf buckets            leaf [start,end) of L rows stored
//                                      coordinate-major: pts[start*ndim + c*L + o]
//   pts32  f32 leaf buckets            same layout, for the FP32 screen
//   idx    u32[n_points]               original row of tree position
//   rows   f64[n_points][ndim]         rows in tree order (row-major), the
//                                      reference's Tree::data
//   rows32c f32[n + 8][ndim]           ndim <= 16 only: rows in tree order as
//                                      fl32(fl64(x - c)), row-major, 8 zero
//                                      padding rows (the tile walk, kdt_walk.hip)
//   pn32c  f32[n + 8]                  ndim <= 16 only: their norms fl32(|row|^2)
//                                      (NaN on the padding rows)
//   boxes32c f32[n_nodes][2][ndim]     ndim <= 16 only: the boxes in the frame as
//                                      FP32 rounded outwards (the tile walk)
//   frame  f64[ndim + 1]               ndim <= 64 only: the screens' frame,
//                                      centre c (middle of the root box per
//                                      coordinate), then the scale sigma = 2^s
//                                      that maps the rows' largest |x - c|
//                                      below 2^8 (the matrix-core screen's)
//   rows_h f16[n_pad][stride]          8 <= ndim <= 64 only: rows in tree order as
//                                      fp16(fl32(sigma (x - c))) (fp16 subnormals
//                                      flushed to zero), zero-padded to stride 16,
//                                      32 or 64 and to n_pad = 32 * ceil(n / 32) rows
//   pnorm_h u32x2[n_pad]               their squared norms pn = fl32(|row_h|^2)
//                                      (+inf on padding rows) as three bf16
//                                      pieces, packed as the screen's norm-block
//                                      operand words (1 | p1 << 16, p2 | p3 << 16)
//   pn32   f32[n]                      ndim <= 16: fl32(|v|^2) per tree position (the FP32 leaf screen)
// All of them are pieces of one device allocation per tree, the tree image
// (layout and order: image_layout in kdt_build.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/kdtree_hip.h"

namespace kdt {

constexpr uint32_t LEAF_BIT = 0x80000000u;
constexpr uint32_t OBB_BIT = 0x40000000u;
constexpr uint32_t DIM_MASK = 0x3FFFFFFFu;

// One node.  Inner: split = split value, a = right child id, b = split_dim
// (| OBB_BIT when children are ordered by box distance).  Leaf: a = first row,
// b = LEAF_BIT, split holds the end row as a 64-bit integer bit pattern.
struct alignas(16) DevNode {
  double split;
  uint32_t a;
  uint32_t b;
};
static_assert(sizeof(DevNode) == 16, "DevNode is one 16-byte load");

enum MetricKind : int { M_L1 = 0, M_L2 = 1, M_LINF = 2, M_LP = 3 };

struct Tree {
  uint64_t n = 0, d = 0, leafsize = 0, n_nodes = 0;
  uint32_t depth = 0;  // number of inner levels on the longest root-leaf path
  int device = 0;
  DevNode* nodes = nullptr;
  double* boxes = nullptr;
  double* pts = nullptr;
  float* pts32 = nullptr;
  uint32_t* idx = nullptr;
  double* rows = nullptr;
  double max_abs = 0.0;  // max |coordinate| over the data (root box)
  double* frame = nullptr;     // matrix-core screen: centre[ndim], sigma
  uint16_t* rows_h = nullptr;  // matrix-core screen rows, fp16 bit patterns
  uint2* pnorm_h = nullptr;    // their norms, norm-block operand words
  double scr_sigma = 1.0;      // the frame's sigma (host copy)
  double scr_pc = 0.0;         // upper bound on max over rows of |x - c| (FP64 rows), ndim <= 64
  float* rows32c = nullptr;    // ndim <= 16: the tile walk's centred FP32 rows
  float* pn32c = nullptr;      // ndim <= 16: their norms
  float* boxes32c = nullptr;   // ndim <= 16: node boxes in the frame, FP32, rounded outwards
  float* pn32 = nullptr;  // ndim <= 16: fl32(|row32|^2) per tree position, the FP32 leaf screen's norms
  double pmax = 0.0;     // ndim <= 16: upper bound on max |row32|
  void* image = nullptr;  // the one device allocation every array above lives in (kdt_build.hip)
  uint64_t image_bytes = 0;
};

// Device scratch of one call (leaf-order keys, screen lists, batch
// descriptors).  Two modes:
//   stream    hipMallocAsync on the call's stream; release() (or the
//             destructor) frees with hipFreeAsync on the same stream, ordered
//             after the work enqueued so far;
//   region    bump allocation inside a caller-given device region: nothing is
//             allocated or freed, so the call can be captured in a HIP graph
//             or made by a caller that cannot allocate.  A request that does
//             not fit fails with hipErrorOutOfMemory; no kernel is launched
//             on a short region.
// Every piece is 256-byte aligned.  SCRATCH_ALIGN pads each piece so that
// size estimates can add `scratch_piece(bytes)` per piece.
constexpr size_t SCRATCH_ALIGN = 256;
inline size_t scratch_piece(size_t bytes) {
  return (std::max<size_t>(bytes, 1) + SCRATCH_ALIGN - 1) & ~(SCRATCH_ALIGN - 1);
}

class Scratch {
 public:
  explicit Scratch(hipStream_t s) : s_(s) {}
  Scratch(void* base, size_t bytes, hipStream_t s) : s_(s), base_((char*)base), cap_(bytes), region_(true) {}
  Scratch(const Scratch&) = delete;
  Scratch& operator=(const Scratch&) = delete;
  ~Scratch() { release(); }

  template <typename T>
  hipError_t get(T** p, size_t count) {
    const size_t bytes = scratch_piece(count * sizeof(T));
    *p = nullptr;
    if (region_) {
      if (bytes > cap_ - used_) return hipErrorOutOfMemory;  // used_ <= cap_ always
      *p = reinterpret_cast<T*>(base_ + used_);
      used_ += bytes;
      return hipSuccess;
    }
    void* q = nullptr;
    const hipError_t e = hipMallocAsync(&q, bytes, s_);
    if (e == hipSuccess) {
      owned_.push_back(q);
      *p = reinterpret_cast<T*>(q);
    }
    return e;
  }
  void release() {
    for (void* q : owned_) (void)hipFreeAsync(q, s_);
    owned_.clear();
  }
  bool region() const { return region_; }
  hipStream_t stream() const { return s_; }
  size_t used() const { return used_; }

 private:
  hipStream_t s_;
  char* base_ = nullptr;
  size_t cap_ = 0, used_ = 0;
  bool region_ = false;
  std::vector<void*> owned_;
};

// Exclusive prefix sum of n uint32 counts into `first` (a separate array), by
// three plain kernels (block scans, a scan of the block sums, the add): no
// memset and no library scratch, so it records cleanly into a HIP graph.
// Scratch: scan_scratch_bytes(n).
size_t scan_scratch_bytes(uint64_t n);
hipError_t exclusive_scan_u32(const uint32_t* count, uint32_t* first, uint64_t n, Scratch& sc, hipStream_t s);
// Counting-sort pieces shared by the single-tree and batch leaf orders:
// clear n counts; order[first[key[i]] + rank[i]] = i for i < nq.
void launch_zero_u32(uint32_t* p, uint64_t n, hipStream_t s);
void launch_order_scatter(const uint32_t* key, const uint32_t* rank, const uint32_t* first, uint64_t nq,
                          uint32_t* order, hipStream_t s);

// Kernel timing (kdt_timing_enable): while it is on, a KernelTimer brackets
// one launch with HIP events on the launch's stream and files the pair under
// `name` with the launch's algorithmic work (FLOPs for the matrix-core
// screen, queries for a traversal) for kdt_timing_read.  Off, it does nothing.
class KernelTimer {
 public:
  explicit KernelTimer(hipStream_t s);
  ~KernelTimer();
  void stop(const char* name, double work);
  KernelTimer(const KernelTimer&) = delete;
  KernelTimer& operator=(const KernelTimer&) = delete;

 private:
  hipStream_t s_;
  hipEvent_t start_ = nullptr;
};

// Build on the device from row-major f64 data already in device memory.
// Returns 0 or a kdt_error code; `msg` receives the reason.
int build_tree(Tree& t, const double* d_data, hipStream_t s, std::string& msg);

struct QueryParams {
  uint64_t nq = 0, k = 0;
  int metric = M_L2;
  double p = 2.0, cutoff = 0.0, eps_factor = 1.0;
  bool fp32_screen = false;
  bool screen_only = false;    // diagnostic: no traversal after the matrix-core screen
  bool matrix_screen = false;  // matrix-core screen also for 8 <= ndim < 32
  int group = 0;               // lanes per query requested (8, 16, 32, 64), 0 = from k
  bool allow_order = true;    // large batches may run in leaf order
  bool allow_scratch = true;  // scratch allowed, stream-ordered or from a region (the matrix-core
                              // screen needs it); false for a pure enqueue with no region
  int slice_shift = 0;        // KDT_QUERY_SLICES_*: matrix-core screen slices x 2^slice_shift (-1 .. 2)
  int walk = 0;               // tile walk for the FP32 screen (ndim <= 16): 0 = for batches of
                              // WALK_MIN_QUERIES or more on rows of 8 or 16, 1 = any batch and
                              // width, -1 = never
};

// Batches from which the FP32 screen of rows of up to 16 coordinates takes the
// tile walk (kdt_walk.hip) rather than the traversal's own leaf screen.
constexpr uint64_t WALK_MIN_QUERIES = 8192;
// The slice scale of KDT_QUERY_SLICES_*: 0 (the rule), -1 (half), 1, 2.
inline int slice_shift_of(uint32_t flags) {
  const uint32_t c = (flags & KDT_QUERY_SLICES_MASK) >> KDT_QUERY_SLICES_SHIFT;
  return c == 1 ? -1 : (int)(c == 0 ? 0 : c - 1);
}

// Whether a batch runs in leaf order (kdt_query.hip): large enough and allowed.
bool order_applies(const QueryParams& qp);

// Lanes per query requested by KDT_QUERY_GROUP bits of the query flags (0 = none).
// Lanes per query of the traversal: the narrowest group that holds the k list
// (16 for k <= 16, 32 for k <= 32, 64 above) -- narrow groups keep more queries
// in flight per wavefront, which measured faster than covering a whole leaf per
// pass -- unless a requested width (query flags or a batch item, picked per
// tree by timing) still holds k.
inline int group_width(uint64_t k, int requested) {
  int g = k <= 16 ? 16 : k <= 32 ? 32 : 64;
  if ((requested == 8 || requested == 16 || requested == 32 || requested == 64) && (uint64_t)requested >= k)
    g = requested;
  return g;
}

inline int requested_group(uint32_t flags) {
  const uint32_t g = (flags >> 4) & 7u;
  return g >= 1 && g <= 4 ? (4 << g) : 0;
}

// Metric, cutoff and eps factor from (p, max_distance, eps) with the checks
// and messages of src/query.rs:49-62 and src/metric.rs:45-71; 0 or an error code.
int query_params(double p, double max_distance, double eps, QueryParams& qp, std::string& msg);

// The matrix-core screen's rounding of one coordinate already moved to the
// tree's frame (y = fl64(x - c), sigma the frame's scale): fp16 of
// fl32(sigma y), round to nearest even, with fp16 subnormals flushed to zero
// (the matrix cores then never see one) and the value clamped to +-2^11 first
// (tree rows stay below 2^8 by the choice of sigma; a query beyond 2^11 is
// not screened, the clamp only keeps its operands finite).  Per coordinate
// |sigma y - result| <= 2^-11 (1 + 2^-12) |sigma y| + 2^-14 within the clamp
// (2^-24 for the FP32 step, 2^-11 for fp16, 2^-14 for what is flushed).
__device__ __forceinline__ _Float16 screen_f16(double y, double sigma) {
  const float v = fminf(fmaxf((float)(y * sigma), -2048.0f), 2048.0f);
  const _Float16 h = (_Float16)v;
  return fabsf((float)h) < 0x1p-14f ? (_Float16)0.0f : h;
}

// The norm-block words of a squared norm pn (FP32): pn as three bf16 pieces
// p1 = bf16(pn), p2 = bf16(pn - p1), p3 = bf16(pn - p1 - p2) (each difference
// exact in FP32; p1 + p2 + p3 = pn up to 2^-27 pn), packed as
// (1 | p1 << 16, p2 | p3 << 16): the bf16 operand (1, p1, p2, p3) of the
// screen's norm block.  pn = +inf gives (inf, 0, 0).
__device__ __forceinline__ uint2 screen_norm_words(float pn) {
  const __bf16 p1 = (__bf16)pn;
  const float r1 = __builtin_isinf(pn) ? 0.0f : pn - (float)p1;
  const __bf16 p2 = (__bf16)r1;
  const __bf16 p3 = (__bf16)(r1 - (float)p2);
  return make_uint2(0x3F80u | ((uint32_t)__builtin_bit_cast(uint16_t, p1) << 16),
                    (uint32_t)__builtin_bit_cast(uint16_t, p2) | ((uint32_t)__builtin_bit_cast(uint16_t, p3) << 16));
}

}  // namespace kdt
struct kdt_tree;
namespace kdt {
// Argument checks of a query call (kdt_api.hip): 0 or an error code (whose
// message is already set); scalar_rc / smsg carry the eps, max_distance and p
// checks, reported after the finiteness check of the queries.
int query_args(const kdt_tree* tree, const double* q, uint64_t nq, uint64_t ndim, uint64_t k, double p,
               double max_distance, double eps, uint32_t flags, const void* out_d, const void* out_i,
               QueryParams& qp, int& scalar_rc, std::string& smsg);
// Sets the thread's last error and returns `code`.
int fail_api(int code, std::string msg);

// Rust's Display of an f64 (shortest round-trip digits, no exponent).
std::string rust_f64(double x);

// Thread-local message behind kdt_last_error().
void set_last_error(std::string msg);

// Launch the exact k-NN query; d_q [nq x d] f64, outputs [nq x k].  Scratch
// comes from `region` when given (no allocation), else from stream-ordered
// allocations on `s`.
int launch_query(const Tree& t, const QueryParams& qp, const double* d_q, double* d_out,
                 int64_t* d_idx, hipStream_t s, std::string& msg, Scratch* region = nullptr);
// Upper bound of the scratch bytes launch_query takes for (t, qp).
size_t query_scratch_bytes(const Tree& t, const QueryParams& qp);

// Leaf order of a query batch (kdt_order.hip): d_order[nq] receives the
// queries grouped by the leaf their near path reaches.
int order_queries(const Tree& t, const double* d_q, uint64_t nq, uint32_t* d_order, Scratch& sc,
                  hipStream_t s, std::string& msg);
size_t order_scratch_bytes(const Tree& t, uint64_t nq);

// Matrix-core screen for 8 <= ndim <= 64 (kdt_screen.hip).
struct KArgs;
bool screen_applies(const Tree& t, const QueryParams& qp);
int launch_screen(const Tree& t, const QueryParams& qp, const KArgs& trav, Scratch& sc, hipStream_t s,
                  std::string& msg);
size_t screen_scratch_bytes(const Tree& t, const QueryParams& qp);
// The FP64 recheck with its visit certificate over per-query candidate lists
// (cnt[qi] tree positions at cand[qi * 64], at most 64; flag[qi] = 1 for a
// query the screen leaves to the traversal, *nflag of them), then the exact
// traversal of every flagged query (kdt_screen.hip).
int finish_candidates(const Tree& t, const QueryParams& qp, const KArgs& trav, const uint32_t* cnt,
                      const uint32_t* cand, uint8_t* flag, uint32_t* nflag, hipStream_t s, std::string& msg);

// Tile walk for rows of up to 16 coordinates (kdt_walk.hip).
bool walk_applies(const Tree& t, const QueryParams& qp);
int launch_walk(const Tree& t, const QueryParams& qp, const KArgs& trav, Scratch& sc, hipStream_t s,
                std::string& msg);
size_t walk_scratch_bytes(const Tree& t, const QueryParams& qp);

// Non-finite check of n doubles; sets *d_flag to 1 if any is NaN or infinite.
void launch_check_finite(const double* d_x, uint64_t n, int* d_flag, hipStream_t s);

// Scatter tree-order rows back to original order: out[idx[i]] = rows[i].
void launch_original_order(const Tree& t, double* d_out, hipStream_t s);

// Unpack the device node array into kdt_node_record-shaped host records.
struct NodeRecord {
  double split_value;
  uint32_t a, b, split_dim;
  uint8_t is_leaf, order_by_box;
  uint16_t pad;
};
static_assert(sizeof(NodeRecord) == 24, "NodeRecord matches kdt_node_record");

}  // namespace kdt

// The opaque handle of the C ABI.
struct kdt_tree {
  kdt::Tree t;
};

#define KDT_HIP_CHECK(expr)                                                      \
  do {                                                                           \
    hipError_t e_ = (expr);                                                      \
    if (e_ != hipSuccess) {                                                      \
      msg = std::string(#expr) + ": " + hipGetErrorString(e_);                   \
      return 100;                                                                \
    }                                                                            \
  } while (0)
