// kdt_batch.hip -- many trees, many query sets, few launches.
//
// No reference counterpart: the reference answers one query matrix against
// one tree per call (Tree::query, src/query.rs:29-78).  Here a batch of
// (tree, queries, k, outputs) items runs in one persistent launch per class
// (row width x FP32 leaf screen).  Every query of those items gets a global
// number; a counting sort over the batch orders them by (item, leaf reached
// by the near path) so that queries sharing a wavefront share a tree and
// nodes; the groups of a persistent grid then take queries from an atomic
// counter until the class is drained and run the same per-query descent as
// the single-tree kernels (knn_ww).  Large items, the matrix-core screen and
// the cases without a compiled width run their own launch sequence, spread
// over the caller's stream and side streams forked from it.  Every result is
// identical to a kdt_query_device call on the same item.
//
// Two ways in: kdt_query_batch_device (one call: classify, upload the item
// descriptors, take stream-ordered scratch, launch) and a batch plan
// (kdt_batch_plan_create classifies once, sizes one device region for every
// piece of scratch and the descriptors, uploads them and makes its own side
// streams; kdt_batch_plan_run then only enqueues kernels and event waits, so
// it can be captured in a HIP graph and replayed).

#include <algorithm>
#include <cstdio>
#include <map>
#include <mutex>
#include <utility>
#include <vector>

#include "../../include/kdtree_hip.h"
#include "kdt_query_kernels.h"

namespace kdt {
namespace {

struct BItem {
  KArgs a;
  uint64_t q_begin;    // first global query number of the item
  uint32_t key_base;   // first leaf-order key of the item (its tree's node ids follow)
};

__device__ __forceinline__ uint32_t find_item(const BItem* __restrict__ items, uint32_t n, uint64_t g) {
  uint32_t lo = 0, hi = n;  // last item with q_begin <= g
  while (hi - lo > 1) {
    const uint32_t m = (lo + hi) / 2;
    if (items[m].q_begin <= g) lo = m; else hi = m;
  }
  return lo;
}

// Leaf-order keys of the batch: query g of item `it` reaches leaf `node` by
// its near path; its key is key_base(it) + node, so keys order the queries by
// (item, leaf) and each class (a run of items) stays one contiguous range.
// The atomic gives each query its rank among the queries of its key.
__global__ __launch_bounds__(256) void k_batch_key(const BItem* __restrict__ items, uint32_t n_items,
                                                   uint64_t total, uint32_t* __restrict__ count,
                                                   uint32_t* __restrict__ key, uint32_t* __restrict__ rank) {
  for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total;
       g += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t it = find_item(items, n_items, g);
    const KArgs& a = items[it].a;
    const double* q = a.q + (g - items[it].q_begin) * a.D;
    uint32_t node = 0;
    for (;;) {
      const DevNode nd = a.nodes[node];
      if (nd.b & LEAF_BIT) break;
      const uint32_t dim = nd.b & DIM_MASK;
      node = q[dim] - nd.split <= 0.0 ? node + 1 : nd.a;  // nodes validated by the build
    }
    const uint32_t k = items[it].key_base + node;
    key[g] = k;
    rank[g] = atomicAdd(&count[k], 1u);
  }
}

__global__ void k_zero(unsigned long long* p, uint32_t n) {
  if (threadIdx.x < n) p[threadIdx.x] = 0;
}

// One width class of the batch: queries [first, first + count) of the batch
// order (or of the global numbering when `order` is null).  Groups take
// queries from `counter` and run them through the phased descent; a group
// switches to the query's item (tree and outputs) with each new query.
template <int DT, int G, int M, bool SCREEN, int STK>
__global__ __launch_bounds__(256, reg_min_waves<DT>()) void knn_batch_kernel(
    const BItem* __restrict__ items, uint32_t n_items, const uint32_t* __restrict__ order, uint64_t first,
    uint64_t count, unsigned long long* __restrict__ counter) {
  constexpr int GPB = 256 / G;
  __shared__ double s_sa[GPB][STK];
  __shared__ double s_sb[GPB][STK];
  __shared__ uint32_t s_sn[GPB][STK];
  __shared__ uint32_t s_sm[GPB][STK];
  __shared__ double s_ax[GPB][DT];
  __shared__ double s_q[GPB][DT];
  constexpr uint32_t CQ = SCREEN ? screen_ring<G>() : 1u;
  __shared__ uint32_t s_cq[GPB][CQ];
  __shared__ uint8_t s_cf[GPB][CQ];
  const uint32_t grp = threadIdx.x / G;
  const uint32_t lane = threadIdx.x & (G - 1);
  RegList<G> best;
  KArgs cur = items[0].a;  // batch-wide scalars (cutoff, eps, p) from the start
  auto next = [&](uint64_t& qi) -> bool {
    unsigned long long mine = 0;
    if (lane == 0) mine = atomicAdd(counter, 1ull);
    const uint64_t gq = (uint64_t)__shfl((long long)mine, 0, G);
    if (gq >= count) return false;
    const uint64_t g = order ? (uint64_t)order[first + gq] : first + gq;
    const uint32_t it = find_item(items, n_items, g);
    cur = items[it].a;
    qi = g - items[it].q_begin;
    return true;
  };
  knn_ww<DT, G, M, SCREEN>(cur, best, s_ax[grp], s_q[grp], s_sa[grp], s_sb[grp], s_sn[grp], s_sm[grp], s_cq[grp],
                           s_cf[grp], next);
}

template <int DT, int M, int STK>
void launch_class_stk(int G, bool screen, uint32_t blocks, const BItem* items, uint32_t n, const uint32_t* order,
                      uint64_t first, uint64_t count, unsigned long long* counter, hipStream_t s) {
#define KDT_CLASS(GG)                                                                                  \
  if (G == GG) {                                                                                       \
    if constexpr (M == M_L2) {                                                                         \
      if (screen) {                                                                                    \
        knn_batch_kernel<DT, GG, M, true, STK><<<blocks, 256, 0, s>>>(items, n, order, first, count, counter); \
        return;                                                                                        \
      }                                                                                                \
    }                                                                                                  \
    knn_batch_kernel<DT, GG, M, false, STK><<<blocks, 256, 0, s>>>(items, n, order, first, count, counter); \
    return;                                                                                            \
  }
  KDT_CLASS(8)
  KDT_CLASS(16)
  KDT_CLASS(32)
  KDT_CLASS(64)
#undef KDT_CLASS
}

template <int DT, int M>
void launch_class(int G, bool screen, const BItem* items, uint32_t n, const uint32_t* order, uint64_t first,
                  uint64_t count, unsigned long long* counter, uint32_t stk, hipStream_t s) {
  const uint32_t blocks = 256u * 8u;  // resident grid: ~8 blocks per CU
  KernelTimer tm(s);
  if (stk <= 32) launch_class_stk<DT, M, 32>(G, screen, blocks, items, n, order, first, count, counter, s);
  else launch_class_stk<DT, M, STK_MAX>(G, screen, blocks, items, n, order, first, count, counter, s);
  char name[48];
  std::snprintf(name, sizeof name, "knn_batch<%d,%d%s>", DT, G, screen ? ",fp32" : "");
  tm.stop(name, (double)count);
}

template <int M>
void launch_class_width(uint64_t d, int G, bool screen, const BItem* items, uint32_t n, const uint32_t* order,
                        uint64_t first, uint64_t count, unsigned long long* counter, uint32_t stk, hipStream_t s) {
  switch (d) {
    case 1: launch_class<1, M>(G, screen, items, n, order, first, count, counter, stk, s); break;
    case 2: launch_class<2, M>(G, screen, items, n, order, first, count, counter, stk, s); break;
    case 3: launch_class<3, M>(G, screen, items, n, order, first, count, counter, stk, s); break;
    case 4: launch_class<4, M>(G, screen, items, n, order, first, count, counter, stk, s); break;
    case 8: launch_class<8, M>(G, screen, items, n, order, first, count, counter, stk, s); break;
    case 16: launch_class<16, M>(G, screen, items, n, order, first, count, counter, stk, s); break;
  }
}


constexpr int NW = 6;       // compiled widths
constexpr int NC = 2 * NW;  // classes: width x FP32 leaf screen (L2 only)
constexpr uint64_t WIDTHS[NW] = {1, 2, 3, 4, 8, 16};
constexpr uint64_t SINGLE_MIN = 8192;
constexpr uint64_t ORDER_MIN = 8192;
constexpr int N_SIDE = 4;

}  // namespace

int build_kargs(const Tree& t, const QueryParams& qp, const double* d_q, double* d_out, int64_t* d_idx,
                KArgs& a);

namespace {

// An item's own screen flags, on top of the call's.
void item_flags(const kdt_batch_item& it, QueryParams& qp) {
  qp.fp32_screen = qp.fp32_screen || (it.flags & KDT_QUERY_FP32_SCREEN) != 0;
  qp.matrix_screen = qp.matrix_screen || (it.flags & KDT_QUERY_MATRIX_SCREEN) != 0;
  if (it.flags & KDT_QUERY_TILE_WALK) qp.walk = 1;
  if (it.flags & KDT_QUERY_NO_TILE_WALK) qp.walk = -1;
  if (it.flags & KDT_QUERY_SLICES_MASK) qp.slice_shift = slice_shift_of(it.flags);
}

// A classified batch: the items that run their own launch sequence and the
// descriptors of the shared class launches.
struct Layout {
  QueryParams base;
  bool no_order = false;
  std::vector<uint64_t> single;        // items with an own launch sequence, largest first
  std::vector<QueryParams> single_qp;  // their parameters
  std::vector<BItem> h;                // shared-class descriptors, class by class
  uint64_t first[NC] = {}, count[NC] = {}, kmax[NC] = {}, gmax[NC] = {};
  uint32_t stk[NC] = {};
  uint64_t total = 0;  // queries in the shared classes
  uint64_t keys = 0;   // leaf-order key space (sum of the shared items' node counts)
  bool order = false;  // the shared classes run in leaf order
};

int fail_msg(std::string& msg, std::string m, int code) {
  msg = std::move(m);
  return code;
}

// Validates the items (the checks of kdt_query_device, in its order) and
// splits them: items whose width has a compiled kernel go to the shared class
// of (width, FP32 leaf screen); general p, other widths, k > 64, the
// matrix-core screen, the tile walk and items of SINGLE_MIN queries or more run one launch
// sequence each (own lanes per query, own leaf ordering, uniform kernel
// arguments: their launch overhead is amortised, and in the shared launch
// they ran slower).  `region` says whether scratch will come from a region
// (a plan) rather than stream-ordered allocations.
int classify(const kdt_batch_item* items, uint64_t n_items, double p, double max_distance, double eps,
             uint32_t flags, bool region, Layout& L, std::string& msg) {
  if (!items && n_items) return fail_msg(msg, "null items", KDT_ERR_ARGUMENT);
  int rc = query_params(p, max_distance, eps, L.base, msg);
  if (rc) return rc;
  QueryParams& base = L.base;
  base.fp32_screen = (flags & KDT_QUERY_FP32_SCREEN) != 0;
  base.matrix_screen = (flags & KDT_QUERY_MATRIX_SCREEN) != 0;
  base.walk = (flags & KDT_QUERY_NO_TILE_WALK) ? -1 : (flags & KDT_QUERY_TILE_WALK) ? 1 : 0;
  base.slice_shift = slice_shift_of(flags);
  base.group = requested_group(flags);
  bool per_tree = (flags & KDT_BATCH_PER_TREE_LAUNCHES) != 0;
  L.no_order = (flags & KDT_BATCH_NO_ORDER) != 0;
  if (L.no_order) {
    // Given order: every item runs its own launch sequence on the caller's
    // stream, without leaf ordering; without a region also without the
    // matrix-core screen (whose scratch would be a stream-ordered allocation),
    // so that the call only enqueues kernels.  Items that would take the
    // screen run the traversal, with the same results.
    per_tree = true;
    base.allow_order = false;
    base.allow_scratch = region;
  }
  std::vector<uint64_t> cls[NC];
  for (uint64_t i = 0; i < n_items; ++i) {
    const kdt_batch_item& it = items[i];
    if (!it.tree) return fail_msg(msg, "null tree in batch item " + std::to_string(i), KDT_ERR_ARGUMENT);
    const Tree& t = it.tree->t;
    if (it.ndim != t.d)
      return fail_msg(msg, "dimension mismatch: expected " + std::to_string(t.d) + ", got " + std::to_string(it.ndim),
                      KDT_ERR_DIMENSION_MISMATCH);
    if (it.k == 0) return fail_msg(msg, "k must be greater than zero", KDT_ERR_INVALID_K);
    if (it.n_queries == 0) continue;
    if (!it.queries || !it.out_dist || !it.out_idx) return fail_msg(msg, "null buffer", KDT_ERR_ARGUMENT);
    if (it.k > 0xFFFFFFFFull || it.n_queries > (~0ull) / it.k)
      return fail_msg(msg, "n_queries * k overflows in batch item " + std::to_string(i), KDT_ERR_ARGUMENT);
    int w = -1;
    for (int j = 0; j < NW; ++j)
      if (t.d == WIDTHS[j]) w = j;
    QueryParams qp = base;
    qp.k = it.k;
    qp.nq = it.n_queries;
    item_flags(it, qp);
    if (it.group) qp.group = (int)it.group;
    if (per_tree || base.metric == M_LP || w < 0 || it.k > 64 || it.n_queries >= SINGLE_MIN ||
        screen_applies(t, qp) || walk_applies(t, qp)) {
      L.single.push_back(i);
    } else {
      const bool scr = qp.fp32_screen && base.metric == M_L2 && base.eps_factor == 1.0;  // the FP32 leaf screen: eps = 0
      cls[2 * w + (scr ? 1 : 0)].push_back(i);
    }
  }
  std::stable_sort(L.single.begin(), L.single.end(),
                   [&](uint64_t x, uint64_t y) { return items[x].n_queries > items[y].n_queries; });
  for (uint64_t i : L.single) {
    const kdt_batch_item& it = items[i];
    QueryParams qp = base;
    qp.k = it.k;
    qp.nq = it.n_queries;
    item_flags(it, qp);
    if (it.group) qp.group = (int)it.group;
    L.single_qp.push_back(qp);
  }
  // Shared classes: items numbered class by class, so each class is one
  // contiguous range of the batch's query numbers, of the leaf-order keys
  // and hence of the sorted order.
  for (int c = 0; c < NC; ++c) {
    L.first[c] = L.total;
    L.stk[c] = 2;
    for (uint64_t i : cls[c]) {
      const kdt_batch_item& it = items[i];
      BItem b{};
      QueryParams qp = base;
      qp.k = it.k;
      qp.nq = it.n_queries;
      build_kargs(it.tree->t, qp, it.queries, it.out_dist, it.out_idx, b.a);
      b.q_begin = L.total;
      b.key_base = (uint32_t)std::min<uint64_t>(L.keys, 0xFFFFFFFFull);
      L.total += it.n_queries;
      L.keys += it.tree->t.n_nodes;
      L.kmax[c] = std::max<uint64_t>(L.kmax[c], it.k);
      L.gmax[c] = std::max<uint64_t>(L.gmax[c], it.group ? it.group : (uint64_t)base.group);
      L.stk[c] = std::max<uint32_t>(L.stk[c], b.a.stk);
      L.h.push_back(b);
    }
    L.count[c] = L.total - L.first[c];
  }
  if (L.total > 0xFFFFFFFFull) return fail_msg(msg, "batch exceeds 2^32 queries", KDT_ERR_ARGUMENT);
  L.order = !L.no_order && L.total >= ORDER_MIN && L.keys <= 0xFFFFFFFFull;
  return KDT_OK;
}

// Scratch of the shared class launches: descriptors, class counters and the
// leaf-order sort.
size_t shared_scratch_bytes(const Layout& L) {
  if (L.h.empty()) return 0;
  size_t b = scratch_piece(L.h.size() * sizeof(BItem)) + scratch_piece(NC * sizeof(unsigned long long));
  if (L.order) b += 3 * scratch_piece(L.total * 4) + 2 * scratch_piece(L.keys * 4) + scan_scratch_bytes(L.keys);
  return b;
}

// Side streams for the items that run their own launch sequence: one item of
// 8192 queries fills about a quarter of the device, so consecutive such items
// on one stream leave it idle.  A call forks the side streams it uses from
// the caller's stream (an event), spreads those items over the caller's
// stream and them, largest first, and joins back only the ones it used (one
// event each) after the shared launches, which stay on the caller's stream
// and overlap with them.  Default priority.
struct SideSet {
  hipStream_t s[N_SIDE] = {};
  hipEvent_t fork = nullptr;
  hipEvent_t done[N_SIDE] = {};
  std::mutex lock;  // one call at a time on this set
  uint32_t rot = 0;
  bool ok = false;

  bool make() {
    ok = hipEventCreateWithFlags(&fork, hipEventDisableTiming) == hipSuccess;
    for (int i = 0; i < N_SIDE && ok; ++i)
      ok = hipStreamCreateWithFlags(&s[i], hipStreamNonBlocking) == hipSuccess &&
           hipEventCreateWithFlags(&done[i], hipEventDisableTiming) == hipSuccess;
    return ok;
  }
  void destroy() {
    for (int i = 0; i < N_SIDE; ++i) {
      if (s[i]) (void)hipStreamDestroy(s[i]);
      if (done[i]) (void)hipEventDestroy(done[i]);
    }
    if (fork) (void)hipEventDestroy(fork);
  }
};

// The side streams of one caller stream on one device: calls from different
// caller streams never queue behind each other's side-stream work (a small
// request's launch on one stream must not wait for a large one issued from
// another).  Made on first use, kept until the process ends.
SideSet* side_set_for(hipStream_t caller) {
  static std::mutex m;
  static std::map<std::pair<int, hipStream_t>, SideSet*> sets;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> g(m);
  SideSet*& ss = sets[{dev, caller}];
  if (!ss) {
    ss = new SideSet;
    ss->make();
  }
  return ss->ok ? ss : nullptr;
}

// One call's use of a side set: which streams it forked, so that it joins
// exactly those.
struct SideUse {
  SideSet* set = nullptr;
  hipStream_t caller = nullptr;
  uint32_t used = 0;
  bool forked = false;
  uint32_t r0 = 0;

  hipError_t stream_for(uint64_t j, hipStream_t* out) {  // j-th own-sequence item
    *out = caller;
    if (!set) return hipSuccess;
    const uint64_t slot = j % (N_SIDE + 1);
    if (slot == 0) return hipSuccess;
    const uint32_t x = (r0 + (uint32_t)slot - 1) % N_SIDE;
    hipError_t e = hipSuccess;
    if (!forked) {
      e = hipEventRecord(set->fork, caller);
      if (e != hipSuccess) return e;
      forked = true;
    }
    if (!(used & (1u << x))) {
      e = hipStreamWaitEvent(set->s[x], set->fork, 0);
      if (e != hipSuccess) return e;
      used |= 1u << x;
    }
    *out = set->s[x];
    return hipSuccess;
  }
  bool join() {
    bool ok = true;
    for (int x = 0; x < N_SIDE; ++x)
      if (used & (1u << x))
        ok = hipEventRecord(set->done[x], set->s[x]) == hipSuccess &&
             hipStreamWaitEvent(caller, set->done[x], 0) == hipSuccess && ok;
    used = 0;
    return ok;
  }
};

// The own-sequence items, spread over the caller's stream and the side
// streams.  Each item's scratch comes from its piece of `region` when given
// (offsets and sizes from the plan), else from stream-ordered allocations on
// the stream it runs on.
int launch_singles(const kdt_batch_item* items, const Layout& L, SideUse& side, char* region,
                   const std::vector<size_t>* off, const std::vector<size_t>* bytes, std::string& msg) {
  for (size_t j = 0; j < L.single.size(); ++j) {
    const kdt_batch_item& it = items[L.single[j]];
    hipStream_t si = side.caller;
    const hipError_t e = side.stream_for(j, &si);
    if (e != hipSuccess) return fail_msg(msg, std::string("batch: side stream fork: ") + hipGetErrorString(e), KDT_ERR_DEVICE);
    int rc;
    if (region) {
      Scratch sc(region + (*off)[j], (*bytes)[j], si);
      rc = launch_query(it.tree->t, L.single_qp[j], it.queries, it.out_dist, it.out_idx, si, msg, &sc);
    } else {
      rc = launch_query(it.tree->t, L.single_qp[j], it.queries, it.out_dist, it.out_idx, si, msg, nullptr);
    }
    if (rc) return rc;
  }
  return KDT_OK;
}

// The shared class launches on the caller's stream.  `d_items` holds the
// descriptors already (a plan), or null: then they are uploaded here from
// L.h into scratch.
int launch_shared(const Layout& L, const BItem* d_items, Scratch& sc, hipStream_t s, std::string& msg) {
  if (L.h.empty()) return KDT_OK;
  hipError_t e = hipSuccess;
  if (!d_items) {
    BItem* up = nullptr;
    e = sc.get(&up, L.h.size());
    if (e == hipSuccess) e = hipMemcpyAsync(up, L.h.data(), L.h.size() * sizeof(BItem), hipMemcpyHostToDevice, s);
    d_items = up;
  }
  unsigned long long* counters = nullptr;
  if (e == hipSuccess) e = sc.get(&counters, NC);
  uint32_t* order = nullptr;
  const uint32_t n = (uint32_t)L.h.size();
  if (e == hipSuccess && L.order) {
    uint32_t *key = nullptr, *rank = nullptr, *count = nullptr, *first = nullptr;
    e = sc.get(&order, L.total);
    if (e == hipSuccess) e = sc.get(&key, L.total);
    if (e == hipSuccess) e = sc.get(&rank, L.total);
    if (e == hipSuccess) e = sc.get(&count, L.keys);
    if (e == hipSuccess) e = sc.get(&first, L.keys);
    if (e == hipSuccess) {
      launch_zero_u32(count, L.keys, s);
      const uint64_t blocks = std::min<uint64_t>((L.total + 255) / 256, 16384);
      k_batch_key<<<(uint32_t)blocks, 256, 0, s>>>(d_items, n, L.total, count, key, rank);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = exclusive_scan_u32(count, first, L.keys, sc, s);
    if (e == hipSuccess) {
      launch_order_scatter(key, rank, first, L.total, order, s);
      e = hipGetLastError();
    }
  }
  if (e == hipSuccess) {
    static_assert(NC <= 64, "one thread per class counter");
    k_zero<<<1, 64, 0, s>>>(counters, NC);
    const QueryParams& base = L.base;
    for (int c = 0; c < NC; ++c) {
      if (!L.count[c]) continue;
      const uint64_t w = WIDTHS[c / 2];
      const bool screen = (c & 1) != 0;
      // Lanes per query: the widest requested by the class's items (an
      // autotuner's per-tree choice) when it holds every item's k, else from k.
      int G = L.kmax[c] <= 16 ? 16 : L.kmax[c] <= 32 ? 32 : 64;
      const uint64_t gm = L.gmax[c];
      if ((gm == 8 || gm == 16 || gm == 32 || gm == 64) && gm >= L.kmax[c]) G = (int)gm;
      switch (base.metric) {
        case M_L1:
          launch_class_width<M_L1>(w, G, false, d_items, n, order, L.first[c], L.count[c], counters + c, L.stk[c], s);
          break;
        case M_L2:
          launch_class_width<M_L2>(w, G, screen, d_items, n, order, L.first[c], L.count[c], counters + c, L.stk[c], s);
          break;
        default:
          launch_class_width<M_LINF>(w, G, false, d_items, n, order, L.first[c], L.count[c], counters + c, L.stk[c],
                                     s);
          break;
      }
    }
    e = hipGetLastError();
  }
  if (e != hipSuccess) return fail_msg(msg, std::string("batch launch: ") + hipGetErrorString(e), KDT_ERR_DEVICE);
  return KDT_OK;
}

}  // namespace
}  // namespace kdt

// A prepared batch (kdt_batch_plan_create).
struct kdt_batch_plan {
  std::vector<kdt_batch_item> items;
  kdt::Layout L;
  kdt::SideSet side;
  bool spread = false;           // own-sequence items spread over the plan's side streams
  char* mem = nullptr;           // one device region: descriptors, shared scratch, per-item scratch
  size_t shared_off = 0, shared_bytes = 0;
  std::vector<size_t> off, bytes;  // per own-sequence item
  kdt::BItem* d_items = nullptr;
  int device = 0;
};

namespace {
int batch_fail(int code, std::string m) {
  kdt::set_last_error(std::move(m));
  return code;
}
}  // namespace

extern "C" int kdt_query_batch_device(const kdt_batch_item* items, uint64_t n_items, double p,
                                      double max_distance, double eps, uint32_t flags, void* stream) {
  using namespace kdt;
  hipStream_t s = (hipStream_t)stream;
  Layout L;
  std::string msg;
  int rc = classify(items, n_items, p, max_distance, eps, flags, false, L, msg);
  if (rc) return batch_fail(rc, msg);
  // Several own-sequence items are spread over side streams of the caller's
  // stream, unless the call is a pure enqueue (KDT_BATCH_NO_ORDER): then every
  // launch stays on the caller's stream.
  SideUse side;
  side.caller = s;
  std::unique_lock<std::mutex> side_lock;
  if (!L.no_order && L.single.size() > 1) {
    side.set = side_set_for(s);
    if (side.set) {
      side_lock = std::unique_lock<std::mutex>(side.set->lock);
      side.r0 = side.set->rot++;
    }
  }
  rc = launch_singles(items, L, side, nullptr, nullptr, nullptr, msg);
  if (rc == KDT_OK) {
    Scratch sc(s);
    rc = launch_shared(L, nullptr, sc, s, msg);
  }
  const bool joined = side.set ? side.join() : true;
  if (rc) return batch_fail(rc, msg);
  if (!joined) return batch_fail(KDT_ERR_DEVICE, "batch: side stream join failed");
  return KDT_OK;
}

extern "C" int kdt_batch_plan_create(const kdt_batch_item* items, uint64_t n_items, double p, double max_distance,
                                     double eps, uint32_t flags, kdt_batch_plan** out) {
  using namespace kdt;
  if (!out) return batch_fail(KDT_ERR_ARGUMENT, "null output");
  *out = nullptr;
  kdt_batch_plan* P = new kdt_batch_plan();
  std::string msg;
  int rc = classify(items, n_items, p, max_distance, eps, flags, true, P->L, msg);
  if (rc) {
    delete P;
    return batch_fail(rc, msg);
  }
  P->items.assign(items, items + n_items);
  (void)hipGetDevice(&P->device);
  // One region: [descriptors | shared scratch | item 0 scratch | item 1 ...].
  size_t total = scratch_piece(P->L.h.size() * sizeof(BItem));
  P->shared_off = total;
  P->shared_bytes = shared_scratch_bytes(P->L);
  total += P->shared_bytes;
  for (size_t j = 0; j < P->L.single.size(); ++j) {
    const kdt_batch_item& it = items[P->L.single[j]];
    const size_t b = query_scratch_bytes(it.tree->t, P->L.single_qp[j]);
    P->off.push_back(total);
    P->bytes.push_back(b);
    total += scratch_piece(b);
  }
  hipError_t e = hipMalloc((void**)&P->mem, std::max<size_t>(total, 1));
  if (e == hipSuccess && !P->L.h.empty()) {
    P->d_items = reinterpret_cast<BItem*>(P->mem);
    e = hipMemcpy(P->d_items, P->L.h.data(), P->L.h.size() * sizeof(BItem), hipMemcpyHostToDevice);
  }
  if (e == hipSuccess && !P->L.no_order && P->L.single.size() > 1) {
    P->spread = P->side.make();
    if (!P->spread) e = hipErrorOutOfMemory;
  }
  if (e != hipSuccess) {
    kdt_batch_plan_free(P);
    return batch_fail(KDT_ERR_DEVICE, std::string("batch plan: ") + hipGetErrorString(e));
  }
  *out = P;
  return KDT_OK;
}

extern "C" int kdt_batch_plan_run(kdt_batch_plan* plan, void* stream) {
  using namespace kdt;
  if (!plan) return batch_fail(KDT_ERR_ARGUMENT, "null plan");
  hipStream_t s = (hipStream_t)stream;
  SideUse side;
  side.caller = s;
  std::unique_lock<std::mutex> side_lock;
  if (plan->spread) {
    side.set = &plan->side;
    side_lock = std::unique_lock<std::mutex>(plan->side.lock);
  }
  std::string msg;
  int rc = launch_singles(plan->items.data(), plan->L, side, plan->mem, &plan->off, &plan->bytes, msg);
  if (rc == KDT_OK) {
    Scratch sc(plan->mem + plan->shared_off, plan->shared_bytes, s);
    rc = launch_shared(plan->L, plan->d_items, sc, s, msg);
  }
  const bool joined = side.set ? side.join() : true;
  if (rc) return batch_fail(rc, msg);
  if (!joined) return batch_fail(KDT_ERR_DEVICE, "batch plan: side stream join failed");
  return KDT_OK;
}

extern "C" uint64_t kdt_batch_plan_bytes(const kdt_batch_plan* plan) {
  if (!plan) return 0;
  uint64_t total = plan->shared_off + plan->shared_bytes;
  for (size_t b : plan->bytes) total += kdt::scratch_piece(b);
  return total;
}

extern "C" void kdt_batch_plan_free(kdt_batch_plan* plan) {
  if (!plan) return;
  if (plan->mem) {
    (void)hipDeviceSynchronize();  // no replay of the plan's work may still read its region
    (void)hipFree(plan->mem);
  }
  if (plan->spread) plan->side.destroy();
  delete plan;
}
