// kdt_batch.hip -- many trees, many query sets, few launches.
//
// No reference counterpart: the reference answers one query matrix against
// one tree per call (Tree::query, src/query.rs:29-78).  Here a batch of
// (tree, queries, k, outputs) items runs in one persistent launch per row
// width.  Every query of those items gets a global number; one radix sort over
// the batch orders them by (item, leaf reached by the near path) so that
// queries sharing a wavefront share a tree and nodes; the groups of a
// persistent grid then take queries from an atomic counter until the class is
// drained and run the same per-query descent as the single-tree kernels
// (knn_ww).  Large items, the matrix-core screen and the cases without a
// compiled width run their own launch sequence, spread over the caller's
// stream and side streams forked from it.  Every result is identical to a
// kdt_query_device call on the same item.

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

#include <algorithm>
#include <map>
#include <mutex>
#include <vector>

#include "../../include/kdtree_hip.h"
#include "kdt_query_kernels.h"

namespace kdt {
namespace {

struct BItem {
  KArgs a;
  uint64_t q_begin;  // first global query number of the item
};

__device__ __forceinline__ uint32_t find_item(const BItem* __restrict__ items, uint32_t n, uint64_t g) {
  uint32_t lo = 0, hi = n;  // last item with q_begin <= g
  while (hi - lo > 1) {
    const uint32_t m = (lo + hi) / 2;
    if (items[m].q_begin <= g) lo = m; else hi = m;
  }
  return lo;
}

__global__ __launch_bounds__(256) void k_batch_key(const BItem* __restrict__ items, uint32_t n_items,
                                                   uint64_t total, uint32_t leaf_bits,
                                                   uint64_t* __restrict__ key) {
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
      node = q[dim] - nd.split <= 0.0 ? node + 1 : nd.a;
    }
    key[g] = ((uint64_t)it << leaf_bits) | node;
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
  knn_ww<DT, G, M, SCREEN>(cur, best, s_ax[grp], s_q[grp], s_sa[grp], s_sb[grp], s_sn[grp], s_sm[grp], next);
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
  if (stk <= 32) launch_class_stk<DT, M, 32>(G, screen, blocks, items, n, order, first, count, counter, s);
  else launch_class_stk<DT, M, STK_MAX>(G, screen, blocks, items, n, order, first, count, counter, s);
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

}  // namespace

int build_kargs(const Tree& t, const QueryParams& qp, const double* d_q, double* d_out, int64_t* d_idx,
                KArgs& a);

}  // namespace kdt

namespace {
int batch_fail(int code, std::string m) {
  kdt::set_last_error(std::move(m));
  return code;
}
}  // namespace

namespace kdt {
namespace {
constexpr uint64_t SINGLE_MIN = 8192;

// An item's own screen flags, on top of the call's.
void item_flags(const kdt_batch_item& it, QueryParams& qp) {
  qp.fp32_screen = qp.fp32_screen || (it.flags & KDT_QUERY_FP32_SCREEN) != 0;
  qp.matrix_screen = qp.matrix_screen || (it.flags & KDT_QUERY_MATRIX_SCREEN) != 0;
}

// Side streams for the items that run their own launch sequence: one item of
// 8192 queries fills about a quarter of the device, so consecutive such items
// on one stream leave it idle.  The call forks the side streams from the
// caller's stream (an event), spreads those items over the caller's stream and
// them, largest first, and joins them back (one event each) after the shared
// launches, which stay on the caller's stream and overlap with them.  Created
// once per device, default priority; `lock` keeps one call at a time on a
// device's side streams and events, and `rot` varies which side stream a
// call's second item takes, so that calls from several caller streams spread.
constexpr int N_SIDE = 4;
struct SideStreams {
  hipStream_t s[N_SIDE] = {};
  hipEvent_t fork = nullptr;
  hipEvent_t done[N_SIDE] = {};
  std::mutex lock;
  uint32_t rot = 0;
  bool ok = false;
};

SideStreams* side_streams() {
  static std::mutex m;
  static std::map<int, SideStreams*> per_device;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> g(m);
  SideStreams*& ss = per_device[dev];
  if (!ss) {
    ss = new SideStreams;  // kept until the process ends, like the device context
    bool ok = hipEventCreateWithFlags(&ss->fork, hipEventDisableTiming) == hipSuccess;
    for (int i = 0; i < N_SIDE && ok; ++i)
      ok = hipStreamCreateWithFlags(&ss->s[i], hipStreamNonBlocking) == hipSuccess &&
           hipEventCreateWithFlags(&ss->done[i], hipEventDisableTiming) == hipSuccess;
    ss->ok = ok;
  }
  return ss->ok ? ss : nullptr;
}
}  // namespace
}  // namespace kdt

extern "C" int kdt_query_batch_device(const kdt_batch_item* items, uint64_t n_items, double p,
                                      double max_distance, double eps, uint32_t flags, void* stream) {
  using namespace kdt;
  hipStream_t s = (hipStream_t)stream;
  if (!items && n_items) return batch_fail(KDT_ERR_ARGUMENT, "null items");
  QueryParams base;
  std::string msg;
  int rc = query_params(p, max_distance, eps, base, msg);
  if (rc) return batch_fail(rc, msg);
  base.fp32_screen = (flags & KDT_QUERY_FP32_SCREEN) != 0;
  base.matrix_screen = (flags & KDT_QUERY_MATRIX_SCREEN) != 0;
  base.group = requested_group(flags);
  const bool per_tree = (flags & KDT_BATCH_PER_TREE_LAUNCHES) != 0;
  const bool no_order = (flags & KDT_BATCH_NO_ORDER) != 0;
  // KDT_BATCH_NO_ORDER makes the call a pure enqueue (capturable in a graph):
  // no leaf-order sort and no matrix-core screen, both of which take
  // stream-ordered scratch whose lifetime ends inside the call.  Items that
  // would take the screen run the traversal, with the same results.
  base.allow_scratch = !no_order;

  // Validate; split items into width classes that run in shared launches and
  // the rest (general p, other widths, k > 64, the matrix-core screen, items of
  // SINGLE_MIN queries or more), which run one launch sequence each.
  static const uint64_t widths[] = {1, 2, 3, 4, 8, 16};
  std::vector<uint64_t> cls_items[6];
  std::vector<uint64_t> single;
  for (uint64_t i = 0; i < n_items; ++i) {
    const kdt_batch_item& it = items[i];
    if (!it.tree) return batch_fail(KDT_ERR_ARGUMENT, "null tree in batch item " + std::to_string(i));
    if (it.k == 0) return batch_fail(KDT_ERR_INVALID_K, "k must be greater than zero");
    if (it.n_queries == 0) continue;
    if (!it.queries || !it.out_dist || !it.out_idx) return batch_fail(KDT_ERR_ARGUMENT, "null buffer");
    const Tree& t = it.tree->t;
    if (it.ndim != t.d)
      return batch_fail(KDT_ERR_DIMENSION_MISMATCH, "dimension mismatch: expected " + std::to_string(t.d) +
                                                        ", got " + std::to_string(it.ndim));
    int c = -1;
    for (int w = 0; w < 6; ++w)
      if (t.d == widths[w]) c = w;
    QueryParams probe = base;
    probe.k = it.k;
    probe.nq = it.n_queries;
    item_flags(it, probe);
    // Items of SINGLE_MIN queries or more take their own launch sequence, as
    // alone (own lanes per query, own leaf ordering, uniform kernel arguments):
    // their launch overhead is amortised, and in the shared launch they ran
    // slower.  The shared persistent launches keep the many small items.
    if (per_tree || base.metric == M_LP || c < 0 || it.k > 64 || it.n_queries >= SINGLE_MIN ||
        screen_applies(t, probe))
      single.push_back(i);
    else cls_items[c].push_back(i);
  }

  // Several own-sequence items are spread over the caller's stream and the
  // side streams, unless the call is a pure enqueue (KDT_BATCH_NO_ORDER, e.g.
  // under graph capture): then every launch stays on the caller's stream.
  SideStreams* side = (!no_order && single.size() > 1) ? side_streams() : nullptr;
  std::unique_lock<std::mutex> side_lock;
  if (side) {
    side_lock = std::unique_lock<std::mutex>(side->lock);
    bool ok = hipEventRecord(side->fork, s) == hipSuccess;
    for (int j = 0; j < N_SIDE && ok; ++j) ok = hipStreamWaitEvent(side->s[j], side->fork, 0) == hipSuccess;
    if (!ok) return batch_fail(KDT_ERR_DEVICE, "batch: side stream fork failed");
  }
  // Joins the side streams back into the caller's stream (also on errors, so
  // that nothing of this call is left unordered behind the caller's stream).
  auto join = [&]() -> bool {
    if (!side) return true;
    bool ok = true;
    for (int j = 0; j < N_SIDE; ++j)
      ok = hipEventRecord(side->done[j], side->s[j]) == hipSuccess &&
           hipStreamWaitEvent(s, side->done[j], 0) == hipSuccess && ok;
    return ok;
  };
  uint32_t r0 = 0;
  if (side) {
    r0 = side->rot++;
    std::stable_sort(single.begin(), single.end(),
                     [&](uint64_t x, uint64_t y) { return items[x].n_queries > items[y].n_queries; });
  }
  uint64_t next_side = 0;
  for (uint64_t i : single) {
    const kdt_batch_item& it = items[i];
    QueryParams qp = base;
    qp.k = it.k;
    qp.nq = it.n_queries;
    item_flags(it, qp);
    qp.allow_order = !no_order;
    if (it.group) qp.group = (int)it.group;
    hipStream_t si = s;
    if (side) {
      const uint64_t j = next_side++ % (N_SIDE + 1);
      if (j) si = side->s[(r0 + j - 1) % N_SIDE];
    }
    rc = launch_query(it.tree->t, qp, it.queries, it.out_dist, it.out_idx, si, msg);
    if (rc) {
      join();
      return batch_fail(rc, msg);
    }
  }

  // Shared launches: items numbered class by class, so each class is one
  // contiguous range of the batch's query numbers (and of the sorted order).
  constexpr int NC = 6;
  std::vector<BItem> h;
  uint64_t first[NC] = {0}, count[NC] = {0}, kmax[NC] = {0}, gmax[NC] = {0};
  uint32_t stk[NC] = {2, 2, 2, 2, 2, 2};
  uint64_t total = 0;
  for (int c = 0; c < NC; ++c) {
    first[c] = total;
    for (uint64_t i : cls_items[c]) {
      const kdt_batch_item& it = items[i];
      BItem b{};
      QueryParams qp = base;
      qp.k = it.k;
      qp.nq = it.n_queries;
      build_kargs(it.tree->t, qp, it.queries, it.out_dist, it.out_idx, b.a);
      b.q_begin = total;
      total += it.n_queries;
      kmax[c] = std::max<uint64_t>(kmax[c], it.k);
      gmax[c] = std::max<uint64_t>(gmax[c], it.group ? it.group : (uint64_t)base.group);
      stk[c] = std::max<uint32_t>(stk[c], b.a.stk);
      h.push_back(b);
    }
    count[c] = total - first[c];
  }
  if (h.empty()) return join() ? KDT_OK : batch_fail(KDT_ERR_DEVICE, "batch: side stream join failed");
  if (total > 0xFFFFFFFFull) {
    join();
    return batch_fail(KDT_ERR_ARGUMENT, "batch exceeds 2^32 queries");
  }

  BItem* d_items = nullptr;
  unsigned long long* counters = nullptr;
  uint32_t* order = nullptr;
  hipError_t e = hipMallocAsync((void**)&d_items, h.size() * sizeof(BItem), s);
  if (e == hipSuccess) e = hipMallocAsync((void**)&counters, NC * sizeof(unsigned long long), s);
  if (e == hipSuccess) e = hipMemcpyAsync(d_items, h.data(), h.size() * sizeof(BItem), hipMemcpyHostToDevice, s);
  if (e == hipSuccess && !no_order && total >= 8192) {
    // One sort over the batch: key = (item, leaf reached by the near path).
    uint32_t leaf_bits = 1;
    uint64_t max_nodes = 1;
    for (const BItem& b : h) max_nodes = std::max<uint64_t>(max_nodes, b.a.n_nodes);
    while (leaf_bits < 32 && (1ull << leaf_bits) < max_nodes) ++leaf_bits;
    uint32_t item_bits = 1;
    while ((1ull << item_bits) < h.size()) ++item_bits;
    uint64_t *keys = nullptr, *keys2 = nullptr;
    void* tmp = nullptr;
    size_t tmp_bytes = 0;
    rocprim::counting_iterator<uint32_t> iota(0);
    e = rocprim::radix_sort_pairs(nullptr, tmp_bytes, keys, keys2, iota, order, (uint32_t)total, 0u,
                                  leaf_bits + item_bits, s);
    if (e == hipSuccess) e = hipMallocAsync((void**)&order, total * 4, s);
    if (e == hipSuccess) e = hipMallocAsync((void**)&keys, total * 8, s);
    if (e == hipSuccess) e = hipMallocAsync((void**)&keys2, total * 8, s);
    if (e == hipSuccess) e = hipMallocAsync(&tmp, tmp_bytes ? tmp_bytes : 1, s);
    if (e == hipSuccess) {
      const uint64_t blocks = std::min<uint64_t>((total + 255) / 256, 16384);
      k_batch_key<<<(uint32_t)blocks, 256, 0, s>>>(d_items, (uint32_t)h.size(), total, leaf_bits, keys);
      e = rocprim::radix_sort_pairs(tmp, tmp_bytes, keys, keys2, iota, order, (uint32_t)total, 0u,
                                    leaf_bits + item_bits, s);
    }
    if (keys) (void)hipFreeAsync(keys, s);
    if (keys2) (void)hipFreeAsync(keys2, s);
    if (tmp) (void)hipFreeAsync(tmp, s);
  }
  if (e == hipSuccess) {
    static_assert(NC <= 64, "one thread per class counter");
    k_zero<<<1, 64, 0, s>>>(counters, NC);
    const bool screen = base.fp32_screen && base.metric == M_L2;
    const uint32_t n = (uint32_t)h.size();
    for (int c = 0; c < NC; ++c) {
      if (!count[c]) continue;
      const uint64_t w = widths[c];
      // Lanes per query: the widest requested by the class's items (an
      // autotuner's per-tree choice) when it holds every item's k, else from k.
      int G = kmax[c] <= 16 ? 16 : kmax[c] <= 32 ? 32 : 64;
      if ((gmax[c] == 8 || gmax[c] == 16 || gmax[c] == 32 || gmax[c] == 64) && gmax[c] >= kmax[c]) G = (int)gmax[c];
      switch (base.metric) {
        case M_L1:
          launch_class_width<M_L1>(w, G, false, d_items, n, order, first[c], count[c], counters + c, stk[c], s);
          break;
        case M_L2:
          launch_class_width<M_L2>(w, G, screen, d_items, n, order, first[c], count[c], counters + c, stk[c], s);
          break;
        default:
          launch_class_width<M_LINF>(w, G, false, d_items, n, order, first[c], count[c], counters + c, stk[c], s);
          break;
      }
    }
    e = hipGetLastError();
  }
  if (order) (void)hipFreeAsync(order, s);
  if (counters) (void)hipFreeAsync(counters, s);
  if (d_items) (void)hipFreeAsync(d_items, s);
  const bool joined = join();
  if (e != hipSuccess) return batch_fail(KDT_ERR_DEVICE, std::string("batch launch: ") + hipGetErrorString(e));
  if (!joined) return batch_fail(KDT_ERR_DEVICE, "batch: side stream join failed");
  return KDT_OK;
}
