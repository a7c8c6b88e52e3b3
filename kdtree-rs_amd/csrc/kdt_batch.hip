// kdt_batch.hip -- many trees, many query sets, one persistent launch.
//
// No reference counterpart: the reference answers one query matrix against
// one tree per call (Tree::query, src/query.rs:29-78).  Here a batch of
// (tree, queries, k, outputs) items runs as one launch.  Every query of every
// item gets a global number; one radix sort over the whole batch orders them
// by (item, leaf reached by the near path) so that queries sharing a wavefront
// share a tree and nodes; the groups of a persistent grid then take queries
// from an atomic counter until the batch is drained and run the same per-query
// descent as the single-tree kernels (knn_one), so every result is identical
// to a kdt_query_device call on the same item.

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

#include <algorithm>
#include <vector>

#include "../../include/kdtree_hip.h"
#include "kdt_query_kernels.h"

namespace kdt {
namespace {

constexpr int DMAX = 16;  // widest compiled row width

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
      const uint32_t next = q[dim] - nd.split <= 0.0 ? node + 1 : nd.a;
      if (next <= node || next >= a.n_nodes || dim >= a.D) break;
      node = next;
    }
    key[g] = ((uint64_t)it << leaf_bits) | node;
  }
}

__global__ void k_zero(unsigned long long* p) { *p = 0; }

template <int G, int M, bool SCREEN>
__global__ __launch_bounds__(256) void knn_batch_kernel(const BItem* __restrict__ items, uint32_t n_items,
                                                        uint64_t total, const uint32_t* __restrict__ order,
                                                        unsigned long long* __restrict__ counter) {
  constexpr int GPB = 256 / G;
  __shared__ double s_sa[GPB][STK];
  __shared__ double s_sb[GPB][STK];
  __shared__ uint32_t s_sn[GPB][STK];
  __shared__ uint32_t s_sm[GPB][STK];
  __shared__ double s_ax[GPB][DMAX];
  __shared__ double s_q[GPB][DMAX];
  const uint32_t grp = threadIdx.x / G;
  const uint32_t lane = threadIdx.x & (G - 1);
  RegList<G> best;
  for (;;) {
    unsigned long long mine = 0;
    if (lane == 0) mine = atomicAdd(counter, 1ull);
    const uint64_t gq = (uint64_t)__shfl((long long)mine, 0, G);
    if (gq >= total) break;
    const uint64_t g = order ? (uint64_t)order[gq] : gq;
    const uint32_t it = find_item(items, n_items, g);
    const KArgs a = items[it].a;
    const uint64_t qi = g - items[it].q_begin;
    switch (a.D) {
#define KDT_BATCH_WIDTH(DW)                                                                        \
  case DW:                                                                                         \
    knn_one<DW, G, M, SCREEN, false>(a, qi, best, s_ax[grp], s_q[grp], s_sa[grp], s_sb[grp],      \
                                     s_sn[grp], s_sm[grp]);                                        \
    break;
      KDT_BATCH_WIDTH(1)
      KDT_BATCH_WIDTH(2)
      KDT_BATCH_WIDTH(3)
      KDT_BATCH_WIDTH(4)
      KDT_BATCH_WIDTH(8)
      KDT_BATCH_WIDTH(16)
#undef KDT_BATCH_WIDTH
    }
  }
}

template <int G, int M>
void launch_batch(const BItem* items, uint32_t n, uint64_t total, const uint32_t* order,
                  unsigned long long* counter, bool screen, uint32_t blocks, hipStream_t s) {
  if (screen)
    knn_batch_kernel<G, M, true><<<blocks, 256, 0, s>>>(items, n, total, order, counter);
  else
    knn_batch_kernel<G, M, false><<<blocks, 256, 0, s>>>(items, n, total, order, counter);
}

}  // namespace

int build_kargs(const Tree& t, const QueryParams& qp, const double* d_q, double* d_out, int64_t* d_idx,
                int* d_err, KArgs& a);
int launch_query(const Tree& t, const QueryParams& qp, const double* d_q, double* d_out, int64_t* d_idx,
                 int* d_err, hipStream_t s, std::string& msg);

}  // namespace kdt

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
  if (!items && n_items) return batch_fail(KDT_ERR_ARGUMENT, "null items");
  QueryParams base;
  std::string msg;
  int rc = query_params(p, max_distance, eps, base, msg);
  if (rc) return batch_fail(rc, msg);
  base.fp32_screen = (flags & KDT_QUERY_FP32_SCREEN) != 0;
  const bool per_tree = (flags & KDT_BATCH_PER_TREE_LAUNCHES) != 0;

  std::vector<BItem> h;
  h.reserve(n_items);
  uint64_t total = 0, kmax = 0;
  bool packable = base.metric != M_LP && n_items < (1u << 20);
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
    const uint64_t d = t.d;
    packable &= (d == 1 || d == 2 || d == 3 || d == 4 || d == 8 || d == 16) && it.k <= 64;
    kmax = std::max<uint64_t>(kmax, it.k);
    BItem b{};
    QueryParams qp = base;
    qp.k = it.k;
    qp.nq = it.n_queries;
    build_kargs(t, qp, it.queries, it.out_dist, it.out_idx, t.err, b.a);
    b.q_begin = total;
    total += it.n_queries;
    h.push_back(b);
  }
  if (h.empty()) return KDT_OK;

  if (per_tree || !packable) {
    for (uint64_t i = 0; i < n_items; ++i) {
      const kdt_batch_item& it = items[i];
      if (it.n_queries == 0) continue;
      QueryParams qp = base;
      qp.k = it.k;
      qp.nq = it.n_queries;
      qp.allow_order = !(flags & KDT_BATCH_NO_ORDER);
      rc = launch_query(it.tree->t, qp, it.queries, it.out_dist, it.out_idx, it.tree->t.err, s, msg);
      if (rc) return batch_fail(rc, msg);
    }
    return KDT_OK;
  }

  BItem* d_items = nullptr;
  unsigned long long* counter = nullptr;
  uint32_t* order = nullptr;
  hipError_t e = hipMallocAsync((void**)&d_items, h.size() * sizeof(BItem), s);
  if (e == hipSuccess) e = hipMallocAsync((void**)&counter, sizeof(unsigned long long), s);
  if (e == hipSuccess) e = hipMemcpyAsync(d_items, h.data(), h.size() * sizeof(BItem), hipMemcpyHostToDevice, s);
  const bool ordered = !(flags & KDT_BATCH_NO_ORDER) && total <= 0xFFFFFFFFull && total >= 8192;
  if (e == hipSuccess && ordered) {
    // One sort over the batch: key = (item, leaf of the near path).
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
    k_zero<<<1, 1, 0, s>>>(counter);
    const int G = kmax <= 16 ? 16 : kmax <= 32 ? 32 : 64;
    const uint64_t want = (total + (256 / G) - 1) / (256 / G);
    const uint32_t blocks = (uint32_t)std::min<uint64_t>(want, 256ull * 16);
    const uint32_t n = (uint32_t)h.size();
    const bool screen = base.fp32_screen && base.metric == M_L2;
    auto go = [&](auto g_tag) {
      constexpr int GG = decltype(g_tag)::value;
      switch (base.metric) {
        case M_L1: launch_batch<GG, M_L1>(d_items, n, total, order, counter, false, blocks, s); break;
        case M_L2: launch_batch<GG, M_L2>(d_items, n, total, order, counter, screen, blocks, s); break;
        default: launch_batch<GG, M_LINF>(d_items, n, total, order, counter, false, blocks, s); break;
      }
    };
    if (G == 16) go(std::integral_constant<int, 16>{});
    else if (G == 32) go(std::integral_constant<int, 32>{});
    else go(std::integral_constant<int, 64>{});
    e = hipGetLastError();
  }
  if (order) (void)hipFreeAsync(order, s);
  if (counter) (void)hipFreeAsync(counter, s);
  if (d_items) (void)hipFreeAsync(d_items, s);
  if (e != hipSuccess) return batch_fail(KDT_ERR_DEVICE, std::string("batch launch: ") + hipGetErrorString(e));
  return KDT_OK;
}
