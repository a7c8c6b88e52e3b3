// kdt_order.hip -- spatial ordering of a query batch.
//
// Results do not depend on the order queries are processed in (each query is
// independent, src/query.rs:93-121), so large batches are processed grouped by
// the leaf each query falls in: queries that share a wavefront then walk the
// same nodes (no divergence between the wave's groups) and neighbouring waves
// reuse the same cache lines.  Key = preorder id of the leaf reached by always
// taking the near child (q[dim] - split <= 0 goes left, src/query.rs:264);
// rocPRIM's radix sort turns the keys into a processing order.  Outputs are
// still written at each query's own row.

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

#include "kdt_internal.h"

namespace kdt {
namespace {

__global__ __launch_bounds__(256) void k_leaf_key(const DevNode* __restrict__ nodes,
                                                  const double* __restrict__ q, uint64_t nq, uint32_t D,
                                                  uint32_t n_nodes, uint32_t* __restrict__ key) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nq;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const double* qi = q + i * D;
    uint32_t node = 0;
    for (;;) {
      const DevNode nd = nodes[node];
      if (nd.b & LEAF_BIT) break;
      const uint32_t dim = nd.b & DIM_MASK;
      node = qi[dim] - nd.split <= 0.0 ? node + 1 : nd.a;  // nodes validated by the build
    }
    key[i] = node;
  }
}

}  // namespace

int order_queries(const Tree& t, const double* d_q, uint64_t nq, uint32_t* d_order, hipStream_t s,
                  std::string& msg) {
  uint32_t *keys = nullptr, *keys_sorted = nullptr;
  void* tmp = nullptr;
  size_t tmp_bytes = 0;
  unsigned end_bit = 1;
  while (end_bit < 32 && (1ull << end_bit) < t.n_nodes) ++end_bit;
  rocprim::counting_iterator<uint32_t> iota(0);
  hipError_t e = rocprim::radix_sort_pairs(nullptr, tmp_bytes, keys, keys_sorted, iota, d_order,
                                           (uint32_t)nq, 0u, end_bit, s);
  if (e == hipSuccess) e = hipMallocAsync((void**)&keys, nq * 4, s);
  if (e == hipSuccess) e = hipMallocAsync((void**)&keys_sorted, nq * 4, s);
  if (e == hipSuccess) e = hipMallocAsync(&tmp, tmp_bytes ? tmp_bytes : 1, s);
  if (e == hipSuccess) {
    const uint64_t blocks = std::min<uint64_t>((nq + 255) / 256, 16384);
    k_leaf_key<<<(uint32_t)blocks, 256, 0, s>>>(t.nodes, d_q, nq, (uint32_t)t.d, (uint32_t)t.n_nodes, keys);
    e = hipGetLastError();
  }
  if (e == hipSuccess)
    e = rocprim::radix_sort_pairs(tmp, tmp_bytes, keys, keys_sorted, iota, d_order, (uint32_t)nq, 0u,
                                  end_bit, s);
  if (keys) (void)hipFreeAsync(keys, s);
  if (keys_sorted) (void)hipFreeAsync(keys_sorted, s);
  if (tmp) (void)hipFreeAsync(tmp, s);
  if (e != hipSuccess) {
    msg = std::string("query ordering: ") + hipGetErrorString(e);
    return 100;
  }
  return 0;
}

}  // namespace kdt
