// kdt_order.hip -- spatial ordering of a query batch.
//
// Results do not depend on the order queries are processed in (each query is
// independent, src/query.rs:93-121), so large batches are processed grouped by
// the leaf each query falls in: queries that share a wavefront then walk the
// same nodes (no divergence between the wave's groups) and neighbouring waves
// reuse the same cache lines.  Key = preorder id of the leaf reached by always
// taking the near child (q[dim] - split <= 0 goes left, src/query.rs:264).
// A counting sort turns the keys into a processing order: the key kernel also
// counts queries per leaf (its atomic returns each query's rank in its leaf),
// an exclusive scan of the counts gives each leaf's first slot, and a scatter
// writes the order.  The order within a leaf is whatever the atomics gave,
// which no result can see.  Outputs are still written at each query's own row.

#include <rocprim/device/device_scan.hpp>

#include "kdt_internal.h"

namespace kdt {
namespace {

__global__ void k_zero_counts(uint32_t* __restrict__ p, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

__global__ __launch_bounds__(256) void k_leaf_key(const DevNode* __restrict__ nodes,
                                                  const double* __restrict__ q, uint64_t nq, uint32_t D,
                                                  uint32_t* __restrict__ count, uint32_t* __restrict__ key,
                                                  uint32_t* __restrict__ rank) {
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
    rank[i] = atomicAdd(&count[node], 1u);
  }
}

__global__ __launch_bounds__(256) void k_order_scatter(const uint32_t* __restrict__ key,
                                                       const uint32_t* __restrict__ rank,
                                                       const uint32_t* __restrict__ first, uint64_t nq,
                                                       uint32_t* __restrict__ order) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nq;
       i += (uint64_t)gridDim.x * blockDim.x)
    order[first[key[i]] + rank[i]] = (uint32_t)i;
}

}  // namespace

int order_queries(const Tree& t, const double* d_q, uint64_t nq, uint32_t* d_order, hipStream_t s,
                  std::string& msg) {
  const uint64_t nn = t.n_nodes;
  uint32_t *count = nullptr, *first = nullptr, *key = nullptr, *rank = nullptr;
  void* tmp = nullptr;
  size_t tmp_bytes = 0;
  hipError_t e = rocprim::exclusive_scan(nullptr, tmp_bytes, count, first, 0u, (size_t)nn,
                                         rocprim::plus<uint32_t>(), s);
  if (e == hipSuccess) e = hipMallocAsync((void**)&count, nn * 4, s);
  if (e == hipSuccess) e = hipMallocAsync((void**)&first, nn * 4, s);
  if (e == hipSuccess) e = hipMallocAsync((void**)&key, nq * 4, s);
  if (e == hipSuccess) e = hipMallocAsync((void**)&rank, nq * 4, s);
  if (e == hipSuccess) e = hipMallocAsync(&tmp, tmp_bytes ? tmp_bytes : 1, s);
  if (e == hipSuccess) {
    k_zero_counts<<<(uint32_t)std::min<uint64_t>((nn + 255) / 256, 4096), 256, 0, s>>>(count, nn);
    const uint32_t blocks = (uint32_t)std::min<uint64_t>((nq + 255) / 256, 16384);
    k_leaf_key<<<blocks, 256, 0, s>>>(t.nodes, d_q, nq, (uint32_t)t.d, count, key, rank);
    e = hipGetLastError();
  }
  if (e == hipSuccess)
    e = rocprim::exclusive_scan(tmp, tmp_bytes, count, first, 0u, (size_t)nn, rocprim::plus<uint32_t>(), s);
  if (e == hipSuccess) {
    const uint32_t blocks = (uint32_t)std::min<uint64_t>((nq + 255) / 256, 16384);
    k_order_scatter<<<blocks, 256, 0, s>>>(key, rank, first, nq, d_order);
    e = hipGetLastError();
  }
  for (void* p : {(void*)count, (void*)first, (void*)key, (void*)rank, tmp})
    if (p) (void)hipFreeAsync(p, s);
  if (e != hipSuccess) {
    msg = std::string("query ordering: ") + hipGetErrorString(e);
    return 100;
  }
  return 0;
}

}  // namespace kdt
