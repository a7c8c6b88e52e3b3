// kdt_order.hip -- spatial ordering of a query batch, and the prefix sum it uses.
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
//
// Every step is a plain kernel launch on the call's stream (the counts are
// cleared by a kernel, the scan is this file's own), so the sequence records
// into a HIP graph with no memset or library node.

#include "kdt_internal.h"

namespace kdt {
namespace {

constexpr uint32_t SCAN_THREADS = 256;
constexpr uint32_t SCAN_PER_THREAD = 8;
constexpr uint32_t SCAN_TILE = SCAN_THREADS * SCAN_PER_THREAD;  // counts per block

__global__ void k_zero_counts(uint32_t* __restrict__ p, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

// Block-wide exclusive scan of one value per thread (256 threads, 4 waves):
// wave scans by shuffles, then the wave totals through LDS.  Returns the
// exclusive prefix; *total gets the block's sum.
__device__ __forceinline__ uint32_t block_exclusive(uint32_t v, uint32_t* s_w, uint32_t* total) {
  const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t incl = v;
#pragma unroll
  for (uint32_t o = 1; o < 64; o <<= 1) {
    const uint32_t t = __shfl_up(incl, o);
    if (lane >= o) incl += t;
  }
  if (lane == 63) s_w[w] = incl;
  __syncthreads();
  uint32_t before = 0, all = 0;
#pragma unroll
  for (uint32_t i = 0; i < SCAN_THREADS / 64; ++i) {
    before += i < w ? s_w[i] : 0u;
    all += s_w[i];
  }
  __syncthreads();  // s_w may be reused by the caller's next round
  *total = all;
  return before + incl - v;
}

// Phase 1: each block scans SCAN_TILE counts (SCAN_PER_THREAD consecutive
// values per thread) into `first` and writes its total to bsum[block].
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_tiles(const uint32_t* __restrict__ count,
                                                             uint32_t* __restrict__ first, uint64_t n,
                                                             uint32_t* __restrict__ bsum) {
  __shared__ uint32_t s_w[SCAN_THREADS / 64];
  const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE + (uint64_t)threadIdx.x * SCAN_PER_THREAD;
  uint32_t v[SCAN_PER_THREAD];
  uint32_t sum = 0;
#pragma unroll
  for (uint32_t i = 0; i < SCAN_PER_THREAD; ++i) {
    v[i] = base + i < n ? count[base + i] : 0u;
    sum += v[i];
  }
  uint32_t total;
  uint32_t run = block_exclusive(sum, s_w, &total);
#pragma unroll
  for (uint32_t i = 0; i < SCAN_PER_THREAD; ++i) {
    const uint32_t c = v[i];
    if (base + i < n) first[base + i] = run;
    run += c;
  }
  if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

// Phase 2: one block scans the nb block totals in place (rounds of 256).
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_sums(uint32_t* __restrict__ bsum, uint32_t nb) {
  __shared__ uint32_t s_w[SCAN_THREADS / 64];
  uint32_t carry = 0;
  for (uint32_t r = 0; r < nb; r += SCAN_THREADS) {
    const uint32_t i = r + threadIdx.x;
    const uint32_t v = i < nb ? bsum[i] : 0u;
    uint32_t total;
    const uint32_t ex = block_exclusive(v, s_w, &total);
    if (i < nb) bsum[i] = carry + ex;
    carry += total;
  }
}

// Phase 3: add each block's offset.
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_add(uint32_t* __restrict__ first, uint64_t n,
                                                           const uint32_t* __restrict__ bsum) {
  const uint32_t off = bsum[blockIdx.x];
  const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE + (uint64_t)threadIdx.x * SCAN_PER_THREAD;
#pragma unroll
  for (uint32_t i = 0; i < SCAN_PER_THREAD; ++i)
    if (base + i < n) first[base + i] += off;
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

uint64_t scan_blocks(uint64_t n) { return (n + SCAN_TILE - 1) / SCAN_TILE; }

}  // namespace

void launch_zero_u32(uint32_t* p, uint64_t n, hipStream_t s) {
  if (n) k_zero_counts<<<(uint32_t)std::min<uint64_t>((n + 255) / 256, 4096), 256, 0, s>>>(p, n);
}

void launch_order_scatter(const uint32_t* key, const uint32_t* rank, const uint32_t* first, uint64_t nq,
                          uint32_t* order, hipStream_t s) {
  if (nq)
    k_order_scatter<<<(uint32_t)std::min<uint64_t>((nq + 255) / 256, 16384), 256, 0, s>>>(key, rank, first, nq,
                                                                                          order);
}

size_t scan_scratch_bytes(uint64_t n) { return scratch_piece(scan_blocks(n) * 4); }

hipError_t exclusive_scan_u32(const uint32_t* count, uint32_t* first, uint64_t n, Scratch& sc, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const uint64_t nb = scan_blocks(n);
  if (nb > 0xFFFFFFFFull) return hipErrorInvalidValue;
  uint32_t* bsum = nullptr;
  hipError_t e = sc.get(&bsum, nb);
  if (e != hipSuccess) return e;
  k_scan_tiles<<<(uint32_t)nb, SCAN_THREADS, 0, s>>>(count, first, n, bsum);
  k_scan_sums<<<1, SCAN_THREADS, 0, s>>>(bsum, (uint32_t)nb);
  k_scan_add<<<(uint32_t)nb, SCAN_THREADS, 0, s>>>(first, n, bsum);
  return hipGetLastError();
}

size_t order_scratch_bytes(const Tree& t, uint64_t nq) {
  return 2 * scratch_piece(t.n_nodes * 4) + 2 * scratch_piece(nq * 4) + scan_scratch_bytes(t.n_nodes);
}

int order_queries(const Tree& t, const double* d_q, uint64_t nq, uint32_t* d_order, Scratch& sc, hipStream_t s,
                  std::string& msg) {
  const uint64_t nn = t.n_nodes;
  uint32_t *count = nullptr, *first = nullptr, *key = nullptr, *rank = nullptr;
  hipError_t e = sc.get(&count, nn);
  if (e == hipSuccess) e = sc.get(&first, nn);
  if (e == hipSuccess) e = sc.get(&key, nq);
  if (e == hipSuccess) e = sc.get(&rank, nq);
  if (e == hipSuccess) {
    launch_zero_u32(count, nn, s);
    const uint32_t blocks = (uint32_t)std::min<uint64_t>((nq + 255) / 256, 16384);
    k_leaf_key<<<blocks, 256, 0, s>>>(t.nodes, d_q, nq, (uint32_t)t.d, count, key, rank);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = exclusive_scan_u32(count, first, nn, sc, s);
  if (e == hipSuccess) {
    launch_order_scatter(key, rank, first, nq, d_order, s);
    e = hipGetLastError();
  }
  if (e != hipSuccess) {
    msg = std::string("query ordering: ") + hipGetErrorString(e);
    return 100;
  }
  return 0;
}

}  // namespace kdt
