// kdt_build.hip -- level-synchronous median-split build on gfx950.
//
// Replaces build::build / build_subtree / partition_rows (src/build.rs:19-314)
// and the bbox kernel (src/kernel.rs:93-151).  The reference recurses node by
// node; here every node of one tree level is processed by the same launches.
// The tree shape is a pure function of (n_points, leafsize) (src/build.rs:1-5,
// :161-173), so the host lays out each level's nodes (id, first row, length)
// and the chunk table that tiles their rows before any data is touched.
//
// Per level, for every node of the level:
//   1. tight box over its rows (src/build.rs:175-181); split_dim = first strict
//      widest axis (src/layout.rs:168-179)                 box_partial/box_final
//   2. pivot = the (len/2)-th smallest split coordinate under f64::total_cmp
//      (src/build.rs:197-205): radix select on the 64-bit total-order key, in
//      LDS for nodes up to SEL_LOCAL rows, else 8 histogram passes over HBM
//   3. lt = #{key < pivot} (src/build.rs:209)
//   4. partition_rows (src/build.rs:266-314), reproduced exactly:
//      A = rows [0,lt) with key >= pivot in ascending order, B = rows [lt,len)
//      with key < pivot ascending; A[i] swaps with B[m-1-i].  The pairs are
//      disjoint, so all swaps run in parallel.  The fix-up loop that pulls
//      pivot-equal rows into [lt,mid) (src/build.rs:303-313) is not a stable
//      partition: its t-th swap exchanges position lt+t-1 with s_t, the t-th
//      pivot-equal position >= lt, so pivot-equal rows land in [lt,mid) in
//      order, rows already in [mid, s_need] stay, and a row at P in [lt,mid)
//      is carried along P -> g(P) -> g(g(P)) ... (g(lt+t-1) = s_t) to the
//      first position >= mid: the rows the loop passes over rotate.  g is
//      increasing, so the chains are resolved in parallel by pointer jumping
//      (k_fix_* below).
// Rows move physically (with their original index), as in the reference, so
// every node's rows stay one contiguous row-major block.
//
// Chunk kernels use one 256-thread workgroup per chunk of at most CH rows of
// one node; counts and offsets per chunk are scanned per node in between.

#include <algorithm>
#include <cstring>
#include <map>
#include <vector>

#include "kdt_internal.h"

namespace kdt {
namespace {

constexpr int BT = 256;           // threads per workgroup
constexpr uint32_t CH = 2048;     // rows per chunk
constexpr uint32_t SEL_LOCAL = 4096;  // nodes up to this many rows select in LDS
constexpr double BOX_ORDER_MAX_SHRINK_RATIO = 0.6;  // src/build.rs:13
constexpr uint32_t BOX_ORDER_MIN_SUBTREE = 64;      // src/build.rs:17

struct LvlNode {
  uint32_t id, start, len, right;
  uint32_t cbeg, cend, pad0, pad1;
};
struct Chunk {
  uint32_t slot, begin, end, pad;
};
struct Sel {             // per node of the current level
  uint64_t prefix, mask; // radix-select state
  uint32_t rank, dim;
  uint32_t lt, m;        // keys below pivot; misplaced pairs
  uint32_t last;         // fix-up window end (node-relative)
  uint32_t pad;
  double pivot;
};

__device__ __forceinline__ uint64_t total_key(double x) {
  // f64::total_cmp order as an unsigned key: flip all bits of negatives,
  // set the sign bit of non-negatives.
  uint64_t u = (uint64_t)__double_as_longlong(x);
  return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
__device__ __forceinline__ double from_total_key(uint64_t k) {
  uint64_t u = (k >> 63) ? (k & 0x7FFFFFFFFFFFFFFFull) : ~k;
  return __longlong_as_double((long long)u);
}

__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Exclusive prefix count of `flag` over the 256 threads of the block, in
// thread order; `total` receives the block count.  Uses s[0..4).
__device__ __forceinline__ uint32_t block_excl_count(bool flag, uint32_t* s, uint32_t& total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t b = __ballot(flag);
  const uint32_t within = (uint32_t)__popcll(b & ((1ull << lane) - 1ull));
  __syncthreads();
  if (lane == 0) s[wave] = (uint32_t)__popcll(b);
  __syncthreads();
  uint32_t off = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < BT / 64; ++w) {
    const uint32_t c = s[w];
    off += w < wave ? c : 0u;
    tot += c;
  }
  total = tot;
  return off + within;
}

__device__ __forceinline__ uint32_t block_sum(uint32_t v, uint32_t* s) {
  v = wave_sum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = v;
  __syncthreads();
  uint32_t t = 0;
#pragma unroll
  for (int w = 0; w < BT / 64; ++w) t += s[w];
  return t;
}

__global__ void k_check_finite(const double* __restrict__ x, uint64_t n, int* flag) {
  bool bad = false;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    bad |= !isfinite(x[i]);
  if (__ballot(bad) && (threadIdx.x & 63) == 0) atomicOr(flag, 1);
}

__global__ void k_iota(uint32_t* idx, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    idx[i] = (uint32_t)i;
}

// Per-chunk partial box, 8 dimensions per pass.  part[chunk][0|1][d].
__global__ __launch_bounds__(BT) void k_box_partial(const double* __restrict__ rows, uint32_t D,
                                                    const Chunk* __restrict__ chunks,
                                                    double* __restrict__ part) {
  __shared__ double s_lo[BT / 64][8], s_hi[BT / 64][8];
  const Chunk c = chunks[blockIdx.x];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double* out = part + (size_t)blockIdx.x * 2 * D;
  for (uint32_t db = 0; db < D; db += 8) {
    double lo[8], hi[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) { lo[j] = INFINITY; hi[j] = -INFINITY; }
    for (uint32_t r = c.begin + threadIdx.x; r < c.end; r += BT) {
      const double* row = rows + (size_t)r * D + db;
#pragma unroll
      for (int j = 0; j < 8; ++j)
        if (db + j < D) {
          const double v = row[j];
          lo[j] = fmin(lo[j], v);
          hi[j] = fmax(hi[j], v);
        }
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) { lo[j] = wave_min(lo[j]); hi[j] = wave_max(hi[j]); }
    __syncthreads();
    if (lane == 0)
#pragma unroll
      for (int j = 0; j < 8; ++j) { s_lo[wave][j] = lo[j]; s_hi[wave][j] = hi[j]; }
    __syncthreads();
    if (threadIdx.x < 8 && db + threadIdx.x < D) {
      double l = s_lo[0][threadIdx.x], h = s_hi[0][threadIdx.x];
      for (int w = 1; w < BT / 64; ++w) { l = fmin(l, s_lo[w][threadIdx.x]); h = fmax(h, s_hi[w][threadIdx.x]); }
      out[db + threadIdx.x] = l;
      out[D + db + threadIdx.x] = h;
    }
  }
}

// Final box of each node from its chunk partials; widest axis for inner nodes.
// One wave per node.
__global__ __launch_bounds__(64) void k_box_final(const double* __restrict__ part,
                                                  const LvlNode* __restrict__ lvl, uint32_t D,
                                                  uint32_t leafsize, double* __restrict__ boxes,
                                                  Sel* __restrict__ sel) {
  const LvlNode nd = lvl[blockIdx.x];
  const uint32_t lane = threadIdx.x;
  double best_span = -INFINITY;
  uint32_t best_dim = 0xFFFFFFFFu;
  double* lo_out = boxes + (size_t)nd.id * 2 * D;
  for (uint32_t c = lane; c < D; c += 64) {
    double lo = INFINITY, hi = -INFINITY;
    for (uint32_t ch = nd.cbeg; ch < nd.cend; ++ch) {
      lo = fmin(lo, part[(size_t)ch * 2 * D + c]);
      hi = fmax(hi, part[(size_t)ch * 2 * D + D + c]);
    }
    lo_out[c] = lo;
    lo_out[D + c] = hi;
    const double span = hi - lo;  // BBox::extent, src/layout.rs:164-166
    if (best_dim == 0xFFFFFFFFu || span > best_span) { best_span = span; best_dim = c; }
  }
  // First strict maximum over dimensions (src/layout.rs:168-179): the largest
  // span wins, equal spans go to the lower dimension.
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const double os = __shfl_xor(best_span, o);
    const uint32_t od = __shfl_xor(best_dim, o);
    const bool take = od != 0xFFFFFFFFu &&
                      (best_dim == 0xFFFFFFFFu || os > best_span || (os == best_span && od < best_dim));
    if (take) { best_span = os; best_dim = od; }
  }
  if (lane == 0 && nd.len > leafsize) {
    Sel s{};
    s.dim = best_dim;
    s.rank = nd.len / 2;  // split_point, src/build.rs:161-164
    sel[blockIdx.x] = s;
  }
}

__device__ __forceinline__ bool is_inner(const LvlNode& nd, uint32_t leafsize) { return nd.len > leafsize; }

__global__ __launch_bounds__(BT) void k_keys(const double* __restrict__ rows, uint32_t D,
                                             const Chunk* __restrict__ chunks,
                                             const LvlNode* __restrict__ lvl, uint32_t leafsize,
                                             const Sel* __restrict__ sel, uint64_t* __restrict__ ukey) {
  const Chunk c = chunks[blockIdx.x];
  if (!is_inner(lvl[c.slot], leafsize)) return;
  const uint32_t dim = sel[c.slot].dim;
  for (uint32_t r = c.begin + threadIdx.x; r < c.end; r += BT) ukey[r] = total_key(rows[(size_t)r * D + dim]);
}

// Radix select of the rank-th smallest key of one node entirely in LDS, plus
// the lt count.  One workgroup per node of at most SEL_LOCAL rows.
__global__ __launch_bounds__(BT) void k_select_local(const uint64_t* __restrict__ ukey,
                                                     const LvlNode* __restrict__ lvl,
                                                     uint32_t leafsize, Sel* __restrict__ sel) {
  __shared__ uint64_t s_key[SEL_LOCAL];
  __shared__ uint32_t s_hist[256];
  __shared__ uint32_t s_red[8];
  __shared__ uint32_t s_pick[2];
  const LvlNode nd = lvl[blockIdx.x];
  if (!is_inner(nd, leafsize) || nd.len > SEL_LOCAL) return;
  const uint32_t len = nd.len;
  for (uint32_t i = threadIdx.x; i < len; i += BT) s_key[i] = ukey[nd.start + i];
  uint64_t prefix = 0, mask = 0;
  uint32_t rank = len / 2;
  for (int shift = 56; shift >= 0; shift -= 8) {
    s_hist[threadIdx.x] = 0;
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < len; i += BT) {
      const uint64_t k = s_key[i];
      if ((k & mask) == prefix) atomicAdd(&s_hist[(k >> shift) & 255u], 1u);
    }
    __syncthreads();
    // Exclusive scan of the 256 bins (one per thread), then the bin holding rank.
    const uint32_t h = s_hist[threadIdx.x];
    uint32_t incl = h;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t v = __shfl_up(incl, o);
      if (lane >= o) incl += v;
    }
    if (lane == 63) s_red[wave] = incl;
    __syncthreads();
    uint32_t base = 0;
    for (int w = 0; w < wave; ++w) base += s_red[w];
    const uint32_t excl = base + incl - h;
    if (h != 0 && excl <= rank && rank < excl + h) { s_pick[0] = threadIdx.x; s_pick[1] = excl; }
    __syncthreads();
    const uint32_t b = s_pick[0];
    rank -= s_pick[1];
    prefix |= (uint64_t)b << shift;
    mask |= 255ull << shift;
    __syncthreads();
  }
  const double pivot = from_total_key(prefix);
  uint32_t lt = 0;
  for (uint32_t i = threadIdx.x; i < len; i += BT) lt += from_total_key(s_key[i]) < pivot ? 1u : 0u;
  lt = block_sum(lt, s_red);
  if (threadIdx.x == 0) {
    Sel s = sel[blockIdx.x];
    s.prefix = prefix;
    s.pivot = pivot;
    s.lt = lt;
    sel[blockIdx.x] = s;
  }
}

// Global radix select for nodes above SEL_LOCAL rows: a histogram pass over
// the chunks, then a per-node pick, repeated for each byte of the key.
__global__ __launch_bounds__(BT) void k_sel_hist(const uint64_t* __restrict__ ukey,
                                                 const Chunk* __restrict__ chunks,
                                                 const LvlNode* __restrict__ lvl, uint32_t leafsize,
                                                 const Sel* __restrict__ sel, uint32_t* __restrict__ hist,
                                                 int shift) {
  __shared__ uint32_t s_hist[256];
  const Chunk c = chunks[blockIdx.x];
  const LvlNode nd = lvl[c.slot];
  if (!is_inner(nd, leafsize) || nd.len <= SEL_LOCAL) return;
  const uint64_t prefix = sel[c.slot].prefix, mask = sel[c.slot].mask;
  s_hist[threadIdx.x] = 0;
  __syncthreads();
  for (uint32_t r = c.begin + threadIdx.x; r < c.end; r += BT) {
    const uint64_t k = ukey[r];
    if ((k & mask) == prefix) atomicAdd(&s_hist[(k >> shift) & 255u], 1u);
  }
  __syncthreads();
  const uint32_t h = s_hist[threadIdx.x];
  if (h) atomicAdd(&hist[(size_t)c.slot * 256 + threadIdx.x], h);
}

__global__ __launch_bounds__(BT) void k_sel_pick(const LvlNode* __restrict__ lvl, uint32_t leafsize,
                                                 Sel* __restrict__ sel, uint32_t* __restrict__ hist,
                                                 int shift) {
  __shared__ uint32_t s_red[4];
  __shared__ uint32_t s_pick[2];
  const LvlNode nd = lvl[blockIdx.x];
  if (!is_inner(nd, leafsize) || nd.len <= SEL_LOCAL) return;
  uint32_t* hh = hist + (size_t)blockIdx.x * 256;
  const uint32_t h = hh[threadIdx.x];
  hh[threadIdx.x] = 0;  // ready for the next pass
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t incl = h;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t v = __shfl_up(incl, o);
    if (lane >= o) incl += v;
  }
  if (lane == 63) s_red[wave] = incl;
  __syncthreads();
  uint32_t base = 0;
  for (int w = 0; w < wave; ++w) base += s_red[w];
  const uint32_t excl = base + incl - h;
  Sel s = sel[blockIdx.x];
  if (h != 0 && excl <= s.rank && s.rank < excl + h) { s_pick[0] = threadIdx.x; s_pick[1] = excl; }
  __syncthreads();
  if (threadIdx.x == 0) {
    s.rank -= s_pick[1];
    s.prefix |= (uint64_t)s_pick[0] << shift;
    s.mask |= 255ull << shift;
    if (shift == 0) { s.pivot = from_total_key(s.prefix); s.lt = 0; }
    sel[blockIdx.x] = s;
  }
}

__global__ __launch_bounds__(BT) void k_sel_lt(const uint64_t* __restrict__ ukey,
                                               const Chunk* __restrict__ chunks,
                                               const LvlNode* __restrict__ lvl, uint32_t leafsize,
                                               Sel* __restrict__ sel) {
  __shared__ uint32_t s_red[4];
  const Chunk c = chunks[blockIdx.x];
  const LvlNode nd = lvl[c.slot];
  if (!is_inner(nd, leafsize) || nd.len <= SEL_LOCAL) return;
  const double pivot = sel[c.slot].pivot;
  uint32_t n = 0;
  for (uint32_t r = c.begin + threadIdx.x; r < c.end; r += BT) n += from_total_key(ukey[r]) < pivot ? 1u : 0u;
  n = block_sum(n, s_red);
  if (threadIdx.x == 0 && n) atomicAdd(&sel[c.slot].lt, n);
}

// Partition phase 1: counts of misplaced rows per chunk.
__global__ __launch_bounds__(BT) void k_part_count(const uint64_t* __restrict__ ukey,
                                                   const Chunk* __restrict__ chunks,
                                                   const LvlNode* __restrict__ lvl, uint32_t leafsize,
                                                   const Sel* __restrict__ sel, uint32_t* __restrict__ cnt) {
  __shared__ uint32_t s_red[4];
  const Chunk c = chunks[blockIdx.x];
  const LvlNode nd = lvl[c.slot];
  if (!is_inner(nd, leafsize)) return;
  const Sel s = sel[c.slot];
  uint32_t na = 0, nb = 0;
  for (uint32_t r = c.begin + threadIdx.x; r < c.end; r += BT) {
    const uint32_t rel = r - nd.start;
    const double key = from_total_key(ukey[r]);
    na += (rel < s.lt && key >= s.pivot) ? 1u : 0u;   // src/build.rs:282-285
    nb += (rel >= s.lt && key < s.pivot) ? 1u : 0u;   // src/build.rs:287-290
  }
  na = block_sum(na, s_red);
  nb = block_sum(nb, s_red);
  if (threadIdx.x == 0) { cnt[2 * blockIdx.x] = na; cnt[2 * blockIdx.x + 1] = nb; }
}

// Per node: exclusive scan of per-chunk counts (npair values per chunk).
__global__ void k_scan_chunks(const LvlNode* __restrict__ lvl, uint32_t nslots, uint32_t leafsize,
                              const uint32_t* __restrict__ cnt, uint32_t* __restrict__ off, int npair,
                              Sel* __restrict__ sel, int* err, int* max_need) {
  const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= nslots) return;
  const LvlNode nd = lvl[slot];
  if (!is_inner(nd, leafsize)) return;
  uint32_t acc0 = 0, acc1 = 0;
  for (uint32_t ch = nd.cbeg; ch < nd.cend; ++ch) {
    off[npair * ch] = acc0;
    acc0 += cnt[npair * ch];
    if (npair == 2) {
      off[2 * ch + 1] = acc1;
      acc1 += cnt[2 * ch + 1];
    }
  }
  if (npair == 2) {
    sel[slot].m = acc0;
    if (acc0 != acc1) atomicOr(err, 2);  // both sides misplace the same count
    const int need = (int)(nd.len / 2 - sel[slot].lt);
    if (need > 0) atomicMax(max_need, need);
  }
}

// Partition phase 1: the two ascending lists of misplaced rows.
__global__ __launch_bounds__(BT) void k_part_lists(const uint64_t* __restrict__ ukey,
                                                   const Chunk* __restrict__ chunks,
                                                   const LvlNode* __restrict__ lvl, uint32_t leafsize,
                                                   const Sel* __restrict__ sel,
                                                   const uint32_t* __restrict__ off,
                                                   uint32_t* __restrict__ alist, uint32_t* __restrict__ blist) {
  __shared__ uint32_t s_a[4], s_b[4];
  const Chunk c = chunks[blockIdx.x];
  const LvlNode nd = lvl[c.slot];
  if (!is_inner(nd, leafsize)) return;
  const Sel s = sel[c.slot];
  uint32_t ra = off[2 * blockIdx.x], rb = off[2 * blockIdx.x + 1];
  for (uint32_t base = c.begin; base < c.end; base += BT) {
    const uint32_t r = base + threadIdx.x;
    bool fa = false, fb = false;
    if (r < c.end) {
      const uint32_t rel = r - nd.start;
      const double key = from_total_key(ukey[r]);
      fa = rel < s.lt && key >= s.pivot;
      fb = rel >= s.lt && key < s.pivot;
    }
    uint32_t ta, tb;
    const uint32_t ea = block_excl_count(fa, s_a, ta);
    const uint32_t eb = block_excl_count(fb, s_b, tb);
    if (fa) alist[nd.start + ra + ea] = r - nd.start;
    if (fb) blist[nd.start + rb + eb] = r - nd.start;
    ra += ta;
    rb += tb;
  }
}

// Partition phase 1: A[i] <-> B[m-1-i] (src/build.rs:296-300).  Thread i of the
// node owns pair i; pairs are disjoint.
__global__ __launch_bounds__(BT) void k_part_swap(double* __restrict__ rows, uint32_t* __restrict__ idx,
                                                  uint64_t* __restrict__ ukey, uint32_t D,
                                                  const Chunk* __restrict__ chunks,
                                                  const LvlNode* __restrict__ lvl, uint32_t leafsize,
                                                  const Sel* __restrict__ sel,
                                                  const uint32_t* __restrict__ alist,
                                                  const uint32_t* __restrict__ blist) {
  const Chunk c = chunks[blockIdx.x];
  const LvlNode nd = lvl[c.slot];
  if (!is_inner(nd, leafsize)) return;
  const uint32_t m = sel[c.slot].m;
  for (uint32_t r = c.begin + threadIdx.x; r < c.end; r += BT) {
    const uint32_t i = r - nd.start;
    if (i >= m) continue;
    const uint32_t ra_ = alist[nd.start + i], rb_ = blist[nd.start + m - 1 - i];
    if (ra_ >= nd.len || rb_ >= nd.len) continue;  // only with inconsistent counts (flagged)
    const uint32_t pa = nd.start + ra_;
    const uint32_t pb = nd.start + rb_;
    double* ra = rows + (size_t)pa * D;
    double* rb = rows + (size_t)pb * D;
    for (uint32_t j = 0; j < D; ++j) { const double t = ra[j]; ra[j] = rb[j]; rb[j] = t; }
    const uint32_t ti = idx[pa]; idx[pa] = idx[pb]; idx[pb] = ti;
    const uint64_t tk = ukey[pa]; ukey[pa] = ukey[pb]; ukey[pb] = tk;
  }
}

// Fix-up (src/build.rs:303-313): count pivot-equal rows at or after lt.
__global__ __launch_bounds__(BT) void k_fix_count(const uint64_t* __restrict__ ukey,
                                                  const Chunk* __restrict__ chunks,
                                                  const LvlNode* __restrict__ lvl, uint32_t leafsize,
                                                  const Sel* __restrict__ sel, uint32_t* __restrict__ cnt) {
  __shared__ uint32_t s_red[4];
  const Chunk c = chunks[blockIdx.x];
  const LvlNode nd = lvl[c.slot];
  if (!is_inner(nd, leafsize)) return;
  const Sel s = sel[c.slot];
  uint32_t ne = 0;
  if (s.lt < nd.len / 2)
    for (uint32_t r = c.begin + threadIdx.x; r < c.end; r += BT)
      ne += (r - nd.start >= s.lt && from_total_key(ukey[r]) == s.pivot) ? 1u : 0u;
  ne = block_sum(ne, s_red);
  if (threadIdx.x == 0) cnt[blockIdx.x] = ne;
}

// Fix-up (src/build.rs:303-313), made parallel.  Let need = mid - lt and let
// s_1 < ... < s_need be the positions of the first `need` pivot-equal rows at
// or after lt.  The t-th swap of the loop exchanges the rows at place =
// lt + t - 1 and scan = s_t, so the loop is the map g(lt + t - 1) = s_t,
// applied in increasing t:
//   * the pivot-equal row at s_t ends at lt + t - 1;
//   * a row in [mid, s_need] that is not pivot-equal never moves;
//   * a row at P in [lt, mid) that is not pivot-equal is carried by each swap
//     whose place it occupies: P -> g(P) -> g(g(P)) -> ..., and stops at the
//     first position >= mid.
// g is increasing with g(x) > x, so the chains are followed by pointer
// jumping.  This kernel writes g (node-relative) and `last` = s_need.
__global__ __launch_bounds__(BT) void k_fix_find(const uint64_t* __restrict__ ukey,
                                                 const Chunk* __restrict__ chunks,
                                                 const LvlNode* __restrict__ lvl, uint32_t leafsize,
                                                 Sel* __restrict__ sel, const uint32_t* __restrict__ off,
                                                 uint32_t* __restrict__ gmap) {
  __shared__ uint32_t s_e[4];
  const Chunk c = chunks[blockIdx.x];
  const LvlNode nd = lvl[c.slot];
  if (!is_inner(nd, leafsize)) return;
  const Sel s = sel[c.slot];
  const uint32_t need = nd.len / 2 - s.lt;
  if (need == 0) return;
  uint32_t re = off[blockIdx.x];
  if (re >= need) return;
  for (uint32_t base = c.begin; base < c.end; base += BT) {
    const uint32_t r = base + threadIdx.x;
    const bool fe = r < c.end && r - nd.start >= s.lt && from_total_key(ukey[r]) == s.pivot;
    uint32_t te;
    const uint32_t e = block_excl_count(fe, s_e, te);
    if (fe && re + e < need) {
      gmap[nd.start + s.lt + re + e] = r - nd.start;  // g(lt + t - 1) = s_t
      if (re + e == need - 1) sel[c.slot].last = r - nd.start;
    }
    re += te;
  }
}

// One pointer-jumping round over [lt, mid): G'(x) = G(G(x)) while G(x) < mid.
__global__ __launch_bounds__(BT) void k_fix_jump(const Chunk* __restrict__ chunks,
                                                 const LvlNode* __restrict__ lvl, uint32_t leafsize,
                                                 const Sel* __restrict__ sel,
                                                 const uint32_t* __restrict__ gin,
                                                 uint32_t* __restrict__ gout) {
  const Chunk c = chunks[blockIdx.x];
  const LvlNode nd = lvl[c.slot];
  if (!is_inner(nd, leafsize)) return;
  const uint32_t lt = sel[c.slot].lt, mid = nd.len / 2;
  if (lt == mid) return;
  for (uint32_t r = c.begin + threadIdx.x; r < c.end; r += BT) {
    const uint32_t rel = r - nd.start;
    if (rel < lt || rel >= mid) continue;
    uint32_t v = gin[r];
    if (v >= lt && v < mid) v = gin[nd.start + v];
    gout[r] = v;
  }
}

// Fix-up: scatter the window [lt, last] to its final order in tmp.
__global__ __launch_bounds__(BT) void k_fix_scatter(const double* __restrict__ rows,
                                                    const uint32_t* __restrict__ idx,
                                                    const uint64_t* __restrict__ ukey, uint32_t D,
                                                    const Chunk* __restrict__ chunks,
                                                    const LvlNode* __restrict__ lvl, uint32_t leafsize,
                                                    const Sel* __restrict__ sel,
                                                    const uint32_t* __restrict__ off,
                                                    const uint32_t* __restrict__ gfinal,
                                                    double* __restrict__ trows, uint32_t* __restrict__ tidx,
                                                    int* err) {
  __shared__ uint32_t s_e[4];
  const Chunk c = chunks[blockIdx.x];
  const LvlNode nd = lvl[c.slot];
  if (!is_inner(nd, leafsize)) return;
  const Sel s = sel[c.slot];
  const uint32_t mid = nd.len / 2;
  if (s.lt == mid) return;
  if (c.begin - nd.start > s.last) return;
  uint32_t re = off[blockIdx.x];
  for (uint32_t base = c.begin; base < c.end; base += BT) {
    const uint32_t r = base + threadIdx.x;
    const uint32_t rel = r - nd.start;
    const bool inwin = r < c.end && rel >= s.lt && rel <= s.last;
    const bool fe = inwin && from_total_key(ukey[r]) == s.pivot;
    uint32_t te;
    const uint32_t e = block_excl_count(fe, s_e, te);
    if (inwin) {
      uint32_t dest = rel;
      if (fe) dest = s.lt + re + e;
      else if (rel < mid) dest = gfinal[r];
      if (dest < mid && !fe) atomicOr(err, 4);  // chain not resolved (never expected)
      if (dest < nd.len) {
        const uint32_t pd = nd.start + dest;
        for (uint32_t j = 0; j < D; ++j) trows[(size_t)pd * D + j] = rows[(size_t)r * D + j];
        tidx[pd] = idx[r];
      }
    }
    re += te;
  }
}

__global__ __launch_bounds__(BT) void k_fix_copy(double* __restrict__ rows, uint32_t* __restrict__ idx,
                                                 uint32_t D, const Chunk* __restrict__ chunks,
                                                 const LvlNode* __restrict__ lvl, uint32_t leafsize,
                                                 const Sel* __restrict__ sel,
                                                 const double* __restrict__ trows,
                                                 const uint32_t* __restrict__ tidx) {
  const Chunk c = chunks[blockIdx.x];
  const LvlNode nd = lvl[c.slot];
  if (!is_inner(nd, leafsize)) return;
  const Sel s = sel[c.slot];
  if (s.lt == nd.len / 2) return;
  for (uint32_t r = c.begin + threadIdx.x; r < c.end; r += BT) {
    const uint32_t rel = r - nd.start;
    if (rel < s.lt || rel > s.last || s.last >= nd.len) continue;
    for (uint32_t j = 0; j < D; ++j) rows[(size_t)r * D + j] = trows[(size_t)r * D + j];
    idx[r] = tidx[r];
  }
}

__global__ void k_write_nodes(const LvlNode* __restrict__ lvl, uint32_t nslots, uint32_t leafsize,
                              const Sel* __restrict__ sel, DevNode* __restrict__ nodes,
                              uint32_t* __restrict__ node_len) {
  const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= nslots) return;
  const LvlNode nd = lvl[slot];
  DevNode out;
  if (is_inner(nd, leafsize)) {
    const Sel s = sel[slot];
    out.split = s.pivot;
    out.a = nd.right;
    out.b = s.dim;
  } else {
    out.split = __longlong_as_double((long long)(nd.start + nd.len));
    out.a = nd.start;
    out.b = LEAF_BIT;
  }
  nodes[nd.id] = out;
  node_len[nd.id] = nd.len;
}

// order_by_box, src/build.rs:234-238 with shrank_off_axis :248-252.
__global__ void k_order_by_box(DevNode* __restrict__ nodes, uint64_t n_nodes,
                               const double* __restrict__ boxes, uint32_t D,
                               const uint32_t* __restrict__ node_len) {
  const uint64_t id = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= n_nodes) return;
  DevNode nd = nodes[id];
  if (nd.b & LEAF_BIT) return;
  const uint32_t dim = nd.b & DIM_MASK;
  if (D <= 1 || node_len[id] < BOX_ORDER_MIN_SUBTREE) return;
  const double* own = boxes + id * 2 * D;
  const double* lb = boxes + (id + 1) * 2 * D;
  const double* rb = boxes + (size_t)nd.a * 2 * D;
  bool shrank = true;
  for (uint32_t j = 0; j < D && shrank; ++j) {
    if (j == dim) continue;
    const double lim = BOX_ORDER_MAX_SHRINK_RATIO * (own[D + j] - own[j]);
    shrank = (lb[D + j] - lb[j]) < lim && (rb[D + j] - rb[j]) < lim;
  }
  if (shrank) nodes[id].b = nd.b | OBB_BIT;
}

// FP32 norms of the FP32 leaf screen (rows of up to 16 coordinates): pn32[r]
// = fl32 of |v|^2 summed in FP64 from v = fl32(row r) (tree order), and
// *pmax_bits = max over rows of |v| rounded up, as f32 bits.
__global__ void k_norms32(const double* __restrict__ rows, uint64_t n, uint32_t D, float* __restrict__ pn32,
                          uint32_t* pmax_bits) {
  float mx = 0.0f;
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (uint64_t)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (uint32_t c = 0; c < D; ++c) {
      const double v = (double)(float)rows[r * D + c];
      s += v * v;
    }
    pn32[r] = (float)s;
    const double nr = sqrt(s);
    float f = (float)nr;
    if ((double)f < nr) f = nextafterf(f, INFINITY);
    mx = fmaxf(mx, f);
  }
  atomicMax(pmax_bits, __float_as_uint(mx));  // non-negative floats order as integers
}

// The tile walk's rows (rows of up to 16 coordinates): row r in tree order as
// rows32c[r] = fl32(fl64(x - c)) (c the centre of the tree's frame), its norm
// pn32c[r] = fl32 of the FP64 sum of those values squared, and 8 padding rows
// after the last (zero coordinates, NaN norm: a score read past a leaf's end
// is NaN, which passes no test).  *pc_bits = max over rows of |x - c| (FP64),
// rounded up, as f32 bits (the same bound k_rows_h forms for wider rows).
__global__ void k_rows32c(const double* __restrict__ rows, uint64_t n, uint32_t D,
                          const double* __restrict__ frame, float* __restrict__ rows32c,
                          float* __restrict__ pn32c, uint32_t* pc_bits) {
  float mx = 0.0f;
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n + 8;
       r += (uint64_t)gridDim.x * blockDim.x) {
    double s = 0.0, pc = 0.0;
    for (uint32_t c = 0; c < D; ++c) {
      const double y = r < n ? rows[r * D + c] - frame[c] : 0.0;
      const float f = (float)y;
      rows32c[r * D + c] = f;
      s += (double)f * (double)f;
      pc += y * y;
    }
    pn32c[r] = r < n ? (float)s : __builtin_nanf("");
    if (r < n) {
      const double nr = sqrt(pc) * (1.0 + 0x1p-44);
      float f = (float)nr;
      if ((double)f < nr) f = nextafterf(f, INFINITY);
      mx = fmaxf(mx, f);
    }
  }
  atomicMax(pc_bits, __float_as_uint(mx));  // non-negative floats order as integers
}

// The tile walk's node boxes (rows of up to 16 coordinates): each box in the
// frame as FP32, rounded outwards -- lo32 <= lo - c and hi32 >= hi - c
// exactly (the FP64 difference is rounded towards the outside in FP32, then
// moved one more FP32 step outwards, which covers the FP64 difference's own
// rounding) -- so that a box distance formed from them never exceeds the
// exact one by more than the rounding of the query and of the arithmetic.
__global__ void k_boxes32c(const double* __restrict__ boxes, uint64_t n_nodes, uint32_t D,
                           const double* __restrict__ frame, float* __restrict__ boxes32c) {
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n_nodes * 2 * D;
       e += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t j = (uint32_t)(e % D);
    const bool hi = (e / D) % 2 == 1;
    const double y = boxes[e] - frame[j];
    boxes32c[e] = hi ? nextafterf(__double2float_ru(y), INFINITY) : nextafterf(__double2float_rd(y), -INFINITY);
  }
}

// Leaf buckets: rows of each leaf stored coordinate-major, f64 and f32.
__global__ __launch_bounds__(64) void k_leaf_buckets(const double* __restrict__ rows, uint32_t D,
                                                     const uint32_t* __restrict__ leaves /*start,len*/,
                                                     double* __restrict__ pts, float* __restrict__ pts32) {
  const uint32_t start = leaves[2 * blockIdx.x], L = leaves[2 * blockIdx.x + 1];
  const size_t base = (size_t)start * D;
  const size_t total = (size_t)L * D;
  for (size_t e = threadIdx.x; e < total; e += 64) {
    const size_t c = e / L, o = e - c * L;
    const double v = rows[base + (size_t)o * D + c];
    pts[base + e] = v;
    if (pts32) pts32[base + e] = (float)v;
  }
}

// One check of every node after the build, so that the query kernels can
// read nodes unchecked: an inner node's right child lies after it and inside
// the node array and its split dim is a dimension; a leaf's rows [start, end)
// lie inside the points.  Sets bit 8 of *err on a violation.
__global__ void k_validate_nodes(const DevNode* __restrict__ nodes, uint64_t n_nodes, uint64_t n_points,
                                 uint32_t D, int* err) {
  bool bad = false;
  for (uint64_t id = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; id < n_nodes;
       id += (uint64_t)gridDim.x * blockDim.x) {
    const DevNode nd = nodes[id];
    if (nd.b & LEAF_BIT) {
      const uint64_t end = (uint64_t)__double_as_longlong(nd.split);
      bad |= end > n_points || end < nd.a;
    } else {
      bad |= nd.a <= id || nd.a >= n_nodes || (nd.b & DIM_MASK) >= D;
      bad |= id + 1 >= n_nodes;
    }
  }
  if (__ballot(bad) && (threadIdx.x & 63) == 0) atomicOr(err, 8);
}

// The matrix-core screen's frame, from the root box (one wave): centre c_j =
// lo_j / 2 + hi_j / 2 and sigma = 2^s with M sigma < 2^8, M = max_j
// max(|fl64(hi_j - c_j)|, |fl64(lo_j - c_j)|) (every row's fl64(x_j - c_j) lies
// between the two, rounding being monotonic), s = 8 - (ilogb(M) + 1), at most
// 1000 (M = 0: s = 0).  frame = [c_0 .. c_{D-1}, sigma].
__global__ void k_screen_frame(const double* __restrict__ boxes, uint32_t D, double* __restrict__ frame) {
  const uint32_t lane = threadIdx.x;
  double m = 0.0;
  for (uint32_t j = lane; j < D; j += 64) {
    const double lo = boxes[j], hi = boxes[D + j];
    const double c = 0.5 * lo + 0.5 * hi;
    frame[j] = c;
    m = fmax(m, fmax(fabs(hi - c), fabs(lo - c)));
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
  if (lane == 0) {
    const int s = m > 0.0 ? min(8 - (ilogb(m) + 1), 1000) : 0;
    frame[D] = ldexp(1.0, s);
  }
}

// Matrix-core screen rows: tree-order rows x as fp16 of fl32(sigma fl64(x -
// c)) (screen_f16), zero-padded to `stride` columns and n_pad rows, in rows_h;
// pnorm_h = the norm-block words of pn = fl32(|row|^2) (the fp16 values
// squared and summed in FP64; +inf on padding rows); *pc_bits = max over rows
// of |x - c| (FP64), rounded up, as f32 bits.
__global__ void k_rows_h(const double* __restrict__ rows, uint64_t n, uint32_t D, uint32_t stride, uint64_t n_pad,
                         const double* __restrict__ frame, uint16_t* __restrict__ rows_h, uint2* __restrict__ pnorm_h,
                         uint32_t* pc_bits) {
  const double sigma = frame[D];
  float mx = 0.0f;
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n_pad;
       r += (uint64_t)gridDim.x * blockDim.x) {
    double s = 0.0, pc = 0.0;
    for (uint32_t c = 0; c < stride; ++c) {
      const double y = (r < n && c < D) ? rows[r * D + c] - frame[c] : 0.0;
      const _Float16 h = screen_f16(y, sigma);
      rows_h[r * stride + c] = __builtin_bit_cast(uint16_t, h);
      s += (double)h * (double)h;
      pc += y * y;
    }
    pnorm_h[r] = screen_norm_words(r < n ? (float)s : INFINITY);
    if (r < n) {
      // |fl64(x_j - c_j)| >= |x_j - c_j| / (1 + 2^-53); the FP64 sum and root
      // add at most (D + 2) 2^-53 more.
      const double nr = sqrt(pc) * (1.0 + 0x1p-44);
      float f = (float)nr;
      if ((double)f < nr) f = nextafterf(f, INFINITY);
      mx = fmaxf(mx, f);
    }
  }
  atomicMax(pc_bits, __float_as_uint(mx));  // non-negative floats order as integers
}

__global__ void k_original_order(const double* __restrict__ rows, const uint32_t* __restrict__ idx,
                                 uint64_t n, uint32_t D, double* __restrict__ out) {
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n * D;
       e += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t r = e / D, c = e - r * D;
    out[(uint64_t)idx[r] * D + c] = rows[e];
  }
}

// Host-side tree shape: node counts per length (src/build.rs:166-173).
struct ShapeCounter {
  uint64_t leafsize;
  std::map<uint64_t, uint64_t> memo;
  uint64_t operator()(uint64_t len) {
    if (len <= leafsize) return 1;
    auto it = memo.find(len);
    if (it != memo.end()) return it->second;
    const uint64_t mid = len / 2;
    const uint64_t c = 1 + (*this)(mid) + (*this)(len - mid);
    memo[len] = c;
    return c;
  }
};

// A device buffer: owned (alloc: its own hipMalloc, freed with it) or a view
// of a piece of the tree image (place: freed with the image).
template <typename T>
struct DevBuf {
  T* p = nullptr;
  bool owned = false;
  ~DevBuf() { if (p && owned) (void)hipFree(p); }
  hipError_t alloc(size_t n) {
    owned = true;
    return hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T));
  }
  void place(char* image, size_t off) { p = reinterpret_cast<T*>(image + off); owned = false; }
  T* release() { T* q = p; p = nullptr; return q; }
};

// The tree image: every array a built tree keeps, in one device allocation,
// each piece 256-byte aligned, in this order (sizes in elements):
//   nodes    DevNode[n_nodes]
//   boxes    f64[n_nodes][2][D]
//   pts      f64[n][D]            leaf buckets, coordinate-major per leaf
//   idx      u32[n]
//   rows     f64[n][D]            tree order
//   pts32    f32[n][D]            D <= 16: FP32 leaf buckets, same layout as pts
//   pn32     f32[n]               D <= 16: fl32(|fl32(row)|^2), tree order
//   rows32c  f32[n + 8][D]        D <= 16: the tile walk's centred FP32 rows, tree order
//   pn32c    f32[n + 8]           D <= 16: their norms (NaN on the 8 padding rows)
//   boxes32c f32[n_nodes][2][D]   D <= 16: the tile walk's boxes (centred, rounded outwards)
//   frame    f64[D + 1]           D <= 64: the screens' frame (centre, sigma)
//   rows_h   f16[n_pad][S]        8 <= D <= 64: matrix-core screen rows
//   pnorm_h  u32x2[n_pad]         8 <= D <= 64: their norms as norm-block words
// (offsets: image_layout below; one hipMalloc and one hipFree per tree).
struct ImageLayout {
  size_t nodes, boxes, pts, idx, rows, pts32, pn32, rows32c, pn32c, boxes32c, frame, rows_h, pnorm_h, total;
};

size_t image_piece(size_t bytes) { return (std::max<size_t>(bytes, 1) + 255) & ~size_t(255); }

ImageLayout image_layout(uint64_t n, uint64_t D, uint64_t n_nodes, bool narrow, bool wide, uint64_t n_pad,
                         uint64_t stride) {
  ImageLayout L{};
  size_t o = 0;
  auto put = [&](size_t& at, size_t bytes) {
    at = o;
    o += image_piece(bytes);
  };
  put(L.nodes, n_nodes * sizeof(DevNode));
  put(L.boxes, n_nodes * 2 * D * 8);
  put(L.pts, n * D * 8);
  put(L.idx, n * 4);
  put(L.rows, n * D * 8);
  if (narrow) {
    put(L.pts32, n * D * 4);
    put(L.pn32, n * 4);
    put(L.rows32c, (n + 8) * D * 4);
    put(L.pn32c, (n + 8) * 4);
    put(L.boxes32c, n_nodes * 2 * D * 4);
  }
  if (narrow || wide) put(L.frame, (D + 1) * 8);
  if (wide) {
    put(L.rows_h, n_pad * stride * 2);
    put(L.pnorm_h, n_pad * 8);
  }
  L.total = o;
  return L;
}

}  // namespace

uint64_t count_tree_nodes(uint64_t n, uint64_t leafsize) {
  ShapeCounter c{leafsize, {}};
  return c(n);
}

void launch_check_finite(const double* d_x, uint64_t n, int* d_flag, hipStream_t s) {
  const uint64_t blocks = std::min<uint64_t>((n + 255) / 256, 4096);
  if (blocks) k_check_finite<<<(uint32_t)blocks, 256, 0, s>>>(d_x, n, d_flag);
}

void launch_original_order(const Tree& t, double* d_out, hipStream_t s) {
  const uint64_t blocks = std::min<uint64_t>((t.n * t.d + 255) / 256, 8192);
  k_original_order<<<(uint32_t)blocks, 256, 0, s>>>(t.rows, t.idx, t.n, (uint32_t)t.d, d_out);
}

int build_tree(Tree& t, const double* d_data, hipStream_t s, std::string& msg) {
  const uint64_t n = t.n, D = t.d, ls = t.leafsize;
  ShapeCounter count{ls, {}};

  // Levels of the preorder tree: node (id, first row, length), children's ids
  // from the subtree sizes.
  std::vector<std::vector<LvlNode>> levels;
  std::vector<std::vector<Chunk>> lchunks;
  std::vector<uint32_t> leaves;  // (start, len) pairs
  levels.push_back({LvlNode{0, 0, (uint32_t)n, 0, 0, 0, 0, 0}});
  while (!levels.back().empty()) {
    std::vector<LvlNode>& cur = levels.back();
    std::vector<LvlNode> next;
    std::vector<Chunk> ch;
    for (uint32_t slot = 0; slot < cur.size(); ++slot) {
      LvlNode& nd = cur[slot];
      nd.cbeg = (uint32_t)ch.size();
      for (uint64_t b = nd.start; b < (uint64_t)nd.start + nd.len; b += CH)
        ch.push_back(Chunk{slot, (uint32_t)b, (uint32_t)std::min<uint64_t>(b + CH, (uint64_t)nd.start + nd.len), 0});
      nd.cend = (uint32_t)ch.size();
      if (nd.len > ls) {
        const uint32_t mid = nd.len / 2;
        const uint32_t right = (uint32_t)(nd.id + 1 + count(mid));
        nd.right = right;
        next.push_back(LvlNode{nd.id + 1, nd.start, mid, 0, 0, 0, 0, 0});
        next.push_back(LvlNode{right, nd.start + mid, nd.len - mid, 0, 0, 0, 0, 0});
      } else {
        leaves.push_back(nd.start);
        leaves.push_back(nd.len);
      }
    }
    lchunks.push_back(std::move(ch));
    levels.push_back(std::move(next));
  }
  levels.pop_back();
  t.depth = (uint32_t)levels.size() - 1;

  // One upload of every level table.
  size_t max_slots = 0, max_chunks = 0, tot_nodes = 0, tot_chunks = 0;
  std::vector<size_t> node_off, chunk_off;
  for (size_t l = 0; l < levels.size(); ++l) {
    node_off.push_back(tot_nodes);
    chunk_off.push_back(tot_chunks);
    tot_nodes += levels[l].size();
    tot_chunks += lchunks[l].size();
    max_slots = std::max(max_slots, levels[l].size());
    max_chunks = std::max(max_chunks, lchunks[l].size());
  }
  std::vector<LvlNode> all_nodes;
  std::vector<Chunk> all_chunks;
  all_nodes.reserve(tot_nodes);
  all_chunks.reserve(tot_chunks);
  for (size_t l = 0; l < levels.size(); ++l) {
    all_nodes.insert(all_nodes.end(), levels[l].begin(), levels[l].end());
    all_chunks.insert(all_chunks.end(), lchunks[l].begin(), lchunks[l].end());
  }

  DevBuf<LvlNode> d_lvl;
  DevBuf<Chunk> d_chunks;
  DevBuf<double> rows, trows, part, boxes, pts;
  DevBuf<float> pts32, pn32, rows32c, pn32c, boxes32c;
  DevBuf<double> frame;
  DevBuf<uint2> pnorm_h;
  DevBuf<uint16_t> rows_h;
  DevBuf<uint32_t> idx, tidx, alist, blist, cnt, off, hist, node_len, d_leaves;
  DevBuf<uint64_t> ukey;
  DevBuf<Sel> sel;
  DevBuf<DevNode> nodes;
  DevBuf<int> err;
  const bool narrow = D <= 16;          // FP32 leaf buckets and norms: the traversal's FP32 screen
  const bool wide = D >= 8 && D <= 64;  // rows for the matrix-core screen
  const uint32_t stride32 = D <= 16 ? 16 : D <= 32 ? 32 : 64;
  const uint64_t n_pad = (n + 31) / 32 * 32;
  // The arrays the tree keeps: one device image (freed by kdt_free), the
  // build's working arrays separately.
  const ImageLayout IL = image_layout(n, D, t.n_nodes, narrow, wide, n_pad, stride32);
  char* image = nullptr;
  KDT_HIP_CHECK(hipMalloc((void**)&image, IL.total));
  struct ImageGuard {
    char*& p;
    ~ImageGuard() { if (p) (void)hipFree(p); }
  } image_guard{image};
  nodes.place(image, IL.nodes);
  boxes.place(image, IL.boxes);
  pts.place(image, IL.pts);
  idx.place(image, IL.idx);
  rows.place(image, IL.rows);
  if (narrow) {
    pts32.place(image, IL.pts32);
    pn32.place(image, IL.pn32);
    rows32c.place(image, IL.rows32c);
    pn32c.place(image, IL.pn32c);
    boxes32c.place(image, IL.boxes32c);
  }
  if (narrow || wide) frame.place(image, IL.frame);
  if (wide) {
    rows_h.place(image, IL.rows_h);
    pnorm_h.place(image, IL.pnorm_h);
  }
  KDT_HIP_CHECK(d_lvl.alloc(tot_nodes));
  KDT_HIP_CHECK(d_chunks.alloc(tot_chunks));
  KDT_HIP_CHECK(trows.alloc(n * D));
  KDT_HIP_CHECK(part.alloc(max_chunks * 2 * D));
  KDT_HIP_CHECK(tidx.alloc(n));
  KDT_HIP_CHECK(alist.alloc(n));
  KDT_HIP_CHECK(blist.alloc(n));
  KDT_HIP_CHECK(cnt.alloc(max_chunks * 2));
  KDT_HIP_CHECK(off.alloc(max_chunks * 2));
  KDT_HIP_CHECK(hist.alloc(max_slots * 256));
  KDT_HIP_CHECK(node_len.alloc(t.n_nodes));
  KDT_HIP_CHECK(d_leaves.alloc(leaves.size()));
  KDT_HIP_CHECK(ukey.alloc(n));
  KDT_HIP_CHECK(sel.alloc(max_slots));
  // [0] error bits, [1] largest fix-up need of a level, [2] pmax bits (FP32
  // leaf screen), [3] max |x - c| bits (matrix-core screen)
  KDT_HIP_CHECK(err.alloc(4));

  KDT_HIP_CHECK(hipMemcpyAsync(d_lvl.p, all_nodes.data(), tot_nodes * sizeof(LvlNode), hipMemcpyHostToDevice, s));
  KDT_HIP_CHECK(hipMemcpyAsync(d_chunks.p, all_chunks.data(), tot_chunks * sizeof(Chunk), hipMemcpyHostToDevice, s));
  KDT_HIP_CHECK(hipMemcpyAsync(d_leaves.p, leaves.data(), leaves.size() * 4, hipMemcpyHostToDevice, s));
  KDT_HIP_CHECK(hipMemcpyAsync(rows.p, d_data, n * D * 8, hipMemcpyDeviceToDevice, s));
  KDT_HIP_CHECK(hipMemsetAsync(hist.p, 0, max_slots * 256 * 4, s));
  KDT_HIP_CHECK(hipMemsetAsync(err.p, 0, 16, s));

  // Finiteness of the input (the root bbox check, src/build.rs:51-57).
  launch_check_finite(rows.p, n * D, err.p, s);
  int h_err = 0;
  KDT_HIP_CHECK(hipMemcpyAsync(&h_err, err.p, 4, hipMemcpyDeviceToHost, s));
  KDT_HIP_CHECK(hipStreamSynchronize(s));
  if (h_err) {
    msg = "all coordinates must be finite";
    return 7;
  }
  k_iota<<<(uint32_t)std::min<uint64_t>((n + 255) / 256, 4096), 256, 0, s>>>(idx.p, n);

  const uint32_t Du = (uint32_t)D, lsu = (uint32_t)ls;
  for (size_t l = 0; l < levels.size(); ++l) {
    const uint32_t ns = (uint32_t)levels[l].size(), nc = (uint32_t)lchunks[l].size();
    const LvlNode* L = d_lvl.p + node_off[l];
    const Chunk* C = d_chunks.p + chunk_off[l];
    bool any_inner = false, any_big = false;
    for (const LvlNode& nd : levels[l]) {
      any_inner |= nd.len > ls;
      any_big |= nd.len > ls && nd.len > SEL_LOCAL;
    }
    k_box_partial<<<nc, BT, 0, s>>>(rows.p, Du, C, part.p);
    k_box_final<<<ns, 64, 0, s>>>(part.p, L, Du, lsu, boxes.p, sel.p);
    if (any_inner) {
      k_keys<<<nc, BT, 0, s>>>(rows.p, Du, C, L, lsu, sel.p, ukey.p);
      k_select_local<<<ns, BT, 0, s>>>(ukey.p, L, lsu, sel.p);
      if (any_big) {
        for (int shift = 56; shift >= 0; shift -= 8) {
          k_sel_hist<<<nc, BT, 0, s>>>(ukey.p, C, L, lsu, sel.p, hist.p, shift);
          k_sel_pick<<<ns, BT, 0, s>>>(L, lsu, sel.p, hist.p, shift);
        }
        k_sel_lt<<<nc, BT, 0, s>>>(ukey.p, C, L, lsu, sel.p);
      }
      const uint32_t sb = (ns + 255) / 256;
      k_part_count<<<nc, BT, 0, s>>>(ukey.p, C, L, lsu, sel.p, cnt.p);
      k_scan_chunks<<<sb, 256, 0, s>>>(L, ns, lsu, cnt.p, off.p, 2, sel.p, err.p, err.p + 1);
      k_part_lists<<<nc, BT, 0, s>>>(ukey.p, C, L, lsu, sel.p, off.p, alist.p, blist.p);
      k_part_swap<<<nc, BT, 0, s>>>(rows.p, idx.p, ukey.p, Du, C, L, lsu, sel.p, alist.p, blist.p);
      // The fix-up only runs where pivot-equal rows must move; the level's
      // largest need also bounds the pointer-jumping rounds.
      int max_need = 0;
      KDT_HIP_CHECK(hipMemcpyAsync(&max_need, err.p + 1, 4, hipMemcpyDeviceToHost, s));
      KDT_HIP_CHECK(hipStreamSynchronize(s));
      if (max_need > 0) {
        KDT_HIP_CHECK(hipMemsetAsync(err.p + 1, 0, 4, s));
        k_fix_count<<<nc, BT, 0, s>>>(ukey.p, C, L, lsu, sel.p, cnt.p);
        k_scan_chunks<<<sb, 256, 0, s>>>(L, ns, lsu, cnt.p, off.p, 1, sel.p, err.p, err.p + 1);
        k_fix_find<<<nc, BT, 0, s>>>(ukey.p, C, L, lsu, sel.p, off.p, alist.p);
        uint32_t* g = alist.p;
        uint32_t* g2 = blist.p;
        for (int span = 1; span < max_need; span *= 2) {
          k_fix_jump<<<nc, BT, 0, s>>>(C, L, lsu, sel.p, g, g2);
          std::swap(g, g2);
        }
        k_fix_scatter<<<nc, BT, 0, s>>>(rows.p, idx.p, ukey.p, Du, C, L, lsu, sel.p, off.p, g, trows.p,
                                         tidx.p, err.p);
        k_fix_copy<<<nc, BT, 0, s>>>(rows.p, idx.p, Du, C, L, lsu, sel.p, trows.p, tidx.p);
      }
    }
    k_write_nodes<<<(ns + 255) / 256, 256, 0, s>>>(L, ns, lsu, sel.p, nodes.p, node_len.p);
  }
  k_order_by_box<<<(uint32_t)((t.n_nodes + 255) / 256), 256, 0, s>>>(nodes.p, t.n_nodes, boxes.p, Du,
                                                                    node_len.p);
  k_leaf_buckets<<<(uint32_t)(leaves.size() / 2), 64, 0, s>>>(rows.p, Du, d_leaves.p, pts.p, pts32.p);
  if (narrow || wide) k_screen_frame<<<1, 64, 0, s>>>(boxes.p, Du, frame.p);
  if (wide) {
    k_rows_h<<<(uint32_t)std::min<uint64_t>((n_pad + 255) / 256, 8192), 256, 0, s>>>(
        rows.p, n, Du, stride32, n_pad, frame.p, rows_h.p, pnorm_h.p, reinterpret_cast<uint32_t*>(err.p) + 3);
  }
  if (narrow) {
    k_norms32<<<(uint32_t)std::min<uint64_t>((n + 255) / 256, 8192), 256, 0, s>>>(
        rows.p, n, Du, pn32.p, reinterpret_cast<uint32_t*>(err.p) + 2);
    k_rows32c<<<(uint32_t)std::min<uint64_t>((n + 8 + 255) / 256, 8192), 256, 0, s>>>(
        rows.p, n, Du, frame.p, rows32c.p, pn32c.p, reinterpret_cast<uint32_t*>(err.p) + 3);
    k_boxes32c<<<(uint32_t)std::min<uint64_t>((t.n_nodes * 2 * D + 255) / 256, 8192), 256, 0, s>>>(
        boxes.p, t.n_nodes, Du, frame.p, boxes32c.p);
  }
  k_validate_nodes<<<(uint32_t)std::min<uint64_t>((t.n_nodes + 255) / 256, 4096), 256, 0, s>>>(
      nodes.p, t.n_nodes, n, Du, err.p);
  KDT_HIP_CHECK(hipGetLastError());

  std::vector<double> rb(2 * D);
  KDT_HIP_CHECK(hipMemcpyAsync(rb.data(), boxes.p, 2 * D * 8, hipMemcpyDeviceToHost, s));
  KDT_HIP_CHECK(hipMemcpyAsync(&h_err, err.p, 4, hipMemcpyDeviceToHost, s));
  uint32_t h_pmax[2] = {0, 0};
  KDT_HIP_CHECK(hipMemcpyAsync(h_pmax, err.p + 2, 8, hipMemcpyDeviceToHost, s));
  double h_sigma = 1.0;
  if (wide) KDT_HIP_CHECK(hipMemcpyAsync(&h_sigma, frame.p + D, 8, hipMemcpyDeviceToHost, s));
  KDT_HIP_CHECK(hipStreamSynchronize(s));
  if (h_err & 8) {
    msg = "internal error: malformed node after the build";
    return 100;
  }
  if (h_err) {
    msg = "internal error: partition counts disagree";
    return 100;
  }
  if (t.depth > 32) {  // the query stack holds two entries per inner level
    msg = "internal error: tree deeper than 32 levels";
    return 100;
  }
  double mx = 0.0;
  for (double v : rb) mx = std::max(mx, std::fabs(v));
  t.max_abs = mx;
  t.nodes = nodes.release();
  t.boxes = boxes.release();
  t.pts = pts.release();
  t.pts32 = pts32.release();
  t.pn32 = pn32.release();
  t.rows32c = rows32c.release();
  t.pn32c = pn32c.release();
  t.boxes32c = boxes32c.release();
  t.idx = idx.release();
  t.rows = rows.release();
  t.frame = frame.release();
  t.rows_h = rows_h.release();
  t.pnorm_h = pnorm_h.release();
  if (narrow) {
    float pm;
    std::memcpy(&pm, &h_pmax[0], 4);
    t.pmax = (double)pm * (1.0 + 1e-6);
  }
  if (narrow || wide) {
    float pc;
    std::memcpy(&pc, &h_pmax[1], 4);
    t.scr_pc = (double)pc;  // rounded up in k_rows_h / k_rows32c
  }
  if (wide) t.scr_sigma = h_sigma;
  t.image = image;
  t.image_bytes = IL.total;
  image = nullptr;  // the tree owns it now
  return 0;
}

}  // namespace kdt
