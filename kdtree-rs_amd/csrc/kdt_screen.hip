// kdt_screen.hip -- exact k-NN for wide rows (32 <= ndim <= 64, and
// 16 <= ndim < 32 on request; p = 2, eps = 0) by an MFMA distance screen, an
// FP64 recheck and a visit certificate.
//
// Replaces, for these widths, the leaf scans of the reference's descent
// (Streamed plan, src/kernel.rs:296-303) without changing any result.  At
// such widths the split planes of a median tree cut few of the dimensions,
// so the descent of src/query.rs:234-341 visits nearly every leaf (all of
// them for 64-d Gaussian rows): the work is a dense query x point
// contraction, which is what the matrix cores do.
//
// Steps for a query matrix against one tree:
//   1. k_screen: for a tile of 32 queries and a slice of the points,
//      v_mfma_f32_32x32x2f32 forms q.p in FP32 and each lane keeps the KP = 16
//      smallest  A = |q|^2 + |p|^2 - 2 q.p  of the points it sees, with the
//      points' tree positions;
//   2. k_screen_select: the k-th smallest A over all of a query's lists gives
//      A_k, and from it a threshold tau_A such that every point whose FP64
//      distance could be among the k best has A <= tau_A (error bound
//      below).  A lane dropped only points with A >= its final 16th value, so
//      when every full list's 16th value exceeds tau_A the list entries with
//      A <= tau_A are all the candidates; otherwise (or beyond 64
//      candidates) the query is flagged;
//   3. k_screen_recheck: FP64 distances of the candidates in the reference's
//      summation order (point_dist_d), lexicographic (distance, index) top-k,
//      then the certificate: for each of the k points, every gate on its
//      root-to-leaf path (plane bound and box bound of each far child, box
//      bound of each child of an order_by_box node, the root's seeded cell)
//      is <= the k-th distance.  The reference's bound never drops below the
//      k-th distance, so it opens every one of those gates: it visits every
//      point of this top-k, and its answer (the top-k of what it visits) is
//      this top-k.  Queries whose certificate fails are flagged;
//   4. the traversal kernels run the flagged queries (KArgs::only).
//
// FP32 error bound (u = 2^-24, q~ and p~ the rows rounded to FP32, P the
// largest |p~| of the tree, N = |q~| + P):
//   |A - |q~ - p~|^2| <= E2 = (2 g + 8u) N^2 + D 2^-125,  g = (2D+4)u/(1-(2D+4)u)
//     (MFMA accumulation of D products, FP32 rounding of |q~|^2, |p~|^2, the
//      sum and the final fma; flushed subnormal products);
//   | |q~ - p~| - |q - p| | <= e1 = u (1 + 2u) N + sqrt(D) 2^-148;
//   the FP64 reduced distance d of the reference satisfies d / s <= |q-p|^2
//     <= d s with s = 1 + 2 (D + 3) 2^-53.
// So U(A) = s (sqrt(A + E2) + e1)^2 bounds d from above, and a point with
// d <= tau has A <= tau_A = (sqrt(s tau) + e1)^2 + E2.  With tau =
// min(U(A_k), cutoff): at least k points have d <= U(A_k), so the k-th best
// distance is <= tau and every point of the top-k is a candidate.
// Rows with a coordinate above 1e18 in magnitude are not screened (flagged).

#include <algorithm>
#include <cmath>
#include <cstdlib>

#include "kdt_query_kernels.h"

namespace kdt {

int launch_knn_lds(const KArgs& a, int metric, hipStream_t s);

namespace {

typedef float floatx16 __attribute__((ext_vector_type(16)));

constexpr uint32_t CAND_CAP = 64;  // candidates per query (one per lane of the recheck)
constexpr int KP = 16;             // largest k on this path
constexpr int LISTS_PER_LANE = 4;  // k_screen_select reads 64 x 4 per-lane lists at most
constexpr uint32_t MAX_SLICES = 32 * LISTS_PER_LANE;  // two lists (lane halves) per slice

struct SArgs {
  const float* __restrict__ rows32;  // [n_pad][2 KS], tree order, zero-padded
  const float* __restrict__ pnorm;   // [n_pad] |p~|^2, +inf on padding rows
  const double* __restrict__ q;      // [nq][D]
  float* __restrict__ part_a;        // [nq][2 S][KP] the kept A values, ascending
  uint32_t* __restrict__ part_i;     // [nq][2 S][KP] their tree positions
  uint64_t nq;
  uint32_t D, n, tiles_per_slice, n_tiles, slices;
};

// One wave: 32 queries (B operand, columns) against 32-row tiles of points
// (A operand, rows).  Lane l holds query j = l & 31, coordinates
// [h KS, h KS + KS) with h = l >> 5, in both operands, so instruction s
// contracts coordinates s and KS + s.  The accumulator holds D[i][j] for
// point rows i = 8 (r / 4) + 4 h + (r % 4), r = 0..15.
template <int KS>
__global__ __launch_bounds__(64) void k_screen(SArgs a) {
  const uint32_t lane = threadIdx.x;
  const uint32_t j = lane & 31, h = lane >> 5;
  const uint64_t qrow = (uint64_t)blockIdx.x * 32 + j;
  const bool qvalid = qrow < a.nq;
  float qf[KS];
  double qn_part = 0.0;
#pragma unroll
  for (int s = 0; s < KS; ++s) {
    const uint32_t c = h * KS + s;
    const float v = (qvalid && c < a.D) ? (float)a.q[qrow * a.D + c] : 0.0f;
    qf[s] = v;
    qn_part += (double)v * (double)v;
  }
  const float qn = (float)(qn_part + __shfl_xor(qn_part, 32));
  const uint32_t t0 = blockIdx.y * a.tiles_per_slice;
  const uint32_t t1 = min(a.n_tiles, t0 + a.tiles_per_slice);

  float la[KP];     // the KP smallest A seen by this lane, ascending
  uint32_t li[KP];  // their tree positions
#pragma unroll
  for (int e = 0; e < KP; ++e) {
    la[e] = INFINITY;
    li[e] = 0xFFFFFFFFu;
  }
  float thr = INFINITY;  // la[KP - 1]

  // Point tiles are loaded two tiles ahead of the contraction (rotating
  // register buffers); the tile's |p~|^2 are loaded as its contraction starts.
  constexpr int NV = KS / 4;
  float4 cur[NV], nx1[NV], nx2[NV];
  auto load = [&](uint32_t t, float4* dst) {
    const float4* src = reinterpret_cast<const float4*>(a.rows32 + ((size_t)t * 32 + j) * (2 * KS) + h * KS);
#pragma unroll
    for (int v = 0; v < NV; ++v) dst[v] = src[v];
  };
  if (t0 < t1) load(t0, cur);
  if (t0 + 1 < t1) load(t0 + 1, nx1);
  for (uint32_t t = t0; t < t1; ++t) {
    if (t + 2 < t1) load(t + 2, nx2);
    float4 pn_cur[4];
    const float4* ps = reinterpret_cast<const float4*>(a.pnorm + (size_t)t * 32 + 4 * h);
#pragma unroll
    for (int g = 0; g < 4; ++g) pn_cur[g] = ps[2 * g];
    floatx16 acc = {};
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(cur[v].x, qf[4 * v + 0], acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(cur[v].y, qf[4 * v + 1], acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(cur[v].z, qf[4 * v + 2], acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(cur[v].w, qf[4 * v + 3], acc, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const float4 pg = pn_cur[r / 4];
      const float pn = (r % 4 == 0) ? pg.x : (r % 4 == 1) ? pg.y : (r % 4 == 2) ? pg.z : pg.w;
      const float A = fmaf(-2.0f, acc[r], qn + pn);  // +inf on padding rows
      if (A < thr) {  // insert (A, row), dropping the largest
        const uint32_t row = t * 32 + 8 * (r / 4) + 4 * h + (r % 4);
#pragma unroll
        for (int e = KP - 1; e > 0; --e) {
          const bool up = A < la[e - 1];
          const bool here = !up && A < la[e];
          la[e] = up ? la[e - 1] : (here ? A : la[e]);
          li[e] = up ? li[e - 1] : (here ? row : li[e]);
        }
        if (A < la[0]) {
          la[0] = A;
          li[0] = row;
        }
        thr = la[KP - 1];
      }
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      cur[v] = nx1[v];
      nx1[v] = nx2[v];
    }
  }
  if (qvalid) {
    const size_t base = (qrow * (2 * a.slices) + 2 * blockIdx.y + h) * KP;
#pragma unroll
    for (int e = 0; e < KP; ++e) {
      a.part_a[base + e] = la[e];
      a.part_i[base + e] = li[e];
    }
  }
}

struct SelArgs {
  const float* __restrict__ part_a;
  const uint32_t* __restrict__ part_i;
  const double* __restrict__ q;
  uint32_t* __restrict__ cnt;
  uint32_t* __restrict__ cand;
  uint8_t* __restrict__ flag;
  uint64_t nq;
  uint32_t D, k, lists;
  double pmax, cutoff, max_abs;
};

// One wave per query; lane l reads lists l, l + 64, ... (at most LISTS_PER_LANE):
// k rounds of wave-min over the list heads give A_k, then tau_A from the error
// bound, the completeness check and the candidates (list entries with
// A <= tau_A, at most CAND_CAP).
__global__ __launch_bounds__(64) void k_screen_select(SelArgs a) {
  const uint64_t qi = blockIdx.x;
  const uint32_t lane = threadIdx.x;
  const float* la0 = a.part_a + qi * a.lists * KP;
  const uint32_t* li0 = a.part_i + qi * a.lists * KP;
  uint32_t head[LISTS_PER_LANE];
  float hv[LISTS_PER_LANE];  // value at each head (+inf when exhausted or absent)
#pragma unroll
  for (int c = 0; c < LISTS_PER_LANE; ++c) {
    const uint32_t L = lane + 64 * c;
    head[c] = 0;
    hv[c] = L < a.lists ? la0[L * KP] : INFINITY;
  }
  float ak = INFINITY;
  for (uint32_t r = 0; r < a.k; ++r) {
    float cur = hv[0];
#pragma unroll
    for (int c = 1; c < LISTS_PER_LANE; ++c) cur = fminf(cur, hv[c]);
    float m = cur;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fminf(m, __shfl_xor(m, o));
    ak = m;
    if (m == INFINITY) break;
    const uint64_t hit = __ballot(cur == m);
    if (lane == (uint32_t)(__ffsll((unsigned long long)hit) - 1)) {
      bool done = false;
#pragma unroll
      for (int c = 0; c < LISTS_PER_LANE; ++c) {
        if (!done && hv[c] == m) {
          done = true;
          ++head[c];
          hv[c] = head[c] < KP ? la0[(lane + 64 * c) * KP + head[c]] : INFINITY;
        }
      }
    }
  }
  double qn = 0.0, qmax = 0.0;
  for (uint32_t c = lane; c < a.D; c += 64) {
    const double v = a.q[qi * a.D + c];
    const double f = (double)(float)v;
    qn += f * f;
    qmax = fmax(qmax, fabs(v));
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    qn += __shfl_xor(qn, o);
    qmax = fmax(qmax, __shfl_xor(qmax, o));
  }
  // tau_A (every lane computes the same value)
  const double u = 0x1p-24;
  const double D = (double)a.D;
  const double m2 = (2.0 * D + 4.0) * u;
  const double g = m2 / (1.0 - m2);
  const double N = (sqrt(qn) * (1.0 + 0x1p-50) + a.pmax);
  const double E2 = ((2.0 * g + 8.0 * u) * N * N + D * 0x1p-125) * (1.0 + 1e-12);
  const double e1 = (u * (1.0 + 2.0 * u) * N + sqrt(D) * 0x1p-148) * (1.0 + 1e-12);
  const double sh = 1.0 + 2.0 * (D + 3.0) * 0x1p-53;
  double tau = a.cutoff;
  if (ak < INFINITY) {
    const double r = sqrt(fmax(0.0, (double)ak + E2)) + e1;
    tau = fmin(tau, sh * r * r * (1.0 + 1e-12));
  }
  float ta = INFINITY;
  if (tau < INFINITY) {
    const double r = sqrt(sh * tau) + e1;
    const double t = (r * r + E2) * (1.0 + 1e-12);
    ta = (float)t;
    if ((double)ta < t) ta = nextafterf(ta, INFINITY);
  }
  // Completeness: a full list whose last value is <= tau_A may have dropped a
  // candidate.  Unfilled slots (A = +inf, position 0xFFFFFFFF) end a list.
  uint32_t cnt_c[LISTS_PER_LANE];
  uint32_t mcount = 0;
  bool complete = true;
#pragma unroll
  for (int c = 0; c < LISTS_PER_LANE; ++c) {
    const uint32_t L = lane + 64 * c;
    uint32_t n = 0;
    if (L < a.lists) {
      const float* la = la0 + L * KP;
      const uint32_t* li = li0 + L * KP;
      const float last = la[KP - 1];
      complete = complete && (last == INFINITY || last > ta);
      while (n < KP && li[n] != 0xFFFFFFFFu && la[n] <= ta) ++n;
    }
    cnt_c[c] = n;
    mcount += n;
  }
  const bool ok = qmax <= 1e18 && a.max_abs <= 1e18 && !__ballot(!complete);
  // Exclusive prefix sum of the per-lane counts.
  uint32_t incl = mcount;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t v = __shfl_up(incl, o);
    if (lane >= (uint32_t)o) incl += v;
  }
  const uint32_t total = __shfl(incl, 63);
  const bool good = ok && total <= CAND_CAP;
  if (good) {
    uint32_t off = incl - mcount;
#pragma unroll
    for (int c = 0; c < LISTS_PER_LANE; ++c) {
      const uint32_t* li = li0 + (lane + 64 * c) * KP;
      for (uint32_t e = 0; e < cnt_c[c]; ++e) a.cand[qi * CAND_CAP + off + e] = li[e];
      off += cnt_c[c];
    }
  }
  if (lane == 0) {
    a.cnt[qi] = good ? total : 0;
    a.flag[qi] = good ? 0 : 1;
  }
}

struct CheckArgs {
  const double* __restrict__ q;
  const double* __restrict__ rows;
  const uint32_t* __restrict__ idx;
  const DevNode* __restrict__ nodes;
  const double* __restrict__ boxes;
  const uint32_t* __restrict__ cnt;
  const uint32_t* __restrict__ cand;
  uint8_t* __restrict__ flag;
  double* __restrict__ out_d;
  int64_t* __restrict__ out_i;
  uint64_t nq;
  uint32_t D, k, n;
  double cutoff;
};

__device__ __forceinline__ bool key_less(double da, uint32_t ia, double db, uint32_t ib) {
  return da < db || (da == db && ia < ib);
}

// One wave per query: FP64 recheck of the candidates, top-k, certificate.
// Dynamic LDS: k x D doubles, the cell-bound axes of each certificate walk.
__global__ __launch_bounds__(64) void k_screen_recheck(CheckArgs a) {
  extern __shared__ double axes_all[];
  const uint64_t qi = blockIdx.x;
  const uint32_t lane = threadIdx.x;
  const uint32_t D = a.D;
  if (a.flag[qi]) return;
  const uint32_t c = a.cnt[qi];
  if (c > CAND_CAP) {
    if (lane == 0) a.flag[qi] = 1;
    return;
  }
  const double* q = a.q + qi * D;
  double d = INFINITY;
  uint32_t id = 0xFFFFFFFFu, pos = 0;
  bool bad_pos = false;
  if (lane < c) {
    pos = a.cand[qi * CAND_CAP + lane];
    bad_pos = pos >= a.n;
  }
  if (__ballot(bad_pos)) {  // never for a list built by k_screen_select
    if (lane == 0) a.flag[qi] = 1;
    return;
  }
  if (lane < c) {
    const double v = point_dist_d<M_L2, 0>(q, a.rows + (size_t)pos * D, 1, D, true, 2.0);
    if (v <= a.cutoff) {
      d = v;
      id = a.idx[pos];
    }
  }
  // Bitonic sort of the 64 (distance, index, position) keys, ascending.
  for (uint32_t size = 2; size <= 64; size <<= 1) {
    for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
      const double od = __shfl_xor(d, stride);
      const uint32_t oid = __shfl_xor(id, stride);
      const uint32_t opos = __shfl_xor(pos, stride);
      const bool up = (lane & size) == 0;
      const bool lower = (lane & stride) == 0;
      const bool other_less = key_less(od, oid, d, id);
      const bool take = (lower == up) ? other_less : !other_less && !(od == d && oid == id);
      if (take) {
        d = od;
        id = oid;
        pos = opos;
      }
    }
  }
  const uint32_t nvalid = (uint32_t)__popcll(__ballot(id != 0xFFFFFFFFu));
  const uint32_t nres = min(nvalid, a.k);
  const double kth = __shfl(d, (int)(a.k - 1));
  const double dcert = nres == a.k ? kth : a.cutoff;

  bool ok = true;
  if (lane < nres) {
    double* axes = axes_all + (size_t)lane * D;
    double cell = 0.0;
    const double root = box_dist_d<M_L2, 0>(q, a.boxes, D, 2.0);
    for (uint32_t j = 0; j < D; ++j) axes[j] = 0.0;
    if (root != 0.0) {  // CellBound::seed, src/query.rs:170-187
      for (uint32_t j = 0; j < D; ++j) {
        const double ax = reduce_d<M_L2>(axis_offset_d(q[j], a.boxes[j], a.boxes[D + j]), 2.0);
        axes[j] = ax;
        cell = fold_d<M_L2>(cell, ax);
      }
    }
    ok = cell <= dcert;  // Descent::run, src/query.rs:227-229
    uint32_t node = 0, start = 0, len = a.n;
    while (ok) {
      const DevNode nd = a.nodes[node];
      if (nd.b & LEAF_BIT) break;
      const uint32_t dim = nd.b & DIM_MASK;
      const double diff = q[dim] - nd.split;
      const bool goleft = diff <= 0.0;
      const uint32_t mid = len / 2;  // split_point, src/build.rs:161-164
      const bool to_left = pos < start + mid;
      const bool is_far = to_left != goleft;
      const uint32_t child = to_left ? node + 1 : nd.a;
      const double far_axis = reduce_d<M_L2>(fabs(diff), 2.0);
      const double far_dist = replace_axis_d<M_L2>(cell, axes[dim], far_axis);
      if (nd.b & OBB_BIT) {  // descend_by_box, src/query.rs:296-327
        ok = box_dist_d<M_L2, 0>(q, a.boxes + (size_t)child * 2 * D, D, 2.0) <= dcert;
      } else if (is_far) {   // descend, src/query.rs:278-286
        ok = far_dist <= dcert && box_dist_d<M_L2, 0>(q, a.boxes + (size_t)child * 2 * D, D, 2.0) <= dcert;
      }
      if (is_far) {          // enter_far, src/query.rs:329-341
        axes[dim] = far_axis;
        cell = far_dist;
      }
      if (to_left) {
        len = mid;
      } else {
        start += mid;
        len -= mid;
      }
      node = child;
    }
  }
  if (__ballot(!ok)) {
    if (lane == 0) a.flag[qi] = 1;
    return;
  }
  if (lane < a.k) {  // write_results, src/query.rs:382-400
    a.out_d[qi * a.k + lane] = lane < nres ? sqrt(d) : INFINITY;
    a.out_i[qi * a.k + lane] = lane < nres ? (int64_t)id : (int64_t)a.n;
  }
}

template <typename T>
int dev_alloc(T** p, size_t count, hipStream_t s) {
  return hipMallocAsync((void**)p, std::max<size_t>(1, count) * sizeof(T), s) == hipSuccess ? 0 : 1;
}

}  // namespace

bool screen_applies(const Tree& t, const QueryParams& qp) {
  return qp.fp32_screen && qp.metric == M_L2 && qp.eps_factor == 1.0 && qp.k <= (uint64_t)KP && t.rows32 &&
         qp.nq < (1ull << 31) && (t.d >= 32 || (qp.matrix_screen && t.d >= 16)) && t.d <= 64 &&
         std::getenv("KDT_NO_MFMA_SCREEN") == nullptr;
}

int launch_screen(const Tree& t, const QueryParams& qp, const KArgs& trav, hipStream_t s, std::string& msg) {
  const uint64_t nq = qp.nq;
  const uint32_t k = (uint32_t)qp.k;
  const uint64_t qtiles = (nq + 31) / 32;
  const uint32_t n_tiles = (uint32_t)((t.n + 31) / 32);
  // Slices of the points per query tile: enough waves to fill the device,
  // at least 8 tiles per slice.
  uint64_t slices = std::max<uint64_t>(1, (4096 + qtiles - 1) / qtiles);
  slices = std::min<uint64_t>({slices, MAX_SLICES, std::max<uint32_t>(1, n_tiles / 8)});
  const uint32_t per = (uint32_t)((n_tiles + slices - 1) / slices);
  slices = (n_tiles + per - 1) / per;

  float* part_a = nullptr;
  uint32_t *part_i = nullptr, *cnt = nullptr, *cand = nullptr;
  uint8_t* flag = nullptr;
  int bad = 0;
  bad |= dev_alloc(&part_a, nq * 2 * slices * KP, s);
  bad |= dev_alloc(&part_i, nq * 2 * slices * KP, s);
  bad |= dev_alloc(&cnt, nq, s);
  bad |= dev_alloc(&cand, nq * CAND_CAP, s);
  bad |= dev_alloc(&flag, nq, s);
  auto release = [&]() {
    for (void* p : {(void*)part_a, (void*)part_i, (void*)cnt, (void*)cand, (void*)flag})
      if (p) (void)hipFreeAsync(p, s);
  };
  if (bad) {
    release();
    msg = "screen scratch allocation failed";
    return 100;
  }
  SArgs sa{};
  sa.rows32 = t.rows32;
  sa.pnorm = t.pnorm;
  sa.q = trav.q;
  sa.part_a = part_a;
  sa.part_i = part_i;
  sa.nq = nq;
  sa.D = (uint32_t)t.d;
  sa.n = (uint32_t)t.n;
  sa.tiles_per_slice = per;
  sa.n_tiles = n_tiles;
  sa.slices = (uint32_t)slices;
  const dim3 grid((uint32_t)qtiles, (uint32_t)slices);
  const int ks = t.d <= 16 ? 8 : t.d <= 32 ? 16 : 32;
  if (ks == 8) k_screen<8><<<grid, 64, 0, s>>>(sa);
  else if (ks == 16) k_screen<16><<<grid, 64, 0, s>>>(sa);
  else k_screen<32><<<grid, 64, 0, s>>>(sa);

  SelArgs se{};
  se.part_a = part_a;
  se.part_i = part_i;
  se.q = trav.q;
  se.cnt = cnt;
  se.cand = cand;
  se.flag = flag;
  se.nq = nq;
  se.D = (uint32_t)t.d;
  se.k = k;
  se.lists = (uint32_t)(2 * slices);
  se.pmax = t.pmax;
  se.cutoff = qp.cutoff;
  se.max_abs = t.max_abs;
  k_screen_select<<<(uint32_t)nq, 64, 0, s>>>(se);

  CheckArgs ca{};
  ca.q = trav.q;
  ca.rows = t.rows;
  ca.idx = t.idx;
  ca.nodes = t.nodes;
  ca.boxes = t.boxes;
  ca.cnt = cnt;
  ca.cand = cand;
  ca.flag = flag;
  ca.out_d = trav.out_d;
  ca.out_i = trav.out_i;
  ca.nq = nq;
  ca.D = (uint32_t)t.d;
  ca.k = k;
  ca.n = (uint32_t)t.n;
  ca.cutoff = qp.cutoff;
  k_screen_recheck<<<(uint32_t)nq, 64, (size_t)k * t.d * 8, s>>>(ca);

  // Flagged queries: the exact traversal.
  KArgs fb = trav;
  fb.order = nullptr;
  fb.only = flag;
  const int rc = qp.screen_only ? 0 : launch_knn_lds(fb, M_L2, s);
  release();
  if (rc) {
    msg = "no fallback kernel";
    return 101;
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    msg = std::string("screen launch: ") + hipGetErrorString(e);
    return 100;
  }
  return 0;
}

}  // namespace kdt
