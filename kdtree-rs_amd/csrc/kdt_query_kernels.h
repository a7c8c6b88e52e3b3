// kdt_query_kernels.h -- exact k-NN traversal kernels for gfx950.
//
// Replaces Tree::query / run / Descent (src/query.rs:29-494) and the leaf
// scans, point and box distances of src/kernel.rs:167-616.
//
// Work split: a group of G lanes (G in {8, 16, 32, 64}, a power of two dividing
// the 64-lane wavefront) owns one query at a time.  The group walks the tree
// together -- every traversal value (node id, cell bound, k-th distance, stack
// pointer) is the same in all G lanes, so a group never diverges internally --
// and spreads each leaf over its lanes: lane o scores point o of the leaf from
// the coordinate-major leaf bucket, so each coordinate is one coalesced load.
//
// The traversal is the reference's depth-first descent made iterative, visit
// for visit: near child first (q[dim] - split <= 0 goes left, src/query.rs:264),
// the far child gated by the split-plane bound and then the box bound, both
// against one bound read after the near subtree (src/query.rs:278-286);
// order_by_box nodes order and gate both children by box distance
// (src/query.rs:296-327).  The per-axis cell bound (CellBound, :156-211) lives
// in LDS; the recursion becomes an explicit stack with four entry kinds:
//   FAR      deferred far child: plane gate, then box gate, then enter
//   BOXFAR   deferred far child of an order_by_box node: box gate, enter
//   BOXNEAR  deferred near child of an order_by_box node: box gate, descend
//   RESTORE  undo of the cell-bound axis set when a far child was entered
// Every inner node on the current path holds at most one deferred entry and
// one RESTORE; a path has at most 32 inner nodes (n < 2^32), so 64 entries
// (32 for trees of depth <= 16).
//
// Because visits, gates and bound updates happen in the reference's order,
// the result is the reference's for every eps, not only eps = 0.  Distances
// use the reference's per-width summation order (plan of src/kernel.rs:244-270):
//   d <= 8  (Packed)   d4 full pairs: (a0+a2)+(a1+a3); d8 full chunks:
//                      ((a0+a1)+(a2+a3))+((a4+a5)+(a6+a7)); all else in order
//   9..16   (Folded)   pairwise hsum of the lane-folded 8-chunk, then the tail
//   >= 17   (Streamed) running sum of per-8-chunk hsums, then the tail
// where "full" means the point lies in a complete chunk of the leaf counted
// from the leaf's first row (src/kernel.rs:447), P = 2 rows for d4, 8 for d8.
//
// The k best (distance, original index) pairs, lexicographic (src/query.rs:
// 462-481), are either one entry per lane (k <= G, RegList) or a sorted array
// in LDS (LdsList, any k that fits).
#pragma once

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <type_traits>

#include "kdt_internal.h"

namespace kdt {

// Stack entries per group: every inner node on the current path holds at
// most one deferred entry and one RESTORE, so 2 x depth entries suffice; the
// build checks every node (k_validate_nodes: preorder child ids, split dims,
// leaf ranges) and that the depth is at most 32, so the traversal reads nodes
// unchecked and never overflows its stack.  The register-list kernels take
// 32 entries for trees of depth <= 16 (less LDS, more resident waves) and
// STK_MAX otherwise; the runtime-width kernel keeps STK_MAX.
constexpr int STK_MAX = 64;

#ifndef KDT_WIDE_WAVES
#define KDT_WIDE_WAVES 4
#endif
constexpr int WIDE_WAVES = KDT_WIDE_WAVES;
// Waves per SIMD the register-list kernels are compiled for.  The descent's
// dependent node loads need many waves to hide them; with enough waves the
// traversal is bound by VALU issue (3-d rows at 8 waves: about 90% of the
// SIMDs' VALU cycles), so occupancy pays as long as the hot loops keep their
// registers: rows of 1..4 coordinates run 8 waves (a few registers of the
// query start-up spill), rows of 8 run 6 (7 or 8 spill in the hot loops), rows
// of 16 take the registers they need (120: 4 waves; at 5 or 6 waves their leaf
// loop spills).
// The FP32-screened kernel for rows of 8 takes 4 waves per SIMD (113 VGPRs,
// no spill; precision workload 132.6 ms vs 134.7 at 5 waves and 242 at 6,
// where 52 registers spill).
#ifndef KDT_SCREEN8_WAVES
#define KDT_SCREEN8_WAVES 4
#endif
template <int DT, bool SCREEN = false>
constexpr int reg_min_waves() {
  return DT >= 1 && DT <= 4 ? 8 : DT == 8 ? (SCREEN ? KDT_SCREEN8_WAVES : 6) : DT >= 32 ? WIDE_WAVES : 1;
}
// Hits in one leaf chunk from which a group inserts them with one bitonic
// merge instead of one by one: for groups of up to 16 lanes (all exchanges DPP
// moves, 14 steps) the merge costs about five single inserts; wider groups
// take more steps, some through the LDS crossbar, and keep 10 (swept 4-16).
template <int G>
constexpr uint32_t bulk_min() { return G <= 16 ? 5u : 10u; }
constexpr uint32_t KIND_FAR = 0, KIND_BOXFAR = 1, KIND_BOXNEAR = 2, KIND_RESTORE = 3;

struct KArgs {
  const DevNode* __restrict__ nodes;
  const double* __restrict__ boxes;
  const double* __restrict__ pts;
  const float* __restrict__ pts32;
  const double* __restrict__ rows;     // tree-order rows, row-major
  const uint32_t* __restrict__ idx;
  const double* __restrict__ q;
  const uint32_t* __restrict__ order;  // processing order of the queries (null: as given)
  const uint8_t* __restrict__ only;    // non-null: run only the queries qi with only[qi] != 0
  const uint32_t* __restrict__ only_count;  // with `only`: how many are set (0: the grid exits at once)
  uint32_t stk;                        // stack entries per group (2 x tree depth, at least 2)
  double* __restrict__ out_d;
  int64_t* __restrict__ out_i;
  uint64_t nq;
  uint32_t D, k, n_points, n_nodes;
  double cutoff, eps_factor, p, inv_p;
  const float* __restrict__ pn32;  // FP32 leaf screen: fl32(|fl32(row)|^2) per tree position
  double pmax;          // FP32 leaf screen: upper bound of |fl32(row)| over the tree
  double screen_shrink; // 1 + 2 (D + 3) 2^-53
  double max_abs;       // max |coordinate| of the tree
};

// ------------------------------------------------------------ metric algebra
template <int M>
__device__ __forceinline__ double reduce_d(double x, double p) {  // Metric::reduce, src/metric.rs:63-71
  if constexpr (M == M_L2) return x * x;
  else if constexpr (M == M_LP) return pow(x, p);
  else return x;
}
template <int M>
__device__ __forceinline__ double restore_d(double x, double inv_p) {  // Metric::restore, :73-80
  if constexpr (M == M_L2) return sqrt(x);
  else if constexpr (M == M_LP) return pow(x, inv_p);
  else return x;
}
template <int M>
__device__ __forceinline__ double fold_d(double dist, double axis) {  // Metric::fold, :82-88
  if constexpr (M == M_LINF) return fmax(dist, axis);
  else return dist + axis;
}
template <int M>
__device__ __forceinline__ double replace_axis_d(double dist, double old, double nw) {  // :90-98
  if constexpr (M == M_LINF) return fmax(dist, nw);
  else return dist - old + nw;
}
// One axis of a leaf distance: |delta|, delta^2 or |delta|^p (axes_lanes,
// src/kernel.rs:22-29; point_dist_lp :207-218).
template <int M>
__device__ __forceinline__ double axis_term(double delta, double p) {
  if constexpr (M == M_L2) return delta * delta;
  else if constexpr (M == M_LP) return pow(fabs(delta), p);
  else return fabs(delta);
}
__device__ __forceinline__ double vmax_d(double a, double b) { return b > a ? b : a; }  // src/simd.rs:12-15
__device__ __forceinline__ double axis_offset_d(double q, double lo, double hi) {      // src/layout.rs:145-149
  return fmax(fmax(lo - q, q - hi), 0.0);
}
__device__ __forceinline__ double hsum8(const double v[8]) {  // src/simd.rs:25-30
  return ((v[0] + v[4]) + (v[2] + v[6])) + ((v[1] + v[5]) + (v[3] + v[7]));
}

// box_dist, src/kernel.rs:565-616.  `lo` points at the node's [lo | hi].
template <int M, int DT>
__device__ __forceinline__ double box_dist_d(const double* q, const double* __restrict__ lo,
                                             uint32_t Drt, double p) {
  const uint32_t D = DT > 0 ? DT : Drt;
  const double* hi = lo + D;
  if constexpr (M == M_L1 || M == M_L2) {
    const uint32_t nch = D / 8;
    double dist = 0.0;
    if (nch) {
      double acc[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll 1
      for (uint32_t ch = 0; ch < nch; ++ch) {
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const uint32_t i = ch * 8 + j;
          const double off = vmax_d(vmax_d(lo[i] - q[i], q[i] - hi[i]), 0.0);
          acc[j] += M == M_L2 ? off * off : fabs(off);
        }
      }
      dist = hsum8(acc);
    }
#pragma unroll
    for (uint32_t i = nch * 8; i < D; ++i) dist += reduce_d<M>(axis_offset_d(q[i], lo[i], hi[i]), p);
    return dist;
  } else if constexpr (M == M_LINF) {
    double worst = 0.0;
#pragma unroll
    for (uint32_t i = 0; i < D; ++i) {
      const double off = axis_offset_d(q[i], lo[i], hi[i]);
      if (off > worst) worst = off;
    }
    return worst;
  } else {
    double dist = 0.0;
    for (uint32_t i = 0; i < D; ++i) {
      const double off = axis_offset_d(q[i], lo[i], hi[i]);
      if (off > 0.0) dist += pow(off, p);
    }
    return dist;
  }
}

// box_dist by a group of G lanes.  For L1/L2 rows of 8 or 16 the lanes own
// the dimensions: lane j forms the reference's lane-fold acc_j (one chunk, or
// two chunks added in order for d = 16) and three xor-shuffles reproduce the
// pairwise hsum ((a0+a4)+(a2+a6)) + ((a1+a5)+(a3+a7)) bit for bit (IEEE
// addition commutes).  Other widths and metrics fall back to box_dist_d.
// Must be called by all G lanes together.
template <int M, int DT, int G>
__device__ __forceinline__ double box_dist_grp(const double* q, const double* s_q, const double* __restrict__ lo,
                                               uint32_t Drt, double p, uint32_t lane) {
  if constexpr ((M == M_L1 || M == M_L2) && (DT == 8 || DT == 16) && G >= 8) {
    const double* hi = lo + DT;
    double v = 0.0;
    if (lane < 8) {
      const double qa = s_q[lane];
      const double o1 = vmax_d(vmax_d(lo[lane] - qa, qa - hi[lane]), 0.0);
      v = M == M_L2 ? o1 * o1 : fabs(o1);
      if constexpr (DT == 16) {
        const double qb = s_q[lane + 8];
        const double o2 = vmax_d(vmax_d(lo[lane + 8] - qb, qb - hi[lane + 8]), 0.0);
        v = v + (M == M_L2 ? o2 * o2 : fabs(o2));
      }
    }
    v = v + __shfl_xor(v, 4, G);
    v = v + __shfl_xor(v, 2, G);
    v = v + __shfl_xor(v, 1, G);
    return __shfl(v, 0, G);
  } else {
    (void)s_q;
    (void)lane;
    return box_dist_d<M, DT>(q, lo, Drt, p);
  }
}

// Reduced distance from q to one point whose coordinate c is at pt[c * sc]
// (sc = L in a coordinate-major leaf bucket, 1 in a row), in the reference's
// order for width D; `full` = the point lies in a complete chunk of its leaf.
template <int M, int DT>
__device__ __forceinline__ double point_dist_d(const double* q, const double* __restrict__ pt,
                                               uint32_t sc, uint32_t Drt, bool full, double p) {
  const uint32_t D = DT > 0 ? DT : Drt;
  auto term = [&](uint32_t c) -> double { return axis_term<M>(pt[(size_t)c * sc] - q[c], p); };
  if constexpr (M == M_LINF) {
    double m = 0.0;
#pragma unroll
    for (uint32_t c = 0; c < D; ++c) {
      const double t = term(c);
      if (t > m) m = t;
    }
    return m;
  } else if constexpr (M == M_LP) {
    double s = 0.0;
    for (uint32_t c = 0; c < D; ++c) s += term(c);
    return s;
  } else {
    if (D <= 8) {
      if (D == 4 && full) {
        const double a0 = term(0), a1 = term(1), a2 = term(2), a3 = term(3);
        return (a0 + a2) + (a1 + a3);
      }
      if (D == 8 && full) {
        double a[8];
#pragma unroll
        for (int c = 0; c < 8; ++c) a[c] = term(c);
        return ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
      }
      double s = term(0);
#pragma unroll
      for (uint32_t c = 1; c < D; ++c) s += term(c);
      return s;
    }
    if (D <= 16) {
      double w[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) w[j] = term(j);
      if (D == 16) {
#pragma unroll
        for (int j = 0; j < 8; ++j) w[j] = w[j] + term(8 + j);
      }
      double s = hsum8(w);
#pragma unroll
      for (uint32_t c = 8; c < (D == 16 ? 8u : D); ++c) s += term(c);
      return s;
    }
    double s = 0.0;
    uint32_t c = 0;
    // One 8-coordinate chunk per iteration, not unrolled: unrolled over 32 or
    // 64 coordinates, the compiler keeps every query coordinate live in
    // registers across the leaf loop (128 VGPRs of them at 64 dimensions).
#pragma unroll 1
    for (; c + 8 <= D; c += 8) {
      double w[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) w[j] = term(c + j);
      s += hsum8(w);
    }
    for (; c < D; ++c) s += term(c);
    return s;
  }
}

// The FP32 leaf screen.  For query q and tree point p, with q~, p~ their rows
// rounded to FP32, m = -2 q~ (exact), qn = fl32(|q~|^2), pn = fl32(|p~|^2)
// (summed in FP64, then rounded once), the screen forms
//   A = fma(m_{D-1}, p~_{D-1}, ... fma(m_0, p~_0, fl32(pn + qn)))
// with D explicit FP32 fused multiply-adds (1 add + D FMAs per point, against
// D subtractions, D multiplications and D - 1 additions for the plain
// difference form).  Error (u = 2^-24, N = |q~| + P, P >= max |p~| over the
// tree): every partial sum is at most N^2 in magnitude, so the D + 1
// roundings and the two rounded norms give
//   |A - |q~ - p~|^2| <= E2 = (D + 3) u N^2 (1 + 2^-10) + (D + 4) 2^-149
// (the last term for results in the subnormal range); rounding the rows costs
//   | |q~ - p~| - |q - p| | <= e1 = u (1 + 2u) N + sqrt(D) 2^-148;
// and the reference's FP64 reduced distance d (src/kernel.rs recipes, no
// fusing) satisfies |q - p|^2 <= s d, s = 1 + 2 (D + 3) 2^-53.  A point the
// reference offers (d <= b, the current bound) therefore has
//   A <= T(b) = (sqrt(s b) + e1)^2 + E2,
// and the screen rejects a point only when A > T(b).  e1 and E2 are per
// query (kept as FP32 values rounded up); T is formed in FP64 when the bound
// changes, widened by 2^-50 (its few FP64 roundings) and rounded up to FP32.
// Rejected points are never offered; the others are rescored in FP64 with the
// reference recipe, so results are identical to the FP64 path.  Coordinates
// above 1e18 in magnitude switch the screen off for the query (FP32 range).
template <int DT>
__device__ __forceinline__ float screen_score32(const float* m32, float qn32, const float* __restrict__ base32,
                                                float pn, uint32_t L, uint32_t o) {
  float acc = pn + qn32;
#pragma unroll
  for (int c = 0; c < DT; ++c) acc = __builtin_fmaf(m32[c], base32[(size_t)c * L + o], acc);
  return acc;
}

// e2 = +inf switches the screen off for the query: every point passes.
__device__ __forceinline__ float screen_t32(double bound, double shrink, float e1, float e2) {
  if (!(bound < INFINITY)) return INFINITY;
  const double r = sqrt(bound * shrink) + (double)e1;
  return __double2float_ru((r * r + (double)e2) * (1.0 + 0x1p-50));
}

// ------------------------------------------------------------------ shuffles
// Shuffles that stay within one 16-lane DPP row are DPP moves, which stay in
// the VALU instead of a round trip through the LDS crossbar: in a group of 16
// lanes the shift up by one (row_shr:1), and in groups of 16 or more lanes the
// xor exchanges with stride 1, 2 (quad_perm) and 8 (row_ror:8).  Other widths
// and strides use __shfl_up / __shfl_xor.  The
// moves carry no old value (bound_ctrl: a lane without a source reads 0), so
// each is one v_mov_b32_dpp with no copy in front; the only such lane is lane 0
// of a row under row_shr:1, and its shifted value is never used (a k list's
// lane 0 never takes the entry above it).
template <int CTRL>
__device__ __forceinline__ uint32_t dpp32(uint32_t x) {
  return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, CTRL, 0xF, 0xF, true);
}
template <int CTRL>
__device__ __forceinline__ double dpp64(double x) {
  const uint64_t b = (uint64_t)__double_as_longlong(x);
  const uint64_t lo = dpp32<CTRL>((uint32_t)b), hi = dpp32<CTRL>((uint32_t)(b >> 32));
  return __longlong_as_double((long long)(lo | (hi << 32)));
}
template <int G, typename T>
__device__ __forceinline__ T shfl_up1(T x) {
  if constexpr (G == 16) {
    if constexpr (sizeof(T) == 8) return dpp64<0x111>(x);
    else return (T)dpp32<0x111>((uint32_t)x);
  } else {
    return __shfl_up(x, 1, G);
  }
}
template <int G, typename T>
__device__ __forceinline__ T shfl_xor_g(T x, uint32_t stride) {
  if constexpr (G >= 16) {
    if (stride == 1 || stride == 2 || stride == 8) {
      if constexpr (sizeof(T) == 8) {
        return stride == 1 ? dpp64<0xB1>(x) : stride == 2 ? dpp64<0x4E>(x) : dpp64<0x128>(x);
      } else {
        const uint32_t u = (uint32_t)x;
        return (T)(stride == 1 ? dpp32<0xB1>(u) : stride == 2 ? dpp32<0x4E>(u) : dpp32<0x128>(u));
      }
    }
  }
  return __shfl_xor(x, stride, G);
}

// ------------------------------------------------------------------ k lists
// One entry per lane (k <= G).  Lane j holds the j-th best.
template <int G>
struct RegList {
  double kd;
  uint32_t ki;
  uint32_t cnt;
  __device__ __forceinline__ void reset() { kd = INFINITY; ki = 0xFFFFFFFFu; cnt = 0; }
  // BestK::consider, src/query.rs:462-481.  Returns the new bound.
  __device__ __forceinline__ double insert(double d, uint32_t x, double bound, uint32_t lane,
                                           uint32_t gshift, uint64_t gmask, const KArgs& a) {
    const bool less = (lane < cnt) & ((kd < d) | ((kd == d) & (ki < x)));
    const uint32_t pos = (uint32_t)__popcll((__ballot(less) >> gshift) & gmask);
    if (pos < a.k) {
      const double ud = shfl_up1<G>(kd);
      const uint32_t ui = shfl_up1<G>(ki);
      const bool at = lane == pos, up = (lane > pos) & (lane < a.k);
      kd = at ? d : up ? ud : kd;
      ki = at ? x : up ? ui : ki;
      if (cnt < a.k) ++cnt;
      if (cnt == a.k) bound = fmin(__shfl(kd, a.k - 1, G), a.cutoff);  // BestK::bound, :453-459
    }
    return bound;
  }
  // Several offers at once: lane o offers (d, x) when `valid`.  The list after
  // BestK::consider of the offers in any order is the lexicographic top-k of
  // (list + offers), so the offers are bitonic-sorted across the group, the
  // lower half of (list, reversed offers) is kept (half-cleaner) and merged
  // into order.  Returns the new bound.
  static constexpr bool kBulk = true;
  // Compare-exchange steps: a lane keeps the smaller of (own, partner) when
  // `lower == up`, else the larger.  "Partner strictly before own" (one
  // lexicographic test, no branch) decides both: the larger is kept when the
  // partner is not before own, which for two equal pairs (only padding
  // entries can be equal: indices are distinct, distances never NaN) swaps
  // identical values.  So take = before(partner, own) XOR (lower != up).
  __device__ __forceinline__ double insert_many(double d, uint32_t x, bool valid, uint32_t nvalid, double bound,
                                                uint32_t lane, const KArgs& a) {
    auto before = [](double d1, uint32_t i1, double d2, uint32_t i2) {
      return (d1 < d2) | ((d1 == d2) & (i1 < i2));
    };
    auto lless = [](double d1, uint32_t i1, double d2, uint32_t i2) {
      return d1 < d2 || (d1 == d2 && i1 < i2);
    };
    double bd = valid ? d : INFINITY;
    uint32_t bi = valid ? x : 0xFFFFFFFFu;
#pragma unroll
    for (uint32_t size = 2; size <= G; size <<= 1) {
#pragma unroll
      for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
        const double pd = shfl_xor_g<G>(bd, stride);
        const uint32_t pi = shfl_xor_g<G>(bi, stride);
        const bool up = (lane & size) == 0 || size == G;
        const bool lower = (lane & stride) == 0;
        const bool take = before(pd, pi, bd, bi) != (lower != up);
        bd = take ? pd : bd;
        bi = take ? pi : bi;
      }
    }
    // lane j: min(list[j], offers[G-1-j]) -- the G smallest of both, bitonic
    const double rd = __shfl(bd, G - 1 - lane, G);
    const uint32_t ri = __shfl(bi, G - 1 - lane, G);
    if (lless(rd, ri, kd, ki)) { kd = rd; ki = ri; }
#pragma unroll
    for (uint32_t stride = G >> 1; stride > 0; stride >>= 1) {
      const double pd = shfl_xor_g<G>(kd, stride);
      const uint32_t pi = shfl_xor_g<G>(ki, stride);
      const bool lower = (lane & stride) == 0;
      const bool take = before(pd, pi, kd, ki) != !lower;
      kd = take ? pd : kd;
      ki = take ? pi : ki;
    }
    cnt = cnt + nvalid < a.k ? cnt + nvalid : a.k;
    if (lane >= cnt) { kd = INFINITY; ki = 0xFFFFFFFFu; }
    if (cnt == a.k) bound = fmin(__shfl(kd, a.k - 1, G), a.cutoff);
    return bound;
  }
  // write_results, src/query.rs:483-493
  template <int M>
  __device__ __forceinline__ void write(uint64_t qi, uint32_t lane, const KArgs& a) const {
    if (lane < a.k) {
      const bool have = lane < cnt;
      // The row offset is formed here, per query: kept opaque so that the
      // compiler does not hoist `out + lane` out of the query loop into two
      // 64-bit registers (at 64 VGPRs they were spilled to scratch at every
      // wave's start: 2^20 3-d queries, 0.4 GB of scratch writes per launch).
      uint64_t o = qi * a.k + lane;
      asm volatile("" : "+v"(o));
      a.out_d[o] = have ? restore_d<M>(kd, a.inv_p) : INFINITY;
      a.out_i[o] = have ? (int64_t)ki : (int64_t)a.n_points;
    }
  }
};

// Sorted array in LDS for any k that fits (one group of 64 lanes per block).
struct LdsList {
  static constexpr bool kBulk = false;
  double* ld;
  uint32_t* li;
  uint32_t cnt;
  __device__ __forceinline__ double insert_many(double, uint32_t, bool, uint32_t, double bound, uint32_t,
                                                const KArgs&) {
    return bound;
  }
  __device__ __forceinline__ void reset() { cnt = 0; }
  __device__ __forceinline__ double insert(double d, uint32_t x, double bound, uint32_t lane,
                                           uint32_t, uint64_t, const KArgs& a) {
    uint32_t pos = 0;
    for (uint32_t j0 = 0; j0 < cnt; j0 += 64) {
      const uint32_t j = j0 + lane;
      bool less = false;
      if (j < cnt) less = ld[j] < d || (ld[j] == d && li[j] < x);
      pos += (uint32_t)__popcll(__ballot(less));
    }
    if (pos >= a.k) return bound;
    const uint32_t ncnt = cnt < a.k ? cnt + 1 : a.k;
    // Shift [pos, ncnt-1) up one slot, highest chunk first: each chunk reads
    // below what it writes and above what the next chunk writes.
    for (int64_t jt = (int64_t)ncnt - 1; jt > (int64_t)pos; jt -= 64) {
      const int64_t j = jt - (int64_t)lane;
      double vd = 0.0;
      uint32_t vi = 0;
      const bool mv = j > (int64_t)pos;
      if (mv) { vd = ld[j - 1]; vi = li[j - 1]; }
      __asm__ volatile("" ::: "memory");
      if (mv) { ld[j] = vd; li[j] = vi; }
      __asm__ volatile("" ::: "memory");
    }
    ld[pos] = d;
    li[pos] = x;
    __asm__ volatile("" ::: "memory");
    cnt = ncnt;
    if (cnt == a.k) bound = fmin(ld[a.k - 1], a.cutoff);
    return bound;
  }
  template <int M>
  __device__ __forceinline__ void write(uint64_t qi, uint32_t lane, const KArgs& a) const {
    for (uint32_t j = lane; j < a.k; j += 64) {
      const bool have = j < cnt;
      a.out_d[qi * a.k + j] = have ? restore_d<M>(ld[j], a.inv_p) : INFINITY;
      a.out_i[qi * a.k + j] = have ? (int64_t)li[j] : (int64_t)a.n_points;
    }
  }
};

// ------------------------------------------------------------------ kernel
// DT: compile-time ndim (0 = runtime a.D).  G: lanes per query.  M: metric.
// SCREEN: FP32 leaf screen + FP64 recheck (L2, DT > 0 only).  LDSLIST: k list
// in LDS (G must be 64, one group per block; dynamic LDS holds the list and,
// for DT == 0, the cell-bound axes).
//
// The descent of Descent::run (src/query.rs:223-232), wave-synchronous.
// Each group runs its own query through exactly the reference's sequence of
// visits, gates and bound updates, but the wavefront advances all of its
// groups phase by phase:
//   1. descend: every group still descending visits one inner node per
//      iteration, until each has reached a leaf or must unwind;
//   2. scan: all groups at a leaf scan it together;
//   3. unwind: every group pops one stack entry per iteration until it has a
//      node to descend into or an empty stack (query done);
// and a group whose query is done writes it and starts its next query at once.
// Grouping the phases keeps one code path active at a time: with groups left
// free to interleave, every iteration would issue the inner-node, leaf and
// unwind code of all groups under divergence masks.  `next(qi)` yields the
// group's next query row (false when there is none).
// Candidates of the FP32 leaf screen wait in a ring of CQ entries per group
// (tree position, and whether it lies in a complete chunk of its leaf).
template <int G>
constexpr uint32_t screen_ring() { return 4u * G; }
constexpr uint32_t SCREEN_DRAIN = 4;

template <int DT, int G, int M, bool SCREEN, typename List, typename Next>
__device__ __forceinline__ void knn_ww(const KArgs& a, List& best, double* axes, double* s_q, double* s_sa,
                                       double* s_sb, uint32_t* s_sn, uint32_t* s_sm, uint32_t* s_cq,
                                       uint8_t* s_cf, Next next) {
  constexpr int DL = DT > 0 ? DT : 1;
  constexpr int ST_DESCEND = 0, ST_SCAN = 1, ST_POP = 2, ST_DONE = 3;
  const uint32_t lane = threadIdx.x & (G - 1);
  const uint32_t gshift = (threadIdx.x & 63) & ~(uint32_t)(G - 1);
  const uint64_t gmask = G == 64 ? ~0ull : ((1ull << G) - 1ull);
  const uint32_t D = DT > 0 ? DT : a.D;
  const double eps = a.eps_factor;

  uint64_t qi = 0;
  const double* qg = a.q;
  // The query's coordinates: in registers for rows of up to 16 coordinates,
  // read from the group's LDS copy for wider rows (64 doubles would take 128
  // VGPRs and force spills at any useful occupancy) and under the FP32 leaf
  // screen (which keeps the FP32 multipliers in registers instead and needs
  // the FP64 coordinates only to rescore the few points that pass).
  constexpr bool Q_IN_LDS = DT >= 32 || SCREEN;
  double qr[Q_IN_LDS ? 1 : DL];
  float m32[SCREEN ? DL : 1];  // FP32 leaf screen: -2 fl32(q)
  const double* q = Q_IN_LDS ? s_q : qr;
  double bound = a.cutoff, cell = 0.0;
  float t32 = INFINITY, qn32 = 0.0f, scr_e1 = 0.0f, scr_e2 = INFINITY;
  uint32_t cq_head = 0, cq_n = 0;  // FP32 screen: the group's candidate ring (group-uniform)
  uint32_t node = 0;
  int sp = 0, state = ST_DONE;
  DevNode nd{};

  auto start_query = [&]() {
    qg = a.q + qi * D;
    double qmax = 0.0;
    if constexpr (DT >= 32) {
#pragma unroll 4
      for (int c = 0; c < DT; ++c) {
        const double v = qg[c];
        s_q[c] = v;
        axes[c] = 0.0;
        qmax = fmax(qmax, fabs(v));
      }
    } else if constexpr (DT > 0) {
      double qn64 = 0.0;
      // A zero formed here, per query: a loop-invariant 0.0 was kept in a
      // register pair across the query loop and spilled at 64 VGPRs.
      double zero = 0.0;
      asm volatile("" : "+v"(zero));
#pragma unroll
      for (int c = 0; c < DT; ++c) {
        const double v = qg[c];
        if constexpr (!Q_IN_LDS) qr[c] = v;
        s_q[c] = v;
        axes[c] = zero;
        if constexpr (SCREEN) {
          const float f = __double2float_rn(v);
          m32[c] = -2.0f * f;
          qn64 += (double)f * (double)f;
        }
        qmax = fmax(qmax, fabs(v));
      }
      if constexpr (SCREEN) {
        qn32 = (float)qn64;
        const double u = 0x1p-24;
        // |q~| from the FP64 sum (low by at most D 2^-53 relative), widened.
        const double N = sqrt(qn64) * (1.0 + 0x1p-45) + a.pmax;
        const double E2 = (DT + 3.0) * u * N * N * (1.0 + 0x1p-10) + (DT + 4.0) * 0x1p-149;
        const double e1 = u * (1.0 + 2.0 * u) * N + sqrt((double)DT) * 0x1p-148;
        scr_e1 = __double2float_ru(e1);
        scr_e2 = __double2float_ru(E2);
      }
    } else {
      q = qg;
      for (uint32_t c = 0; c < D; ++c) axes[c] = 0.0;
    }
    best.reset();
    bound = a.cutoff;  // BestK::bound with an empty list
    cq_head = 0;
    cq_n = 0;
    if constexpr (SCREEN) {
      if (!(qmax <= 1e18 && a.max_abs <= 1e18)) scr_e2 = INFINITY;  // FP32 range: no screen
      t32 = screen_t32(bound, a.screen_shrink, scr_e1, scr_e2);
    }
    // CellBound::start / seed, src/query.rs:170-187
    cell = 0.0;
    if (box_dist_grp<M, DT, G>(q, s_q, a.boxes, D, a.p, lane) != 0.0) {
      for (uint32_t c = 0; c < D; ++c) {
        const double ax = reduce_d<M>(axis_offset_d(q[c], a.boxes[c], a.boxes[D + c]), a.p);
        axes[c] = ax;
        cell = fold_d<M>(cell, ax);
      }
    }
    sp = 0;
    node = 0;
    state = cell * eps <= bound ? ST_DESCEND : ST_DONE;  // Descent::run, src/query.rs:227-229
  };

  bool alive = next(qi);
  if (alive) start_query();
  while (__ballot(alive)) {
    // 1. descend (src/query.rs:248-289, :296-327)
    for (;;) {
      if (alive && state == ST_DESCEND) {
        nd = a.nodes[node];
        const bool leaf = (nd.b & LEAF_BIT) != 0;
        if (leaf) {
          state = ST_SCAN;
        } else {
          const uint32_t dim = nd.b & DIM_MASK;
          const double qd = DT > 0 ? s_q[dim] : qg[dim];
          const double diff = qd - nd.split;
          const bool goleft = diff <= 0.0;
          const uint32_t near = goleft ? node + 1 : nd.a;
          const uint32_t far = goleft ? nd.a : node + 1;
          const double far_axis = reduce_d<M>(fabs(diff), a.p);
          const double old = axes[dim];
          const double far_dist = replace_axis_d<M>(cell, old, far_axis);
          if (nd.b & OBB_BIT) {
            const double nb = box_dist_grp<M, DT, G>(q, s_q, a.boxes + (size_t)near * 2 * D, D, a.p, lane);
            const double fb = box_dist_grp<M, DT, G>(q, s_q, a.boxes + (size_t)far * 2 * D, D, a.p, lane);
            if (nb <= fb) {
              s_sa[sp] = far_axis; s_sb[sp] = far_dist; s_sn[sp] = far; s_sm[sp] = KIND_BOXFAR | (dim << 2);
              ++sp;
              if (nb * eps <= bound) node = near;
              else state = ST_POP;
            } else {
              s_sa[sp] = cell; s_sn[sp] = near; s_sm[sp] = KIND_BOXNEAR;  // (s_sb unused)
              ++sp;
              if (fb * eps <= bound) {  // enter_far, src/query.rs:329-341
                s_sa[sp] = old; s_sn[sp] = 0; s_sm[sp] = KIND_RESTORE | (dim << 2);
                ++sp;
                axes[dim] = far_axis;
                node = far;
                cell = far_dist;
              } else {
                state = ST_POP;
              }
            }
          } else {
            s_sa[sp] = far_axis; s_sb[sp] = far_dist; s_sn[sp] = far; s_sm[sp] = KIND_FAR | (dim << 2);
            ++sp;
            node = near;
          }
        }
      }
      // A group without a query stays in ST_DONE (state changes only under
      // `alive`), so `state == X` alone is the phase test: a ballot of one
      // compare is its lane mask, with no 0/1 value formed and re-tested.
      if (!__ballot(state == ST_DESCEND)) break;
    }
    // 2. scan_leaf, src/query.rs:343-353; leaf scans, src/kernel.rs:359-563
    // The group's offers go in lane order; after each insert an offer stays
    // pending while it is not yet taken and still within the new bound.
    // `gm` is the group's window of the wave's ballot of pending offers;
    // `pos` is the offered point's tree position.
    using GM = typename std::conditional<G == 64, uint64_t, uint32_t>::type;
    auto offer = [&](bool hit, double dist, uint32_t pos) {
      bool pending = hit && dist <= bound;
      GM gm = (GM)((__ballot(pending) >> gshift) & gmask);
      if (!gm) return;
      const uint32_t myidx = hit ? a.idx[pos] : 0u;
      if constexpr (List::kBulk) {
        const uint32_t nv = (uint32_t)__popcll((uint64_t)gm);
        if (nv >= bulk_min<G>()) {
          bound = best.insert_many(dist, myidx, pending, nv, bound, lane, a);
          if constexpr (SCREEN) t32 = screen_t32(bound, a.screen_shrink, scr_e1, scr_e2);
          return;
        }
      }
      do {
        const uint32_t f = G == 64 ? (uint32_t)__ffsll((unsigned long long)gm) - 1u
                                   : (uint32_t)__ffs((uint32_t)gm) - 1u;
        const double df = __shfl(dist, (int)f, G);
        const uint32_t xf = __shfl(myidx, (int)f, G);
        bound = best.insert(df, xf, bound, lane, gshift, gmask, a);
        pending = pending & (lane != f) & (dist <= bound);
        gm = (GM)((__ballot(pending) >> gshift) & gmask);
      } while (gm);
      if constexpr (SCREEN) t32 = screen_t32(bound, a.screen_shrink, scr_e1, scr_e2);
    };
    // FP32 screen: the next m (<= G) candidates of the ring, one per lane,
    // rescored in FP64 with the reference recipe from the tree-order rows and
    // offered against the current bound.
    auto rescore = [&](uint32_t m) {
      constexpr uint32_t CQ = screen_ring<G>();
      const bool mine = lane < m;
      double dist = INFINITY;
      uint32_t pos = 0;
      if (mine) {
        const uint32_t slot = (cq_head + lane) & (CQ - 1);
        pos = s_cq[slot];
        dist = point_dist_d<M, DT>(q, a.rows + (size_t)pos * D, 1, D, s_cf[slot] != 0, a.p);
      }
      cq_head = (cq_head + m) & (CQ - 1);
      cq_n -= m;
      offer(mine, dist, pos);
    };
    if (alive && state == ST_SCAN) {
      const uint32_t start = nd.a;
      const uint32_t L = (uint32_t)__double_as_longlong(nd.split) - start;
      const uint32_t full = D == 4 ? (L & ~1u) : D == 8 ? (L & ~7u) : L;
      if constexpr (SCREEN) {
        // FP32 scores for every point; the points that pass join the group's
        // ring and are rescored G at a time, so one lane's pass no longer makes
        // the whole group form FP64 distances.  Offers come later than in the
        // reference, against a bound that is never lower than it would have
        // been, which with eps = 0 (the only case these kernels run) changes
        // no result: the returned list is the lexicographic top k of the
        // points visited within the cutoff, every point the screen rejects
        // lies above the bound of its leaf and hence above the final k-th
        // distance, and a larger bound only visits more nodes.
        constexpr uint32_t CQ = screen_ring<G>();
        const float* base32 = a.pts32 + (size_t)start * D;
        for (uint32_t c0 = 0; c0 < L; c0 += 2 * G) {
#pragma unroll
          for (uint32_t h = 0; h < 2; ++h) {
            const uint32_t o = c0 + h * G + lane;
            const bool pass =
                o < L && screen_score32<DL>(m32, qn32, base32, a.pn32[start + o], L, o) <= t32;
            const GM gm = (GM)((__ballot(pass) >> gshift) & gmask);
            if (pass) {
              const uint32_t rank = (uint32_t)__popcll((uint64_t)gm & ((1ull << lane) - 1ull));
              const uint32_t slot = (cq_head + cq_n + rank) & (CQ - 1);
              s_cq[slot] = start + o;
              s_cf[slot] = o < full ? 1 : 0;
            }
            cq_n += (uint32_t)__popcll((uint64_t)gm);
          }
          while (cq_n >= (uint32_t)G) rescore(G);  // at most G - 1 stay: the ring never overflows
        }
        // A leaf that leaves SCREEN_DRAIN or more candidates rescoring them at
        // once keeps the bound from lagging far behind (precision workload:
        // 128.7 ms per call, 132.7 ms draining only full batches of G, 130.1
        // draining every leaf).
        if (cq_n >= SCREEN_DRAIN) rescore(cq_n);
      } else {
        const double* base = a.pts + (size_t)start * D;
        auto score = [&](uint32_t o, double& dist) -> bool {
          if (o >= L) return false;
          dist = point_dist_d<M, DT>(q, base + o, L, D, o < full, a.p);
          return dist <= bound;
        };
        for (uint32_t c0 = 0; c0 < L; c0 += 2 * G) {
          const uint32_t o0 = c0 + lane, o1 = c0 + G + lane;
          double d0 = INFINITY, d1 = INFINITY;
          const bool h0 = score(o0, d0);
          const bool h1 = score(o1, d1);
          offer(h0, d0, start + o0);
          offer(h1, d1, start + o1);
        }
      }
      state = ST_POP;
    }
    // 3. unwind to the next deferred visit.  RESTORE entries and far children
    // whose plane bound already fails need no box distance: each group
    // consumes those in its own inner loop, so one wave iteration costs at
    // most one box gate per group.  Order and bound reads are unchanged.
    for (;;) {
      bool test = false;
      uint32_t en = 0, kind = 0, edim = 0;
      double ea = 0.0, eb = 0.0;
      if (alive && state == ST_POP) {
        for (;;) {
          if (sp == 0) {
            state = ST_DONE;
            break;
          }
          --sp;
          const uint32_t meta = s_sm[sp];
          kind = meta & 3u;
          edim = meta >> 2;
          ea = s_sa[sp];
          if (kind == KIND_RESTORE) {
            axes[edim] = ea;
            continue;
          }
          eb = s_sb[sp];  // written for FAR and BOXFAR entries only (BOXNEAR leaves it unused)
          if (kind == KIND_FAR && !(eb * eps <= bound)) continue;
          en = s_sn[sp];
          test = true;
          break;
        }
      }
      if (test) {
        if (box_dist_grp<M, DT, G>(q, s_q, a.boxes + (size_t)en * 2 * D, D, a.p, lane) * eps <= bound) {
          if (kind == KIND_BOXNEAR) {
            node = en;
            cell = ea;
          } else {  // enter_far, src/query.rs:329-341
            s_sa[sp] = axes[edim]; s_sn[sp] = 0; s_sm[sp] = KIND_RESTORE | (edim << 2);
            ++sp;
            axes[edim] = ea;
            node = en;
            cell = eb;
          }
          state = ST_DESCEND;
        }
      }
      if (!__ballot(state == ST_POP)) break;
    }
    // 4. finished queries: results out, next query in
    if (alive && state == ST_DONE) {
      if constexpr (SCREEN) {
        if (cq_n) rescore(cq_n);  // the last screen candidates
      }
      best.template write<M>(qi, lane, a);
      alive = next(qi);
      if (alive) start_query();
    }
  }
}

template <int DT, int G, int M, bool SCREEN, bool LDSLIST>
__device__ __forceinline__ void knn_body(const KArgs& a, double* axes, double* s_q, double* s_sa,
                                         double* s_sb, uint32_t* s_sn, uint32_t* s_sm, uint32_t* s_cq,
                                         uint8_t* s_cf, double* l_d, uint32_t* l_i) {
  constexpr int GPB = 256 / G;
  const uint32_t grp = threadIdx.x / G;
  const uint32_t gpb = LDSLIST ? 1u : (uint32_t)GPB;
  const uint64_t ngroups = (uint64_t)gridDim.x * gpb;
  typename std::conditional<LDSLIST, LdsList, RegList<G>>::type best;
  if constexpr (LDSLIST) { best.ld = l_d; best.li = l_i; }
  // Workgroups are handed to the 8 XCDs in turn (workgroup b to XCD b mod 8),
  // each XCD with its own L2: XCD x gets a contiguous eighth of the blocks, so
  // that neighbouring queries (large batches run in leaf order) share one L2
  // (C3 117 -> 103 ms, C2 -2%).  A bijection on the blocks below the largest
  // multiple of 8; the rest keep their index.
  const uint32_t full = gridDim.x & ~7u;
  const uint32_t blk = blockIdx.x < full ? (blockIdx.x & 7u) * (full >> 3) + (blockIdx.x >> 3) : blockIdx.x;
  uint64_t qs = (uint64_t)blk * gpb + grp;
  if (a.only && a.only_count && *a.only_count == 0) return;  // nothing left to the traversal (uniform)
  bool first = true;
  auto next = [&](uint64_t& qi) -> bool {
    for (;;) {
      if (!first) qs += ngroups;
      first = false;
      if (qs >= a.nq) return false;
      qi = a.order ? (uint64_t)a.order[qs] : qs;
      if (!a.only || a.only[qi]) return true;
    }
  };
  knn_ww<DT, G, M, SCREEN>(a, best, axes, s_q, s_sa, s_sb, s_sn, s_sm, s_cq, s_cf, next);
}

// STK: stack entries per group, 32 for trees of depth <= 16, else 64.
template <int DT, int G, int M, bool SCREEN, int STK>
__global__ __launch_bounds__(256, (reg_min_waves<DT, SCREEN>())) void knn_reg_kernel(KArgs a) {
  constexpr int GPB = 256 / G;
  constexpr int DL = DT > 0 ? DT : 1;
  __shared__ double s_sa[GPB][STK];
  __shared__ double s_sb[GPB][STK];
  __shared__ uint32_t s_sn[GPB][STK];
  __shared__ uint32_t s_sm[GPB][STK];
  __shared__ double s_ax[GPB][DL];
  __shared__ double s_q[GPB][DL];
  constexpr uint32_t CQ = SCREEN ? screen_ring<G>() : 1u;
  __shared__ uint32_t s_cq[GPB][CQ];
  __shared__ uint8_t s_cf[GPB][CQ];
  const uint32_t grp = threadIdx.x / G;
  knn_body<DT, G, M, SCREEN, false>(a, s_ax[grp], s_q[grp], s_sa[grp], s_sb[grp], s_sn[grp], s_sm[grp],
                                    s_cq[grp], s_cf[grp], nullptr, nullptr);
}

// Runtime ndim, k list in LDS; one 64-lane group per block.  Dynamic LDS:
// [axes D x f64][list k x f64][list k x u32].
template <int M>
__global__ __launch_bounds__(64) void knn_lds_kernel(KArgs a) {
  extern __shared__ double dyn[];
  __shared__ double s_sa[STK_MAX];
  __shared__ double s_sb[STK_MAX];
  __shared__ uint32_t s_sn[STK_MAX];
  __shared__ uint32_t s_sm[STK_MAX];
  double* axes = dyn;
  double* l_d = dyn + a.D;
  uint32_t* l_i = reinterpret_cast<uint32_t*>(l_d + a.k);
  knn_body<0, 64, M, false, true>(a, axes, nullptr, s_sa, s_sb, s_sn, s_sm, nullptr, nullptr, l_d, l_i);
}

// Dispatch entry points, one per metric translation unit.
int launch_knn_l1(const KArgs& a, int G, hipStream_t s);
int launch_knn_l2(const KArgs& a, int G, bool screen, hipStream_t s);
int launch_knn_linf(const KArgs& a, int G, hipStream_t s);
int launch_knn_lds(const KArgs& a, int metric, hipStream_t s);

template <int DT, int G, int M, bool SCREEN>
inline void launch_reg(const KArgs& a, hipStream_t s) {
  constexpr int GPB = 256 / G;
  // One group per query (measured faster than a resident grid that hands
  // each group several queries); a larger batch loops (knn_body's next()).
  uint64_t blocks = std::min<uint64_t>((a.nq + GPB - 1) / GPB, 1u << 30);
  // Only the queries a screen left over (a.only): usually none or a few, so
  // the groups of 2048 workgroups (about what the device holds at once) each
  // look through a share of the flags; a launch with nothing flagged then
  // starts few workgroups beside the other streams' work.
  if (a.only && blocks > (1u << 11)) blocks = 1u << 11;
  KernelTimer tm(s);
  if (a.stk <= 32) knn_reg_kernel<DT, G, M, SCREEN, 32><<<(uint32_t)blocks, 256, 0, s>>>(a);
  else knn_reg_kernel<DT, G, M, SCREEN, STK_MAX><<<(uint32_t)blocks, 256, 0, s>>>(a);
  char name[48];
  std::snprintf(name, sizeof name, "knn_reg<%d,%d,%d%s>", DT, G, M, SCREEN ? ",fp32" : "");
  tm.stop(name, (double)a.nq);
}

// Widths with a compiled kernel: the reference's own monomorphised list
// (src/build.rs:60); other widths run the runtime-ndim kernel.
template <int G, int M, bool SCREEN>
inline int launch_reg_width(const KArgs& a, hipStream_t s) {
  static_assert(G == 8 || G == 16 || G == 32 || G == 64, "group width");
  switch (a.D) {
    case 1: launch_reg<1, G, M, SCREEN>(a, s); return 0;
    case 2: launch_reg<2, G, M, SCREEN>(a, s); return 0;
    case 3: launch_reg<3, G, M, SCREEN>(a, s); return 0;
    case 4: launch_reg<4, G, M, SCREEN>(a, s); return 0;
    case 8: launch_reg<8, G, M, SCREEN>(a, s); return 0;
    case 16: launch_reg<16, G, M, SCREEN>(a, s); return 0;
    default: break;
  }
  // Wide rows (no FP32 leaf buckets, so no screened instances; no 8-lane
  // groups): the Streamed recipe of point_dist_d, compiled for the width.
  if constexpr (!SCREEN && G >= 16) {
    switch (a.D) {
      case 32: launch_reg<32, G, M, false>(a, s); return 0;
      case 64: launch_reg<64, G, M, false>(a, s); return 0;
      default: break;
    }
  }
  return -1;
}

}  // namespace kdt
