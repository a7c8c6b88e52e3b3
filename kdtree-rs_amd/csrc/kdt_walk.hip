// kdt_walk.hip -- exact k-NN for rows of 1-4, 8 or 16 coordinates (p = 2,
// eps = 0, k <= 32) by one tree walk per wavefront of 64 queries, an FP32
// screen of the points it reaches with a proven margin, then the FP64 recheck
// and visit certificate of finish_candidates (kdt_screen.hip).
//
// Replaces, for batches of queries, the per-query descent of
// src/query.rs:234-341 and its leaf scans (src/kernel.rs:244-563) without
// changing any result.  At 8 dimensions the reference's descent makes about
// 700 node visits and 270 box tests per query to scan about 8,000 points
// (Gaussian mixture, 4M points, k = 32), and on the GPU that control flow,
// not the distance arithmetic, is what the per-query traversal spends its time
// on.  Batches run in leaf order, so the 64 queries of a wavefront are
// neighbours: they walk the tree together, once, and every point the walk
// reaches is scored for all 64 of them at once (lane = query, the point read
// by wave-uniform loads).  The walk reaches more points per query than the
// reference visits (about 45,000 instead of 8,000 on that workload), but each
// costs one FP32 add and D fused multiply-adds in the lane and no control flow.
//
// Per lane (query q), in the tree's frame (centre c, kdt_build.hip): the
// screen value of tree point p is
//   A = fma(m_{D-1}, p~_{D-1}, ... fma(m_0, p~_0, fl32(pn + qn))),
// p~ = fl32(fl64(p - c)) (the build's rows32c), q~ likewise, m = -2 q~ (exact),
// pn, qn = fl32 of the FP64 sums of the squares (pn32c for the points).
// Error: with a = 2^-24 + 2^-52 (the rounding of fl64(x - c), then of fl32,
// relative to |x - c|), Q >= |q - c|, P_c >= |p - c| over the tree, t =
// 2 sqrt(D) 2^-150 (values in the FP32 subnormal range),
//   | |q~ - p~| - |q - p| | <= e1 = a (Q + P_c) + t,
//   N = (1 + a)(Q + P_c) + t >= |q~| + |p~|,
//   |A - |q~ - p~|^2| <= E2 = (D + 3) 2^-24 N^2 (1 + 2^-10) + (D + 4) 2^-149
// (every partial sum of the chain is at most N^2; the D + 1 roundings and the
// two rounded norms), and the reference's FP64 reduced distance d satisfies
// d / s <= |q - p|^2 <= s d, s = 1 + 2 (D + 3) 2^-53.  Hence
//   d <= U(A) = s (sqrt(A + E2) + e1)^2            (every point), and
//   d <= b  =>  A <= T(b) = (sqrt(s b) + e1)^2 + E2.
// (e1 and E2 are rounded up to FP32; T(cutoff) is formed in FP64 and widened
// by 2^-50, then rounded up to FP32: screen_t32.)  After an insert the lane
// needs T(U(A_k)) and TB(U(A_k)) (box test, below) only as upper bounds, and
// forms them in FP32 from one square root:
//   x = fl(A + E2) (1 + 2^-22),  r = sqrt(x) (1 + 2^-21),
//   t = ((r + e1) + e1) (1 + 2^-21),
//   T' = (t^2 + E2 + 2^-147)(1 + 2^-21),
//   TB' = (t^2 (1 + (D + 3) u) + (D + 1) 2^-149 + 2^-147)(1 + 2^-21),
// every product and sum rounded to nearest: each factor 1 + 2^-21 (eight
// FP32 roundings) covers the roundings of its stage and the sqrt's (at most
// four units in the last place: v_sqrt_f32 is within one), and sqrt(s U) + e1
// = s (sqrt(A + E2) + e1) + e1 <= t (s < 1 + 2^-47); the 2^-147 covers the
// absolute rounding (2^-150 per operation) of a subnormal t^2.  So T' >= T(U(A_k)) and TB' >= TB(U(A_k)): a larger
// threshold only appends or visits more.
//
// Box test.  The build stores every node box in the frame as FP32 rounded
// outwards (lo32 <= lo - c, hi32 >= hi - c; boxes32c), and the lane forms
//   bd32 = sum_j fl(g_j^2),  g_j = max(fl(lo32_j - q~_j), fl(q~_j - hi32_j), 0)
// (fl(lo32_j - q~_j) = fma(0.5, m_j, lo32_j): m = -2 q~ exactly).  With
// delta = q~ - (q - c) (|delta| <= e1 as for the rows) and gap the exact gap
// vector of q to the box, g_j <= (1 + u)(gap_j + |delta_j|), so
//   bd32 <= (1 + (D + 3) u) (sqrt(bd) + e1)^2 + (D + 1) 2^-149,
// bd the exact box distance; a point p inside has bd <= |q - p|^2 <= s d.  A
// node holding a point with d <= b therefore has bd32 <= TB(b) =
// (1 + (D + 3) u)(sqrt(s b) + e1)^2 + (D + 1) 2^-149 (formed like T, from
// the same t after an insert, rounded up to FP32).
//
// The lane keeps the K smallest A it has seen (la, K >= k), its bound
// b = min(U(la[k-1]), cutoff) -- at least k points have d <= b, so b is never
// below the k-th reference distance d_k -- and appends every reached point
// with A <= T(b) to its buffer.  The wave skips a node when no lane has
// bd32 <= TB(b).  b only decreases.  So every point of the
// reference's top-k (d <= d_k <= b at all times) lies in a node the walk
// enters, is appended, and has A <= T(b_final): the candidates -- the appended
// points with A <= T(b_final) -- contain the top-k, and the recheck's FP64
// top-k of the candidates is the top-k over the whole tree.  The certificate
// then shows that the reference visits each of those points (with eps = 0 its
// answer is the lexicographic top-k of the points it visits), so the result is
// the reference's, bit for bit.  A lane whose buffer fills is compacted
// (entries above the current T dropped: they cannot be final candidates); a
// lane with more than 64 candidates, an overflowing buffer or a query outside
// the FP32 range is flagged and the traversal answers it.

#include <algorithm>
#include <cmath>
#include <mutex>

#include "kdt_query_kernels.h"

namespace kdt {
namespace {

constexpr uint32_t WALK_CB = 256;    // buffer entries per lane (compacted when full)
constexpr uint32_t WALK_CAND = 64;   // candidates per query: the recheck's 64 lanes
constexpr uint32_t WALK_KMAX = 32;   // largest k on this path
constexpr int WALK_WAVES = 4;        // independent wavefronts per workgroup
constexpr uint32_t WALK_LP = 64;     // leaf points staged in LDS at a time
// Rows from which the walk is taken without the explicit flag.  Below 8
// coordinates a query's k-NN ball is small beside a tile's extent, so the
// tile's union costs more than the per-query traversal saves: uniform 3-d
// rows, 2^20 points and queries, k = 16: walk 9.0 ms + recheck 4.0 ms against
// 2.6 ms for the FP64 traversal.
constexpr uint64_t WALK_MIN_DIM = 8;

// Waves per SIMD the walk is compiled for.  Rows of up to 8 coordinates: 4
// (k <= 32: 128 VGPRs; the few values that do not fit are the list's starting
// entries, reloaded once per tile, outside the walk's loops); rows of 16: as
// many as the registers allow.
#ifndef KDT_WALK_MIN_WAVES
#define KDT_WALK_MIN_WAVES 4
#endif
template <int D>
constexpr int walk_min_waves() { return D <= 8 ? KDT_WALK_MIN_WAVES : 1; }

struct WParams {
  uint64_t nq, n_tiles, n_waves;
  uint32_t k;
  double pc, cutoff, shrink;
};

// Inserts A into the ascending list la (the largest drops out): entry e
// becomes the median of (la[e-1], la[e], A), i.e. A clamped to
// [la[e-1], la[e]] -- one v_med3_f32 per entry, from the old values (e runs
// downwards).  A = +inf leaves the list as it is.  The list starts as K - k
// entries -inf and k entries +inf, so that its last entry is always the k-th
// smallest value inserted (a constant index: a runtime one would put the list
// in scratch memory).
template <int K>
__device__ __forceinline__ void list_insert(float (&la)[K], float A) {
#pragma unroll
  for (int e = K - 1; e > 0; --e) la[e] = __builtin_amdgcn_fmed3f(la[e - 1], la[e], A);
  la[0] = fminf(la[0], A);
}

// One wavefront per tile of 64 consecutive queries of the processing order,
// a persistent grid of waves drawing tiles from per-XCD counters.  Pointers are kernel
// arguments (not a struct) so that the compiler knows they do not alias and
// reads the tree through wave-uniform scalar loads.
template <int D, int K>
__global__ __launch_bounds__(64 * WALK_WAVES, walk_min_waves<D>()) void k_walk(
    const DevNode* __restrict__ nodes, const float* __restrict__ boxes32c, const float* __restrict__ rows32c,
    const float* __restrict__ pn32c, const double* __restrict__ frame, const double* __restrict__ q,
    const uint32_t* __restrict__ order, uint2* __restrict__ buf, uint32_t* __restrict__ cnt_out,
    uint32_t* __restrict__ cand, uint8_t* __restrict__ flag, uint32_t* __restrict__ nflag,
    uint32_t* __restrict__ tile_ctr, WParams p) {
  __shared__ double s_rep[WALK_WAVES][D];  // the tile's representative query (near-child choice)
  __shared__ float s_pts[WALK_WAVES][WALK_LP * D];  // a staged piece of a leaf: rows, then
  __shared__ float s_pn[WALK_WAVES][WALK_LP];       // their norms
  const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint64_t wave = (uint64_t)blockIdx.x * WALK_WAVES + w;
  if (wave >= p.n_waves) return;  // wave-uniform; the kernel has no barrier
  uint2* lb = buf + wave * WALK_CB * 64 + lane;  // entry e of this lane at lb[e * 64]
  const double shrink = p.shrink;

  // Tiles are taken from one counter per XCD: XCD x (workgroups go to the 8
  // XCDs in turn, workgroup b to XCD b mod 8) owns the x-th eighth of the
  // tiles, so that its L2 holds one region of the tree, and its waves draw
  // them one at a time (tiles in dense regions cost more); a wave whose XCD's
  // eighth is done helps the others, so every tile is taken whatever the grid.
  const uint32_t xcd = blockIdx.x & 7u;
  uint32_t xi = 0;
  auto next_tile = [&](uint64_t& tile) -> bool {
    while (xi < 8) {
      const uint32_t x = (xcd + xi) & 7u;
      const uint64_t lo = p.n_tiles * x / 8, hi = p.n_tiles * (x + 1) / 8;
      uint32_t t = 0;
      if (lane == 0) t = atomicAdd(tile_ctr + x, 1u);
      t = (uint32_t)__builtin_amdgcn_readfirstlane((int)t);
      if (lo + t < hi) {
        tile = lo + t;
        return true;
      }
      ++xi;
    }
    return false;
  };

  uint64_t tile = 0;
  while (next_tile(tile)) {
    const uint64_t slot = tile * 64 + lane;
    const bool live = slot < p.nq;
    const uint64_t qi = live ? (order ? (uint64_t)order[slot] : slot) : 0;
    float m[D];
    double qn64 = 0.0, qc2 = 0.0;
    const uint32_t rep = (uint32_t)__popcll(__builtin_amdgcn_ballot_w64(live)) / 2;  // the middle live lane
#pragma unroll
    for (int j = 0; j < D; ++j) {
      const double qj = q[qi * D + j];
      if (lane == rep) s_rep[w][j] = qj;
      const double y = qj - frame[j];
      const float f = (float)y;
      m[j] = -2.0f * f;
      qn64 += (double)f * (double)f;
      qc2 += y * y;
    }
    const float qn = (float)qn64;
    // Margins (header): Q = |q - c| from the FP64 sum, widened; all rounded up.
    const double Q = sqrt(qc2) * (1.0 + 0x1p-44);
    const double ar = 0x1p-24 + 0x1p-52;
    const double tt = 2.0 * sqrt((double)D) * 0x1p-150;
    const double e1 = (ar * (Q + p.pc) + tt) * (1.0 + 0x1p-50);
    const double N = ((1.0 + ar) * (Q + p.pc) + tt) * (1.0 + 0x1p-50);
    const double E2 = (D + 3.0) * 0x1p-24 * N * N * (1.0 + 0x1p-10) + (D + 4.0) * 0x1p-149;
    const float e1f = __double2float_ru(e1), e2f = __double2float_ru(E2);
    bool ok = live && Q <= 1e18;  // FP32 range (the tree's P_c <= 1e18 is checked by the host)
    float la[K];
#pragma unroll
    for (int e = 0; e < K; ++e) la[e] = (uint32_t)e + p.k < (uint32_t)K ? -INFINITY : INFINITY;
    const double tbf = 1.0 + (D + 3.0) * 0x1p-24;   // TB's rounding factor
    const double tbt = (D + 1.0) * 0x1p-149;        // and its subnormal term
    constexpr float tbf32 = 1.0f + (D + 3.0f) * 0x1p-24f, tbt32 = (D + 1.0f) * 0x1p-149f;  // exact in FP32
    float T = -INFINITY, TB = -INFINITY;
    if (ok) {  // the cutoff's thresholds (FP64, rounded up)
      T = screen_t32(p.cutoff, shrink, e1f, e2f);
      if (p.cutoff < INFINITY) {
        const double t = sqrt(p.cutoff * shrink) + (double)e1f;
        TB = __double2float_ru((tbf * t * t + tbt) * (1.0 + 0x1p-50));
      } else {
        TB = INFINITY;
      }
    }
    uint32_t nb = 0;
    __builtin_amdgcn_wave_barrier();  // s_rep written above

    // The lane's thresholds from its list (after an insert): FP32, one square
    // root (header).  They only decrease: T and TB of the cutoff bound them.
    auto update = [&]() {
      const float ak = la[K - 1];  // the k-th smallest A so far (+inf until k values)
      if (ak < INFINITY) {
        constexpr float up1 = 1.0f + 0x1p-22f, up2 = 1.0f + 0x1p-21f;
        const float x = (fmaxf(ak, 0.0f) + e2f) * up1;
        const float r = __builtin_amdgcn_sqrtf(x) * up2;
        const float t = ((r + e1f) + e1f) * up2;
        const float tt2 = t * t;
        T = fminf(T, (tt2 + e2f + 0x1p-147f) * up2);
        TB = fminf(TB, (tt2 * tbf32 + tbt32 + 0x1p-147f) * up2);
      }
    };
    // A reached point: append under the threshold, insert into the list.
    auto offer = [&](float A, uint32_t pos) {
      const bool app = A <= T;  // NaN (padding) and lanes without a query (T = -inf): no
      if (__builtin_amdgcn_ballot_w64(app)) {
        if (app) {
          lb[(size_t)nb * 64] = make_uint2(__float_as_uint(A), pos);
          ++nb;
        }
        const float Ai = app && A < la[K - 1] ? A : INFINITY;
        if (__builtin_amdgcn_ballot_w64(Ai < INFINITY)) {
          list_insert<K>(la, Ai);
          if (Ai < INFINITY) update();
        }
        if (__builtin_amdgcn_ballot_w64(nb >= WALK_CB)) {
          if (nb >= WALK_CB) {
            // Compaction: entries above the current threshold can never be
            // final candidates (T only decreases).
            uint32_t keep = 0;
            for (uint32_t e = 0; e < WALK_CB; ++e) {
              const uint2 v = lb[(size_t)e * 64];
              if (__uint_as_float(v.x) <= T) {
                lb[(size_t)keep * 64] = v;
                ++keep;
              }
            }
            nb = keep;
            if (nb > WALK_CB / 2) {  // still crowded: leave the query to the traversal
              ok = false;
              T = -INFINITY;
              TB = -INFINITY;
            }
          }
        }
      }
    };

    // The walk is wave-uniform; readfirstlane states it, so that the node,
    // box and row reads are scalar loads.
    // The walk is wave-uniform; readfirstlane states it, so that the node,
    // box and row reads are scalar loads.
    uint32_t node = 0, sp = 0, stk = 0;  // stack entry i in lane i's stk (depth <= 32)
    for (;;) {
      node = (uint32_t)__builtin_amdgcn_readfirstlane((int)node);
      const float* bx = boxes32c + (size_t)node * 2 * D;
      float bd = 0.0f;
#pragma unroll
      for (int j = 0; j < D; ++j) {
        const float g = fmaxf(fmaxf(__builtin_fmaf(0.5f, m[j], bx[j]), -__builtin_fmaf(0.5f, m[j], bx[D + j])), 0.0f);
        bd = __builtin_fmaf(g, g, bd);
      }
      if (__builtin_amdgcn_ballot_w64(bd <= TB)) {
        const DevNode nd = nodes[node];
        if (!(nd.b & LEAF_BIT)) {
          const bool left = __builtin_amdgcn_readfirstlane(s_rep[w][nd.b & DIM_MASK] <= nd.split ? 1 : 0) != 0;
          const uint32_t far = left ? nd.a : node + 1;
          node = left ? node + 1 : nd.a;
          stk = lane == sp ? far : stk;
          ++sp;
          continue;
        }
        const uint32_t start = nd.a, end = (uint32_t)__double_as_longlong(nd.split);
        // The leaf in pieces of WALK_LP points: one coalesced load of the
        // piece into the wave's LDS (one memory round trip), then every lane
        // reads each point by a broadcast LDS read.  Four points per step
        // (reads past the piece's end stay inside the staging array and are
        // not offered).
        for (uint32_t o0 = start; o0 < end; o0 += WALK_LP) {
          const uint32_t mp = min(WALK_LP, end - o0);
          float* sp_ = s_pts[w];
          if constexpr (D % 4 == 0) {
            const float4* src = reinterpret_cast<const float4*>(rows32c + (size_t)o0 * D);
            for (uint32_t f = lane; f < mp * (D / 4); f += 64) reinterpret_cast<float4*>(sp_)[f] = src[f];
          } else {
            const float* src = rows32c + (size_t)o0 * D;
            for (uint32_t f = lane; f < mp * D; f += 64) sp_[f] = src[f];
          }
          if (lane < mp) s_pn[w][lane] = pn32c[o0 + lane];
          __builtin_amdgcn_wave_barrier();
          for (uint32_t i0 = 0; i0 < mp; i0 += 4) {
            float A[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
              const float* pr = sp_ + (size_t)(i0 + i) * D;
              float acc = s_pn[w][i0 + i] + qn;
#pragma unroll
              for (int j = 0; j < D; ++j) acc = __builtin_fmaf(m[j], pr[j], acc);
              A[i] = acc;
            }
            // One test for the four points (most groups have no lane under
            // its threshold; min ignores the NaN of padding rows).
            const float amin = fminf(fminf(A[0], A[1]), fminf(A[2], A[3]));
            if (__builtin_amdgcn_ballot_w64(amin <= T)) {
#pragma unroll
              for (int i = 0; i < 4; ++i)
                if (i0 + i < mp) offer(A[i], o0 + i0 + i);
            }
          }
          __builtin_amdgcn_wave_barrier();
        }
      }
      if (sp == 0) break;
      --sp;
      node = (uint32_t)__builtin_amdgcn_readlane((int)stk, (int)sp);
    }

    // Candidates: the appended points with A <= the final threshold.
    bool good = ok;
    uint32_t c = 0;
    if (good) {
      for (uint32_t e = 0; e < nb; e += 4) {
        uint2 v[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) v[i] = e + i < nb ? lb[(size_t)(e + i) * 64] : make_uint2(0x7FC00000u, 0u);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          if (__uint_as_float(v[i].x) <= T) {
            if (c < WALK_CAND) cand[qi * WALK_CAND + c] = v[i].y;
            ++c;
          }
        }
      }
      good = c <= WALK_CAND;
    }
    if (live) {
      cnt_out[qi] = good ? c : 0;
      flag[qi] = good ? 0 : 1;
      if (!good) atomicAdd(nflag, 1u);
    }
  }
}

template <int D, int K>
const void* walk_kernel() {
  return reinterpret_cast<const void*>(k_walk<D, K>);
}

const void* walk_kernel_for(uint64_t d, uint32_t K) {
#define KDT_WALK_K(DD)                                                                            \
  case DD:                                                                                        \
    return K <= 8 ? walk_kernel<DD, 8>() : K <= 16 ? walk_kernel<DD, 16>() : walk_kernel<DD, 32>();
  switch (d) {
    KDT_WALK_K(1)
    KDT_WALK_K(2)
    KDT_WALK_K(3)
    KDT_WALK_K(4)
    KDT_WALK_K(8)
    KDT_WALK_K(16)
    default: return nullptr;
  }
#undef KDT_WALK_K
}

// Wavefronts of the walk the device holds at once, per (width, list length),
// looked up once per device: the grid's size, and with it the buffer scratch.
uint64_t walk_waves(uint64_t d, uint32_t k, uint64_t n_tiles) {
  static std::mutex mu;
  static int cache[64][6][3] = {};
  const int di = d == 1 ? 0 : d == 2 ? 1 : d == 3 ? 2 : d == 4 ? 3 : d == 8 ? 4 : 5;
  const int ki = k <= 8 ? 0 : k <= 16 ? 1 : 2;
  int dev = 0;
  int per_dev = 0;
  if (hipGetDevice(&dev) == hipSuccess && dev >= 0 && dev < 64) {
    std::lock_guard<std::mutex> g(mu);
    int& c = cache[dev][di][ki];
    if (!c) {
      int cus = 0, per = 0;
      if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, walk_kernel_for(d, k), 64 * WALK_WAVES, 0) !=
              hipSuccess ||
          per <= 0)
        per = 1;
      c = cus * per * WALK_WAVES;
    }
    per_dev = c;
  }
  if (per_dev <= 0) per_dev = 1024;
#ifdef KDT_WALK_WAVES_PER_CU
  per_dev = std::min(per_dev, 256 * KDT_WALK_WAVES_PER_CU);
#endif
  return std::max<uint64_t>(1, std::min<uint64_t>(n_tiles, (uint64_t)per_dev));
}

bool walk_width(uint64_t d) { return d == 1 || d == 2 || d == 3 || d == 4 || d == 8 || d == 16; }

}  // namespace

bool walk_applies(const Tree& t, const QueryParams& qp) {
  return qp.fp32_screen && qp.allow_scratch && qp.metric == M_L2 && qp.eps_factor == 1.0 &&
         qp.k <= (uint64_t)WALK_KMAX && t.rows32c && t.pn32c && t.boxes32c && t.frame && walk_width(t.d) &&
         t.scr_pc <= 1e18 && t.n < (1ull << 32) && qp.nq < (1ull << 40) &&
         (qp.walk > 0 || (qp.walk == 0 && qp.nq >= WALK_MIN_QUERIES && t.d >= WALK_MIN_DIM));
}

size_t walk_scratch_bytes(const Tree& t, const QueryParams& qp) {
  const uint64_t nq = qp.nq;
  const uint64_t waves = walk_waves(t.d, (uint32_t)qp.k, (nq + 63) / 64);
  size_t b = scratch_piece(waves * WALK_CB * 64 * 8) + scratch_piece(9 * 4) + scratch_piece(nq * 4) +
             scratch_piece(nq * WALK_CAND * 4) + scratch_piece(nq);
  if (order_applies(qp)) b += scratch_piece(nq * 4) + order_scratch_bytes(t, nq);
  return b;
}

int launch_walk(const Tree& t, const QueryParams& qp, const KArgs& trav, Scratch& sc, hipStream_t s,
                std::string& msg) {
  const uint64_t nq = qp.nq;
  const uint32_t k = (uint32_t)qp.k;
  const uint64_t n_tiles = (nq + 63) / 64;
  const uint64_t waves = walk_waves(t.d, k, n_tiles);
  uint32_t* order = nullptr;
  if (order_applies(qp)) {
    const hipError_t e = sc.get(&order, nq);
    if (e != hipSuccess) {
      msg = std::string("query ordering: ") + hipGetErrorString(e);
      return 100;
    }
    const int rc = order_queries(t, trav.q, nq, order, sc, s, msg);
    if (rc) return rc;
  }
  uint2* buf = nullptr;
  uint32_t *cnt = nullptr, *cand = nullptr, *nflag = nullptr;
  uint8_t* flag = nullptr;
  hipError_t e = sc.get(&buf, waves * WALK_CB * 64);
  if (e == hipSuccess) e = sc.get(&nflag, 9);  // the flagged count, then the 8 tile counters
  if (e == hipSuccess) e = sc.get(&cnt, nq);
  if (e == hipSuccess) e = sc.get(&cand, nq * WALK_CAND);
  if (e == hipSuccess) e = sc.get(&flag, nq);
  if (e != hipSuccess) {
    msg = std::string("walk scratch: ") + hipGetErrorString(e);
    return 100;
  }
  launch_zero_u32(nflag, 9, s);
  WParams wp{};
  wp.nq = nq;
  wp.n_tiles = n_tiles;
  wp.n_waves = waves;
  wp.k = k;
  wp.pc = t.scr_pc;
  wp.cutoff = qp.cutoff;
  wp.shrink = 1.0 + 2.0 * ((double)t.d + 3.0) * 0x1p-53;
  const uint32_t blocks = (uint32_t)((waves + WALK_WAVES - 1) / WALK_WAVES);
  {
    KernelTimer tm(s);
#define KDT_WALK_LAUNCH(DD, KK)                                                                              \
  k_walk<DD, KK><<<blocks, 64 * WALK_WAVES, 0, s>>>(t.nodes, t.boxes32c, t.rows32c, t.pn32c, t.frame, trav.q, \
                                                    order, buf, cnt, cand, flag, nflag, nflag + 1, wp)
#define KDT_WALK_D(DD)                                  \
  case DD:                                              \
    if (k <= 8) KDT_WALK_LAUNCH(DD, 8);                 \
    else if (k <= 16) KDT_WALK_LAUNCH(DD, 16);          \
    else KDT_WALK_LAUNCH(DD, 32);                       \
    break;
    switch (t.d) {
      KDT_WALK_D(1)
      KDT_WALK_D(2)
      KDT_WALK_D(3)
      KDT_WALK_D(4)
      KDT_WALK_D(8)
      KDT_WALK_D(16)
      default:
        msg = "no tile-walk kernel for this width";
        return 101;
    }
#undef KDT_WALK_D
#undef KDT_WALK_LAUNCH
    char name[40];
    std::snprintf(name, sizeof name, "k_walk<%d,%d>", (int)t.d, k <= 8 ? 8 : k <= 16 ? 16 : 32);
    tm.stop(name, (double)nq);
  }
  return finish_candidates(t, qp, trav, cnt, cand, flag, nflag, s, msg);
}

}  // namespace kdt
