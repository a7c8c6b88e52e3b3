// kdt_screen.hip -- exact k-NN for wide rows (32 <= ndim <= 64, and
// 8 <= ndim < 32 on request; p = 2, eps = 0) by an MFMA distance screen, an
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
//   1. k_screen: for a tile of 32 queries and a slice of the points, the
//      matrix cores form  A = |q^|^2 + |p^|^2 - 2 q^.p^  from the rows rounded
//      to fp16 in the tree's frame (one v_mfma_f32_32x32x16_f16 per 16
//      coordinates for -2 q^.p^, one v_mfma_f32_32x32x16_bf16 more that adds
//      the two norms) and each lane keeps the KP = 16 smallest A of the points
//      it sees, with the points' tree positions;
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
// Error bound.  The screen works in a frame fixed per tree at the build:
// centre c (the middle of the root box) and a power of two sigma with
// sigma |fl64(x_j - c_j)| < 2^8 for every row.  A row x becomes
// x^_j = screen_f16(fl64(x_j - c_j), sigma) (fp16 of the FP32 value, round to
// nearest even, fp16 subnormals flushed), so per coordinate
//   |sigma (x_j - c_j) - x^_j| <= a sigma |x_j - c_j| + b,
//   a = 2^-11 (1 + 2^-12),  b = 2^-14 (1 + 2^-60)
// (the FP64 difference adds 2^-53 relative), and with R = sigma |q - p|
// (both rows moved by the same c) the triangle inequality gives
//   | |q^ - p^| - R | <= e1 = a sigma (|q - c| + P_c) + 2 b sqrt(D),
// P_c >= |p - c| for every row p of the tree.  The matrix cores sum in FP32
// the D exact products (-2 q^_j) p^_j (an fp16 product is exact in FP32) and
// the norms qn = fl32(|q^|^2), pn = fl32(|p^|^2) as three bf16 pieces each
// (sum = the norm up to 2^-27 of it); with any order and any rounding of at
// most 2^-23 per addition (D + 5 additions, the pieces' magnitudes at most
// (1 + 2^-8) times their norm),
//   |A - |q^ - p^|^2| <= E2 = ((1 + 2^-7) gamma + 1.13 2^-24) Nh^2 + 2^-118,
//   gamma = (D + 5) 2^-23 / (1 - (D + 5) 2^-23),
//   Nh = sigma (|q - c| + P_c) (1 + a) + 2 b sqrt(D)  >=  |q^| + |p^|.
// The FP64 reduced distance d of the reference satisfies d / s <= |q - p|^2
// <= d s with s = 1 + 2 (D + 3) 2^-53.  So U(A) = s (sqrt(A + E2) + e1)^2 /
// sigma^2 bounds d from above, and a point with d <= tau has A <= tau_A =
// (sigma sqrt(s tau) + e1)^2 + E2.  With tau = min(U(A_k), cutoff): at least
// k points have d <= U(A_k), so the k-th best distance is <= tau and every
// point of the top-k is a candidate.  A query with a coordinate of
// sigma |fl64(q_j - c_j)| above 2^11 is not screened (flagged): below that
// its fp16 values, -2 q^ and |q^|^2 stay in range.

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <mutex>

#include "kdt_query_kernels.h"

namespace kdt {

namespace {

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));

constexpr uint32_t CAND_CAP = 64;  // candidates per query (one per lane of the recheck)
constexpr int KP = 16;             // entries per lane list
constexpr uint32_t KMAX = 32;      // largest k on this path (k > KP takes more point slices, below)
// k_screen_select reads 64 x 16 per-lane lists at most: up to 512 point slices
// (two lists, lane halves, per slice), so that a batch of a few query tiles still
// spreads over enough waves (32 queries: 0.22 ms with 512, 0.36 ms with 128).
constexpr int LISTS_PER_LANE = 16;
constexpr uint32_t MAX_SLICES = 32 * LISTS_PER_LANE;

struct SArgs {
  const uint16_t* __restrict__ rows_h;  // [n_pad][2 KS] fp16, tree order, zero-padded
  const uint2* __restrict__ pnorm_h;    // [n_pad] their norms, norm-block words (+inf on padding rows)
  const uint4* __restrict__ qrows;      // [nq][KS / 4] the queries as -2 q^, fp16 (k_screen_queries)
  const uint2* __restrict__ qnorm;      // [nq] their norm-block words, query side
  float* __restrict__ part_a;        // [nq][2 S][KP] the kept A values, ascending
  uint32_t* __restrict__ part_i;     // [nq][2 S][KP] their tree positions
  // [nq] per query: the seed (k_screen_seed), lowered during the main pass to
  // the smallest published last entry of the query's lists (below); or null:
  // +inf and no sharing.
  float* __restrict__ thr0;
  // [nq] the queries' (e1, E2) (k_screen_queries): a list's k-th entry gives
  // a bound of tau_A at each merge (screen_tau_a32); k = 0 for none.
  const float2* __restrict__ qterm;
  uint32_t k;
  uint64_t nq;
  uint32_t D, n, n_tiles, slices;
  // Slice y takes point tiles y tb_mul + i ts for i < tmax (below n_tiles).
  uint32_t tb_mul, ts, tmax;
};

// Query tiles (32 queries each) per wave: with two, each point operand read
// from LDS feeds two accumulator chains (half the LDS traffic per matrix
// instruction) at the cost of the registers of a second query tile.
#ifndef KDT_SCREEN_QT32
#define KDT_SCREEN_QT32 1
#endif
template <int KS>
constexpr int screen_qt() { return KS <= 16 ? 1 : KDT_SCREEN_QT32; }

// Waves per workgroup sharing each point tile.  Rows of 16 and 32 columns: 4
// (the per-tile barrier waits on fewer waves; 16-d uniform rows 1.8 vs 2.3 ms,
// 16-d normal rows 3% faster; 24-d uniform rows at k = 32 5.1 vs 6.2 ms, 32-d
// normal rows at k = 10 3% slower).  Rows of 64 columns: 8 query tiles per
// workgroup (with 4, each point tile is read twice as often from L2).
#ifndef KDT_SCREEN_SW32
#define KDT_SCREEN_SW32 (8 / KDT_SCREEN_QT32)
#endif
template <int KS>
constexpr int screen_waves() { return KS <= 16 ? 4 : KDT_SCREEN_SW32; }

// Point tiles per pipeline step: the TS tiles of a step are staged together
// and contracted between two barriers, which amortises the barrier and the
// staging round over TS tiles.  Rows of 64 columns: 4 (5 matrix instructions
// per tile; 2^20 points, k = 10, 65,536 queries: 10.44 vs 10.93 ms with 2 and
// 10.63 with 3; with 2 against 1 earlier, 13.07 vs 13.48 ms; 4 fits beside
// two workgroups per CU with the rows loaded straight into LDS and 8 insert
// slots, below).  Rows of 16
// columns: 4 (2 matrix instructions per tile; 2^21 16-d normal rows, k = 10:
// 65,536 queries 13.3 ms vs 14.7 with 1 and 13.9 with 2, 8,192 queries 2.27
// vs 2.60 and 2.39 ms; 2 and 8 again with the rows loaded straight into LDS:
// 1-10% slower or mixed).  Rows of 32 columns: 2 (2^20 32-d normal rows,
// 32,768 queries, k = 10: 4.09 ms vs 4.38-4.57 with 1 and 4.14 with 4).
#ifndef KDT_SCREEN_TS8
#define KDT_SCREEN_TS8 4
#endif
#ifndef KDT_SCREEN_TS16
#define KDT_SCREEN_TS16 2
#endif
#ifndef KDT_SCREEN_TS32
#define KDT_SCREEN_TS32 4
#endif
template <int KS>
constexpr int screen_tiles_per_step() { return KS <= 8 ? KDT_SCREEN_TS8 : KS <= 16 ? KDT_SCREEN_TS16 : KDT_SCREEN_TS32; }

// One step: TS point tiles, each 32 rows of 2 KS fp16 (KS / 4 uint4), staged
// by the whole workgroup of SW = screen_waves waves: thread u loads uint4 u,
// u + 64 SW of the step straight into LDS (stage_dma), and threads
// 0 .. 32 TS - 1 the norm words through registers.
// (Nothing is read past a tile: TU, SU and the clamps bound every index.)
template <int KS>
struct Stage {
  static constexpr int RU = KS / 4;                          // uint4 per row
  static constexpr int TU = 32 * RU;                         // uint4 per tile
  static constexpr int TS = screen_tiles_per_step<KS>();
  static constexpr int SU = TS * TU;                         // uint4 per step
  static constexpr int SW = screen_waves<KS>();
  static constexpr int NT = 64 * SW;                         // threads per workgroup
  static constexpr int PT = (SU + NT - 1) / NT;              // uint4 per thread
  static_assert(PT <= 2, "two loads per thread");
  static_assert(SU % 64 == 0, "whole waves per load");
  static_assert(TS * 32 <= NT, "one thread per split norm of the step");
};

// Row layout in LDS: unpadded, TU uint4 per tile, with the uint4 columns of
// row r stored XOR-ed with swz(r) = (r RU / 16) mod RU, so that the 16 lanes
// of one LDS read cycle (16 rows, one column) hit distinct banks.
template <int KS>
__device__ __forceinline__ uint32_t swz(uint32_t r) {
  constexpr uint32_t RU = KS / 4;
  return (r * RU / 16) % RU;
}

// The step's rows go global -> LDS directly (global_load_lds_dwordx4: lane L
// of a wave writes 16 bytes at the wave's LDS block + 16 L): thread u of the
// workgroup fills LDS uint4 u (and u + NT) with the source uint4 that the
// swizzle puts there.  Loads at clamped (in-bounds) indices: the step's local
// tiles lt0 .. lt0 + TS - 1 are clamped to the slice's last tile t1 - 1 (a
// harmless re-read), and local tile i is global tile tb + i ts.  Completion
// is tracked by vmcnt: the caller waits for it before the barrier that hands
// the stage to its readers.
template <int KS>
__device__ __forceinline__ void stage_dma(const SArgs& a, uint32_t lt0, uint32_t t1, uint32_t tb, uint32_t ts,
                                          uint32_t tid, uint4* rows) {
  using St = Stage<KS>;
  const uint4* src = reinterpret_cast<const uint4*>(a.rows_h);
  auto gtile = [&](uint32_t i) -> size_t { return (size_t)tb + (size_t)min(lt0 + i, t1 - 1) * ts; };
  const uint32_t wbase = tid & ~63u;
#pragma unroll
  for (int p = 0; p < St::PT; ++p) {
    const uint32_t ub = wbase + (uint32_t)p * St::NT;  // wave-uniform (SU is a multiple of 64)
    if (ub < (uint32_t)St::SU) {
      const uint32_t u = tid + (uint32_t)p * St::NT;
      const uint32_t w = u % St::TU, r = w / St::RU, x = w % St::RU;
      const uint4* g = src + gtile(u / St::TU) * St::TU + r * St::RU + (x ^ swz<KS>(r));
      __builtin_amdgcn_global_load_lds(g, (__attribute__((address_space(3))) void*)(rows + ub), 16, 0, 0);
    }
  }
}

template <int KS>
__device__ __forceinline__ void pn_load(const SArgs& a, uint32_t lt0, uint32_t t1, uint32_t tb, uint32_t ts,
                                        uint32_t tid, uint2& pn) {
  using St = Stage<KS>;
  const uint32_t pi = min(tid, (uint32_t)St::TS * 32 - 1);
  pn = a.pnorm_h[((size_t)tb + (size_t)min(lt0 + pi / 32, t1 - 1) * ts) * 32 + pi % 32];
}

// Tile i of the step starts at rows + i TU and pn + i 64: its 32 norm words,
// then 32 zero words (written once per launch) that the lanes of half 1
// read, so every lane forms its norm operand from one LDS read and no select.
template <int KS>
__device__ __forceinline__ void pn_store(const uint2& pnv, uint32_t tid, uint2* pn) {
  using St = Stage<KS>;
  if (tid < (uint32_t)St::TS * 32) pn[(tid / 32) * 64 + tid % 32] = pnv;
}

// One wave: 32 queries (B operand, columns) against 32-row tiles of points
// (A operand, rows); the SW waves of a workgroup take SW consecutive query
// tiles and share each point tile through LDS (one HBM read per workgroup
// instead of one per wave).  Lane l holds query j = l & 31, coordinates
// [h KS, h KS + KS) with h = l >> 5, in both operands; instruction v of the
// product contracts coordinates 8 v .. 8 v + 7 of each half.  The query
// operand is -2 q^ (exact in fp16), so the NS = KS / 8 v_mfma_f32_32x32x16_f16
// sum -2 q^.p^ in FP32; one v_mfma_f32_32x32x16_bf16 more adds the norms
// (lanes of half 0: query side (q1, q2, q3, 1, 1, 1, 0, 0), point side
// (1, 1, 1, p1, p2, p3, 0, 0), the norms' bf16 pieces).  Element r of the
// accumulator is then A for query j and point row i = 8 (r / 4) + 4 h +
// (r % 4), r = 0..15.
//
// A value under the lane's threshold is not inserted at once: it is appended
// to the lane's buffer of IB slots in LDS, and after every CHK elements (8, or
// 4 for rows of 64 columns), if any lane of the wave holds more than IB - CHK
// entries, all lanes merge their
// buffers into their sorted lists together (the insert network then runs once
// for many lanes, not once for each lane that inserts).  The threshold is the
// list's 16th value as of the last merge, never below the current one, so a
// value that skips the buffer is >= the final 16th value: after the last merge
// each list holds the 16 smallest A of its lane's points, as with immediate
// inserts.  IB = 12 slots for rows of 16 and 32 columns: for 16 columns (4
// waves, four point tiles per barrier) 40 KiB of LDS per workgroup, four per
// CU (16 slots: 48 KiB, three per CU; 2^21 16-d rows, 8,192 queries: 1.49 vs
// 1.69 ms).  Rows of 64 columns: 8 slots with a check after every group of
// four rows (CHK = 4; merges start at the same fill, more than 4 entries), so
// that four point tiles per barrier (32 KiB of rows) and the buffers (32 KiB)
// leave room for two workgroups per CU.
#ifndef KDT_SCREEN_IB8
#define KDT_SCREEN_IB8 12
#endif
#ifndef KDT_SCREEN_IB32
#define KDT_SCREEN_IB32 8
#endif
#ifndef KDT_SCREEN_CHK32
#define KDT_SCREEN_CHK32 4
#endif
template <int KS>
constexpr uint32_t insert_slots() { return KS <= 8 ? KDT_SCREEN_IB8 : KS <= 16 ? 12 : KDT_SCREEN_IB32; }
template <int KS>
constexpr int chk() { return KS >= 32 ? KDT_SCREEN_CHK32 : 8; }
static_assert(KDT_SCREEN_IB8 > 8 && KDT_SCREEN_IB32 > KDT_SCREEN_CHK32, "a buffer holds more than one check interval");

// x rounded up to float (a bound stays a bound).
__device__ __forceinline__ float round_up_f(double x) {
  float f = (float)x;
  if ((double)f < x) f = nextafterf(f, INFINITY);
  return f;
}

// The per-tree factors of the screen's error terms (the bound in the file
// header), formed on the host: c1 = a sigma, eb = 2 b sqrt(D) and
// g = (1 + 2^-7) gamma + 1.13 2^-24, in the operation order of k_screen_select.
struct TermK {
  double c1, eb, g, sg, pc;
};

TermK screen_term_k(uint32_t D_, double sg, double pc) {
  const double D = (double)D_;
  const double ea = 0x1p-11 * (1.0 + 0x1p-12);
  const double m = (D + 5.0) * 0x1p-23;
  const double gam = m / (1.0 - m);
  return TermK{ea * sg, 0x1p-14 * (1.0 + 0x1p-60) * 2.0 * sqrt(D) * (1.0 + 0x1p-50),
               (1.0 + 0x1p-7) * gam + 1.13 * 0x1p-24, sg, pc};
}

// One query's (e1, E2) from qc2 = the sum of fl64(q_j - c_j)^2 in any order
// (|q - c| rounded up by a relative 2^-44, as P_c in the build), each rounded
// up to float: e1 >= | |q^ - p^| - sigma |q - p| | and E2 >= |A - |q^ - p^|^2|
// for every row p.  k_screen_select forms the same bounds from its own sum of
// the squares, so the two agree to a relative 2^-40.
__device__ __forceinline__ float2 screen_terms(double qc2, const TermK& t) {
  const double ea = 0x1p-11 * (1.0 + 0x1p-12);
  const double qc = sqrt(qc2) * (1.0 + 0x1p-44);
  const double e1 = (t.c1 * (qc + t.pc) + t.eb) * (1.0 + 1e-12);
  const double Nh = (t.sg * (qc + t.pc) * (1.0 + ea) + t.eb) * (1.0 + 1e-12);
  const double E2 = (t.g * Nh * Nh + 0x1p-118) * (1.0 + 1e-12);
  return make_float2(round_up_f(e1), round_up_f(E2));
}

// A float strictly above the tau_A that k_screen_select derives from an
// upper bound ak of A_k (with no cutoff, which only lowers it; e1, E2 as
// screen_terms gives them): with the select's own terms that value is
//   T = ((s' (sqrt(ak + E2) + e1) + e1)^2 + E2) (1 + 1e-12),  s' <= 1 + 2^-44,
// rounded up to float (one ulp, <= 2^-23 relative), so with these terms
// (2^-40 of the select's) T <= (w^2 + E2)(1 + 2^-22) with
// w = sqrt(max(ak + E2, 0)) + 2 e1.  Here the seven FP32
// operations each err by at most 2^-23 relatively (sqrtf within one ulp, the
// rest within half an ulp; E2 >= 2^-118 keeps the radicand's rounding, when
// it is subnormal, far below e1's share of w), so the result is >= (w^2 +
// E2)(1 + 2^-18)(1 - 2^-20) > (w^2 + E2)(1 + 2^-22).  Overflow gives +inf.
__device__ __forceinline__ float screen_tau_a32(float ak, float2 terms) {
  const float r = sqrtf(fmaxf(ak + terms.y, 0.0f)) + terms.x;
  const float w = r + terms.x;
  return (w * w + terms.y) * (1.0f + 0x1p-18f);
}

// The queries in the screen's operand form, once per call (every point
// slice reads them): row i as -2 q^ (q^_j = screen_f16(fl64(q_j - c_j),
// sigma), exact in fp16) over 2 KS columns (zero past D), and the norm block's
// query-side words of qn = fl32(|q^|^2) (summed in FP64): (q1 | q2 << 16,
// q3 | 1 << 16), from the three bf16 pieces of screen_norm_words; the
// query's error terms (e1, E2) of screen_terms, for the main pass's bounds;
// and, when thr0 is given, +inf as the query's shared threshold for a
// seeding pass.
template <int KS>
__global__ void k_screen_queries(const double* __restrict__ q, uint64_t nq, uint32_t D,
                                 const double* __restrict__ frame, TermK tk, uint16_t* __restrict__ qrows,
                                 uint2* __restrict__ qnorm, float2* __restrict__ qterm, float* __restrict__ thr0) {
  const double sigma = frame[D];
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nq; i += (uint64_t)gridDim.x * blockDim.x) {
    double qn = 0.0, qc2 = 0.0;
    // (Unrolled by 4 at most: a fully unrolled 16-column loop holds every load
    // in flight and takes 68 registers, too many to share a compute unit with
    // a running screen pass.)
#pragma unroll 4
    for (uint32_t c = 0; c < 2 * KS; ++c) {
      const double y = c < D ? q[i * D + c] - frame[c] : 0.0;
      const _Float16 v = c < D ? screen_f16(y, sigma) : (_Float16)0.0f;
      qrows[i * 2 * KS + c] = __builtin_bit_cast(uint16_t, (_Float16)(-2.0f * (float)v));
      qn += (double)v * (double)v;
      qc2 += y * y;
    }
    const uint2 w = screen_norm_words((float)qn);
    qnorm[i] = make_uint2((w.x >> 16) | (w.y << 16), (w.y >> 16) | 0x3F800000u);
    qterm[i] = screen_terms(qc2, tk);
    if (thr0) thr0[i] = INFINITY;
  }
}

// One wave's state for its QT query tiles: per tile the query operands, the
// lane's list and its insert buffer (tile u, slot e at buf[(u IB + e) 64 SW]).
// A tile's 16 values per lane are tested in four groups of four rows: a
// group's per-element tests run only when some lane of the wave has a value
// in it under its threshold (branch-free minima first).  The minima are taken
// over the values' bit patterns as signed integers: that order is the float
// order for values >= +0 and puts every negative value (and -0) below 0, so
// with t = max(bits(thr), 0), A < thr implies bits(A) < t (thr > 0: A >= 0
// compares alike, A < 0 has bits < 0 < t; thr <= 0: A < thr is negative,
// bits(A) < 0 = t).  The group tests therefore never miss a value under the
// threshold, the per-element tests stay in float, and integer minima need no
// canonicalising instruction before them (float minima of matrix results do,
// one per operand).
template <int KS, int KL>
struct Screen {
  static constexpr int NS = KS / 8;  // f16x8 operands per lane half
  static constexpr int RU = KS / 4;
  static constexpr int QT = screen_qt<KS>();
  static constexpr int SW = screen_waves<KS>();
  static constexpr uint32_t IB = insert_slots<KS>();
  f16x8 qh[QT][NS];     // -2 q^
  bf16x8 qnb[QT];       // norm block, query side: (q1, q2, q3, 1, 1, 1, 0, 0) in half 0, zero in half 1
  float la[QT][KL];     // the KL smallest A seen by this lane, ascending
  uint32_t li[QT][KL];  // their tree positions
  float thr[QT];        // min(thr0, la[KL - 1] and the query's shared threshold at the last merge)
  int thri[QT];         // max(bits(thr), 0), the group tests' integer threshold
  float thr0[QT];       // the query's starting threshold (+inf without a seed)
  uint32_t cnt[QT];     // entries in this lane's buffer
  int* gthr[QT];        // the query's shared threshold (bits of a float >= +0), or null
  float2 term[QT];      // the query's (e1, E2)
  uint32_t kk;          // k when a list of KL holds a k-th entry, else 0 (wave-uniform)
  uint2* buf;

  __device__ __forceinline__ void insert(int u, float A, uint32_t row) {  // dropping the largest
#pragma unroll
    for (int e = KL - 1; e > 0; --e) {
      const bool up = A < la[u][e - 1];
      const bool here = !up && A < la[u][e];
      la[u][e] = up ? la[u][e - 1] : (here ? A : la[u][e]);
      li[u][e] = up ? li[u][e - 1] : (here ? row : li[u][e]);
    }
    if (A < la[u][0]) {
      la[u][0] = A;
      li[u][0] = row;
    }
  }

  // Merge every lane's buffer of tile u into its list; the whole wave calls it.
  __device__ __forceinline__ void merge(int u) {
#pragma unroll 1
    for (uint32_t e = 0; e < IB; ++e) {
      if (!__builtin_amdgcn_ballot_w64(e < cnt[u])) break;
      if (e < cnt[u]) {
        const uint2 v = buf[(u * IB + e) * 64 * SW];
        const float A = __uint_as_float(v.x);
        if (A < la[u][KL - 1]) insert(u, A, v.y);
      }
    }
    // The list's k-th entry ak is the A of a point, so A_k <= ak and the
    // select's tau_A is below screen_tau_a32(ak): values at or above that are
    // no candidates.  (The k-th entry as the largest of the first k: an index
    // k - 1 into the list would put the list in scratch memory.)
    float lim = la[u][KL - 1];
    if (kk) {
      float ak = -INFINITY;
#pragma unroll
      for (int e = 0; e < KL; ++e) ak = (uint32_t)e < kk ? fmaxf(ak, la[u][e]) : ak;
      lim = fminf(lim, screen_tau_a32(ak, term[u]));
    }
    thr[u] = fminf(thr0[u], lim);
    if (gthr[u]) {
      // Publish max(lim, +0) to the query's shared threshold and take the
      // smallest one published so far by any list of the query (any slice or
      // half): a value >= some list's last entry is not among the query's KL
      // smallest, and if it were <= tau_A that list's final last entry would
      // be too, which the select flags; a value >= the other bound is no
      // candidate.  Bits of floats >= +0 order as integers.
      // The value is read first and the atomic issued only by a lane that
      // lowers it (hundreds of lists of a query publish when a call has few
      // queries and many point slices); its result is not waited for.
      const int pub = max(__float_as_int(lim), 0);
      const int cur = __hip_atomic_load(gthr[u], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (pub < cur) atomicMin(gthr[u], pub);
      thr[u] = fminf(thr[u], __int_as_float(min(cur, pub)));
    }
    thri[u] = max(__float_as_int(thr[u]), 0);
    cnt[u] = 0;
  }

  // Test the 16 values of query tile u.  Element r of the accumulator is row
  // rowbase | (8 (r / 4) + r % 4) (disjoint bits, one immediate OR).
  __device__ __forceinline__ void test(int u, const floatx16& acc, uint32_t rowbase) {
    int g[4];
#pragma unroll
    for (int q = 0; q < 4; ++q)
      g[q] = min(min(__float_as_int(acc[4 * q]), __float_as_int(acc[4 * q + 1])),
                 min(__float_as_int(acc[4 * q + 2]), __float_as_int(acc[4 * q + 3])));
    if (!__builtin_amdgcn_ballot_w64(min(min(g[0], g[1]), min(g[2], g[3])) < thri[u])) return;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (__builtin_amdgcn_ballot_w64(g[q] < thri[u])) {
#pragma unroll
        for (int r = 4 * q; r < 4 * q + 4; ++r) {
          const float A = acc[r];
          // A wave-uniform branch round the append: a plain divergent `if`
          // here runs the append under an empty exec mask on every element
          // instead of skipping it (the barrier keeps the two tests apart).
          if (__builtin_amdgcn_ballot_w64(A < thr[u])) {
            __builtin_amdgcn_sched_barrier(0);
            if (A < thr[u]) {
              const uint32_t row = rowbase | (8 * (r / 4) + (r % 4));
              buf[(u * IB + cnt[u]) * 64 * SW] = make_uint2(__float_as_uint(A), row);
              ++cnt[u];
            }
          }
        }
      }
      // cnt <= IB - CHK before each check interval of CHK elements (one or two
      // groups), so <= IB after it.
      constexpr int CHK = chk<KS>();
      if ((CHK == 4 || q % 2 == 1) && __builtin_amdgcn_ballot_w64(cnt[u] > IB - CHK)) merge(u);
    }
  }

  // The wave's QT x 32 queries against point tile t (32 rows in LDS): -2 q^.p^
  // from the NS fp16 products (each point operand, read once, feeds the QT
  // chains), then the norms, then the tests.
  __device__ __forceinline__ void tile(uint32_t t, uint32_t j, uint32_t h, const uint4* rows, const uint2* pns) {
    const uint4* rb = rows + j * RU;
    const uint32_t sj = swz<KS>(j);
    // Norm block, point side, of row j in half 0 (built with the rows), zero
    // in half 1.  Padding rows: p1 = +inf, p2 = p3 = 0, A = +inf.
    const uint2 pw = pns[j + 32 * h];
    const uint4 pnw = make_uint4(h == 0 ? 0x3F803F80u : 0u, pw.x, pw.y, 0u);
    floatx16 acc[QT];
#pragma unroll
    for (int u = 0; u < QT; ++u) acc[u] = floatx16{};
#pragma unroll
    for (int v = 0; v < NS; ++v) {
      const f16x8 pv = __builtin_bit_cast(f16x8, rb[(h * NS + v) ^ sj]);
#pragma unroll
      for (int u = 0; u < QT; ++u) acc[u] = __builtin_amdgcn_mfma_f32_32x32x16_f16(pv, qh[u][v], acc[u], 0, 0, 0);
    }
#pragma unroll
    for (int u = 0; u < QT; ++u)
      acc[u] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, pnw), qnb[u], acc[u], 0, 0, 0);
#pragma unroll
    for (int u = 0; u < QT; ++u) test(u, acc[u], t * 32 + 4 * h);
  }
};

// KL: entries per lane list (KP in the main pass; the seeding pass may keep
// fewer, its lists only bound A_k from above).
template <int KS, int KL>
__global__ __launch_bounds__(64 * screen_waves<KS>()) void k_screen(SArgs a) {
  constexpr int SW = screen_waves<KS>();
  constexpr int TS = Stage<KS>::TS;
  constexpr int RU = Stage<KS>::RU;
  __shared__ uint4 s_rows[2][TS * 32 * RU];
  __shared__ uint2 s_pn[2][TS * 64];
  constexpr int QT = screen_qt<KS>();
  __shared__ uint2 s_buf[QT * insert_slots<KS>() * 64 * SW];  // per-lane insert buffers, slot-major
  const uint32_t tid = threadIdx.x, lane = tid & 63;
  const uint32_t j = lane & 31, h = lane >> 5;
  // Wave w of the workgroup takes query tiles (blockIdx.x SW + w) QT + u.
  const uint64_t qtile0 = ((uint64_t)blockIdx.x * SW + (tid >> 6)) * QT;
  const bool active = qtile0 * 32 < a.nq;  // wave-uniform: a wave past the last query only stages tiles
  Screen<KS, KL> sc;
  constexpr int NS = KS / 8, RQ = KS / 4;
#pragma unroll
  for (int u = 0; u < QT; ++u) {
    // Rows past nq read the last query (their lists are not written).
    const uint64_t qrow = min((qtile0 + u) * 32 + j, a.nq - 1);
#pragma unroll
    for (int v = 0; v < NS; ++v) sc.qh[u][v] = __builtin_bit_cast(f16x8, a.qrows[qrow * RQ + h * NS + v]);
    const uint2 w = a.qnorm[qrow];
    sc.qnb[u] = __builtin_bit_cast(bf16x8, h == 0 ? make_uint4(w.x, w.y, 0x3F803F80u, 0u) : make_uint4(0u, 0u, 0u, 0u));
    sc.thr0[u] = a.thr0 ? a.thr0[qrow] : INFINITY;
    // Only lanes whose lists the select checks publish (rows past nq repeat
    // the last query and are not written).
    sc.gthr[u] = a.thr0 && (qtile0 + u) * 32 + j < a.nq ? reinterpret_cast<int*>(a.thr0) + qrow : nullptr;
    sc.term[u] = a.qterm ? a.qterm[qrow] : make_float2(0.0f, 0.0f);
#pragma unroll
    for (int e = 0; e < KL; ++e) {
      sc.la[u][e] = INFINITY;
      sc.li[u][e] = 0xFFFFFFFFu;
    }
    sc.thr[u] = sc.thr0[u];
    sc.thri[u] = max(__float_as_int(sc.thr[u]), 0);
    sc.cnt[u] = 0;
  }
  sc.kk = a.qterm && a.k <= (uint32_t)KL ? a.k : 0u;
  sc.buf = s_buf + tid;

  // Local tile i of slice y is global point tile tb + i ts.  The rows are in
  // tree order, so a contiguous range of tiles is one region of space and holds
  // most of a query's near neighbours: fine while its two lane lists of KP hold
  // them (k <= KP: contiguous, tb = y * tiles_per_slice, ts = 1), not beyond
  // (k > KP: every S-th tile, tb = y, ts = S, so that each slice samples all of
  // space; 16-d uniform data, k = 32: 12-28% of the queries overflowed a list
  // and took the traversal with contiguous slices, none interleaved).  The
  // seeding pass (k_screen_seed) takes every G-th tile: tb = y G, ts = G S.
  // The loop below runs over local indices.
  const uint32_t sy = blockIdx.y;
  const uint32_t tb = sy * a.tb_mul, ts = a.ts;
  const uint32_t t1 = tb < a.n_tiles ? min(a.tmax, (a.n_tiles - tb + ts - 1) / ts) : 0;
#define KDT_TILE(i) (tb + (i) * ts)
  // Steps of TS local tiles: step q holds local tiles q TS .. q TS + TS - 1
  // (those from t1 on are re-reads of tile t1 - 1 and are not contracted).
  // Step q sits in LDS stage q & 1 while it is contracted; step q + 1 is
  // written to the other stage in the same iteration (its readers, step
  // q - 1, finished before the barrier that ended the previous iteration) and
  // step q + 2 is in flight to registers.  x and y alternate as the register
  // stage of steps q + 1 and q + 2.  Every path issues the same loads (indices
  // clamped), so the compiler's wait before each LDS store covers only the
  // step it stores and the next step's loads stay in flight.
  // (A macro rather than a lambda: the two register stages must stay named
  // registers, never a stack object behind a reference.)
  const uint32_t nsteps = (t1 + TS - 1) / TS;
#define KDT_SCREEN_STEP(Q, B, NP)                                                                \
  do {                                                                                           \
    if ((Q) + 1 < nsteps) {                                                                      \
      pn_store<KS>(NP, tid, s_pn[(B) ^ 1]);                                                      \
      stage_dma<KS>(a, ((Q) + 1) * TS, t1, tb, ts, tid, s_rows[(B) ^ 1]);                        \
    }                                                                                            \
    pn_load<KS>(a, min((Q) + 3, nsteps - 1) * TS, t1, tb, ts, tid, NP);                          \
    if (active && (Q) < nsteps) {                                                                \
      _Pragma("unroll") for (int i_ = 0; i_ < TS; ++i_) {                                        \
        const uint32_t lt_ = (Q) * TS + i_;                                                      \
        if (lt_ < t1) sc.tile(KDT_TILE(lt_), j, h, s_rows[B] + i_ * 32 * RU, s_pn[B] + i_ * 64); \
      }                                                                                          \
    }                                                                                            \
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");                                             \
    __syncthreads();                                                                             \
  } while (0)
  if (nsteps > 0) {  // uniform over the workgroup
    uint2 xp = {}, yp = {};
    stage_dma<KS>(a, 0, t1, tb, ts, tid, s_rows[0]);
    pn_load<KS>(a, 0, t1, tb, ts, tid, xp);
    for (uint32_t u = tid; u < 2 * TS * 32; u += 64 * SW)  // the zero halves of both stages
      s_pn[u / (TS * 32)][((u % (TS * 32)) / 32) * 64 + 32 + u % 32] = make_uint2(0u, 0u);
    pn_store<KS>(xp, tid, s_pn[0]);
    pn_load<KS>(a, min(1u, nsteps - 1) * TS, t1, tb, ts, tid, xp);
    pn_load<KS>(a, min(2u, nsteps - 1) * TS, t1, tb, ts, tid, yp);
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    for (uint32_t q = 0; q < nsteps; q += 2) {
      KDT_SCREEN_STEP(q, 0, xp);
      KDT_SCREEN_STEP(q + 1, 1, yp);
    }
  }
#undef KDT_SCREEN_STEP
#undef KDT_TILE
#pragma unroll
  for (int u = 0; u < QT; ++u) {
    if (active) sc.merge(u);  // wave-uniform
    const uint64_t qrow = (qtile0 + u) * 32 + j;
    if (qrow < a.nq) {
      const size_t base = (qrow * (2 * a.slices) + 2 * blockIdx.y + h) * KL;
#pragma unroll
      for (int e = 0; e < KL; ++e) {
        a.part_a[base + e] = sc.la[u][e];
        a.part_i[base + e] = sc.li[u][e];
      }
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
  uint32_t* __restrict__ nflag;  // queries flagged so far (cleared before the select)
  float* __restrict__ seed;      // seeding mode: [nq] tau_A of the lists' k-th A, and nothing else
  uint64_t nq;
  const double* __restrict__ frame;  // [D + 1] the tree's centre and sigma
  uint32_t D, k, lists;
  uint32_t kl;               // entries per list (KP; fewer for the seeding pass's lists)
  double sigma, pc, cutoff;  // sigma, P_c >= max over rows of |p - c|, the reduced cutoff
};

// One wave per query; lane l reads lists l, l + 64, ... (at most LPL <=
// LISTS_PER_LANE; the launch takes the smallest LPL of 1, 4, 16 that covers the
// lists, so that a call with few lists neither loops over nor holds registers
// for absent ones):
// k rounds of wave-min over the list heads give A_k, then tau_A from the error
// bound, the completeness check and the candidates (list entries with
// A <= tau_A, at most CAND_CAP).
template <int LPL>
__global__ __launch_bounds__(64) void k_screen_select(SelArgs a) {
  const uint64_t qi = blockIdx.x;
  const uint32_t lane = threadIdx.x;
  const uint32_t kl = a.kl;
  const float* la0 = a.part_a + qi * a.lists * kl;
  const uint32_t* li0 = a.part_i + qi * a.lists * kl;
  uint32_t head[LPL];
  float hv[LPL];  // value at each head (+inf when exhausted or absent)
#pragma unroll
  for (int c = 0; c < LPL; ++c) {
    const uint32_t L = lane + 64 * c;
    head[c] = 0;
    hv[c] = L < a.lists ? la0[L * kl] : INFINITY;
  }
  float ak = INFINITY;
  for (uint32_t r = 0; r < a.k; ++r) {
    float cur = hv[0];
#pragma unroll
    for (int c = 1; c < LPL; ++c) cur = fminf(cur, hv[c]);
    float m = cur;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fminf(m, __shfl_xor(m, o));
    ak = m;
    if (m == INFINITY) break;
    const uint64_t hit = __ballot(cur == m);
    if (lane == (uint32_t)(__ffsll((unsigned long long)hit) - 1)) {
      bool done = false;
#pragma unroll
      for (int c = 0; c < LPL; ++c) {
        if (!done && hv[c] == m) {
          done = true;
          ++head[c];
          hv[c] = head[c] < kl ? la0[(lane + 64 * c) * kl + head[c]] : INFINITY;
        }
      }
    }
  }
  // |q - c| and the largest |fl64(q_j - c_j)|, in the frame of the rows.
  double qc2 = 0.0, qmax = 0.0;
  for (uint32_t c = lane; c < a.D; c += 64) {
    const double y = a.q[qi * a.D + c] - a.frame[c];
    qc2 += y * y;
    qmax = fmax(qmax, fabs(y));
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    qc2 += __shfl_xor(qc2, o);
    qmax = fmax(qmax, __shfl_xor(qmax, o));
  }
  // tau_A (every lane computes the same value).  Each bound is rounded up by
  // a relative 1e-12 (far above the FP64 error of computing it).
  const double D = (double)a.D;
  const double sg = a.sigma;
  const double qc = sqrt(qc2) * (1.0 + 0x1p-44);  // >= |q - c| (as P_c in the build)
  const double ea = 0x1p-11 * (1.0 + 0x1p-12);
  const double eb = 0x1p-14 * (1.0 + 0x1p-60) * 2.0 * sqrt(D) * (1.0 + 0x1p-50);
  const double e1 = (ea * sg * (qc + a.pc) + eb) * (1.0 + 1e-12);
  const double Nh = (sg * (qc + a.pc) * (1.0 + ea) + eb) * (1.0 + 1e-12);
  const double m = (D + 5.0) * 0x1p-23;
  const double gam = m / (1.0 - m);
  const double E2 = (((1.0 + 0x1p-7) * gam + 1.13 * 0x1p-24) * Nh * Nh + 0x1p-118) * (1.0 + 1e-12);
  const double sh = 1.0 + 2.0 * (D + 3.0) * 0x1p-53;
  double tau = a.cutoff;
  if (ak < INFINITY) {
    const double r = (sqrt(fmax(0.0, (double)ak + E2)) + e1) / sg;
    tau = fmin(tau, sh * r * r * (1.0 + 1e-12));
  }
  float ta = INFINITY;
  if (tau < INFINITY) {
    const double r = sg * sqrt(sh * tau) + e1;
    const double t = (r * r + E2) * (1.0 + 1e-12);
    ta = (float)t;
    if ((double)ta < t) ta = nextafterf(ta, INFINITY);
  }
  if (a.seed) {
    // Any k points' A values bound A_k from above, so tau_A of the sampled
    // lists' k-th value is >= the final tau_A: a lane may drop every value at
    // or above the next float up (lanes drop A >= threshold, and a point with
    // A = tau_A can be a candidate).  Out-of-range queries: no seed.
    if (lane == 0) a.seed[qi] = qmax * sg <= 2048.0 ? nextafterf(ta, INFINITY) : INFINITY;
    return;
  }
  // Completeness: a full list whose last value is <= tau_A may have dropped a
  // candidate.  Unfilled slots (A = +inf, position 0xFFFFFFFF) end a list.
  uint32_t cnt_c[LPL];
  uint32_t mcount = 0;
  bool complete = true;
#pragma unroll
  for (int c = 0; c < LPL; ++c) {
    const uint32_t L = lane + 64 * c;
    uint32_t n = 0;
    if (L < a.lists) {
      const float* la = la0 + L * kl;
      const uint32_t* li = li0 + L * kl;
      const float last = la[kl - 1];
      complete = complete && (last == INFINITY || last > ta);
      while (n < kl && li[n] != 0xFFFFFFFFu && la[n] <= ta) ++n;
    }
    cnt_c[c] = n;
    mcount += n;
  }
  const bool ok = qmax * sg <= 2048.0 && !__ballot(!complete);  // in range of the fp16 rounding (NaN: not)
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
    for (int c = 0; c < LPL; ++c) {
      const uint32_t* li = li0 + (lane + 64 * c) * kl;
      for (uint32_t e = 0; e < cnt_c[c]; ++e) a.cand[qi * CAND_CAP + off + e] = li[e];
      off += cnt_c[c];
    }
  }
  if (lane == 0) {
    a.cnt[qi] = good ? total : 0;
    a.flag[qi] = good ? 0 : 1;
    if (!good) atomicAdd(a.nflag, 1u);
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
  uint32_t* __restrict__ nflag;
  double* __restrict__ out_d;
  int64_t* __restrict__ out_i;
  uint64_t nq;
  uint32_t D, k, n;
  uint32_t leafsize;
  double cutoff;
};

__device__ __forceinline__ bool key_less(double da, uint32_t ia, double db, uint32_t ib) {
  return da < db || (da == db && ia < ib);
}

// Whether tree position p lies in a complete chunk of P rows counted from the
// first row of its leaf: the Packed recipes of 4- and 8-column rows use the
// pair / chunk order only there (src/kernel.rs:447).  The tree shape depends on
// (n, leafsize) only: a node of more than leafsize rows keeps its first len / 2
// rows on the left (src/build.rs:166-173), so the leaf follows from p alone.
__device__ __forceinline__ bool in_full_chunk(uint32_t p, uint32_t n, uint32_t leafsize, uint32_t P) {
  uint32_t start = 0, len = n;
  while (len > leafsize) {
    const uint32_t mid = len / 2;
    if (p < start + mid) {
      len = mid;
    } else {
      start += mid;
      len -= mid;
    }
  }
  return p - start < len / P * P;
}

// Marks query qi for the traversal (called once per query, by one lane: the
// recheck returns at once for a query the select already flagged).
__device__ __forceinline__ void flag_query(const CheckArgs& a, uint64_t qi) {
  a.flag[qi] = 1;
  atomicAdd(a.nflag, 1u);
}

// One wave per query: FP64 recheck of the candidates, top-k, certificate.
// Dynamic LDS: k x D doubles, the cell-bound axes of each certificate walk.
__global__ __launch_bounds__(64) void k_screen_recheck(CheckArgs a) {
  extern __shared__ double axes_all[];
  constexpr uint32_t PATH_INNER = 32;         // inner nodes on a root-to-leaf path (the build checks depth <= 32)
  __shared__ uint32_t s_chk[KMAX * PATH_INNER];  // per result lane: the children whose box gate is still to check
  __shared__ uint32_t s_off[KMAX + 1];
  const uint64_t qi = blockIdx.x;
  const uint32_t lane = threadIdx.x;
  const uint32_t D = a.D;
  if (a.flag[qi]) return;
  const uint32_t c = a.cnt[qi];
  if (c > CAND_CAP) {
    if (lane == 0) flag_query(a, qi);
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
    if (lane == 0) flag_query(a, qi);
    return;
  }
  if (lane < c) {
    const bool full = D == 8 ? in_full_chunk(pos, a.n, a.leafsize, 8) : D == 4 ? in_full_chunk(pos, a.n, a.leafsize, 2)
                                                                        : true;
    const double v = point_dist_d<M_L2, 0>(q, a.rows + (size_t)pos * D, 1, D, full, 2.0);
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

  // The certificate in two parts.  Each result lane walks its point's path
  // with the plane gates and cell bounds inline (cheap) and records the
  // children whose box gate must also hold; then the whole wave evaluates
  // those box gates, spread over its 64 lanes (each one the same box_dist_d
  // as before, so the same bits).
  bool ok = true;
  uint32_t nchk = 0;
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
      // descend_by_box (src/query.rs:296-327): the child's box gate; descend
      // (src/query.rs:278-286): a far child's plane gate, then its box gate.
      const bool box_gate = (nd.b & OBB_BIT) || is_far;
      if (!(nd.b & OBB_BIT) && is_far) ok = far_dist <= dcert;
      if (ok && box_gate) {
        if (nchk < PATH_INNER) s_chk[lane * PATH_INNER + nchk] = child;
        else ok = false;  // deeper than the build allows: leave the query to the traversal
        ++nchk;
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
    if (lane == 0) flag_query(a, qi);
    return;
  }
  // Box gates, all paths at once: offsets of each lane's list, then lane w
  // takes checks w, w + 64, ...
  const uint32_t mine = lane < nres ? nchk : 0u;
  uint32_t incl = mine;
#pragma unroll
  for (uint32_t o = 1; o < 64; o <<= 1) {
    const uint32_t v = __shfl_up(incl, o);
    if (lane >= o) incl += v;
  }
  if (lane < KMAX) s_off[lane] = incl - mine;
  const uint32_t total = __shfl(incl, 63);
  __syncthreads();
  bool bad = false;
  for (uint32_t w = lane; w < total; w += 64) {
    uint32_t L = 0;
    while (L + 1 < nres && s_off[L + 1] <= w) ++L;
    const uint32_t child = s_chk[L * PATH_INNER + (w - s_off[L])];
    bad |= !(box_dist_d<M_L2, 0>(q, a.boxes + (size_t)child * 2 * D, D, 2.0) <= dcert);
  }
  if (__ballot(bad)) {
    if (lane == 0) flag_query(a, qi);
    return;
  }
  if (lane < a.k) {  // write_results, src/query.rs:382-400
    a.out_d[qi * a.k + lane] = lane < nres ? sqrt(d) : INFINITY;
    a.out_i[qi * a.k + lane] = lane < nres ? (int64_t)id : (int64_t)a.n;
  }
}

// Workgroups of k_screen<KS> the device holds at once (compute units x
// workgroups per unit), looked up once per device.
template <int KS>
uint64_t screen_resident_ks() {
  static std::mutex mu;
  static int cache[64] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 512;
  std::lock_guard<std::mutex> g(mu);
  if (!cache[dev]) {
    int cus = 0, per = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, reinterpret_cast<const void*>(k_screen<KS, KP>),
                                                     64 * screen_waves<KS>(), 0) != hipSuccess || per <= 0)
      per = 1;
    cache[dev] = cus * per;
  }
  return (uint64_t)cache[dev];
}

uint64_t screen_resident(int ks) {
  return ks == 8 ? screen_resident_ks<8>() : ks == 16 ? screen_resident_ks<16>() : screen_resident_ks<32>();
}

int screen_ks(uint64_t d) { return d <= 16 ? 8 : d <= 32 ? 16 : 32; }

// Point slices per query block (the SW query tiles of a workgroup).  Each
// workgroup walks its slice's tiles, so a launch takes ceil(workgroups /
// resident) rounds of about ceil(tiles / slices) + W0 tile steps, W0 = 256 the
// per-workgroup cost of filling its lane lists (a short slice keeps the lists'
// thresholds high: more appends and merges).  The count minimises that
// (fewest slices among equals), between at least 8 tiles per slice and, for
// k > KP, at least k / 4 interleaved slices: a query's top k then spread over
// at least k / 2 lists of KP, about two per list, so a list rarely fills below
// tau_A (the completeness check flags the query when one does).  Measured on
// 2^20 64-d points (k_screen ms by slice count): 20,000 queries at k = 32,
// 8: 7.19, 12: 5.67, 19: 5.74, 32: 6.23, 64: 7.49, 512: 15.15; 8,192 queries
// at k = 10, 8: 3.36, 16: 2.26, 32: 2.55; 1,024 queries, 128: 0.54, 256: 0.63.
// Returns the slice count; *per = tiles per slice.
uint64_t screen_slices(const Tree& t, const QueryParams& qp, uint32_t* per) {
  const int ks = screen_ks(t.d);
  const uint64_t block = (uint64_t)32 * (ks == 8 ? screen_waves<8>() * screen_qt<8>()
                                                 : ks == 16 ? screen_waves<16>() * screen_qt<16>()
                                                            : screen_waves<32>() * screen_qt<32>());
  const uint64_t gx = (qp.nq + block - 1) / block;
  const uint64_t n_tiles = (t.n + 31) / 32;
  const uint64_t resident = screen_resident(ks);
  const uint64_t hi = std::min<uint64_t>(MAX_SLICES, std::max<uint64_t>(1, n_tiles / 8));
  const uint64_t lo = std::min<uint64_t>(hi, qp.k > (uint64_t)KP ? (qp.k + 3) / 4 : 1);
  uint64_t best_s = lo, best_cost = UINT64_MAX, best_per = (n_tiles + lo - 1) / lo;
  for (uint64_t S = lo; S <= hi; ++S) {
    const uint64_t p = (n_tiles + S - 1) / S;
    const uint64_t used = (n_tiles + p - 1) / p;  // slices that hold tiles
    const uint64_t cost = (gx * used + resident - 1) / resident * (p + 256);
    if (cost < best_cost) {
      best_cost = cost;
      best_s = used;
      best_per = p;
    }
  }
  if (qp.slice_shift != 0) {  // KDT_QUERY_SLICES_*: a per-tree work-tile shape
    uint64_t S = qp.slice_shift > 0 ? best_s << qp.slice_shift : (best_s + 1) / 2;
    S = std::min(std::max(S, lo), hi);
    best_per = (n_tiles + S - 1) / S;
    best_s = (n_tiles + best_per - 1) / best_per;
  }
  *per = (uint32_t)best_per;
  return best_s;
}

// Threshold seeding.  A lane's list starts empty, so its threshold is +inf
// until 16 values have passed and then falls slowly: over a slice of m values
// a lane keeps about 16 (1 + ln(m / 16)) of them, each an LDS append and a
// merge through the insert network, most of them early.  A seeding pass over
// every SEED_STRIDE-th point tile gives each query an upper bound of A_k (the
// k-th A among the sampled points) and from it a threshold that the full pass
// starts from in every lane.  Trees of fewer than SEED_MIN_TILES tiles are
// screened without it.
#ifndef KDT_SCREEN_SEED_MIN_TILES
#define KDT_SCREEN_SEED_MIN_TILES 1024
#endif
constexpr uint32_t SEED_STRIDE = 64;
constexpr uint32_t SEED_MIN_TILES = KDT_SCREEN_SEED_MIN_TILES;
bool screen_seeded(uint32_t n_tiles) { return n_tiles >= SEED_MIN_TILES; }

// Lists of the seeding pass.  Its seed only has to bound A_k from above, and
// the k-th smallest A over any set of kept entries does, however short the
// lists: a lane that keeps its SEED_KL smallest values instead of KP gives a
// seed no lower (a looser one only when more than SEED_KL of the sampled top
// k fall into one lane).  Its slices interleave the sampled tiles, so a
// query's sampled top k spread over all 2 s0 lanes: with 2 s0 SEED_KL >= 2 k
// entries a lane rarely holds more than SEED_KL of them.  Short lists are what
// makes the pass cheap: a lane's list starts empty, so nearly every value of
// its first tiles goes through the insert network (about L (1 + ln(m / L))
// inserts for m values and lists of L), and each insert moves L entries.
constexpr uint32_t SEED_KL = 4;
uint32_t seed_list_len(uint32_t s0, uint32_t k) { return 2 * s0 * SEED_KL >= 2 * k ? SEED_KL : (uint32_t)KP; }

}  // namespace

bool screen_applies(const Tree& t, const QueryParams& qp) {
  return qp.fp32_screen && qp.allow_scratch && qp.metric == M_L2 && qp.eps_factor == 1.0 &&
         qp.k <= (uint64_t)KMAX && t.rows_h &&
         qp.nq < (1ull << 31) && (t.d >= 32 || (qp.matrix_screen && t.d >= 8)) && t.d <= 64;
}

size_t screen_scratch_bytes(const Tree& t, const QueryParams& qp) {
  uint32_t per = 0;
  const uint64_t slices = screen_slices(t, qp, &per);
  const uint64_t nq = qp.nq;
  const uint64_t ks = (uint64_t)screen_ks(t.d);
  const uint32_t n_tiles = (uint32_t)((t.n + 31) / 32);
  return 2 * scratch_piece(nq * 2 * slices * KP * 4) + scratch_piece(nq * 4) + scratch_piece(nq * CAND_CAP * 4) +
         scratch_piece(nq) + scratch_piece(4) + scratch_piece(nq * 2 * ks * 2) + 2 * scratch_piece(nq * 8) +
         (screen_seeded(n_tiles) ? scratch_piece(nq * 4) : 0);
}

void launch_select(const SelArgs& a, hipStream_t s) {
  if (a.lists <= 64) k_screen_select<1><<<(uint32_t)a.nq, 64, 0, s>>>(a);
  else if (a.lists <= 256) k_screen_select<4><<<(uint32_t)a.nq, 64, 0, s>>>(a);
  else k_screen_select<LISTS_PER_LANE><<<(uint32_t)a.nq, 64, 0, s>>>(a);
}

int launch_screen(const Tree& t, const QueryParams& qp, const KArgs& trav, Scratch& sc, hipStream_t s,
                  std::string& msg) {
  const uint64_t nq = qp.nq;
  const uint32_t k = (uint32_t)qp.k;
  const uint64_t qtiles = (nq + 31) / 32;
  const uint32_t n_tiles = (uint32_t)((t.n + 31) / 32);
  uint32_t per = 0;
  const uint64_t slices = screen_slices(t, qp, &per);

  // Scratch: every piece is written in full before it is read (the lists by
  // k_screen, counts, candidates and flags by k_screen_select), except the
  // flagged-query count, which a kernel clears before the select.
  float* part_a = nullptr;
  uint32_t *part_i = nullptr, *cnt = nullptr, *cand = nullptr;
  uint8_t* flag = nullptr;
  uint32_t* nflag = nullptr;
  hipError_t e = sc.get(&part_a, nq * 2 * slices * KP);
  if (e == hipSuccess) e = sc.get(&part_i, nq * 2 * slices * KP);
  if (e == hipSuccess) e = sc.get(&cnt, nq);
  if (e == hipSuccess) e = sc.get(&cand, nq * CAND_CAP);
  if (e == hipSuccess) e = sc.get(&flag, nq);
  if (e == hipSuccess) e = sc.get(&nflag, 1);
  const int ks = screen_ks(t.d);
  uint16_t* qrows = nullptr;
  uint2* qnorm = nullptr;
  float2* qterm = nullptr;
  if (e == hipSuccess) e = sc.get(&qrows, nq * 2 * ks);
  if (e == hipSuccess) e = sc.get(&qnorm, nq);
  if (e == hipSuccess) e = sc.get(&qterm, nq);
  const bool seeded = screen_seeded(n_tiles);
  float* thr0 = nullptr;
  if (e == hipSuccess && seeded) e = sc.get(&thr0, nq);
  if (e != hipSuccess) {
    msg = std::string("screen scratch: ") + hipGetErrorString(e);
    return 100;
  }
  // Seeding pass geometry (below): S0 slices over every SEED_STRIDE-th tile.
  const uint32_t ns = (n_tiles + SEED_STRIDE - 1) / SEED_STRIDE;
  const uint32_t s0 = (uint32_t)std::min<uint64_t>(slices, std::max<uint32_t>(1, ns / 8));
  const uint32_t seed_kl = seed_list_len(s0, k);
  // The seeding lists share a threshold per query (the seed's slot, +inf
  // until then) when they hold k entries or more: a value at or above some
  // list's k-th entry, or above a bound of tau_A from it, is not among the k
  // smallest entries of all the lists, so the seed is the same.  Shorter
  // lists keep their own thresholds (sharing their last entry could drop
  // some of those k smallest and weaken the seed).
  const bool seed_shared = seeded && seed_kl >= k;
  {
    const uint32_t pb = (uint32_t)std::min<uint64_t>((nq + 255) / 256, 4096);
    const uint32_t D = (uint32_t)t.d;
    const TermK tk = screen_term_k(D, t.scr_sigma, t.scr_pc);
    if (ks == 8) k_screen_queries<8><<<pb, 256, 0, s>>>(trav.q, nq, D, t.frame, tk, qrows, qnorm, qterm, seed_shared ? thr0 : nullptr);
    else if (ks == 16) k_screen_queries<16><<<pb, 256, 0, s>>>(trav.q, nq, D, t.frame, tk, qrows, qnorm, qterm, seed_shared ? thr0 : nullptr);
    else k_screen_queries<32><<<pb, 256, 0, s>>>(trav.q, nq, D, t.frame, tk, qrows, qnorm, qterm, seed_shared ? thr0 : nullptr);
  }
  SArgs sa{};
  sa.rows_h = t.rows_h;
  sa.pnorm_h = t.pnorm_h;
  sa.qrows = reinterpret_cast<const uint4*>(qrows);
  sa.qnorm = qnorm;
  sa.qterm = qterm;
  sa.k = k;
  sa.part_a = part_a;
  sa.part_i = part_i;
  sa.nq = nq;
  sa.D = (uint32_t)t.d;
  sa.n = (uint32_t)t.n;
  sa.n_tiles = n_tiles;
  sa.slices = (uint32_t)slices;
  if (k > (uint32_t)KP) {  // interleaved
    sa.tb_mul = 1;
    sa.ts = (uint32_t)slices;
    sa.tmax = 0xFFFFFFFFu;
  } else {  // contiguous
    sa.tb_mul = per;
    sa.ts = 1;
    sa.tmax = per;
  }
  const int sw = ks == 8 ? screen_waves<8>() : ks == 16 ? screen_waves<16>() : screen_waves<32>();
  const int qt = ks == 8 ? screen_qt<8>() : ks == 16 ? screen_qt<16>() : screen_qt<32>();
  auto screen = [&](const SArgs& args, uint32_t ny, uint32_t kl, const char* name, double work) {
    const dim3 grid((uint32_t)((qtiles + (uint64_t)sw * qt - 1) / ((uint64_t)sw * qt)), ny);
    KernelTimer tm(s);
    if (kl == SEED_KL) {
      if (ks == 8) k_screen<8, SEED_KL><<<grid, 64 * sw, 0, s>>>(args);
      else if (ks == 16) k_screen<16, SEED_KL><<<grid, 64 * sw, 0, s>>>(args);
      else k_screen<32, SEED_KL><<<grid, 64 * sw, 0, s>>>(args);
    } else {
      if (ks == 8) k_screen<8, KP><<<grid, 64 * sw, 0, s>>>(args);
      else if (ks == 16) k_screen<16, KP><<<grid, 64 * sw, 0, s>>>(args);
      else k_screen<32, KP><<<grid, 64 * sw, 0, s>>>(args);
    }
    tm.stop(name, work);
  };

  SelArgs se{};
  se.part_a = part_a;
  se.part_i = part_i;
  se.q = trav.q;
  se.cnt = cnt;
  se.cand = cand;
  se.flag = flag;
  se.nflag = nflag;
  se.nq = nq;
  se.D = (uint32_t)t.d;
  se.k = k;
  se.lists = (uint32_t)(2 * slices);
  se.kl = KP;
  se.frame = t.frame;
  se.sigma = t.scr_sigma;
  se.pc = t.scr_pc;
  se.cutoff = qp.cutoff;

  if (seeded) {
    // Seeding pass: every SEED_STRIDE-th point tile, in S0 slices, then each
    // query's tau_A of the k-th A over those lists as every lane's starting
    // threshold in the full pass (the lists' scratch is reused).
    SArgs sp = sa;
    sp.thr0 = seed_shared ? thr0 : nullptr;
    sp.slices = s0;
    sp.tb_mul = SEED_STRIDE;
    sp.ts = SEED_STRIDE * s0;
    sp.tmax = 0xFFFFFFFFu;
    const uint32_t kl = seed_kl;
    screen(sp, s0, kl, "k_screen_seed", 2.0 * (double)t.d * 32.0 * (double)ns * (double)nq);
    SelArgs ss = se;
    ss.lists = 2 * s0;
    ss.kl = kl;
    ss.seed = thr0;
    launch_select(ss, s);
    sa.thr0 = thr0;
  }
  // Algorithmic work: 2 D FLOPs per (query, point) pair of the contraction.
  screen(sa, (uint32_t)slices, KP, ks == 8 ? "k_screen<8>" : ks == 16 ? "k_screen<16>" : "k_screen<32>",
         2.0 * (double)t.d * (double)t.n * (double)nq);

  launch_zero_u32(nflag, 1, s);
  {
    KernelTimer tm(s);
    launch_select(se, s);
    tm.stop("k_screen_select", (double)nq);
  }

  return finish_candidates(t, qp, trav, cnt, cand, flag, nflag, s, msg);
}

int finish_candidates(const Tree& t, const QueryParams& qp, const KArgs& trav, const uint32_t* cnt,
                      const uint32_t* cand, uint8_t* flag, uint32_t* nflag, hipStream_t s, std::string& msg) {
  CheckArgs ca{};
  ca.q = trav.q;
  ca.rows = t.rows;
  ca.idx = t.idx;
  ca.nodes = t.nodes;
  ca.boxes = t.boxes;
  ca.cnt = cnt;
  ca.cand = cand;
  ca.flag = flag;
  ca.nflag = nflag;
  ca.out_d = trav.out_d;
  ca.out_i = trav.out_i;
  ca.nq = qp.nq;
  ca.D = (uint32_t)t.d;
  ca.k = (uint32_t)qp.k;
  ca.n = (uint32_t)t.n;
  ca.leafsize = (uint32_t)std::min<uint64_t>(t.leafsize, 0xFFFFFFFFu);
  ca.cutoff = qp.cutoff;
  {
    KernelTimer tm(s);
    k_screen_recheck<<<(uint32_t)qp.nq, 64, (size_t)qp.k * t.d * 8, s>>>(ca);
    tm.stop("k_screen_recheck", (double)qp.nq);
  }

  // Flagged queries: the exact traversal (the compiled register-list kernel
  // for its widths, the runtime-width one otherwise), in the given order.
  KArgs fb = trav;
  fb.order = nullptr;
  fb.only = flag;
  fb.only_count = nflag;
  int rc = 0;
  if (!qp.screen_only) {
    rc = -1;
    if (qp.k <= 64) {
      int G = group_width(qp.k, qp.group);
      if (t.d >= 32 && G < 32) G = 32;
      rc = launch_knn_l2(fb, G, false, s);
    }
    if (rc) rc = launch_knn_lds(fb, M_L2, s);
  }
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
