// kdtree_oracle.cpp -- CPU restatement of the reference KD-tree build and k-NN
// query (FujishigeTemma/kdtree-rs), written from the reference's Rust source.
//
// TEST INFRASTRUCTURE ONLY.  This file is the oracle: it is loaded by tests/,
// by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and nowhere
// else.  The product (kdtree-rs_amd/) never links, loads or calls it.
//
// Every function below cites the reference file:line it restates.  The
// arithmetic follows the reference statement for statement so that reduced
// distances carry the same bits: Rust never contracts a*b+c into an FMA and
// never reassociates, so this file is compiled with -ffp-contract=off and
// without -ffast-math (see oracle/Makefile).
//
// Parity status per part (how each is pinned is listed in tests/test_oracle*.py):
//   tree shape / node ids ........ closed form of count_nodes + invariants
//   split values / split dims .... brute-force invariants on small inputs
//   row permutation, leaf order .. literal restatement of partition_rows only
//                                  (the reference API never exposes it)
//   query indices ................ SciPy fixtures, brute force, lattice ties
//   distances (L1/L2/Linf) ....... SciPy fixtures (1e-12) + per-width
//                                  summation closed forms (terms 1 and
//                                  2^-53: a wrong pairing flips the last bit)
//   distances (general p) ........ SciPy fixtures (1e-12); bits of powf are
//                                  platform libm, so bit-level parity unpinned

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <map>
#include <thread>
#include <vector>

namespace {

constexpr double INF = std::numeric_limits<double>::infinity();

// src/build.rs:13, src/build.rs:17
constexpr double BOX_ORDER_MAX_SHRINK_RATIO = 0.6;
constexpr size_t BOX_ORDER_MIN_SUBTREE = 64;
// src/query.rs:26
constexpr size_t PARALLEL_QUERY_MIN_CHUNK = 16;
// src/simd.rs:6
constexpr size_t LANES = 8;
// src/kernel.rs:235, src/kernel.rs:239
constexpr size_t MAX_PACKED_WIDTH = LANES;
constexpr size_t MAX_FOLDED_WIDTH = 2 * LANES;

// Error codes, one per KDTreeError variant (src/error.rs:3-16).
enum Err : int {
  OK = 0,
  EMPTY_DATA = 1,
  INVALID_LEAFSIZE = 2,
  INVALID_K = 3,
  INVALID_METRIC = 4,
  INVALID_EPS = 5,
  INVALID_MAX_DISTANCE = 6,
  NON_FINITE = 7,
  TOO_MANY_POINTS = 8,
  TOO_MANY_NODES = 9,
  INVALID_SHAPE = 10,
  DIMENSION_MISMATCH = 11,
};

// ---------------------------------------------------------------- metric.rs
enum class MK { L1, L2, LInf, LP };
struct Metric {
  MK kind;
  double p;
};

// Metric::new, src/metric.rs:45-59
int metric_new(double p, Metric* out) {
  if (std::isinf(p) && p > 0) { *out = {MK::LInf, p}; return OK; }
  if (!std::isfinite(p) || p < 1.0) return INVALID_METRIC;
  if (p == 1.0) *out = {MK::L1, p};
  else if (p == 2.0) *out = {MK::L2, p};
  else *out = {MK::LP, p};
  return OK;
}

// Rust f64::max / f64::min on non-NaN operands (IEEE maxNum/minNum).
inline double fmax_(double a, double b) { return std::fmax(a, b); }
inline double fmin_(double a, double b) { return std::fmin(a, b); }

// Metric::reduce, src/metric.rs:63-71
inline double reduce(const Metric& m, double x) {
  switch (m.kind) {
    case MK::L2: return x * x;
    case MK::LP: return std::pow(x, m.p);
    default: return x;
  }
}
// Metric::restore, src/metric.rs:73-80
inline double restore(const Metric& m, double x) {
  switch (m.kind) {
    case MK::L2: return std::sqrt(x);
    case MK::LP: return std::pow(x, 1.0 / m.p);
    default: return x;
  }
}
// Metric::fold, src/metric.rs:82-88
inline double fold(const Metric& m, double dist, double axis) {
  return m.kind == MK::LInf ? fmax_(dist, axis) : dist + axis;
}
// Metric::replace_axis, src/metric.rs:90-98
inline double replace_axis(const Metric& m, double dist, double old, double nw) {
  return m.kind == MK::LInf ? fmax_(dist, nw) : dist - old + nw;
}

// ------------------------------------------------------------------ simd.rs
// vmax / vmin: compare-select, src/simd.rs:12-21
inline double vmax(double a, double b) { return b > a ? b : a; }
inline double vmin(double a, double b) { return b < a ? b : a; }
// hsum: pairwise tree, src/simd.rs:25-30
inline double hsum(const double v[8]) {
  double s0 = v[0] + v[4], s1 = v[1] + v[5], s2 = v[2] + v[6], s3 = v[3] + v[7];
  double t0 = s0 + s2, t1 = s1 + s3;
  return t0 + t1;
}
// hmax, src/simd.rs:34-42
inline double hmax(const double v[8]) {
  double m0 = vmax(v[0], v[4]), m1 = vmax(v[1], v[5]), m2 = vmax(v[2], v[6]), m3 = vmax(v[3], v[7]);
  double n0 = vmax(m0, m2), n1 = vmax(m1, m3);
  return n1 > n0 ? n1 : n0;
}

// ---------------------------------------------------------------- layout.rs
// axis_offset, src/layout.rs:145-149
inline double axis_offset(double q, double lo, double hi) {
  return fmax_(fmax_(lo - q, q - hi), 0.0);
}

// --------------------------------------------------------------- kernel.rs
// axes_lanes, src/kernel.rs:22-29 (one lane)
inline double axes_lane(const Metric& m, double delta) {
  return m.kind == MK::L2 ? delta * delta : std::fabs(delta);
}
// combine, src/kernel.rs:31-37 (one lane)
inline double combine(const Metric& m, double a, double b) {
  return m.kind == MK::LInf ? vmax(a, b) : a + b;
}
// fold_scalar, src/kernel.rs:39-65
inline double fold_scalar(const Metric& m, double dist, const double* lhs, const double* rhs, size_t n) {
  for (size_t i = 0; i < n; ++i) {
    double delta = lhs[i] - rhs[i];
    if (m.kind == MK::L1) dist += std::fabs(delta);
    else if (m.kind == MK::L2) dist += delta * delta;
    else {
      double a = std::fabs(delta);
      if (a > dist) dist = a;
    }
  }
  return dist;
}

// all_finite, src/kernel.rs:67-74
bool all_finite(const double* v, size_t n) {
  for (size_t i = 0; i < n; ++i)
    if (!std::isfinite(v[i])) return false;
  return true;
}

// point_dist_sum, src/kernel.rs:177-188
double point_dist_sum(const Metric& m, const double* l, const double* r, size_t d, double bound) {
  size_t nch = d / LANES;
  double dist = 0.0;
  for (size_t c = 0; c < nch; ++c) {
    double a[8];
    for (size_t j = 0; j < 8; ++j) a[j] = axes_lane(m, l[c * 8 + j] - r[c * 8 + j]);
    dist += hsum(a);
    if (dist > bound) return dist;
  }
  return fold_scalar(m, dist, l + nch * 8, r + nch * 8, d - nch * 8);
}
// point_dist_linf, src/kernel.rs:190-205
double point_dist_linf(const double* l, const double* r, size_t d, double bound) {
  size_t nch = d / LANES;
  Metric linf{MK::LInf, INF};
  if (nch == 0) return fold_scalar(linf, 0.0, l, r, d);
  double vm[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (size_t c = 0; c < nch; ++c) {
    bool any = false;
    for (size_t j = 0; j < 8; ++j) {
      vm[j] = vmax(vm[j], std::fabs(l[c * 8 + j] - r[c * 8 + j]));
      any |= vm[j] > bound;
    }
    if (any) return hmax(vm);
  }
  return fold_scalar(linf, hmax(vm), l + nch * 8, r + nch * 8, d - nch * 8);
}
// point_dist_lp, src/kernel.rs:207-218
double point_dist_lp(double p, const double* l, const double* r, size_t d, double bound) {
  double dist = 0.0;
  for (size_t c = 0; c < d; c += LANES) {
    size_t e = std::min(d, c + LANES);
    for (size_t i = c; i < e; ++i) dist += std::pow(std::fabs(l[i] - r[i]), p);
    if (dist > bound) return dist;
  }
  return dist;
}
// point_dist, src/kernel.rs:167-175
double point_dist(const Metric& m, const double* l, const double* r, size_t d, double bound) {
  switch (m.kind) {
    case MK::L1:
    case MK::L2: return point_dist_sum(m, l, r, d, bound);
    case MK::LInf: return point_dist_linf(l, r, d, bound);
    default: return point_dist_lp(m.p, l, r, d, bound);
  }
}
// point_dist_folded, src/kernel.rs:343-354
double point_dist_folded(const Metric& m, const double* l, const double* r, size_t d) {
  size_t nch = d / LANES;
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (size_t c = 0; c < nch; ++c)
    for (size_t j = 0; j < 8; ++j) acc[j] += axes_lane(m, l[c * 8 + j] - r[c * 8 + j]);
  double dist = nch == 0 ? 0.0 : hsum(acc);
  return fold_scalar(m, dist, l + nch * 8, r + nch * 8, d - nch * 8);
}

// box_dist_sum / box_dist_linf / box_dist_lp / box_dist, src/kernel.rs:565-616
double box_dist(const Metric& m, const double* q, const double* lo, const double* hi, size_t d) {
  if (m.kind == MK::L1 || m.kind == MK::L2) {
    size_t nch = d / LANES;
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (size_t c = 0; c < nch; ++c)
      for (size_t j = 0; j < 8; ++j) {
        size_t i = c * 8 + j;
        double off = vmax(vmax(lo[i] - q[i], q[i] - hi[i]), 0.0);
        acc[j] += axes_lane(m, off);
      }
    double dist = nch == 0 ? 0.0 : hsum(acc);
    for (size_t i = nch * 8; i < d; ++i) dist += reduce(m, axis_offset(q[i], lo[i], hi[i]));
    return dist;
  }
  if (m.kind == MK::LInf) {
    double worst = 0.0;
    for (size_t i = 0; i < d; ++i) {
      double off = axis_offset(q[i], lo[i], hi[i]);
      if (off > worst) worst = off;
    }
    return worst;
  }
  double dist = 0.0;
  for (size_t i = 0; i < d; ++i) {
    double off = axis_offset(q[i], lo[i], hi[i]);
    if (off > 0.0) dist += std::pow(off, m.p);
  }
  return dist;
}

// ------------------------------------------------------------------ tree.rs
struct Node {  // src/tree.rs:11-27
  bool leaf;
  uint32_t start, end;         // leaf
  uint32_t left, right;        // inner
  uint32_t split_dim;
  bool order_by_box;
  double split_value;
};

struct Tree {  // src/tree.rs:29-36
  std::vector<double> data;
  std::vector<uint32_t> indices;
  std::vector<Node> nodes;
  std::vector<double> boxes;  // [lo(ndim) | hi(ndim)] per node, src/layout.rs:197-217
  size_t ndim, leafsize;
  size_t n_points() const { return data.size() / ndim; }
  const double* lo(uint32_t id) const { return &boxes[2 * ndim * id]; }
  const double* hi(uint32_t id) const { return &boxes[2 * ndim * id + ndim]; }
};

// ----------------------------------------------------------------- build.rs
// split_point, src/build.rs:161-164
inline size_t split_point(size_t len) { return len / 2; }

// count_nodes, src/build.rs:166-173 (memoised: the value depends on len only)
struct NodeCounter {
  size_t leafsize;
  std::map<size_t, uint64_t> memo;
  uint64_t operator()(size_t len) {
    if (len <= leafsize) return 1;
    auto it = memo.find(len);
    if (it != memo.end()) return it->second;
    size_t mid = split_point(len);
    uint64_t c = 1 + (*this)(mid) + (*this)(len - mid);
    memo[len] = c;
    return c;
  }
};

// bbox_kernel, src/kernel.rs:93-151 (min/max are order-free on finite values;
// the compare-select of vmin/vmax is kept for the choice between equal values)
bool bbox(const double* rows, size_t len, size_t d, double* lo, double* hi, bool check) {
  for (size_t j = 0; j < d; ++j) { lo[j] = INF; hi[j] = -INF; }
  for (size_t r = 0; r < len; ++r)
    for (size_t j = 0; j < d; ++j) {
      double v = rows[r * d + j];
      if (check && !std::isfinite(v)) return false;
      lo[j] = vmin(v, lo[j]);
      hi[j] = vmax(v, hi[j]);
    }
  return true;
}

// BBox::widest_axis, src/layout.rs:168-179
size_t widest_axis(const double* lo, const double* hi, size_t d) {
  size_t best_dim = 0;
  double best_span = hi[0] - lo[0];
  for (size_t dim = 1; dim < d; ++dim) {
    double span = hi[dim] - lo[dim];
    if (span > best_span) { best_span = span; best_dim = dim; }
  }
  return best_dim;
}

// shrank_off_axis, src/build.rs:248-252
bool shrank_off_axis(const double* plo, const double* phi, const double* clo, const double* chi,
                     size_t d, size_t split_dim) {
  for (size_t j = 0; j < d; ++j) {
    if (j == split_dim) continue;
    if (!((chi[j] - clo[j]) < BOX_ORDER_MAX_SHRINK_RATIO * (phi[j] - plo[j]))) return false;
  }
  return true;
}

// f64::total_cmp, as used by select_nth_unstable_by at src/build.rs:204
inline int64_t total_key(double x) {
  int64_t b;
  std::memcpy(&b, &x, 8);
  b ^= (int64_t)(((uint64_t)(b >> 63)) >> 1);
  return b;
}

struct Builder {
  Tree& t;
  NodeCounter& count;
  std::vector<double>& keys;
  std::vector<uint32_t>& marks;

  void swap_rows(size_t base, size_t a, size_t b) {  // RowsMut::swap, src/layout.rs:129-142
    if (a == b) return;
    size_t d = t.ndim;
    double* p = t.data.data();
    for (size_t j = 0; j < d; ++j) std::swap(p[(base + a) * d + j], p[(base + b) * d + j]);
    std::swap(t.indices[base + a], t.indices[base + b]);
  }

  // partition_rows, src/build.rs:266-314
  void partition_rows(size_t first, size_t len, size_t split_dim, double pivot, size_t lt, size_t mid) {
    const size_t d = t.ndim;
    const double* data = t.data.data();
    auto coord = [&](size_t row) { return t.data[(first + row) * d + split_dim]; };
    (void)data;
    uint32_t* right_of_lt = marks.data() + first;       // marks[..lt]
    uint32_t* left_of_ge = marks.data() + first + lt;   // marks[lt..len]
    size_t n_right = 0;
    for (size_t row = 0; row < lt; ++row) {
      right_of_lt[n_right] = (uint32_t)row;
      n_right += coord(row) >= pivot ? 1 : 0;
    }
    size_t n_left = 0;
    for (size_t row = lt; row < len; ++row) {
      left_of_ge[n_left] = (uint32_t)row;
      n_left += coord(row) < pivot ? 1 : 0;
    }
    for (size_t pair = 0; pair < n_right; ++pair) {
      size_t a = right_of_lt[pair];
      size_t b = left_of_ge[n_left - 1 - pair];
      swap_rows(first, a, b);
    }
    size_t place = lt, scan = lt;
    while (place < mid) {
      if (coord(scan) == pivot) {
        swap_rows(first, scan, place);
        place += 1;
      }
      scan += 1;
    }
  }

  // build_subtree, src/build.rs:184-246 (own box already written)
  void build_subtree(size_t first, size_t len, uint32_t id) {
    const size_t d = t.ndim;
    if (len <= t.leafsize) {
      Node& nd = t.nodes[id];
      nd = Node{};
      nd.leaf = true;
      nd.start = (uint32_t)first;
      nd.end = (uint32_t)(first + len);
      return;
    }
    size_t split_dim = widest_axis(t.lo(id), t.hi(id), d);
    size_t mid = split_point(len);
    for (size_t r = 0; r < len; ++r) keys[first + r] = t.data[(first + r) * d + split_dim];
    double* kb = keys.data() + first;
    std::nth_element(kb, kb + mid, kb + len,
                     [](double a, double b) { return total_key(a) < total_key(b); });
    double pivot = kb[mid];
    size_t lt = 0;
    for (size_t r = 0; r < mid; ++r) lt += kb[r] < pivot ? 1 : 0;
    partition_rows(first, len, split_dim, pivot, lt, mid);

    uint64_t left_nodes = count(mid);
    uint32_t left_id = id + 1;
    uint32_t right_id = (uint32_t)(id + 1 + left_nodes);
    bound_and_build(first, mid, left_id);            // bound_and_build, src/build.rs:175-181
    bound_and_build(first + mid, len - mid, right_id);

    bool order_by_box = d > 1 && len >= BOX_ORDER_MIN_SUBTREE &&
                        shrank_off_axis(t.lo(id), t.hi(id), t.lo(left_id), t.hi(left_id), d, split_dim) &&
                        shrank_off_axis(t.lo(id), t.hi(id), t.lo(right_id), t.hi(right_id), d, split_dim);
    Node& nd = t.nodes[id];
    nd = Node{};
    nd.leaf = false;
    nd.left = left_id;
    nd.right = right_id;
    nd.split_dim = (uint32_t)split_dim;
    nd.order_by_box = order_by_box;
    nd.split_value = pivot;
  }

  void bound_and_build(size_t first, size_t len, uint32_t id) {
    double* lo = &t.boxes[2 * t.ndim * id];
    bbox(&t.data[first * t.ndim], len, t.ndim, lo, lo + t.ndim, false);
    build_subtree(first, len, id);
  }
};

// build, src/build.rs:19-82
int build(const double* data, uint64_t n, uint64_t d, uint64_t leafsize, Tree** out) {
  if (leafsize == 0) return INVALID_LEAFSIZE;
  if (d == 0 || n == 0) return EMPTY_DATA;
  if (n > 0xFFFFFFFFull) return TOO_MANY_POINTS;
  NodeCounter count{(size_t)leafsize, {}};
  uint64_t n_nodes = count((size_t)n);
  if (n_nodes > 0xFFFFFFFFull) return TOO_MANY_NODES;

  Tree* t = new Tree();
  t->ndim = d;
  t->leafsize = leafsize;
  t->data.assign(data, data + n * d);
  t->indices.resize(n);
  for (uint64_t i = 0; i < n; ++i) t->indices[i] = (uint32_t)i;
  t->nodes.assign(n_nodes, Node{});
  t->boxes.assign(2 * d * n_nodes, 0.0);
  std::vector<double> keys(n, 0.0);
  std::vector<uint32_t> marks(n, 0);
  if (!bbox(t->data.data(), n, d, &t->boxes[0], &t->boxes[d], true)) {
    delete t;
    return NON_FINITE;
  }
  Builder b{*t, count, keys, marks};
  b.build_subtree(0, n, 0);
  *out = t;
  return OK;
}

// ----------------------------------------------------------------- query.rs
struct Counters {
  uint64_t points = 0;  // leaf rows scanned
  uint64_t nodes = 0;   // descend() calls
  uint64_t boxes = 0;   // box_dist() evaluations
};

// BestK, src/query.rs:432-494 (Best1, :390-430, is the k == 1 case of it)
struct BestK {
  size_t k;
  std::vector<double> dists;
  std::vector<uint32_t> ids;
  void reset() { dists.clear(); ids.clear(); }
  double bound(double cutoff) const {
    return dists.size() < k ? cutoff : fmin_(dists[k - 1], cutoff);
  }
  void consider(double dist, uint32_t idx) {
    if (dists.size() == k) {
      double worst = dists[k - 1];
      if (dist > worst || (dist == worst && ids[k - 1] <= idx)) return;
      dists.pop_back();
      ids.pop_back();
    }
    size_t pos = dists.size();
    while (pos > 0) {
      double prev = dists[pos - 1];
      if (prev < dist || (prev == dist && ids[pos - 1] < idx)) break;
      pos -= 1;
    }
    dists.insert(dists.begin() + pos, dist);
    ids.insert(ids.begin() + pos, idx);
  }
  void write_results(double* od, int64_t* oi, size_t kk, size_t n_points, const Metric& m) const {
    for (size_t j = 0; j < kk; ++j) {
      if (j < dists.size()) { od[j] = restore(m, dists[j]); oi[j] = ids[j]; }
      else { od[j] = INF; oi[j] = (int64_t)n_points; }
    }
  }
};

enum class Plan { Packed, FoldedAbs, FoldedSquared, Streamed };

// with_plan!, src/kernel.rs:244-270
Plan plan_for(const Metric& m, size_t d) {
  if ((m.kind == MK::L1 || m.kind == MK::L2 || m.kind == MK::LInf) && d <= MAX_PACKED_WIDTH) return Plan::Packed;
  if (m.kind == MK::L1 && d <= MAX_FOLDED_WIDTH) return Plan::FoldedAbs;
  if (m.kind == MK::L2 && d <= MAX_FOLDED_WIDTH) return Plan::FoldedSquared;
  return Plan::Streamed;
}

struct Descent {
  const Tree& t;
  Metric m;
  double cutoff, eps_factor;
  Plan plan;
  const double* q;
  BestK& best;
  std::vector<double>& axes;  // CellBound::axes, src/query.rs:156-159
  Counters& ctr;

  double bound() const { return best.bound(cutoff); }
  bool admits_at(double dist, double b) const { return dist * eps_factor <= b; }  // :242-244
  bool admits(double dist) const { return admits_at(dist, bound()); }            // :235-237
  double boxd(uint32_t id) {
    ctr.boxes++;
    return box_dist(m, q, t.lo(id), t.hi(id), t.ndim);
  }

  // LeafSink::offer, src/query.rs:371-374
  double offer(size_t start, size_t offset, double dist) {
    best.consider(dist, t.indices[start + offset]);
    return bound();
  }

  // Scan::unrolled_from, src/kernel.rs:420-435
  double unrolled_from(size_t start, size_t base, size_t count, double b) {
    const size_t d = t.ndim;
    for (size_t j = 0; j < count; ++j) {
      const double* p = &t.data[(start + base + j) * d];
      double dist = fold_scalar(m, 0.0, q, p, d);
      if (dist <= b) b = offer(start, base + j, dist);
    }
    return b;
  }

  // Scan::packed + Scan::emit + the width recipes d2/d3/d4/d8, src/kernel.rs:441-562
  double packed(size_t start, size_t len, double b) {
    const size_t d = t.ndim;
    size_t P;
    switch (d) {
      case 2: P = 4; break;
      case 3: P = 8; break;
      case 4: P = 2; break;
      case 8: P = 8; break;
      default: return unrolled_from(start, 0, len, b);  // unrolled::<1|5|6|7>, :281-287
    }
    size_t nchunks = len / P;
    for (size_t c = 0; c < nchunks; ++c) {
      double dists[8];
      for (size_t j = 0; j < P; ++j) {
        const double* p = &t.data[(start + c * P + j) * d];
        double a[8];
        for (size_t i = 0; i < d; ++i) a[i] = axes_lane(m, p[i] - q[i]);
        switch (d) {
          case 2: dists[j] = combine(m, a[0], a[1]); break;                              // d2 :472-481
          case 3: dists[j] = combine(m, combine(m, a[0], a[1]), a[2]); break;             // d3 :487-513
          case 4: dists[j] = combine(m, combine(m, a[0], a[2]), combine(m, a[1], a[3]));  // d4 :517-533
            break;
          default:                                                                         // d8 :540-562
            dists[j] = combine(m, combine(m, combine(m, a[0], a[1]), combine(m, a[2], a[3])),
                               combine(m, combine(m, a[4], a[5]), combine(m, a[6], a[7])));
        }
      }
      bool any = false;
      for (size_t j = 0; j < P; ++j) any |= dists[j] <= b;
      if (any)
        for (size_t j = 0; j < P; ++j)
          if (dists[j] <= b) b = offer(start, c * P + j, dists[j]);
    }
    return unrolled_from(start, nchunks * P, len - nchunks * P, b);
  }

  // scan_leaf, src/query.rs:343-353, with Strategy::scan, src/kernel.rs:277-341
  void scan_leaf(size_t start, size_t end) {
    const size_t d = t.ndim, len = end - start;
    ctr.points += len;
    double b = bound();
    switch (plan) {
      case Plan::Packed: packed(start, len, b); return;
      case Plan::FoldedAbs:
      case Plan::FoldedSquared:  // Scan::folded, :388-397
        for (size_t j = 0; j < len; ++j) {
          double dist = point_dist_folded(m, q, &t.data[(start + j) * d], d);
          if (dist <= b) b = offer(start, j, dist);
        }
        return;
      case Plan::Streamed:  // Scan::streamed, :400-409
        for (size_t j = 0; j < len; ++j) {
          double dist = point_dist(m, q, &t.data[(start + j) * d], d, b);
          if (dist <= b) b = offer(start, j, dist);
        }
        return;
    }
  }

  // enter_far, src/query.rs:329-341
  void enter_far(uint32_t far, size_t dim, double far_axis, double far_dist) {
    double saved = axes[dim];
    axes[dim] = far_axis;
    descend(far, far_dist);
    axes[dim] = saved;
  }

  // descend_by_box, src/query.rs:296-327
  void descend_by_box(uint32_t near, uint32_t far, size_t dim, double cell_dist, double far_axis,
                      double far_dist) {
    double near_box = boxd(near);
    double far_box = boxd(far);
    if (near_box <= far_box) {
      if (admits(near_box)) descend(near, cell_dist);
      if (admits(far_box)) enter_far(far, dim, far_axis, far_dist);
    } else {
      if (admits(far_box)) enter_far(far, dim, far_axis, far_dist);
      if (admits(near_box)) descend(near, cell_dist);
    }
  }

  // descend, src/query.rs:248-289
  void descend(uint32_t node_id, double cell_dist) {
    ctr.nodes++;
    const Node& nd = t.nodes[node_id];
    if (nd.leaf) { scan_leaf(nd.start, nd.end); return; }
    size_t dim = nd.split_dim;
    double diff = q[dim] - nd.split_value;
    uint32_t near = diff <= 0.0 ? nd.left : nd.right;
    uint32_t far = diff <= 0.0 ? nd.right : nd.left;
    double far_axis = reduce(m, std::fabs(diff));
    double far_dist = replace_axis(m, cell_dist, axes[dim], far_axis);
    if (nd.order_by_box) {
      descend_by_box(near, far, dim, cell_dist, far_axis, far_dist);
      return;
    }
    descend(near, cell_dist);
    double b = bound();
    if (admits_at(far_dist, b)) {
      if (admits_at(boxd(far), b)) enter_far(far, dim, far_axis, far_dist);
    }
  }

  // Descent::run, src/query.rs:223-232, with CellBound::start/seed/finish, :170-194
  void run(double* od, int64_t* oi, size_t k) {
    best.reset();
    double seed = 0.0;
    bool seeded = false;
    if (boxd(0) != 0.0) {
      for (size_t j = 0; j < t.ndim; ++j) {
        double axis = reduce(m, axis_offset(q[j], t.lo(0)[j], t.hi(0)[j]));
        axes[j] = axis;
        seed = fold(m, seed, axis);
      }
      seeded = true;
    }
    if (admits(seed)) descend(0, seed);
    if (seeded) std::fill(axes.begin(), axes.end(), 0.0);
    best.write_results(od, oi, k, t.n_points(), m);
  }
};

}  // namespace

// ------------------------------------------------------------------- C ABI
extern "C" {

struct kdo_node {  // 24 bytes, the same packing budget as src/tree.rs:97-100
  double split_value;
  uint32_t a;  // leaf: start, inner: left
  uint32_t b;  // leaf: end,   inner: right
  uint32_t split_dim;
  uint8_t is_leaf;
  uint8_t order_by_box;
  uint16_t pad;
};

int kdo_build(const double* data, uint64_t n, uint64_t d, uint64_t leafsize, void** out) {
  Tree* t = nullptr;
  int rc = build(data, n, d, leafsize, &t);
  *out = t;
  return rc;
}

void kdo_free(void* t) { delete static_cast<Tree*>(t); }

uint64_t kdo_n_nodes(const void* t) { return static_cast<const Tree*>(t)->nodes.size(); }

uint64_t kdo_count_nodes(uint64_t n, uint64_t leafsize) {
  NodeCounter c{(size_t)leafsize, {}};
  return c((size_t)n);
}

void kdo_export(const void* tp, double* data, uint32_t* indices, kdo_node* nodes, double* boxes) {
  const Tree* t = static_cast<const Tree*>(tp);
  if (data) std::memcpy(data, t->data.data(), t->data.size() * 8);
  if (indices) std::memcpy(indices, t->indices.data(), t->indices.size() * 4);
  if (boxes) std::memcpy(boxes, t->boxes.data(), t->boxes.size() * 8);
  if (nodes)
    for (size_t i = 0; i < t->nodes.size(); ++i) {
      const Node& s = t->nodes[i];
      kdo_node r{};
      r.split_value = s.leaf ? 0.0 : s.split_value;
      r.a = s.leaf ? s.start : s.left;
      r.b = s.leaf ? s.end : s.right;
      r.split_dim = s.leaf ? 0 : s.split_dim;
      r.is_leaf = s.leaf;
      r.order_by_box = s.leaf ? 0 : s.order_by_box;
      nodes[i] = r;
    }
}

// Tree::query, src/query.rs:29-78, plus the query-dimension check that the
// binding performs first (src/lib.rs:173-174).  threads > 1 splits queries in
// chunks of at least 16 (src/query.rs:104-113).  counters (may be null)
// receives {leaf rows scanned, nodes visited, box tests} summed over queries.
int kdo_query(const void* tp, const double* q, uint64_t nq, uint64_t d, uint64_t k, double p,
              double max_distance, double eps, int threads, double* out_d, int64_t* out_i,
              uint64_t* counters) {
  const Tree* t = static_cast<const Tree*>(tp);
  if (d != t->ndim) return DIMENSION_MISMATCH;
  if (k == 0) return INVALID_K;
  if (nq == 0) return INVALID_SHAPE;
  if (!all_finite(q, nq * d)) return NON_FINITE;
  if (!std::isfinite(eps) || eps < 0.0) return INVALID_EPS;
  if (std::isnan(max_distance) || max_distance < 0.0) return INVALID_MAX_DISTANCE;
  Metric m;
  int rc = metric_new(p, &m);
  if (rc) return rc;
  const double cutoff = reduce(m, max_distance);
  const double eps_factor = reduce(m, 1.0 + eps);
  const Plan plan = plan_for(m, d);

  std::atomic<uint64_t> next{0};
  std::atomic<uint64_t> tot_points{0}, tot_nodes{0}, tot_boxes{0};
  auto worker = [&]() {
    BestK best{(size_t)k, {}, {}};
    best.dists.reserve(k);
    best.ids.reserve(k);
    std::vector<double> axes(d, 0.0);
    Counters ctr;
    const uint64_t chunk = PARALLEL_QUERY_MIN_CHUNK;
    for (;;) {
      uint64_t b = next.fetch_add(chunk);
      if (b >= nq) break;
      uint64_t e = std::min<uint64_t>(nq, b + chunk);
      for (uint64_t i = b; i < e; ++i) {
        Descent ds{*t, m, cutoff, eps_factor, plan, q + i * d, best, axes, ctr};
        ds.run(out_d + i * k, out_i + i * k, k);
      }
    }
    tot_points += ctr.points;
    tot_nodes += ctr.nodes;
    tot_boxes += ctr.boxes;
  };
  if (threads <= 1 || nq <= 1) {
    worker();
  } else {
    std::vector<std::thread> pool;
    for (int i = 0; i < threads; ++i) pool.emplace_back(worker);
    for (auto& th : pool) th.join();
  }
  if (counters) {
    counters[0] = tot_points;
    counters[1] = tot_nodes;
    counters[2] = tot_boxes;
  }
  return OK;
}

// Single-point reduced distance with the leaf-scan recipe the reference uses
// for (metric, ndim) at leaf offset `offset` of a leaf of `len` rows, for the
// per-width known-answer tests.  `bound` = +inf.
double kdo_point_dist(const double* q, const double* pt, uint64_t d, double p, uint64_t offset,
                      uint64_t len) {
  Metric m;
  if (metric_new(p, &m)) return std::numeric_limits<double>::quiet_NaN();
  Plan plan = plan_for(m, d);
  if (plan == Plan::Packed) {
    size_t P = d == 2 ? 4 : d == 3 ? 8 : d == 4 ? 2 : d == 8 ? 8 : 0;
    if (P && offset < P * (len / P)) {
      double a[8];
      for (size_t i = 0; i < d; ++i) a[i] = axes_lane(m, pt[i] - q[i]);
      switch (d) {
        case 2: return combine(m, a[0], a[1]);
        case 3: return combine(m, combine(m, a[0], a[1]), a[2]);
        case 4: return combine(m, combine(m, a[0], a[2]), combine(m, a[1], a[3]));
        default:
          return combine(m, combine(m, combine(m, a[0], a[1]), combine(m, a[2], a[3])),
                         combine(m, combine(m, a[4], a[5]), combine(m, a[6], a[7])));
      }
    }
    return fold_scalar(m, 0.0, q, pt, d);
  }
  if (plan == Plan::Streamed) return point_dist(m, q, pt, d, INF);
  return point_dist_folded(m, q, pt, d);
}

}  // extern "C"
