// kdt_query.hip -- kernel selection for one exact k-NN launch.
//
// Replaces the plan dispatch of Tree::query (src/query.rs:68-75, with_plan!
// src/kernel.rs:244-270): the metric and the row width pick a kernel once per
// call.  Widths 1, 2, 3, 4, 8 and 16 (the reference's monomorphised list,
// src/build.rs:60), 32 and 64 have compiled kernels with the k list in
// registers when k <= 64; everything else runs the runtime-width kernel whose
// k list is in LDS.

#include <charconv>
#include <cmath>
#include <cstdlib>

#include "kdt_query_kernels.h"

namespace kdt {

namespace {
constexpr size_t LDS_LIST_LIMIT = 60000;  // bytes of dynamic LDS per block
constexpr uint64_t ORDER_MIN_QUERIES = 8192; // smaller batches run in the given order

// Widths with compiled register-list kernels: the reference's monomorphised
// list (src/build.rs:60) plus 32 and 64, the wide rows of the serving
// workload, whose runtime-width kernel (one 64-lane group per query, half of
// it idle on 32-row leaves, the k list in LDS) ran several times slower.
bool compiled_width(uint64_t d) {
  return d == 1 || d == 2 || d == 3 || d == 4 || d == 8 || d == 16 || d == 32 || d == 64;
}

}  // namespace

std::string rust_f64(double x) {
  if (std::isnan(x)) return "NaN";
  if (std::isinf(x)) return x > 0 ? "inf" : "-inf";
  std::string out = std::signbit(x) ? "-" : "";
  x = std::fabs(x);
  if (x == 0.0) return out + "0";
  char buf[64];
  auto r = std::to_chars(buf, buf + sizeof buf, x, std::chars_format::scientific);
  std::string sci(buf, r.ptr);
  const size_t epos = sci.find('e');
  std::string mant = sci.substr(0, epos);
  const int exp10 = std::stoi(sci.substr(epos + 1));
  std::string digits;
  for (char c : mant)
    if (c != '.') digits.push_back(c);
  // value = 0.digits * 10^(exp10 + 1)
  const int point = exp10 + 1;
  if (point <= 0) {
    out += "0." + std::string((size_t)(-point), '0') + digits;
  } else if ((size_t)point >= digits.size()) {
    out += digits + std::string((size_t)point - digits.size(), '0');
  } else {
    out += digits.substr(0, (size_t)point) + "." + digits.substr((size_t)point);
  }
  return out;
}

int query_params(double p, double max_distance, double eps, QueryParams& qp, std::string& msg) {
  if (!std::isfinite(eps) || eps < 0.0) {
    msg = "eps must be finite and non-negative, got " + rust_f64(eps);
    return 5;
  }
  if (std::isnan(max_distance) || max_distance < 0.0) {
    msg = "max_distance must be finite and non-negative, or infinity, got " + rust_f64(max_distance);
    return 6;
  }
  // Metric::new, src/metric.rs:45-59
  int metric;
  if (std::isinf(p) && p > 0) metric = M_LINF;
  else if (!std::isfinite(p) || p < 1.0) {
    msg = "p must be finite and >= 1, or infinity, got " + rust_f64(p);
    return 4;
  } else if (p == 1.0) metric = M_L1;
  else if (p == 2.0) metric = M_L2;
  else metric = M_LP;
  auto reduce = [&](double x) {  // Metric::reduce, src/metric.rs:63-71
    if (metric == M_L2) return x * x;
    if (metric == M_LP) return std::pow(x, p);
    return x;
  };
  qp.metric = metric;
  qp.p = p;
  qp.cutoff = reduce(max_distance);
  qp.eps_factor = reduce(1.0 + eps);
  return 0;
}

int build_kargs(const Tree& t, const QueryParams& qp, const double* d_q, double* d_out, int64_t* d_idx,
                KArgs& a) {
  a = KArgs{};
  a.nodes = t.nodes;
  a.boxes = t.boxes;
  a.pts = t.pts;
  a.pts32 = t.pts32;
  a.rows = t.rows;
  a.idx = t.idx;
  a.q = d_q;
  a.out_d = d_out;
  a.out_i = d_idx;
  a.nq = qp.nq;
  a.D = (uint32_t)t.d;
  a.k = (uint32_t)qp.k;
  a.n_points = (uint32_t)t.n;
  a.n_nodes = (uint32_t)t.n_nodes;
  a.cutoff = qp.cutoff;
  a.eps_factor = qp.eps_factor;
  a.p = qp.p;
  a.inv_p = 1.0 / qp.p;
  const double D = (double)t.d;
  a.pn32 = t.pn32;
  a.pmax = t.pmax;
  a.screen_shrink = 1.0 + 2.0 * (D + 3.0) * std::ldexp(1.0, -53);
  a.max_abs = t.max_abs;
  a.stk = std::max<uint32_t>(2, 2 * t.depth);
  return 0;
}

bool order_applies(const QueryParams& qp) {
  return qp.allow_order && qp.nq >= ORDER_MIN_QUERIES && qp.nq <= 0xFFFFFFFFull;
}

size_t query_scratch_bytes(const Tree& t, const QueryParams& qp) {
  if (qp.nq == 0) return 0;
  if (screen_applies(t, qp)) return screen_scratch_bytes(t, qp);
  if (walk_applies(t, qp)) return walk_scratch_bytes(t, qp);
  if (order_applies(qp)) return scratch_piece(qp.nq * 4) + order_scratch_bytes(t, qp.nq);
  return 0;
}

int launch_query(const Tree& t, const QueryParams& qp, const double* d_q, double* d_out,
                 int64_t* d_idx, hipStream_t s, std::string& msg, Scratch* region) {
  KArgs a;
  build_kargs(t, qp, d_q, d_out, d_idx, a);
  if (qp.nq == 0) return 0;
  Scratch local(s);  // stream-ordered scratch when no region is given
  Scratch& sc = region ? *region : local;
  if (screen_applies(t, qp)) return launch_screen(t, qp, a, sc, s, msg);
  if (walk_applies(t, qp)) return launch_walk(t, qp, a, sc, s, msg);

  // Large batches run grouped by leaf (see kdt_order.hip).
  if (order_applies(qp)) {
    uint32_t* order = nullptr;
    const hipError_t e = sc.get(&order, qp.nq);
    if (e != hipSuccess) {
      msg = std::string("query ordering: ") + hipGetErrorString(e);
      return 100;
    }
    const int rc = order_queries(t, d_q, qp.nq, order, sc, s, msg);
    if (rc) return rc;
    a.order = order;
  }

  const bool reg_ok = qp.metric != M_LP && compiled_width(t.d) && qp.k <= 64;
  int rc = -1;
  if (reg_ok) {
    int G = group_width(qp.k, qp.group);
    // Wide rows: no 8-lane instances (compile time, LDS per block), and 32
    // lanes by default for k <= 32 (a 32-row leaf in one pass: 64-d rows,
    // 10^6 points, k = 10: 9.6 K queries/s vs 6.7 K with 16 lanes, 32-d 18.6 K
    // vs 15.7 K).
    if (t.d >= 32 && (G == 8 || (qp.group == 0 && G == 16))) G = qp.group == 0 ? 32 : 16;
    // The traversal's FP32 leaf screen reads the FP32 leaf buckets, which the
    // build keeps for rows of up to 16 coordinates, and defers its FP64
    // rescoring, which leaves results unchanged only for eps = 0.
    const bool screen = qp.fp32_screen && qp.metric == M_L2 && t.pts32 != nullptr && qp.eps_factor == 1.0;
    switch (qp.metric) {
      case M_L1: rc = launch_knn_l1(a, G, s); break;
      case M_L2: rc = launch_knn_l2(a, G, screen, s); break;
      case M_LINF: rc = launch_knn_linf(a, G, s); break;
    }
  } else {
    if ((size_t)t.d * 8 + (size_t)qp.k * 12 > LDS_LIST_LIMIT) {
      msg = "k and ndim too large for the device k list (8*ndim + 12*k must not exceed " +
            std::to_string(LDS_LIST_LIMIT) + " bytes)";
      return 101;
    }
    rc = launch_knn_lds(a, qp.metric, s);
  }
  if (rc != 0) {
    msg = "no kernel for this query configuration";
    return 101;
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    msg = std::string("query launch: ") + hipGetErrorString(e);
    return 100;
  }
  return 0;
}

}  // namespace kdt
