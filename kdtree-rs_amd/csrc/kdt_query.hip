// kdt_query.hip -- kernel selection for one exact k-NN launch.
//
// Replaces the plan dispatch of Tree::query (src/query.rs:68-75, with_plan!
// src/kernel.rs:244-270): the metric and the row width pick a kernel once per
// call.  Widths 1, 2, 3, 4, 8 and 16 (the reference's monomorphised list,
// src/build.rs:60) have compiled kernels with the k list in registers when
// k <= 64; everything else runs the runtime-width kernel whose k list is in LDS.

#include <cmath>

#include "kdt_query_kernels.h"

namespace kdt {

namespace {
constexpr size_t LDS_LIST_LIMIT = 60000;  // bytes of dynamic LDS per block

bool compiled_width(uint64_t d) { return d == 1 || d == 2 || d == 3 || d == 4 || d == 8 || d == 16; }

// Lanes per query: enough for the k list and for one leaf (up to 64) per pass.
int group_width(uint64_t k, uint64_t leafsize) {
  const uint64_t need = std::max<uint64_t>(k, std::min<uint64_t>(leafsize, 64));
  return need <= 16 ? 16 : need <= 32 ? 32 : 64;
}
}  // namespace

int launch_query(const Tree& t, const QueryParams& qp, const double* d_q, double* d_out,
                 int64_t* d_idx, int* d_err, hipStream_t s, std::string& msg) {
  KArgs a{};
  a.nodes = t.nodes;
  a.boxes = t.boxes;
  a.pts = t.pts;
  a.pts32 = t.pts32;
  a.idx = t.idx;
  a.q = d_q;
  a.out_d = d_out;
  a.out_i = d_idx;
  a.nq = qp.nq;
  a.D = (uint32_t)t.d;
  a.k = (uint32_t)qp.k;
  a.n_points = (uint32_t)t.n;
  a.n_nodes = (uint32_t)t.n_nodes;
  a.err = d_err;
  a.cutoff = qp.cutoff;
  a.eps_factor = qp.eps_factor;
  a.p = qp.p;
  a.inv_p = 1.0 / qp.p;
  const double u = std::ldexp(1.0, -24);
  const double D = (double)t.d;
  a.screen_e = 4.0001 * std::sqrt(D) * u;
  a.screen_gam = (1.0 + D * u / (1.0 - D * u)) * (1.0 + 1e-14);
  a.screen_shrink = 1.0 + 2.0 * (D + 3.0) * std::ldexp(1.0, -53);
  a.max_abs = t.max_abs;
  if (qp.nq == 0) return 0;

  const bool reg_ok = qp.metric != M_LP && compiled_width(t.d) && qp.k <= 64;
  int rc = -1;
  if (reg_ok) {
    const int G = group_width(qp.k, t.leafsize);
    const bool screen = qp.fp32_screen && qp.metric == M_L2 && t.d <= 64;
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
