// kdt_query_linf.hip -- Chebyshev (p = inf) traversal kernels.
#include "kdt_query_kernels.h"

namespace kdt {

int launch_knn_linf(const KArgs& a, int G, hipStream_t s) {
  switch (G) {
    case 8: return launch_reg_width<8, M_LINF, false>(a, s);
    case 16: return launch_reg_width<16, M_LINF, false>(a, s);
    case 32: return launch_reg_width<32, M_LINF, false>(a, s);
    case 64: return launch_reg_width<64, M_LINF, false>(a, s);
  }
  return -1;
}

}  // namespace kdt
