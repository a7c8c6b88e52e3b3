// kdt_query_l2.hip -- Euclidean (p = 2) traversal kernels, FP64 and FP32-screened.
#include "kdt_query_kernels.h"

namespace kdt {

int launch_knn_l2(const KArgs& a, int G, bool screen, hipStream_t s) {
  if (screen) {
    switch (G) {
      case 8: return launch_reg_width<8, M_L2, true>(a, s);
      case 16: return launch_reg_width<16, M_L2, true>(a, s);
      case 32: return launch_reg_width<32, M_L2, true>(a, s);
      case 64: return launch_reg_width<64, M_L2, true>(a, s);
    }
    return -1;
  }
  switch (G) {
    case 8: return launch_reg_width<8, M_L2, false>(a, s);
    case 16: return launch_reg_width<16, M_L2, false>(a, s);
    case 32: return launch_reg_width<32, M_L2, false>(a, s);
    case 64: return launch_reg_width<64, M_L2, false>(a, s);
  }
  return -1;
}

}  // namespace kdt
