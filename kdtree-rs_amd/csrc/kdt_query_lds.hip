// kdt_query_lds.hip -- runtime-width kernels with the k list in LDS: every
// metric (general p included), any ndim, any k whose list fits in LDS.
#include "kdt_query_kernels.h"

namespace kdt {

int launch_knn_lds(const KArgs& a, int metric, hipStream_t s) {
  const size_t shm = (size_t)a.D * 8 + (size_t)a.k * 12;
  uint64_t blocks = a.nq < (1u << 20) ? a.nq : (1u << 20);
  // Only the queries a screen left over (a.only): usually none or a few, so a
  // grid of one workgroup per query would mostly start workgroups that find
  // nothing to do; 8192 single-wave groups (more than the device holds at
  // once) each look through a share of the flags.
  if (a.only && blocks > (1u << 13)) blocks = 1u << 13;
  KernelTimer tm(s);
  switch (metric) {
    case M_L1: knn_lds_kernel<M_L1><<<(uint32_t)blocks, 64, shm, s>>>(a); break;
    case M_L2: knn_lds_kernel<M_L2><<<(uint32_t)blocks, 64, shm, s>>>(a); break;
    case M_LINF: knn_lds_kernel<M_LINF><<<(uint32_t)blocks, 64, shm, s>>>(a); break;
    case M_LP: knn_lds_kernel<M_LP><<<(uint32_t)blocks, 64, shm, s>>>(a); break;
    default: return -1;
  }
  char name[48];
  std::snprintf(name, sizeof name, "knn_lds<%d%s>", metric, a.only ? ",flagged" : "");
  tm.stop(name, (double)a.nq);
  return 0;
}

}  // namespace kdt
