// kdt_noalloc.hip -- kdt_query_device_noalloc: the device query with every
// piece of scratch taken from a fixed array in device memory, for callers that
// cannot allocate (code under HIP graph capture, tight serving loops,
// harnesses that forbid allocation).  No reference counterpart.
//
// The scratch is one file-scope __device__ array per GPU (`scratch`, below),
// allocated with the code object when the library loads on that device and
// held until the process ends.  Host code never uses the array's name as an
// address: at the first call on a device (with that device current) the
// array's device address is looked up once with hipGetSymbolAddress and kept
// per device; every kernel argument and every region handed to the query is
// that address.  Each call carves its pieces (finiteness flag, leaf-order
// sort, matrix-core screen lists, tile-walk buffers) from the start of the
// array (a tile walk that does not fit runs as the traversal's FP32 screen); a call whose
// pieces need more than KDT_NOALLOC_SCRATCH_BYTES is rejected before any
// launch.  Calls that overlap on one GPU (from several streams or threads,
// before the previous call's work has finished) share the array and race:
// the caller must order them, e.g. on one stream.

#include <mutex>

#include "../../include/kdtree_hip.h"
#include "kdt_internal.h"

namespace {

constexpr size_t NOALLOC_BYTES = KDT_NOALLOC_SCRATCH_BYTES;
static_assert(NOALLOC_BYTES % 256 == 0, "scratch size is a multiple of the piece alignment");

__device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
void* g_ptr[MAX_DEVICES] = {};  // device address of `scratch` on each device, once looked up

// The array's address on the current device (looked up once per device).
int scratch_ptr(void** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return kdt::fail_api(KDT_ERR_DEVICE, std::string("hipGetDevice: ") + hipGetErrorString(e));
  if (dev < 0 || dev >= MAX_DEVICES) return kdt::fail_api(KDT_ERR_DEVICE, "device ordinal out of range");
  std::lock_guard<std::mutex> g(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return kdt::fail_api(KDT_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_ptr[dev] = p;
  }
  *out = g_ptr[dev];
  return KDT_OK;
}

}  // namespace

extern "C" int kdt_query_device_noalloc(const kdt_tree* tree, const double* d_queries, uint64_t n_queries,
                                        uint64_t ndim, uint64_t k, double p, double max_distance, double eps,
                                        uint32_t flags, double* d_out_dist, int64_t* d_out_idx, void* stream) {
  kdt::QueryParams qp;
  int scalar_rc = 0;
  std::string smsg;
  const int arc = kdt::query_args(tree, d_queries, n_queries, ndim, k, p, max_distance, eps, flags, d_out_dist,
                                  d_out_idx, qp, scalar_rc, smsg);
  if (arc) return arc;
  const bool check = (flags & KDT_QUERY_ASSUME_FINITE) == 0;
  if (!check && scalar_rc) return kdt::fail_api(scalar_rc, smsg);
  const kdt::Tree& t = tree->t;
  size_t need = (check ? kdt::scratch_piece(4) : 0) + kdt::query_scratch_bytes(t, qp);
  if (need > NOALLOC_BYTES && qp.walk == 0 && kdt::walk_applies(t, qp)) {
    // The tile walk's buffers (per resident wavefront) do not fit: the
    // traversal's own FP32 screen gives the same results with less scratch.
    qp.walk = -1;
    need = (check ? kdt::scratch_piece(4) : 0) + kdt::query_scratch_bytes(t, qp);
  }
  if (need > NOALLOC_BYTES)
    return kdt::fail_api(KDT_ERR_ARGUMENT, "query needs " + std::to_string(need) +
                                               " bytes of scratch; kdt_query_device_noalloc holds " +
                                               std::to_string(NOALLOC_BYTES) + " (use fewer queries per call)");
  void* base = nullptr;
  int rc = scratch_ptr(&base);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  kdt::Scratch sc(base, NOALLOC_BYTES, s);
  if (check) {
    // all_finite (src/kernel.rs:67-74) first, as kdt_query_device: the flag is
    // cleared by a kernel, read back after a stream synchronisation, and the
    // query runs only for finite queries.
    uint32_t* flag = nullptr;
    hipError_t e = sc.get(&flag, 1);
    if (e != hipSuccess) return kdt::fail_api(KDT_ERR_ARGUMENT, "scratch: " + std::string(hipGetErrorString(e)));
    kdt::launch_zero_u32(flag, 1, s);
    kdt::launch_check_finite(d_queries, n_queries * ndim, reinterpret_cast<int*>(flag), s);
    uint32_t h = 0;
    e = hipMemcpyAsync(&h, flag, sizeof h, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return kdt::fail_api(KDT_ERR_DEVICE, std::string("finiteness check: ") + hipGetErrorString(e));
    if (h) return kdt::fail_api(KDT_ERR_NON_FINITE, "all coordinates must be finite");
    if (scalar_rc) return kdt::fail_api(scalar_rc, smsg);
  }
  std::string msg;
  rc = kdt::launch_query(t, qp, d_queries, d_out_dist, d_out_idx, s, msg, &sc);
  return rc ? kdt::fail_api(rc, msg) : KDT_OK;
}
