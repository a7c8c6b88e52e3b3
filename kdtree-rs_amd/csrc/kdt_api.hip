// kdt_api.hip -- the C ABI (include/kdtree_hip.h): validation in the
// reference's order, its error text, device memory and copies.
//
// Replaces the Rust core's public surface: Tree::new / build (src/tree.rs:39-46,
// src/build.rs:19-82), Tree::query (src/query.rs:29-78), Tree::original_data
// (src/tree.rs:60-68) and KDTreeError's Display (src/error.rs:18-51).

#include <algorithm>
#include <atomic>
#include <charconv>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/kdtree_hip.h"
#include "kdt_internal.h"

namespace kdt {
uint64_t count_tree_nodes(uint64_t n, uint64_t leafsize);
}

namespace {
thread_local std::string g_last_error;

// Kernel timing records (kdt_timing_enable / kdt_timing_read).
struct TimedLaunch {
  std::string name;
  double work;
  hipEvent_t a, b;
};
std::atomic<bool> g_timing{false};
std::mutex g_timing_mu;
std::vector<TimedLaunch> g_timed;
hipEvent_t g_timing_ref = nullptr;  // recorded when timing is enabled: the origin of the busy intervals
}  // namespace

void kdt::set_last_error(std::string msg) { g_last_error = std::move(msg); }

kdt::KernelTimer::KernelTimer(hipStream_t s) : s_(s) {
  if (!g_timing.load(std::memory_order_relaxed)) return;
  if (hipEventCreate(&start_) != hipSuccess) {
    start_ = nullptr;
    return;
  }
  if (hipEventRecord(start_, s_) != hipSuccess) {
    (void)hipEventDestroy(start_);
    start_ = nullptr;
  }
}

kdt::KernelTimer::~KernelTimer() {
  if (start_) (void)hipEventDestroy(start_);
}

void kdt::KernelTimer::stop(const char* name, double work) {
  if (!start_) return;
  hipEvent_t b = nullptr;
  if (hipEventCreate(&b) != hipSuccess) return;
  if (hipEventRecord(b, s_) != hipSuccess) {
    (void)hipEventDestroy(b);
    return;
  }
  std::lock_guard<std::mutex> g(g_timing_mu);
  g_timed.push_back({name, work, start_, b});
  start_ = nullptr;
}

namespace {

int fail(int code, std::string msg) {
  g_last_error = std::move(msg);
  return code;
}

hipStream_t host_stream() { return hipStreamPerThread; }

template <typename T>
struct DevMem {
  T* p = nullptr;
  hipStream_t s = nullptr;
  ~DevMem() {
    if (p) (void)hipFreeAsync(p, s);
  }
  hipError_t alloc(size_t n, hipStream_t st) {
    s = st;
    return hipMallocAsync((void**)&p, std::max<size_t>(n, 1) * sizeof(T), st);
  }
};

// Copies between device memory and the caller's pageable host arrays.  A
// pageable copy goes through the runtime's own staging at a fraction of the
// link rate, so large copies are staged here instead: STAGE_WORKERS host
// threads each take STAGE_BYTES chunks and move them through their own pinned
// buffer on their own stream (device -> pinned -> caller's array, or the
// reverse), so that one worker's DMA overlaps the others' host copies.  The
// pinned buffers (hipHostMalloc) and streams (default priority) are made once
// per device and kept until the process ends; one call at a time uses them.
constexpr size_t STAGE_BYTES = 8u << 20;
constexpr int STAGE_WORKERS = 8;

struct Staging {
  std::mutex lock;
  void* buf[STAGE_WORKERS] = {};
  hipStream_t ws[STAGE_WORKERS] = {};
  bool ok = false;
};

Staging* staging_for(int dev) {
  static std::mutex m;
  static std::map<int, Staging*> per_device;
  std::lock_guard<std::mutex> g(m);
  Staging*& st = per_device[dev];
  if (!st) {
    st = new Staging;
    bool ok = true;
    for (int w = 0; w < STAGE_WORKERS && ok; ++w)
      ok = hipHostMalloc(&st->buf[w], STAGE_BYTES, hipHostMallocDefault) == hipSuccess &&
           hipStreamCreateWithFlags(&st->ws[w], hipStreamNonBlocking) == hipSuccess;
    st->ok = ok;
  }
  return st;
}

struct HostSeg {
  void* dst;
  const void* src;
  size_t bytes;
};

// Copies every segment (device -> host when `to_host`, else host -> device).
// `s` is the stream that wrote (or will read) the device side: the copy is
// ordered after its work so far, and small copies are simply enqueued on it;
// staged copies have finished when the call returns.  Returns the first HIP error.
hipError_t staged_copy(const std::vector<HostSeg>& segs, bool to_host, hipStream_t s) {
  const hipMemcpyKind kind = to_host ? hipMemcpyDeviceToHost : hipMemcpyHostToDevice;
  size_t total = 0;
  for (const HostSeg& g : segs) total += g.bytes;
  if (total < 2 * STAGE_BYTES) {
    for (const HostSeg& g : segs) {
      const hipError_t e = hipMemcpyAsync(g.dst, g.src, g.bytes, kind, s);
      if (e != hipSuccess) return e;
    }
    return hipSuccess;
  }
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  Staging& st = *staging_for(dev);
  std::lock_guard<std::mutex> guard(st.lock);
  e = hipStreamSynchronize(s);  // results complete, or device buffers free to overwrite
  if (e != hipSuccess) return e;
  if (!st.ok) {  // no staging memory: the runtime's own path
    for (const HostSeg& g : segs) {
      e = hipMemcpy(g.dst, g.src, g.bytes, kind);
      if (e != hipSuccess) return e;
    }
    return hipSuccess;
  }
  std::vector<std::pair<size_t, size_t>> chunks;  // (segment, offset)
  for (size_t i = 0; i < segs.size(); ++i)
    for (size_t off = 0; off < segs[i].bytes; off += STAGE_BYTES) chunks.push_back({i, off});
  std::atomic<size_t> next{0};
  std::atomic<int> err{(int)hipSuccess};
  auto worker = [&](int w) {
    if (hipSetDevice(dev) != hipSuccess) {
      err = (int)hipErrorInvalidDevice;
      return;
    }
    for (size_t c = next++; c < chunks.size() && err == (int)hipSuccess; c = next++) {
      const HostSeg& g = segs[chunks[c].first];
      const size_t off = chunks[c].second, len = std::min(STAGE_BYTES, g.bytes - off);
      hipError_t we;
      if (to_host) {
        we = hipMemcpyAsync(st.buf[w], (const char*)g.src + off, len, kind, st.ws[w]);
        if (we == hipSuccess) we = hipStreamSynchronize(st.ws[w]);
        if (we == hipSuccess) std::memcpy((char*)g.dst + off, st.buf[w], len);
      } else {
        std::memcpy(st.buf[w], (const char*)g.src + off, len);
        we = hipMemcpyAsync((char*)g.dst + off, st.buf[w], len, kind, st.ws[w]);
        if (we == hipSuccess) we = hipStreamSynchronize(st.ws[w]);
      }
      if (we != hipSuccess) {
        err = (int)we;
        return;
      }
    }
  };
  const int nw = (int)std::min<size_t>(STAGE_WORKERS, chunks.size());
  std::vector<std::thread> pool;
  try {
    for (int w = 1; w < nw; ++w) pool.emplace_back(worker, w);
  } catch (...) {  // fewer threads: the started ones and this one take every chunk
  }
  worker(0);
  for (std::thread& t : pool) t.join();
  return (hipError_t)err.load();
}

#define API_HIP(expr)                                                             \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess) return fail(KDT_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

}  // namespace

// The argument checks of Tree::query in the reference's order (dimension,
// src/lib.rs:74-84; k; shape; buffers and sizes at the ABI), then the query
// parameters.  The scalar checks (eps, max_distance, p, src/query.rs:49-62)
// come after the finiteness of the queries, which needs the device, so their
// outcome is returned in scalar_rc / smsg for the caller to report in order.
int kdt::query_args(const kdt_tree* tree, const double* q, uint64_t nq, uint64_t ndim, uint64_t k, double p,
                    double max_distance, double eps, uint32_t flags, const void* out_d, const void* out_i,
                    QueryParams& qp, int& scalar_rc, std::string& smsg) {
  if (!tree) return fail(KDT_ERR_ARGUMENT, "null tree");
  const kdt::Tree& t = tree->t;
  if (ndim != t.d)
    return fail(KDT_ERR_DIMENSION_MISMATCH,
                "dimension mismatch: expected " + std::to_string(t.d) + ", got " + std::to_string(ndim));
  if (k == 0) return fail(KDT_ERR_INVALID_K, "k must be greater than zero");
  if (nq == 0) return fail(KDT_ERR_INVALID_SHAPE, "queries must be a contiguous row-major matrix");
  if (!q || !out_d || !out_i) return fail(KDT_ERR_ARGUMENT, "null buffer");
  if (k > 0xFFFFFFFFull || nq > (~0ull) / k) return fail(KDT_ERR_ARGUMENT, "n_queries * k overflows");
  qp = kdt::QueryParams{};
  qp.k = k;
  qp.nq = nq;
  qp.fp32_screen = (flags & KDT_QUERY_FP32_SCREEN) != 0;
  qp.screen_only = (flags & KDT_QUERY_SCREEN_ONLY) != 0;
  qp.matrix_screen = (flags & KDT_QUERY_MATRIX_SCREEN) != 0;
  qp.group = kdt::requested_group(flags);
  qp.allow_order = (flags & KDT_QUERY_NO_ORDER) == 0;
  qp.walk = (flags & KDT_QUERY_NO_TILE_WALK) ? -1 : (flags & KDT_QUERY_TILE_WALK) ? 1 : 0;
  qp.slice_shift = kdt::slice_shift_of(flags);
  scalar_rc = kdt::query_params(p, max_distance, eps, qp, smsg);
  return KDT_OK;
}

int kdt::fail_api(int code, std::string msg) { return fail(code, std::move(msg)); }

namespace {

int query_impl(const kdt_tree* tree, const double* q, bool q_on_device, uint64_t nq, uint64_t ndim,
               uint64_t k, double p, double max_distance, double eps, uint32_t flags, double* out_d,
               int64_t* out_i, bool out_on_device, hipStream_t s) {
  kdt::QueryParams qp;
  int scalar_rc = 0;
  std::string smsg;
  const int arc = kdt::query_args(tree, q, nq, ndim, k, p, max_distance, eps, flags, out_d, out_i, qp, scalar_rc,
                                  smsg);
  if (arc) return arc;
  const kdt::Tree& t = tree->t;
  if (q_on_device && out_on_device && (flags & KDT_QUERY_ASSUME_FINITE)) {
    if (scalar_rc) return fail(scalar_rc, smsg);
    qp.nq = nq;
    std::string msg;
    const int rc = kdt::launch_query(t, qp, q, out_d, out_i, s, msg);
    return rc ? fail(rc, msg) : KDT_OK;
  }

  // Queries per launch: bounded so that device buffers stay modest for huge
  // host batches.  Device-pointer calls run in one launch.
  uint64_t tile = nq;
  if (!q_on_device || !out_on_device) {
    const uint64_t per = k * 16 + ndim * 8;
    tile = std::max<uint64_t>(1, std::min<uint64_t>(nq, (1ull << 30) / per));
  }
  DevMem<int> flag;
  API_HIP(flag.alloc(1, s));
  DevMem<double> dq, dd;
  DevMem<int64_t> di;
  if (!q_on_device) API_HIP(dq.alloc(tile * ndim, s));
  if (!out_on_device) {
    API_HIP(dd.alloc(tile * k, s));
    API_HIP(di.alloc(tile * k, s));
  }
  // Each tile: queries in, all_finite (src/kernel.rs:67-74, checked before
  // eps, max_distance and p, src/query.rs:46-62), and only when the tile is
  // finite the query and the results out -- a call that fails leaves the
  // caller's output buffers untouched by this tile.
  for (uint64_t b = 0; b < nq; b += tile) {
    const uint64_t m = std::min(tile, nq - b);
    const double* qsrc = q + b * ndim;
    const double* dqp = qsrc;
    if (!q_on_device) {
      API_HIP(staged_copy({{dq.p, qsrc, m * ndim * 8}}, false, s));
      dqp = dq.p;
    }
    API_HIP(hipMemsetAsync(flag.p, 0, sizeof(int), s));
    kdt::launch_check_finite(dqp, m * ndim, flag.p, s);
    int h = 0;
    API_HIP(hipMemcpyAsync(&h, flag.p, sizeof(int), hipMemcpyDeviceToHost, s));
    API_HIP(hipStreamSynchronize(s));
    if (h) return fail(KDT_ERR_NON_FINITE, "all coordinates must be finite");
    if (scalar_rc) {
      if (b + m >= nq) return fail(scalar_rc, smsg);
      continue;
    }
    double* od = out_on_device ? out_d + b * k : dd.p;
    int64_t* oi = out_on_device ? out_i + b * k : di.p;
    if (qp.screen_only && !out_on_device) {  // rows the screen skips keep the caller's contents
      API_HIP(hipMemcpyAsync(dd.p, out_d + b * k, m * k * 8, hipMemcpyHostToDevice, s));
      API_HIP(hipMemcpyAsync(di.p, out_i + b * k, m * k * 8, hipMemcpyHostToDevice, s));
    }
    kdt::QueryParams qpt = qp;
    qpt.nq = m;
    std::string msg;
    const int rc = kdt::launch_query(t, qpt, dqp, od, oi, s, msg);
    if (rc) return fail(rc, msg);
    if (!out_on_device) {
      API_HIP(staged_copy({{out_d + b * k, dd.p, m * k * 8}, {out_i + b * k, di.p, m * k * 8}}, true, s));
      API_HIP(hipStreamSynchronize(s));
    }
  }
  return KDT_OK;
}

int build_impl(const double* data, bool on_device, uint64_t n, uint64_t d, uint64_t leafsize,
               hipStream_t s, kdt_tree** out) {
  if (!out) return fail(KDT_ERR_ARGUMENT, "null output");
  *out = nullptr;
  // Validation order of build::build, src/build.rs:25-43.
  if (leafsize == 0) return fail(KDT_ERR_INVALID_LEAFSIZE, "leafsize must be greater than zero");
  if (d == 0 || n == 0) return fail(KDT_ERR_EMPTY_DATA, "data must contain at least one point");
  if (!data) return fail(KDT_ERR_ARGUMENT, "null data");
  if (n > 0xFFFFFFFFull)
    return fail(KDT_ERR_TOO_MANY_POINTS, "n_points must fit in a 32-bit index, got " + std::to_string(n));
  const uint64_t n_nodes = kdt::count_tree_nodes(n, leafsize);
  if (n_nodes > 0xFFFFFFFFull)
    return fail(KDT_ERR_TOO_MANY_NODES, "tree would need " + std::to_string(n_nodes) +
                                            " nodes, which exceeds the 32-bit node index; increase leafsize");
  if (d > (~0ull) / n / 8) return fail(KDT_ERR_ARGUMENT, "n_points * ndim overflows");
  kdt_tree* h = new kdt_tree();
  h->t.n = n;
  h->t.d = d;
  h->t.leafsize = leafsize;
  h->t.n_nodes = n_nodes;
  (void)hipGetDevice(&h->t.device);
  const double* src = data;
  double* tmp = nullptr;
  if (!on_device) {
    hipError_t e = hipMalloc(&tmp, n * d * 8);
    if (e == hipSuccess) e = staged_copy({{tmp, data, n * d * 8}}, false, s);
    if (e != hipSuccess) {
      if (tmp) (void)hipFree(tmp);
      delete h;
      return fail(KDT_ERR_DEVICE, std::string("device upload: ") + hipGetErrorString(e));
    }
    src = tmp;
  }
  std::string msg;
  const int rc = kdt::build_tree(h->t, src, s, msg);
  if (tmp) {
    (void)hipStreamSynchronize(s);
    (void)hipFree(tmp);
  }
  if (rc) {
    kdt_free(h);
    return fail(rc, msg);
  }
  *out = h;
  return KDT_OK;
}

}  // namespace

extern "C" {

int kdt_build(const double* data, uint64_t n_points, uint64_t ndim, uint64_t leafsize, int parallel,
              kdt_tree** out) {
  (void)parallel;
  return build_impl(data, false, n_points, ndim, leafsize, host_stream(), out);
}

int kdt_build_device(const double* d_data, uint64_t n_points, uint64_t ndim, uint64_t leafsize,
                     void* stream, kdt_tree** out) {
  return build_impl(d_data, true, n_points, ndim, leafsize, (hipStream_t)stream, out);
}

int kdt_query(const kdt_tree* tree, const double* queries, uint64_t n_queries, uint64_t ndim, uint64_t k,
              double p, double max_distance, double eps, uint32_t flags, double* out_dist, int64_t* out_idx) {
  return query_impl(tree, queries, false, n_queries, ndim, k, p, max_distance, eps, flags, out_dist, out_idx,
                    false, host_stream());
}

int kdt_query_device(const kdt_tree* tree, const double* d_queries, uint64_t n_queries, uint64_t ndim,
                     uint64_t k, double p, double max_distance, double eps, uint32_t flags,
                     double* d_out_dist, int64_t* d_out_idx, void* stream) {
  return query_impl(tree, d_queries, true, n_queries, ndim, k, p, max_distance, eps, flags, d_out_dist,
                    d_out_idx, true, (hipStream_t)stream);
}

int kdt_original_data(const kdt_tree* tree, double* out) {
  if (!tree || !out) return fail(KDT_ERR_ARGUMENT, "null argument");
  const kdt::Tree& t = tree->t;
  hipStream_t s = host_stream();
  DevMem<double> d;
  API_HIP(d.alloc(t.n * t.d, s));
  kdt::launch_original_order(t, d.p, s);
  API_HIP(staged_copy({{out, d.p, t.n * t.d * 8}}, true, s));
  API_HIP(hipStreamSynchronize(s));
  return KDT_OK;
}

uint64_t kdt_ndim(const kdt_tree* tree) { return tree ? tree->t.d : 0; }
uint64_t kdt_n_points(const kdt_tree* tree) { return tree ? tree->t.n : 0; }
uint64_t kdt_leafsize(const kdt_tree* tree) { return tree ? tree->t.leafsize : 0; }
uint64_t kdt_n_nodes(const kdt_tree* tree) { return tree ? tree->t.n_nodes : 0; }

int kdt_export(const kdt_tree* tree, double* data, uint32_t* indices, kdt_node_record* nodes, double* boxes) {
  if (!tree) return fail(KDT_ERR_ARGUMENT, "null tree");
  const kdt::Tree& t = tree->t;
  hipStream_t s = host_stream();
  if (data) API_HIP(hipMemcpyAsync(data, t.rows, t.n * t.d * 8, hipMemcpyDeviceToHost, s));
  if (indices) API_HIP(hipMemcpyAsync(indices, t.idx, t.n * 4, hipMemcpyDeviceToHost, s));
  if (boxes) API_HIP(hipMemcpyAsync(boxes, t.boxes, t.n_nodes * 2 * t.d * 8, hipMemcpyDeviceToHost, s));
  std::vector<kdt::DevNode> dn;
  if (nodes) {
    dn.resize(t.n_nodes);
    API_HIP(hipMemcpyAsync(dn.data(), t.nodes, t.n_nodes * sizeof(kdt::DevNode), hipMemcpyDeviceToHost, s));
  }
  API_HIP(hipStreamSynchronize(s));
  if (nodes) {
    for (uint64_t id = 0; id < t.n_nodes; ++id) {
      const kdt::DevNode& x = dn[id];
      kdt_node_record r{};
      if (x.b & kdt::LEAF_BIT) {
        uint64_t end;
        std::memcpy(&end, &x.split, 8);
        r.is_leaf = 1;
        r.a = x.a;
        r.b = (uint32_t)end;
      } else {
        r.split_value = x.split;
        r.a = (uint32_t)(id + 1);
        r.b = x.a;
        r.split_dim = x.b & kdt::DIM_MASK;
        r.order_by_box = (x.b & kdt::OBB_BIT) ? 1 : 0;
      }
      nodes[id] = r;
    }
  }
  return KDT_OK;
}

void kdt_free(kdt_tree* tree) {
  if (!tree) return;
  kdt::Tree& t = tree->t;
  if (t.image) (void)hipFree(t.image);  // every array of the tree is a piece of it
  delete tree;
}

const char* kdt_last_error(void) { return g_last_error.c_str(); }

int kdt_timing_enable(int on) {
  if (on) {
    std::lock_guard<std::mutex> g(g_timing_mu);
    if (!g_timing_ref) (void)hipEventCreate(&g_timing_ref);
    if (g_timing_ref) (void)hipEventRecord(g_timing_ref, nullptr);
  }
  return g_timing.exchange(on != 0) ? 1 : 0;
}

int kdt_timing_read(kdt_kernel_time* out, uint32_t capacity, uint32_t* count) {
  if (!count || (!out && capacity)) return fail(KDT_ERR_ARGUMENT, "null argument");
  std::vector<TimedLaunch> recs;
  hipEvent_t ref = nullptr;
  {
    std::lock_guard<std::mutex> g(g_timing_mu);
    recs.swap(g_timed);
    ref = g_timing_ref;
  }
  std::vector<kdt_kernel_time> agg;
  std::vector<std::vector<std::pair<float, float>>> spans;  // per name: [start, end) from `ref`
  hipError_t err = hipSuccess;
  for (TimedLaunch& r : recs) {
    float ms = 0.0f, t0 = 0.0f, t1 = 0.0f;
    hipError_t e = hipEventSynchronize(r.b);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, r.a, r.b);
    bool have_span = false;
    if (e == hipSuccess && ref)
      have_span = hipEventElapsedTime(&t0, ref, r.a) == hipSuccess && hipEventElapsedTime(&t1, ref, r.b) == hipSuccess;
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
    if (e != hipSuccess) {
      err = e;
      continue;
    }
    size_t j = 0;
    while (j < agg.size() && std::strncmp(agg[j].name, r.name.c_str(), sizeof agg[j].name) != 0) ++j;
    if (j == agg.size()) {
      kdt_kernel_time z{};
      std::strncpy(z.name, r.name.c_str(), sizeof z.name - 1);
      agg.push_back(z);
      spans.emplace_back();
    }
    agg[j].launches += 1;
    agg[j].ms += ms;
    agg[j].work += r.work;
    if (have_span) spans[j].push_back({t0, t1});
  }
  // Busy time per kernel: the length of the union of its launches' intervals
  // (launches on several streams overlap; each one's own duration then
  // includes the time it shared the device).
  for (size_t j = 0; j < agg.size(); ++j) {
    std::vector<std::pair<float, float>>& v = spans[j];
    std::sort(v.begin(), v.end());
    double busy = 0.0, cs = 0.0, ce = -1.0;
    bool open = false;
    for (const auto& iv : v) {
      if (!open || iv.first > ce) {
        if (open) busy += ce - cs;
        cs = iv.first;
        ce = iv.second;
        open = true;
      } else if (iv.second > ce) {
        ce = iv.second;
      }
    }
    if (open) busy += ce - cs;
    agg[j].busy_ms = v.size() == agg[j].launches ? busy : -1.0;
  }
  // "*": every timed kernel together (the union of all intervals): the time
  // the device ran at least one of them.
  if (!agg.empty()) {
    std::vector<std::pair<float, float>> all;
    bool complete = true;
    for (size_t j = 0; j < agg.size(); ++j) {
      complete = complete && spans[j].size() == agg[j].launches;
      all.insert(all.end(), spans[j].begin(), spans[j].end());
    }
    std::sort(all.begin(), all.end());
    double busy = 0.0, cs = 0.0, ce = -1.0;
    bool open = false;
    kdt_kernel_time z{};
    std::strncpy(z.name, "*", sizeof z.name - 1);
    for (const auto& iv : all) {
      if (!open || iv.first > ce) {
        if (open) busy += ce - cs;
        cs = iv.first;
        ce = iv.second;
        open = true;
      } else if (iv.second > ce) {
        ce = iv.second;
      }
    }
    if (open) busy += ce - cs;
    for (const kdt_kernel_time& a : agg) {
      z.launches += a.launches;
      z.ms += a.ms;
      z.work += a.work;
    }
    z.busy_ms = complete ? busy : -1.0;
    agg.push_back(z);
  }
  for (size_t j = 0; j < agg.size() && j < capacity; ++j) out[j] = agg[j];
  *count = (uint32_t)agg.size();
  if (err != hipSuccess) return fail(KDT_ERR_DEVICE, std::string("kernel timing: ") + hipGetErrorString(err));
  return KDT_OK;
}

}  // extern "C"
