// sanitize_main.cpp -- drives the oracle (kdtree_oracle.cpp) under
// AddressSanitizer and UndefinedBehaviorSanitizer on the host.
//
// TEST INFRASTRUCTURE ONLY (built by `make -C oracle sanitize`, run by
// tests/test_oracle_sanitized.py).  Builds trees over uniform, duplicate-heavy,
// lattice, single-point and wide inputs (n up to 300, widths 1-33) at several
// leafsizes, queries them
// with every metric, eps, max_distance and k from 1 to beyond n, single- and
// multi-threaded, and walks every error path; checks that each result row is
// sorted by (distance, index) with indices in range or the padding n.  Exits
// non-zero on a failed check; the sanitizers abort on any memory or UB error.

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <limits>
#include <random>
#include <vector>

extern "C" {
int kdo_build(const double* data, uint64_t n, uint64_t d, uint64_t leafsize, void** out);
void kdo_free(void* t);
uint64_t kdo_n_nodes(const void* t);
int kdo_query(const void* tp, const double* q, uint64_t nq, uint64_t d, uint64_t k, double p, double max_distance,
              double eps, int threads, double* out_d, int64_t* out_i, uint64_t* counters);
}

namespace {

int failures = 0;

void check(bool ok, const char* what, uint64_t n, uint64_t d, uint64_t k) {
  if (!ok) {
    std::fprintf(stderr, "FAILED: %s (n=%llu d=%llu k=%llu)\n", what, (unsigned long long)n, (unsigned long long)d,
                 (unsigned long long)k);
    ++failures;
  }
}

std::vector<double> make(const char* kind, uint64_t n, uint64_t d, std::mt19937_64& rng) {
  std::vector<double> x(n * d);
  std::uniform_real_distribution<double> u(0.0, 1.0);
  for (uint64_t i = 0; i < n * d; ++i) {
    if (kind[0] == 'u') x[i] = u(rng);
    else if (kind[0] == 'l') x[i] = std::floor(u(rng) * 5.0);  // lattice: ties everywhere
    else x[i] = (double)((i * 2654435761ull) % 10007ull);      // hashed: repeated coordinates
  }
  return x;
}

void run_case(const char* kind, uint64_t n, uint64_t d, uint64_t leafsize, std::mt19937_64& rng) {
  std::vector<double> data = make(kind, n, d, rng);
  void* t = nullptr;
  if (kdo_build(data.data(), n, d, leafsize, &t) != 0 || !t) {
    check(false, "build", n, d, 0);
    return;
  }
  const uint64_t nq = 37;
  std::vector<double> q = make("u", nq, d, rng);
  const double inf = std::numeric_limits<double>::infinity();
  const double ps[] = {1.0, 2.0, 3.0, inf};
  const uint64_t ks[] = {1, 5, 33, n + 3};
  for (double p : ps)
    for (uint64_t k : ks)
      for (double eps : {0.0, 0.5})
        for (double md : {inf, 0.4})
          for (int threads : {1, 3}) {
            std::vector<double> od(nq * k);
            std::vector<int64_t> oi(nq * k);
            uint64_t ctr[3] = {0, 0, 0};
            const int rc = kdo_query(t, q.data(), nq, d, k, p, md, eps, threads, od.data(), oi.data(), ctr);
            check(rc == 0, "query rc", n, d, k);
            if (rc) continue;
            for (uint64_t r = 0; r < nq; ++r)
              for (uint64_t j = 0; j < k; ++j) {
                const double dj = od[r * k + j];
                const int64_t ij = oi[r * k + j];
                check(ij >= 0 && (uint64_t)ij <= n, "index range", n, d, k);
                check((uint64_t)ij == n ? std::isinf(dj) : dj <= md * (1.0 + 1e-12), "padding / cutoff", n, d, k);
                if (j) {
                  // Rows are ordered by (reduced distance, index); the restored
                  // distance (sqrt, pow) may map two reduced values to one
                  // double, so the index order is checked on equal distances
                  // only where restore is the identity (p = 1, inf).
                  const double dp = od[r * k + j - 1];
                  const int64_t ip = oi[r * k + j - 1];
                  check(dp <= dj, "row order", n, d, k);
                  if (dp == dj && (p == 1.0 || std::isinf(p)))
                    check(ip < ij || ij == (int64_t)n, "tie order", n, d, k);
                }
              }
          }
  kdo_free(t);
}

void error_paths() {
  std::vector<double> x = {0.0, 1.0, 2.0, 3.0};
  void* t = nullptr;
  check(kdo_build(x.data(), 2, 2, 0, &t) != 0, "leafsize 0", 2, 2, 0);
  check(kdo_build(x.data(), 0, 2, 4, &t) != 0, "empty", 0, 2, 0);
  x[1] = std::numeric_limits<double>::quiet_NaN();
  check(kdo_build(x.data(), 2, 2, 4, &t) != 0, "NaN data", 2, 2, 0);
  x[1] = 1.0;
  check(kdo_build(x.data(), 2, 2, 4, &t) == 0 && t, "valid build", 2, 2, 0);
  double qn[2] = {std::numeric_limits<double>::infinity(), 0.0};
  double od[2];
  int64_t oi[2];
  uint64_t ctr[3];
  check(kdo_query(t, qn, 1, 2, 2, 2.0, 1e300, 0.0, 1, od, oi, ctr) != 0, "non-finite query", 2, 2, 2);
  double q[2] = {0.5, 0.5};
  check(kdo_query(t, q, 1, 2, 0, 2.0, 1e300, 0.0, 1, od, oi, ctr) != 0, "k 0", 2, 2, 0);
  check(kdo_query(t, q, 1, 2, 2, 0.5, 1e300, 0.0, 1, od, oi, ctr) != 0, "p < 1", 2, 2, 2);
  check(kdo_query(t, q, 1, 2, 2, 2.0, -1.0, 0.0, 1, od, oi, ctr) != 0, "max_distance < 0", 2, 2, 2);
  check(kdo_query(t, q, 1, 2, 2, 2.0, 1e300, -1.0, 1, od, oi, ctr) != 0, "eps < 0", 2, 2, 2);
  kdo_free(t);
}

}  // namespace

int main() {
  std::mt19937_64 rng(2026);
  const char* kinds[] = {"uniform", "lattice", "hashed"};
  for (const char* kind : kinds)
    for (uint64_t d : {1ull, 3ull, 4ull, 8ull, 9ull, 16ull, 33ull})
      for (uint64_t leafsize : {1ull, 3ull, 32ull})
        for (uint64_t n : {1ull, 7ull, 300ull}) run_case(kind, n, d, leafsize, rng);
  error_paths();
  if (failures) {
    std::fprintf(stderr, "%d checks failed\n", failures);
    return 1;
  }
  std::printf("sanitized oracle run: all checks passed\n");
  return 0;
}
