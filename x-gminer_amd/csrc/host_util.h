// Host-side helpers shared by the libxgm translation units: a small
// deterministic thread pool loop, counter-based random streams and the
// library-owned graph container behind the opaque xgm_hgraph handle.
#pragma once

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstdlib>
#include <thread>
#include <vector>

#include "../../include/xgm_graph.h"

struct xgm_hgraph {
  int64_t n = 0;
  int64_t nnz = 0;
  std::vector<int64_t> rowptr;
  std::vector<int32_t> colidx;
  uint32_t flags = 0;
};

namespace xgm {

// Number of host worker threads: XGM_NUM_THREADS if set, else all cores.
inline int host_threads() {
  if (const char *s = std::getenv("XGM_NUM_THREADS")) {
    int v = std::atoi(s);
    if (v > 0) return v;
  }
  unsigned hc = std::thread::hardware_concurrency();
  return hc ? (int)std::min(hc, 64u) : 1;
}

// Run fn(begin, end) over [0, n) in dynamically claimed chunks.  Results
// must not depend on which thread runs which chunk (callers only write
// disjoint outputs or use commutative integer atomics).
template <class F>
void parallel_for(int64_t n, int64_t chunk, F &&fn) {
  if (n <= 0) return;
  int nt = host_threads();
  if (nt <= 1 || n <= chunk) {
    fn((int64_t)0, n);
    return;
  }
  std::atomic<int64_t> next{0};
  auto worker = [&]() {
    for (;;) {
      int64_t b = next.fetch_add(chunk, std::memory_order_relaxed);
      if (b >= n) break;
      fn(b, std::min(n, b + chunk));
    }
  };
  int64_t nchunks = (n + chunk - 1) / chunk;
  nt = (int)std::min<int64_t>(nt, nchunks);
  std::vector<std::thread> pool;
  pool.reserve(nt - 1);
  for (int t = 1; t < nt; ++t) pool.emplace_back(worker);
  worker();
  for (auto &th : pool) th.join();
}

// splitmix64 finaliser; stream element k of key s is mix(s + (k+1)*gamma).
inline uint64_t mix64(uint64_t x) {
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
  return x ^ (x >> 31);
}
constexpr uint64_t kGamma = 0x9E3779B97F4A7C15ULL;
inline uint64_t stream_at(uint64_t key, uint64_t k) { return mix64(key + (k + 1) * kGamma); }

// Uniform integer in [0, range) from a 64-bit draw (multiply-shift).
inline uint64_t bounded_draw(uint64_t r, uint64_t range) {
  return (uint64_t)(((unsigned __int128)r * range) >> 64);
}

// Exclusive prefix sum of counts[0..n) into out[0..n]; returns total.
int64_t exclusive_scan(const std::vector<int64_t> &counts, std::vector<int64_t> &out);

// xgm_csr_validate, optionally without the O(m log d) mirror check of a
// symmetric CSR (the upload path runs that check on the device instead).
int validate_csr(const xgm_csr *g, bool check_mirrors);

}  // namespace xgm
