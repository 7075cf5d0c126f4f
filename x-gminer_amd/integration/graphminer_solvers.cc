// GraphMiner binding of libxgm: defines the three solver entry points that
// GraphMiner's drivers resolve at link time (codes/GraphMiner/src/triangle/main.cc:5,
// src/clique/main.cc:6, src/sgl/main.cc:7), on top of the C ABI in include/xgm.h.
// Link it instead of omp_base.o / gpu_base.o.  Errors print and exit, as the
// reference's CUDA path does.
#include <cstdlib>
#include <iostream>

#include "graph.h"
#include "pattern.hh"
#include "xgm.h"

namespace {

xgm_ctx *ctx() {
  static xgm_ctx *c = nullptr;
  if (!c) {
    int rc = xgm_ctx_create(0, &c);
    if (rc != XGM_OK) {
      std::cerr << "xgm: " << xgm_strerror(rc) << "\n";
      std::exit(EXIT_FAILURE);
    }
  }
  return c;
}

uint64_t count(Graph &g, uint32_t flags, int pattern) {
  xgm_csr v{(int64_t)g.V(), (int64_t)g.E(), g.out_rowptr(), g.out_colidx(), flags};
  uint64_t c = 0;
  int rc = xgm_count_csr(ctx(), &v, pattern, nullptr, &c);
  if (rc != XGM_OK) {
    std::cerr << "xgm: " << xgm_strerror(rc) << ": " << xgm_ctx_last_error(ctx()) << "\n";
    std::exit(EXIT_FAILURE);
  }
  xgm_profile p;
  xgm_last_profile(ctx(), &p);
  std::cout << "runtime [xgm gfx950] = " << p.total_ms * 1e-3 << " sec\n";
  return c;
}

}  // namespace

void TCSolver(Graph &g, uint64_t &total, int, int) { total = count(g, XGM_CSR_SORTED | XGM_CSR_DAG, XGM_TRIANGLE); }

void CliqueSolver(Graph &g, int k, uint64_t &total, int, int) {
  int p = k == 3 ? XGM_TRIANGLE : k == 4 ? XGM_CLIQUE4 : k == 5 ? XGM_CLIQUE5 : -1;
  if (p < 0) {
    std::cout << "Not implemented yet\n";
    std::exit(0);
  }
  total = count(g, XGM_CSR_SORTED | XGM_CSR_DAG, p);
}

void SglSolver(Graph &g, Pattern &pat, uint64_t &total, int, int) {
  int p = pat.is_diamond() ? XGM_DIAMOND_EDGE : pat.is_rectangle() ? XGM_CYCLE4_EDGE : -1;
  if (p < 0) {
    std::cout << "Not implemented\n";
    total = 0;
    return;
  }
  total = count(g, XGM_CSR_SORTED | XGM_CSR_SYMMETRIC, p);
}
