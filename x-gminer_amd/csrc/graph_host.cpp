// Host graph substrate: generators, CSR cleaning, degree orientation,
// on-disk format and content hashing.  Declared in include/xgm_graph.h.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <new>
#include <string>

#include "host_util.h"

namespace xgm {

int64_t exclusive_scan(const std::vector<int64_t> &counts, std::vector<int64_t> &out) {
  const int64_t n = (int64_t)counts.size();
  out.assign(n + 1, 0);
  // Two-pass blocked scan: per-block totals, then per-block local scans.
  const int64_t block = 1 << 16;
  const int64_t nb = (n + block - 1) / block;
  std::vector<int64_t> bsum(nb + 1, 0);
  parallel_for(nb, 1, [&](int64_t b0, int64_t b1) {
    for (int64_t b = b0; b < b1; ++b) {
      int64_t s = 0;
      for (int64_t i = b * block, e = std::min(n, (b + 1) * block); i < e; ++i) s += counts[i];
      bsum[b + 1] = s;
    }
  });
  for (int64_t b = 0; b < nb; ++b) bsum[b + 1] += bsum[b];
  parallel_for(nb, 1, [&](int64_t b0, int64_t b1) {
    for (int64_t b = b0; b < b1; ++b) {
      int64_t s = bsum[b];
      for (int64_t i = b * block, e = std::min(n, (b + 1) * block); i < e; ++i) {
        out[i] = s;
        s += counts[i];
      }
    }
  });
  out[n] = bsum[nb];
  return out[n];
}

namespace {

// Symmetrize + dedup + drop self-loops.  Fill order inside a list depends
// on thread timing, but every list is sorted and uniqued afterwards, so the
// result is a pure function of the input multiset of pairs.
int build_clean(int64_t n, int64_t m, const int32_t *src, const int32_t *dst, xgm_hgraph **out) {
  if (n < 0 || m < 0 || n > INT32_MAX || (m > 0 && (!src || !dst)) || !out) return XGM_EINVAL;
  std::unique_ptr<xgm_hgraph> g(new (std::nothrow) xgm_hgraph);
  if (!g) return XGM_ENOMEM;
  try {
    std::atomic<bool> bad{false};
    std::vector<int64_t> deg(n, 0);
    parallel_for(m, 1 << 16, [&](int64_t b, int64_t e) {
      for (int64_t i = b; i < e; ++i) {
        int32_t u = src[i], v = dst[i];
        if (u < 0 || v < 0 || u >= n || v >= n) { bad = true; continue; }
        if (u == v) continue;
        __atomic_fetch_add(&deg[u], 1, __ATOMIC_RELAXED);
        __atomic_fetch_add(&deg[v], 1, __ATOMIC_RELAXED);
      }
    });
    if (bad) return XGM_EINVAL;
    std::vector<int64_t> off;
    int64_t total = exclusive_scan(deg, off);
    std::vector<int32_t> tmp(total);
    std::vector<int64_t> pos(off.begin(), off.end() - 1);
    parallel_for(m, 1 << 16, [&](int64_t b, int64_t e) {
      for (int64_t i = b; i < e; ++i) {
        int32_t u = src[i], v = dst[i];
        if (u == v) continue;
        tmp[__atomic_fetch_add(&pos[u], 1, __ATOMIC_RELAXED)] = v;
        tmp[__atomic_fetch_add(&pos[v], 1, __ATOMIC_RELAXED)] = u;
      }
    });
    // sort + unique each list in place; deg[u] becomes the cleaned degree
    parallel_for(n, 1024, [&](int64_t b, int64_t e) {
      for (int64_t u = b; u < e; ++u) {
        int32_t *first = tmp.data() + off[u], *last = tmp.data() + off[u + 1];
        std::sort(first, last);
        deg[u] = std::unique(first, last) - first;
      }
    });
    g->n = n;
    g->nnz = exclusive_scan(deg, g->rowptr);
    g->colidx.resize(g->nnz);
    parallel_for(n, 4096, [&](int64_t b, int64_t e) {
      for (int64_t u = b; u < e; ++u)
        std::memcpy(g->colidx.data() + g->rowptr[u], tmp.data() + off[u], sizeof(int32_t) * deg[u]);
    });
    g->flags = XGM_CSR_SORTED | XGM_CSR_SYMMETRIC;
  } catch (const std::bad_alloc &) {
    return XGM_ENOMEM;
  }
  *out = g.release();
  return XGM_OK;
}

}  // namespace
}  // namespace xgm

using namespace xgm;

extern "C" {

const char *xgm_strerror(int status) {
  switch (status) {
    case XGM_OK: return "ok";
    case XGM_EINVAL: return "invalid argument";
    case XGM_ENOMEM: return "out of memory";
    case XGM_EUNSUPPORTED: return "unsupported";
    case XGM_EHIP: return "HIP runtime error";
    case XGM_ENODEV: return "no usable gfx950 device";
    case XGM_EIO: return "I/O error";
    default: return "unknown status";
  }
}

int xgm_hgraph_view(const xgm_hgraph *g, xgm_csr *view) {
  if (!g || !view) return XGM_EINVAL;
  view->n = g->n;
  view->nnz = g->nnz;
  view->rowptr = g->rowptr.data();
  view->colidx = g->colidx.data();
  view->flags = g->flags;
  return XGM_OK;
}

void xgm_hgraph_free(xgm_hgraph *g) { delete g; }

int xgm_build_undirected(int64_t n, int64_t m, const int32_t *src, const int32_t *dst,
                         xgm_hgraph **out) {
  return build_clean(n, m, src, dst, out);
}

int xgm_gen_rmat(int scale, int edge_factor, double a, double b, double c, uint64_t seed,
                 xgm_hgraph **out) {
  if (scale < 1 || scale > 30 || edge_factor < 1 || !out) return XGM_EINVAL;
  if (a < 0 || b < 0 || c < 0 || a + b + c > 1.0) return XGM_EINVAL;
  const int64_t n = int64_t(1) << scale;
  const int64_t m = (int64_t)edge_factor * n;
  // 32-bit cumulative thresholds for quadrants a | b | c | d
  const double two32 = 4294967296.0;
  const uint64_t ta = (uint64_t)std::floor(a * two32);
  const uint64_t tab = (uint64_t)std::floor((a + b) * two32);
  const uint64_t tabc = (uint64_t)std::floor((a + b + c) * two32);
  const uint64_t key = mix64(seed ^ 0x524D4154ULL /* "RMAT" */);
  std::vector<int32_t> src, dst;
  try {
    src.resize(m);
    dst.resize(m);
  } catch (const std::bad_alloc &) {
    return XGM_ENOMEM;
  }
  const uint64_t draws = (uint64_t)(scale + 1) / 2;  // two 32-bit levels per draw
  parallel_for(m, 1 << 15, [&](int64_t b0, int64_t b1) {
    for (int64_t i = b0; i < b1; ++i) {
      uint32_t u = 0, v = 0;
      int level = 0;
      for (uint64_t k = 0; k < draws; ++k) {
        uint64_t r = stream_at(key, (uint64_t)i * draws + k);
        for (int h = 0; h < 2 && level < scale; ++h, ++level) {
          uint64_t x = h ? (r >> 32) : (r & 0xffffffffULL);
          uint32_t ub = x >= tab ? 1u : 0u;                       // c or d
          uint32_t vb = (x >= ta && x < tab) || x >= tabc ? 1u : 0u;  // b or d
          u |= ub << (scale - 1 - level);
          v |= vb << (scale - 1 - level);
        }
      }
      src[i] = (int32_t)u;
      dst[i] = (int32_t)v;
    }
  });
  return build_clean(n, m, src.data(), dst.data(), out);
}

int xgm_gen_chung_lu(int64_t n, double avg_degree, double beta, double w_max, uint64_t seed,
                     xgm_hgraph **out) {
  if (n < 2 || n > INT32_MAX || !(avg_degree > 0) || !(beta > 1.0) || !(w_max >= avg_degree) || !out)
    return XGM_EINVAL;
  const double gam = 1.0 / (beta - 1.0);
  // weights w_i = min(w_max, A*(i+1)^-gam); bisection on A for mean = avg_degree
  std::vector<double> base(n);
  parallel_for(n, 1 << 16, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i) base[i] = std::pow((double)(i + 1), -gam);
  });
  auto mean_for = [&](double A) {
    // ordered sum over blocks keeps the result independent of thread count
    const int64_t block = 1 << 16;
    const int64_t nb = (n + block - 1) / block;
    std::vector<double> part(nb, 0.0);
    parallel_for(nb, 1, [&](int64_t b0, int64_t b1) {
      for (int64_t bb = b0; bb < b1; ++bb) {
        double s = 0;
        for (int64_t i = bb * block, e = std::min(n, (bb + 1) * block); i < e; ++i)
          s += std::min(w_max, A * base[i]);
        part[bb] = s;
      }
    });
    double s = 0;
    for (double p : part) s += p;
    return s / (double)n;
  };
  double lo = 0.0, hi = w_max * std::pow((double)n, gam) + 1.0;
  for (int it = 0; it < 200; ++it) {
    double mid = 0.5 * (lo + hi);
    if (mean_for(mid) < avg_degree) lo = mid; else hi = mid;
  }
  const double A = 0.5 * (lo + hi);
  // integer weights (units of 1/1024) and their cumulative sums
  std::vector<int64_t> wi(n);
  parallel_for(n, 1 << 16, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i) wi[i] = std::max<int64_t>(1, std::llround(std::min(w_max, A * base[i]) * 1024.0));
  });
  std::vector<int64_t> cum;
  const int64_t total = exclusive_scan(wi, cum);
  const int64_t m = std::llround((double)n * avg_degree / 2.0);
  const uint64_t key = mix64(seed ^ 0x43484C55ULL /* "CHLU" */);
  std::vector<int32_t> src(m), dst(m);
  parallel_for(m, 1 << 15, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i) {
      for (int side = 0; side < 2; ++side) {
        uint64_t r = bounded_draw(stream_at(key, 2 * (uint64_t)i + side), (uint64_t)total);
        // largest j with cum[j] <= r
        int64_t j = std::upper_bound(cum.begin(), cum.begin() + n, (int64_t)r) - cum.begin() - 1;
        (side ? dst : src)[i] = (int32_t)j;
      }
    }
  });
  return build_clean(n, m, src.data(), dst.data(), out);
}

int xgm_gen_erdos_renyi(int64_t n, double avg_degree, uint64_t seed, xgm_hgraph **out) {
  if (n < 2 || n > INT32_MAX || !(avg_degree >= 0) || !out) return XGM_EINVAL;
  const int64_t m = std::llround((double)n * avg_degree / 2.0);
  const uint64_t key = mix64(seed ^ 0x45524E44ULL /* "ERND" */);
  std::vector<int32_t> src(m), dst(m);
  parallel_for(m, 1 << 15, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i) {
      src[i] = (int32_t)bounded_draw(stream_at(key, 2 * (uint64_t)i), (uint64_t)n);
      dst[i] = (int32_t)bounded_draw(stream_at(key, 2 * (uint64_t)i + 1), (uint64_t)n);
    }
  });
  return build_clean(n, m, src.data(), dst.data(), out);
}

}  // extern "C"

namespace xgm {

int validate_csr(const xgm_csr *g, bool check_mirrors) {
  if (!g || g->n < 0 || g->nnz < 0 || g->n > INT32_MAX) return XGM_EINVAL;
  if (!g->rowptr || (g->nnz > 0 && !g->colidx)) return XGM_EINVAL;
  if (g->rowptr[0] != 0 || g->rowptr[g->n] != g->nnz) return XGM_EINVAL;
  const int64_t n = g->n;
  std::atomic<bool> bad{false};
  parallel_for(n, 4096, [&](int64_t b, int64_t e) {
    for (int64_t u = b; u < e && !bad; ++u) {
      int64_t s = g->rowptr[u], t = g->rowptr[u + 1];
      if (t < s) { bad = true; break; }
      for (int64_t k = s; k < t; ++k) {
        int32_t v = g->colidx[k];
        if (v < 0 || v >= n) { bad = true; break; }
        if ((g->flags & XGM_CSR_SORTED) && k > s && g->colidx[k - 1] >= v) { bad = true; break; }
        if ((g->flags & (XGM_CSR_SYMMETRIC | XGM_CSR_DAG)) && v == u) { bad = true; break; }
      }
    }
  });
  if (bad) return XGM_EINVAL;
  if ((g->flags & XGM_CSR_SYMMETRIC) && !(g->flags & XGM_CSR_SORTED)) return XGM_EINVAL;
  if ((g->flags & XGM_CSR_SYMMETRIC) && check_mirrors) {
    // every (u,v) has its mirror (v,u): binary search in the sorted list of v
    parallel_for(n, 4096, [&](int64_t b, int64_t e) {
      for (int64_t u = b; u < e && !bad; ++u)
        for (int64_t k = g->rowptr[u]; k < g->rowptr[u + 1]; ++k) {
          int32_t v = g->colidx[k];
          const int32_t *f = g->colidx + g->rowptr[v], *l = g->colidx + g->rowptr[v + 1];
          if (!std::binary_search(f, l, (int32_t)u)) { bad = true; break; }
        }
    });
  }
  return bad ? XGM_EINVAL : XGM_OK;
}

}  // namespace xgm

extern "C" {

int xgm_csr_validate(const xgm_csr *g) { return xgm::validate_csr(g, true); }

int xgm_orient_dag(const xgm_csr *g, xgm_hgraph **out) {
  if (!g || !out) return XGM_EINVAL;
  if ((g->flags & (XGM_CSR_SORTED | XGM_CSR_SYMMETRIC)) != (XGM_CSR_SORTED | XGM_CSR_SYMMETRIC))
    return XGM_EINVAL;
  const int64_t n = g->n;
  std::unique_ptr<xgm_hgraph> d(new (std::nothrow) xgm_hgraph);
  if (!d) return XGM_ENOMEM;
  try {
    auto keep = [&](int64_t u, int32_t v) {
      int64_t du = g->rowptr[u + 1] - g->rowptr[u];
      int64_t dv = g->rowptr[v + 1] - g->rowptr[v];
      return dv > du || (dv == du && v > u);
    };
    std::vector<int64_t> cnt(n, 0);
    parallel_for(n, 4096, [&](int64_t b, int64_t e) {
      for (int64_t u = b; u < e; ++u) {
        int64_t c = 0;
        for (int64_t k = g->rowptr[u]; k < g->rowptr[u + 1]; ++k) c += keep(u, g->colidx[k]);
        cnt[u] = c;
      }
    });
    d->n = n;
    d->nnz = exclusive_scan(cnt, d->rowptr);
    d->colidx.resize(d->nnz);
    parallel_for(n, 4096, [&](int64_t b, int64_t e) {
      for (int64_t u = b; u < e; ++u) {
        int64_t o = d->rowptr[u];
        for (int64_t k = g->rowptr[u]; k < g->rowptr[u + 1]; ++k)
          if (keep(u, g->colidx[k])) d->colidx[o++] = g->colidx[k];
      }
    });
    d->flags = XGM_CSR_SORTED | XGM_CSR_DAG;
  } catch (const std::bad_alloc &) {
    return XGM_ENOMEM;
  }
  *out = d.release();
  return XGM_OK;
}

int xgm_save_csr(const char *prefix, const xgm_csr *g) {
  if (!prefix || !g || !g->rowptr) return XGM_EINVAL;
  std::string p(prefix);
  int64_t maxd = 0;
  for (int64_t u = 0; u < g->n; ++u) maxd = std::max(maxd, g->rowptr[u + 1] - g->rowptr[u]);
  {
    std::ofstream m(p + ".meta.txt");
    if (!m) return XGM_EIO;
    // n, nnz, sizeof(vid) sizeof(eid) sizeof(vlabel) sizeof(elabel), max degree,
    // feature length, #vertex classes, #edge classes
    m << g->n << "\n" << g->nnz << "\n4 8 1 2\n" << maxd << "\n0\n0\n0\n";
    if (!m) return XGM_EIO;
  }
  FILE *fv = std::fopen((p + ".vertex.bin").c_str(), "wb");
  if (!fv) return XGM_EIO;
  size_t wv = std::fwrite(g->rowptr, sizeof(int64_t), (size_t)g->n + 1, fv);
  std::fclose(fv);
  FILE *fe = std::fopen((p + ".edge.bin").c_str(), "wb");
  if (!fe) return XGM_EIO;
  size_t we = g->nnz ? std::fwrite(g->colidx, sizeof(int32_t), (size_t)g->nnz, fe) : 0;
  std::fclose(fe);
  return (wv == (size_t)g->n + 1 && we == (size_t)g->nnz) ? XGM_OK : XGM_EIO;
}

int xgm_load_csr(const char *prefix, uint32_t flags, xgm_hgraph **out) {
  if (!prefix || !out) return XGM_EINVAL;
  std::string p(prefix);
  std::ifstream m(p + ".meta.txt");
  if (!m) return XGM_EIO;
  int64_t n = -1, nnz = -1;
  int vid_size = 0, eid_size = 0;
  m >> n >> nnz >> vid_size >> eid_size;
  if (!m || n < 0 || nnz < 0 || n > INT32_MAX) return XGM_EIO;
  if (vid_size != 4 || eid_size != 8) return XGM_EUNSUPPORTED;
  std::unique_ptr<xgm_hgraph> g(new (std::nothrow) xgm_hgraph);
  if (!g) return XGM_ENOMEM;
  try {
    g->rowptr.resize(n + 1);
    g->colidx.resize(nnz);
  } catch (const std::bad_alloc &) {
    return XGM_ENOMEM;
  }
  FILE *fv = std::fopen((p + ".vertex.bin").c_str(), "rb");
  if (!fv) return XGM_EIO;
  size_t rv = std::fread(g->rowptr.data(), sizeof(int64_t), (size_t)n + 1, fv);
  std::fclose(fv);
  FILE *fe = std::fopen((p + ".edge.bin").c_str(), "rb");
  if (!fe) return XGM_EIO;
  size_t re = nnz ? std::fread(g->colidx.data(), sizeof(int32_t), (size_t)nnz, fe) : 0;
  std::fclose(fe);
  if (rv != (size_t)n + 1 || re != (size_t)nnz) return XGM_EIO;
  g->n = n;
  g->nnz = nnz;
  g->flags = flags;
  xgm_csr v;
  xgm_hgraph_view(g.get(), &v);
  if (xgm_csr_validate(&v) != XGM_OK) return XGM_EINVAL;
  *out = g.release();
  return XGM_OK;
}

int xgm_load_edgelist(const char *path, xgm_hgraph **out) {
  if (!path || !out) return XGM_EINVAL;
  *out = nullptr;
  FILE *f = std::fopen(path, "r");
  if (!f) return XGM_EIO;
  std::unique_ptr<FILE, int (*)(FILE *)> hold(f, std::fclose);
  long long n = -1, m = -1;
  if (std::fscanf(f, "%lld %lld", &n, &m) != 2 || n < 0 || m < 0 || n > INT32_MAX) return XGM_EIO;
  std::vector<int32_t> src, dst;
  try {
    src.reserve((size_t)m);
    dst.reserve((size_t)m);
    long long lines = 0, x, y;
    std::vector<long long> rx, ry;
    rx.reserve((size_t)m);
    ry.reserve((size_t)m);
    while (std::fscanf(f, "%lld %lld", &x, &y) == 2) {
      ++lines;
      if (x == y) continue;  // self-loops are dropped
      rx.push_back(x);
      ry.push_back(y);
    }
    if (!std::feof(f) || lines != m) return XGM_EIO;
    // ids in order of first appearance (x before y on each line)
    std::vector<long long> seq;
    seq.reserve(2 * rx.size());
    for (size_t i = 0; i < rx.size(); ++i) {
      seq.push_back(rx[i]);
      seq.push_back(ry[i]);
    }
    std::vector<size_t> idx(seq.size());
    for (size_t i = 0; i < idx.size(); ++i) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) { return seq[a] < seq[b]; });
    std::vector<size_t> first_pos;  // per distinct raw id, its first position
    std::vector<int32_t> group(seq.size());
    for (size_t k = 0; k < idx.size(); ++k) {
      if (k == 0 || seq[idx[k]] != seq[idx[k - 1]]) first_pos.push_back(idx[k]);
      group[idx[k]] = (int32_t)(first_pos.size() - 1);
    }
    if ((long long)first_pos.size() != n) return XGM_EIO;
    std::vector<size_t> by_first(first_pos.size());
    for (size_t i = 0; i < by_first.size(); ++i) by_first[i] = i;
    std::sort(by_first.begin(), by_first.end(), [&](size_t a, size_t b) { return first_pos[a] < first_pos[b]; });
    std::vector<int32_t> new_id(first_pos.size());
    for (size_t r = 0; r < by_first.size(); ++r) new_id[by_first[r]] = (int32_t)r;
    for (size_t i = 0; i < rx.size(); ++i) {
      src.push_back(new_id[group[2 * i]]);
      dst.push_back(new_id[group[2 * i + 1]]);
    }
  } catch (const std::bad_alloc &) {
    return XGM_ENOMEM;
  }
  return build_clean(n, (int64_t)src.size(), src.data(), dst.data(), out);
}

uint64_t xgm_csr_hash(const xgm_csr *g) {
  if (!g || !g->rowptr) return 0;
  // FNV-1a over 64-bit words inside fixed-size chunks, chunk digests folded in order
  const int64_t chunk = 1 << 18;
  auto hash_words = [&](const unsigned char *base, int64_t nbytes) {
    const int64_t nch = (nbytes + chunk - 1) / chunk;
    std::vector<uint64_t> dig(nch);
    parallel_for(nch, 1, [&](int64_t c0, int64_t c1) {
      for (int64_t c = c0; c < c1; ++c) {
        uint64_t h = 0xcbf29ce484222325ULL ^ (uint64_t)c;
        const int64_t b = c * chunk, e = std::min(nbytes, b + chunk);
        int64_t i = b;
        for (; i + 8 <= e; i += 8) {
          uint64_t w;
          std::memcpy(&w, base + i, 8);
          h = (h ^ w) * 0x100000001b3ULL;
        }
        for (; i < e; ++i) h = (h ^ base[i]) * 0x100000001b3ULL;
        dig[c] = mix64(h);
      }
    });
    uint64_t h = mix64((uint64_t)nbytes);
    for (uint64_t d : dig) h = mix64(h ^ d) + kGamma;
    return h;
  };
  uint64_t h = mix64((uint64_t)g->n * kGamma ^ (uint64_t)g->nnz);
  h = mix64(h ^ hash_words((const unsigned char *)g->rowptr, (g->n + 1) * 8));
  if (g->nnz) h = mix64(h ^ hash_words((const unsigned char *)g->colidx, g->nnz * 4));
  return h;
}

int xgm_csr_get_stats(const xgm_csr *g, xgm_csr_stats *st) {
  if (!g || !st || !g->rowptr) return XGM_EINVAL;
  const int64_t n = g->n;
  std::vector<int64_t> indeg;
  if (g->flags & XGM_CSR_DAG) {
    indeg.assign(n, 0);
    parallel_for(g->nnz, 1 << 16, [&](int64_t b, int64_t e) {
      for (int64_t k = b; k < e; ++k) __atomic_fetch_add(&indeg[g->colidx[k]], 1, __ATOMIC_RELAXED);
    });
  }
  st->n = n;
  st->nnz = g->nnz;
  int64_t maxd = 0;
  long double s2 = 0, sio = 0;
  for (int64_t u = 0; u < n; ++u) {
    int64_t d = g->rowptr[u + 1] - g->rowptr[u];
    maxd = std::max(maxd, d);
    s2 += (long double)d * d;
    if (!indeg.empty()) sio += (long double)d * indeg[u];
  }
  st->max_degree = maxd;
  st->sum_deg_sq = (double)s2;
  st->sum_in_out = (double)sio;
  return XGM_OK;
}

}  // extern "C"
