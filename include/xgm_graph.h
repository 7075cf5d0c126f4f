/*
 * xgm_graph.h - host-side graph substrate of libxgm (C ABI).
 *
 * Everything here runs on host cores and touches no GPU: deterministic
 * synthetic generators, CSR cleaning (symmetrize, dedup, self-loop removal),
 * degree orientation, the reference's on-disk CSR format, and a content hash
 * that ties a count to the exact graph it was computed on.
 *
 * CSR layout used everywhere: int64 rowptr[n+1], int32 colidx[nnz], every
 * neighbour list strictly increasing (sorted, no duplicates, no self-loops).
 *
 * Ownership: a generated or loaded graph is an opaque xgm_hgraph owned by the
 * library until xgm_hgraph_free(); xgm_hgraph_view() exposes read-only
 * pointers into it.  Nothing allocated by the caller is ever freed here.
 */
#ifndef XGM_GRAPH_H
#define XGM_GRAPH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes shared by every libxgm entry point (0 = success). */
enum xgm_status {
  XGM_OK = 0,
  XGM_EINVAL = 1,       /* bad argument or malformed CSR */
  XGM_ENOMEM = 2,       /* host or device allocation failed */
  XGM_EUNSUPPORTED = 3, /* pattern / size / option not supported */
  XGM_EHIP = 4,         /* a HIP runtime call failed */
  XGM_ENODEV = 5,       /* no usable gfx950 device */
  XGM_EIO = 6           /* file could not be read or written */
};

/* CSR property flags (xgm_csr.flags). */
#define XGM_CSR_SORTED    0x1u /* every list strictly increasing */
#define XGM_CSR_SYMMETRIC 0x2u /* undirected: v in N(u) <=> u in N(v) */
#define XGM_CSR_DAG       0x4u /* degree-oriented DAG (see xgm_orient_dag) */

/* Borrowed view of a CSR graph.  Replaces Graph_V2's rowptr/colidx pair
 * (reference: xgminer/include/graph_v2.h:237-238, 287-288) and GraphMiner's
 * Graph::vertices/edges (codes/GraphMiner/include/graph.h). */
typedef struct xgm_csr {
  int64_t n;              /* number of vertices */
  int64_t nnz;            /* number of stored (directed) entries */
  const int64_t *rowptr;  /* n+1 entries, rowptr[0]=0, rowptr[n]=nnz */
  const int32_t *colidx;  /* nnz entries */
  uint32_t flags;         /* XGM_CSR_* */
} xgm_csr;

/* Library-owned host graph. */
typedef struct xgm_hgraph xgm_hgraph;

/* Fill *view with pointers into g (valid until xgm_hgraph_free(g)). */
int xgm_hgraph_view(const xgm_hgraph *g, xgm_csr *view);
void xgm_hgraph_free(xgm_hgraph *g);

/* R-MAT generator: 2^scale vertices, edge_factor*2^scale samples; each
 * sample descends `scale` levels choosing quadrant a/b/c/d by a 32-bit
 * threshold test on a counter-based splitmix64 stream keyed by (seed,
 * sample index).  The result is symmetrized, deduplicated and stripped of
 * self-loops (flags SORTED|SYMMETRIC).  No reference counterpart: the
 * reference only ships networkx generators (scripts/graph_generator.py:10-101). */
int xgm_gen_rmat(int scale, int edge_factor, double a, double b, double c,
                 uint64_t seed, xgm_hgraph **out);

/* Chung-Lu generator: expected-degree weights w_i = min(w_max, A*(i+1)^(-1/(beta-1)))
 * with A fitted so that mean(w) = avg_degree; round(n*avg_degree/2) edge
 * samples, each endpoint drawn with probability proportional to w.  Same
 * cleaning as xgm_gen_rmat. */
int xgm_gen_chung_lu(int64_t n, double avg_degree, double beta, double w_max,
                     uint64_t seed, xgm_hgraph **out);

/* Erdos-Renyi G(n, M) style: M = round(n*avg_degree/2) uniform samples.
 * Same cleaning as xgm_gen_rmat. */
int xgm_gen_erdos_renyi(int64_t n, double avg_degree, uint64_t seed,
                        xgm_hgraph **out);

/* Build a cleaned undirected CSR (SORTED|SYMMETRIC) from an edge list of
 * m pairs; out-of-range ids give XGM_EINVAL. */
int xgm_build_undirected(int64_t n, int64_t m, const int32_t *src,
                         const int32_t *dst, xgm_hgraph **out);

/* Degree orientation: keep u->v iff deg(v) > deg(u) or (deg(v) == deg(u)
 * and v > u), deg = undirected degree; lists stay id-sorted.  Replaces
 * Graph::orientation (codes/GraphMiner/src/common/graph.cc:233-279) and
 * Graph_V2::orientation (xgminer/src/graph_v2.cpp:243-289).  Input must be
 * SORTED|SYMMETRIC; output flags SORTED|DAG. */
int xgm_orient_dag(const xgm_csr *g, xgm_hgraph **out);

/* Reference on-disk format: <prefix>.meta.txt, <prefix>.vertex.bin (int64
 * rowptr), <prefix>.edge.bin (int32 colidx).  Replaces Graph_V2's loader
 * (xgminer/src/graph_v2.cpp:5-134) and the writer in
 * scripts/graph_format_conversion.py:47-79.  `flags` is stored into the
 * returned graph's view (the format does not record it). */
int xgm_load_csr(const char *prefix, uint32_t flags, xgm_hgraph **out);
int xgm_save_csr(const char *prefix, const xgm_csr *g);

/* Edge-list text (the SNAP form GraphPi reads, xgminer/src/dataloader.cpp:
 * 29-175): a header "n m", then m lines "u v" of arbitrary integer ids.
 * Ids are renumbered 0..n-1 in order of first appearance, self-loops are
 * dropped, duplicates merged, and the result is the sorted symmetric CSR.
 * XGM_EIO if the file is unreadable, the line count differs from m, or the
 * number of distinct endpoints differs from n (the reference's own
 * consistency checks).  The format is detected as in
 * xgminer/src/common.cpp:66-103 by the xgminer front end. */
int xgm_load_edgelist(const char *path, xgm_hgraph **out);

/* Check the CSR invariants implied by g->flags; XGM_OK or XGM_EINVAL. */
int xgm_csr_validate(const xgm_csr *g);

/* 64-bit content hash of (n, nnz, rowptr, colidx); independent of thread
 * count.  Used to key cached oracle counts to the exact graph. */
uint64_t xgm_csr_hash(const xgm_csr *g);

/* Summary statistics used by the schedulers and the benchmark. */
typedef struct xgm_csr_stats {
  int64_t n, nnz;
  int64_t max_degree;
  double sum_deg_sq;        /* sum_v deg(v)^2 */
  double sum_in_out;        /* sum_v indeg(v)*outdeg(v) (DAG only, else 0) */
} xgm_csr_stats;
int xgm_csr_get_stats(const xgm_csr *g, xgm_csr_stats *st);

/* Human-readable text for a status code (static storage). */
const char *xgm_strerror(int status);

#ifdef __cplusplus
}
#endif
#endif /* XGM_GRAPH_H */
