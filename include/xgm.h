/*
 * xgm.h - pattern-counting C ABI of libxgm (HIP / gfx950).
 *
 * Contract kept from the reference driver: a CSR graph and a pattern go in,
 * one uint64 count comes out.  The reference binds the same contract by
 * link-time solver substitution -- TCSolver(Graph&, uint64_t&, int, int)
 * (codes/GraphMiner/src/triangle/main.cc:5), CliqueSolver(Graph&, int k,
 * uint64_t&, int, int) (codes/GraphMiner/src/clique/main.cc:6),
 * SglSolver(Graph&, Pattern&, uint64_t&, int, int)
 * (codes/GraphMiner/src/sgl/main.cc:7) -- and, in xgminer, by the GLUMIN
 * solver methods CliqueSolver_on_G2Miner / CliqueSolver_LUT_on_G2Miner /
 * PatternSolver_LUT_on_G2Miner(Graph_V2&) (xgminer/include/glumin.h:82-86)
 * that leave the count in `std::vector<uint64_t> total`
 * (xgminer/include/glumin.h:108).
 *
 * Every entry point returns an xgm_status (0 = XGM_OK); nothing exits or
 * throws across the ABI (the reference's CUDA_SAFE_CALL exits instead,
 * xgminer/include/cuda_utils/gpu_utils.h:4-12).  Ownership: the caller owns
 * host CSR arrays; an xgm_ctx owns every device buffer, stream and event it
 * creates and frees them in xgm_ctx_destroy().  A context is not thread-safe;
 * use one per host thread.
 */
#ifndef XGM_H
#define XGM_H

#include <stdint.h>
#include "xgm_graph.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Patterns.  Mapping from the reference's ids:
 *   XGM_TRIANGLE     <- GraphMiner tc (triangle/omp_base.cc:6-26); xgminer
 *                       XGMinerPatternType::TRIANGLE (common.h:187)
 *   XGM_CLIQUE4      <- GLUMIN --patternID 20 (local id 4, glumin.h:36-39),
 *                       GraphMiner clique k=4 (clique/cpu_kernels/automine_omp.h:67-83)
 *   XGM_CLIQUE5      <- GLUMIN --patternID 21 (local id 5),
 *                       GraphMiner clique k=5 (automine_omp.h:138-157)
 *   XGM_DIAMOND_EDGE <- GraphMiner sgl "diamond" (sgl/cpu_kernels/diamond.h:1-14)
 *   XGM_CYCLE4_EDGE  <- GraphMiner sgl "rectangle" (sgl/cpu_kernels/rectangle.h:1-11)
 * TRIANGLE / CLIQUE4 / CLIQUE5 count on a degree-oriented DAG
 * (XGM_CSR_DAG); DIAMOND_EDGE / CYCLE4_EDGE on the undirected graph
 * (XGM_CSR_SYMMETRIC).  Counts are unique embeddings (GraphMiner semantics). */
enum xgm_pattern {
  XGM_TRIANGLE = 1,
  XGM_CLIQUE4 = 2,
  XGM_CLIQUE5 = 3,
  XGM_DIAMOND_EDGE = 4,
  XGM_CYCLE4_EDGE = 5,
  /* Vertex-induced 4-vertex motifs on the undirected graph, each induced
   * embedding counted once (GraphMiner's unique-embedding semantics,
   * src/motif/cpu_kernels/automine_base.h:23-74):
   *   XGM_STAR3_V      <- GLUMIN --patternID 17 (local id 1, P1; GM_kernels/P1-GM.yml.cuh)
   *   XGM_DIAMOND_V    <- GLUMIN --patternID 18 (local id 2, P2; P2-GM.yml.cuh).  GLUMIN's P2
   *                       walks both directions of every edge and reports twice this count.
   *   XGM_TAILEDTRI_V  <- GLUMIN --patternID 19 (local id 3, P3; P3-GM.yml.cuh)
   * Strategies: LUT -- LDS LUT over N(v0) with AND-NOT row popcounts for
   * degree <= 1024 (GM_LUT_kernels/P1.yml.cuh, codegen_bitmap.cuh:392-407), set
   * difference and hybrid merge / binary-search intersections above; LIST --
   * the set-difference path for every vertex.  The enumeration is cubic in the
   * degree, so AUTO is LUT while every vertex has at most 1024 neighbours and
   * otherwise the census identities of xgm_motifs (src/motif/omp_formula.cc:
   * 40-45: per-edge sums of the motif edge kernel and the 4-cliques of the
   * degree-oriented copy).  All give the same counts. */
  XGM_STAR3_V = 6,
  XGM_DIAMOND_V = 7,
  XGM_TAILEDTRI_V = 8
};

/* Intersection / counting strategies (xgm_opts.strategy).  All produce
 * identical counts; they differ only in speed. */
enum xgm_strategy {
  XGM_STRAT_AUTO = 0,   /* library default: triangles -- LDS hash; cliques -- rank-ordered LUT
                           (as XGM_STRAT_LUT_RANK, but tasks of d+ <= 64 on the packed id-ordered
                           LUT; XGM_STRAT_LUT on a DAG not in (degree, id) orientation);
                           diamonds / 4-cycles -- the tiered kernels */
  XGM_STRAT_HASH = 1,   /* vertex task, LDS hash of the staged list, flattened probes */
  XGM_STRAT_BSEARCH = 2,/* vertex task, binary search in the staged sorted list */
  XGM_STRAT_LUT = 3,    /* cliques: vertex tasks with a bitmap LUT in LDS, edge LUT above lut_max */
  XGM_STRAT_LIST = 4,   /* cliques: edge tasks, nested list intersections (G2Miner plain);
                           4-cycle: bounded pairwise intersections in the reference loop order */
  XGM_STRAT_ELUT = 5,   /* cliques: edge tasks, bitmap LUT over N+(v0) ∩ N+(v1) */
  XGM_STRAT_LUT_RANK = 6,/* cliques: vertex LUT with rows in DAG order (upper-triangular) */
  XGM_STRAT_BUCKET = 7, /* triangles and vertex-induced diamonds: X-GMiner bucket-bitmap
                           prefilter -- lists split into 64 buckets by id & 63, only buckets
                           non-empty on both sides intersected (xgminer/include/xgminer_bitmap.cuh:
                           161-320, 661-822; the reference's P2 kernel,
                           xgminer/src/gpu_kernels/xgminer/P2_kernel.cuh:110-230); degrees < 65536 */
  XGM_STRAT_MERGE = 8   /* triangles: merge-based intersection -- one lane merges N+(u) with N+(v)
                           per DAG edge, the wave64 merge path when both lists exceed 1024
                           (intersect_num_merge / intersect_num_hybrid,
                           xgminer/include/glumin/set_intersect.cuh:468-515, operations.cuh:208-300) */
};

typedef struct xgm_opts {
  int strategy;          /* xgm_strategy; 0 = auto */
  int lut_max_degree;    /* cliques: vertex LUT for out-degree <= this; 0 = 1024 */
  uint32_t flags;        /* reserved, must be 0 */
  int hash_shift;        /* triangles: LDS hash load factor 2^-shift, 1 or 2 (LDS per wave); 0 = 2 */
  int wave_max_small;    /* triangles: staged out-degree up to this runs one wave per item in the
                            64-slot kernel (1..64); 0 = 64 */
  int wave_max_large;    /* triangles: ... in the 256-slot wave kernel (wave_max_small..256), larger
                            staged lists run one block per item; 0 = 64 (no 256-slot tier) */
  int hub_rows;          /* 4-cycles: HBM counter rows shared by hub tasks in one round (each row
                            spans the largest instance); 0 = as many as fit in 8 GiB */
  int wedge_range;       /* 4-cycles: counters of tasks above 4096 wedges in LDS by id range
                            (0 = where the range cuts are at most as many as the wedges,
                            1 = every task above 512 wedges, 2 = never, 3 = as 1 with 32-bit
                            counters only; otherwise ranges of tasks with < 65536 lower
                            neighbours use 16-bit counters over twice the ids) */
} xgm_opts;

typedef struct xgm_ctx xgm_ctx;
typedef struct xgm_dgraph xgm_dgraph;

/* Create a context on HIP device `device` (must be gfx950). */
int xgm_ctx_create(int device, xgm_ctx **out);
void xgm_ctx_destroy(xgm_ctx *ctx);

/* Last error message recorded by ctx (static until the next call). */
const char *xgm_ctx_last_error(const xgm_ctx *ctx);

/* Where a count builds the task lists it has not cached yet: its work items,
 * orders and relabelled copies.  XGM_PREP_AUTO (the default): on the host
 * when a host copy of the graph's CSR is at hand (a DAG uploaded from the
 * host), else in HBM with no copy of the CSR back (undirected uploads, DAGs
 * from xgm_graph_orient, batches).  XGM_PREP_HOST: always on the host
 * (copying the CSR back once when needed).  XGM_PREP_DEVICE: always in HBM.
 * Both routes give the same lists.  Replaces the host task preparation of
 * xgminer/src/graph_v2.cpp:307-358 and glumin_GM_solver.cu:326-335.
 * Lists cached on a graph stay as built. */
enum { XGM_PREP_AUTO = 0, XGM_PREP_HOST = 1, XGM_PREP_DEVICE = 2 };
int xgm_ctx_set_prep(xgm_ctx *ctx, int route);

/* Copy a CSR into HBM and build its task lists (degree bins).  The CSR must
 * satisfy xgm_csr_validate() for its flags; nnz must be below 2^31.  A device
 * graph is used only with the context that uploaded it; xgm_graph_free() may
 * run before or after xgm_ctx_destroy().  Replaces GraphGPU's H2D copy
 * (xgminer/include/graph_v2_gpu.h:81-134) and the host task prep of
 * glumin_GM_solver.cu:326-335. */
int xgm_graph_upload(xgm_ctx *ctx, const xgm_csr *g, xgm_dgraph **out);
void xgm_graph_free(xgm_dgraph *g);

/* Degree-orient an uploaded undirected graph (sorted, symmetric, not a
 * batch) in HBM: the DAG, its reverse (in-neighbour) lists and its vertex
 * task order are built by kernels, with no host pass over the edges; only
 * d+ and the task order (two ints per vertex) are copied back -- the CSR
 * stays in HBM and the triangle items are built there too.  *out equals
 * xgm_graph_upload() of xgm_orient_dag() of the same CSR.  Replaces
 * Graph::orientation (codes/GraphMiner/src/common/graph.cc:233-279) and
 * Graph_V2::orientation (xgminer/src/graph_v2.cpp:243-289). */
int xgm_graph_orient(xgm_ctx *ctx, const xgm_dgraph *und, xgm_dgraph **out);

/* Copy a device graph's CSR from HBM into a new host graph (flags kept). */
int xgm_graph_download(xgm_ctx *ctx, const xgm_dgraph *g, xgm_hgraph **out);

/* Vertex task order of a DAG device graph: vertices with d+ >= 2 by
 * decreasing sum over N+(u) of d+, ties by increasing id.  Writes
 * min(cap, len) ids to out; *len receives the full length. */
int xgm_graph_task_order(const xgm_dgraph *g, int32_t *out, int64_t cap, int64_t *len);

/* Triangle task items of a DAG device graph under opts (strategy AUTO /
 * HASH / BSEARCH and its tier split): int4 (t, k0, k1, 0) per item, tiers
 * concatenated (4 of them, items per tier in tier_counts, which may be
 * NULL), each tier by decreasing work.  Built on the host for a graph
 * uploaded from a host CSR and in HBM for one built by xgm_graph_orient;
 * both give the same items.  Writes min(cap, len) ints; *len = 4 x items. */
int xgm_graph_tc_items(xgm_ctx *ctx, const xgm_dgraph *g, const xgm_opts *opts, int32_t *out, int64_t cap,
                       int64_t *len, int64_t *tier_counts);

/* Task lists cached on a device graph, for inspection (no counterpart in
 * the reference): xgm_graph_task_list_names writes the names of the cached
 * lists, each ended by '\n', into names (min(cap, len) bytes; *len = the full
 * length); xgm_graph_task_list copies the list `name` (ints, tiers
 * concatenated: *len ints in all, tier sizes in units of tasks into
 * counts[0 .. min(counts_cap, *n_counts))).  Names: "tc_items_<s>_<l>"
 * (triangles, int4 items), "diamond" / "motif_edges" (int4 owner items),
 * "vmotif" / "vmotif_list" (vertex ids), "cycle4_wedge_<m>" (vertex ids),
 * "cycle4_range_<m>" / "cycle4_hubs_<rows>_<m>" (int4 items). */
int xgm_graph_task_list_names(const xgm_dgraph *g, char *names, int64_t cap, int64_t *len);
int xgm_graph_task_list(xgm_ctx *ctx, const xgm_dgraph *g, const char *name, int32_t *out, int64_t cap,
                        int64_t *len, int64_t *counts, int counts_cap, int *n_counts);

/* Count `pattern` on a device graph; *out receives the uint64 count.
 * opts may be NULL (auto).  Synchronous. */
int xgm_count(xgm_ctx *ctx, const xgm_dgraph *g, int pattern, const xgm_opts *opts,
              uint64_t *out);

/* One-shot convenience: upload, count, free. */
int xgm_count_csr(xgm_ctx *ctx, const xgm_csr *g, int pattern, const xgm_opts *opts,
                  uint64_t *out);

/* ---- set operations on neighbour lists --------------------------------
 * For every pair i, A = N(a[i]) and B = N(b[i]) (the lists of g), both
 * restricted to elements < bounds[i] (bounds may be NULL: no bound):
 *   XGM_SETOP_INTERSECT  |A ∩ B|          intersect_num_bs_cache (+ upper bound),
 *                                        xgminer/include/glumin/set_intersect.cuh:440-465, 596-637
 *   XGM_SETOP_DIFF       |A \ B|          GLUMIN difference_num_bs_cache (no owner
 *                                        exclusion), set_difference.cuh:20-35, 110-129
 *   XGM_SETOP_DIFF_EXCL  |A \ B \ {b[i]}| GraphMiner VertexSet::difference_num,
 *                                        codes/GraphMiner/src/common/VertexSet.cc:69-88
 * counts[i] receives the size.  With values != NULL the sets are also
 * written in increasing order (intersect_bs_cache, set_intersect.cuh:99-132;
 * difference_set_bs_cache, set_difference.cuh:218-241): offsets (n_pairs + 1,
 * caller-allocated) receives the exclusive scan of counts and
 * values[offsets[i] .. offsets[i+1]) set i.  If offsets[n_pairs] > values_cap
 * values is not written and XGM_EINVAL is returned (offsets still filled).
 * a, b, bounds, counts, offsets and values are host arrays. */
enum { XGM_SETOP_INTERSECT = 1, XGM_SETOP_DIFF = 2, XGM_SETOP_DIFF_EXCL = 3 };
/* Strategy, OR-ed into op (same results; default = binary search of the
 * streamed list's elements in the other list, staged in LDS up to 2,048
 * elements, else narrowed by a 64-sample LDS index):
 *   XGM_SETOP_MERGE   wave64 merge path over both lists (tiles of 512 path
 *                     elements, 8 per lane) -- intersect_merge /
 *                     intersect_num_merge, xgminer/include/glumin/
 *                     operations.cuh:151-204, set_intersect.cuh:468-515
 *   XGM_SETOP_HYBRID  merge path when both lists exceed 1024 elements, else
 *                     binary search -- intersect_num_hybrid,
 *                     operations.cuh:208-300 (ADJ_SIZE_THREASHOLD, common.h:61) */
enum { XGM_SETOP_MERGE = 0x100, XGM_SETOP_HYBRID = 0x200 };
int xgm_setop_pairs(xgm_ctx *ctx, const xgm_dgraph *g, int op, const int32_t *a, const int32_t *b,
                    const int32_t *bounds, int64_t n_pairs, uint32_t *counts, int64_t *offsets, int32_t *values,
                    int64_t values_cap);

/* Counts of op (with its strategy flag) excluding ancestor vertices: for
 * pair i, the elements x of the result with x != ancestors[i * n_anc + q]
 * for q < n_anc (n_anc <= XGM_SETOP_MAX_ANCESTORS; -1 entries are unused).
 * For XGM_SETOP_INTERSECT this is intersect_num(a, b, [upper_bound,]
 * ancestor) / intersect_num(a, b, ancestor[], n)
 * (xgminer/include/glumin/set_intersect.cuh:640-707).  Host arrays. */
#define XGM_SETOP_MAX_ANCESTORS 4
int xgm_setop_pairs_anc(xgm_ctx *ctx, const xgm_dgraph *g, int op, const int32_t *a, const int32_t *b,
                        const int32_t *bounds, const int32_t *ancestors, int n_anc, int64_t n_pairs, uint32_t *counts);

/* Allocation-free variant of xgm_setop_pairs (counts only, values are not
 * materialised).  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  The pair arrays are
 * staged in a static device array of the library (one per GPU, 16 MiB,
 * held until the process ends) instead of buffers allocated per call; no
 * hipMalloc / hipFree is called.  n_pairs above XGM_NOALLOC_MAX_PAIRS ->
 * XGM_EINVAL.  Calls that overlap on one GPU (several threads or contexts)
 * share that array and race: serialise them.  Same counts as
 * xgm_setop_pairs for the same arguments. */
#define XGM_NOALLOC_MAX_PAIRS (1 << 20)
int xgm_setop_pairs_noalloc(xgm_ctx *ctx, const xgm_dgraph *g, int op, const int32_t *a, const int32_t *b,
                            const int32_t *bounds, int64_t n_pairs, uint32_t *counts);

/* ---- vertex-induced motif census -------------------------------------
 * g undirected (or a batch of undirected instances).  k = 3: out[2 * i ..] =
 * {wedges, triangles}; k = 4: out[6 * i ..] = {3-stars, 4-paths, tailed
 * triangles, 4-cycles, diamonds, 4-cliques} per instance i, in the order and
 * with the counts of the reference's formula path (GraphMiner
 * src/motif/omp_formula.cc:8-52, cpu_kernels/automine_formula.h:1-58): the
 * per-edge sums of that formula on the GPU, 4-cliques and triangles on a
 * degree-oriented copy (built on first use, kept with g), and the induced
 * 4-cycles from the edge-induced count (C4 = I(C4) + I(diamond) + 3 K4). */
int xgm_motifs(xgm_ctx *ctx, const xgm_dgraph *g, int k, uint64_t *out);

/* ---- batches of instances (disjoint union) ---------------------------
 * Upload n_inst graphs of one kind (same flags) as one device graph; every
 * count on it can return one uint64 per instance.  Replaces running the
 * reference solver once per graph. */
int xgm_graph_upload_batch(xgm_ctx *ctx, const xgm_csr *graphs, int n_inst, xgm_dgraph **out);
int xgm_graph_instances(const xgm_dgraph *g);
/* out[i] = count of `pattern` in instance i (n_inst entries). */
int xgm_count_instances(xgm_ctx *ctx, const xgm_dgraph *g, int pattern, const xgm_opts *opts, uint64_t *out);
/* One call for many (graph, pattern) instances: gs[i] is an undirected
 * sorted CSR (XGM_CSR_SORTED | XGM_CSR_SYMMETRIC), patterns[i] its pattern;
 * triangle and clique instances are oriented by the library.  Instances of
 * one pattern are counted as one batch.  out[i] = count of instance i. */
int xgm_count_batch(xgm_ctx *ctx, const xgm_csr *gs, const int *patterns, int n_inst, uint64_t *out);

/* A plan captures the counts of several patterns -- triangles / cliques on a
 * DAG batch, diamonds / 4-cycles on an undirected batch (either may be NULL
 * if unused) -- into one hipGraph of persistent work-queue kernels that is
 * replayed by xgm_plan_run.  opts: one per pattern (NULL = auto).  A plan
 * holds its graphs: xgm_graph_free() on a graph a live plan uses only marks
 * it, and the graph's device memory is released by the last xgm_plan_free()
 * of a plan using it.  Later counts on those graphs leave the plan valid
 * (scratch buffers they outgrow are kept until the graph is released). */
typedef struct xgm_plan xgm_plan;
int xgm_plan_create(xgm_ctx *ctx, const xgm_dgraph *dag, const xgm_dgraph *und, const int *patterns, int n_pat,
                    const xgm_opts *opts, xgm_plan **out);
/* As xgm_plan_create, with an explicit clique schedule per pattern: scheds
 * (NULL, or one entry per pattern, NULL = the pattern's opts) gives a
 * schedule made by xgm_sched_create on `dag` for that pattern -- for
 * instance the learned codes of xgm_sched_learned -- so the learned
 * schedule replays inside the captured batch.  The plan keeps its own copy
 * of each schedule's task lists: a schedule may be freed after this call.
 * No counterpart in the reference (its GLUMIN clique path has one fixed
 * edge-parallel schedule, glumin_GM_solver.cu:1136-1145). */
typedef struct xgm_sched xgm_sched;
int xgm_plan_create_sched(xgm_ctx *ctx, const xgm_dgraph *dag, const xgm_dgraph *und, const int *patterns,
                          int n_pat, const xgm_opts *opts, const xgm_sched *const *scheds, xgm_plan **out);
/* out: n_pat x n_inst counts (pattern-major); *ms = replay time (HIP events). */
int xgm_plan_run(xgm_ctx *ctx, xgm_plan *plan, uint64_t *out, float *ms);
void xgm_plan_free(xgm_plan *plan);
/* Autotuner: times the candidate option sets of `pattern` on g (every
 * candidate gives the same count; a disagreement is an error) and writes the
 * fastest into *best.  Candidates: triangles -- LDS hash load 1/2 or 1/4 x
 * five wave/block tier splits, binary search, the merge strategy and (out-
 * degrees below 65536) the bucket bitmap; cliques -- auto, rank LUT,
 * id-ordered LUT with threshold 1024 or 256, edge LUT; 4-cycles -- the wedge
 * aggregation with each wedge_range mode.  Each candidate runs 1 + reps
 * times. */
int xgm_autotune(xgm_ctx *ctx, const xgm_dgraph *g, int pattern, int reps, xgm_opts *best, float *best_ms);

/* Measured HBM bandwidth: a plain streaming copy kernel (k_copy_u4) over
 * `bytes` (rounded down to 16 B) timed with HIP events, best of `reps`;
 * *gbps = 2 * bytes / time (read + write).  No counterpart in the
 * reference: the reference point for the roofline and for scaling HBM
 * traffic counters. */
int xgm_copy_bandwidth(xgm_ctx *ctx, size_t bytes, int reps, float *gbps);

/* Timing of the most recent xgm_count on ctx, from HIP events recorded on
 * the stream the kernels run on. */
#define XGM_MAX_PROFILE_KERNELS 16
typedef struct xgm_kernel_time {
  char name[40];
  float ms;              /* elapsed time of this launch */
  int64_t tasks;         /* tasks the launch processed */
} xgm_kernel_time;
typedef struct xgm_profile {
  float total_ms;        /* first launch start to last launch end */
  int n_kernels;
  xgm_kernel_time k[XGM_MAX_PROFILE_KERNELS];
} xgm_profile;
int xgm_last_profile(const xgm_ctx *ctx, xgm_profile *out);

/* ---------------------------------------------------------------------
 * Learned scheduler.  No counterpart in the reference (which uses fixed
 * degree thresholds, xgminer/include/common.h:61,114-116).  A 16-64-64-5
 * MLP predicts per-task log-costs of the clique strategies; it is trained
 * and evaluated on the GPU by the library's own BF16-MFMA kernels and only
 * chooses schedules: counts are identical under every schedule.
 * --------------------------------------------------------------------- */
#define XGM_MLP_IN 16
#define XGM_MLP_OUT 5
#define XGM_MLP_PARAMS 5573 /* W1[16x64] b1[64] W2[64x64] b2[64] W3[64x5] b3[5] */
typedef struct xgm_mlp xgm_mlp;
int xgm_mlp_create(int device, uint64_t seed, xgm_mlp **out);
void xgm_mlp_free(xgm_mlp *m);
const char *xgm_mlp_last_error(const xgm_mlp *m);
int xgm_mlp_set_params(xgm_mlp *m, const float *params);
int xgm_mlp_get_params(const xgm_mlp *m, float *params);
/* Y[n x 5] = model(X[n x 16]) (host arrays, batched on the device). */
int xgm_mlp_predict(xgm_mlp *m, const float *X, int64_t n, float *Y);
/* Masked mean-squared-error training (Adam, shuffled mini-batches);
 * T, M are n x 5 targets and 0/1 masks.  *final_loss = mean batch loss of
 * the last epoch. */
int xgm_mlp_train(xgm_mlp *m, const float *X, const float *T, const float *M, int64_t n, int epochs, int batch,
                  float lr, uint64_t seed, float *final_loss);
/* One full-batch loss and gradient (XGM_MLP_PARAMS floats), for checking
 * the backward pass against a CPU reference. */
int xgm_mlp_grads(xgm_mlp *m, const float *X, const float *T, const float *M, int64_t n, float *grads,
                  float *loss);
/* Clique schedules: one code per vertex -- 0 id-ordered LUT, 1 rank-ordered
 * LUT, 2 edge LUT, 3 edge nested lists, 4 rank-ordered LUT one capacity bin
 * up; -1 for vertices that are not tasks (d+ < k-1).  On vertices with
 * 1024 < d+ <= 4096 the rank-ordered codes (1, 4) run the hub LUT (rows
 * packed upper-triangular in LDS up to 1344, in HBM above); other LUT codes
 * above 1024 run as edge LUT.  Every schedule gives the same
 * count.  (xgm_sched is declared with the plans above.) */
/* Degree-threshold codes of a strategy (XGM_STRAT_*), no model: AUTO and
 * LUT_RANK = rank-ordered LUT up to lut_max_degree and for hubs (1024 < d+ <=
 * 4096), edge LUT between lut_max_degree and 1024 and above 4096, where AUTO
 * gives tasks of d+ <= min(64, lut_max_degree) the id-ordered LUT (code 0,
 * 8 tasks per wave); LUT = id-ordered LUT up to lut_max_degree, edge LUT
 * above. */
int xgm_sched_heuristic(const xgm_dgraph *g, int pattern, int strategy, int lut_max_degree, int8_t *codes);
/* Per-vertex scheduler features (n x XGM_MLP_IN floats, computed on the GPU):
 * log out-degree, log sum / max of 2-hop out-degrees, sampled neighbour
 * overlap, LUT fill density, log estimated sub-DAG edges / triangles /
 * 4-cliques, capacity bin, row words, log in-degree, bias. */
int xgm_sched_features(xgm_ctx *ctx, const xgm_dgraph *g, int pattern, float *feats);
/* Learned codes: features, model inference and argmin on the GPU (the
 * model predicts each task's share of its kernel's time under every code;
 * a task leaves the AUTO code only for a predicted saving above the model's
 * error, and rank-LUT codes on tasks of d+ <= 64 become code 0, the packed
 * id-ordered LUT), then a critical-path guard on the host: the lists run as consecutive
 * kernels of about max(sum of shares, longest task) each, so a vertex task
 * predicted to be its kernel's tail moves to the edge LUT (single-edge
 * items) when that lowers the estimate. */
int xgm_sched_learned(xgm_ctx *ctx, const xgm_dgraph *g, int pattern, xgm_mlp *m, int8_t *codes);
int xgm_sched_create(xgm_ctx *ctx, const xgm_dgraph *g, int pattern, const int8_t *codes, xgm_sched **out);
void xgm_sched_free(xgm_sched *s);
/* Count with an explicit schedule.  vertex_us (optional, n floats) receives
 * each vertex's share of the kernel time: the wall-clock microseconds of its
 * task(s) divided by the resident blocks of the kernel that ran them. */
int xgm_count_sched(xgm_ctx *ctx, const xgm_dgraph *g, const xgm_sched *s, float *vertex_us, uint64_t *out);

/* C = A*B through the MFMA tile GEMM (A, B rounded to BF16), for checking
 * the matrix-core operand layout. */
int xgm_mfma_gemm_check(int device, int M, int N, int K, const float *A, const float *B, float *C);

#ifdef __cplusplus
}
#endif
#endif /* XGM_H */
