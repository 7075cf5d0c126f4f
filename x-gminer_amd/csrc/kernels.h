// Internal interface between the C-ABI layer (xgm_api.hip) and the kernel
// translation units.  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <functional>
#include <string>
#include <vector>

typedef unsigned long long u64;

struct xgm_mlp;

namespace xgm {

// Device-resident CSR (HBM): int64 rowptr[n+1], int32 colidx[nnz].
struct DevCSR {
  const int64_t *rowptr = nullptr;
  const int32_t *colidx = nullptr;
  int64_t n = 0, nnz = 0;
  // reverse (in-neighbour) lists of a DAG; null for undirected graphs
  const int64_t *rrowptr = nullptr;
  const int32_t *rcolidx = nullptr;
  // (degree, id)-relabelled undirected graphs (4-cycle wedges): for entry e
  // of row x with v = colidx[e] < x, cut[e] = |N(v) ∩ [0, x)|; nlow[x] =
  // |N(x) ∩ [0, x)|; null elsewhere
  const int32_t *cut = nullptr;
  const int32_t *nlow = nullptr;
};

// A list of vertex tasks (int32 ids) resident in HBM, ordered by
// decreasing estimated work so dynamic claiming approximates LPT.
struct TaskList {
  const int32_t *tasks = nullptr;
  int64_t count = 0;
};

// Everything a launcher needs from its context.
struct Launch {
  hipStream_t stream = nullptr;
  int num_cu = 256;
  u64 *d_out = nullptr;   // accumulator slot the launch adds into (pre-zeroed)
  u64 *d_err = nullptr;   // error word kernels may OR into (pre-zeroed)
  // batched instances: per-vertex output slot (d_out[slots[v]]), null = one count
  const int32_t *slots = nullptr;
  int *d_queue = nullptr; // work counter the launch claims from (pre-zeroed)
  // Bracket one kernel launch with events: begin(name, tasks) / end().
  std::function<void(const char *, int64_t)> begin;
  std::function<void()> end;
};

// Tiers of triangle-count items by the staged out-degree d+(t).
enum TcTier { TIER_W64 = 0, TIER_W256 = 1, TIER_B2048 = 2, TIER_BIG = 3, NUM_TIERS = 4 };
constexpr int kTierMaxDeg[NUM_TIERS] = {64, 256, 2048, 1 << 30};
// wave tiers take at most this many partners per item (one segment table)
constexpr int kTcWaveCap[2] = {64, 256};
// Default wave_max_large: staged lists above 64 go to the block kernel,
// whose LDS hash spans a 2048 << shift table (load <= d+/8192); on R-MAT-22
// that beats the 256-slot wave tier (28.2 vs 29.3 ms per count).
constexpr int kTcWaveLargeDefault = 64;

// Triangle counting over a degree-oriented DAG with its reverse lists.
// tiers[] hold int4 items (t, k0, k1, 0) (TaskList.count = items).
// strategy: 1 = LDS hash, 2 = LDS binary search.  Adds into l.d_out.
hipError_t launch_tc(const DevCSR &g, const TaskList tiers[NUM_TIERS], int strategy, int hash_shift, Launch &l,
                     int *d_queues /* NUM_TIERS counters */);

// Merge triangle strategy: tasks = vertices with d+ >= 2; one lane merges
// N+(u) with N+(v) per DAG edge u->v, the wave merge path for pairs with
// both lists above 1024.  Adds into l.d_out.  One queue counter.
hipError_t launch_tc_merge(const DevCSR &g, const TaskList &tasks, Launch &l, int *queue);

// Bucket-bitmap triangle strategy (k_bucket.hip): bucket-major copy of the
// DAG's lists (64 buckets by w & 63), per-vertex uint16 bucket offsets and
// non-empty masks; then one wave per task vertex (d+ >= 2) intersecting
// only the buckets set in both masks.  Adds into l.d_out.
hipError_t launch_bucket_build(const int64_t *rowptr, const int32_t *colidx, int64_t n, int32_t *bcol, uint16_t *boff,
                               u64 *bmap, int num_cu, hipStream_t s);
// Vertex-induced diamonds with the bucket prefilter over the tasks of
// degree >= 3; scratch holds vmotif_bucket_grid(num_cu) * 4 * md ints, md
// at least the largest degree.
int vmotif_bucket_grid(int num_cu);
hipError_t launch_vmotif_diamond_bucket(const DevCSR &g, const int32_t *bcol, const uint16_t *boff, const u64 *bmap,
                                        const TaskList &tasks, int32_t *scratch, int md, Launch &l, int *queue);
hipError_t launch_tc_bucket(const DevCSR &g, const int32_t *bcol, const uint16_t *boff, const u64 *bmap,
                            const TaskList &tasks, Launch &l, int *queue);

// Edge tasks (v0, v1) of a k-clique count, as int32 pairs in HBM.
struct EdgeItems {
  const int32_t *items = nullptr;  // 2 * count ints
  int64_t count = 0;
};

// Tasks of a k-clique count.  Vertex tasks run the id-ordered LUT (lut) or
// the rank-ordered LUT (rlut) in bins by capacity (64/256/512/1024 rows);
// edge tasks (v0, v1) run with the LUT (elut) or nested lists (list), split
// by min(d+(v0), d+(v1)) <= 128.  Optional per-item wall-clock logs
// (100 MHz ticks) have the same shapes.
struct CliqueTasks {
  TaskList lut[4], rlut[4];
  EdgeItems elut[2], list[2];
  TaskList rlut_hbm;        // rank-ordered LUT for hubs (1024 < d+ <= kHubLutMax): LDS up to kTriLutMax, HBM above
  u64 *hub_slots = nullptr; // clique_hub_slots() * kHubLutMax^2 / 64 words
  u64 *log_lut[4] = {}, *log_rlut[4] = {}, *log_elut[2] = {}, *log_list[2] = {}, *log_rlut_hbm = nullptr;
};
constexpr int kHubLutMax = 4096;
// Triangle items end where their probes reach kTcSlice (or at a tier's
// position cap): long enough that staging N+(t) once per item costs little
// next to its probes (2^16 -> 2^18: tc_block_2048 on R-MAT-22 12.0 -> 11.5 ms).
constexpr int64_t kTcSlice = 1 << 18;
// Diamond / motif owners of more than 1024 neighbours are cut into items of
// about kDiaSlice probes or kDiaSlicePositions partners: every item stages
// N(t) again (an LDS hash, or the binned set of a hub), so slices 16 times
// the triangle items' length keep that cost small next to the probes.
constexpr int64_t kDiaSlice = 1 << 20;
constexpr int kDiaSlicePositions = 1 << 16;
// Hubs up to kTriLutMax (21 row words) keep the rank LUT in LDS as packed
// upper-triangular rows; larger ones up to kHubLutMax use rows in HBM.
constexpr int kTriLutMax = 1344;

// k-clique counting (k = 4 or 5) over a DAG.  scratch holds
// clique_edge_grid() * 2 * md ints (md = max out-degree).  Adds into l.d_out.
hipError_t launch_clique(const DevCSR &g, const CliqueTasks &ct, int k, int *scratch, int md, Launch &l,
                         int *d_queues /* 13 counters */);
// Scheduler features of every vertex (n x 16 floats) for k-clique tasks.
hipError_t launch_clique_features(const DevCSR &g, int k, float *F, hipStream_t s);
int clique_edge_grid(int num_cu);
int clique_list_capacity(int num_cu, int list);  // list 12 = rank LUT in HBM
int clique_hub_slots(int num_cu);

// Vertex-induced 4-motifs (which: 0 3-star, 1 diamond, 2 tailed triangle)
// over an undirected CSR, each embedding once.  Vertex tasks of degree >= 3:
// tl[0..2] LUT in LDS (degree <= 64 / 256 / 1024), tl[3] the list path by set
// difference (any degree); scratch holds vmotif_list_grid() * 8 * md ints
// (md = max degree) when tl[3] is not empty.  Adds into l.d_out.  Queue
// counters 0-3.
hipError_t launch_vmotif(const DevCSR &g, int which, const TaskList tl[4], int32_t *scratch, int md, Launch &l,
                         int *d_queues);
int vmotif_list_grid(int num_cu);

// Edge-induced diamond over an undirected CSR; items are int4 (owner t,
// k0, k1, 0) (TaskList.tasks points at 4*count ints, count = items) in
// owner-degree tiers (<=64, <=1024, <=8192, larger).
hipError_t launch_diamond(const DevCSR &g, const TaskList tl[4], Launch &l, int *d_queues);
// Same items, the four per-edge sums of the 4-motif formula added into
// l.d_out[0..3] (l.d_out[4 * slot + j] with instance slots).
hipError_t launch_motif_edges(const DevCSR &g, const TaskList tl[4], Launch &l, int *d_queues);

// HBM counter rows of the 4-cycle wedge kernels: one row of `stride` ints
// spans the vertex ids of one instance (every instance occupies the
// contiguous ids [inst_lo[i], inst_lo[i] + n_i), n_i <= stride); inst_lo is
// null for a single graph (stride = n).
struct CounterRows {
  int64_t stride = 0;
  const int64_t *inst_lo = nullptr;
  // 4-cycle range items: cycle4_range_grid() rows of cursor_stride int4,
  // cursor_stride >= the lower-neighbour count of every range task; no
  // initial contents needed
  int4 *cursor = nullptr;
  int64_t cursor_stride = 0;
};

// Edge-induced 4-cycle, wedge aggregation over the (degree, id)-relabelled
// CSR; vertex tasks by wedge work (tl[0] <= 512, tl[1] <= 4096, tl[2]
// larger, tl[3] <= 128 wedges with <= 64 lower neighbours, packed); dense holds
// cycle4_dense_grid() rows, all zero on entry (left zero on exit).
// hub_rounds[r] are int4 (x, k0, k1, slot) slices of hub tasks adding into
// row `slot` of hub; the first hub_round_u64 words of hub are zeroed before
// every round.  Uses queue counters 0-2 and 4-7 of d_queues.
// range_items are int4 (x, r0, r1, half): the wedges of x ending in
// [r0, r1), counted in LDS one span after another, a span being
// kRangeCounters ids (32-bit counters, half = 0) or 2 * kRangeCounters ids
// (half = 1: 16-bit counters, two per word; only for tasks with fewer than
// 65536 lower neighbours, which bound every count).  Queue counter 6.
constexpr int kRangeCounters = 32768;
hipError_t launch_cycle4_wedge(const DevCSR &g, const TaskList tl[4], int *dense, const TaskList *hub_rounds,
                               int n_rounds, int *hub, size_t hub_round_u64, const TaskList &range_items,
                               const CounterRows &rows, Launch &l, int *d_queues);
int cycle4_dense_grid(int num_cu, int64_t tasks);
int cycle4_range_grid(int num_cu, int64_t items);

// Edge-induced 4-cycle, bounded pairwise intersections in the reference's
// loop order (original ids).
hipError_t launch_cycle4_pair(const DevCSR &g, const TaskList &tl, Launch &l, int *d_queues);

// Scheduler cost model inference on device arrays (n x 16 -> n x 5).
int mlp_predict_device(xgm_mlp *m, const float *dX, int64_t n, float *dY);
// codes[v] = argmin over the valid clique strategy codes of the predicted
// cost Y[v] (-1 where F marks v as no task).
hipError_t launch_sched_argmin(const DevCSR &g, const float *F, const float *Y, int8_t *codes, hipStream_t s);

// Small utility kernel: zero `n` 64-bit words (graph-capture safe, unlike memset nodes).
// Batched set operations on neighbour lists (a wave per pair): op is one
// of SETOP_*; counts[i] always written; values (at offsets[i]) in order when
// non-null.  Pair ids must be < g.n (checked by the caller).
enum SetOp { SETOP_INTERSECT = 1, SETOP_DIFF = 2, SETOP_DIFF_EXCL = 3 };
// Allocation-free set operations: largest pair count of one call, and the
// device address (on `device`, which must be current) of the library's
// static staging array of 4 * kNoAllocPairs int32: a, b, bounds, counts.
constexpr int64_t kNoAllocPairs = 1 << 20;
hipError_t setop_scratch_address(int device, int32_t **ptr);
// strategy: SETOP_STRAT_*; anc: n_anc (<= 4) ancestor ids per pair excluded
// from the result (-1 = none), or null.
enum SetOpStrategy { SETOP_STRAT_BSEARCH = 0, SETOP_STRAT_MERGE = 1, SETOP_STRAT_HYBRID = 2 };
hipError_t launch_setop_pairs(const DevCSR &g, int op, int strategy, const int32_t *a, const int32_t *b,
                              const int32_t *bounds, const int32_t *anc, int n_anc, int64_t npairs, uint32_t *counts,
                              const int64_t *offsets, int32_t *values, Launch &l);

// On-device degree orientation (k_prep.hip): deg[v]; out/in flag per entry
// of the undirected CSR (flag[nnz] = 0); after an exclusive scan pos of the
// nnz + 1 flags, the out/in lists and row pointers; the task sort keys; and
// the hipCUB scan / descending radix sort behind them (temp == nullptr sizes
// the temporary storage).
hipError_t launch_prep_degree(const int64_t *rowptr, int64_t n, int32_t *deg, int num_cu, hipStream_t s);
hipError_t launch_prep_split(const int64_t *rowptr, const int32_t *colidx, const int32_t *deg, int64_t n, int64_t nnz,
                             int32_t *flag, int num_cu, hipStream_t s);
hipError_t launch_prep_scatter(const int64_t *rowptr, const int32_t *colidx, const int32_t *flag, const int32_t *pos,
                               int64_t n, int64_t nnz, int64_t *orow, int32_t *ocol, int64_t *irow, int32_t *icol,
                               int num_cu, hipStream_t s);
hipError_t launch_prep_work(const int64_t *orow, const int32_t *ocol, int64_t n, uint32_t *key, int32_t *ids,
                            int32_t *dplus, int num_cu, hipStream_t s);
// bad[0] |= 1 if some entry v > u of N(u) lacks its mirror; bad[1] / bad[2]
// += entries above / below the diagonal (bad zeroed by the caller).
hipError_t launch_prep_mirror(const int64_t *rowptr, const int32_t *colidx, int64_t n, int64_t nnz, u64 *bad,
                              int num_cu, hipStream_t s);
hipError_t prep_scan(void *temp, size_t &bytes, const int32_t *in, int32_t *out, int64_t n, hipStream_t s);
hipError_t prep_sort(void *temp, size_t &bytes, const uint32_t *key_in, uint32_t *key_out, const int32_t *ids_in,
                     int32_t *ids_out, int64_t n, hipStream_t s);

// Triangle task items built in HBM from the DAG and its reverse lists,
// identical to the host builder (xgm_api.hip build_tc_items): *out (int4
// items, tiers concatenated, owned by the caller) and the items per tier.
hipError_t build_tc_items_device(const DevCSR &g, int max_small, int max_large, int num_cu, hipStream_t s,
                                 int32_t **out, int64_t counts[NUM_TIERS]);

// Task lists of an undirected graph built in HBM; each equals the host
// builder's list of the same name in xgm_api.hip.
// Diamond / motif owner items (owners of degree >= min_deg): *out
// (hipMalloc'd, the caller frees) holds counts[0..3] int4 items, tier-major.
hipError_t build_dia_items_device(const DevCSR &g, int min_deg, int num_cu, hipStream_t s, int32_t **out,
                                  int64_t counts[4]);
// Vertex-induced motif tasks: degree >= 3 by decreasing degree, lists of
// degree <= 64 / 256 / 1024 / above (list_all: all in the last); *max_deg =
// the largest degree in the last list (0 if empty).
hipError_t build_deg3_order_device(const DevCSR &g, bool list_all, int num_cu, hipStream_t s, int32_t **out,
                                   int64_t counts[4], int64_t *max_deg);
// host_out[i] = sum of d(d-1) over the vertices of instance i (slot may be
// null: one instance)
hipError_t degree_pair_sums_device(const DevCSR &g, const int32_t *slot, int n_inst, hipStream_t s, u64 *host_out);
// Relabel by ascending (degree, id) inside each instance: the relabelled CSR
// in HBM (*rl_rowptr, *rl_colidx, hipMalloc'd), and on the host the vertices
// of positive wedge work by decreasing work, the work and the lower-neighbour
// count of every vertex.
hipError_t relabel_by_degree_device(const DevCSR &g, const int32_t *slot, int num_cu, hipStream_t s,
                                    int64_t **rl_rowptr, int32_t **rl_colidx, std::vector<int32_t> &order,
                                    std::vector<int64_t> &work, std::vector<int32_t> &nlow);
// cut[e] and nlow[x] of a relabelled CSR (DevCSR::cut / nlow), hipMalloc'd
// (the caller frees)
hipError_t build_rl_cuts_device(const int64_t *rl_rowptr, const int32_t *rl_colidx, int64_t n, int64_t nnz,
                                int num_cu, hipStream_t s, int32_t **cut, int32_t **nlow);
// cnt[hoff[h] + k] = |N(v_k) ∩ [0, x)| for hub x = hubs[h] of a relabelled
// CSR and its k-th lower neighbour v_k
hipError_t cycle4_hub_counts_device(const int64_t *rl_rowptr, const int32_t *rl_colidx,
                                    const std::vector<int32_t> &hubs, const std::vector<int64_t> &hoff, int num_cu,
                                    hipStream_t s, std::vector<int32_t> &cnt);

hipError_t launch_zero(void *ptr, size_t n_u64, hipStream_t s);
hipError_t launch_copy(void *dst, const void *src, size_t n16, int num_cu, hipStream_t s);

}  // namespace xgm
