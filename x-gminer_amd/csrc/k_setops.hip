// Batched set operations on the neighbour lists of a resident graph: for
// every pair (a, b) with A = N(a), B = N(b), truncated below an optional
// bound, the intersection count, the difference count with or without the
// owner exclusion, or the ordered result set; optionally excluding given
// ancestor vertices.
//
// Reference semantics: intersect_num_bs_cache (+ upper bound) and
// intersect_bs_cache (xgminer/include/glumin/set_intersect.cuh:99-132,
// 440-465, 596-637); intersect_num with ancestors (:640-707);
// difference_num_bs_cache without owner exclusion (set_difference.cuh:20-35,
// 110-129) and difference_set_bs_cache (:218-241, 308-333); GraphMiner's
// VertexSet::difference_num, which also drops the owner vid(B) = b
// (codes/GraphMiner/src/common/VertexSet.cc:69-88); intersect_merge /
// intersect_num_merge / intersect_num_hybrid (operations.cuh:151-204,
// 208-300; set_intersect.cuh:468-515).
//
// A wave per pair, three strategies with identical results:
//  * binary search: the list searched into (L) sits in LDS when it has at
//    most kStage elements, else a 64-sample LDS index narrows a binary
//    search in HBM to |L|/64 elements; the other list is streamed 64
//    elements per window;
//  * merge path: both lists are merged in tiles of 512 path elements; the
//    tile boundaries of 64 tiles at a time are found by one merge-path
//    search per lane in HBM, a tile's two list pieces are staged in LDS with
//    coalesced loads, and every lane merges its own 8 path elements after a
//    merge-path search in LDS (work per lane fixed, whatever the list
//    lengths);
//  * hybrid: merge path when both lists exceed 1024 elements, else binary
//    search (the reference's ADJ_SIZE_THREASHOLD rule, common.h:61).
// Results are written in order with a ballot prefix (binary search) or a
// wave scan of per-lane counts (merge path).
#include <algorithm>
#include <map>
#include <mutex>

#include "device_common.h"
#include "kernels.h"
#include "merge_path.h"

namespace xgm {
namespace {

constexpr int kStage = 2048;
constexpr int kSetopWaves = 4;  // waves per block
constexpr int kMaxAnc = 4;

struct WaveStage {
  int list[kStage];
  int sample[64];
  int split[64];
};
static_assert(2 * kMergeTile + 1 <= kStage, "merge tile fits the stage");

// Membership of key in L (length ll, sorted): LDS copy when staged, else
// the sample index then HBM.
__device__ __forceinline__ bool in_list(const WaveStage &S, bool staged, const int32_t *__restrict__ L, int ll,
                                        int key) {
  if (ll <= 0) return false;
  if (staged) {
    const int i = lower_bound_dev(S.list, ll, key);
    return i < ll && S.list[i] == key;
  }
  // sample[j] = L[j * ll / 64]: the key lies in [L(j), L(j+1)) of the last
  // sample <= key
  const int j = upper_bound_dev(S.sample, 64, key) - 1;
  if (j < 0) return false;
  const int lo = (int)((int64_t)j * ll / 64);
  const int hi = j == 63 ? ll : (int)((int64_t)(j + 1) * ll / 64);
  const int i = lo + lower_bound_dev(L + lo, hi - lo, key);
  return i < hi && L[i] == key;
}

__device__ __forceinline__ void stage(WaveStage &S, const int32_t *__restrict__ L, int ll, int lane, bool &staged) {
  staged = ll <= kStage;
  if (staged) {
    for (int i = lane; i < ll; i += 64) S.list[i] = L[i];
  } else {
    S.sample[lane] = L[(int64_t)lane * ll / 64];
  }
  wave_sync();
}

__global__ __launch_bounds__(64 * kSetopWaves) void k_setop_pairs(
    const int64_t *__restrict__ rowptr, const int32_t *__restrict__ colidx, const int32_t *__restrict__ pa,
    const int32_t *__restrict__ pb, const int32_t *__restrict__ bounds, const int32_t *__restrict__ anc, int n_anc,
    int64_t npairs, int op, int strategy, uint32_t *__restrict__ counts, const int64_t *__restrict__ offsets,
    int32_t *__restrict__ values) {
  __shared__ WaveStage stages[kSetopWaves];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  WaveStage &S = stages[wave];
  const int64_t nw = (int64_t)gridDim.x * kSetopWaves;
  for (int64_t p = (int64_t)blockIdx.x * kSetopWaves + wave; p < npairs; p += nw) {
    const int a = pa[p], b = pb[p];
    const int ub = bounds ? bounds[p] : INT32_MAX;
    // ancestors of this pair (at most kMaxAnc; -1 = unused), wave-uniform
    int an[kMaxAnc];
#pragma unroll
    for (int q = 0; q < kMaxAnc; ++q) an[q] = (anc && q < n_anc) ? anc[p * n_anc + q] : -1;
    const int32_t *A = colidx + rowptr[a];
    const int32_t *B = colidx + rowptr[b];
    // truncate both lists below the bound
    const int la = lower_bound_dev(A, (int)(rowptr[a + 1] - rowptr[a]), ub);
    const int lb = lower_bound_dev(B, (int)(rowptr[b + 1] - rowptr[b]), ub);
    const bool intersect = op == SETOP_INTERSECT;
    const bool excl = op == SETOP_DIFF_EXCL;
    auto keep = [&](int x, bool found) {
      bool k = intersect ? found : (!found && !(excl && x == b));
#pragma unroll
      for (int q = 0; q < kMaxAnc; ++q) k = k && x != an[q];
      return k;
    };
    const bool merge = strategy == SETOP_STRAT_MERGE ||
                       (strategy == SETOP_STRAT_HYBRID && la > kHybridMin && lb > kHybridMin);
    if (merge) {
      // A is the list whose elements are kept (either list for an intersection)
      wave_sync();  // the previous pair's reads of S are done
      const uint32_t hits = merge_walk(A, la, B, lb, S.list, S.list + kMergeTile, S.split, lane, keep,
                                       values ? values + offsets[p] : nullptr);
      if (lane == 0) counts[p] = hits;
      continue;
    }
    // intersections stream the shorter list; differences stream A
    const bool swap = intersect && la > lb;
    const int32_t *T = swap ? B : A;   // streamed
    const int lt = swap ? lb : la;
    const int32_t *L = swap ? A : B;   // searched
    const int ll = swap ? la : lb;
    bool staged = true;
    wave_sync();  // the previous pair's reads of S are done
    if (ll > 0) stage(S, L, ll, lane, staged);
    uint32_t hits = 0;
    int64_t out = values ? offsets[p] : 0;
    for (int k0 = 0; k0 < lt; k0 += 64) {
      const int k = k0 + lane;
      bool kept = false;
      if (k < lt) {
        const int x = T[k];
        kept = keep(x, in_list(S, staged, L, ll, x));
      }
      const u64 m = __ballot(kept);
      if (values && kept) values[out + __popcll(m & lanemask_lt(lane))] = T[k];
      out += __popcll(m);
      hits += (uint32_t)__popcll(m);
    }
    if (lane == 0) counts[p] = hits;
  }
}

}  // namespace

// Staging of the allocation-free entry point: pair ids a, b, bounds and
// the counts, kNoAllocPairs each, in one file-scope device array (every GPU
// the library runs on holds its own copy until the process ends).  Host code
// reaches it only through the address hipGetSymbolAddress returns on that
// device, never through the name.
static __device__ __align__(256) int32_t scratch[4 * kNoAllocPairs];

hipError_t setop_scratch_address(int device, int32_t **ptr) {
  // looked up once per device, with that device current, by the caller's
  // context set-up; later lookups return the cached address
  static std::mutex mu;
  static std::map<int, int32_t *> addr;
  std::lock_guard<std::mutex> lock(mu);
  auto it = addr.find(device);
  if (it != addr.end()) {
    *ptr = it->second;
    return hipSuccess;
  }
  void *p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  addr[device] = static_cast<int32_t *>(p);
  *ptr = static_cast<int32_t *>(p);
  return hipSuccess;
}

hipError_t launch_setop_pairs(const DevCSR &g, int op, int strategy, const int32_t *a, const int32_t *b,
                              const int32_t *bounds, const int32_t *anc, int n_anc, int64_t npairs, uint32_t *counts,
                              const int64_t *offsets, int32_t *values, Launch &l) {
  if (n_anc < 0 || n_anc > kMaxAnc) return hipErrorInvalidValue;
  if (npairs <= 0) return hipSuccess;
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_setop_pairs, 64 * kSetopWaves, 0) != hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  const int64_t want = (npairs + kSetopWaves - 1) / kSetopWaves;
  const int grid = (int)std::max<int64_t>(1, std::min<int64_t>(want, (int64_t)per_cu * l.num_cu));
  static const char *names[3][2] = {{"setop_pairs_count", "setop_pairs_materialise"},
                                    {"setop_merge_count", "setop_merge_materialise"},
                                    {"setop_hybrid_count", "setop_hybrid_materialise"}};
  l.begin(names[strategy][values ? 1 : 0], npairs);
  k_setop_pairs<<<grid, 64 * kSetopWaves, 0, l.stream>>>(g.rowptr, g.colidx, a, b, bounds, anc, n_anc, npairs, op,
                                                         strategy, counts, offsets, values);
  l.end();
  return hipGetLastError();
}

}  // namespace xgm
