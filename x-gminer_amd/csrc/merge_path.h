// Wave64 merge path over two sorted int32 lists (Odeh, Green, Mwassi,
// Shmueli, Birk, "Merge Path - Parallel Merging Made Simple", IPDPS
// workshops 2012; Green, McColl, Bader, "GPU Merge Path", ICS 2012): the
// merge of A and B is cut into tiles of kMergeTile path elements; the tile
// boundaries of 64 tiles at a time come from one merge-path search per lane,
// a tile's two list pieces are staged in LDS with coalesced loads, and every
// lane merges its own kMergeE consecutive path elements after a second
// merge-path search in LDS.  Equal keys take A first, so the partner of an
// element of A in B, if any, is the next unconsumed element of B.
#pragma once

#include "device_common.h"

constexpr int kMergeE = 8;                // merge-path elements per lane per tile
constexpr int kMergeTile = 64 * kMergeE;  // path elements per tile
// hybrid rule: the wave merge path when both lists exceed this many
// elements (the reference's ADJ_SIZE_THREASHOLD, xgminer/include/common.h:61)
constexpr int kHybridMin = 1024;

// Number of A elements among the first d elements of the merge of A and B
// (A before B on equal keys): the first i in [max(0, d - lb), min(d, la)]
// with A[i] > B[d - 1 - i].
template <class PA, class PB>
__device__ __forceinline__ int merge_split(PA A, int la, PB B, int lb, int d) {
  int lo = max(0, d - lb), hi = min(d, la);
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (A[mid] <= B[d - 1 - mid]) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// Merge-path walk of A[0..la) (LDS or HBM) and B[0..lb) (HBM): every
// element x of A is visited once with found = (x in B); the lanes' counts of
// keep(x, found) are summed over the wave, and with out != null the kept
// elements are written in increasing order to out[0..).  As / Bs: this
// wave's LDS tile buffers of kMergeTile and kMergeTile + 1 ints, split: 64
// ints of LDS.  All 64 lanes call; the result is the same on every lane.
template <class PA, class Keep>
__device__ uint32_t merge_walk(PA A, int la, const int32_t *__restrict__ B, int lb, int *As, int *Bs, int *split,
                               int lane, Keep keep, int32_t *__restrict__ out) {
  const int L = la + lb;
  uint32_t total = 0;
  int ia0 = 0, ib0 = 0;  // split at the start of the current tile
  for (int t0 = 0; t0 < L; t0 += 64 * kMergeTile) {
    // tile ends of the next 64 tiles, one merge-path search per lane
    wave_sync();
    {
      const int d = (int)min((int64_t)L, (int64_t)t0 + (int64_t)(lane + 1) * kMergeTile);
      split[lane] = merge_split(A, la, B, lb, d);
    }
    wave_sync();
    const int ntiles = min(64, (L - t0 + kMergeTile - 1) / kMergeTile);
    for (int q = 0; q < ntiles; ++q) {
      const int d1 = (int)min((int64_t)L, (int64_t)t0 + (int64_t)(q + 1) * kMergeTile);
      const int ia1 = split[q], ib1 = d1 - ia1;
      const int na = ia1 - ia0, nb = ib1 - ib0;
      wave_sync();  // the previous tile's reads of As / Bs are done
      for (int i = lane; i < na; i += 64) As[i] = A[ia0 + i];
      // B one past the tile: the partner of the tile's last A element may open the next tile
      for (int i = lane; i <= nb; i += 64) Bs[i] = (ib0 + i < lb) ? B[ib0 + i] : INT32_MAX;
      wave_sync();
      const int dl = min(lane * kMergeE, na + nb);
      const int ia_s = merge_split(As, na, Bs, nb, dl);
      const int steps = min(kMergeE, na + nb - dl);
      // pass 1: count this lane's kept elements
      uint32_t c = 0;
      {
        int ia = ia_s, ib = dl - ia_s;
#pragma unroll
        for (int st = 0; st < kMergeE; ++st) {
          if (st < steps) {
            if (ia < na && (ib >= nb || As[ia] <= Bs[ib])) {
              const int x = As[ia];
              c += keep(x, Bs[ib] == x) ? 1u : 0u;
              ++ia;
            } else {
              ++ib;
            }
          }
        }
      }
      const uint32_t incl = (uint32_t)wave_incl_scan((int)c, lane);
      if (out && c) {
        // pass 2: write them at this lane's place in the wave's order
        uint32_t pos = total + incl - c;
        int ia = ia_s, ib = dl - ia_s;
#pragma unroll
        for (int st = 0; st < kMergeE; ++st) {
          if (st < steps) {
            if (ia < na && (ib >= nb || As[ia] <= Bs[ib])) {
              const int x = As[ia];
              if (keep(x, Bs[ib] == x)) out[pos++] = x;
              ++ia;
            } else {
              ++ib;
            }
          }
        }
      }
      total += (uint32_t)__shfl((int)incl, 63, 64);
      ia0 = ia1;
      ib0 = ib1;
    }
  }
  wave_sync();
  return total;
}

