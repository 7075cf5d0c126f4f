// Wave64 building blocks shared by every gfx950 kernel in libxgm:
// wave-scope synchronisation, reductions and scans over 64 lanes, and the
// two exact membership structures used on staged neighbour lists (an
// open-addressing hash and a branch-free binary search).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

typedef unsigned long long u64;

#define XGM_WAVE 64

// Order LDS traffic between lanes of one wave (no s_barrier needed).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

__device__ __forceinline__ u64 lanemask_lt(int lane) { return (1ull << lane) - 1ull; }
__device__ __forceinline__ u64 lanemask_le(int lane) { return lane == 63 ? ~0ull : ((2ull << lane) - 1ull); }

__device__ __forceinline__ u64 wave_sum_u64(u64 v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ int wave_sum_i32(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ u64 wave_or_u64(u64 v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v |= __shfl_xor(v, o, 64);
  return v;
}

// Inclusive prefix sum over the 64 lanes.
__device__ __forceinline__ int wave_incl_scan(int v, int lane) {
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int t = __shfl_up(v, o, 64);
    if (lane >= o) v += t;
  }
  return v;
}

// First index in a[0..n) with a[idx] > key (a non-decreasing).  All lanes
// usually pass the same (a, n, key) so the loads broadcast.
template <class P>
__device__ __forceinline__ int upper_bound_dev(P a, int n, int key) {
  int lo = 0, hi = n;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (a[mid] <= key) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// First index in a[0..n) with a[idx] >= key.
template <class P>
__device__ __forceinline__ int lower_bound_dev(P a, int n, int key) {
  int lo = 0, hi = n;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (a[mid] < key) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// Membership in a strictly increasing list; fixed trip count per list size.
template <class P>
__device__ __forceinline__ bool sorted_contains(P a, int n, int key) {
  if (n <= 0) return false;
  int pos = 0;
  int step = 1 << (31 - __clz(n));  // largest power of two <= n
  for (; step > 0; step >>= 1) {
    int nxt = pos + step;
    if (nxt <= n && a[nxt - 1] < key) pos = nxt;
  }
  return pos < n && a[pos] == key;
}

// Open-addressing hash over int32 keys >= 0; empty slot = -1; table size
// 2^bits with bits chosen so the load factor stays <= 1/2.
__device__ __forceinline__ uint32_t hash_slot(int key, int bits) {
  return ((uint32_t)key * 0x9E3779B1u) >> (32 - bits);
}

__device__ __forceinline__ void hash_insert(int *tab, int bits, int key) {
  const uint32_t mask = (1u << bits) - 1u;
  uint32_t h = hash_slot(key, bits);
  while (atomicCAS(&tab[h], -1, key) != -1) h = (h + 1) & mask;
}

__device__ __forceinline__ bool hash_contains(const int *tab, int bits, int key) {
  const uint32_t mask = (1u << bits) - 1u;
  uint32_t h = hash_slot(key, bits);
  for (;;) {
    int k = tab[h];
    if (k == key) return true;
    if (k < 0) return false;
    h = (h + 1) & mask;
  }
}

// Bucketed variant: 2^bbits buckets of 4 slots (int4, 16-byte aligned),
// empty slot = -1.  A key goes into the first empty slot of its home bucket
// or, when that bucket is full, of the next buckets in turn; slots of a
// bucket fill in order, so a bucket whose last slot is empty ends a lookup.
// A lookup reads one 16-byte bucket (one LDS access) unless the bucket is
// full: at load <= 1/4 that is rare, so the probe loop is almost never
// taken and a wave's lanes stay converged.
__device__ __forceinline__ void bhash_insert(int *tab, int bbits, int key) {
  const uint32_t mask = (1u << bbits) - 1u;
  uint32_t h = hash_slot(key, bbits);
  for (;;) {
    int *b = tab + 4 * h;
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (atomicCAS(&b[q], -1, key) == -1) return;
    h = (h + 1) & mask;
  }
}

// Minimum of the four XORs of key with bucket q: zero iff the bucket holds
// key (straight-line integer ops, no per-slot branches or lane masks).
__device__ __forceinline__ uint32_t bucket_dist(const int4 &q, int key) {
  const uint32_t k = (uint32_t)key;
  return min(min((uint32_t)q.x ^ k, (uint32_t)q.y ^ k), min((uint32_t)q.z ^ k, (uint32_t)q.w ^ k));
}

__device__ __forceinline__ bool bucket_has(const int4 &q, int key) { return bucket_dist(q, key) == 0u; }

// Probe key of a walker's invalid lanes where the probed set is a bucketed
// hash: ids are >= 0 and an empty slot is -1, so no bucket holds it and a
// match implies a valid lane.
constexpr int kNoKey = -2;

// Rest of a lookup whose home bucket h was full and lacked the key.
__device__ __noinline__ bool bhash_contains_after(const int *tab, int bbits, uint32_t h, int key) {
  const uint32_t mask = (1u << bbits) - 1u;
  for (;;) {
    h = (h + 1) & mask;
    const int4 q = reinterpret_cast<const int4 *>(tab)[h];
    if (bucket_has(q, key)) return true;
    if (q.w < 0) return false;
  }
}

__device__ __forceinline__ bool bhash_contains(const int *tab, int bbits, int key) {
  const uint32_t mask = (1u << bbits) - 1u;
  uint32_t h = hash_slot(key, bbits);
  int4 q = reinterpret_cast<const int4 *>(tab)[h];
  bool hit = bucket_has(q, key);
  bool more = !hit && q.w >= 0;
  while (__builtin_expect(more, 0)) {
    h = (h + 1) & mask;
    q = reinterpret_cast<const int4 *>(tab)[h];
    hit = bucket_has(q, key);
    more = !hit && q.w >= 0;
  }
  return hit;
}

// Table bits for n keys: smallest b >= 6 with 2^b >= n << shift (load
// factor <= 2^-shift; misses probe ~1.4 slots at 1/4, ~2.5 at 1/2).
__device__ __forceinline__ int hash_bits_for(int n, int shift = 1) {
  int b = 6;
  while ((1 << b) < (n << shift)) ++b;
  return b;
}

// Add a task's per-lane partial count.  Without slots it joins the
// kernel-wide accumulator (flushed once at the end); with slots (batched
// instances) the wave reduces it and adds it to out[slots[v]].  Every lane
// of the wave must call.
__device__ __forceinline__ void task_flush(u64 &acc, u64 part, const int32_t *__restrict__ slots, int v,
                                           u64 *__restrict__ out, int lane) {
  if (!slots) {
    acc += part;
    return;
  }
  part = wave_sum_u64(part);
  if (lane == 0 && part) atomicAdd(out + slots[v], part);
}
