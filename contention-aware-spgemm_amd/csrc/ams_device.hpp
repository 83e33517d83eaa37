// ams_device.hpp -- device building blocks shared by the SpGEMM kernels.
//
// Wave64 scans and reductions, the block-level product expansion that every
// row kernel uses, LDS hash probes and an LDS bitonic sort.  gfx950 only.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ams {

constexpr int kWave = 64;
constexpr uint32_t kHashMul = 2654435761u;  // Knuth's multiplier, the reference's HASH_CONSTANT
constexpr int kEmpty = -1;

__device__ __forceinline__ int lane_id() { return threadIdx.x & (kWave - 1); }
__device__ __forceinline__ int wave_id() { return threadIdx.x >> 6; }

template <typename I>
__device__ __forceinline__ I wave_incl_scan(I x) {
  const int lane = lane_id();
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    I y = __shfl_up(x, d, kWave);
    if (lane >= d) x += y;
  }
  return x;
}

template <typename I>
__device__ __forceinline__ I wave_sum(I x) {
#pragma unroll
  for (int d = kWave / 2; d > 0; d >>= 1) x += __shfl_xor(x, d, kWave);
  return x;
}

// Exclusive scan over the T threads of a block.  s_wave holds T/64 values.
// Contains two barriers; every thread must call it.
template <int T, typename I>
__device__ __forceinline__ I block_excl_scan(I x, I* s_wave, I& total) {
  static_assert(T % kWave == 0 && T <= 1024, "block size");
  const I incl = wave_incl_scan(x);
  if (lane_id() == kWave - 1) s_wave[wave_id()] = incl;
  __syncthreads();
  I base = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < T / kWave; ++w) {
    const I v = s_wave[w];
    base += (w < wave_id()) ? v : I(0);
    tot += v;
  }
  __syncthreads();
  total = tot;
  return base + incl - x;
}

template <int T, typename I>
__device__ __forceinline__ I block_sum(I x, I* s_wave) {
  I total;
  (void)block_excl_scan<T, I>(x, s_wave, total);
  return total;
}

// LDS scratch of one block-level expansion window.
template <int T>
struct ExpandLds {
  int32_t off[T];     // exclusive prefix of segment lengths in the window
  int64_t start[T];   // first B position of each segment
  int64_t wtmp[T / kWave];
};

// Enumerates the products of A nonzeros [p0, p1): nonzero p contributes the
// B positions [start(p), start(p) + len(p)).  Windows of T nonzeros are
// flattened with a block scan; each thread then takes products t = tid,
// tid + T, ... of the window and finds its nonzero by binary search over the
// window's prefix.  Consecutive products of one segment go to consecutive
// lanes, so B reads coalesce whatever the segment lengths are.
//   seg(p, start&, len&)          fills the segment of nonzero p (p < p1);
//                                 it runs on thread p - window base
//   body(j, p, q, t)              j = index in window, q = B position,
//                                 t = rank of the product in A-row order
// All threads of the block must call it (it holds barriers).
template <int T, class Seg, class Body>
__device__ __forceinline__ void block_expand(int64_t p0, int64_t p1, ExpandLds<T>& L, Seg seg,
                                             Body body) {
  int64_t base = 0;
  for (int64_t w0 = p0; w0 < p1; w0 += T) {
    const int64_t p = w0 + threadIdx.x;
    int64_t st = 0;
    int32_t len = 0;
    if (p < p1) seg(p, st, len);
    int64_t total64;
    const int64_t ex = block_excl_scan<T, int64_t>((int64_t)len, L.wtmp, total64);
    L.off[threadIdx.x] = (int32_t)ex;
    L.start[threadIdx.x] = st;
    __syncthreads();
    const int32_t total = (int32_t)total64;
    const int nvalid = (int)((p1 - w0) < T ? (p1 - w0) : T);
    for (int32_t t = threadIdx.x; t < total; t += T) {
      int lo = 0, hi = nvalid;
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (L.off[mid] <= t) lo = mid; else hi = mid;
      }
      body(lo, w0 + lo, L.start[lo] + (int64_t)(t - L.off[lo]), base + t);
    }
    base += total;
    __syncthreads();
  }
}

__device__ __forceinline__ uint32_t hash_slot(int32_t key, int shift) {
  return ((uint32_t)key * kHashMul) >> shift;
}

// Key-only insert into an LDS linear-probing table of 2^(32-shift) slots.
// Returns 1 when this call placed the key, 0 when it was present, -1 when
// the table is full (a sizing bug: the caller raises an error flag).
__device__ __forceinline__ int lds_insert_key(int32_t* keys, int shift, int32_t key) {
  const uint32_t mask = (0xffffffffu >> shift);
  uint32_t h = hash_slot(key, shift);
  for (uint32_t probe = 0; probe <= mask; ++probe) {
    const int32_t k = __hip_atomic_load(&keys[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (k == key) return 0;
    if (k == kEmpty) {
      const int32_t old = atomicCAS(&keys[h], kEmpty, key);
      if (old == kEmpty) return 1;
      if (old == key) return 0;
    }
    h = (h + 1) & mask;
  }
  return -1;
}

// Insert-or-accumulate into an LDS table of keys + values (values start 0).
// Returns false when the table is full.
template <typename V>
__device__ __forceinline__ bool lds_insert_add(int32_t* keys, V* vals, int shift, int32_t key, V v) {
  const uint32_t mask = (0xffffffffu >> shift);
  uint32_t h = hash_slot(key, shift);
  for (uint32_t probe = 0; probe <= mask; ++probe) {
    const int32_t k = __hip_atomic_load(&keys[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (k != key && k == kEmpty) {
      const int32_t old = atomicCAS(&keys[h], kEmpty, key);
      if (old != kEmpty && old != key) { h = (h + 1) & mask; continue; }
    } else if (k != key) {
      h = (h + 1) & mask;
      continue;
    }
    atomicAdd(&vals[h], v);
    return true;
  }
  return false;
}

// Lookup of a key that should be present; returns the table size when it
// is not (callers treat that as an error).
__device__ __forceinline__ uint32_t lds_find(const int32_t* keys, int shift, int32_t key) {
  const uint32_t mask = (0xffffffffu >> shift);
  uint32_t h = hash_slot(key, shift);
  for (uint32_t probe = 0; probe <= mask; ++probe) {
    if (keys[h] == key) return h;
    h = (h + 1) & mask;
  }
  return mask + 1;
}

// Ascending bitonic sort of n (power of two) keys in LDS by the T threads
// of the block.  Ends with a barrier.
template <int T, typename K>
__device__ __forceinline__ void block_bitonic_sort(K* s, int n) {
  for (int k = 2; k <= n; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < (n >> 1); i += T) {
        const int a = ((i & ~(j - 1)) << 1) | (i & (j - 1));
        const int b = a + j;
        const K x = s[a], y = s[b];
        const bool up = (a & k) == 0;
        if ((x > y) == up) { s[a] = y; s[b] = x; }
      }
      __syncthreads();
    }
  }
}

}  // namespace ams
