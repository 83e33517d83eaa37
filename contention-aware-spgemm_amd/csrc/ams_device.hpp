// ams_device.hpp -- device building blocks shared by the SpGEMM kernels.
//
// Wave64 scans and reductions, the block-level product expansion that every
// row kernel uses, LDS hash probes and an LDS bitonic sort.  gfx950 only.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>

namespace ams {

constexpr int kWave = 64;
constexpr uint32_t kHashMul = 2654435761u;  // Knuth's multiplier, the reference's HASH_CONSTANT
constexpr int kEmpty = -1;

__device__ __forceinline__ int lane_id() { return threadIdx.x & (kWave - 1); }
__device__ __forceinline__ int wave_id() { return threadIdx.x >> 6; }

// A value that is the same on every lane, moved to a scalar register (so
// that loads addressed by it become scalar loads).
__device__ __forceinline__ unsigned long long uniform_u64(unsigned long long v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return ((unsigned long long)hi << 32) | lo;
}

// Store of write-once output (C entries stored in coalesced runs by the
// dense chunks, ESC, the LDS-hash numeric and the light kernel's values;
// the kept row bitmaps): with
// AMS_NT_STORES set it is a non-temporal store, which streams
// past the caches instead of evicting the B blocks and columns the kernels
// re-read.  Only for runs of consecutive entries: with it the light kernel's
// per-bit column stores (one entry a lane, scattered) took 20 -> 57 ms on
// G500 s19 ef32.
#ifndef AMS_NT_STORES
#define AMS_NT_STORES 1
#endif
template <class T>
struct same_type {
  using type = T;
};
template <class T>
__device__ __forceinline__ void stream_store(T* p, typename same_type<T>::type v) {
#if AMS_NT_STORES
  __builtin_nontemporal_store(v, p);
#else
  *p = v;
#endif
}

// DPP lane move of a wave64 (GFX9 controls): row_shr:n reads lane i - n of
// the same 16-lane row, row_bcast:15 / row_bcast:31 read the last lane of
// the previous row / of row 1 into the rows of kRowMask; lanes without a
// source (or outside kRowMask) get 0.
template <int kCtrl, int kRowMask>
__device__ __forceinline__ uint32_t dpp_move(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, kCtrl, kRowMask, 0xf, false);
}

template <int kCtrl, int kRowMask, typename I>
__device__ __forceinline__ I dpp_move_int(I v) {
  if constexpr (sizeof(I) == 8) {
    const uint64_t u = (uint64_t)v;
    const uint32_t lo = dpp_move<kCtrl, kRowMask>((uint32_t)u);
    const uint32_t hi = dpp_move<kCtrl, kRowMask>((uint32_t)(u >> 32));
    return (I)(((uint64_t)hi << 32) | lo);
  } else {
    return (I)dpp_move<kCtrl, kRowMask>((uint32_t)v);
  }
}

// Inclusive prefix sum over the 64 lanes (integers): Hillis-Steele inside
// each row of 16 lanes, then the row totals carried forward by broadcasts.
// No LDS traffic.
template <typename I>
__device__ __forceinline__ I wave_incl_scan(I x) {
  static_assert(std::is_integral<I>::value && (sizeof(I) == 4 || sizeof(I) == 8), "32/64-bit integers");
  x += dpp_move_int<0x111, 0xf>(x);  // row_shr:1
  x += dpp_move_int<0x112, 0xf>(x);  // row_shr:2
  x += dpp_move_int<0x114, 0xf>(x);  // row_shr:4
  x += dpp_move_int<0x118, 0xf>(x);  // row_shr:8
  x += dpp_move_int<0x142, 0xa>(x);  // row_bcast:15 into rows 1 and 3
  x += dpp_move_int<0x143, 0xc>(x);  // row_bcast:31 into rows 2 and 3
  return x;
}

// Value of lane 63, in a scalar register.
template <typename I>
__device__ __forceinline__ I wave_last(I x) {
  if constexpr (sizeof(I) == 8) {
    const uint64_t u = (uint64_t)x;
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, kWave - 1);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), kWave - 1);
    return (I)(((uint64_t)hi << 32) | lo);
  } else {
    return (I)__builtin_amdgcn_readlane((int)x, kWave - 1);
  }
}

template <typename I>
__device__ __forceinline__ I wave_sum(I x) {
  if constexpr (std::is_integral<I>::value) {
    return wave_last(wave_incl_scan(x));
  } else {
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) x += __shfl_xor(x, d, kWave);
    return x;
  }
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

// block_excl_scan without its closing barrier, for scratch that is not
// written again before every thread has passed another barrier.
template <int T, typename I>
__device__ __forceinline__ I block_excl_scan_once(I x, I* s_wave, I& total) {
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
  total = tot;
  return base + incl - x;
}

// Inclusive prefix sum over lanes 0 .. T/64 - 1 (one 16-lane DPP row):
// combines one value per wave of a T-thread block.
template <int T>
__device__ __forceinline__ uint32_t wave_row_scan(uint32_t v) {
  static_assert(T % kWave == 0 && T / kWave <= 16, "one lane per wave, inside one DPP row");
  if constexpr (T / kWave > 1) v += dpp_move<0x111, 0xf>(v);  // row_shr:1
  if constexpr (T / kWave > 2) v += dpp_move<0x112, 0xf>(v);  // row_shr:2
  if constexpr (T / kWave > 4) v += dpp_move<0x114, 0xf>(v);  // row_shr:4
  if constexpr (T / kWave > 8) v += dpp_move<0x118, 0xf>(v);  // row_shr:8
  return v;
}

// Exclusive prefix (at the calling wave) and total of the per-wave values
// s[0 .. T/64) that every wave wrote before a barrier: lane w reads wave w's
// value, so the VALU work does not grow with the wave count.
template <int T>
__device__ __forceinline__ uint32_t block_wave_prefix(const uint32_t* s, uint32_t& total) {
  const uint32_t v = lane_id() < T / kWave ? s[lane_id()] : 0u;
  const uint32_t incl = wave_row_scan<T>(v);
  total = (uint32_t)__builtin_amdgcn_readlane((int)incl, T / kWave - 1);
  const int wv = __builtin_amdgcn_readfirstlane(wave_id());
  return (uint32_t)__builtin_amdgcn_readlane((int)(incl - v), wv);
}

// Segment scan of a window of T segment lengths (one per thread, each >= 0):
// off = products before this segment, ci = non-empty segments before it
// (its compact index), total = products of the window (must fit 31 bits),
// nonempty = non-empty segments of the window.
// A 32-bit DPP scan and a ballot per wave, then one barrier and a per-lane
// combine of the waves' (products, non-empty) pairs in s_wave[T/64].  No
// closing barrier: s_wave must not be written again before every thread has
// passed another barrier.
template <int T>
__device__ __forceinline__ void block_seg_scan(int32_t len, uint2* s_wave, int32_t& off, int32_t& ci,
                                               int32_t& total, int32_t& nonempty) {
  const uint32_t incl = (uint32_t)wave_incl_scan(len);
  const unsigned long long ne = __ballot(len > 0);
  const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(ne >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)ne, 0u));
  if (lane_id() == kWave - 1) s_wave[wave_id()] = make_uint2(incl, (uint32_t)__popcll(ne));
  __syncthreads();
  const uint2 v = lane_id() < T / kWave ? s_wave[lane_id()] : make_uint2(0u, 0u);
  const uint32_t ia = wave_row_scan<T>(v.x), ib = wave_row_scan<T>(v.y);
  const int wv = __builtin_amdgcn_readfirstlane(wave_id());
  total = __builtin_amdgcn_readlane((int)ia, T / kWave - 1);
  nonempty = __builtin_amdgcn_readlane((int)ib, T / kWave - 1);
  off = __builtin_amdgcn_readlane((int)(ia - v.x), wv) + (int32_t)incl - len;
  ci = __builtin_amdgcn_readlane((int)(ib - v.y), wv) + (int32_t)below;
}

template <int T, typename I>
__device__ __forceinline__ I block_sum(I x, I* s_wave) {
  I total;
  (void)block_excl_scan<T, I>(x, s_wave, total);
  return total;
}

// LDS scratch of one block-level expansion window of NW <= T nonzeros.
template <int T, int NW = T>
struct ExpandLds {
  static_assert(NW <= T && NW % kWave == 0, "window of whole waves, at most the block");
  int32_t off[NW];    // exclusive prefix of segment lengths in the window
  int64_t start[NW];  // first B position of each segment
  int64_t wtmp[T / kWave];
};

// Largest j in [cur, n) with off[j] <= t, given off[cur] <= t and off
// non-decreasing: gallop forward from cur, then bisect.  One LDS read when
// t is still inside segment cur, which is the common case on long B rows.
__device__ __forceinline__ int seek_segment(const int32_t* off, int n, int cur, int32_t t) {
  int lo = cur, hi = n, step = 1;
  for (;;) {
    const int nx = lo + step;
    if (nx >= n) break;
    if (off[nx] > t) { hi = nx; break; }
    lo = nx;
    step <<= 1;
  }
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (off[mid] <= t) lo = mid; else hi = mid;
  }
  return lo;
}

// Moves a lane's segment cursor forward to the segment holding product t
// (t >= seg_end, t < total) of a window of n segments with exclusive offsets
// off[] and first B positions start[]; afterwards B position = seg_bias + t
// for t < seg_end.  Start with cur = -1, seg_end = 0 (off[0] = 0).  Mostly
// the next segment, which starts at seg_end: its end and start position are
// two independent reads; past it (an empty or short segment), gallop from
// the one after, whose offset is <= t.
__device__ __forceinline__ void segment_advance(const int32_t* off, const int64_t* start, int n, int32_t total,
                                                int32_t t, int& cur, int32_t& seg_end, int64_t& seg_bias) {
  const int nx = cur + 1;
  const int32_t nend = nx + 1 < n ? off[nx + 1] : total;
  const int64_t nst = start[nx];
  if (t < nend) {
    cur = nx;
    seg_bias = nst - (int64_t)seg_end;
    seg_end = nend;
  } else {
    cur = seek_segment(off, n, nx + 1, t);
    seg_end = cur + 1 < n ? off[cur + 1] : total;
    seg_bias = start[cur] - (int64_t)off[cur];
  }
}

// Enumerates the products of A nonzeros [p0, p1): nonzero p contributes the
// B positions [start(p), start(p) + len(p)).  Windows of NW nonzeros (the
// block size T unless the kernel saves LDS with fewer) are flattened with a
// block scan.  In every T*U-wide stride of the window,
// wave w owns the 64*U consecutive products starting at w*64*U and lane l
// takes products l, l+64, ... of them, so consecutive products of one B row
// go to consecutive lanes and B reads coalesce whatever the segment lengths
// are, while each lane's cursor moves only 64 products per step.  Each
// thread keeps the segment of its last product and seeks forward from it.
// The U products' global reads (load) are issued before their LDS updates
// (apply), giving every lane U independent loads in flight.
//   seg(p, start&, len&)     segment of nonzero p (p < p1), on thread p - w0 < NW
//   load(j, p, q, t) -> item j = index in window, q = B position,
//                            t = rank of the product in A-row order; it must
//                            only gather (no arithmetic on loaded values), so
//                            that the U loads stay in flight together
//   apply(item)              runs for the lanes whose product exists
// All threads of the block must call it (it holds barriers).
// block_expand_batch hands each thread's U items of a stride to
// apply_batch(item[U], ok[U]) at once, so that it can issue its LDS reads
// for all of them before it uses the first.
template <int T, int U, int NW, class Seg, class Load, class ApplyBatch>
__device__ __forceinline__ void block_expand_batch(int64_t p0, int64_t p1, ExpandLds<T, NW>& L, Seg seg, Load load,
                                                   ApplyBatch apply_batch) {
  using Item = decltype(load(0, int64_t(0), int64_t(0), int64_t(0)));
  int64_t base = 0;
  const bool loader = NW == T || (int)threadIdx.x < NW;
  for (int64_t w0 = p0; w0 < p1; w0 += NW) {
    const int64_t p = w0 + threadIdx.x;
    int64_t st = 0;
    int32_t len = 0;
    if (loader && p < p1) seg(p, st, len);
    // 32-bit segment scan (a window's products fit 31 bits): one barrier;
    // L.wtmp is next written after this window's closing barrier
    int32_t ex, ci, total, nne;
    block_seg_scan<T>(len, (uint2*)L.wtmp, ex, ci, total, nne);
    if (loader) {
      L.off[threadIdx.x] = ex;
      L.start[threadIdx.x] = st;
    }
    __syncthreads();
    const int nvalid = (int)((p1 - w0) < NW ? (p1 - w0) : NW);
    // the lane's current segment stays in registers: [.., seg_end) and
    // B position = seg_bias + t; LDS is searched only when t leaves it.
    // Before the first product: segment -1, which ends at off[0] = 0
    int cur = -1;
    int32_t seg_end = 0;
    int64_t seg_bias = 0;
    for (int32_t t0 = 0; t0 < total; t0 += T * U) {
      Item item[U];
      bool ok[U];
      // branch-free issue: lanes past the end re-load the last product and
      // are masked in apply, so all U loads go out back to back
#pragma unroll
      for (int u = 0; u < U; ++u) {
        int32_t t = t0 + wave_id() * (kWave * U) + u * kWave + lane_id();
        ok[u] = t < total;
        t = ok[u] ? t : total - 1;
        if (t >= seg_end) segment_advance(L.off, L.start, nvalid, total, t, cur, seg_end, seg_bias);
        item[u] = load(cur, w0 + cur, seg_bias + t, base + t);
      }
      apply_batch(item, ok);
    }
    base += total;
    __syncthreads();
  }
}

template <int T, int U, int NW, class Seg, class Load, class Apply>
__device__ __forceinline__ void block_expand(int64_t p0, int64_t p1, ExpandLds<T, NW>& L, Seg seg, Load load,
                                             Apply apply) {
  using Item = decltype(load(0, int64_t(0), int64_t(0), int64_t(0)));
  block_expand_batch<T, U>(p0, p1, L, seg, load, [&](Item (&item)[U], bool (&ok)[U]) {
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (ok[u]) apply(item[u]);
  });
}

// Reads an LDS accumulator and leaves 0 in it, in one exchange.
template <typename V>
__device__ __forceinline__ V lds_take(V* p) {
  static_assert(sizeof(V) == 4 || sizeof(V) == 8, "32- or 64-bit values");
  if constexpr (sizeof(V) == 8) {
    return __builtin_bit_cast(V, atomicExch((unsigned long long*)p, 0ull));
  } else {
    return __builtin_bit_cast(V, atomicExch((unsigned int*)p, 0u));
  }
}

// Makes this wave's LDS writes visible to its other lanes.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// block_expand for one wave working alone: windows of 64 nonzeros, wave
// scan by shuffles, per-wave LDS prefix (w_off[64], w_start[64]); no block
// barriers, so the waves of a block can work on different rows.
template <int U, class Seg, class Load, class Apply>
__device__ __forceinline__ void wave_expand(int64_t p0, int64_t p1, int32_t* w_off, int64_t* w_start, Seg seg,
                                            Load load, Apply apply) {
  using Item = decltype(load(0, int64_t(0), int64_t(0), int64_t(0)));
  const int lane = lane_id();
  int64_t base = 0;
  for (int64_t w0 = p0; w0 < p1; w0 += kWave) {
    const int64_t p = w0 + lane;
    int64_t st = 0;
    int32_t len = 0;
    if (p < p1) seg(p, st, len);
    const int32_t incl = wave_incl_scan(len);
    const int32_t total = wave_last(incl);
    w_off[lane] = incl - len;
    w_start[lane] = st;
    wave_lds_sync();
    const int nvalid = (int)((p1 - w0) < kWave ? (p1 - w0) : kWave);
    int cur = -1;
    int32_t seg_end = 0;
    int64_t seg_bias = 0;
    for (int32_t t0 = 0; t0 < total; t0 += kWave * U) {
      Item item[U];
      bool ok[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        int32_t t = t0 + u * kWave + lane;
        ok[u] = t < total;
        t = ok[u] ? t : total - 1;
        if (t >= seg_end) segment_advance(w_off, w_start, nvalid, total, t, cur, seg_end, seg_bias);
        item[u] = load(cur, w0 + cur, seg_bias + t, base + t);
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (ok[u]) apply(item[u]);
    }
    base += total;
    wave_lds_sync();
  }
}

__device__ __forceinline__ uint32_t hash_slot(int32_t key, int shift) {
  return ((uint32_t)key * kHashMul) >> shift;
}

// Key-only insert into an LDS linear-probing table of 2^(32-shift) slots.
// One compare-and-swap per probe: its old value tells empty (placed), equal
// (present) or another key (probe on).  Returns 1 when this call placed the
// key, 0 when it was present, -1 when the table is full (a sizing bug: the
// caller raises an error flag).
__device__ __forceinline__ int lds_insert_key(int32_t* keys, int shift, int32_t key) {
  const uint32_t mask = (0xffffffffu >> shift);
  uint32_t h = hash_slot(key, shift);
  for (uint32_t probe = 0; probe <= mask; ++probe) {
    const int32_t old = atomicCAS(&keys[h], kEmpty, key);
    if (old == kEmpty) return 1;
    if (old == key) return 0;
    h = (h + 1) & mask;
  }
  return -1;
}

// Adds v into an LDS accumulator.  FP64 takes ds_add_f64.  FP32 takes a
// 32-bit compare-and-swap loop over an FP32 add: on gfx950 ds_add_f32 runs
// at about 200 G lane operations a second against 1.5 T for this loop (and
// for ds_add_f64) with a few colliding lanes, and 340 G against 200 G with
// 64 hot slots.  Each add is one rounded FP32 add either way; the order of
// concurrent adds is unspecified in both.
template <typename V>
__device__ __forceinline__ void lds_add(V* p, V v) {
  if constexpr (sizeof(V) == 8) {
    atomicAdd(p, v);
  } else {
    unsigned* u = (unsigned*)p;
    unsigned old = *u;
    for (;;) {
      const unsigned assumed = old;
      old = atomicCAS(u, assumed, __float_as_uint(__uint_as_float(assumed) + v));
      if (old == assumed) break;
    }
  }
}

// G adds into LDS accumulators base[idx[j]] += v[j].  FP32: every slot's
// read, then every slot's compare-and-swap, go out back to back, and only
// the slots whose swap met another add loop again.
template <int G, typename V>
__device__ __forceinline__ void lds_add_group(V* base, const uint32_t (&idx)[G], const V (&v)[G]) {
  if constexpr (sizeof(V) == 8) {
#pragma unroll
    for (int j = 0; j < G; ++j) atomicAdd(&base[idx[j]], v[j]);
  } else {
    unsigned* u = (unsigned*)base;
    unsigned old[G];
#pragma unroll
    for (int j = 0; j < G; ++j) old[j] = u[idx[j]];
    bool miss[G];
#pragma unroll
    for (int j = 0; j < G; ++j) {
      const unsigned got = atomicCAS(&u[idx[j]], old[j], __float_as_uint(__uint_as_float(old[j]) + v[j]));
      miss[j] = got != old[j];
      old[j] = got;
    }
#pragma unroll
    for (int j = 0; j < G; ++j) {
      while (miss[j]) {
        const unsigned assumed = old[j];
        old[j] = atomicCAS(&u[idx[j]], assumed, __float_as_uint(__uint_as_float(assumed) + v[j]));
        miss[j] = old[j] != assumed;
      }
    }
  }
}

// Insert-or-accumulate into an LDS table of keys + values (values start 0).
// Returns false when the table is full.
template <typename V>
__device__ __forceinline__ bool lds_insert_add(int32_t* keys, V* vals, int shift, int32_t key, V v) {
  const uint32_t mask = (0xffffffffu >> shift);
  uint32_t h = hash_slot(key, shift);
  for (uint32_t probe = 0; probe <= mask; ++probe) {
    const int32_t old = atomicCAS(&keys[h], kEmpty, key);
    if (old == kEmpty || old == key) {
      lds_add(&vals[h], v);
      return true;
    }
    h = (h + 1) & mask;
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

// Ascending bitonic sort of n (power of two, n <= 2*T*MAXP) keys in LDS by
// the T threads of the block.  Each thread loads all of its MAXP pairs of a
// stage before comparing and storing any, so the LDS reads of a stage are in
// flight together.  Ends with a barrier.
template <int T, int MAXP, typename K>
__device__ __forceinline__ void block_bitonic_sort(K* s, int n) {
  const int half = n >> 1;
  for (int k = 2; k <= n; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      K x[MAXP], y[MAXP];
      int a[MAXP];
#pragma unroll
      for (int p = 0; p < MAXP; ++p) {
        const int i = threadIdx.x + p * T;
        a[p] = ((i & ~(j - 1)) << 1) | (i & (j - 1));
        if (i < half) {
          x[p] = s[a[p]];
          y[p] = s[a[p] + j];
        }
      }
#pragma unroll
      for (int p = 0; p < MAXP; ++p) {
        const int i = threadIdx.x + p * T;
        if (i < half) {
          const bool up = (a[p] & k) == 0;
          if ((x[p] > y[p]) == up) {
            s[a[p]] = y[p];
            s[a[p] + j] = x[p];
          }
        }
      }
      __syncthreads();
    }
  }
}

// Lane i ^ M's value (M < 64, a power of two).  M = 1, 2 and 8 are one DPP
// move (quad_perm [1,0,3,2], quad_perm [2,3,0,1], row_ror:8), M = 4 two DPP
// row shifts and a select; only 16 and 32 cross rows through ds_bpermute.
template <int M>
__device__ __forceinline__ uint32_t lane_xor(uint32_t x) {
  if constexpr (M == 1) {
    return dpp_move<0xb1, 0xf>(x);
  } else if constexpr (M == 2) {
    return dpp_move<0x4e, 0xf>(x);
  } else if constexpr (M == 4) {
    const uint32_t up = dpp_move<0x104, 0xf>(x);    // row_shl:4, lane i + 4
    const uint32_t down = dpp_move<0x114, 0xf>(x);  // row_shr:4, lane i - 4
    return (lane_id() & 4) ? down : up;
  } else if constexpr (M == 8) {
    return dpp_move<0x128, 0xf>(x);
  } else {
    return (uint32_t)__shfl_xor((int)x, M, kWave);
  }
}

template <int M, typename K>
__device__ __forceinline__ K lane_xor_key(K x) {
  if constexpr (sizeof(K) == 8) {
    const uint32_t lo = lane_xor<M>((uint32_t)x);
    const uint32_t hi = lane_xor<M>((uint32_t)(x >> 32));
    return (K)(((uint64_t)hi << 32) | lo);
  } else {
    return (K)lane_xor<M>((uint32_t)x);
  }
}

// Compare-exchange of every register with lane tid ^ M (stride j = M*IPT,
// merge size k): all exchanges of the stage go out before any compare.
template <int M, int IPT, typename K>
__device__ __forceinline__ void lane_cx_stage(K (&v)[IPT], int base, int k, int j) {
  K o[IPT];
#pragma unroll
  for (int r = 0; r < IPT; ++r) o[r] = lane_xor_key<M>(v[r]);
#pragma unroll
  for (int r = 0; r < IPT; ++r) {
    const int e = base + r;
    const bool asc = (e & k) == 0, lower = (e & j) == 0;
    const K lo = v[r] < o[r] ? v[r] : o[r], hi = v[r] < o[r] ? o[r] : v[r];
    v[r] = lower == asc ? lo : hi;
  }
}

// One compare-exchange stage of stride J < IPT inside every thread's
// registers (element e = base + r sits in v[r]).
template <int J, int IPT, typename K>
__device__ __forceinline__ void reg_cx_stage(K (&v)[IPT], int base, int k) {
#pragma unroll
  for (int r = 0; r < IPT; ++r) {
    if ((r & J) == 0) {
      const bool asc = ((base + r) & k) == 0;
      const K a = v[r], b = v[r + J];
      const bool sw = (a > b) == asc;
      v[r] = sw ? b : a;
      v[r + J] = sw ? a : b;
    }
  }
}

// Ascending bitonic sort of T*IPT keys held IPT per thread, thread-contiguous
// (element e = threadIdx.x*IPT + r in v[r]).  Strides below IPT compare
// registers, strides below 64*IPT exchange with lane (tid ^ stride/IPT) of
// the same wave by shuffles, and only the longer ones go through LDS (s,
// T*IPT keys) with barriers: a 2048-key sort by 256 threads has 3 such
// stages out of 66.
// Stage (merge size KS, stride J) of the bitonic network, with both known at
// compile time: the stride picks registers, lanes or LDS, and the direction
// masks fold to constants.
template <int T, int IPT, int KS, int J, typename K>
__device__ __forceinline__ void bitonic_stage(K (&v)[IPT], K* s, int base) {
  if constexpr (J < IPT) {
    reg_cx_stage<J, IPT>(v, base, KS);
  } else if constexpr (J / IPT < kWave) {
    lane_cx_stage<J / IPT, IPT>(v, base, KS, J);
  } else {  // another wave's registers: through LDS
#pragma unroll
    for (int r = 0; r < IPT; ++r) s[base + r] = v[r];
    __syncthreads();
#pragma unroll
    for (int r = 0; r < IPT; ++r) {
      const int e = base + r;
      const K o = s[e ^ J];
      const bool asc = (e & KS) == 0, lower = (e & J) == 0;
      const K lo = v[r] < o ? v[r] : o, hi = v[r] < o ? o : v[r];
      v[r] = lower == asc ? lo : hi;
    }
    __syncthreads();
  }
  if constexpr (J > 1) bitonic_stage<T, IPT, KS, J / 2>(v, s, base);
}

template <int T, int IPT, int KS, typename K>
__device__ __forceinline__ void bitonic_levels(K (&v)[IPT], K* s, int base) {
  bitonic_stage<T, IPT, KS, KS / 2>(v, s, base);
  if constexpr (KS < T * IPT) bitonic_levels<T, IPT, KS * 2>(v, s, base);
}

template <int T, int IPT, typename K>
__device__ __forceinline__ void reg_bitonic_sort(K (&v)[IPT], K* s) {
  static_assert(IPT == 1 || IPT == 2 || IPT == 4 || IPT == 8, "1, 2, 4 or 8 keys per thread");
  static_assert(T * IPT >= 2, "at least two keys");
  bitonic_levels<T, IPT, 2>(v, s, (int)threadIdx.x * IPT);
}

// Lane i ^ M's value for M = 2^k - 1 (the mirror partner inside groups of
// M + 1 lanes): quad_perm [1,0,3,2], quad_perm [3,2,1,0], row_half_mirror
// and row_mirror are one DPP move each; 31 and 63 go through ds_bpermute.
template <int M>
__device__ __forceinline__ uint32_t lane_mirror(uint32_t x) {
  if constexpr (M == 1) {
    return dpp_move<0xb1, 0xf>(x);
  } else if constexpr (M == 3) {
    return dpp_move<0x1b, 0xf>(x);
  } else if constexpr (M == 7) {
    return dpp_move<0x141, 0xf>(x);
  } else if constexpr (M == 15) {
    return dpp_move<0x140, 0xf>(x);
  } else {
    return (uint32_t)__shfl_xor((int)x, M, kWave);
  }
}

// One stage of the flip (mirror) bitonic network on T*IPT 32-bit keys held
// IPT per thread, thread-contiguous (element e = base + r in v[r], base =
// tid*IPT).  MIR: the first stage of merge size KS, partner e ^ (KS - 1);
// otherwise a half-cleaner, partner e ^ J.  The lower element of a pair
// always takes the minimum, so no stage has a direction.  Elements at and
// past nact are padding (0xffffffff) in waves that hold nothing else; such
// a wave (act false) skips the stage, and an element whose partner lies
// there keeps its key, which is what the comparator would leave (the
// partner is the higher element and the larger key).  Every wave still
// meets the barriers of the stages that cross waves through LDS.
template <int T, int IPT, int KS, int J, bool MIR>
__device__ __forceinline__ void flip_stage(uint32_t (&v)[IPT], uint32_t* s, int base, bool act, int nact) {
  if constexpr (MIR ? KS <= IPT : J < IPT) {  // inside the thread's registers
    if (act) {
#pragma unroll
      for (int r = 0; r < IPT; ++r) {
        const int p = MIR ? (r ^ (KS - 1)) : (r ^ J);
        if (r < p) {
          const uint32_t lo = min(v[r], v[p]), hi = max(v[r], v[p]);
          v[r] = lo;
          v[p] = hi;
        }
      }
    }
  } else if constexpr (MIR ? KS <= kWave * IPT : J < kWave * IPT) {  // lanes of the same wave
    if (act) {
      uint32_t o[IPT];
#pragma unroll
      for (int r = 0; r < IPT; ++r) {
        if constexpr (MIR) o[r] = lane_mirror<KS / IPT - 1>(v[IPT - 1 - r]);
        else o[r] = lane_xor<J / IPT>(v[r]);
      }
      const bool lower = (lane_id() & (MIR ? KS / (2 * IPT) : J / IPT)) == 0;
#pragma unroll
      for (int r = 0; r < IPT; ++r) v[r] = lower ? min(v[r], o[r]) : max(v[r], o[r]);
    }
  } else {  // another wave's registers: through LDS
    // The partner of every element of wave w lies in one wave, w ^ (KS/WK - 1)
    // (mirror) or w ^ (J/WK); a partner wave in the padding suffix leaves
    // this wave's keys as they are.
    constexpr int WK = kWave * IPT;
    const int pw = MIR ? (wave_id() ^ (KS / WK - 1)) : (wave_id() ^ (J / WK));
    const bool both = act && __builtin_amdgcn_readfirstlane(pw * WK) < nact;
    if (act) {
#pragma unroll
      for (int r = 0; r < IPT; ++r) s[base + r] = v[r];
    }
    __syncthreads();
    if (both) {
      uint32_t o[IPT];  // every partner read goes out before any compare
#pragma unroll
      for (int r = 0; r < IPT; ++r) o[r] = s[MIR ? ((base + r) ^ (KS - 1)) : ((base + r) ^ J)];
#pragma unroll
      for (int r = 0; r < IPT; ++r) {
        const int e = base + r, p = MIR ? (e ^ (KS - 1)) : (e ^ J);
        v[r] = e < p ? min(v[r], o[r]) : max(v[r], o[r]);
      }
    }
    __syncthreads();
  }
}

template <int T, int IPT, int KS, int J>
__device__ __forceinline__ void flip_halves(uint32_t (&v)[IPT], uint32_t* s, int base, bool act, int nact) {
  if constexpr (J >= 1) {
    flip_stage<T, IPT, KS, J, false>(v, s, base, act, nact);
    flip_halves<T, IPT, KS, J / 2>(v, s, base, act, nact);
  }
}

constexpr int next_pow2_c(int x) { return x <= 1 ? 1 : 2 * next_pow2_c((x + 1) / 2); }

template <int T, int IPT, int KS>
__device__ __forceinline__ void flip_levels(uint32_t (&v)[IPT], uint32_t* s, int base, bool act, int nact) {
  flip_stage<T, IPT, KS, KS / 2, true>(v, s, base, act, nact);
  flip_halves<T, IPT, KS, KS / 4>(v, s, base, act, nact);
  if constexpr (KS < next_pow2_c(T * IPT)) flip_levels<T, IPT, KS * 2>(v, s, base, act, nact);
}

// Ascending sort of T*IPT 32-bit keys, IPT per thread, thread-contiguous,
// by the flip bitonic network over the next power of two of T*IPT keys (the
// positions past T*IPT are virtual padding), of which only the first f
// elements (in whole waves of 64*IPT) take part: the rest must be
// 0xffffffff, and those waves skip every stage.  s: T*IPT keys of LDS
// (stages that cross waves).  All threads call it: it holds barriers when
// T > 64.
template <int T, int IPT>
__device__ __forceinline__ void reg_flip_sort(uint32_t (&v)[IPT], uint32_t* s, int f) {
  static_assert(IPT == 1 || IPT == 2 || IPT == 4 || IPT == 8, "1, 2, 4 or 8 keys per thread");
  static_assert(T * IPT >= 2 && T % kWave == 0, "whole waves, at least two keys");
  constexpr int kWaveKeys = kWave * IPT;
  const int nact = (f + kWaveKeys - 1) / kWaveKeys * kWaveKeys;
  const bool act = __builtin_amdgcn_readfirstlane(wave_id() * kWaveKeys) < nact;
  flip_levels<T, IPT, 2>(v, s, (int)threadIdx.x * IPT, act, nact);
}

// Sorts perm[0..n) = 0..n-1 by the composite (key[i], i) ascending: a
// counting pass over nb (power of two, >= n) order-preserving buckets of
// the key range [kmin, kmin + span), a scatter, and an insertion sort of
// every bucket by its owner thread.  Keys spread evenly over the range
// (rows of uniform graphs) leave O(1) elements per bucket; when a bucket
// holds more than kMaxBucket elements the whole array is re-sorted with an
// indirect bitonic sort instead.  cnt holds nb ints.  Ends with a barrier.
constexpr int kMaxBucket = 32;

template <int T>
__device__ __forceinline__ void block_sort_perm(const uint32_t* key, uint16_t* perm, int32_t* cnt, int n, int nb,
                                                uint32_t kmin, uint64_t span, int32_t* s_flag) {
  // monotone bucket map: float conversion, scaling and truncation all keep
  // order, so bucket(k1) <= bucket(k2) whenever k1 <= k2
  const float scale = (float)nb / (float)span;
  auto bucket = [&](uint32_t k) -> int {
    const int b = (int)((float)(k - kmin) * scale);
    return b < nb ? b : nb - 1;
  };
  for (int b = threadIdx.x; b < nb; b += T) cnt[b] = 0;
  if (threadIdx.x == 0) *s_flag = 0;
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += T) atomicAdd(&cnt[bucket(key[i])], 1);
  __syncthreads();
  // exclusive scan of cnt: each thread owns nb / T consecutive buckets
  __shared__ int32_t s_w[T / kWave];
  const int per = nb >= T ? nb / T : 1;
  const int b0 = threadIdx.x * per;
  int32_t mine = 0;
  if (b0 < nb)
    for (int b = b0; b < b0 + per; ++b) mine += cnt[b];
  int32_t tot;
  int32_t run = block_excl_scan<T, int32_t>(mine, s_w, tot);
  if (b0 < nb)
    for (int b = b0; b < b0 + per; ++b) {
      const int32_t c = cnt[b];
      cnt[b] = run;
      run += c;
    }
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += T) perm[atomicAdd(&cnt[bucket(key[i])], 1)] = (uint16_t)i;
  __syncthreads();
  // cnt[b] is now the end of bucket b
  for (int b = threadIdx.x; b < nb; b += T) {
    const int e = cnt[b], s = b == 0 ? 0 : cnt[b - 1];
    if (e - s > kMaxBucket) {
      *s_flag = 1;
      continue;
    }
    for (int x = s + 1; x < e; ++x) {
      const uint16_t v = perm[x];
      const uint64_t kv = ((uint64_t)key[v] << 16) | v;
      int y = x - 1;
      while (y >= s && (((uint64_t)key[perm[y]] << 16) | perm[y]) > kv) {
        perm[y + 1] = perm[y];
        --y;
      }
      perm[y + 1] = v;
    }
  }
  __syncthreads();
  if (*s_flag) {
    const int n2 = 1 << (32 - __clz(n - 1 > 0 ? n - 1 : 1));
    for (int i = threadIdx.x; i < n2; i += T) perm[i] = (uint16_t)i;  // identity, padded
    __syncthreads();
    for (int k = 2; k <= n2; k <<= 1) {
      for (int j = k >> 1; j > 0; j >>= 1) {
        for (int i = threadIdx.x; i < (n2 >> 1); i += T) {
          const int a = ((i & ~(j - 1)) << 1) | (i & (j - 1));
          const int c = a + j;
          const uint16_t x = perm[a], y = perm[c];
          const uint64_t kx = x < n ? (((uint64_t)key[x] << 16) | x) : ~0ull;
          const uint64_t ky = y < n ? (((uint64_t)key[y] << 16) | y) : ~0ull;
          const bool up = (a & k) == 0;
          if ((kx > ky) == up) { perm[a] = y; perm[c] = x; }
        }
        __syncthreads();
      }
    }
  }
}

}  // namespace ams
