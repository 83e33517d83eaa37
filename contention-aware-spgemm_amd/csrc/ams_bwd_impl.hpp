// ams_bwd_impl.hpp -- masked-SpGEMM backward of C = A * B (included once,
// by ams_spgemm.hip).
//
// With the upstream gradient G on pattern(C):
//   dA[i,k] = sum_j G[i,j] B[k,j]   over (i,k) in pattern(A)
//   dB[k,j] = sum_i A[i,k] G[i,j]   over (k,j) in pattern(B)
// Both come from one pass over the forward's products (i, k, j): every
// product j of row i lies in pattern(C) row i, so G's row i is the
// pre-seeded key set of the row's accumulator, looked up instead of
// inserted.  dA[p] (p = A nonzero) is a sum over one B row: reduced in
// registers and LDS and written once.  dB[q] (q = B nonzero) is shared
// between rows: one global atomic per product at the known CSR slot q.
// Cost: P lookups + P atomics, against P inserts of the forward.
#pragma once

#include "ams_kernels_impl.hpp"

namespace ams {

// Value of lane i - d inside 16-lane rows (row_shr), or of the last lane of
// the previous row / of row 1 (row_bcast:15 / :31) for the rows in kRowMask;
// 0 where there is no source.
template <int kCtrl, int kRowMask, typename V>
__device__ __forceinline__ V dpp_val(V v) {
  if constexpr (sizeof(V) == 8) {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = dpp_move<kCtrl, kRowMask>((uint32_t)u);
    const uint32_t hi = dpp_move<kCtrl, kRowMask>((uint32_t)(u >> 32));
    return __builtin_bit_cast(V, ((uint64_t)hi << 32) | lo);
  } else {
    return __builtin_bit_cast(V, dpp_move<kCtrl, kRowMask>(__builtin_bit_cast(uint32_t, v)));
  }
}

// Wave-segmented sum of v over runs of consecutive lanes with equal key: a
// segmented inclusive scan by DPP lane moves (the scan of wave_incl_scan
// with a run-start flag that stops each addition), so the last lane of each
// run ends with the run's total; returns true on that lane.
template <typename V>
__device__ __forceinline__ bool wave_seg_sum(int key, V& v) {
  const int lane = lane_id();
  const int kp = __shfl_up(key, 1, kWave), kn = __shfl_down(key, 1, kWave);
  uint32_t f = (lane == 0 || kp != key) ? 1u : 0u;  // a run starts in the lanes covered so far
  auto step = [&](auto ctrl_c, auto mask_c) {
    constexpr int C = decltype(ctrl_c)::value, M = decltype(mask_c)::value;
    const V vs = dpp_val<C, M>(v);
    const uint32_t fs = dpp_move<C, M>(f);
    v = f ? v : v + vs;
    f |= fs;
  };
  step(std::integral_constant<int, 0x111>{}, std::integral_constant<int, 0xf>{});  // row_shr:1
  step(std::integral_constant<int, 0x112>{}, std::integral_constant<int, 0xf>{});  // row_shr:2
  step(std::integral_constant<int, 0x114>{}, std::integral_constant<int, 0xf>{});  // row_shr:4
  step(std::integral_constant<int, 0x118>{}, std::integral_constant<int, 0xf>{});  // row_shr:8
  step(std::integral_constant<int, 0x142>{}, std::integral_constant<int, 0xa>{});  // row_bcast:15
  step(std::integral_constant<int, 0x143>{}, std::integral_constant<int, 0xc>{});  // row_bcast:31
  return lane == kWave - 1 || kn != key;
}

// One window-flattened pass over the products of A nonzeros [p0, p1).  In
// every stride of T*U products of a window, wave w owns the 64*U consecutive
// products from w*64*U and lane l takes products l, l + 64, ... of them, so
// each group of 64 (one per u) covers 64 consecutive products, over which
// the body may run wave-wide operations (every lane runs every u).  Each
// lane keeps the segment of its last product and seeks forward from it.
//   seg(p, st&, len&)          segment of nonzero p, on thread p - w0
//   body(ok[U], j[U], q[U])    j = index of the nonzero in the window, q =
//                              B position (lanes past the end: ok false, q
//                              and j those of the window's last product)
//   flush(j, p)                per nonzero after each window
template <int T, int U, class Seg, class Body, class Flush>
__device__ __forceinline__ void block_expand_flush(int64_t p0, int64_t p1, ExpandLds<T>& L, Seg seg, Body body,
                                                   Flush flush) {
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
    int cur = -1;  // segment_advance's start: segment -1 ends at off[0] = 0
    int32_t seg_end = 0;
    int64_t seg_bias = 0;
    for (int32_t t0 = 0; t0 < total; t0 += T * U) {
      bool ok[U];
      int j[U];
      int64_t q[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        int32_t t = t0 + wave_id() * (kWave * U) + u * kWave + lane_id();
        ok[u] = t < total;
        t = ok[u] ? t : total - 1;
        if (t >= seg_end) segment_advance(L.off, L.start, nvalid, total, t, cur, seg_end, seg_bias);
        j[u] = cur;
        q[u] = seg_bias + t;
      }
      body(ok, j, q);
    }
    __syncthreads();
    if (threadIdx.x < nvalid) flush(threadIdx.x, w0 + threadIdx.x);
    __syncthreads();
  }
}

// Products in flight per lane in the backward expansion.
constexpr int kBwdU = 4;

// Rows whose G row fits an LDS table: one row per block.
template <int S, int T, typename V>
__global__ __launch_bounds__(T) void k_bwd_hash(Csr A, Csr B, Csr G, const int32_t* list, const int32_t* counts,
                                                int bin, V* dA, V* dB, int32_t* err) {
  __shared__ int32_t keys[S];
  __shared__ V gv[S];
  __shared__ ExpandLds<T> L;
  __shared__ V aval[T];
  __shared__ V sda[T];
  constexpr int shift = 32 - __builtin_ctz(S);
  int32_t start, cnt;
  bin_range(counts, bin, start, cnt);
  const V* av = (const V*)A.val;
  const V* bv = (const V*)B.val;
  const V* gval = (const V*)G.val;
  for (int32_t idx = blockIdx.x; idx < cnt; idx += gridDim.x) {
    const int32_t row = list[start + idx];
    for (int s = threadIdx.x; s < S; s += T) keys[s] = kEmpty;
    __syncthreads();
    for (int64_t e = G.rp[row] + threadIdx.x; e < G.rp[row + 1]; e += T) {
      const int32_t key = G.col[e];
      const uint32_t mask = (0xffffffffu >> shift);
      uint32_t h = hash_slot(key, shift);
      uint32_t probe = 0;
      while (atomicCAS(&keys[h], kEmpty, key) != kEmpty && probe <= mask) {
        h = (h + 1) & mask;
        ++probe;
      }
      if (probe > mask) atomicOr(err, 32);
      else gv[h] = gval[e];
    }
    __syncthreads();
    block_expand_flush<T, kBwdU>(
        A.rp[row], A.rp[row + 1], L,
        [&](int64_t p, int64_t& st, int32_t& len) {
          const int32_t k = A.col[p];
          st = B.rp[k];
          len = (int32_t)(B.rp[k + 1] - st);
          aval[threadIdx.x] = av[p];
          sda[threadIdx.x] = V(0);
        },
        [&](bool (&ok)[kBwdU], int (&j)[kBwdU], int64_t (&q)[kBwdU]) {
          int32_t bc[kBwdU];
          V bq[kBwdU];
#pragma unroll
          for (int u = 0; u < kBwdU; ++u) {
            bc[u] = B.col[q[u]];
            bq[u] = bv[q[u]];
          }
#pragma unroll
          for (int u = 0; u < kBwdU; ++u) {
            V contrib = V(0);
            if (ok[u]) {
              const uint32_t slot = lds_find(keys, shift, bc[u]);
              const V g = slot < (uint32_t)S ? gv[slot] : V(0);
              if (slot >= (uint32_t)S) atomicOr(err, 64);
              unsafeAtomicAdd(&dB[q[u]], aval[j[u]] * g);
              contrib = g * bq[u];
            }
            if (wave_seg_sum<V>(ok[u] ? j[u] : -1, contrib) && ok[u]) lds_add(&sda[j[u]], contrib);
          }
        },
        [&](int j, int64_t p) { dA[p] = sda[j]; });
  }
}

// dB[perm[q]] += dBb[q]: the dense bin's dB partials, kept in the order of
// B's column-blocked copy (q), added at B's own slots (perm is a bijection,
// so no two threads write one slot).
template <typename V>
__global__ __launch_bounds__(256) void k_bwd_unpermute(const int64_t* perm, const V* dBb, V* dB, int64_t n) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
    dB[perm[q]] += dBb[q];
}

// First G entry of every (dense-bin row h, column chunk c), c = 0..R:
// gsplit[h * (R + 1) + c] = lower bound of the chunk's first column in G's
// row (G rows are sorted), one thread per entry, so the morsels of the
// backward read their range instead of searching G's row.
__global__ __launch_bounds__(256) void k_bwd_gsplit(Csr G, const int32_t* list, const int32_t* counts, int bin, int R,
                                                    int W, Probs pr, int64_t* gsplit) {
  int32_t start, H;
  bin_range(counts, bin, start, H);
  const int64_t n = (int64_t)H * (R + 1);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t h = (int32_t)(i / (R + 1));
    const int c = (int)(i - (int64_t)h * (R + 1));
    const int32_t row = list[start + h];
    int64_t lo, width;
    prob_cols(pr, row, lo, width);
    const int64_t c0 = lo + (int64_t)c * W;
    int64_t a = G.rp[row], b = G.rp[row + 1];
    while (a < b) {
      const int64_t mid = (a + b) >> 1;
      if (G.col[mid] < c0) a = mid + 1;
      else b = mid;
    }
    gsplit[i] = a;
  }
}

// Rows whose G row is too wide for a table: persistent blocks take
// (row, column chunk) tickets like the forward's dense bin; G's chunk is
// scattered into a dense LDS array, products come from the column-blocked
// B, and dB accumulates in the blocked order (dBb[q], consecutive for the
// consecutive products of a segment) for k_bwd_unpermute.
template <int T, typename V>
__global__ __launch_bounds__(T) void k_bwd_dense(Csr A, Csr G, int64_t mB, const int64_t* bptr, const uint16_t* bcol,
                                                 const V* bval, const int32_t* list,
                                                 const int32_t* counts, int bin, int R, int W, Probs pr,
                                                 const int64_t* gsplit, unsigned long long* ticket, V* dA,
                                                 V* dBb) {
  extern __shared__ __align__(16) unsigned char dyn_raw[];
  V* gd = (V*)dyn_raw;
  __shared__ ExpandLds<T> L;
  __shared__ V aval[T];
  __shared__ V sda[T];
  __shared__ unsigned long long s_ticket;
  int32_t start, H;
  bin_range(counts, bin, start, H);
  for (int i = threadIdx.x; i < W; i += T) gd[i] = V(0);
  const unsigned long long ntask = (unsigned long long)H * (unsigned long long)R;
  const V* av = (const V*)A.val;
  const V* gval = (const V*)G.val;
  const int64_t ld = mB + 1;
  // thread 0 claims the next morsel while the block works on the current
  // one, so the claim's latency is not waited for
  unsigned long long nxt = 0;
  if (threadIdx.x == 0) nxt = atomicAdd(ticket, 1ull);
  for (;;) {
    __syncthreads();  // every thread has read the last ticket
    if (threadIdx.x == 0) s_ticket = nxt;
    __syncthreads();
    const unsigned long long tk = s_ticket;
    if (tk >= ntask) break;
    if (threadIdx.x == 0) nxt = atomicAdd(ticket, 1ull);
    const int c = (int)(tk / (unsigned long long)H);
    const int32_t h = (int32_t)(tk - (unsigned long long)c * H);
    const int32_t row = list[start + h];
    int64_t lo, width;
    prob_cols(pr, row, lo, width);
    const int64_t c0 = lo + (int64_t)c * W;
    // G's entries of this chunk
    const int64_t e0 = gsplit[(int64_t)h * (R + 1) + c], e1 = gsplit[(int64_t)h * (R + 1) + c + 1];
    if (e0 == e1) continue;
    for (int64_t e = e0 + threadIdx.x; e < e1; e += T) gd[G.col[e] - c0] = gval[e];
    __syncthreads();
    const int64_t* bp = bptr + (int64_t)c * ld;
    block_expand_flush<T, kBwdU>(
        A.rp[row], A.rp[row + 1], L,
        [&](int64_t p, int64_t& st, int32_t& len) {
          const int32_t k = A.col[p];
          st = bp[k];
          len = (int32_t)(bp[k + 1] - st);
          aval[threadIdx.x] = av[p];
          sda[threadIdx.x] = V(0);
        },
        [&](bool (&ok)[kBwdU], int (&j)[kBwdU], int64_t (&q)[kBwdU]) {
          // every lane's loads go out before the first is used
          uint32_t bc[kBwdU];
          V bq[kBwdU];
#pragma unroll
          for (int u = 0; u < kBwdU; ++u) {
            bc[u] = bcol[q[u]];
            bq[u] = bval[q[u]];
          }
#pragma unroll
          for (int u = 0; u < kBwdU; ++u) {
            const V g = gd[bc[u]];
            V contrib = V(0);
            if (ok[u]) {
              unsafeAtomicAdd(&dBb[q[u]], aval[j[u]] * g);
              contrib = g * bq[u];
            }
            if (wave_seg_sum<V>(ok[u] ? j[u] : -1, contrib) && ok[u]) lds_add(&sda[j[u]], contrib);
          }
        },
        [&](int j, int64_t p) {
          if (sda[j] != V(0)) unsafeAtomicAdd(&dA[p], sda[j]);
        });
    for (int64_t e = e0 + threadIdx.x; e < e1; e += T) gd[G.col[e] - c0] = V(0);
  }
}

// Backward route of every row: by nnz of G's row.
__global__ __launch_bounds__(256) void k_classify_bwd(int64_t m, const int64_t* g_rp, int64_t hash_max,
                                                      uint8_t* bin, int32_t* counts) {
  for (int64_t i0 = blockIdx.x * (int64_t)blockDim.x; i0 < m; i0 += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = i0 + threadIdx.x;
    int b = NB_NONE;
    if (i < m) {
      const int64_t z = g_rp[i + 1] - g_rp[i];
      if (z > 0) {
        if (z > hash_max) {
          b = NB_DENSE;
        } else {
          const int64_t s = next_pow2_i64(z + ((z + 1) >> 1));
          b = s <= 64 ? 0 : ilog2_pow2(s) - 6;
        }
      }
      bin[i] = (uint8_t)b;
    }
    (void)wave_append(counts, b, b != NB_NONE);
  }
}

}  // namespace ams
