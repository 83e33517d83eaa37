// ams_kernels_impl.hpp -- HIP/CDNA4 kernels of the hybrid-accumulation SpGEMM.
//
// Pipeline (one stream, no MFMA: the work is sparse gathers and LDS atomics):
//   k_row_flops      row_flp[i] = sum_{k in A(i,:)} |B(k,:)|
//                    (flp_populate, reference src/mm/mm_utils.c:1063-1125)
//   k_classify_sym   symbolic route per row: LDS hash table sized by flops,
//                    or the LDS bitmap for wide rows
//   k_scatter_bins   rows grouped by route (replaces the morsel packing of
//                    pack_ams_pack, src/ams/ams_utils.c:1410-1596)
//   k_symb_hash      distinct columns per row with an LDS hash
//                    (ams_hash_symb_dnr, src/ams/ams_utils.c:1600-1752)
//   k_symb_bitmap    distinct columns per row and per dense column chunk
//   k_scan_*         exclusive scan (scan, src/mm/mm_utils.c:866-957)
//   k_classify_num   numeric route: ESC for low compression, LDS hash, or
//                    dense column chunks (the cf routing of
//                    ams_hybacc_numc_krnl, src/ams/ams_utils.c:1141-1149)
//   k_numc_hash      LDS hash accumulate, sort keys, write sorted row
//                    (hybacc_hash_krnl, src/ams/ams_utils.c:956-1019)
//   k_numc_esc       expand to LDS, bitonic sort (key, rank), compress runs
//                    left to right (hybacc_hsersc_krnl, :1023-1075)
//   k_bblk_*         column-blocked copy of B for the dense bin
//   k_numc_dense     persistent kernel: a global ticket queue hands out
//                    (heavy row, column chunk) morsels; each morsel is
//                    accumulated in a dense LDS array and written sorted
#pragma once
#include "ams_kernels.hpp"

namespace ams {

// ---------------------------------------------------------------- helpers

__global__ void k_fill_i64(int64_t* p, int64_t n, int64_t v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void k_fill_i32(int32_t* p, int64_t n, int32_t v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void k_i64_to_i32(const int64_t* in, int32_t* out, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = (int32_t)in[i];
}

__global__ void k_i32_to_i64(const int32_t* in, int64_t* out, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = (int64_t)in[i];
}

__device__ __forceinline__ int64_t next_pow2_i64(int64_t x) {
  return x <= 1 ? 1 : (int64_t)1 << (64 - __clzll((unsigned long long)(x - 1)));
}

__device__ __forceinline__ int ilog2_pow2(int64_t x) { return 63 - __clzll((unsigned long long)x); }

// Segment of A nonzero p in the original B: B row A.col[p].
struct SegB {
  const int32_t* acol;
  const int64_t* brp;
  __device__ __forceinline__ void operator()(int64_t p, int64_t& st, int32_t& len) const {
    const int32_t k = acol[p];
    st = brp[k];
    len = (int32_t)(brp[k + 1] - st);
  }
};

// The same with 32-bit byte offsets from the arrays' bases (A with fewer
// than 2^30 nonzeros, B with fewer than 2^28 rows): the loads take a scalar
// base and one vector offset register, and no 64-bit address stays live.
struct SegB32 {
  const int32_t* acol;
  const int64_t* brp;
  __device__ __forceinline__ void operator()(int64_t p, int64_t& st, int32_t& len) const {
    const int32_t k = *(const int32_t*)((const char*)acol + ((uint32_t)p << 2));
    const int64_t* b = (const int64_t*)((const char*)brp + ((uint32_t)k << 3));
    st = b[0];
    len = (int32_t)(b[1] - st);
  }
};

// Products in flight per lane in the expansion (loads issued before the
// LDS updates they feed), per kernel.
#ifndef AMS_SYMB_U
#define AMS_SYMB_U 16
#endif
constexpr int kUSymBitmap = AMS_SYMB_U;
// products per thread and pass of the dense numeric kernel (register bound)
#ifndef AMS_DENSE_U1024
#define AMS_DENSE_U1024 6
#endif
#ifndef AMS_DENSE_U512
#define AMS_DENSE_U512 12
#endif
template <int T>
constexpr int dense_u() { return T >= 1024 ? AMS_DENSE_U1024 : AMS_DENSE_U512; }

// f(integral_constant<int, i>) for i in [B, E), unrolled at compile time.
template <int B, int E, class F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

// A non-empty B segment of a window of A nonzeros: product t of the window
// (t in [off, off + len)) reads B position bias + t and is scaled by a.
template <typename V>
struct __align__(16) SegBias {
  int64_t bias;
  V a;
};

template <typename V>
struct KeyVal {
  int32_t key;
  V val;
};

template <typename V>
struct RankKeyVal {
  int32_t rank;  // product's rank in A-row order
  int32_t key;   // column
  V val;
};

// What the expansion gathers for one product: column, window index of the
// A nonzero, B value (the product itself is formed in apply).
template <typename V>
struct Gather {
  int32_t key;
  int32_t j;
  V val;
};

template <typename V>
struct RankGather {
  int32_t rank;
  int32_t key;
  int32_t j;
  V val;
};

// A product of the expansion with its A value gathered alongside (no LDS
// staging of the window's A values).
template <typename V>
struct RankGatherA {
  int32_t rank;
  int32_t key;
  V a;
  V val;
};

// ------------------------------------------------------------- row flops

// Row flops in groups of 64 rows per wave: rows of at most kFlpShort
// nonzeros are summed by their own lane, longer ones by the whole wave, one
// at a time.
constexpr int64_t kFlpShort = 16;
// Also accumulates P into *total.
__global__ __launch_bounds__(256) void k_row_flops(Csr A, Csr B, int64_t* row_flp,
                                                   unsigned long long* total) {
  __shared__ int64_t s_w[4];
  const int64_t waves = (int64_t)gridDim.x * 4;
  int64_t acc = 0;
  for (int64_t r0 = (blockIdx.x * 4 + wave_id()) * (int64_t)kWave; r0 < A.m; r0 += waves * kWave) {
    const int64_t r = r0 + lane_id();
    const bool valid = r < A.m;
    const int64_t s = valid ? A.rp[r] : 0, e = valid ? A.rp[r + 1] : 0;
    const bool is_long = e - s > kFlpShort;
    if (valid && !is_long) {
      int64_t f = 0;
      for (int64_t p = s; p < e; ++p) {
        const int32_t k = A.col[p];
        f += B.rp[k + 1] - B.rp[k];
      }
      if (row_flp) row_flp[r] = f;
      acc += f;
    }
    unsigned long long pending = __ballot(valid && is_long);
    while (pending) {
      const int lead = __ffsll(pending) - 1;
      pending &= pending - 1;
      const int64_t ss = (int64_t)uniform_u64(__shfl((unsigned long long)s, lead, kWave));
      const int64_t ee = (int64_t)uniform_u64(__shfl((unsigned long long)e, lead, kWave));
      int64_t f = 0;
      for (int64_t p = ss + lane_id(); p < ee; p += kWave) {
        const int32_t k = A.col[p];
        f += B.rp[k + 1] - B.rp[k];
      }
      f = wave_sum(f);
      if (lane_id() == 0) {
        if (row_flp) row_flp[r0 + lead] = f;
        acc += f;
      }
    }
  }
  const int64_t t = block_sum<256, int64_t>(acc, s_w);
  if (threadIdx.x == 0 && t) atomicAdd(total, (unsigned long long)t);
}

// One wave per row of A (rows longer than kFlpShort on average); also
// accumulates P into *total.  A row's nonzeros go in strides of 8*64: the
// stride's A columns, then its B bounds, are all loaded before any is used,
// so a hub row of 10^5 nonzeros (which sets the kernel's time) takes a few
// hundred load latencies instead of a few thousand.
#ifndef AMS_FLP_UNROLL
#define AMS_FLP_UNROLL 8
#endif
constexpr int kFlpUnroll = AMS_FLP_UNROLL;
__global__ __launch_bounds__(256) void k_row_flops_wave(Csr A, Csr B, int64_t* row_flp,
                                                        unsigned long long* total) {
  __shared__ int64_t s_w[4];
  const int64_t waves = (int64_t)gridDim.x * 4;
  int64_t acc = 0;
  for (int64_t r = blockIdx.x * 4 + wave_id(); r < A.m; r += waves) {
    const int64_t s = A.rp[r], e = A.rp[r + 1];
    int64_t f = 0;
    for (int64_t p0 = s + lane_id(); p0 < e; p0 += kWave * kFlpUnroll) {
      int32_t k[kFlpUnroll];
#pragma unroll
      for (int j = 0; j < kFlpUnroll; ++j) {
        const int64_t p = p0 + (int64_t)j * kWave;
        k[j] = p < e ? A.col[p] : -1;
      }
      int64_t lo[kFlpUnroll], hi[kFlpUnroll];
#pragma unroll
      for (int j = 0; j < kFlpUnroll; ++j) {
        const int32_t kk = k[j] < 0 ? 0 : k[j];
        lo[j] = B.rp[kk];
        hi[j] = B.rp[kk + 1];
      }
#pragma unroll
      for (int j = 0; j < kFlpUnroll; ++j) f += k[j] < 0 ? 0 : hi[j] - lo[j];
    }
    f = wave_sum(f);
    if (lane_id() == 0) {
      if (row_flp) row_flp[r] = f;
      acc += f;
    }
  }
  const int64_t t = block_sum<256, int64_t>(acc, s_w);
  if (threadIdx.x == 0 && t) atomicAdd(total, (unsigned long long)t);
}

// Nonzero-split row flops for matrices with very long rows: each thread
// takes kFlpIpt consecutive nonzeros of A, finds its first row by binary
// search over A.rp, and adds its partial sums to row_flp with one atomic per
// row it touches.  row_flp must start at zero.
constexpr int kFlpIpt = 64;
__global__ __launch_bounds__(256) void k_row_flops_nz(Csr A, Csr B, unsigned long long* row_flp,
                                                      unsigned long long* total) {
  __shared__ int64_t s_w[4];
  int64_t acc_all = 0;
  const int64_t span = (int64_t)gridDim.x * blockDim.x * kFlpIpt;
  for (int64_t z0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * kFlpIpt; z0 < A.nnz; z0 += span) {
    int64_t lo = 0, hi = A.m;  // largest r with rp[r] <= z0
    while (hi - lo > 1) {
      const int64_t mid = (lo + hi) >> 1;
      if (A.rp[mid] <= z0) lo = mid; else hi = mid;
    }
    int64_t r = lo, rend = A.rp[r + 1], acc = 0;
    const int64_t z1 = z0 + kFlpIpt < A.nnz ? z0 + kFlpIpt : A.nnz;
    for (int64_t p = z0; p < z1; ++p) {
      while (p >= rend) {
        if (acc) atomicAdd(&row_flp[r], (unsigned long long)acc);
        acc = 0;
        ++r;
        rend = A.rp[r + 1];
      }
      const int32_t k = A.col[p];
      const int64_t d = B.rp[k + 1] - B.rp[k];
      acc += d;
      acc_all += d;
    }
    if (acc) atomicAdd(&row_flp[r], (unsigned long long)acc);
  }
  const int64_t t = block_sum<256, int64_t>(acc_all, s_w);
  if (threadIdx.x == 0 && t) atomicAdd(total, (unsigned long long)t);
}

// ------------------------------------------------------------ binning

// Wave-aggregated append of this lane's row (bin b, b != 0xff) into the
// per-bin counters; returns the rank within the bin.
__device__ __forceinline__ int32_t wave_append(int32_t* counters, int b, bool active) {
  int32_t rank = -1;
  unsigned long long pending = __ballot(active);
  while (pending) {
    const int lead = __ffsll(pending) - 1;
    const int bb = __shfl(b, lead, kWave);
    const unsigned long long mask = __ballot(active && b == bb);
    int32_t base = 0;
    if (lane_id() == lead) base = atomicAdd(&counters[bb], (int32_t)__popcll(mask));
    base = __shfl(base, lead, kWave);
    if (active && b == bb) {
      rank = base + (int32_t)__popcll(mask & ((1ull << lane_id()) - 1));
      active = false;
    }
    pending &= ~mask;
  }
  return rank;
}

// Block-aggregated bin counting: lanes add into LDS counters (one LDS
// atomic per distinct bin per wave), the block adds its totals to the
// global counters once at the end.  s_cnt holds kMaxBins ints, zeroed by
// block_count_begin; every thread must call begin and end.
constexpr int kMaxBins = 32;

__device__ __forceinline__ void block_count_begin(int32_t* s_cnt) {
  for (int i = threadIdx.x; i < kMaxBins; i += blockDim.x) s_cnt[i] = 0;
  __syncthreads();
}

__device__ __forceinline__ int32_t wave_count_lds(int32_t* s_cnt, int b, bool active) {
  int32_t rank = -1;
  unsigned long long pending = __ballot(active);
  while (pending) {
    const int lead = __ffsll(pending) - 1;
    const int bb = __shfl(b, lead, kWave);
    const unsigned long long mask = __ballot(active && b == bb);
    int32_t base = 0;
    if (lane_id() == lead) base = atomicAdd(&s_cnt[bb], (int32_t)__popcll(mask));
    base = __shfl(base, lead, kWave);
    if (active && b == bb) {
      rank = base + (int32_t)__popcll(mask & ((1ull << lane_id()) - 1));
      active = false;
    }
    pending &= ~mask;
  }
  return rank;
}

__device__ __forceinline__ void block_count_end(int32_t* s_cnt, int32_t* counters, int nbins) {
  __syncthreads();
  for (int i = threadIdx.x; i < nbins; i += blockDim.x)
    if (s_cnt[i]) atomicAdd(&counters[i], s_cnt[i]);
}

// Per-row problem tables of a batch (one binary search per row, once per
// plan): A row -> (column base, width), B row -> column base.
__global__ __launch_bounds__(256) void k_prob_rows(Probs pr, int64_t m, int64_t mB, int2* arow_cols,
                                                   int32_t* brow_lo) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m || i < mB;
       i += (int64_t)gridDim.x * blockDim.x) {
    if (i < m) {
      const int pi = find_prob(pr.arow, pr.n, i);
      arow_cols[i] = make_int2((int32_t)pr.ccol[pi], (int32_t)(pr.ccol[pi + 1] - pr.ccol[pi]));
    }
    if (i < mB) brow_lo[i] = (int32_t)pr.ccol[find_prob(pr.brow, pr.n, i)];
  }
}

__device__ __forceinline__ int64_t row_width(const Probs& pr, int64_t row) {
  int64_t lo, width;
  prob_cols(pr, row, lo, width);
  return width;
}

// Symbolic route of every row.  Rows without products get nnz 0 here.
// Rows too long for an LDS table whose products are few next to their
// column range (f < ghash_ratio * words) take the global-hash route: a
// bitmap pass would walk the whole range for them.  Its table bytes are
// summed into *ghash_bytes.
__global__ __launch_bounds__(256) void k_classify_sym(int64_t m, const int64_t* row_flp, Probs pr,
                                                      SymParams sp, uint8_t* sbin, int64_t* rownz,
                                                      int32_t* counts, unsigned long long* ghash_bytes) {
  __shared__ int32_t s_cnt[kMaxBins];
  __shared__ int64_t s_w[256 / kWave];
  block_count_begin(s_cnt);
  int64_t gb = 0;
  for (int64_t i0 = blockIdx.x * (int64_t)blockDim.x; i0 < m; i0 += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = i0 + threadIdx.x;
    const bool valid = i < m;
    int b = SB_NONE;
    if (valid) {
      const int64_t f = row_flp[i];
      if (f == 0) {
        rownz[i] = 0;
      } else {
        const int64_t words = (row_width(pr, i) + 31) >> 5;
        if (f > sp.sym_hash_max && (double)f < sp.ghash_ratio * (double)words) {
          b = SB_GHASH;
          gb += 4 * ghash_slots(f);
        } else if (f > sp.sym_hash_max || (double)f >= sp.bitmap_ratio * (double)words) {
          b = SB_BITMAP;
        } else {
          const int64_t s = next_pow2_i64(f + ((f + 1) >> 1));
          b = s <= 64 ? 0 : ilog2_pow2(s) - 6;
        }
      }
      sbin[i] = (uint8_t)b;
    }
    (void)wave_count_lds(s_cnt, b, b != SB_NONE);
  }
  block_count_end(s_cnt, counts, kSymBins);
  const int64_t t = block_sum<256, int64_t>(gb, s_w);
  if (threadIdx.x == 0 && t) atomicAdd(ghash_bytes, (unsigned long long)t);
}

// Numeric route of every row (needs C's row pointers).  Rows of the
// global-hash symbolic route that neither ESC nor an LDS table takes stay
// in global hash tables; their region bytes are summed into *ghash_bytes.
__global__ __launch_bounds__(256) void k_classify_num(int64_t m, const int64_t* row_flp,
                                                      const int64_t* rp_c, const uint8_t* sbin,
                                                      Probs pr, NumParams np, uint8_t* nbin,
                                                      int32_t* counts, unsigned long long* ghash_bytes) {
  __shared__ int32_t s_cnt[kMaxBins];
  __shared__ int64_t s_w[256 / kWave];
  block_count_begin(s_cnt);
  int64_t gb = 0;
  for (int64_t i0 = blockIdx.x * (int64_t)blockDim.x; i0 < m; i0 += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = i0 + threadIdx.x;
    const bool valid = i < m;
    int b = NB_NONE;
    if (valid) {
      const int64_t f = row_flp[i];
      const int64_t z = rp_c[i + 1] - rp_c[i];
      if (f > 0 && z > 0) {
        const bool bitmap = sbin[i] == SB_BITMAP;
        const double width = (double)row_width(pr, i);
        if (bitmap && (z > np.num_hash_max || (double)z >= np.dense_density * width)) {
          b = NB_DENSE;
        } else if (np.mode != 1 && f <= np.esc_max &&
                   (np.mode == 2 || (double)f / (double)z < np.acc_thres)) {
          b = kNumHashBins + esc_bin_of(f);
        } else if (z <= np.num_hash_max || sbin[i] < SB_BITMAP) {
          // LDS-hash symbolic rows have no chunk counts for the dense route;
          // their z <= f <= sym_hash_max fits an LDS table whatever
          // num_hash_max is
          const int64_t s = next_pow2_i64(z + ((z + 1) >> 1));
          b = s <= 64 ? 0 : ilog2_pow2(s) - 6;
        } else if (sbin[i] == SB_GHASH) {
          b = NB_GHASH;
          gb += ghash_num_bytes(z, np.vbytes);
        } else {
          b = NB_DENSE;
        }
      }
      nbin[i] = (uint8_t)b;
    }
    (void)wave_count_lds(s_cnt, b, b != NB_NONE);
  }
  block_count_end(s_cnt, counts, kNumBins);
  const int64_t t = block_sum<256, int64_t>(gb, s_w);
  if (threadIdx.x == 0 && t) atomicAdd(ghash_bytes, (unsigned long long)t);
}

// Rows grouped by bin: bin b occupies list[sum(counts[0..b)) ...).
__global__ __launch_bounds__(256) void k_scatter_bins(int64_t m, const uint8_t* bin, int nbins,
                                                      const int32_t* counts, int32_t* cursors,
                                                      int32_t* list) {
  __shared__ int32_t s_off[kMaxBins];
  __shared__ int32_t s_cnt[kMaxBins];
  __shared__ int32_t s_base[kMaxBins];
  if (threadIdx.x == 0) {
    int32_t acc = 0;
    for (int b = 0; b < nbins; ++b) { s_off[b] = acc; acc += counts[b]; }
  }
  // per tile of 256 rows: ranks inside the block, then one global cursor
  // claim per bin and block
  for (int64_t i0 = blockIdx.x * (int64_t)blockDim.x; i0 < m; i0 += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = i0 + threadIdx.x;
    const int b = i < m ? (int)bin[i] : 0xff;
    const bool active = b != 0xff;
    block_count_begin(s_cnt);
    const int32_t r = wave_count_lds(s_cnt, b, active);
    __syncthreads();
    if ((int)threadIdx.x < nbins) s_base[threadIdx.x] = s_cnt[threadIdx.x] ? atomicAdd(&cursors[threadIdx.x], s_cnt[threadIdx.x]) : 0;
    __syncthreads();
    if (active) list[s_off[b] + s_base[b] + r] = (int32_t)i;
  }
}

__device__ __forceinline__ void bin_range(const int32_t* counts, int bin, int32_t& start, int32_t& cnt) {
  int32_t s = 0;
  for (int b = 0; b < bin; ++b) s += counts[b];
  start = s;
  cnt = counts[bin];
}

// ----------------------------------------------------------------- scan

constexpr int kScanT = 256;
constexpr int kScanIpt = 8;
constexpr int kScanTile = kScanT * kScanIpt;

__global__ __launch_bounds__(kScanT) void k_scan_tiles(const int64_t* in, int64_t n, int64_t* tsum) {
  __shared__ int64_t s_w[kScanT / kWave];
  const int64_t base = blockIdx.x * (int64_t)kScanTile + threadIdx.x * kScanIpt;
  int64_t acc = 0;
#pragma unroll
  for (int j = 0; j < kScanIpt; ++j) acc += (base + j < n) ? in[base + j] : 0;
  const int64_t t = block_sum<kScanT, int64_t>(acc, s_w);
  if (threadIdx.x == 0) tsum[blockIdx.x] = t;
}

__global__ __launch_bounds__(1024) void k_scan_top(int64_t* tsum, int64_t nt) {
  __shared__ int64_t s_w[16];
  int64_t carry = 0;
  for (int64_t b = 0; b < nt; b += 1024) {
    const int64_t i = b + threadIdx.x;
    const int64_t v = i < nt ? tsum[i] : 0;
    int64_t tot;
    const int64_t ex = block_excl_scan<1024, int64_t>(v, s_w, tot);
    if (i < nt) tsum[i] = carry + ex;
    carry += tot;
  }
}

__global__ __launch_bounds__(kScanT) void k_scan_apply(const int64_t* in, int64_t* out, int64_t n,
                                                       const int64_t* tsum) {
  __shared__ int64_t s_w[kScanT / kWave];
  const int64_t base = blockIdx.x * (int64_t)kScanTile + threadIdx.x * kScanIpt;
  int64_t v[kScanIpt];
  int64_t acc = 0;
#pragma unroll
  for (int j = 0; j < kScanIpt; ++j) {
    v[j] = (base + j < n) ? in[base + j] : 0;
    acc += v[j];
  }
  int64_t tot;
  int64_t run = block_excl_scan<kScanT, int64_t>(acc, s_w, tot) + tsum[blockIdx.x];
#pragma unroll
  for (int j = 0; j < kScanIpt; ++j) {
    if (base + j < n) out[base + j] = run;
    run += v[j];
  }
}

// ------------------------------------------------------------- symbolic

// LDS-hash symbolic, one row per block; table of S int32 keys (the row has
// at most 2S/3 products).  Products are handed to threads in index order
// (x = tid, tid + T, ...: 64 consecutive products of a B row per load) and
// find their A nonzero without searching: the non-empty B segments of a
// window of T A nonzeros are numbered, each sets the bit of its first
// product in an S-bit map, and product x belongs to segment
// (set bits at positions <= x) - 1, read from one (bits, bits before) word.
#ifndef AMS_SYMB_HASH_U
#define AMS_SYMB_HASH_U 4
#endif
template <int S, int T>
__global__ __launch_bounds__(T) void k_symb_hash(Csr A, Csr B, const int32_t* list,
                                                 const int32_t* counts, int bin, int64_t* rownz,
                                                 int32_t* err) {
  constexpr int U = AMS_SYMB_HASH_U;
  constexpr int kWords = S / 32;
  constexpr int kWpl = kWords >= kWave ? kWords / kWave : 1;
  __shared__ int32_t keys[S];
  __shared__ uint2 wpre[kWords];
  __shared__ int64_t sbias[T];
  __shared__ int64_t s_w[T / kWave];
  constexpr int shift = 32 - __builtin_ctz(S);
  int32_t start, cnt;
  bin_range(counts, bin, start, cnt);
  for (int32_t idx = blockIdx.x; idx < cnt; idx += gridDim.x) {
    const int32_t row = list[start + idx];
    const int64_t p0 = A.rp[row], p1 = A.rp[row + 1];
    for (int s = threadIdx.x; s < S; s += T) keys[s] = kEmpty;
    int64_t mine = 0;
    for (int64_t w0 = p0; w0 < p1; w0 += T) {
      for (int i = threadIdx.x; i < kWords; i += T) wpre[i] = make_uint2(0u, 0u);
      const int64_t p = w0 + threadIdx.x;
      int64_t st = 0;
      int32_t len = 0;
      if (p < p1) {
        const int32_t k = A.col[p];
        st = B.rp[k];
        len = (int32_t)(B.rp[k + 1] - st);
      }
      // the scan's barrier also orders the clearing above before the bits
      int32_t o, ci, total, nne;
      block_seg_scan<T>(len, (uint2*)s_w, o, ci, total, nne);
      if (len > 0) {
        sbias[ci] = st - o;
        atomicOr(&wpre[o >> 5].x, 1u << (o & 31));
      }
      __syncthreads();
      if (wave_id() == 0) {
        uint32_t pc[kWpl];
        int32_t sum = 0;
#pragma unroll
        for (int q = 0; q < kWpl; ++q) {
          const int wi = lane_id() * kWpl + q;
          pc[q] = wi < kWords ? (uint32_t)__popc(wpre[wi].x) : 0u;
          sum += (int32_t)pc[q];
        }
        int32_t run = wave_incl_scan(sum) - sum;
#pragma unroll
        for (int q = 0; q < kWpl; ++q) {
          const int wi = lane_id() * kWpl + q;
          if (wi < kWords) wpre[wi].y = (uint32_t)run;
          run += (int32_t)pc[q];
        }
      }
      __syncthreads();
      for (int32_t x0 = 0; x0 < total; x0 += T * U) {
        int32_t col[U];
        bool ok[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int32_t x = x0 + u * T + (int32_t)threadIdx.x;
          ok[u] = x < total;
          const int32_t xc = ok[u] ? x : 0;
          const uint2 wp = wpre[xc >> 5];
          const int32_t seg = (int32_t)wp.y + __popc(wp.x & ((2u << (xc & 31)) - 1u)) - 1;
          const int64_t q = ok[u] ? sbias[seg < 0 ? 0 : seg] + xc : 0;
          col[u] = ok[u] ? B.col[q] : 0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
          if (ok[u]) {
            const int r = lds_insert_key(keys, shift, col[u]);
            if (r > 0) ++mine;
            if (r < 0) atomicOr(err, 8);
          }
        }
      }
      if (w0 + T < p1) __syncthreads();  // the next window rewrites the table
    }
    const int64_t tot = block_sum<T, int64_t>(mine, s_w);
    if (threadIdx.x == 0) rownz[row] = tot;
  }
}

// Global-hash symbolic: one block per row, distinct columns counted in a
// table of ghash_slots(f) int32 keys in global memory (the row's products
// are few next to its column range, so an LDS bitmap pass would mostly walk
// empty words, and the row is too long for an LDS table).  The restatement
// of the reference's per-row table sized from the row's products
// (src/ams/ams_utils.c:1600-1752) without an LDS bound.  Tables come from
// gtab by an atomic cursor (in keys); the host sized gtab from the classify
// kernel's sum.  Linear probing, insert by compare-and-swap.
__device__ __forceinline__ int ghash_insert_key(int32_t* tab, int lg, int32_t key) {
  const uint32_t mask = (1u << lg) - 1u;  // lg <= 31
  uint32_t h = lg == 0 ? 0u : ((uint32_t)key * kHashMul) >> (32 - lg);
  for (uint32_t probe = 0; probe <= mask; ++probe) {
    const int32_t old = atomicCAS(&tab[h], kEmpty, key);
    if (old == kEmpty) return 1;
    if (old == key) return 0;
    h = (h + 1) & mask;
  }
  return -1;
}

template <int T>
__global__ __launch_bounds__(T) void k_symb_ghash(Csr A, Csr B, const int64_t* row_flp, const int32_t* list,
                                                  const int32_t* counts, int bin, int64_t* rownz, int32_t* gtab,
                                                  unsigned long long* cursor, int32_t* err) {
  __shared__ ExpandLds<T> L;
  __shared__ int64_t s_w[T / kWave];
  __shared__ unsigned long long s_base;
  int32_t start, cnt;
  bin_range(counts, bin, start, cnt);
  const SegB seg{A.col, B.rp};
  for (int32_t idx = blockIdx.x; idx < cnt; idx += gridDim.x) {
    const int32_t row = list[start + idx];
    const int64_t S = ghash_slots(row_flp[row]);
    const int lg = ilog2_pow2(S);
    if (threadIdx.x == 0) s_base = atomicAdd(cursor, (unsigned long long)S);
    __syncthreads();
    int32_t* tab = gtab + s_base;
    for (int64_t i = threadIdx.x; i < S; i += T) tab[i] = kEmpty;
    __syncthreads();
    int64_t mine = 0;
    block_expand<T, 4>(A.rp[row], A.rp[row + 1], L, seg,
                       [&](int, int64_t, int64_t q, int64_t) { return B.col[q]; },
                       [&](int32_t key) {
                         const int r = ghash_insert_key(tab, lg, key);
                         if (r > 0) ++mine;
                         if (r < 0) atomicOr(err, 0x40);
                       });
    const int64_t tot = block_sum<T, int64_t>(mine, s_w);
    if (threadIdx.x == 0) rownz[row] = tot;
  }
}

// Bitmap rows of narrow problems with few products do not fill a block: a
// wave takes each of them, with a private LDS bitmap of the row's width and
// no block barriers.  k_symb_bitmap skips these rows when told (wave_rows)
// that this kernel runs; the host launches it whenever a problem of at most
// kWaveBitmapBits columns may exist (every batch, and single problems that
// narrow), so the two kernels always split the bin between them.
constexpr int kWaveBitmapBits = 16384;
constexpr int64_t kWaveBitmapFlops = 8192;

template <int T>
__global__ __launch_bounds__(T) void k_symb_bitmap_wave(Csr A, Csr B, Probs pr, const int64_t* row_flp,
                                                        const int32_t* list, const int32_t* counts, int bin,
                                                        int64_t* rownz, int32_t* coff, int R, int W) {
  constexpr int NW = T / kWave;
  __shared__ uint32_t s_bm[NW][kWaveBitmapBits / 32];
  __shared__ int32_t s_off[NW][kWave];
  __shared__ int64_t s_start[NW][kWave];
  uint32_t* bm = s_bm[wave_id()];
  int32_t start, cnt;
  bin_range(counts, bin, start, cnt);
  const SegB seg{A.col, B.rp};
  const int wchunk = W >> 5;
  for (int32_t h = blockIdx.x * NW + wave_id(); h < cnt; h += gridDim.x * NW) {
    const int32_t row = list[start + h];
    int64_t lo, width;
    prob_cols(pr, row, lo, width);
    if (width > kWaveBitmapBits || row_flp[row] > kWaveBitmapFlops) continue;
    const int words = (int)((width + 31) >> 5);
    for (int w = lane_id(); w < words; w += kWave) bm[w] = 0u;
    wave_lds_sync();
    wave_expand<8>(A.rp[row], A.rp[row + 1], s_off[wave_id()], s_start[wave_id()], seg,
                   [&](int, int64_t, int64_t q, int64_t) { return B.col[q]; },
                   [&](int32_t col) {
                     const uint32_t u = (uint32_t)((int64_t)col - lo);
                     const uint32_t bit = 1u << (u & 31);
                     if (!(bm[u >> 5] & bit)) atomicOr(&bm[u >> 5], bit);
                   });
    wave_lds_sync();
    if (pr.rowbits) {
      uint32_t* rb = pr.rowbits + (int64_t)h * pr.bwords;
      for (int w = lane_id(); w < words; w += kWave) stream_store(&rb[w], bm[w]);
    }
    int32_t* crow = coff + (int64_t)h * (R + 1);
    const int nch = (int)((width + W - 1) / W);
    int32_t running = 0;
    for (int c = 0; c < nch; ++c) {
      const int w0 = c * wchunk;
      const int w1 = (w0 + wchunk) < words ? (w0 + wchunk) : words;
      int32_t pc = 0;
      for (int w = w0 + lane_id(); w < w1; w += kWave) pc += __popc(bm[w]);
      pc = wave_sum(pc);
      if (lane_id() == 0) crow[c] = running;
      running += pc;
    }
    if (lane_id() == 0) {
      for (int c = nch; c <= R; ++c) crow[c] = running;
      rownz[row] = running;
    }
    wave_lds_sync();  // every lane has read bm before the next row clears it
  }
}

// Heavy rows of the bitmap bin.  A row with more than kPiece products
// (kPiece = max(kMinPiece, P / kMaxHeavy)) is split into q <= 64 pieces of
// equal A-nonzero ranges.  Blocks OR each piece's LDS bitmap into the row's
// global bitmap (gbm, words of the row's column range), and
// k_heavy_finish counts it per chunk.  Since every heavy row has more than
// P / kMaxHeavy products there are at most kMaxHeavy of them, and at most
// 2 * kMaxHeavy pieces.
constexpr int kMaxHeavy = 2048;
constexpr int kMaxPieces = 2 * kMaxHeavy;
constexpr int64_t kMinPiece = 1 << 20;
constexpr int kMaxPiecesPerRow = 64;

struct HeavyPiece {
  int32_t h;    // index in the bitmap bin
  int32_t hid;  // heavy row slot (its gbm row)
  int32_t i, q; // piece i of q
};

__global__ __launch_bounds__(256) void k_heavy_plan(const Probs pr, const int64_t* row_flp, const int32_t* list,
                                                    const int32_t* counts, int bin, const unsigned long long* flops,
                                                    uint8_t* hflag, int32_t* hrow, HeavyPiece* pieces,
                                                    int32_t* nheavy, int32_t* npieces) {
  int32_t start, cnt;
  bin_range(counts, bin, start, cnt);
  const int64_t per = (int64_t)(*flops / (unsigned long long)kMaxHeavy);
  const int64_t kPiece = per > kMinPiece ? per : kMinPiece;
  for (int32_t h = blockIdx.x * blockDim.x + threadIdx.x; h < cnt; h += gridDim.x * blockDim.x) {
    const int32_t row = list[start + h];
    const int64_t f = row_flp[row];
    if (f <= kPiece) {
      hflag[h] = 0;
      continue;
    }
    int64_t q64 = (f + kPiece - 1) / kPiece;
    const int32_t q = (int32_t)(q64 < kMaxPiecesPerRow ? q64 : kMaxPiecesPerRow);
    // the counts above bound both reservations; should a total of products
    // ever undercount, a row that does not fit stays whole and the slots it
    // took are marked empty (h = -1)
    const int32_t hid = atomicAdd(nheavy, 1);
    if (hid >= kMaxHeavy) {
      hflag[h] = 0;
      continue;
    }
    const int32_t base = atomicAdd(npieces, q);
    const bool fits = base + q <= kMaxPieces;
    hflag[h] = fits ? 1 : 0;
    hrow[hid] = fits ? h : -1;
    for (int32_t i = 0; i < q && base + i < kMaxPieces; ++i)
      pieces[base + i] = fits ? HeavyPiece{h, hid, i, q} : HeavyPiece{-1, -1, 0, 1};
  }
}

// Per heavy row: population of every W-column chunk of its global bitmap
// into coff (exclusive prefix, total at index R) and rownz; the bitmap is
// cleared for the next call.
template <int T>
__global__ __launch_bounds__(T) void k_heavy_finish(const Probs pr, const int32_t* list, const int32_t* counts,
                                                    int bin, const int32_t* hrow, const int32_t* nheavy,
                                                    uint32_t* gbm, int64_t gwords, int64_t* rownz, int32_t* coff,
                                                    int R, int W) {
  __shared__ int64_t s_w[T / kWave];
  int32_t start, cnt;
  bin_range(counts, bin, start, cnt);
  const int32_t nh = *nheavy < kMaxHeavy ? *nheavy : kMaxHeavy;
  const int wchunk = W >> 5;
  for (int32_t hid = blockIdx.x; hid < nh; hid += gridDim.x) {
    const int32_t h = hrow[hid];
    if (h < 0) continue;
    const int32_t row = list[start + h];
    int64_t lo, width;
    prob_cols(pr, row, lo, width);
    const int64_t words = (width + 31) >> 5;
    uint32_t* g = gbm + (int64_t)hid * gwords;
    int32_t* crow = coff + (int64_t)h * (R + 1);
    const int nch = (int)((width + W - 1) / W);
    int32_t running = 0;
    for (int c = 0; c < nch; ++c) {
      const int64_t w0 = (int64_t)c * wchunk;
      const int64_t w1 = (w0 + wchunk) < words ? (w0 + wchunk) : words;
      int64_t pc = 0;
      uint32_t* rb = pr.rowbits ? pr.rowbits + (int64_t)h * pr.bwords : nullptr;
      for (int64_t w = w0 + threadIdx.x; w < w1; w += T) {
        const uint32_t x = g[w];
        pc += __popc(x);
        if (rb) stream_store(&rb[w], x);
        g[w] = 0u;
      }
      const int64_t tot = block_sum<T, int64_t>(pc, s_w);
      if (threadIdx.x == 0) crow[c] = running;
      running += (int32_t)tot;
    }
    if (threadIdx.x == 0) {
      for (int c = nch; c <= R; ++c) crow[c] = running;
      rownz[row] = running;
    }
  }
}

// Sets the bits [base, base + pw) of the columns of the products of A
// nonzeros [p0, p1) in the LDS bitmap bm.  Hub columns recur: the bit is
// tested first, so that only unset bits take the (same-word serialised)
// atomic; a stride's bit reads all go out before the first test.
// Q32: A and B have fewer than 2^30 nonzeros and B fewer than 2^28 rows,
// so every byte offset into A's columns and B's row pointers and columns
// fits 32 bits and each load takes the scalar base plus a 32-bit vector
// offset.
// FULL: the pass covers the row's whole column range, so every product's
// column is inside it (the clamp only keeps LDS addresses in bounds for
// columns outside the problem), and lanes past the window's end, which
// re-read its last product, may set that product's bit again: OR is
// idempotent, so neither needs a mask.
#ifndef AMS_SYMB_FULL_GROUP
#define AMS_SYMB_FULL_GROUP 8
#endif
constexpr int kSymFullGroup = AMS_SYMB_FULL_GROUP;
static_assert(kUSymBitmap % kSymFullGroup == 0, "whole groups");
template <int T, bool Q32, bool FULL, class Seg>
__device__ __forceinline__ void symb_bitmap_expand(int64_t p0, int64_t p1, ExpandLds<T>& L, const Seg& seg,
                                                   const Csr& B, uint32_t* bm, int64_t base, int32_t pw) {
  block_expand_batch<T, kUSymBitmap>(
      p0, p1, L, seg,
      [&](int, int64_t, int64_t q, int64_t) {
        if constexpr (Q32) {
          const uint32_t boff = (uint32_t)q << 2;
          return *(const int32_t*)((const char*)B.col + boff);
        } else {
          return B.col[q];
        }
      },
      [&](int32_t (&col)[kUSymBitmap], bool (&ok)[kUSymBitmap]) {
        if constexpr (FULL) {
          // groups of kSymFullGroup reads, then their ORs: the ORs order the
          // next group's reads after them, which bounds the live registers
#pragma unroll
          for (int h = 0; h < kUSymBitmap; h += kSymFullGroup) {
            uint32_t u[kSymFullGroup], w[kSymFullGroup];
#pragma unroll
            for (int j = 0; j < kSymFullGroup; ++j) {
              u[j] = min((uint32_t)(col[h + j] - (int32_t)base), (uint32_t)pw - 1u);
              w[j] = bm[u[j] >> 5];
            }
#pragma unroll
            for (int j = 0; j < kSymFullGroup; ++j) {
              const uint32_t bit = 1u << (u[j] & 31);
              if (!(w[j] & bit)) atomicOr(&bm[u[j] >> 5], bit);
            }
          }
          return;
        }
        uint32_t u[kUSymBitmap], w[kUSymBitmap];
#pragma unroll
        for (int j = 0; j < kUSymBitmap; ++j) {
          u[j] = (uint32_t)((int64_t)col[j] - base);
          const bool in = ok[j] && u[j] < (uint32_t)pw;
          u[j] = in ? u[j] : 0u;
          w[j] = in ? bm[u[j] >> 5] : ~0u;
        }
#pragma unroll
        for (int j = 0; j < kUSymBitmap; ++j) {
          const uint32_t bit = 1u << (u[j] & 31);
          if (!(w[j] & bit)) atomicOr(&bm[u[j] >> 5], bit);
        }
      });
}

// LDS-bitmap symbolic for wide rows, one row per block.  The row's output
// columns [0, width) are covered in passes of Wb bits; the population of
// every W-column chunk goes to coff (exclusive prefix, total at index R).
// Dynamic LDS: Wb/32 words of bitmap + Wb/W chunk counters.
// Eight waves per SIMD (64 VGPRs): two 1024-thread blocks share a CU's LDS,
// twice the waves of 512-thread blocks, 15% faster on G500 s19 ef32.
#ifndef AMS_SYMB_WPE
#define AMS_SYMB_WPE 8
#endif
constexpr int kSymGrab = 4;
// SW > 0: the bitmap and chunk counters live in a static LDS array of SW
// words (its address is known when compiling, so the word offsets need no
// base register), used by 1024-thread blocks, which two waves-limited
// blocks per CU leave 80 KB of LDS each.
constexpr int kSymStaticWords = 17152;
// FULL: Wb covers the widest problem, so every row takes one pass of its
// whole column range (pw == width).
template <int T, bool Q32, int SW, bool FULL>
__global__ __launch_bounds__(T) __attribute__((amdgpu_waves_per_eu(AMS_SYMB_WPE))) void k_symb_bitmap(Csr A, Csr B, Probs pr, const int64_t* row_flp,
                                                   const int32_t* list, const int32_t* counts, int bin,
                                                   int64_t* rownz, int32_t* coff, int R, int W, int Wb,
                                                   const uint8_t* hflag, const HeavyPiece* pieces,
                                                   const int32_t* npieces, uint32_t* gbm, int64_t gwords,
                                                   unsigned long long* row_ticket, int wave_rows) {
  extern __shared__ uint32_t dyn[];
  __shared__ uint32_t s_bm[SW > 0 ? SW : 1];
  uint32_t* bm = SW > 0 ? s_bm : dyn;
  int32_t* ccount = (int32_t*)(bm + (Wb >> 5));
  __shared__ ExpandLds<T> L;
  int32_t start, cnt;
  bin_range(counts, bin, start, cnt);
  const std::conditional_t<Q32, SegB32, SegB> seg{A.col, B.rp};
  const int wchunk = W >> 5;
  // the bitmap is cleared once here; afterwards each pass's reader (the
  // pieces' OR into the global bitmap, the rows' chunk count) zeroes every
  // word it reads, and those are the only words a pass can set
  for (int w = threadIdx.x; w < (Wb >> 5); w += T) bm[w] = 0u;
  __syncthreads();
  // the pieces of heavy rows first (they are the longest work items): each
  // ORs its bitmap into the row's global bitmap, word by non-zero word
  const int32_t np = hflag ? (*npieces < kMaxPieces ? *npieces : kMaxPieces) : 0;
  for (int32_t pi = blockIdx.x; pi < np; pi += gridDim.x) {
    const HeavyPiece hp = pieces[pi];
    if (hp.h < 0) continue;
    // uniform: in scalar registers, so that row, lo and width do not take
    // vector registers (and spill) across the expansion
    const int32_t row = __builtin_amdgcn_readfirstlane(list[start + hp.h]);
    int64_t lo, width;
    prob_cols(pr, row, lo, width);
    lo = (int64_t)uniform_u64((unsigned long long)lo);
    width = (int64_t)uniform_u64((unsigned long long)width);
    const int64_t p0 = (int64_t)uniform_u64((unsigned long long)A.rp[row]);
    const int64_t deg = (int64_t)uniform_u64((unsigned long long)A.rp[row + 1]) - p0;
    const int64_t q0 = p0 + deg * hp.i / hp.q, q1 = p0 + deg * (hp.i + 1) / hp.q;
    uint32_t* g = gbm + (int64_t)hp.hid * gwords;
    for (int64_t c0 = 0; c0 < width; c0 += Wb) {
      const int32_t pw = (int32_t)((width - c0) < Wb ? (width - c0) : Wb);
      const int32_t words = (pw + 31) >> 5;
      const int64_t base = lo + c0;
      symb_bitmap_expand<T, Q32, FULL>(q0, q1, L, seg, B, bm, base, pw);
      // the thread index passes through an empty asm statement, so that the
      // compiler cannot hoist gbm + 4 * threadIdx.x out of the loops: held
      // across every row's product loop, that address spilled
      int tid = (int)threadIdx.x;
      asm volatile("" : "+v"(tid));
      for (int w = tid; w < words; w += T) {
        const uint32_t x = bm[w];
        bm[w] = 0u;
        if (x) atomicOr(&g[(c0 >> 5) + w], x);
      }
      __syncthreads();
    }
  }
  // then the rows, up to kSymGrab tickets at a time from the global counter
  // (the bin list's order, and so the rows' sizes, vary from run to run); one
  // at a time when the bin has few rows per block
  __shared__ int32_t s_next;
  // thread 0's batch of tickets, kept in LDS: in registers it would stay
  // live (and spill) across every row's product loop
  __shared__ unsigned long long s_qtk, s_qend;
  const unsigned long long per_block = (unsigned long long)cnt / ((unsigned long long)gridDim.x * 16ull);
  const unsigned long long grab = per_block < 1 ? 1 : (per_block > kSymGrab ? kSymGrab : per_block);
  if (threadIdx.x == 0) s_qtk = s_qend = 0;
  for (;;) {
    if (threadIdx.x == 0) {
      unsigned long long q_tk = s_qtk;
      if (q_tk == s_qend) {
        q_tk = atomicAdd(row_ticket, grab);
        s_qend = q_tk + grab;
      }
      s_next = q_tk < (unsigned long long)cnt ? (int32_t)q_tk : cnt;
      s_qtk = q_tk + 1;
    }
    __syncthreads();
    const int32_t h = __builtin_amdgcn_readfirstlane(s_next);
    __syncthreads();  // every thread has read s_next before thread 0 writes it again
    if (h >= cnt) break;
    // uniform values in scalar registers, as for the pieces above
    const int32_t row = __builtin_amdgcn_readfirstlane(list[start + h]);
    int64_t lo, width;
    prob_cols(pr, row, lo, width);
    lo = (int64_t)uniform_u64((unsigned long long)lo);
    width = (int64_t)uniform_u64((unsigned long long)width);
    // k_symb_bitmap_wave's rows, when the host launches it (wave_rows)
    if (wave_rows && width <= kWaveBitmapBits && row_flp[row] <= kWaveBitmapFlops) continue;
    if (hflag && hflag[h]) continue;  // split into pieces above
    int32_t running = 0;
    int32_t* crow = coff + (int64_t)h * (R + 1);
    for (int64_t c0 = 0; c0 < width; c0 += Wb) {
      const int32_t pw = (int32_t)((width - c0) < Wb ? (width - c0) : Wb);
      const int32_t words = (pw + 31) >> 5;
      const int32_t nch = (pw + W - 1) / W;
      for (int c = threadIdx.x; c < nch; c += T) ccount[c] = 0;
      __syncthreads();
      const int64_t base = lo + c0;
      symb_bitmap_expand<T, Q32, FULL>((int64_t)uniform_u64((unsigned long long)A.rp[row]),
                                       (int64_t)uniform_u64((unsigned long long)A.rp[row + 1]), L, seg, B, bm, base,
                                       pw);
      // the row's column bitmap (for the numeric phase) and the populations
      // of its chunks
      uint32_t* rb = pr.rowbits ? pr.rowbits + (int64_t)h * pr.bwords + (c0 >> 5) : nullptr;
      if (wchunk >= kWave) {
        // wave v takes the words [v*per, (v+1)*per) (per a multiple of 64) in
        // coalesced 64-word steps, stores them to rb and sums their
        // populations per lane; the sum goes to the chunk's counter (one wave
        // sum, one LDS add) only when the range leaves the chunk, and a step of
        // 64 <= wchunk words crosses at most one chunk boundary
        constexpr int kNw = T / kWave;
        const int per = ((words + kNw - 1) / kNw + kWave - 1) / kWave * kWave;
        const int k0 = __builtin_amdgcn_readfirstlane(wave_id()) * per;
        const int k1 = k0 + per < words ? k0 + per : words;
        int cur = k0 / wchunk;
        int cend = (cur + 1) * wchunk;  // first word of the next chunk
        int acc = 0;
        auto flush = [&](int c, int v) {
          const int tot = wave_sum(v);
          if (lane_id() == 0 && tot) atomicAdd(&ccount[c], tot);
        };
        for (int k = k0; k < k1; k += kWave) {
          if (k >= cend) {  // the step starts the next chunk
            flush(cur, acc);
            acc = 0;
            ++cur;
            cend += wchunk;
          }
          const int w = k + lane_id();
          uint32_t x = 0u;
          if (w < k1) {
            x = bm[w];
            bm[w] = 0u;
            if (rb) stream_store(&rb[w], x);
          }
          const int pc = __popc(x);
          if (k + kWave > cend) {  // the step's words from cend on are the next chunk's
            const int nx = w >= cend ? pc : 0;
            flush(cur, acc + pc - nx);
            acc = nx;
            ++cur;
            cend += wchunk;
          } else {
            acc += pc;
          }
        }
        if (k0 < k1) flush(cur, acc);
      } else {
        if (rb) {
          int tid = (int)threadIdx.x;  // opaque, as above: no hoisted rowbits + 4 * threadIdx.x
          asm volatile("" : "+v"(tid));
          for (int w = tid; w < words; w += T) stream_store(&rb[w], bm[w]);
        }
        // words_pass rounded up to whole waves so every lane joins the vote
        for (int w0 = 0; w0 < words; w0 += T) {
          const int w = w0 + threadIdx.x;
          uint32_t x = 0u;
          if (w < words) {
            x = bm[w];
            bm[w] = 0u;
          }
          const int pc = __popc(x);
          const int cidx = w < words ? w / wchunk : -1;
          const int c_lane0 = __shfl(cidx, 0, kWave);
          if (__all(cidx == c_lane0)) {
            const int tot = wave_sum(pc);
            if (lane_id() == 0 && tot && c_lane0 >= 0) atomicAdd(&ccount[c_lane0], tot);
          } else if (pc) {
            atomicAdd(&ccount[cidx], pc);
          }
        }
      }
      __syncthreads();
      if (threadIdx.x == 0) {
        const int c_first = (int)(c0 / W);
        for (int c = 0; c < nch; ++c) {
          crow[c_first + c] = running;
          running += ccount[c];
        }
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      const int rrow = (int)((width + W - 1) / W);
      for (int c = rrow; c <= R; ++c) crow[c] = running;
      rownz[row] = running;
    }
    __syncthreads();
  }
}

// Blocked symbolic for wide rows: every wave takes (row h, chunk c)
// morsels, chunk-major, and counts the distinct columns of C's row h in
// chunk c with a W-bit LDS bitmap of its own, reading the uint16 columns of
// the column-blocked B.  Concurrent waves share one column block of B, which
// stays in L2.  Output: ccnt[h * R + c].  Dynamic LDS per wave:
//   uint32 bitmap[W/32] | int32 off[64] | int64 start[64]
template <int T>
__global__ __launch_bounds__(T) void k_symb_blocked(Csr A, int64_t mB, const int64_t* bptr, const uint16_t* bcol,
                                                    const int32_t* list, const int32_t* counts, int bin, int R,
                                                    int W, int32_t* ccnt, unsigned long long* ticket) {
  extern __shared__ __align__(16) unsigned char dyn_raw[];
  const int words = W >> 5;
  const size_t per_wave = ((size_t)words * 4 + 64 * 4 + 64 * 8 + 15) & ~(size_t)15;
  unsigned char* mine = dyn_raw + per_wave * wave_id();
  uint32_t* bm = (uint32_t*)mine;
  int32_t* w_off = (int32_t*)(mine + (size_t)words * 4);
  int64_t* w_start = (int64_t*)(mine + (size_t)words * 4 + 64 * 4);
  int32_t start, H;
  bin_range(counts, bin, start, H);
  const unsigned long long ntask = (unsigned long long)H * (unsigned long long)R;
  const int64_t ld = mB + 1;
  const unsigned long long waves = (unsigned long long)gridDim.x * (T / kWave);
  const unsigned long long per_wave_tasks = ntask / (waves * 64ull);
  const unsigned long long grab = per_wave_tasks < 1 ? 1 : (per_wave_tasks > 16 ? 16 : per_wave_tasks);
  for (int w = lane_id(); w < words; w += kWave) bm[w] = 0u;
  unsigned long long tk = 0, tk_end = 0;
  for (;;) {
    if (tk >= tk_end) {
      unsigned long long got = 0;
      if (lane_id() == 0) got = atomicAdd(ticket, grab);
      tk = __shfl(got, 0, kWave);
      tk_end = tk + grab;
    }
    const unsigned long long cur = tk++;
    if (cur >= ntask) break;
    const int c = (int)(cur / (unsigned long long)H);
    const int32_t h = (int32_t)(cur - (unsigned long long)c * H);
    const int32_t row = list[start + h];
    const int64_t* bp = bptr + (int64_t)c * ld;
    wave_lds_sync();
    wave_expand<kUSymBitmap>(A.rp[row], A.rp[row + 1], w_off, w_start,
                             [&](int64_t p, int64_t& st, int32_t& len) {
                               const int32_t k = A.col[p];
                               st = bp[k];
                               len = (int32_t)(bp[k + 1] - st);
                             },
                             [&](int, int64_t, int64_t q, int64_t) { return (uint32_t)bcol[q]; },
                             [&](uint32_t u) { atomicOr(&bm[u >> 5], 1u << (u & 31)); });
    int32_t pc = 0;
    for (int w = lane_id(); w < words; w += kWave) {
      pc += __popc(bm[w]);
      bm[w] = 0u;
    }
    pc = wave_sum(pc);
    if (lane_id() == 0) ccnt[(int64_t)h * R + c] = pc;
  }
}

// Per wide row: coff = exclusive prefix of its chunk counts, total into
// rownz and coff[R].
__global__ __launch_bounds__(256) void k_coff_rows(const int32_t* list, const int32_t* counts, int bin, int R,
                                                   const int32_t* ccnt, int32_t* coff, int64_t* rownz) {
  int32_t start, H;
  bin_range(counts, bin, start, H);
  for (int64_t h = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; h < H; h += (int64_t)gridDim.x * blockDim.x) {
    int32_t run = 0;
    const int32_t* cr = ccnt + h * R;
    int32_t* o = coff + h * (R + 1);
    for (int c = 0; c < R; ++c) {
      o[c] = run;
      run += cr[c];
    }
    o[R] = run;
    rownz[list[start + h]] = run;
  }
}

// -------------------------------------------------------------- numeric

// LDS hash accumulate of one row per block (S slots of int32 key + V),
// then the keys are sorted and each is looked up again for its value, so
// the row is written with ascending columns in coalesced stores.
template <int S, int T, typename V>
__global__ __launch_bounds__(T) void k_numc_hash(Csr A, Csr B, const int32_t* list,
                                                 const int32_t* counts, int bin, const int64_t* rp_c,
                                                 int32_t* col_c, V* val_c, int64_t cap, int32_t* err) {
  __shared__ int32_t keys[S];
  __shared__ V vals[S];
  __shared__ int32_t sk[S];
  __shared__ ExpandLds<T> L;
  __shared__ V aval[T];
  __shared__ int32_t s_n;
  constexpr int shift = 32 - __builtin_ctz(S);
  if (rp_c[A.m] > cap) return;
  int32_t start, cnt;
  bin_range(counts, bin, start, cnt);
  const V* av = (const V*)A.val;
  const V* bv = (const V*)B.val;
  for (int32_t idx = blockIdx.x; idx < cnt; idx += gridDim.x) {
    const int32_t row = list[start + idx];
    const int64_t off = rp_c[row];
    const int32_t z = (int32_t)(rp_c[row + 1] - off);
    for (int s = threadIdx.x; s < S; s += T) { keys[s] = kEmpty; vals[s] = V(0); }
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    block_expand<T, 4>(A.rp[row], A.rp[row + 1], L,
                    [&](int64_t p, int64_t& st, int32_t& len) {
                      const int32_t k = A.col[p];
                      st = B.rp[k];
                      len = (int32_t)(B.rp[k + 1] - st);
                      aval[threadIdx.x] = av[p];
                    },
                    [&](int j, int64_t, int64_t q, int64_t) { return Gather<V>{B.col[q], j, bv[q]}; },
                    [&](Gather<V> g) {
                      if (!lds_insert_add<V>(keys, vals, shift, g.key, aval[g.j] * g.val)) atomicOr(err, 16);
                    });
    for (int s0 = 0; s0 < S; s0 += T) {
      const int s = s0 + threadIdx.x;
      const bool occ = s < S && keys[s] != kEmpty;
      const unsigned long long mask = __ballot(occ);
      int32_t base = 0;
      if (lane_id() == 0 && mask) base = atomicAdd(&s_n, (int32_t)__popcll(mask));
      base = __shfl(base, 0, kWave);
      if (occ) sk[base + __popcll(mask & ((1ull << lane_id()) - 1))] = keys[s];
    }
    __syncthreads();
    const int32_t n = (int32_t)next_pow2_i64(z);
    if (threadIdx.x == 0 && s_n != z) atomicOr(err, 1);
    const int32_t zz = s_n < z ? s_n : z;  // never sort unset keys
    for (int i = zz + threadIdx.x; i < n; i += T) sk[i] = 0x7fffffff;
    __syncthreads();
    block_bitonic_sort<T, (S / 2 + T - 1) / T, int32_t>(sk, n);
    for (int r = threadIdx.x; r < z; r += T) {
      const int32_t key = sk[r];
      const uint32_t slot = lds_find(keys, shift, key);
      stream_store(&col_c[off + r], key);
      stream_store(&val_c[off + r], slot < (uint32_t)S ? vals[slot] : V(0));
    }
    __syncthreads();
  }
}

// ESC of one row per block for columns too wide for 32-bit (column, rank)
// keys, and the bucket sort: expand the row's f <= E products into LDS
// (column and value, indexed by the product's rank in A-row order), sort the
// ranks by (column, rank), then every run of equal columns is summed left to
// right by the thread that owns its head.  The rank keeps equal columns in
// A-row order, so each value is the same left-to-right sum of separately
// rounded products as the reference's.  The sort is a bitonic sort of
// (column << 32 | rank) words in registers (esc_sort 0 or 2) or a bucket
// sort of a rank permutation (esc_sort 1); both share one LDS region:
//   bucket:  cnt int32[E] | key uint32[E] | perm uint16[E]
//   bitonic: sk uint64[E] (over cnt and key)
// ESC expansion windows of at most 128 A nonzeros: for E = 2048 the block
// then needs 26.2 KB of LDS instead of 27.7 KB.
template <int T>
constexpr int kEscWin = T < 128 ? T : 128;

template <int E, int T, typename V>
__global__ __launch_bounds__(T) void k_numc_esc(Csr A, Csr B, const int32_t* list,
                                                const int32_t* counts, int bin, const int64_t* row_flp,
                                                const int64_t* rp_c, int32_t* col_c, V* val_c,
                                                int64_t cap, int esc_sort, int32_t* err) {
  static_assert(E <= 65536, "ranks are uint16");
  constexpr int IPT = (E + T - 1) / T;
  __shared__ __align__(16) unsigned char region[E * 10];
  int32_t* cnt = (int32_t*)region;
  uint32_t* key = (uint32_t*)(region + 4 * E);
  uint16_t* perm = (uint16_t*)(region + 8 * E);  // bucket sort only
  uint64_t* sk = (uint64_t*)region;             // 64-bit sorted keys
  __shared__ V vals[E];
  __shared__ ExpandLds<T, kEscWin<T>> L;  // rows of this bin have few A nonzeros
  __shared__ int64_t s_w[T / kWave];
  __shared__ uint32_t s_min[T / kWave], s_max[T / kWave];
  __shared__ int32_t s_flag;
  if (rp_c[A.m] > cap) return;
  int32_t start, cnt_rows;
  bin_range(counts, bin, start, cnt_rows);
  const V* av = (const V*)A.val;
  const V* bv = (const V*)B.val;
  for (int32_t idx = blockIdx.x; idx < cnt_rows; idx += gridDim.x) {
    const int32_t row = list[start + idx];
    const int32_t f = (int32_t)row_flp[row];
    const int64_t off = rp_c[row];
    const int32_t z = (int32_t)(rp_c[row + 1] - off);
    block_expand<T, 4>(A.rp[row], A.rp[row + 1], L,
                       [&](int64_t p, int64_t& st, int32_t& len) {
                         const int32_t k = A.col[p];
                         st = B.rp[k];
                         len = (int32_t)(B.rp[k + 1] - st);
                       },
                       [&](int, int64_t p, int64_t q, int64_t t) {
                         return RankGatherA<V>{(int32_t)t, B.col[q], av[p], bv[q]};
                       },
                       [&](RankGatherA<V> r) {
                         key[r.rank] = (uint32_t)r.key;
                         vals[r.rank] = r.a * r.val;
                       });
    const int nb = (int)next_pow2_i64(f);
    if (esc_sort != 1) {
      // (column, rank) keys sorted in registers: packed in 32 bits as
      // column << log2(E) | rank when the columns allow, else 64 bits
      static_assert(T * IPT == E, "the bin fills the threads' registers");
      const int i0 = threadIdx.x * IPT;
      {
        uint64_t kr[IPT];
#pragma unroll
        for (int r = 0; r < IPT; ++r)
          kr[r] = i0 + r < f ? (((uint64_t)key[i0 + r] << 32) | (uint64_t)(i0 + r)) : ~0ull;
        __syncthreads();
        reg_bitonic_sort<T, IPT>(kr, sk);
#pragma unroll
        for (int r = 0; r < IPT; ++r) sk[i0 + r] = kr[r];
      }
      __syncthreads();
    } else {
      uint32_t kmin = 0xffffffffu, kmax = 0u;
      for (int i = threadIdx.x; i < f; i += T) {
        kmin = min(kmin, key[i]);
        kmax = max(kmax, key[i]);
      }
#pragma unroll
      for (int d = kWave / 2; d > 0; d >>= 1) {
        kmin = min(kmin, (uint32_t)__shfl_xor((int)kmin, d, kWave));
        kmax = max(kmax, (uint32_t)__shfl_xor((int)kmax, d, kWave));
      }
      if (lane_id() == 0) { s_min[wave_id()] = kmin; s_max[wave_id()] = kmax; }
      __syncthreads();
      kmin = s_min[0];
      kmax = s_max[0];
      for (int w = 1; w < T / kWave; ++w) { kmin = min(kmin, s_min[w]); kmax = max(kmax, s_max[w]); }
      block_sort_perm<T>(key, perm, cnt, f, nb, kmin, (uint64_t)(kmax - kmin) + 1, &s_flag);
    }
    // sorted element e: column and rank
    auto col_at = [&](int e) -> uint32_t {
      return esc_sort != 1 ? (uint32_t)(sk[e] >> 32) : key[perm[e]];
    };
    auto rank_at = [&](int e) -> uint32_t {
      return esc_sort != 1 ? (uint32_t)sk[e] : (uint32_t)perm[e];
    };
    const int32_t ipt = nb >= T ? nb / T : 1;
    const int32_t e0 = threadIdx.x * ipt;
    const int32_t e1 = (e0 + ipt) < f ? (e0 + ipt) : f;
    int64_t heads = 0;
    for (int32_t e = e0; e < e1; ++e) heads += (e == 0) || (col_at(e) != col_at(e - 1));
    int64_t tot;
    int64_t rank = block_excl_scan<T, int64_t>(heads, s_w, tot);
    if (threadIdx.x == 0 && tot != z) atomicOr(err, 2);
    for (int32_t e = e0; e < e1; ++e) {
      const uint32_t k = col_at(e);
      if (e != 0 && k == col_at(e - 1)) continue;
      V sum = vals[rank_at(e)];
      for (int32_t g = e + 1; g < f && col_at(g) == k; ++g) sum += vals[rank_at(g)];
      if (rank < z) {
        col_c[off + rank] = (int32_t)k;
        val_c[off + rank] = sum;
      }
      ++rank;
    }
    __syncthreads();
  }
}

// ESC of one row per block with (column << log2(E) | rank) keys in 32 bits.
// Products are numbered by their rank t in A-row order (the order of the
// reference's left-to-right sum) and held wave-contiguously: product t sits
// in wave t / (64 IPT), register slot (t / 64) % IPT, lane t % 64, so each
// load instruction reads 64 consecutive products of the row.  A product
// finds its A nonzero without searching: the non-empty B segments of a
// window of T A nonzeros are numbered, each sets the bit of its first
// product in a bitmap, and product x belongs to segment
//   (set bits at positions <= x) - 1,
// one LDS read of the (bits, bits before) word pair.  The keys go straight
// into the register bitonic sort; products a*b (separately rounded) wait in
// LDS by rank.  After the sort each thread holds IPT consecutive keys; runs
// of equal columns are summed left to right by the thread that owns their
// head, and the sums go through LDS in rank order to coalesced stores.
// LDS: vals V[E] | region of max(4E, 16T + E/4) bytes: during the expansion
// the segment table (16 T bytes) and the bitmap word pairs (E / 4 bytes),
// then the sort's cross-wave exchanges and the sorted keys, then the staged
// columns.
// Stable LSD radix sort of the first f of E keys by bits [lo, lo + bits),
// 7 bits a pass, by the T threads of a block.  Element i lives in wave
// i / (64 IPT), register slot (i / 64) % IPT, lane i % 64 ("wave-contiguous"),
// so each wave's elements, chunk after chunk, are in index order.  A pass
// ranks every element among the same-digit elements of its wave before it
// (the lanes of a 64-element chunk with the same digit are found by seven
// ballots; each wave keeps running counts per digit in LDS), one block scan
// over (digit, wave) turns the counts into offsets, and the keys are
// scattered through buf.  Equal digits keep their order, so equal columns
// stay in rank order.  Only the f real keys take part; at the end buf[0, f)
// holds the sorted keys.  cnt: 128 * (T / 64) ints.  Holds barriers.
template <int T, int IPT>
__device__ __forceinline__ void block_radix_sort(uint32_t (&kr)[IPT], int f, int lo, int bits, uint32_t* buf,
                                                 int32_t* cnt, int64_t* s_w) {
  constexpr int NW = T / kWave;
  constexpr int kDigits = 128;
  static_assert(kDigits * NW == 2 * T, "two counters a thread in the scan");
  const int w = wave_id(), l = lane_id();
  const unsigned long long below = (1ull << l) - 1ull;
  for (int sh = lo; sh < lo + bits; sh += 7) {
    cnt[2 * threadIdx.x] = 0;
    cnt[2 * threadIdx.x + 1] = 0;
    __syncthreads();
    int32_t pos[IPT];
#pragma unroll
    for (int s = 0; s < IPT; ++s) {
      const int i = (w * IPT + s) * kWave + l;
      const bool act = i < f;
      const uint32_t d = (kr[s] >> sh) & (kDigits - 1);
      unsigned long long peers = __ballot(act);
#pragma unroll
      for (int b = 0; b < 7; ++b) {
        const bool bit = (d >> b) & 1u;
        const unsigned long long m = __ballot(act && bit);
        peers &= bit ? m : ~m;
      }
      const int32_t before = (int32_t)__popcll(peers & below);
      int32_t* c = &cnt[d * NW + w];
      const int32_t run = act ? *c : 0;
      pos[s] = run + before;
      if (act && before == 0) *c = run + (int32_t)__popcll(peers);
      wave_lds_sync();  // the next chunk reads the counts this one wrote
    }
    __syncthreads();
    // offsets of (digit, wave) in digit-major order
    const int32_t c0 = cnt[2 * threadIdx.x], c1 = cnt[2 * threadIdx.x + 1];
    int64_t tot;
    const int32_t ex = (int32_t)block_excl_scan<T, int64_t>((int64_t)(c0 + c1), s_w, tot);
    cnt[2 * threadIdx.x] = ex;
    cnt[2 * threadIdx.x + 1] = ex + c0;
    __syncthreads();
#pragma unroll
    for (int s = 0; s < IPT; ++s) {
      const int i = (w * IPT + s) * kWave + l;
      if (i < f) {
        const uint32_t d = (kr[s] >> sh) & (kDigits - 1);
        buf[cnt[d * NW + w] + pos[s]] = kr[s];
      }
    }
    __syncthreads();
    if (sh + 7 < lo + bits) {
#pragma unroll
      for (int s = 0; s < IPT; ++s) {
        const int i = (w * IPT + s) * kWave + l;
        kr[s] = i < f ? buf[i] : 0xffffffffu;
      }
    }
  }
}

// Sorts the f keys of one row (column << lo | rank: distinct, one per
// product slot i < f, slot i of this thread in kr[s] with
// i = (wave * IPT + s) * 64 + lane) into buf[0, f) without a sorting
// network.  A counting pass spreads the keys over NB = T * IPT buckets by an
// order-preserving map of the row's column range (every key of a bucket is
// smaller than every key of the next one, equal columns share a bucket), a
// scan turns the counts into starts, the keys are scattered, and each key's
// place is its bucket's start plus the keys of its bucket that are smaller
// than it.  Rows of uniform graphs leave about f / NB keys a bucket, so a
// key costs a few LDS operations instead of the log2(E)(log2(E)+1)/2 stages
// of the bitonic network over the padded bin.  Returns false, with kr
// untouched and buf clobbered, when some bucket holds more than
// kEscBucketMax keys (clustered columns); the caller then sorts kr itself.
// buf: NB words; s_w: T / 64 int64 of scratch.  Holds barriers; every thread
// of the block calls it.
constexpr int kEscBucketMax = 16;
template <int T, int IPT>
__device__ __forceinline__ bool block_bucket_sort(const uint32_t (&kr)[IPT], int f, int lo, uint32_t* buf,
                                                  int64_t* s_w) {
  constexpr int NW = T / kWave;
  constexpr int NB = T * IPT;
  static_assert(NB <= 4096, "12-bit bucket indices and positions");
  const int w = wave_id(), l = lane_id();
  uint32_t cmin = 0xffffffffu, cmax = 0u;
#pragma unroll
  for (int s = 0; s < IPT; ++s) {
    if ((w * IPT + s) * kWave + l < f) {
      const uint32_t c = kr[s] >> lo;
      cmin = min(cmin, c);
      cmax = max(cmax, c);
    }
  }
#pragma unroll
  for (int d = kWave / 2; d > 0; d >>= 1) {
    cmin = min(cmin, (uint32_t)__shfl_xor((int)cmin, d, kWave));
    cmax = max(cmax, (uint32_t)__shfl_xor((int)cmax, d, kWave));
  }
  uint32_t* s_mm = (uint32_t*)s_w;  // (min, max) of each wave
  if (l == 0) {
    s_mm[2 * w] = cmin;
    s_mm[2 * w + 1] = cmax;
  }
#pragma unroll
  for (int q = 0; q < IPT; ++q) buf[threadIdx.x * IPT + q] = 0u;
  __syncthreads();
#pragma unroll
  for (int v = 0; v < NW; ++v) {
    cmin = min(cmin, s_mm[2 * v]);
    cmax = max(cmax, s_mm[2 * v + 1]);
  }
  // bucket = floor((column - cmin) * NB / span), computed in float: the
  // conversion, the product with a positive scale and the truncation are
  // each non-decreasing, so the map keeps order
  const float scale = (float)NB / ((float)(cmax - cmin) + 1.0f);
  auto bucket = [&](uint32_t k) -> uint32_t {
    const uint32_t b = (uint32_t)((float)((k >> lo) - cmin) * scale);
    return b < (uint32_t)NB ? b : (uint32_t)(NB - 1);
  };
  // pk[s] packs a key's bucket size (bits 24-31, read only while at most
  // kEscBucketMax), its bucket and later its scattered position (bits
  // 12-23) and its position in the bucket from the counting atomic (0-11),
  // so that a key costs two registers
  uint32_t pk[IPT];
#pragma unroll
  for (int s = 0; s < IPT; ++s) {
    pk[s] = 0u;
    if ((w * IPT + s) * kWave + l < f) {
      const uint32_t b = bucket(kr[s]);
      pk[s] = (b << 12) | (uint32_t)atomicAdd(&buf[b], 1u);
    }
  }
  __syncthreads();
  // this thread's NB / T counters, the sizes of its keys' buckets, then the
  // starts written over the counts (after the scan's barrier, so every size
  // read is done)
  uint32_t sum = 0, big = 0;
#pragma unroll
  for (int q = 0; q < IPT; ++q) {
    const uint32_t c = buf[threadIdx.x * IPT + q];
    sum += c;
    big |= c > (uint32_t)kEscBucketMax;
  }
#pragma unroll
  for (int s = 0; s < IPT; ++s)
    if ((w * IPT + s) * kWave + l < f) pk[s] |= min(buf[pk[s] >> 12], 255u) << 24;
  int64_t tot;
  uint32_t run = (uint32_t)block_excl_scan_once<T, int64_t>((int64_t)sum, s_w, tot);
#pragma unroll
  for (int q = 0; q < IPT; ++q) {
    const uint32_t c = buf[threadIdx.x * IPT + q];
    buf[threadIdx.x * IPT + q] = run;
    run += c;
  }
  if (__syncthreads_or((int)big)) return false;
#pragma unroll
  for (int s = 0; s < IPT; ++s) {
    if ((w * IPT + s) * kWave + l < f) {
      const uint32_t pos = buf[(pk[s] >> 12) & 0xfffu] + (pk[s] & 0xfffu);
      pk[s] = (pk[s] & 0xff000fffu) | (pos << 12);
    }
  }
  __syncthreads();  // every start is read: the keys go over the counters
#pragma unroll
  for (int s = 0; s < IPT; ++s)
    if ((w * IPT + s) * kWave + l < f) buf[(pk[s] >> 12) & 0xfffu] = kr[s];
  __syncthreads();
  // place = bucket start + smaller keys of the bucket; a key of a bucket of
  // one stays where it is.  (Keys are read back from their scattered
  // slots, so that kr is dead from the scatter on.)
  uint32_t kv[IPT];
#pragma unroll
  for (int s = 0; s < IPT; ++s) {
    kv[s] = 0u;
    if ((w * IPT + s) * kWave + l < f && (pk[s] >> 24) > 1u) {
      const uint32_t p = (pk[s] >> 12) & 0xfffu;
      const uint32_t st = p - (pk[s] & 0xfffu);
      const uint32_t en = st + (pk[s] >> 24);
      const uint32_t k = buf[p];
      uint32_t r = st;
      for (uint32_t j = st; j < en; ++j) r += buf[j] < k ? 1u : 0u;
      kv[s] = k;
      pk[s] = r | 0x80000000u;
    } else {
      pk[s] = 0u;
    }
  }
  __syncthreads();  // every bucket read is done
#pragma unroll
  for (int s = 0; s < IPT; ++s)
    if (pk[s]) buf[pk[s] & 0xfffu] = kv[s];
  __syncthreads();
  return true;
}

// Sorts the f keys key[0, f) of one row, made of k ascending runs (run c is
// key[roff[c], roff[c + 1]), roff[k] = f), by rounds of pairwise merges:
// round L merges groups of 2^L runs two by two.  Thread i writes the sorted
// positions [i IPT, i IPT + IPT): it finds how many keys of each of its two
// groups come before its first position by a binary search along the merge
// path, merges IPT keys from there into registers and, after a barrier,
// stores them.  Keys are distinct (the rank is in their low bits).  The work
// follows f and the number of runs instead of the power of two a sorting
// network needs, and it uses that B's rows arrive sorted.  Returns false,
// with key[] untouched, when some run does not ascend (B rows with unsorted
// columns); otherwise out[r] is sorted key i IPT + r (0xffffffff past f) and
// key[] holds the sorted keys.  All threads call it: it holds barriers.
template <int T, int IPT>
__device__ __forceinline__ bool block_merge_runs(uint32_t* key, const uint16_t* roff, int k, int f,
                                                 uint32_t (&out)[IPT]) {
  const int i0 = (int)threadIdx.x * IPT;
  int c = 0;  // run of position i0
  bool bad = false;
  if (i0 < f) {
    int lo = 0, hi = k;  // roff[lo] <= i0 < roff[hi]
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if ((int)roff[mid] <= i0) lo = mid;
      else hi = mid;
    }
    c = lo;
    // this thread's keys and the next thread's first ascend inside each run
    int cc = c + 1, nst = roff[cc];
    uint32_t prev = key[i0];
    for (int p = i0 + 1; p <= i0 + IPT && p < f; ++p) {
      const uint32_t v = key[p];
      if (p == nst) nst = roff[++cc];  // p < f = roff[k], so cc < k here
      else bad |= !(prev < v);
      prev = v;
    }
  }
  if (__syncthreads_or(bad)) return false;
  for (int L = 0; (1 << L) < k; ++L) {
    if (i0 < f) {
      // the pair of groups holding position i0: runs [a0, a1) and [a1, b1)
      int a0 = (c >> (L + 1)) << (L + 1);
      int a1 = a0 + (1 << L) < k ? a0 + (1 << L) : k;
      int b1 = a0 + (2 << L) < k ? a0 + (2 << L) : k;
      int ae = roff[a1], be = roff[b1];
      const int as = roff[a0];
      const int d = i0 - as;  // outputs of the pair before i0
      int lo = d - (be - ae) > 0 ? d - (be - ae) : 0, hi = d < ae - as ? d : ae - as;
      while (lo < hi) {  // keys of the first group among them
        const int mid = (lo + hi) >> 1;
        if (key[as + mid] < key[ae + d - 1 - mid]) lo = mid + 1;
        else hi = mid;
      }
      int ia = as + lo, jb = ae + d - lo;
      uint32_t ka = ia < ae ? key[ia] : 0u, kb = jb < be ? key[jb] : 0u;
#pragma unroll
      for (int r = 0; r < IPT; ++r) {
        out[r] = 0xffffffffu;
        if (i0 + r < f) {
          if (ia == ae && jb == be) {  // the pair ends at roff[b1] = i0 + r < f: the next one starts
            a0 = b1;
            a1 = a0 + (1 << L) < k ? a0 + (1 << L) : k;
            b1 = a0 + (2 << L) < k ? a0 + (2 << L) : k;
            ia = roff[a0];
            ae = roff[a1];
            jb = ae;
            be = roff[b1];
            ka = key[ia];
            kb = jb < be ? key[jb] : 0u;
          }
          const bool ta = ia < ae && (jb >= be || ka < kb);
          out[r] = ta ? ka : kb;
          const int nx = (ta ? ia : jb) + 1;
          const uint32_t v = nx < (ta ? ae : be) ? key[nx] : 0u;
          if (ta) {
            ia = nx;
            ka = v;
          } else {
            jb = nx;
            kb = v;
          }
        }
      }
    }
    __syncthreads();  // every read of this round is done
#pragma unroll
    for (int r = 0; r < IPT; ++r)
      if (i0 + r < f) key[i0 + r] = out[r];
    __syncthreads();
  }
  if (k <= 1) {
#pragma unroll
    for (int r = 0; r < IPT; ++r) out[r] = i0 + r < f ? key[i0 + r] : 0xffffffffu;
  }
  return true;
}

#ifndef AMS_ESC_GROUP
#define AMS_ESC_GROUP 4
#endif
constexpr int kEscGroup = AMS_ESC_GROUP;  // products per thread in flight in the ESC expansion

#ifndef AMS_ESC_WPE
#define AMS_ESC_WPE 6
#endif
// SORT: 0 register bitonic network, 2 block LSD radix sort of the column
// bits, 3 merge of the row's B segments (rows of one window of A nonzeros
// whose segments ascend; other rows take the bitonic network), 4 bucket
// sort over the row's column range (rows with a crowded bucket take the
// bitonic network).
#ifndef AMS_ESC_WPE_MERGE
#define AMS_ESC_WPE_MERGE 5
#endif
#ifndef AMS_ESC_WPE_COUNT
#define AMS_ESC_WPE_COUNT 5
#endif
// 512-thread blocks of 49 KB of LDS: three a CU at six waves a SIMD (80
// VGPRs, with spills), two at four (128 VGPRs, none)
#ifndef AMS_ESC_WPE_4096
#define AMS_ESC_WPE_4096 4
#endif
template <int E, int T, typename V, int SORT>
__global__ __launch_bounds__(T) __attribute__((amdgpu_waves_per_eu(
    E >= 4096 ? AMS_ESC_WPE_4096 : (SORT == 3 ? AMS_ESC_WPE_MERGE : (SORT == 4 ? AMS_ESC_WPE_COUNT : AMS_ESC_WPE))))) void k_numc_esc32(
    Csr A, Csr B, const int32_t* list, const int32_t* counts, int bin, const int64_t* row_flp, const int64_t* rp_c,
    int32_t* col_c, V* val_c, int64_t cap, int col_bits, int32_t* err) {
  constexpr int IPT = E / T;
  constexpr int RB = __builtin_ctz(next_pow2_c(E));  // rank bits of a key
  constexpr int kWords = E / 32;                               // bitmap words of a window
  constexpr int kWpl = (kWords + kWave - 1) / kWave;           // words per lane of wave 0
  constexpr int kRegion = 4 * E > 16 * T + 8 * kWords ? 4 * E : 16 * T + 8 * kWords;
  static_assert(T * IPT == E && IPT <= 16, "the bin fills the threads' registers");
  __shared__ __align__(16) unsigned char region[kRegion];
  __shared__ V vals[E];
  __shared__ int64_t s_w[T / kWave];
  constexpr bool RADIX = SORT == 2, MERGE = SORT == 3, BUCKET = SORT == 4;
  __shared__ int32_t s_cnt[RADIX ? 2 * T : 1];
  __shared__ uint16_t roff[MERGE ? T + 1 : 1];  // run starts of a one-window row (merge)
  __shared__ int32_t s_nrun;
  uint32_t* key = (uint32_t*)region;
  SegBias<V>* sb = (SegBias<V>*)region;
  uint2* wpre = (uint2*)(region + 16 * T);
  if (rp_c[A.m] > cap) return;
  int32_t start, cnt_rows;
  bin_range(counts, bin, start, cnt_rows);
  const V* av = (const V*)A.val;
  const V* bv = (const V*)B.val;
  const int w = wave_id(), l = lane_id();
  // Rows idx, idx + gridDim.x, ... of the bin.  A block works on one row at
  // a time, so the next row's bounds, first A entries and B segment bounds
  // are loaded while the current row is expanded, sorted and compressed.
  // Row bounds travel as one dword per lane (lanes 0-3 C's row pointers,
  // 4-7 A's, 8-9 the row's products), read out by readlane.
  // (the lane index goes through an empty asm statement, so that the three
  // per-lane addresses are formed at each call rather than held, and
  // spilled, across the kernel)
  auto meta_load = [&](int32_t r) -> uint32_t {
    int ll = l;
    asm volatile("" : "+v"(ll));
    uint32_t v = 0;
    if (r >= 0 && ll < 10) {
      const uint32_t* src = ll < 4 ? (const uint32_t*)(rp_c + r) + ll
                                   : (ll < 8 ? (const uint32_t*)(A.rp + r) + (ll - 4)
                                             : (const uint32_t*)(row_flp + r) + (ll - 8));
      v = *src;
    }
    return v;
  };
  auto rd64 = [](uint32_t v, int lo) -> int64_t {
    const uint32_t a0 = (uint32_t)__builtin_amdgcn_readlane((int)v, lo);
    const uint32_t a1 = (uint32_t)__builtin_amdgcn_readlane((int)v, lo + 1);
    return (int64_t)(((uint64_t)a1 << 32) | a0);
  };
  int32_t idx = blockIdx.x;
  if (idx >= cnt_rows) return;
  uint32_t mv = meta_load(list[start + idx]);
  // this thread's A entry of the row's first window and its B segment
  int32_t pk = 0;
  V pa = V(0);
  // the next row's segment [pst, pen): its length is taken at use, so that
  // the bounds' load is not waited for where it is issued
  int64_t pst = 0, pen = 0;
  {
    const int64_t q0 = rd64(mv, 4) + threadIdx.x;
    if (q0 < rd64(mv, 6)) {
      pk = A.col[q0];
      pa = av[q0];
      pst = B.rp[pk];
      pen = B.rp[pk + 1];
    }
  }
  for (; idx < cnt_rows; idx += gridDim.x) {
    // the thread index passes through an empty asm statement once per row,
    // so that the compiler cannot hold (array + thread offset) addresses
    // across the whole kernel: they spilled, and each reload waited for
    // every load in flight, the next row's prefetch included
    int tid = (int)threadIdx.x;
    asm volatile("" : "+v"(tid));
    const int64_t off = rd64(mv, 0);
    const int32_t z = (int32_t)(rd64(mv, 2) - off);
    const int64_t p0 = rd64(mv, 4), p1 = rd64(mv, 6);
    const int32_t f = (int32_t)rd64(mv, 8);
    const int32_t nidx = idx + (int32_t)gridDim.x;
    const int32_t nrow = nidx < cnt_rows ? list[start + nidx] : -1;  // used after the first scan
    uint32_t nmv = 0;
    uint32_t kr[IPT];
#pragma unroll
    for (int s = 0; s < IPT; ++s) kr[s] = 0xffffffffu;
    int32_t base = 0;  // rank of the window's first product
    for (int64_t w0 = p0; w0 < p1; w0 += T) {
      for (int i = threadIdx.x; i < kWords; i += T) wpre[i] = make_uint2(0u, 0u);
      int64_t st = pst;
      int32_t len = (int32_t)(pen - pst);
      V a = pa;
      if (w0 != p0) {  // windows after the first (rows of more than T A nonzeros)
        const int64_t p = w0 + tid;
        st = 0;
        len = 0;
        a = V(0);
        if (p < p1) {
          const int32_t k = A.col[p];
          a = av[p];
          st = B.rp[k];
          len = (int32_t)(B.rp[k + 1] - st);
        }
      }
      // offsets and compact indices of the non-empty segments (the scan's
      // barrier also orders the bitmap clearing before the bits below)
      int32_t o, ci, total, nne;
      block_seg_scan<T>(len, (uint2*)s_w, o, ci, total, nne);
      if (w0 == p0) nmv = meta_load(nrow);  // the next row's bounds
      if (len > 0) {
        sb[ci] = SegBias<V>{st - o, a};
        atomicOr(&wpre[o >> 5].x, 1u << (o & 31));
        if constexpr (MERGE) roff[ci] = (uint16_t)o;
      }
      if constexpr (MERGE) {
        if (threadIdx.x == 0) {
          s_nrun = nne;
          roff[nne] = (uint16_t)total;
        }
      }
      __syncthreads();
      if (w == 0) {  // set bits before each word
        uint32_t pc[kWpl];
        int32_t sum = 0;
#pragma unroll
        for (int q = 0; q < kWpl; ++q) {
          const int wi = l * kWpl + q;
          pc[q] = wi < kWords ? (uint32_t)__popc(wpre[wi].x) : 0u;
          sum += (int32_t)pc[q];
        }
        int32_t run = wave_incl_scan(sum) - sum;
#pragma unroll
        for (int q = 0; q < kWpl; ++q) {
          const int wi = l * kWpl + q;
          if (wi < kWords) wpre[wi].y = (uint32_t)run;
          run += (int32_t)pc[q];
        }
      }
      __syncthreads();
      // this thread's products of the window, in groups of kEscGroup slots:
      // a group's LDS reads and loads are all issued before any is used.  A
      // wave whose slots [w*IPT*64, (w+1)*IPT*64) miss [base, base + total)
      // has nothing to expand.
      const int32_t wslot = __builtin_amdgcn_readfirstlane(w * IPT * kWave);
      if (wslot < base + total && wslot + IPT * kWave > base)
      static_for<0, (IPT + kEscGroup - 1) / kEscGroup>([&](auto gc) {
        constexpr int g0 = decltype(gc)::value * kEscGroup;
        constexpr int G = IPT - g0 < kEscGroup ? IPT - g0 : kEscGroup;
        int32_t xq[G];
        SegBias<V> sg[G];
        int32_t cq[G];
        V bq[G];
        bool in[G];
#pragma unroll
        for (int j = 0; j < G; ++j) {
          const int32_t t = (w * IPT + g0 + j) * kWave + l;
          in[j] = t >= base && t < base + total;
          xq[j] = in[j] ? t - base : 0;
        }
#pragma unroll
        for (int j = 0; j < G; ++j) {
          const uint2 wp = wpre[xq[j] >> 5];
          const int32_t seg = (int32_t)wp.y + __popc(wp.x & ((2u << (xq[j] & 31)) - 1u)) - 1;
          sg[j] = sb[seg < 0 ? 0 : seg];
        }
        // unconditional loads (slots outside the window read B position 0,
        // which exists when the row has products): straight-line code, so
        // the first wait comes at the first use below, not after each load
#pragma unroll
        for (int j = 0; j < G; ++j) {
          const int64_t q = in[j] ? sg[j].bias + xq[j] : 0;
          cq[j] = B.col[q];
          bq[j] = bv[q];
        }
#pragma unroll
        for (int j = 0; j < G; ++j) {
          if (in[j]) {
            const uint32_t t = (uint32_t)(base + xq[j]);
            vals[t] = sg[j].a * bq[j];
            kr[g0 + j] = ((uint32_t)cq[j] << RB) | t;
          }
        }
      });
      base += total;
      if (w0 + T < p1) __syncthreads();  // the next window rewrites the table
    }
    // the next row's first A entries
    pk = 0;
    pa = V(0);
    pst = 0;
    pen = 0;
    const int64_t q0 = rd64(nmv, 4) + tid;
    const bool pn = q0 < rd64(nmv, 6);
    if (pn) {
      pk = A.col[q0];
      pa = av[q0];
    }
    __syncthreads();  // products in LDS; the region is free for the sort
    const int i0 = tid * IPT;  // from the row's opaque thread index: no hoisted address
    bool merged = false;
    if constexpr (MERGE) {
      if (p1 - p0 <= T) {  // one window: the runs are its B segments, keys go to their ranks
        // every slot, without a branch (slots past f hold 0xffffffff and
        // land past f): branches made the compiler reload the base address
        // from scratch before each store, each reload waiting for the next
        // row's loads in flight
#pragma unroll
        for (int r = 0; r < IPT; ++r) key[(w * IPT + r) * kWave + l] = kr[r];
        __syncthreads();
        merged = block_merge_runs<T, IPT>(key, roff, __builtin_amdgcn_readfirstlane(s_nrun), f, kr);
      }
    }
    bool in_lds = RADIX;  // sorted keys in key[0, f) rather than in the registers
    if (!merged) {
      if constexpr (RADIX) {
        block_radix_sort<T, IPT>(kr, f, RB, col_bits, key, s_cnt, s_w);
      } else if constexpr (BUCKET) {
        in_lds = block_bucket_sort<T, IPT>(kr, f, RB, key, s_w);
        if (!in_lds) reg_flip_sort<T, IPT>(kr, key, f);
      } else {
        reg_flip_sort<T, IPT>(kr, key, f);
      }
    }
    if (pn) {  // and their B segments, while the runs are compressed
      pst = B.rp[pk];
      pen = B.rp[pk + 1];
    }
    if (!merged) {
      if (in_lds) {  // this thread's IPT consecutive sorted keys
#pragma unroll
        for (int r = 0; r < IPT; ++r) kr[r] = i0 + r < f ? key[i0 + r] : 0xffffffffu;
      } else {
#pragma unroll
        for (int r = 0; r < IPT; ++r) key[i0 + r] = kr[r];
        __syncthreads();
      }
    }
    // heads of runs of equal columns; the previous thread's last key from LDS
    const uint32_t pcol = (i0 > 0 && i0 <= f) ? (key[i0 - 1] >> RB) : 0u;
    uint32_t hd = 0;
    int64_t heads = 0;
#pragma unroll
    for (int r = 0; r < IPT; ++r) {
      const bool head = i0 + r < f &&
                        (r == 0 ? (i0 == 0 || (kr[0] >> RB) != pcol) : (kr[r] >> RB) != (kr[r - 1] >> RB));
      hd |= (uint32_t)head << r;
      heads += head;
    }
    int32_t tot;
    const int32_t rank0 = block_excl_scan<T, int32_t>(heads, (int32_t*)s_w, tot);
    if (threadIdx.x == 0 && tot != z) atomicOr(err, 2);
    // Runs are summed in registers.  The run that ends at element r (the last
    // one before the next head, or the thread's last element when the run
    // goes on into the next threads) keeps its column and sum in slot r, so
    // every register index is a constant.
    uint32_t ocol[IPT];
    V osum[IPT];
    uint32_t ends = 0;
    bool open = false;
    uint32_t rc = 0;
    V sum = V(0);
    static_for<0, IPT>([&](auto rr) {
      constexpr int r = decltype(rr)::value;
      ocol[r] = 0u;
      osum[r] = V(0);
      if (i0 + r < f) {
        const V x = vals[kr[r] & ((1u << RB) - 1u)];
        if ((hd >> r) & 1u) {
          if constexpr (r > 0) {
            if (open) {
              ocol[r - 1] = rc;
              osum[r - 1] = sum;
              ends |= 1u << (r - 1);
            }
          }
          open = true;
          rc = kr[r] >> RB;
          sum = x;
        } else if (open) {
          sum += x;
        }
      }
    });
    if (open) {
      for (int32_t g = i0 + IPT; g < f; ++g) {
        const uint32_t kg = key[g];
        if ((kg >> RB) != rc) break;
        sum += vals[kg & ((1u << RB) - 1u)];
      }
      ocol[IPT - 1] = rc;
      osum[IPT - 1] = sum;
      ends |= 1u << (IPT - 1);
    }
    __syncthreads();  // every read of the sorted keys and the products is done
#pragma unroll
    for (int r = 0; r < IPT; ++r) {
      if ((ends >> r) & 1u) {
        const int64_t i = rank0 + __popc(ends & ((1u << r) - 1u));
        if (i < z) {
          key[i] = ocol[r];
          vals[i] = osum[r];
        }
      }
    }
    __syncthreads();
    for (int32_t i = tid; i < z; i += T) {
      stream_store(&col_c[off + i], (int32_t)key[i]);
      stream_store(&val_c[off + i], vals[i]);
    }
    __syncthreads();  // before the next row reuses the LDS
    mv = nmv;
  }
}

// Stable LSD radix sort of n (key, value) pairs in global memory by key
// bits [0, bits), 7 bits a pass, by one block: wave w owns the index range
// [w n / NW, (w + 1) n / NW) and walks it 64 pairs at a time, so a pass
// counts each (digit, wave) over its range (same-digit lanes of a step found
// by seven ballots), one block scan turns the counts into offsets, and a
// second walk scatters the pairs into the other buffer.  Returns true when
// the result ended in (k1, v1).  cnt: 128 * NW ints of LDS.  Holds barriers.
template <int T, typename V>
__device__ bool block_radix_sort_global(int32_t* k0, V* v0, int32_t* k1, V* v1, int64_t n, int bits, int32_t* cnt,
                                        int64_t* s_w) {
  constexpr int NW = T / kWave;
  constexpr int kDigits = 128;
  const int w = wave_id(), l = lane_id();
  const unsigned long long below = (1ull << l) - 1ull;
  const int64_t lo = n * w / NW, hi = n * (w + 1) / NW;
  bool flip = false;
  auto peers_of = [&](bool act, uint32_t d) {
    unsigned long long peers = __ballot(act);
#pragma unroll
    for (int b = 0; b < 7; ++b) {
      const bool bit = (d >> b) & 1u;
      const unsigned long long mb = __ballot(act && bit);
      peers &= bit ? mb : ~mb;
    }
    return peers;
  };
  for (int sh = 0; sh < bits; sh += 7) {
    const int32_t* sk = flip ? k1 : k0;
    const V* sv = flip ? v1 : v0;
    int32_t* dk = flip ? k0 : k1;
    V* dv = flip ? v0 : v1;
    for (int i = threadIdx.x; i < kDigits * NW; i += T) cnt[i] = 0;
    __syncthreads();
    for (int64_t i0 = lo; i0 < hi; i0 += kWave) {
      const int64_t i = i0 + l;
      const bool act = i < hi;
      const uint32_t d = act ? ((uint32_t)sk[i] >> sh) & (kDigits - 1) : 0u;
      const unsigned long long peers = peers_of(act, d);
      if (act && (peers & below) == 0) cnt[d * NW + w] += (int32_t)__popcll(peers);
      wave_lds_sync();
    }
    __syncthreads();
    constexpr int PER = (kDigits * NW + T - 1) / T;
    int32_t part[PER];
    int32_t sum = 0;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int e = threadIdx.x * PER + q;
      part[q] = e < kDigits * NW ? cnt[e] : 0;
      sum += part[q];
    }
    int64_t tot;
    int32_t run = (int32_t)block_excl_scan<T, int64_t>((int64_t)sum, s_w, tot);
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int e = threadIdx.x * PER + q;
      if (e < kDigits * NW) cnt[e] = run;
      run += part[q];
    }
    __syncthreads();
    for (int64_t i0 = lo; i0 < hi; i0 += kWave) {
      const int64_t i = i0 + l;
      const bool act = i < hi;
      int32_t k = 0;
      V v = V(0);
      if (act) {
        k = sk[i];
        v = sv[i];
      }
      const uint32_t d = ((uint32_t)k >> sh) & (kDigits - 1);
      const unsigned long long peers = peers_of(act, d);
      const int32_t before = (int32_t)__popcll(peers & below);
      int32_t* c = &cnt[d * NW + w];
      const int32_t base = act ? *c : 0;
      if (act) {
        dk[base + before] = k;
        dv[base + before] = v;
      }
      if (act && before == 0) *c = base + (int32_t)__popcll(peers);
      wave_lds_sync();
    }
    __syncthreads();  // the pass's pairs are written before the next one reads them
    flip = !flip;
  }
  return flip;
}

// Global-hash numeric: one block per row of the global-hash route.  Products
// accumulate into a table of ghash_slots(z) keys and values in global memory
// (compare-and-swap of the key, atomic add of the separately rounded
// product), the occupied slots are compacted into C's row, and the row is
// radix-sorted by column through the region's scratch pairs.  Region bytes
// come from gbuf by an atomic cursor (ghash_num_bytes per row).
template <int T, typename V>
__global__ __launch_bounds__(T) void k_numc_ghash(Csr A, Csr B, const int32_t* list, const int32_t* counts, int bin,
                                                  const int64_t* rp_c, int32_t* col_c, V* val_c, int64_t cap,
                                                  unsigned char* gbuf, unsigned long long* cursor, int col_bits,
                                                  int32_t* err) {
  __shared__ ExpandLds<T> L;
  __shared__ V aval[T];
  __shared__ int64_t s_w[T / kWave];
  __shared__ unsigned long long s_base;
  __shared__ int32_t s_n;
  __shared__ int32_t s_cnt[128 * (T / kWave)];
  if (rp_c[A.m] > cap) return;
  int32_t start, cnt;
  bin_range(counts, bin, start, cnt);
  const V* av = (const V*)A.val;
  const V* bv = (const V*)B.val;
  for (int32_t idx = blockIdx.x; idx < cnt; idx += gridDim.x) {
    const int32_t row = list[start + idx];
    const int64_t off = rp_c[row];
    const int64_t z = rp_c[row + 1] - off;
    const int64_t S = ghash_slots(z);
    const int lg = ilog2_pow2(S);
    const uint32_t mask = (uint32_t)(S - 1);
    if (threadIdx.x == 0) {
      s_base = atomicAdd(cursor, (unsigned long long)ghash_num_bytes(z, (int)sizeof(V)));
      s_n = 0;
    }
    __syncthreads();
    int32_t* keys = (int32_t*)(gbuf + s_base);
    V* vals = (V*)(gbuf + s_base + (((uint64_t)S * 4 + 7) & ~7ull));
    int32_t* tk = (int32_t*)(vals + S);
    V* tv = (V*)(gbuf + s_base + (((((uint64_t)S * 4 + 7) & ~7ull) + (uint64_t)S * sizeof(V) + (uint64_t)z * 4 + 7) & ~7ull));
    for (int64_t i = threadIdx.x; i < S; i += T) {
      keys[i] = kEmpty;
      vals[i] = V(0);
    }
    __syncthreads();
    block_expand<T, 4>(A.rp[row], A.rp[row + 1], L,
                       [&](int64_t p, int64_t& st, int32_t& len) {
                         const int32_t k = A.col[p];
                         st = B.rp[k];
                         len = (int32_t)(B.rp[k + 1] - st);
                         aval[threadIdx.x] = av[p];
                       },
                       [&](int j, int64_t, int64_t q, int64_t) { return Gather<V>{B.col[q], j, bv[q]}; },
                       [&](Gather<V> g) {
                         const V x = aval[g.j] * g.val;
                         uint32_t h = lg == 0 ? 0u : ((uint32_t)g.key * kHashMul) >> (32 - lg);
                         for (uint32_t probe = 0; probe <= mask; ++probe) {
                           const int32_t old = atomicCAS(&keys[h], kEmpty, g.key);
                           if (old == kEmpty || old == g.key) {
                             atomicAdd(&vals[h], x);
                             return;
                           }
                           h = (h + 1) & mask;
                         }
                         atomicOr(err, 0x80);
                       });
    // compact the occupied slots into C's row (any order; sorted below)
    for (int64_t s0 = 0; s0 < S; s0 += T) {
      const int64_t sl = s0 + threadIdx.x;
      const int32_t key = sl < S ? keys[sl] : kEmpty;
      const bool occ = key != kEmpty;
      const unsigned long long mk = __ballot(occ);
      int32_t base = 0;
      if (lane_id() == 0 && mk) base = atomicAdd(&s_n, (int32_t)__popcll(mk));
      base = __shfl(base, 0, kWave);
      const int32_t r = base + (int32_t)__popcll(mk & ((1ull << lane_id()) - 1ull));
      if (occ && r < z) {
        col_c[off + r] = key;
        val_c[off + r] = vals[sl];
      }
    }
    __syncthreads();
    if (threadIdx.x == 0 && s_n != z) atomicOr(err, 0x20);
    if (s_n == z && z > 1) {
      const bool in_t = block_radix_sort_global<T, V>(col_c + off, val_c + off, tk, tv, z, col_bits, s_cnt, s_w);
      if (in_t) {
        for (int64_t i = threadIdx.x; i < z; i += T) {
          col_c[off + i] = tk[i];
          val_c[off + i] = tv[i];
        }
      }
    }
    __syncthreads();
  }
}

// --------------------------------------------------- column-blocked B

// Column-blocked copy of B for the dense bin: chunk c of row k holds the
// entries with (column - problem column base) / W == c.  One wave per B
// row.  Rows with ascending columns (the usual case) have every chunk as
// one contiguous run: an entry starts a run when the previous entry lies in
// another chunk, a DPP max-scan hands every entry its run's start, and the
// runs are counted / copied without LDS or barriers.  Rows in any other
// order take a ballot loop over the chunks present in each 64-entry group,
// with per-chunk cursors in LDS (R ints).

// Lane i - 1's value (wave_shr:1); lane 0 gets `first`.
__device__ __forceinline__ int32_t wave_prev(int32_t x, int32_t first) {
  const int32_t y = (int32_t)__builtin_amdgcn_update_dpp(0, x, 0x138, 0xf, 0xf, false);
  return lane_id() == 0 ? first : y;
}

// Lane i + 1's value (wave_shl:1); lane 63 gets `last`.
__device__ __forceinline__ int32_t wave_next(int32_t x, int32_t last) {
  const int32_t y = (int32_t)__builtin_amdgcn_update_dpp(0, x, 0x130, 0xf, 0xf, false);
  return lane_id() == kWave - 1 ? last : y;
}

// True when the columns of B row [s, e) never decrease.
// Hub rows of B have about 10^5 entries: both walks below load kRowGroups
// 64-entry groups before they use the first, so that a long row takes a
// quarter of the load latencies in a row.
constexpr int kRowGroups = 4;

__device__ __forceinline__ bool row_ascending(const int32_t* col, int64_t s, int64_t e) {
  bool bad = false;
  int32_t before = INT32_MIN;  // last column of the previous group
  for (int64_t q0 = s; q0 < e; q0 += kWave * kRowGroups) {
    int32_t x[kRowGroups];
#pragma unroll
    for (int g = 0; g < kRowGroups; ++g) {
      const int64_t q = q0 + (int64_t)g * kWave + lane_id();
      x[g] = q < e ? col[q] : INT32_MAX;
    }
#pragma unroll
    for (int g = 0; g < kRowGroups; ++g) {
      bad |= wave_prev(x[g], before) > x[g];
      before = __builtin_amdgcn_readlane(x[g], kWave - 1);
    }
  }
  return !__any(bad);
}

// Inclusive prefix max over the 64 lanes of non-negative values.
__device__ __forceinline__ int32_t wave_incl_max_nonneg(int32_t x) {
  x = max(x, (int32_t)dpp_move<0x111, 0xf>((uint32_t)x));
  x = max(x, (int32_t)dpp_move<0x112, 0xf>((uint32_t)x));
  x = max(x, (int32_t)dpp_move<0x114, 0xf>((uint32_t)x));
  x = max(x, (int32_t)dpp_move<0x118, 0xf>((uint32_t)x));
  x = max(x, (int32_t)dpp_move<0x142, 0xa>((uint32_t)x));
  x = max(x, (int32_t)dpp_move<0x143, 0xc>((uint32_t)x));
  return x;
}

// Walks an ascending row in 64-entry groups; for every entry calls
// f(q, chunk, local column, first index of its run in the row, whether it
// ends the run).  One 32-bit division per entry; neighbours' chunks come
// from lane shifts.
template <class F>
__device__ __forceinline__ void row_runs(const int32_t* col, int64_t s, int64_t e, int32_t lo, int W, F f) {
  int32_t carry = 0;   // 1 + row-local index of the latest run start
  int32_t cprev = -1;  // chunk of the previous group's last entry
  for (int64_t g0 = s; g0 < e; g0 += kWave * kRowGroups) {
    int32_t x[kRowGroups], xn[kRowGroups];  // columns, and lane 63's next one
#pragma unroll
    for (int g = 0; g < kRowGroups; ++g) {
      const int64_t q = g0 + (int64_t)g * kWave + lane_id();
      x[g] = q < e ? col[q] : 0;
      xn[g] = (lane_id() == kWave - 1 && q + 1 < e) ? col[q + 1] : 0;
    }
#pragma unroll
    for (int g = 0; g < kRowGroups; ++g) {
      const int64_t q = g0 + (int64_t)g * kWave + lane_id();
      const bool valid = q < e;
      const int32_t loc = valid ? x[g] - lo : 0;
      const int32_t c = valid ? loc / W : -1;
      int32_t cn = -1;
      if (lane_id() == kWave - 1 && q + 1 < e) cn = (xn[g] - lo) / W;
      const int32_t cp = wave_prev(c, cprev);
      cn = wave_next(c, cn);
      const bool start = valid && cp != c;
      int32_t m = wave_incl_max_nonneg(start ? (int32_t)(q - s + 1) : 0);
      m = m > carry ? m : carry;
      carry = __builtin_amdgcn_readlane(m, kWave - 1);
      cprev = __builtin_amdgcn_readlane(c, kWave - 1);
      if (valid) f(q, c, loc, (int64_t)(m - 1), q + 1 == e || cn != c);
    }
  }
}

// Rows of at most kBblkShort entries: one thread per row, walking its
// entries in order (64 rows in flight per wave).  The wave kernels below
// skip these rows.
constexpr int64_t kBblkShort = 64;

__global__ __launch_bounds__(256) void k_bblk_count_short(Csr B, Probs pr, int W, int R, int64_t* cnt,
                                                          int32_t* lrows, int32_t* nlong) {
  const int64_t ld = B.m + 1;
  for (int64_t k0 = blockIdx.x * (int64_t)blockDim.x; k0 < B.m; k0 += (int64_t)gridDim.x * blockDim.x) {
    const int64_t k = k0 + threadIdx.x;
    const int64_t s = k < B.m ? B.rp[k] : 0, e = k < B.m ? B.rp[k + 1] : 0;
    // long rows go to the list of the wave-per-row kernels
    const bool is_long = e - s > kBblkShort;
    const unsigned long long lm = __ballot(is_long);
    if (lm) {
      int32_t base = 0;
      if (lane_id() == __ffsll(lm) - 1) base = atomicAdd(nlong, (int32_t)__popcll(lm));
      base = __shfl(base, __ffsll(lm) - 1, kWave);
      if (is_long) lrows[base + __popcll(lm & ((1ull << lane_id()) - 1))] = (int32_t)k;
    }
    if (is_long || e == s) continue;
    const int32_t lo = (int32_t)prob_bcol_lo(pr, k);
    int32_t prev_col = INT32_MIN, prev_c = -1, run = 0;
    bool sorted = true;
    for (int64_t q = s; q < e; ++q) {
      const int32_t col = B.col[q];
      const int32_t c = (col - lo) / W;
      sorted &= col >= prev_col;
      prev_col = col;
      if (c != prev_c) {
        if (prev_c >= 0) cnt[prev_c * ld + k] = run;
        prev_c = c;
        run = 0;
      }
      ++run;
    }
    if (sorted) {
      cnt[prev_c * ld + k] = run;
      continue;
    }
    // any other order: recount (this thread alone owns row k's counters)
    for (int c = 0; c < R; ++c) cnt[c * ld + k] = 0;
    for (int64_t q = s; q < e; ++q) cnt[((B.col[q] - lo) / W) * ld + k] += 1;
  }
}

template <typename V>
__global__ __launch_bounds__(256) void k_bblk_scatter_short(Csr B, Probs pr, int W, const int64_t* bptr,
                                                            uint16_t* bcol, V* bval, int64_t* perm) {
  // rows longer than kBblkShort are k_bblk_scatter's
  const int64_t ld = B.m + 1;
  const V* bv = (const V*)B.val;
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < B.m; k += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = B.rp[k], e = B.rp[k + 1];
    if (e - s > kBblkShort || e == s) continue;
    const int32_t lo = (int32_t)prob_bcol_lo(pr, k);
    bool sorted = true;
    for (int64_t q = s + 1; q < e; ++q) sorted &= B.col[q] >= B.col[q - 1];
    int32_t prev_c = -1;
    int64_t pos = 0;
    for (int64_t q = s; q < e; ++q) {
      const int32_t loc = B.col[q] - lo;
      const int32_t c = loc / W;
      if (sorted) {
        if (c != prev_c) {
          pos = bptr[c * ld + k];
          prev_c = c;
        }
      } else {  // rank among the row's earlier entries of the same chunk
        pos = bptr[c * ld + k];
        for (int64_t j = s; j < q; ++j) pos += ((B.col[j] - lo) / W) == c;
      }
      bcol[pos] = (uint16_t)(loc - c * W);
      bval[pos] = bv[q];
      if (perm) perm[pos] = q;
      ++pos;
    }
  }
}

// Long rows (more than kBblkShort entries, listed by k_bblk_count_short):
// a wave each, persistent over the list.  Per-wave LDS: R chunk cursors for
// rows whose columns are not ascending.
template <class Row>
__device__ __forceinline__ void long_rows(const Csr& B, const int32_t* lrows, const int32_t* nlong, Row row) {
  const int32_t n = *nlong;
  const int32_t waves = (int32_t)(gridDim.x * (blockDim.x / kWave));
  for (int32_t i = (int32_t)(blockIdx.x * (blockDim.x / kWave) + wave_id()); i < n; i += waves) {
    const int64_t k = lrows[i];
    row(k, (int64_t)B.rp[k], (int64_t)B.rp[k + 1]);
  }
}

__global__ __launch_bounds__(256) void k_bblk_count(Csr B, Probs pr, int W, int R, int64_t* cnt,
                                                    const int32_t* lrows, const int32_t* nlong) {
  extern __shared__ int32_t cw_all[];
  int32_t* cw = cw_all + (size_t)R * wave_id();
  const int64_t ld = B.m + 1;
  long_rows(B, lrows, nlong, [&](int64_t k, int64_t s, int64_t e) {
    const int32_t lo = (int32_t)prob_bcol_lo(pr, k);
    if (row_ascending(B.col, s, e)) {
      // cnt starts at zero; the last entry of every run writes its length
      row_runs(B.col, s, e, lo, W, [&](int64_t q, int c, int32_t, int64_t first, bool last) {
        if (last) cnt[c * ld + k] = (q - s) - first + 1;
      });
      return;
    }
    for (int c = lane_id(); c < R; c += kWave) cw[c] = 0;
    wave_lds_sync();
    for (int64_t q = s + lane_id(); q < e; q += kWave) atomicAdd(&cw[(B.col[q] - lo) / W], 1);
    wave_lds_sync();
    for (int c = lane_id(); c < R; c += kWave) cnt[c * ld + k] = cw[c];
    wave_lds_sync();
  });
}

// Scatter B into the blocked layout; entries of one (chunk, row) segment
// keep B's order.  perm (optional) records each entry's original position.
template <typename V>
__global__ __launch_bounds__(256) void k_bblk_scatter(Csr B, Probs pr, int W, int R, const int64_t* bptr,
                                                      uint16_t* bcol, V* bval, int64_t* perm,
                                                      const int32_t* lrows, const int32_t* nlong) {
  extern __shared__ int32_t cw_all[];
  int32_t* cw = cw_all + (size_t)R * wave_id();
  const int64_t ld = B.m + 1;
  const V* bv = (const V*)B.val;
  long_rows(B, lrows, nlong, [&](int64_t k, int64_t s, int64_t e) {
    const int32_t lo = (int32_t)prob_bcol_lo(pr, k);
    if (row_ascending(B.col, s, e)) {
      row_runs(B.col, s, e, lo, W, [&](int64_t q, int c, int32_t loc, int64_t first, bool) {
        const int64_t pos = bptr[c * ld + k] + (q - s) - first;
        bcol[pos] = (uint16_t)(loc - c * W);
        bval[pos] = bv[q];
        if (perm) perm[pos] = q;
      });
      return;
    }
    for (int c = lane_id(); c < R; c += kWave) cw[c] = 0;
    wave_lds_sync();
    for (int64_t q0 = s; q0 < e; q0 += kWave) {
      const int64_t q = q0 + lane_id();
      bool active = q < e;
      int32_t loc = 0;
      int c = 0;
      if (active) {
        loc = B.col[q] - lo;
        c = loc / W;
      }
      unsigned long long todo = __ballot(active);
      while (todo) {
        const int lead = __ffsll(todo) - 1;
        const int cc = __shfl(c, lead, kWave);
        const unsigned long long mask = __ballot(active && c == cc);
        const int32_t prev = cw[cc];
        if (active && c == cc) {
          const int64_t pos = bptr[cc * ld + k] + prev + __popcll(mask & ((1ull << lane_id()) - 1));
          bcol[pos] = (uint16_t)(loc - cc * W);
          bval[pos] = bv[q];
          if (perm) perm[pos] = q;
          active = false;
        }
        wave_lds_sync();
        if (lane_id() == lead) cw[cc] = prev + (int32_t)__popcll(mask);
        wave_lds_sync();
        todo &= ~mask;
      }
    }
    wave_lds_sync();
  });
}

// One morsel of the dense bin: wide row `row`, column chunk `c`.
struct DenseTask {
  int32_t h;        // the row's index in the bitmap bin (its coff and rowbits row)
  int32_t cnt;      // nnz of C in this chunk
  int32_t c;
  int32_t deg;      // nnz of A's row
  int64_t off;      // first C slot of the chunk's entries
  int64_t p0;       // first A nonzero of the row
  int32_t colbase;  // first output column of the chunk
  int32_t cwidth;   // output columns in the chunk (<= W)
};
static_assert(sizeof(DenseTask) == 40, "descriptor dwords");

// Light morsels: at most this many A nonzeros (two a lane of one wave; a
// longer row stays with the block kernel).
constexpr int kLightDeg = 128;

// flag[c*H + h] for morsel (h, c) routed dense and not empty: 1 for the
// block kernel, 1 << 32 for the wave kernel of light morsels (at most
// light_max C entries and kLightDeg A nonzeros; light_max 0: none).
__global__ __launch_bounds__(256) void k_dense_flags(const int32_t* list, const int32_t* counts, int bin,
                                                     const uint8_t* nbin, const int32_t* coff, int R,
                                                     const int64_t* arp, int light_max, int64_t* flag) {
  int32_t start, H;
  bin_range(counts, bin, start, H);
  const int64_t n = (int64_t)H * R;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(i / H);
    const int32_t h = (int32_t)(i - (int64_t)c * H);
    const int32_t row = list[start + h];
    const int32_t* crow = coff + (int64_t)h * (R + 1);
    const int32_t cnt = crow[c + 1] - crow[c];
    int64_t f = 0;
    if (nbin[row] == NB_DENSE && cnt > 0)
      f = (cnt <= light_max && arp[row + 1] - arp[row] <= kLightDeg) ? (1ll << 32) : 1;
    flag[i] = f;
  }
}

// Writes the descriptors of the flagged morsels at their scanned positions
// (pos = exclusive scan of the flags, n_all + 1 entries; block-kernel
// morsels counted in the low 32 bits, light ones in the high 32), keeping
// the chunk-major order: the block kernel's morsels first, then the light
// ones.  Their counts go to ntask[0] and ntask[1].
__global__ __launch_bounds__(256) void k_dense_fill(Csr A, Probs pr, int W, const int32_t* list,
                                                    const int32_t* counts, int bin, const int32_t* coff, int R,
                                                    const int64_t* rp_c, const int64_t* pos, DenseTask* tasks,
                                                    int64_t* ntask) {
  int32_t start, H;
  bin_range(counts, bin, start, H);
  const int64_t n = (int64_t)H * R;
  const int64_t nheavy = pos[n] & 0xffffffffll;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t pv = pos[i], pn = pos[i + 1];
    if (pn == pv) continue;  // not flagged
    const int64_t p = (pn & 0xffffffffll) != (pv & 0xffffffffll) ? (pv & 0xffffffffll) : nheavy + (pv >> 32);
    const int c = (int)(i / H);
    const int32_t h = (int32_t)(i - (int64_t)c * H);
    const int32_t row = list[start + h];
    const int32_t* crow = coff + (int64_t)h * (R + 1);
    DenseTask d;
    d.h = h;
    d.cnt = crow[c + 1] - crow[c];
    d.c = c;
    d.p0 = A.rp[row];
    d.deg = (int32_t)(A.rp[row + 1] - d.p0);
    d.off = rp_c[row] + crow[c];
    int64_t lo, width;
    prob_cols(pr, row, lo, width);
    d.colbase = (int32_t)(lo + (int64_t)c * W);
    const int64_t cw = width - (int64_t)c * W;
    d.cwidth = (int32_t)(cw < W ? cw : W);
    tasks[p] = d;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ntask[0] = nheavy;
    ntask[1] = pos[n] >> 32;
  }
}

constexpr unsigned long long kGrab = 8;
// Emit groups whose non-zero bitmap words average fewer set bits than this
// take one set bit per lane; denser groups take a word per half-wave, one
// bit per lane (G500 s19 ef32: 336.6 -> 319.9 ms; 4: 328, 16: 326 ms).
constexpr int kSparseEmitBits = 8;
// Staging entries a wave for the sparse emission (a sparse group has fewer
// than kSparseEmitBits * 64 set bits); T of them fill the segment
// descriptors' LDS (T * sizeof(SegBias) bytes).
constexpr int kSparseStage = kSparseEmitBits * 64;
static_assert(kSparseStage * 2 <= kWave * 16, "a wave's staging row fits its share of the descriptors");

// A morsel descriptor held as one dword per lane (lanes 0..7 of every
// wave): it stays a vector load, so waits on LDS results never also wait
// for it, and its fields are read out only when the morsel starts.
__device__ __forceinline__ uint32_t task_load(const DenseTask* tasks, unsigned long long t,
                                              unsigned long long ntask) {
  uint32_t v = 0;
  if (t < ntask && lane_id() < (int)(sizeof(DenseTask) / 4)) v = ((const uint32_t*)(tasks + t))[lane_id()];
  return v;
}

__device__ __forceinline__ DenseTask task_get(uint32_t v) {
  auto lane = [&](int i) { return (uint32_t)__builtin_amdgcn_readlane((int)v, i); };
  DenseTask d;
  d.h = (int32_t)lane(0);
  d.cnt = (int32_t)lane(1);
  d.c = (int32_t)lane(2);
  d.deg = (int32_t)lane(3);
  d.off = (int64_t)(((uint64_t)lane(5) << 32) | lane(4));
  d.p0 = (int64_t)(((uint64_t)lane(7) << 32) | lane(6));
  d.colbase = (int32_t)lane(8);
  d.cwidth = (int32_t)lane(9);
  return d;
}

template <typename V>
struct DenseItem {
  uint32_t key;  // column inside the chunk
  V val;         // B value
  V a;           // A value
};

// Dense-chunk numeric: persistent blocks take morsels (wide row, column
// chunk) from the compacted descriptor array through a global ticket
// counter; the array is chunk-major, so concurrently running morsels read
// the same column block of B.  A morsel is accumulated in a dense LDS array
// and emitted by bitmap rank as one sorted, coalesced run of C.
//
// Per window of T nonzeros of the A row, products are numbered by an
// exclusive scan of the segment lengths and handed out in passes of up to
// T*U: wave w takes pass positions [w*64*ue, (w+1)*64*ue) with ue =
// ceil(n/T), lane l every 64th of them.  A pass locates the segment of each
// position without searching: the non-empty segments are numbered
// (compact index), each one starting inside the pass sets its start bit in
// a pass bitmap, and the segment of position x is
//   cover + (bits set at positions <= x),
// with cover the segment holding the pass's first product.  The bitmaps are
// double buffered (each pass clears the other one), so a pass costs two
// barriers.  The next morsel's descriptor, A columns and B segment bounds
// are loaded while the current one runs.
//
// Dynamic LDS: V dense[W + 64] | uint32 bitmap[(W + 64)/32] (padded to 16
//              bytes) | SegBias<V> seg[T]; entry W + lane is lane's sink
#ifndef AMS_DENSE_WPE
#define AMS_DENSE_WPE 4
#endif
// PB: the symbolic phase left each row's column bitmap in pr.rowbits (one
// word a thread: W <= 32 T): the products skip the bitmap test and OR, and
// each morsel's words are loaded while its products run and stored into the
// LDS bitmap before its output.
template <int T, typename V, bool PB>
__global__ __launch_bounds__(T) __attribute__((amdgpu_waves_per_eu(T >= 1024 ? AMS_DENSE_WPE : 4))) void k_numc_dense(
                                                  Csr A, int64_t mB, const int64_t* bptr,
                                                  const uint16_t* bcol, const V* bval,
                                                  const DenseTask* tasks,
                                                  const int64_t* __restrict__ d_ntask, int W,
                                                  const int64_t* rp_c, int32_t* col_c, V* val_c,
                                                  int64_t cap, Probs pr, unsigned long long* ticket,
                                                  int32_t* err) {
  constexpr int U = dense_u<T>();
  constexpr int kPassWords = T * U / 32;
  constexpr int kWpl = kPassWords / kWave;  // pass bitmap words per lane
  static_assert(kPassWords % kWave == 0, "one wave scans the pass bitmap");
  extern __shared__ __align__(16) unsigned char dyn_raw[];
  V* dense = (V*)dyn_raw;                           // W columns, then one sink per lane
  uint32_t* bm = (uint32_t*)(dense + W + kWave);    // their bits: (W + 64) / 32 words
  unsigned char* uni = dyn_raw + (((size_t)(W + kWave) * sizeof(V) + (size_t)((W + kWave) >> 5) * 4 + 15) & ~(size_t)15);
  SegBias<V>* sb = (SegBias<V>*)uni;
  __shared__ uint32_t pbits[3][kPassWords];
  __shared__ uint2 wpword[T / kWave][2 * U];  // per wave: (start bits, set bits before) of its pass words
  __shared__ int32_t cover[3];
  __shared__ uint2 s_w[2][T / kWave];
  __shared__ uint32_t s_rank[T / kWave];
  __shared__ unsigned long long s_ring[4];
  if (rp_c[A.m] > cap) return;
  const int words = (W + kWave) >> 5;
  for (int i = threadIdx.x; i < W + kWave; i += T) dense[i] = V(0);
  for (int i = threadIdx.x; i < words; i += T) bm[i] = 0u;
  for (int i = threadIdx.x; i < 3 * kPassWords; i += T) (&pbits[0][0])[i] = 0u;
  const unsigned long long ntask = (unsigned long long)*d_ntask;
  const V* av = (const V*)A.val;
  const int64_t ld = mB + 1;
  // Ticket stream of this block, kept by thread 0: batches of `grab`
  // tickets (up to kGrab when morsels are many, one at a time when few and
  // uneven), the next batch claimed one batch ahead.  s_ring[i & 3] is the
  // ticket of the block's i-th morsel, published two morsels ahead.
  const unsigned long long per_block = ntask / ((unsigned long long)gridDim.x * 64ull);
  const unsigned long long grab = per_block < 1 ? 1 : (per_block > kGrab ? kGrab : per_block);
  unsigned long long q_tk = 0, q_end = 0, q_nb = 0;
  auto take = [&]() -> unsigned long long {
    if (q_tk == q_end) {
      q_tk = q_nb;
      q_end = q_nb + grab;
      q_nb = atomicAdd(ticket, grab);
    }
    return q_tk++;
  };
  if (threadIdx.x == 0) {
    q_nb = atomicAdd(ticket, grab);
    s_ring[0] = take();
    s_ring[1] = take();
    s_ring[2] = take();
  }
  __syncthreads();
  unsigned long long t_cur = uniform_u64(s_ring[0]), t_nxt = uniform_u64(s_ring[1]);
  uint32_t dv = task_load(tasks, t_cur, ntask), ndv = task_load(tasks, t_nxt, ntask);
  DenseTask d = task_get(dv);
  // this thread's segment [st, en) of the current morsel's first window; its
  // length is taken where it is first used, so that the load's wait comes
  // there (after the previous morsel's output) and not right after the load
  int64_t st = 0, en = 0;
  V a = V(0);
  if (t_cur < ntask && (int)threadIdx.x < d.deg) {
    const int32_t k = A.col[d.p0 + threadIdx.x];
    a = av[d.p0 + threadIdx.x];
    const int64_t* bp = bptr + (int64_t)d.c * ld;
    st = bp[k];
    en = bp[k + 1];
  }
  int pb = 0;       // pass bitmap of the next pass (three in turn)
  unsigned wc = 0;  // window counter: parity picks the scan scratch
  for (unsigned mc = 0; t_cur < ntask; ++mc) {
    // slot (mc+3)&3 was last read at morsel mc-1 and is read next at mc+1;
    // the barriers of every morsel order both
    // everything waited on here was issued a morsel ago; the new prefetches
    // go out after it, so that no wait also covers them
    const DenseTask nd = task_get(ndv);
    const unsigned long long t_nn = uniform_u64(s_ring[(mc + 2) & 3]);
    if (threadIdx.x == 0) s_ring[(mc + 3) & 3] = take();
    const uint32_t nndv = task_load(tasks, t_nn, ntask);  // descriptor two morsels ahead
    // the morsel's column bitmap word of this thread (PB)
    uint32_t pbw = 0u;
    if constexpr (PB) {
      if ((int)threadIdx.x < ((d.cwidth + 31) >> 5))
        pbw = pr.rowbits[(int64_t)d.h * pr.bwords + (int64_t)d.c * (W >> 5) + threadIdx.x];
    }
    const int64_t p1 = d.p0 + d.deg;
    for (int64_t w0 = d.p0; w0 < p1; w0 += T) {
      const int32_t len = (int32_t)(en - st);
      // scan of the segment lengths: offsets and compact index
      int32_t off, ci, total, nne;
      block_seg_scan<T>(len, s_w[wc++ & 1], off, ci, total, nne);
      // prefetch the A entries of the next window: this row's next one, or
      // the next morsel's first; their B bounds follow after the passes
      const bool last = w0 + T >= p1;
      int64_t np = 0;
      int32_t nn = 0, ncw = 0;
      if (!last) {
        np = w0 + T;
        nn = (int32_t)((p1 - np) < T ? (p1 - np) : T);
        ncw = d.c;
      } else if (t_nxt < ntask) {
        np = nd.p0;
        nn = nd.deg < T ? nd.deg : T;
        ncw = nd.c;
      }
      const bool npre = (int)threadIdx.x < nn;
      int32_t nk = 0;
      V na = V(0);
      if (npre) {
        nk = A.col[np + threadIdx.x];
        na = av[np + threadIdx.x];
      }
      if (len > 0) sb[ci] = SegBias<V>{st - off, a};
      for (int32_t t0 = 0; t0 < total; t0 += T * U) {
        const int Bf = pb;
        pb = pb == 2 ? 0 : pb + 1;
        const int32_t n = (total - t0) < T * U ? (total - t0) : T * U;
        if (len > 0) {
          if (off <= t0 && t0 < off + len) {
            cover[Bf] = ci;
          } else if (off > t0 && off < t0 + n) {
            const int32_t x = off - t0;
            atomicOr(&pbits[Bf][x >> 5], 1u << (x & 31));
          }
        }
        __syncthreads();
        // the bitmap after next was last read in the previous pass, before
        // the barrier above, and is next written after the next one
        {
          const int Bc = pb == 2 ? 0 : pb + 1;
          for (int i = threadIdx.x; i < kPassWords; i += T) pbits[Bc][i] = 0u;
        }
        // every wave ranks the pass's start bits itself: lane l holds words
        // [l*kWpl, l*kWpl + kWpl) and the count of set bits before them
        uint32_t wb[kWpl];
        int32_t wpre[kWpl];
        {
          int32_t sum = 0;
#pragma unroll
          for (int q = 0; q < kWpl; ++q) {
            wb[q] = pbits[Bf][lane_id() * kWpl + q];
            sum += __popc(wb[q]);
          }
          int32_t run = wave_incl_scan(sum) - sum;
#pragma unroll
          for (int q = 0; q < kWpl; ++q) {
            wpre[q] = run;
            run += __popc(wb[q]);
          }
        }
        const int ue = (n + T - 1) / T;
        const int xw = wave_id() * kWave * ue;  // the wave's first position
        const int xb = xw + lane_id();
        const int cov = cover[Bf];
        // the wave's own pass words [xw/32, xw/32 + 2*ue) go to its LDS row
        {
          const int k0 = xw >> 5;
#pragma unroll
          for (int q = 0; q < kWpl; ++q) {
            const int k = lane_id() * kWpl + q - k0;
            if (k >= 0 && k < 2 * ue) wpword[wave_id()][k] = make_uint2(wb[q], (uint32_t)wpre[q]);
          }
          wave_lds_sync();
        }
        // A pass runs as one of U straight-line bodies, chosen by its slot
        // count ue.  Its slots go in groups of 2 to 4: a group's LDS reads,
        // then its segment reads, then its loads are each issued for the
        // whole group before the first result is used, and each group is
        // issued before the one before it is applied, so that two groups'
        // latencies always overlap.  With no branch between them the
        // compiler counts the waits exactly.
        auto body = [&](auto ue_c) {
          constexpr int UE = decltype(ue_c)::value;
          constexpr int NG = (UE + 3) / 4;  // groups of 3 or 4 slots (2 for 5 slots)
          // group i: G(i) slots from g0(i), sizes as even as possible
          auto gsize = [](int i) constexpr { return UE / NG + (i < UE % NG ? 1 : 0); };
          auto gstart = [](int i) constexpr { return i * (UE / NG) + (i < UE % NG ? i : UE % NG); };
          DenseItem<V> item[UE];
          bool ok[UE];
          auto issue = [&](auto gi_c) {
            constexpr int g0 = gstart(decltype(gi_c)::value);
            constexpr int G = gsize(decltype(gi_c)::value);
            // no segment starts at or past n, so a position past n finds
            // the segment of n - 1, which it then reads
            int xs[G];
            uint2 pw[G];
#pragma unroll
            for (int j = 0; j < G; ++j) {
              const int x = xb + (g0 + j) * kWave;
              ok[g0 + j] = x < n;  // only the last slot row runs past n
              xs[j] = x < n ? x : n - 1;
              pw[j] = wpword[wave_id()][((g0 + j) * kWave + lane_id()) >> 5];
            }
            SegBias<V> sg[G];
#pragma unroll
            for (int j = 0; j < G; ++j) {
              const int x = xb + (g0 + j) * kWave;
              const int c = cov + (int)pw[j].y + __popc(pw[j].x & ((2u << (x & 31)) - 1u));
              sg[j] = sb[c];
            }
#pragma unroll
            for (int j = 0; j < G; ++j) {
              const int32_t tx = t0 + xs[j];  // 32-bit: the clamp above stays 32-bit too
              const int64_t q = sg[j].bias + (int64_t)tx;
              item[g0 + j] = DenseItem<V>{(uint32_t)bcol[q], bval[q], sg[j].a};
            }
          };
          // hub columns recur: the bit is tested first, so that mostly unset
          // bits take the (same-word serialised) atomic OR; a group's adds
          // and bit reads go out before its first test.  Slots past n add
          // into the lane's sink column W + lane, outside every chunk, so
          // that no branch guards the adds
          auto apply = [&](auto gi_c) {
            constexpr int g0 = gstart(decltype(gi_c)::value);
            constexpr int G = gsize(decltype(gi_c)::value);
            uint32_t kk[G], bw[G];
            if constexpr (sizeof(V) == 8) {  // ds_add_f64, each add beside its bitmap read
#pragma unroll
              for (int j = 0; j < G; ++j) {
                kk[j] = ok[g0 + j] ? item[g0 + j].key : (uint32_t)(W + lane_id());
                atomicAdd(&dense[kk[j]], item[g0 + j].a * item[g0 + j].val);
                if constexpr (!PB) bw[j] = bm[kk[j] >> 5];
              }
            } else {  // the group's FP32 swaps back to back
              V pv[G];
#pragma unroll
              for (int j = 0; j < G; ++j) {
                kk[j] = ok[g0 + j] ? item[g0 + j].key : (uint32_t)(W + lane_id());
                pv[j] = item[g0 + j].a * item[g0 + j].val;
              }
              lds_add_group<G>(dense, kk, pv);
              if constexpr (!PB) {
#pragma unroll
                for (int j = 0; j < G; ++j) bw[j] = bm[kk[j] >> 5];
              }
            }
            if constexpr (!PB) {
#pragma unroll
              for (int j = 0; j < G; ++j) {
                const uint32_t bit = 1u << (kk[j] & 31);
                if (!(bw[j] & bit)) atomicOr(&bm[kk[j] >> 5], bit);
              }
            }
          };
          issue(std::integral_constant<int, 0>{});
          static_for<1, NG>([&](auto gi_c) {
            issue(gi_c);
            apply(std::integral_constant<int, decltype(gi_c)::value - 1>{});
          });
          apply(std::integral_constant<int, NG - 1>{});
        };
        static_assert(U >= 1 && U <= 16, "pass bodies for 1..16 slots");
        static_for<1, U + 1>([&](auto uc) {
          if (ue == decltype(uc)::value) body(uc);
        });
      }
      // the next window's B segment bounds (its A columns have arrived)
      st = 0;
      en = 0;
      a = na;
      if (npre) {
        const int64_t* nbp = bptr + (int64_t)ncw * ld;
        st = nbp[nk];
        en = nbp[nk + 1];
      }
    }
    if constexpr (PB) {
      if ((int)threadIdx.x < ((d.cwidth + 31) >> 5)) bm[threadIdx.x] = pbw;
    }
    __syncthreads();  // every product of the morsel is in dense/bm
    // rank and emit by waves: wave w owns words [w*wpw, w*wpw + wpw) of this
    // chunk's bitmap, 64 words at a time; its popcount goes through LDS
    // (one barrier).  A group of sparse words is emitted one set bit per
    // lane; otherwise the wave takes its non-zero words (found by a ballot)
    // two at a time, one bit per lane, and set bits get their rank by
    // ballot.  Either way each step stores a coalesced run of C, and each
    // value is read and cleared by one LDS exchange.
    const int mwords = (d.cwidth + 31) >> 5;
    const int wpw = (mwords + T / kWave - 1) / (T / kWave);
    const int wv0 = wave_id() * wpw;
    const int wv1 = (wv0 + wpw) < mwords ? (wv0 + wpw) : mwords;
    const uint32_t word0 = (wv0 + lane_id() < wv1) ? bm[wv0 + lane_id()] : 0u;
    int32_t pc = __popc(word0);
    for (int w = wv0 + kWave + lane_id(); w < wv1; w += kWave) pc += __popc(bm[w]);
    pc = wave_sum(pc);
    if (lane_id() == 0) s_rank[wave_id()] = (uint32_t)pc;
    __syncthreads();
    uint32_t utot;
    int32_t base = (int32_t)block_wave_prefix<T>(s_rank, utot);
    if (threadIdx.x == 0 && (int32_t)utot != d.cnt) atomicOr(err, 4);
    const int half = lane_id() >> 5, bit = lane_id() & 31;
    const unsigned long long below = (1ull << lane_id()) - 1;
    for (int g = wv0; g < wv1; g += kWave) {
      const uint32_t word = g == wv0 ? word0 : ((g + lane_id() < wv1) ? bm[g + lane_id()] : 0u);
      unsigned long long nz = __ballot(word != 0u);
      // lowest non-zero word of the group left in nz (its lane), 64 if none
      auto take = [&]() -> int {
        const int i = (int)__builtin_ctzg(nz, kWave);
        nz &= nz - 1;
        return i;
      };
      int32_t* colp = col_c + d.off;
      V* valp = val_c + d.off;
      const int32_t wcnt = __popc(word);
      const int32_t winc = wave_incl_scan(wcnt);
      const int32_t gtot = wave_last(winc);  // set bits of the group
      if (gtot < kSparseEmitBits * __popcll(nz)) {
        // sparse words (fewer than 512 set bits in the group): every lane
        // writes the (word, bit) codes of its word's set bits at their
        // ranks into the wave's staging row, then lane l emits the group's
        // set bit o0 + l from it.  The staging rows are the segment
        // descriptors' LDS, unused from the morsel's last pass until the
        // next morsel's first window scan, which every wave reaches only
        // after its emission.  Every step stores 64 consecutive entries.
        uint16_t* stg = (uint16_t*)sb + wave_id() * kSparseStage;
        {
          uint32_t y = word;
          int32_t r = winc - wcnt;
          while (y) {
            stg[r++] = (uint16_t)((lane_id() << 5) | __builtin_ctz(y));
            y &= y - 1u;
          }
        }
        wave_lds_sync();
        for (int32_t o0 = 0; o0 < gtot; o0 += kWave) {
          const int32_t o = o0 + lane_id();
          const bool valid = o < gtot;
          const uint32_t e = valid ? stg[o] : 0u;
          const int u = (g << 5) + (int)e;
          const int32_t r = base + o;
          V v = V(0);
          if (valid) v = lds_take(&dense[u]);
          if (valid && r < d.cnt) {
            stream_store(&colp[r], d.colbase + u);
            stream_store(&valp[r], v);
          }
        }
        wave_lds_sync();  // the staging row is read before the next group writes it
        base += gtot;
        continue;
      }
      while (nz) {
        const int i0 = take(), i1 = take(), i2 = take(), i3 = take();
        const int ia = half ? i1 : i0, ib = half ? i3 : i2;
        const uint32_t xa = ia < kWave ? bm[g + ia] : 0u, xb = ib < kWave ? bm[g + ib] : 0u;
        const bool seta = (xa >> bit) & 1u, setb = (xb >> bit) & 1u;
        const unsigned long long ma = __ballot(seta), mb = __ballot(setb);
        const int32_t ra = base + (int32_t)__popcll(ma & below);
        const int32_t rb = base + (int32_t)__popcll(ma) + (int32_t)__popcll(mb & below);
        base += (int32_t)(__popcll(ma) + __popcll(mb));
        const int ua = ((g + ia) << 5) + bit, ub = ((g + ib) << 5) + bit;
        V va = V(0), vb = V(0);
        if (seta) va = lds_take(&dense[ua]);
        if (setb) vb = lds_take(&dense[ub]);
        if (seta && ra < d.cnt) {
          stream_store(&colp[ra], d.colbase + ua);
          stream_store(&valp[ra], va);
        }
        if (setb && rb < d.cnt) {
          stream_store(&colp[rb], d.colbase + ub);
          stream_store(&valp[rb], vb);
        }
      }
    }
    wave_lds_sync();  // the wave's reads of its words are done
    for (int w = wv0 + lane_id(); w < wv1; w += kWave) bm[w] = 0u;
    d = nd;
    ndv = nndv;
    t_cur = t_nxt;
    t_nxt = t_nn;
  }
}

// Light morsels, one wave each.  A (row, chunk) morsel of few C entries
// costs the block kernel about as much as its fixed steps (window scan, pass
// set-up, three barriers, ranking): most of a G500 s19 ef32 row's 62 chunks
// hold a few hundred products.  Here a wave takes such a morsel alone, with
// no block barrier:
//   * the morsel's A nonzeros (at most kLightDeg, two a lane) and their B
//     segments in the chunk: a wave scan of the lengths, (bias, a) of the
//     non-empty segments in a per-wave table;
//   * products in rounds of 64*kLightU positions: the segments starting in
//     the round set bits in per-wave mark words, so position x belongs to
//     segment (segments begun before the round) + (marks <= x) - 1;
//   * each product a*b goes into a per-wave LDS hash table (linear probing,
//     claim by compare-and-swap, a lane's kLightU first probes issued
//     together; FP64 ds_add_f64, FP32 the compare-and-swap add) of the
//     fewest slots, 64 to kLightS, that the morsel's C entries fill at most
//     two-thirds, and its column into a chunk bitmap;
//   * output: prefix popcounts of the bitmap words give each present
//     column its rank, so every slot stores its entry at rank order (sorted
//     C without a sort); slots, bitmap and marks are cleared as they are
//     read.
// Morsels go through a three-stage pipeline: while morsel i runs, the B
// segment bounds of morsel i+1, the A entries of morsel i+2 and the
// descriptor of morsel i+3 (of the wave's stride) are in flight.  Per-wave
// dynamic LDS (light_wave_lds):
//   uint32 keys[kLightS] | V vals[kLightS] | uint32 bm[BW rounded to 4] |
//   2 KB: SegBias<V> seg[kLightDeg] (products), uint16 prefix[BW] (output) | u64 marks[kLightU]
#ifndef AMS_LIGHT_LG
#define AMS_LIGHT_LG 9
#endif
#ifndef AMS_LIGHT_CAP_LG
#define AMS_LIGHT_CAP_LG 9
#endif
constexpr int kLightT = 256;
constexpr int kLightLg = AMS_LIGHT_LG;
constexpr int kLightS = 1 << kLightLg;
constexpr int kLightCap = 1 << AMS_LIGHT_CAP_LG;  // values of a rank-addressed (PB) light morsel
constexpr int kLightMaxCnt = (kLightS * 2) / 3;
constexpr int kLightMaxWords = 512;  // chunks of up to 16384 columns
constexpr int kLightU = 4;
constexpr int kLightSegBytes = kLightDeg * 16;
static_assert(kLightDeg == 2 * kWave, "two A nonzeros a lane");
static_assert(kLightMaxWords * 2 <= kLightSegBytes, "word prefixes fit the segment table's bytes");

// PB (the symbolic phase kept the rows' column bitmaps): no hash table; the
// wave loads its chunk's bitmap words and their rank prefixes into a uint2
// array, and each product adds at its column's rank:
//   uint2 wp[BW rounded to 2] | V vals[kLightCap] | 2 KB segment table | u64 marks[kLightU]
__host__ __device__ constexpr size_t light_wave_lds(int W, int vbytes, bool pb) {
  return pb ? (size_t)((((W + 31) >> 5) + 1) & ~1) * 8 + (size_t)kLightCap * vbytes + kLightSegBytes +
                  (size_t)kLightU * 8
            : (size_t)kLightS * 4 + (size_t)kLightS * vbytes + (size_t)((((W + 31) >> 5) + 3) & ~3) * 4 +
                  kLightSegBytes + (size_t)kLightU * 8;
}
__host__ __device__ constexpr size_t light_block_lds(int W, int vbytes, bool pb) {
  return light_wave_lds(W, vbytes, pb) * (kLightT / kWave);
}
constexpr int kLightPbWpl = kLightMaxWords / kWave;  // bitmap words a lane (PB)

template <typename V, bool PB>
__global__ __launch_bounds__(kLightT) void k_numc_light(Csr A, int64_t mB, const int64_t* bptr,
                                                        const uint16_t* bcol, const V* bval,
                                                        const DenseTask* tasks, const int64_t* __restrict__ d_ntask,
                                                        int W, const int64_t* rp_c, int32_t* col_c, V* val_c,
                                                        int64_t cap, Probs pr, int32_t* err) {
  extern __shared__ __align__(16) unsigned char dyn_raw[];
  if (rp_c[A.m] > cap) return;
  const int BW = (W + 31) >> 5;
  unsigned char* mine = dyn_raw + light_wave_lds(W, (int)sizeof(V), PB) * wave_id();
  // !PB: keys | vals | bm | tail; PB: wp | vals | tail
  uint32_t* keys = (uint32_t*)mine;
  uint2* wp = (uint2*)mine;
  V* vals = PB ? (V*)(wp + ((BW + 1) & ~1)) : (V*)(keys + kLightS);
  uint32_t* bm = (uint32_t*)(vals + kLightS);
  unsigned char* tail = PB ? (unsigned char*)(vals + kLightCap) : (unsigned char*)(bm + ((BW + 3) & ~3));
  SegBias<V>* seg = (SegBias<V>*)tail;
  uint16_t* pre = (uint16_t*)tail;
  unsigned long long* marks = (unsigned long long*)(tail + kLightSegBytes);
  const int lane = lane_id();
  for (int i = lane; i < (PB ? kLightCap : kLightS); i += kWave) {
    if constexpr (!PB) keys[i] = 0xffffffffu;
    vals[i] = V(0);
  }
  if constexpr (!PB) {
    for (int i = lane; i < BW; i += kWave) bm[i] = 0u;
  }
  if (lane < kLightU) marks[lane] = 0ull;
  wave_lds_sync();
  // PB: lane l holds the chunk's bitmap words [l*wpl, l*wpl + wpl)
  const int wpl = (BW + kWave - 1) / kWave;
  auto load_w = [&](const DenseTask& m, uint32_t (&ww)[kLightPbWpl]) {
    const int mw = (m.cwidth + 31) >> 5;
    const uint32_t* rb = pr.rowbits + (int64_t)m.h * pr.bwords + (int64_t)m.c * (W >> 5);
#pragma unroll
    for (int q = 0; q < kLightPbWpl; ++q) {
      const int w = lane * wpl + q;
      ww[q] = 0u;
      if (q < wpl && w < mw && m.deg > 0) ww[q] = rb[w];
    }
  };
  const int64_t nheavy = d_ntask[0], nlight = d_ntask[1];
  const unsigned long long nend = (unsigned long long)(nheavy + nlight);
  const int64_t ld = mB + 1;
  const V* av = (const V*)A.val;
  const int64_t nw = (int64_t)gridDim.x * (kLightT / kWave);
  int64_t t = (int64_t)blockIdx.x * (kLightT / kWave) + wave_id();
  if (t >= nlight) return;
  // descriptor of the wave's morsel t + i*nw (all zero past the end: deg 0)
  auto desc = [&](int64_t i) {
    const int64_t x = t + i * nw;
    return task_load(tasks, (unsigned long long)(nheavy + (x < nlight ? x : nlight)), nend);
  };
  // a morsel's A entries 2*lane + s and their B segments in its chunk
  auto load_a = [&](const DenseTask& m, int32_t (&k)[2], V (&a)[2]) {
#pragma unroll
    for (int s2 = 0; s2 < 2; ++s2) {
      const int e = 2 * lane + s2;
      k[s2] = 0;
      a[s2] = V(0);
      if (e < m.deg) {
        k[s2] = A.col[m.p0 + e];
        a[s2] = av[m.p0 + e];
      }
    }
  };
  // (the segment [st, en); lengths are taken at use, so that no wait
  // follows the load)
  auto load_b = [&](const DenseTask& m, const int32_t (&k)[2], int64_t (&st)[2], int64_t (&en)[2]) {
    const int64_t* bp = bptr + (int64_t)m.c * ld;
#pragma unroll
    for (int s2 = 0; s2 < 2; ++s2) {
      st[s2] = 0;
      en[s2] = 0;
      if (2 * lane + s2 < m.deg) {
        st[s2] = bp[k[s2]];
        en[s2] = bp[k[s2] + 1];
      }
    }
  };
  // the morsel's table: 2^lg slots, the fewest that its C entries fill at
  // most two-thirds
  auto table_lg = [](int32_t cnt) {
    int lg = 6;
    while (lg < kLightLg && (1 << lg) * 2 < cnt * 3) ++lg;
    return lg;
  };
  // stage set-up: morsel 0 with its B bounds, morsel 1 with its A entries,
  // morsel 2's descriptor
  DenseTask d0 = task_get(desc(0));
  int32_t k0[2], k1[2];
  V a0[2], a1[2];
  int64_t st0[2], en0[2];
  load_a(d0, k0, a0);
  load_b(d0, k0, st0, en0);
  uint32_t ww0[kLightPbWpl], ww1[kLightPbWpl];
  if constexpr (PB) load_w(d0, ww0);
  DenseTask d1 = task_get(desc(1));
  load_a(d1, k1, a1);
  uint32_t dv2 = desc(2);
  for (; t < nlight; t += nw) {
    // in flight while this morsel runs: morsel 1's B bounds, morsel 2's A
    // entries, morsel 3's descriptor
    // (morsel 2's descriptor is read first: its wait then covers only
    // loads issued a morsel ago)
    const DenseTask d2 = task_get(dv2);
    int64_t st1[2], en1[2];
    load_b(d1, k1, st1, en1);
    if constexpr (PB) load_w(d1, ww1);
    int32_t k2[2];
    V a2[2];
    load_a(d2, k2, a2);
    const uint32_t dv3 = desc(3);
    const DenseTask& d = d0;
    const int lg = table_lg(d.cnt);
    const uint32_t tmask = (1u << lg) - 1u;
    int32_t nbits = 0;  // PB: set bits of the chunk
    if constexpr (PB) {
      // each word's rank prefix: a wave scan of the lanes' popcounts
      int32_t sum = 0;
#pragma unroll
      for (int q = 0; q < kLightPbWpl; ++q) sum += __popc(ww0[q]);
      const int32_t inc = wave_incl_scan(sum);
      int32_t run = inc - sum;
#pragma unroll
      for (int q = 0; q < kLightPbWpl; ++q) {
        const int w = lane * wpl + q;
        if (q < wpl && w < BW) wp[w] = make_uint2(ww0[q], (uint32_t)run);
        run += __popc(ww0[q]);
      }
      nbits = wave_last(inc);
      wave_lds_sync();
    }
    {
      int32_t ln0[2];
#pragma unroll
      for (int s2 = 0; s2 < 2; ++s2) ln0[s2] = (int32_t)(en0[s2] - st0[s2]);
      const int32_t lt = ln0[0] + ln0[1];
      const int32_t incl = wave_incl_scan(lt);
      const int32_t total = wave_last(incl);
      int32_t off[2];
      off[0] = incl - lt;
      off[1] = off[0] + ln0[0];
      const unsigned long long ne0 = __ballot(ln0[0] > 0), ne1 = __ballot(ln0[1] > 0);
      const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(ne0 >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)ne0, 0u)) +
                             __builtin_amdgcn_mbcnt_hi((uint32_t)(ne1 >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)ne1, 0u));
      if (ln0[0] > 0) seg[below] = SegBias<V>{st0[0] - off[0], a0[0]};
      if (ln0[1] > 0) seg[below + (ln0[0] > 0 ? 1 : 0)] = SegBias<V>{st0[1] - off[1], a0[1]};
      for (int32_t r0 = 0; r0 < total; r0 += kWave * kLightU) {
#pragma unroll
        for (int s2 = 0; s2 < 2; ++s2) {
          if (ln0[s2] > 0 && off[s2] >= r0 && off[s2] < r0 + kWave * kLightU) {
            const int32_t x = off[s2] - r0;
            atomicOr(&marks[x >> 6], 1ull << (x & 63));
          }
        }
        // non-empty segments begun before the round
        const int32_t cov = (int32_t)(__popcll(__ballot(ln0[0] > 0 && off[0] < r0)) +
                                      __popcll(__ballot(ln0[1] > 0 && off[1] < r0)));
        wave_lds_sync();
        unsigned long long mk[kLightU];
#pragma unroll
        for (int u = 0; u < kLightU; ++u) mk[u] = marks[u];
        wave_lds_sync();
        if (lane < kLightU) marks[lane] = 0ull;
        const unsigned long long le = ~0ull >> (63 - lane);  // lanes 0..lane
        uint32_t key[kLightU];
        V pv[kLightU];
        bool ok[kLightU];
        {
          SegBias<V> sg[kLightU];
          int32_t cum = cov;
#pragma unroll
          for (int u = 0; u < kLightU; ++u) {
            const int32_t x = r0 + u * kWave + lane;
            ok[u] = x < total;
            int32_t c = cum + (int32_t)__popcll(mk[u] & le) - 1;
            c = ok[u] ? c : 0;
            cum += (int32_t)__popcll(mk[u]);
            sg[u] = seg[c];
          }
#pragma unroll
          for (int u = 0; u < kLightU; ++u) {
            const int32_t x = r0 + u * kWave + lane;
            const int64_t q = ok[u] ? sg[u].bias + (int64_t)x : 0;
            key[u] = bcol[q];
            pv[u] = bval[q];
          }
#pragma unroll
          for (int u = 0; u < kLightU; ++u) pv[u] = sg[u].a * pv[u];
        }
        if constexpr (PB) {  // each product adds at its column's rank
          uint2 wr[kLightU];
#pragma unroll
          for (int u = 0; u < kLightU; ++u) wr[u] = wp[ok[u] ? key[u] >> 5 : 0];
#pragma unroll
          for (int u = 0; u < kLightU; ++u) {
            const uint32_t r = wr[u].y + __popc(wr[u].x & ((1u << (key[u] & 31)) - 1u));
            if (ok[u] && r < (uint32_t)kLightCap) lds_add(&vals[r], pv[u]);
          }
          continue;
        }
        // first probes of all slots at once, then the collisions one by one
        uint32_t h[kLightU], old[kLightU];
#pragma unroll
        for (int u = 0; u < kLightU; ++u) {
          h[u] = (key[u] * kHashMul) >> (32 - lg);
          old[u] = ok[u] ? atomicCAS(&keys[h[u]], 0xffffffffu, key[u]) : 0u;
        }
#pragma unroll
        for (int u = 0; u < kLightU; ++u) {
          if (!ok[u]) continue;
          uint32_t hh = h[u], o = old[u];
          bool done = false;
          for (uint32_t probe = 0; probe <= tmask; ++probe) {
            if (o == 0xffffffffu) atomicOr(&bm[key[u] >> 5], 1u << (key[u] & 31));
            if (o == 0xffffffffu || o == key[u]) {
              lds_add(&vals[hh], pv[u]);
              done = true;
              break;
            }
            hh = (hh + 1) & tmask;
            o = atomicCAS(&keys[hh], 0xffffffffu, key[u]);
          }
          if (!done) atomicOr(err, 0x10);
        }
      }
      wave_lds_sync();  // the segment table is read before the ranks overwrite it
    }
    if constexpr (PB) {
      // values are at their ranks already; the columns from the bitmap words
      if (lane == 0 && nbits != d.cnt) atomicOr(err, 4);
      int32_t* colp = col_c + d.off;
      V* valp = val_c + d.off;
      for (int i = lane; i < d.cnt; i += kWave) {
        stream_store(&valp[i], vals[i]);
        vals[i] = V(0);
      }
      int32_t r = 0;
      {
        int32_t sum = 0;
#pragma unroll
        for (int q = 0; q < kLightPbWpl; ++q) sum += __popc(ww0[q]);
        r = wave_incl_scan(sum) - sum;
      }
      // each set bit's chunk column goes to its rank in the staging row (the
      // segment table, read by now), then the lanes store consecutive
      // entries: coalesced, where the bit loop's own stores were scattered
#pragma unroll
      for (int q = 0; q < kLightPbWpl; ++q) {
        uint32_t x = ww0[q];
        const int32_t cl = (lane * wpl + q) * 32;
        while (x) {
          if (r < d.cnt) pre[r] = (uint16_t)(cl + __builtin_ctz(x));
          ++r;
          x &= x - 1u;
        }
      }
      wave_lds_sync();
      for (int i = lane; i < d.cnt; i += kWave) stream_store(&colp[i], d.colbase + (int32_t)pre[i]);
      wave_lds_sync();  // vals, wp and the staging row are read before the next morsel writes them
      d0 = d1;
      d1 = d2;
      dv2 = dv3;
#pragma unroll
      for (int s2 = 0; s2 < 2; ++s2) {
        st0[s2] = st1[s2];
        en0[s2] = en1[s2];
        a0[s2] = a1[s2];
        k1[s2] = k2[s2];
        a1[s2] = a2[s2];
      }
#pragma unroll
      for (int q = 0; q < kLightPbWpl; ++q) ww0[q] = ww1[q];
      continue;
    }
    // ranks: lane l sums the popcounts of bitmap words [l*wpl, l*wpl + wpl),
    // one wave scan, then each word's prefix goes over the segment table
    const int wb = lane * wpl;
    int32_t sum = 0;
    for (int q = 0; q < wpl; ++q)
      if (wb + q < BW) sum += __popc(bm[wb + q]);
    const int32_t inc = wave_incl_scan(sum);
    int32_t run = inc - sum;
    for (int q = 0; q < wpl; ++q) {
      if (wb + q < BW) {
        pre[wb + q] = (uint16_t)run;
        run += __popc(bm[wb + q]);
      }
    }
    if (lane == 0 && wave_last(inc) != d.cnt) atomicOr(err, 4);
    wave_lds_sync();
    int32_t* colp = col_c + d.off;
    V* valp = val_c + d.off;
    const int nu = 1 << (lg - 6);  // 64-slot rows of the morsel's table
    uint32_t kx[kLightS / kWave];
#pragma unroll
    for (int u = 0; u < kLightS / kWave; ++u) kx[u] = u < nu ? keys[u * kWave + lane] : 0xffffffffu;
#pragma unroll
    for (int u = 0; u < kLightS / kWave; ++u) {
      const uint32_t key = kx[u];
      if (key != 0xffffffffu) {
        const int i = u * kWave + lane;
        const uint32_t wd = bm[key >> 5];
        const int32_t r = (int32_t)pre[key >> 5] + __popc(wd & ((1u << (key & 31)) - 1u));
        const V v = vals[i];
        if (r < d.cnt) {
          colp[r] = d.colbase + (int32_t)key;
          valp[r] = v;
        }
        keys[i] = 0xffffffffu;
        vals[i] = V(0);
      }
    }
    wave_lds_sync();  // every rank is read before the bitmap is cleared
    for (int i = lane; i < BW; i += kWave) bm[i] = 0u;
    wave_lds_sync();
    // the stages move on (their loads have had this morsel's time)
    d0 = d1;
    d1 = d2;
    dv2 = dv3;
#pragma unroll
    for (int s2 = 0; s2 < 2; ++s2) {
      st0[s2] = st1[s2];
      en0[s2] = en1[s2];
      a0[s2] = a1[s2];
      k1[s2] = k2[s2];
      a1[s2] = a2[s2];
    }
  }
}

}  // namespace ams

namespace ams {

__global__ void k_copy_i64(const int64_t* src, int64_t* dst) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *dst = *src;
}

__global__ void k_copy_i32(const int32_t* src, int32_t* dst) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *dst = *src;
}

// Work per route, for reports and per-kernel rooflines:
//   sums[r]      products of numeric route r (0 hash, 1 ESC, 2 dense, 3 global hash)
//   sums[4 + r]  nnz(C) of route r, sums[8 + r] nnz(A) of route r
//   sums[12 + s] products of symbolic route s (0 hash, 1 bitmap, 2 global hash)
//   sums[15 + s] nnz(A) of symbolic route s
constexpr int kRouteSums = 18;
__global__ __launch_bounds__(256) void k_route_sums(int64_t m, const int64_t* row_flp, const int64_t* rp_c,
                                                    const int64_t* rp_a, const uint8_t* sbin,
                                                    const uint8_t* nbin, unsigned long long* sums) {
  int64_t acc[kRouteSums];
#pragma unroll
  for (int i = 0; i < kRouteSums; ++i) acc[i] = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t f = row_flp[i], z = rp_c[i + 1] - rp_c[i], d = rp_a[i + 1] - rp_a[i];
    const int nb = nbin[i], sb = sbin[i];
    const int r = nb == NB_NONE ? -1
                                : (nb == NB_GHASH ? 3 : (nb == NB_DENSE ? 2 : (nb >= kNumHashBins ? 1 : 0)));
    if (r >= 0) {
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        acc[k] += (r == k) ? f : 0;
        acc[4 + k] += (r == k) ? z : 0;
        acc[8 + k] += (r == k) ? d : 0;
      }
    }
    if (sb != SB_NONE) {
      const int q = sb == SB_GHASH ? 2 : (sb == SB_BITMAP ? 1 : 0);
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        acc[12 + k] += (q == k) ? f : 0;
        acc[15 + k] += (q == k) ? d : 0;
      }
    }
  }
  // waves add into the block's LDS sums, then one global add a sum a block:
  // every block's adds go to the same two cache lines, where they serialise
  __shared__ unsigned long long s_sum[kRouteSums];
  if (threadIdx.x < kRouteSums) s_sum[threadIdx.x] = 0ull;
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kRouteSums; ++k) {
    const int64_t v = wave_sum(acc[k]);
    if (lane_id() == 0 && v) atomicAdd(&s_sum[k], (unsigned long long)v);
  }
  __syncthreads();
  if (threadIdx.x < kRouteSums && s_sum[threadIdx.x]) atomicAdd(&sums[threadIdx.x], s_sum[threadIdx.x]);
}

// Checksums of C: out[0] += sum of columns, out[1] += rows whose columns
// are not strictly ascending; vsum += sum of values.
template <typename V>
__global__ __launch_bounds__(256) void k_verify(Csr C, unsigned long long* out, double* vsum) {
  int64_t scol = 0, bad = 0;
  double sval = 0.0;
  const V* cv = (const V*)C.val;
  for (int64_t r = blockIdx.x * 4 + wave_id(); r < C.m; r += (int64_t)gridDim.x * 4) {
    const int64_t s = C.rp[r], e = C.rp[r + 1];
    int64_t unsorted = 0;
    for (int64_t q = s + lane_id(); q < e; q += kWave) {
      scol += C.col[q];
      sval += (double)cv[q];
      if (q > s && C.col[q] <= C.col[q - 1]) unsorted = 1;
    }
    bad += wave_sum(unsorted) > 0 ? 1 : 0;
  }
  scol = wave_sum(scol);
  sval = wave_sum(sval);
  if (lane_id() == 0) {
    atomicAdd(&out[0], (unsigned long long)scol);
    if (bad) atomicAdd(&out[1], (unsigned long long)bad);
    atomicAdd(vsum, sval);
  }
}

}  // namespace ams
