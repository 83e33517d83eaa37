// ams_kernels.hpp -- the SpGEMM kernels (declarations of the templates are
// instantiated in ams_spgemm.hip and ams_bwd.hip).
//
// Data layout in HBM (all device arrays):
//   A, B, C   CSR: int64 row_ptr[m+1], int32 col[nnz], V val[nnz] (V = double|float)
//   row_flp   int64[m]   products per row of A*B
//   sym_bin   uint8[m]   symbolic route of each row, num_bin uint8[m] numeric route
//   lists     int32      rows grouped by bin (one segment per bin)
//   coff      int32[H*(R+1)]  per heavy row: exclusive prefix of nnz over the
//             R column chunks of width W, coff[h*(R+1)+R] = nnz of the row
//   bptr      int64[R*(mB+1)] chunk-major row pointers of the column-blocked
//             copy of B (bcol: uint16 column within chunk, bval: V)
//
// Problem table: a stacked batch is one block-diagonal product.  Problem i
// owns A rows [arow[i], arow[i+1]), B rows (= A columns) [brow[i], brow[i+1])
// and output columns [ccol[i], ccol[i+1]).  A single product is nprob = 1.
#pragma once

#include "ams_device.hpp"

namespace ams {

// Symbolic routes.
enum : uint8_t { SB_NONE = 0xff, SB_BITMAP = 8, SB_GHASH = 9 };  // 0..7: LDS hash, table 64 << b
constexpr int kSymHashBins = 8;
constexpr int kSymBins = 10;
// Numeric routes.
// 0..7 LDS hash (table 64 << b); 8..16 ESC, products per row at most
// esc_bin_cap(b - 8): 64, 128, 256, 512, 1024, 1536, 2048, 3072, 4096
enum : uint8_t { NB_NONE = 0xff, NB_DENSE = 17, NB_GHASH = 18 };
constexpr int kNumHashBins = 8;
constexpr int kNumEscBins = 9;
constexpr int kNumBins = 19;
constexpr int kMaxHashSlots = 64 << (kNumHashBins - 1);  // 8192
constexpr int kMaxEsc = 4096;
__host__ __device__ constexpr int esc_bin_cap(int i) { return i <= 4 ? 64 << i : (i & 1 ? 1536 : 2048) << ((i - 5) >> 1); }
// ESC bin of a row of f products, 1 <= f <= kMaxEsc
__host__ __device__ inline int esc_bin_of(int64_t f) {
  int i = 0;
  while (esc_bin_cap(i) < f) ++i;
  return i;
}

struct Csr {
  int64_t m, n, nnz;
  const int64_t* rp;
  const int32_t* col;
  const void* val;
};

struct Probs {
  int32_t n;              // number of problems; n <= 1: one product, the
                          // table pointers are unused
  const int64_t* arow;    // n + 1
  const int64_t* brow;    // n + 1
  const int64_t* ccol;    // n + 1
  int64_t width0;         // output columns when n <= 1
  int64_t max_width;      // max output columns of one problem
  const int2* arow_cols;  // per A row: (column base, width) of its problem, or null
  const int32_t* brow_lo; // per B row: column base of its problem, or null
  uint32_t* rowbits;      // column bitmaps of the bitmap bin's rows (row h at h * bwords), or null
  int64_t bwords;         // words per row in rowbits
};

__device__ __forceinline__ int find_prob(const int64_t* off, int n, int64_t r) {
  if (n <= 1) return 0;
  int lo = 0, hi = n;  // off[lo] <= r < off[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (off[mid] <= r) lo = mid; else hi = mid;
  }
  return lo;
}

// Column base and width of the problem that owns A row `row`.
__device__ __forceinline__ void prob_cols(const Probs& pr, int64_t row, int64_t& lo, int64_t& width) {
  if (pr.n <= 1) { lo = 0; width = pr.width0; return; }
  if (pr.arow_cols) {
    const int2 v = pr.arow_cols[row];
    lo = v.x;
    width = v.y;
    return;
  }
  const int pi = find_prob(pr.arow, pr.n, row);
  lo = pr.ccol[pi];
  width = pr.ccol[pi + 1] - lo;
}

// Column base of the problem that owns B row k.
__device__ __forceinline__ int64_t prob_bcol_lo(const Probs& pr, int64_t k) {
  if (pr.n <= 1) return 0;
  if (pr.brow_lo) return pr.brow_lo[k];
  return pr.ccol[find_prob(pr.brow, pr.n, k)];
}

struct SymParams {
  int64_t sym_hash_max;
  double bitmap_ratio;
  int64_t bitmap_bits;  // usable LDS bitmap width (0 disables the bitmap route)
  double ghash_ratio;   // global hash for rows with sym_hash_max < f < ratio * words (0: off)
};

struct NumParams {
  int64_t num_hash_max;
  int64_t esc_max;
  double acc_thres;
  double dense_density;
  int32_t mode;    // 0 hybrid, 1 hash, 2 esc
  int32_t vbytes;  // bytes of a value (global-hash table sizes)
};

// Slots of a global hash table for n keys: a power of two >= 1.5 n.
__device__ __forceinline__ int64_t ghash_slots(int64_t n) {
  const int64_t want = n + ((n + 1) >> 1);
  return want <= 2 ? 2 : (int64_t)1 << (64 - __clzll((unsigned long long)(want - 1)));
}

// Bytes of a row's numeric global-hash region: keys and values of the
// table, then the z columns and values the sort passes through.
__device__ __forceinline__ int64_t ghash_num_bytes(int64_t z, int vbytes) {
  const int64_t s = ghash_slots(z);
  return ((s * (4 + vbytes) + z * (4 + vbytes) + 8) + 15) & ~(int64_t)15;
}

}  // namespace ams
