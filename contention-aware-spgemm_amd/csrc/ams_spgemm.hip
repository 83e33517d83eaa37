// ams_spgemm.hip -- host side of the device API (include/ams_hip.h): the
// context, scratch management, routing and launch of every phase.
//
// Plain product, one stream:
//   symbolic: flops -> classify -> bins -> [1 sync: bin sizes] -> hash /
//             bitmap counts -> scan -> numeric classify -> [1 sync: nnz]
//   numeric:  [column-blocked B + dense queue kernel] + hash bins + ESC bins
// The batched form runs the same kernels with grids sized from upper
// bounds, so no host synchronisation is needed after its plan.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <mutex>
#include <vector>

#include "ams_hip.h"
#include "ams_kernels_impl.hpp"
#include "ams_bwd_impl.hpp"

using namespace ams;

namespace {

struct Buf {
  void* p = nullptr;
  size_t cap = 0;
};

// Control block in device memory (zeroed by a fill kernel, never memset,
// so that the sequence can be captured into a hipGraph).
struct Ctrl {
  int32_t cnt_sym[32];
  int32_t cur_sym[32];
  int32_t cnt_num[32];
  int32_t cur_num[32];
  int32_t err;
  int32_t pad;
  unsigned long long flops;
  unsigned long long ticket;
  int64_t ntask[2];  // morsels of the dense block kernel, light morsels
  int32_t nlong;  // long B rows listed for the wave-per-row blocked copy
  int32_t pad1;
  unsigned long long sums[24];     // route sums [0, 18), checksums of C [20, 23)
  int32_t nheavy;   // heavy rows of the bitmap bin split into pieces
  int32_t npieces;
  unsigned long long sym_ticket;  // next row of the bitmap bin to count
  unsigned long long ghash_sym_bytes;  // global-hash tables of the symbolic route
  unsigned long long ghash_num_bytes;  // global-hash regions of the numeric route
  unsigned long long ghash_cursor;     // next free table slot / byte
};
static_assert(sizeof(Ctrl) % 8 == 0, "ctrl");

#ifndef AMS_SYMB_T
#define AMS_SYMB_T 1024
#endif
constexpr int kSymBitmapT = AMS_SYMB_T;
// Batches of narrow problems (a bitmap pass of at most 64 Ki bits, 8 KiB of
// LDS) count their wide rows in 256-thread blocks: eight rows per CU in
// flight instead of two, for the many rows of a few thousand products that
// do not fill a 1024-thread block (the 512-product batch of R-MAT scales
// 8..14: 83.5 -> 77.3 ms).  A single problem keeps the 1024-thread blocks:
// with few such rows the longest row's latency decides (G500 s16 ef16
// 9.0 ms against 9.4 ms with 256 threads).
#ifndef AMS_SYMB_NARROW_T
#define AMS_SYMB_NARROW_T 256
#endif
#ifndef AMS_SYMB_NARROW_BITS
#define AMS_SYMB_NARROW_BITS 65536
#endif
constexpr int kSymBitmapNarrowT = AMS_SYMB_NARROW_T;
constexpr int kSymNarrowBits = AMS_SYMB_NARROW_BITS;
constexpr int kDenseT = 512;
constexpr size_t kLdsPerCu = 160 * 1024;  // gfx950

// Threads of an LDS-hash block: S/8, at most AMS_HASH_TMAX.  The 8192-slot
// FP64 table takes 141 KB of LDS, one block a CU: 1024 threads give it 16
// waves instead of 8 (G500 s19 ef32 hash numeric 6.48 -> 5.24 ms).
#ifndef AMS_HASH_TMAX
#define AMS_HASH_TMAX 1024
#endif
#ifndef AMS_HASH_TDIV
#define AMS_HASH_TDIV 8
#endif
template <int S>
constexpr int hash_threads() {
  return S / AMS_HASH_TDIV < 64 ? 64 : (S / AMS_HASH_TDIV > AMS_HASH_TMAX ? AMS_HASH_TMAX : S / AMS_HASH_TDIV);
}
// Threads of an ESC block: eight keys a thread (E / 8), at least one wave.
template <int E>
constexpr int esc_threads() { return E / 8 < 64 ? 64 : E / 8; }

// The symbolic hash keeps keys only (4S bytes of LDS): half the threads of
// the numeric table's block, so twice the rows per CU are in flight at the
// 32-wave limit (ER s20 ef16: 7.0 -> 6.0 ms; a quarter: 9.8 ms).
template <int S>
constexpr int symb_hash_threads() { return S / 16 < 64 ? 64 : (S / 16 > 512 ? 512 : S / 16); }

}  // namespace

// Fixed scratch of ams_spgemm_fixed, for callers that cannot allocate: one
// array per GPU in device memory, reserved when the library's code object
// loads on that GPU and held until the process ends.  Host code never uses
// this name as a device address (on the host it is a stand-in symbol): it
// takes the GPU's address from hipGetSymbolAddress once per device, with the
// device current, and uses that address only.
constexpr size_t kFixedScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kFixedScratchBytes];

constexpr int kMaxDevices = 64;
static void* g_scratch_addr[kMaxDevices];
static std::mutex g_scratch_mu;

// Device address of the fixed scratch on `device`, which must be current.
static hipError_t fixed_scratch_address(int device, void** out) {
  *out = nullptr;
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_addr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_addr[device] = p;
  }
  *out = g_scratch_addr[device];
  return hipSuccess;
}

struct ams_ctx {
  hipStream_t stream = nullptr;
  int device = 0;
  int num_cu = 256;
  size_t lds_max = 65536;
  ams_params_t prm{};
  ams_stats_t stats{};
  std::string err;
  Buf row_flp, sbin, nbin, lsym, lnum, ctrl, coff, bptr, bcol, bval, bperm, tsum;
  Buf pool_rp, pool_col, pool_val, probs, prow, pbrow, scratch_a, dflag, dtask, ccnt, lrows;
  Buf hflag, hrow, hpiece, gbm;  // heavy bitmap rows: flags, slots, pieces, global bitmaps
  Buf gtab, gnum;                // global-hash tables: symbolic keys, numeric regions
  Buf bgrad;                     // backward: dB partials of the dense bin in blocked-B order
  Buf rowbits;                   // column bitmaps of the bitmap bin's rows, kept for the numeric phase
  size_t mem_total = 0;          // the GPU's memory
  int64_t ghash_num_bytes = 0;   // of the last symbolic plan
  size_t gbm_zero = 0;           // bytes of gbm known to be zero
  bool bblk_ready = false;  // blocked B of the current plan already built
  Ctrl* h_ctrl = nullptr;       // pinned mirror of the control block
  int64_t* h_pin = nullptr;     // pinned scratch
  // plan of the last symbolic call
  bool plan_valid = false;
  bool plan_counts = false;     // bin sizes known on the host
  int64_t plan_m = 0, plan_mB = 0, plan_nnzB = 0, plan_H = 0;
  int plan_R = 0, plan_W = 0, plan_Wb = 0, plan_dtype = 0;
  Probs plan_pr{};
  int32_t cnt_sym[kSymBins] = {}, cnt_num[kNumBins] = {};
  // event timing of phases
  bool timing = false;
  std::vector<hipEvent_t> ev_free;
  struct Rec { int ph; hipEvent_t a, b; };
  std::vector<Rec> recs;
  double acc_ms[AMS_PH_COUNT] = {};
  int64_t acc_n[AMS_PH_COUNT] = {};
  ams_params_t plan_prm{};       // parameters the plan was made with
  // operands of the plan: the numeric phase must see the same matrices
  int64_t plan_nnzA = 0, plan_nB = 0;
  const void *plan_arp = nullptr, *plan_acol = nullptr, *plan_brp = nullptr, *plan_bcol = nullptr;
  // batched plan
  bool batch_ready = false;
  Probs batch_pr{};
  int64_t batch_m = 0, batch_mB = 0, batch_nnzA = 0, batch_nnzB = 0;
  int batch_dtype = 0;
  // Generation of the scratch layout: bumped whenever a buffer moves
  // (reallocation) or a batched plan rewrites its problem tables.  A
  // captured graph bakes in buffer addresses and replays only while the
  // generation it was captured at is current.
  uint64_t gen = 1;
  bool capturing = false;  // inside ams_batched_capture: buffers must not move
  bool fixed = false;      // scratch carved from the fixed device array: nothing may grow
  unsigned char* fixed_base = nullptr;  // device address of this GPU's fixed scratch
};

// Handle of a captured batched run: the executable graph, the context it
// belongs to and the scratch generation its kernel arguments point into.
struct AmsGraph {
  hipGraphExec_t exec;
  const ams_ctx_t* ctx;
  uint64_t gen;
};

static int fail(ams_ctx_t* ctx, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (ctx) ctx->err = buf;
  return code;
}

#define HIPCHK(x)                                                                        \
  do {                                                                                   \
    hipError_t e_ = (x);                                                                 \
    if (e_ != hipSuccess)                                                                \
      return fail(ctx, AMS_ERR_HIP, "%s: %s (%s:%d)", #x, hipGetErrorString(e_), __FILE__, \
                  __LINE__);                                                             \
  } while (0)

#define CHK(x)              \
  do {                      \
    int r_ = (x);           \
    if (r_ != AMS_OK) return r_; \
  } while (0)

static int grow(ams_ctx_t* ctx, Buf& b, size_t bytes) {
  if (bytes == 0) bytes = 8;
  if (b.cap >= bytes) return AMS_OK;
  if (ctx->fixed)
    return fail(ctx, AMS_ERR_LIMIT, "a scratch buffer of %zu bytes exceeds its share of the fixed scratch", bytes);
  if (ctx->capturing)
    return fail(ctx, AMS_ERR_STATE, "a scratch buffer of %zu bytes would move during graph capture; "
                "call ams_batched_plan for these inputs first", bytes);
  if (b.p) {
    ++ctx->gen;  // the buffer moves: graphs captured before are stale
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipFree(b.p));
    b.p = nullptr;
    b.cap = 0;
  }
  const size_t want = bytes < ((size_t)1 << 30) ? bytes + bytes / 8 : bytes;  // no slack on huge buffers
  hipError_t e = hipMalloc(&b.p, want);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    e = hipMalloc(&b.p, bytes);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      b.p = nullptr;
      return fail(ctx, AMS_ERR_ALLOC, "hipMalloc of %zu bytes failed", bytes);
    }
    b.cap = bytes;
  } else {
    b.cap = want;
  }
  return AMS_OK;
}

static Csr to_csr(const ams_dcsr_t* m) {
  return Csr{m->num_row, m->num_col, m->num_nnz, m->row_ptr, m->col, m->val};
}

static int check_csr(ams_ctx_t* ctx, const ams_dcsr_t* m, const char* name) {
  if (!m) return fail(ctx, AMS_ERR_ARG, "%s is NULL", name);
  if (m->num_row < 0 || m->num_col < 0 || m->num_nnz < 0)
    return fail(ctx, AMS_ERR_ARG, "%s has a negative size", name);
  if (m->num_row >= INT32_MAX || m->num_col >= INT32_MAX)
    return fail(ctx, AMS_ERR_LIMIT, "%s: rows and columns must be < 2^31", name);
  if (!m->row_ptr) return fail(ctx, AMS_ERR_ARG, "%s has no row pointers", name);
  if (m->num_nnz > 0 && (!m->col || !m->val)) return fail(ctx, AMS_ERR_ARG, "%s has no entries", name);
  return AMS_OK;
}

template <typename K, typename... Args>
static int occupancy(K kernel, int threads, size_t dyn, int* out) {
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernel, threads, dyn) != hipSuccess) {
    (void)hipGetLastError();
    n = 1;
  }
  *out = n < 1 ? 1 : n;
  return AMS_OK;
}

static int grid_for(ams_ctx_t* ctx, int occ, int64_t work, bool known) {
  int64_t g = (int64_t)ctx->num_cu * occ;
  if (known) g = std::min<int64_t>(g, work);
  return (int)std::max<int64_t>(g, 1);
}

static int scan_i64(ams_ctx_t* ctx, const int64_t* in, int64_t* out, int64_t n) {
  if (n <= 0) return AMS_OK;
  const int64_t nt = (n + kScanTile - 1) / kScanTile;
  CHK(grow(ctx, ctx->tsum, sizeof(int64_t) * nt));
  int64_t* ts = (int64_t*)ctx->tsum.p;
  hipLaunchKernelGGL(k_scan_tiles, dim3((unsigned)nt), dim3(kScanT), 0, ctx->stream, in, n, ts);
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, ctx->stream, ts, nt);
  hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nt), dim3(kScanT), 0, ctx->stream, in, out, n,
                     (const int64_t*)ts);
  HIPCHK(hipGetLastError());
  return AMS_OK;
}

static int fill_i64(ams_ctx_t* ctx, int64_t* p, int64_t n, int64_t v) {
  if (n <= 0) return AMS_OK;
  const int g = (int)std::min<int64_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_fill_i64, dim3(g), dim3(256), 0, ctx->stream, p, n, v);
  HIPCHK(hipGetLastError());
  return AMS_OK;
}

static int fill_i32(ams_ctx_t* ctx, int32_t* p, int64_t n, int32_t v) {
  if (n <= 0) return AMS_OK;
  const int g = (int)std::min<int64_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_fill_i32, dim3(g), dim3(256), 0, ctx->stream, p, n, v);
  HIPCHK(hipGetLastError());
  return AMS_OK;
}

static Ctrl* dctrl(ams_ctx_t* ctx) { return (Ctrl*)ctx->ctrl.p; }

static hipEvent_t take_event(ams_ctx_t* ctx) {
  if (!ctx->ev_free.empty()) {
    hipEvent_t e = ctx->ev_free.back();
    ctx->ev_free.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  return e;
}

// Brackets the launches of one phase with events when timing is on.
struct Phase {
  ams_ctx_t* ctx;
  int ph;
  hipEvent_t a = nullptr;
  Phase(ams_ctx_t* c, int p, bool on) : ctx(c), ph(p) {
    if (on && ctx->timing) {
      a = take_event(ctx);
      if (a) (void)hipEventRecord(a, ctx->stream);
    }
  }
  ~Phase() {
    if (!a) return;
    hipEvent_t b = take_event(ctx);
    if (!b) return;
    (void)hipEventRecord(b, ctx->stream);
    ctx->recs.push_back({ph, a, b});
  }
};

// Blocks a CU the backward's dense chunks are narrowed to fit.
#ifndef AMS_BWD_MIN_BLOCKS
#define AMS_BWD_MIN_BLOCKS 2
#endif

// Geometry of the dense column chunks and the symbolic bitmap passes.
// dense_width is the FP64 chunk width; FP32 chunks are 7/4 as wide (the
// 4-byte accumulators leave the LDS room, and fewer, larger morsels pay
// less per-morsel overhead: 95.9 -> 85.0 ms of FP32 dense numeric on G500
// s18 ef16; twice as wide no longer fits two blocks per CU).
static void dense_geometry(ams_ctx_t* ctx, int dtype, int64_t max_width, int& W, int& R, int& Wb) {
  W = ctx->prm.dense_width;
  if (dtype == AMS_F32) W = (int)((int64_t)W * 7 / 4 / 64 * 64);
  if (max_width > 0 && W > max_width) W = (int)((max_width + 63) / 64 * 64);
  W = std::max(64, std::min(W, 65536));
  R = (int)((max_width + W - 1) / W);
  if (R < 1) R = 1;
  // one bitmap pass covers the whole row when that fits bitmap_bits
  int64_t wb = std::min<int64_t>((int64_t)R * W, (int64_t)ctx->prm.bitmap_bits / W * W);
  if (wb < W) wb = W;
  Wb = (int)wb;
}

// Words of a heavy row's global bitmap (the widest problem's columns), or 0
// when the kMaxHeavy bitmaps would take more than 512 MiB (no splitting).
static int64_t heavy_words(const Probs& pr) {
  const int64_t gwords = (std::max<int64_t>(pr.max_width, 1) + 31) / 32;
  return (size_t)gwords * 4 * kMaxHeavy <= ((size_t)512 << 20) ? gwords : 0;
}

// Buffers of the heavy-row split for a bitmap bin of up to H rows; the
// global bitmaps start zeroed and k_heavy_finish leaves them zero.
static int grow_heavy(ams_ctx_t* ctx, const Probs& pr, int64_t H) {
  CHK(grow(ctx, ctx->hflag, (size_t)std::max<int64_t>(H, 1)));
  CHK(grow(ctx, ctx->hrow, sizeof(int32_t) * kMaxHeavy));
  CHK(grow(ctx, ctx->hpiece, sizeof(HeavyPiece) * kMaxPieces));
  CHK(grow(ctx, ctx->gbm, (size_t)heavy_words(pr) * 4 * kMaxHeavy));
  if (ctx->gbm_zero < ctx->gbm.cap) {
    CHK(fill_i32(ctx, (int32_t*)ctx->gbm.p, (int64_t)(ctx->gbm.cap / 4), 0));
    ctx->gbm_zero = ctx->gbm.cap;
  }
  return AMS_OK;
}

// Dynamic LDS of k_numc_dense<T>: dense values, column bitmap (padded to
// 16 bytes), then the segment table of a window.
template <int T>
static size_t dense_lds(int W, int vbytes) {
  const size_t head = ((size_t)(W + 64) * vbytes + (size_t)((W + 64) / 32) * 4 + 15) & ~(size_t)15;
  return head + sizeof(SegBias<double>) * T;
}
// Static LDS of k_numc_dense<T> (pass bitmaps, prefix counts, scan and
// ticket scratch), bounded from above.
template <int T>
constexpr size_t dense_static_lds() { return 5 * T * dense_u<T>() / 8 + T / 2 + 64; }

// Dense numeric block: 512 threads while two blocks fit a CU's LDS, else
// 1024 threads so that one block still brings 16 waves.
static int dense_block(int W, int vbytes) {
#ifdef AMS_DENSE_FORCE_T
  return AMS_DENSE_FORCE_T;
#endif
  return dense_lds<512>(W, vbytes) + dense_static_lds<512>() <= (size_t)kLdsPerCu / 2 ? 512 : 1024;
}

// Row flops: a wave per row, or the nonzero split when rows are long on
// average (a wave per row would leave most of the GPU idle).
#ifndef AMS_FLP_NZ_AVG
#define AMS_FLP_NZ_AVG 256
#endif
static int launch_row_flops(ams_ctx_t* ctx, const Csr& A, const Csr& B, int64_t* row_flp,
                            unsigned long long* total) {
  if (A.m == 0) return AMS_OK;
  if (row_flp && A.nnz > (int64_t)AMS_FLP_NZ_AVG * A.m) {
    CHK(fill_i64(ctx, row_flp, A.m, 0));
    const int64_t threads = (A.nnz + kFlpIpt - 1) / kFlpIpt;
    const int g = (int)std::min<int64_t>((threads + 255) / 256, (int64_t)ctx->num_cu * 32);
    hipLaunchKernelGGL(k_row_flops_nz, dim3(g), dim3(256), 0, ctx->stream, A, B, (unsigned long long*)row_flp,
                       total);
  } else {
    if (A.nnz <= kFlpShort * A.m) {  // short rows: a lane each
      const int g = (int)std::min<int64_t>((A.m + 255) / 256, (int64_t)ctx->num_cu * 16);
      hipLaunchKernelGGL(k_row_flops, dim3(g), dim3(256), 0, ctx->stream, A, B, row_flp, total);
    } else {
      const int g = (int)std::min<int64_t>((A.m + 3) / 4, (int64_t)ctx->num_cu * 16);
      hipLaunchKernelGGL(k_row_flops_wave, dim3(g), dim3(256), 0, ctx->stream, A, B, row_flp, total);
    }
  }
  HIPCHK(hipGetLastError());
  return AMS_OK;
}

// Column-blocked copy of B (chunk-major CSR of width-W column blocks).
template <typename V>
static int build_bblk(ams_ctx_t* ctx, const Csr& B, const Probs& pr, int W, int R, bool perm) {
  const int64_t ld = B.m + 1;
  CHK(grow(ctx, ctx->bptr, sizeof(int64_t) * (size_t)R * ld));
  CHK(grow(ctx, ctx->bcol, sizeof(uint16_t) * std::max<int64_t>(B.nnz, 1)));
  CHK(grow(ctx, ctx->bval, sizeof(V) * std::max<int64_t>(B.nnz, 1)));
  if (perm) CHK(grow(ctx, ctx->bperm, sizeof(int64_t) * std::max<int64_t>(B.nnz, 1)));
  int64_t* bptr = (int64_t*)ctx->bptr.p;
  CHK(fill_i64(ctx, bptr, (int64_t)R * ld, 0));
  if (B.m > 0) {
    const int g = ctx->num_cu * 8;  // persistent waves over the long-row list
    const size_t dyn = sizeof(int32_t) * R * 4;  // per wave of a 256-thread block
    if (dyn > ctx->lds_max)
      return fail(ctx, AMS_ERR_ARG, "%d column chunks of width %d exceed the LDS cursors; raise dense_width", R, W);
    const int gs = (int)std::min<int64_t>((B.m + 255) / 256, (int64_t)ctx->num_cu * 8);
    CHK(grow(ctx, ctx->lrows, sizeof(int32_t) * B.m));
    Ctrl* c = dctrl(ctx);
    CHK(fill_i32(ctx, &c->nlong, 1, 0));
    int32_t* lrows = (int32_t*)ctx->lrows.p;
    hipLaunchKernelGGL(k_bblk_count_short, dim3(gs), dim3(256), 0, ctx->stream, B, pr, W, R, bptr, lrows,
                       &c->nlong);
    hipLaunchKernelGGL(k_bblk_count, dim3(g), dim3(256), dyn, ctx->stream, B, pr, W, R, bptr,
                       (const int32_t*)lrows, (const int32_t*)&c->nlong);
    CHK(scan_i64(ctx, bptr, bptr, (int64_t)R * ld));
    int64_t* bperm = perm ? (int64_t*)ctx->bperm.p : (int64_t*)nullptr;
    hipLaunchKernelGGL(k_bblk_scatter_short<V>, dim3(gs), dim3(256), 0, ctx->stream, B, pr, W,
                       (const int64_t*)bptr, (uint16_t*)ctx->bcol.p, (V*)ctx->bval.p, bperm);
    hipLaunchKernelGGL(k_bblk_scatter<V>, dim3(g), dim3(256), dyn, ctx->stream, B, pr, W, R,
                       (const int64_t*)bptr, (uint16_t*)ctx->bcol.p, (V*)ctx->bval.p, bperm,
                       (const int32_t*)lrows, (const int32_t*)&c->nlong);
    HIPCHK(hipGetLastError());
  } else {
    CHK(scan_i64(ctx, bptr, bptr, (int64_t)R * ld));
  }
  return AMS_OK;
}

// ------------------------------------------------------------ symbolic

template <int S>
static void launch_symb_hash(ams_ctx_t* ctx, const Csr& A, const Csr& B, int64_t* rownz, int bin) {
  constexpr int T = symb_hash_threads<S>();
  const int64_t cnt = ctx->plan_counts ? ctx->cnt_sym[bin] : A.m;
  if (ctx->plan_counts && cnt == 0) return;
  int occ;
  occupancy(k_symb_hash<S, T>, T, 0, &occ);
  const int g = grid_for(ctx, occ, cnt, ctx->plan_counts);
  Ctrl* c = dctrl(ctx);
  hipLaunchKernelGGL((k_symb_hash<S, T>), dim3(g), dim3(T), 0, ctx->stream, A, B,
                     (const int32_t*)ctx->lsym.p, (const int32_t*)c->cnt_sym, bin, rownz, &c->err);
}

// The bitmap bin's wide rows (and heavy-row pieces, when hflag is set), one
// row per T-thread block with dyn bytes of bitmap and chunk counters.
template <int T, bool Q32, int SW, bool FULL = false>
static int launch_symb_bitmap_q(ams_ctx_t* ctx, const Csr& A, const Csr& B, const Probs& pr, int64_t H, size_t dyn,
                                int64_t* rownz, int R, int W, int Wb, const uint8_t* hflag, int64_t gwords,
                                int wave_rows) {
  if (SW > 0) dyn = 0;  // the bitmap is the kernel's static array
  if (dyn + sizeof(ExpandLds<T>) + (size_t)SW * 4 > ctx->lds_max)
    return fail(ctx, AMS_ERR_ARG, "bitmap_bits %d needs %zu B of LDS", Wb, dyn);
  HIPCHK(hipFuncSetAttribute((const void*)k_symb_bitmap<T, Q32, SW, FULL>,
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn));
  Ctrl* c = dctrl(ctx);
  int occ;
  occupancy(k_symb_bitmap<T, Q32, SW, FULL>, T, dyn, &occ);
  const int g = grid_for(ctx, occ, H, ctx->plan_counts);
  // c->sym_ticket starts at zero: the control block is cleared at the start
  // of every symbolic phase, which launches this kernel once
  hipLaunchKernelGGL((k_symb_bitmap<T, Q32, SW, FULL>), dim3(g), dim3(T), dyn, ctx->stream, A, B, pr,
                     (const int64_t*)ctx->row_flp.p,
                     (const int32_t*)ctx->lsym.p, (const int32_t*)c->cnt_sym, (int)SB_BITMAP, rownz,
                     (int32_t*)ctx->coff.p, R, W, Wb, hflag, (const HeavyPiece*)ctx->hpiece.p,
                     (const int32_t*)&c->npieces, (uint32_t*)ctx->gbm.p, gwords, &c->sym_ticket, wave_rows);
  return AMS_OK;
}

template <int T>
static int launch_symb_bitmap(ams_ctx_t* ctx, const Csr& A, const Csr& B, const Probs& pr, int64_t H, size_t dyn,
                              int64_t* rownz, int R, int W, int Wb, const uint8_t* hflag, int64_t gwords,
                              int wave_rows) {
  const bool q32 = A.nnz < ((int64_t)1 << 30) && B.nnz < ((int64_t)1 << 30) && B.m < ((int64_t)1 << 28);
  if constexpr (T == 1024) {
    // every row in one pass of its whole width: no per-product range test
    if ((int64_t)(Wb >> 5) + Wb / W + 1 <= kSymStaticWords && Wb >= pr.max_width)
      return q32 ? launch_symb_bitmap_q<T, true, kSymStaticWords, true>(ctx, A, B, pr, H, dyn, rownz, R, W, Wb, hflag,
                                                                        gwords, wave_rows)
                 : launch_symb_bitmap_q<T, false, kSymStaticWords, true>(ctx, A, B, pr, H, dyn, rownz, R, W, Wb,
                                                                         hflag, gwords, wave_rows);
    if ((int64_t)(Wb >> 5) + Wb / W + 1 <= kSymStaticWords)
      return q32 ? launch_symb_bitmap_q<T, true, kSymStaticWords>(ctx, A, B, pr, H, dyn, rownz, R, W, Wb, hflag, gwords,
                                                                  wave_rows)
                 : launch_symb_bitmap_q<T, false, kSymStaticWords>(ctx, A, B, pr, H, dyn, rownz, R, W, Wb, hflag,
                                                                   gwords, wave_rows);
  }
  return q32 ? launch_symb_bitmap_q<T, true, 0>(ctx, A, B, pr, H, dyn, rownz, R, W, Wb, hflag, gwords, wave_rows)
             : launch_symb_bitmap_q<T, false, 0>(ctx, A, B, pr, H, dyn, rownz, R, W, Wb, hflag, gwords, wave_rows);
}

// Numeric route of every row (k_classify_num) from the symbolic phase's
// bins and C's row pointers, rows grouped by route; with sync, the route
// counts, nnz(C) and the statistics come back to the host.  The control
// block's numeric counters and route sums must be zero.
static int route_numeric(ams_ctx_t* ctx, const Csr& A, const Probs& pr, const int64_t* rp_c, int R, bool sync,
                         int64_t* nnz) {
  const int64_t m = A.m;
  Ctrl* c = dctrl(ctx);
  const NumParams np{ctx->prm.num_hash_max, ctx->prm.esc_max, ctx->prm.acc_thres, ctx->prm.dense_density,
                     ctx->prm.mode, ctx->plan_dtype == AMS_F64 ? 8 : 4};
  if (m > 0) {
    Phase ph(ctx, AMS_PH_CLASSIFY, sync);
    const int gc = (int)std::min<int64_t>((m + 255) / 256, (int64_t)ctx->num_cu * 8);
    hipLaunchKernelGGL(k_classify_num, dim3(gc), dim3(256), 0, ctx->stream, m,
                       (const int64_t*)ctx->row_flp.p, (const int64_t*)rp_c, (const uint8_t*)ctx->sbin.p, pr,
                       np, (uint8_t*)ctx->nbin.p, c->cnt_num, &c->ghash_num_bytes);
    hipLaunchKernelGGL(k_scatter_bins, dim3(gc), dim3(256), 0, ctx->stream, m, (const uint8_t*)ctx->nbin.p,
                       kNumBins, (const int32_t*)c->cnt_num, c->cur_num, (int32_t*)ctx->lnum.p);
    if (sync)  // one block a CU: its sums go to global memory once a block
      hipLaunchKernelGGL(k_route_sums, dim3((unsigned)std::min<int64_t>(gc, ctx->num_cu)), dim3(256), 0, ctx->stream, m,
                         (const int64_t*)ctx->row_flp.p,
                         (const int64_t*)rp_c, A.rp, (const uint8_t*)ctx->sbin.p, (const uint8_t*)ctx->nbin.p,
                         c->sums);
    HIPCHK(hipGetLastError());
  }
  if (sync) {
    HIPCHK(hipMemcpyAsync(ctx->h_ctrl, c, sizeof(Ctrl), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->h_pin, rp_c + m, sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    std::memcpy(ctx->cnt_num, ctx->h_ctrl->cnt_num, sizeof(ctx->cnt_num));
    if (ctx->h_ctrl->err) return fail(ctx, AMS_ERR_STATE, "symbolic kernels raised error flags 0x%x", ctx->h_ctrl->err);
    if (nnz) *nnz = ctx->h_pin[0];
    ams_stats_t& s = ctx->stats;
    s.rows_total = m;
    s.rows_sym_bitmap = ctx->cnt_sym[SB_BITMAP];
    s.rows_sym_hash = 0;
    for (int i = 0; i < kSymHashBins; ++i) s.rows_sym_hash += ctx->cnt_sym[i];
    s.rows_empty = m - s.rows_sym_hash - s.rows_sym_bitmap;
    s.rows_num_hash = s.rows_num_esc = 0;
    for (int i = 0; i < kNumHashBins; ++i) s.rows_num_hash += ctx->cnt_num[i];
    for (int i = 0; i < kNumEscBins; ++i) s.rows_num_esc += ctx->cnt_num[kNumHashBins + i];
    s.rows_num_dense = ctx->cnt_num[NB_DENSE];
    s.rows_sym_ghash = ctx->cnt_sym[SB_GHASH];
    s.rows_num_ghash = ctx->cnt_num[NB_GHASH];
    s.rows_empty -= s.rows_sym_ghash;
    s.dense_chunks = R;
    s.nnz_c = ctx->h_pin[0];
    for (int k = 0; k < 4; ++k) {
      s.flops_route[k] = (int64_t)ctx->h_ctrl->sums[k];
      s.nnzc_route[k] = (int64_t)ctx->h_ctrl->sums[4 + k];
      s.nnza_route[k] = (int64_t)ctx->h_ctrl->sums[8 + k];
    }
    for (int k = 0; k < 3; ++k) {
      s.flops_sym[k] = (int64_t)ctx->h_ctrl->sums[12 + k];
      s.nnza_sym[k] = (int64_t)ctx->h_ctrl->sums[15 + k];
    }
    ctx->ghash_num_bytes = (int64_t)ctx->h_ctrl->ghash_num_bytes;
  }
  return AMS_OK;
}

// The routing parameters of the numeric phase changed since the symbolic
// phase (same geometry): route again without counting again.
static int reroute_numeric(ams_ctx_t* ctx, const ams_dcsr_t* a, const int64_t* rp_c) {
  Ctrl* c = dctrl(ctx);
  CHK(fill_i32(ctx, c->cnt_num, 2 * 32, 0));  // cnt_num and cur_num
  CHK(fill_i64(ctx, (int64_t*)c->sums, 18, 0));
  CHK(fill_i64(ctx, (int64_t*)&c->ghash_num_bytes, 1, 0));
  CHK(route_numeric(ctx, to_csr(a), ctx->plan_pr, rp_c, ctx->plan_R, true, nullptr));
  ctx->plan_prm = ctx->prm;
  return AMS_OK;
}

static int run_symbolic(ams_ctx_t* ctx, const ams_dcsr_t* a, const ams_dcsr_t* b, int dtype,
                        const Probs& pr_in, int64_t* rp_c, bool sync, int64_t* nnz) {
  const Csr A = to_csr(a), B = to_csr(b);
  Probs pr = pr_in;
  pr.rowbits = nullptr;
  pr.bwords = 0;
  const int64_t m = A.m;
  ctx->plan_valid = false;
  CHK(grow(ctx, ctx->ctrl, sizeof(Ctrl)));
  CHK(grow(ctx, ctx->row_flp, sizeof(int64_t) * std::max<int64_t>(m, 1)));
  CHK(grow(ctx, ctx->sbin, std::max<int64_t>(m, 1)));
  CHK(grow(ctx, ctx->nbin, std::max<int64_t>(m, 1)));
  CHK(grow(ctx, ctx->lsym, sizeof(int32_t) * std::max<int64_t>(m, 1)));
  CHK(grow(ctx, ctx->lnum, sizeof(int32_t) * std::max<int64_t>(m, 1)));
  Ctrl* c = dctrl(ctx);
  CHK(fill_i32(ctx, (int32_t*)c, sizeof(Ctrl) / 4, 0));

  int W, R, Wb;
  dense_geometry(ctx, dtype, pr.max_width, W, R, Wb);
  ctx->plan_W = W;
  ctx->plan_R = R;
  ctx->plan_Wb = Wb;
  ctx->plan_m = m;
  ctx->plan_mB = B.m;
  ctx->plan_nnzB = B.nnz;
  ctx->plan_dtype = dtype;
  ctx->plan_pr = pr;
  ctx->plan_nnzA = A.nnz;
  ctx->plan_nB = B.n;
  ctx->plan_arp = A.rp;
  ctx->plan_acol = A.col;
  ctx->plan_brp = B.rp;
  ctx->plan_bcol = B.col;

  // the global-hash route needs the sizes of its tables on the host: only
  // in synchronised calls
  SymParams sp{ctx->prm.sym_hash_max, ctx->prm.bitmap_ratio, Wb, sync ? ctx->prm.ghash_ratio : 0.0};
  if (m > 0) {
    {
      Phase ph(ctx, AMS_PH_FLOPS, sync);
      CHK(launch_row_flops(ctx, A, B, (int64_t*)ctx->row_flp.p, &c->flops));
    }
    Phase ph(ctx, AMS_PH_CLASSIFY, sync);
    const int gc = (int)std::min<int64_t>((m + 255) / 256, (int64_t)ctx->num_cu * 8);
    hipLaunchKernelGGL(k_classify_sym, dim3(gc), dim3(256), 0, ctx->stream, m,
                       (const int64_t*)ctx->row_flp.p, pr, sp, (uint8_t*)ctx->sbin.p, rp_c, c->cnt_sym,
                       &c->ghash_sym_bytes);
    hipLaunchKernelGGL(k_scatter_bins, dim3(gc), dim3(256), 0, ctx->stream, m, (const uint8_t*)ctx->sbin.p,
                       kSymBins, (const int32_t*)c->cnt_sym, c->cur_sym, (int32_t*)ctx->lsym.p);
    HIPCHK(hipGetLastError());
  }
  ctx->plan_counts = false;
  if (sync) {
    HIPCHK(hipMemcpyAsync(ctx->h_ctrl, c, sizeof(Ctrl), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    std::memcpy(ctx->cnt_sym, ctx->h_ctrl->cnt_sym, sizeof(ctx->cnt_sym));
    ctx->plan_counts = true;
    ctx->stats.flops = (int64_t)ctx->h_ctrl->flops;
  }
  const int64_t H = ctx->plan_counts ? ctx->cnt_sym[SB_BITMAP] : m;
  ctx->plan_H = H;
  CHK(grow(ctx, ctx->coff, sizeof(int32_t) * std::max<int64_t>(H, 1) * (R + 1)));
  // The bitmap bin's column bitmaps are kept for the numeric phase (its
  // dense chunks then skip the per-product bitmap test and OR) in
  // synchronised calls of the row-bitmap symbolic, when they take at most a
  // twelfth of the GPU's memory (G500 s19 ef32: 20 GB next to C's 246 GB).
  if (ctx->prm.row_bitmaps && ctx->plan_counts && !ctx->fixed && !ctx->prm.sym_blocked && H > 0 &&
      ctx->mem_total > 0) {
    const int64_t bwords = (std::max<int64_t>(pr.max_width, 1) + 31) / 32;
    const size_t bytes = (size_t)H * (size_t)bwords * 4;
    if (bytes <= ctx->mem_total / 12) {
      CHK(grow(ctx, ctx->rowbits, bytes));
      pr.rowbits = (uint32_t*)ctx->rowbits.p;
      pr.bwords = bwords;
    }
  }
  ctx->plan_pr = pr;

  if (m > 0) {
    int64_t* rownz = rp_c;
    Phase ph(ctx, AMS_PH_SYMB_HASH, sync);
    launch_symb_hash<64>(ctx, A, B, rownz, 0);
    launch_symb_hash<128>(ctx, A, B, rownz, 1);
    launch_symb_hash<256>(ctx, A, B, rownz, 2);
    launch_symb_hash<512>(ctx, A, B, rownz, 3);
    launch_symb_hash<1024>(ctx, A, B, rownz, 4);
    launch_symb_hash<2048>(ctx, A, B, rownz, 5);
    launch_symb_hash<4096>(ctx, A, B, rownz, 6);
    launch_symb_hash<8192>(ctx, A, B, rownz, 7);
    HIPCHK(hipGetLastError());
  }
  if (m > 0 && ctx->plan_counts && ctx->cnt_sym[SB_GHASH] > 0) {
    Phase ph(ctx, AMS_PH_SYMB_GHASH, sync);
    CHK(grow(ctx, ctx->gtab, (size_t)ctx->h_ctrl->ghash_sym_bytes));
    constexpr int TG = 256;
    int occ;
    occupancy(k_symb_ghash<TG>, TG, 0, &occ);
    const int g = grid_for(ctx, occ, ctx->cnt_sym[SB_GHASH], true);
    hipLaunchKernelGGL(k_symb_ghash<TG>, dim3(g), dim3(TG), 0, ctx->stream, A, B, (const int64_t*)ctx->row_flp.p,
                       (const int32_t*)ctx->lsym.p, (const int32_t*)c->cnt_sym, (int)SB_GHASH, rp_c,
                       (int32_t*)ctx->gtab.p, &c->ghash_cursor, &c->err);
    HIPCHK(hipGetLastError());
  }
  ctx->bblk_ready = false;
  if (m > 0 && ctx->prm.sym_blocked && (!ctx->plan_counts || H > 0)) {
    // wide rows counted per (row, chunk) morsel on the column-blocked B
    {
      Phase ph(ctx, AMS_PH_BBLK, sync);
      if (dtype == AMS_F64) CHK(build_bblk<double>(ctx, B, pr, W, R, false));
      else CHK(build_bblk<float>(ctx, B, pr, W, R, false));
      ctx->bblk_ready = true;
    }
    Phase ph(ctx, AMS_PH_SYMB_BITMAP, sync);
    CHK(grow(ctx, ctx->ccnt, sizeof(int32_t) * std::max<int64_t>(H, 1) * R));
    CHK(fill_i64(ctx, (int64_t*)&c->ticket, 1, 0));
    constexpr int TB = 512;
    const size_t per_wave = ((size_t)(W / 32) * 4 + 64 * 4 + 64 * 8 + 15) & ~(size_t)15;
    const size_t dyn = per_wave * (TB / 64);
    HIPCHK(hipFuncSetAttribute((const void*)k_symb_blocked<TB>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)dyn));
    int occ;
    occupancy(k_symb_blocked<TB>, TB, dyn, &occ);
    const int g = grid_for(ctx, occ, H * R, ctx->plan_counts);
    hipLaunchKernelGGL(k_symb_blocked<TB>, dim3(g), dim3(TB), dyn, ctx->stream, A, B.m, (const int64_t*)ctx->bptr.p,
                       (const uint16_t*)ctx->bcol.p, (const int32_t*)ctx->lsym.p, (const int32_t*)c->cnt_sym,
                       (int)SB_BITMAP, R, W, (int32_t*)ctx->ccnt.p, &c->ticket);
    const int gr = (int)std::max<int64_t>(1, std::min<int64_t>((H + 255) / 256, (int64_t)ctx->num_cu * 8));
    hipLaunchKernelGGL(k_coff_rows, dim3(gr), dim3(256), 0, ctx->stream, (const int32_t*)ctx->lsym.p,
                       (const int32_t*)c->cnt_sym, (int)SB_BITMAP, R, (const int32_t*)ctx->ccnt.p,
                       (int32_t*)ctx->coff.p, rp_c);
    HIPCHK(hipGetLastError());
  } else if (m > 0) {
    int64_t* rownz = rp_c;
    Phase ph(ctx, AMS_PH_SYMB_BITMAP, sync);
    if (!ctx->plan_counts || H > 0) {
      const size_t dyn = (size_t)(Wb / 32) * 4 + (size_t)(Wb / W + 1) * 4;
      // heavy rows (more than max(2^20, P/2048) products) are split into
      // pieces when their global bitmaps fit 512 MiB
      // (not in the fixed scratch, which has no room for the global bitmaps:
      // heavy rows are then counted whole)
      const int64_t gwords = ctx->fixed ? 0 : heavy_words(pr);
      const bool split = gwords > 0;
      const uint8_t* hflag = nullptr;
      if (split) {
        CHK(grow_heavy(ctx, pr, H));
        const int gp = (int)std::max<int64_t>(1, std::min<int64_t>((H + 255) / 256, (int64_t)ctx->num_cu * 8));
        hipLaunchKernelGGL(k_heavy_plan, dim3(gp), dim3(256), 0, ctx->stream, pr, (const int64_t*)ctx->row_flp.p,
                           (const int32_t*)ctx->lsym.p, (const int32_t*)c->cnt_sym, (int)SB_BITMAP,
                           (const unsigned long long*)&c->flops, (uint8_t*)ctx->hflag.p, (int32_t*)ctx->hrow.p,
                           (HeavyPiece*)ctx->hpiece.p, &c->nheavy, &c->npieces);
        HIPCHK(hipGetLastError());
        hflag = (const uint8_t*)ctx->hflag.p;
      }
      // light rows of problems of at most kWaveBitmapBits columns go to the
      // wave kernel below; in a batch any problem may be that narrow
      const int wave_rows = (pr.n > 1 || pr.max_width <= kWaveBitmapBits) ? 1 : 0;
      if (pr.n > 1 && Wb <= kSymNarrowBits)
        CHK(launch_symb_bitmap<kSymBitmapNarrowT>(ctx, A, B, pr, H, dyn, rownz, R, W, Wb, hflag, gwords, wave_rows));
      else
        CHK(launch_symb_bitmap<kSymBitmapT>(ctx, A, B, pr, H, dyn, rownz, R, W, Wb, hflag, gwords, wave_rows));
      if (split) {
        // k_heavy_finish clears what the pieces set; should either launch
        // fail, the global bitmaps are refilled with zeros on the next call
        ctx->gbm_zero = 0;
        HIPCHK(hipGetLastError());
        constexpr int TF = 256;
        const int gf = std::min(kMaxHeavy, ctx->num_cu * 2);
        hipLaunchKernelGGL(k_heavy_finish<TF>, dim3(gf), dim3(TF), 0, ctx->stream, pr, (const int32_t*)ctx->lsym.p,
                           (const int32_t*)c->cnt_sym, (int)SB_BITMAP, (const int32_t*)ctx->hrow.p,
                           (const int32_t*)&c->nheavy, (uint32_t*)ctx->gbm.p, gwords, rownz,
                           (int32_t*)ctx->coff.p, R, W);
        HIPCHK(hipGetLastError());
        ctx->gbm_zero = ctx->gbm.cap;
      }
      HIPCHK(hipGetLastError());
      if (wave_rows) {  // light rows of narrow problems: a wave each
        constexpr int TW = 256;
        int occw;
        occupancy(k_symb_bitmap_wave<TW>, TW, 0, &occw);
        const int gw = grid_for(ctx, occw, (H + TW / 64 - 1) / (TW / 64), ctx->plan_counts);
        hipLaunchKernelGGL(k_symb_bitmap_wave<TW>, dim3(gw), dim3(TW), 0, ctx->stream, A, B, pr,
                           (const int64_t*)ctx->row_flp.p, (const int32_t*)ctx->lsym.p,
                           (const int32_t*)c->cnt_sym, (int)SB_BITMAP, rownz, (int32_t*)ctx->coff.p, R, W);
        HIPCHK(hipGetLastError());
      }
    }
  }
  {
    Phase ph(ctx, AMS_PH_SCAN, sync);
    CHK(fill_i64(ctx, rp_c + m, 1, 0));
    CHK(scan_i64(ctx, rp_c, rp_c, m + 1));
  }

  CHK(route_numeric(ctx, A, pr, rp_c, R, sync, nnz));
  ctx->plan_prm = ctx->prm;
  ctx->plan_valid = true;
  return AMS_OK;
}

// ------------------------------------------------------------- numeric

template <int S, typename V>
static void launch_numc_hash(ams_ctx_t* ctx, const Csr& A, const Csr& B, const int64_t* rp_c, int32_t* col_c,
                             V* val_c, int64_t cap, int bin) {
  constexpr int T = hash_threads<S>();
  const int64_t cnt = ctx->plan_counts ? ctx->cnt_num[bin] : A.m;
  if (ctx->plan_counts && cnt == 0) return;
  int occ;
  occupancy(k_numc_hash<S, T, V>, T, 0, &occ);
  const int g = grid_for(ctx, occ, cnt, ctx->plan_counts);
  Ctrl* c = dctrl(ctx);
  hipLaunchKernelGGL((k_numc_hash<S, T, V>), dim3(g), dim3(T), 0, ctx->stream, A, B,
                     (const int32_t*)ctx->lnum.p, (const int32_t*)c->cnt_num, bin, rp_c, col_c, val_c, cap,
                     &c->err);
}

template <int E, typename V, bool K32>
static void launch_numc_esc_k(ams_ctx_t* ctx, const Csr& A, const Csr& B, const int64_t* rp_c, int32_t* col_c,
                              V* val_c, int64_t cap, int bin, int64_t cnt) {
  constexpr int T = esc_threads<E>();
  if constexpr (K32) {
    {
      // register bitonic sort (esc_sort 0), LSD radix sort by the column bits
      // (2) or merge of the B segments (3)
      const int col_bits = B.n > 1 ? 64 - __builtin_clzll((unsigned long long)(B.n - 1)) : 1;
      auto go = [&](auto sort_c) {
        constexpr int SORT = decltype(sort_c)::value;
        Ctrl* c = dctrl(ctx);
        int occ;
        occupancy(k_numc_esc32<E, T, V, SORT>, T, 0, &occ);
        const int g = grid_for(ctx, occ, cnt, ctx->plan_counts);
        hipLaunchKernelGGL((k_numc_esc32<E, T, V, SORT>), dim3(g), dim3(T), 0, ctx->stream, A, B,
                           (const int32_t*)ctx->lnum.p, (const int32_t*)c->cnt_num, bin,
                           (const int64_t*)ctx->row_flp.p, rp_c, col_c, val_c, cap, col_bits, &c->err);
      };
      if (ctx->prm.esc_sort == 2) go(std::integral_constant<int, 2>{});
      else if (ctx->prm.esc_sort == 3) go(std::integral_constant<int, 3>{});
      else if (ctx->prm.esc_sort == 4) go(std::integral_constant<int, 4>{});
      else go(std::integral_constant<int, 0>{});
    }
  } else {
    int occ;
    occupancy(k_numc_esc<E, T, V>, T, 0, &occ);
    const int g = grid_for(ctx, occ, cnt, ctx->plan_counts);
    Ctrl* c = dctrl(ctx);
    hipLaunchKernelGGL((k_numc_esc<E, T, V>), dim3(g), dim3(T), 0, ctx->stream, A, B,
                       (const int32_t*)ctx->lnum.p, (const int32_t*)c->cnt_num, bin,
                       (const int64_t*)ctx->row_flp.p, rp_c, col_c, val_c, cap, ctx->prm.esc_sort, &c->err);
  }
}

// ESC bin of rows of at most E products.  E may be 3/4 of a power of two:
// the 32-bit-key kernel then runs T = E/8 threads and sorts with the next
// power of two's network, whose last quarter is virtual; the 64-bit-key
// kernel takes such a bin with the next power of two's geometry.
template <int E, typename V>
static void launch_numc_esc(ams_ctx_t* ctx, const Csr& A, const Csr& B, const int64_t* rp_c, int32_t* col_c,
                            V* val_c, int64_t cap, int bin) {
  const int64_t cnt = ctx->plan_counts ? ctx->cnt_num[bin] : A.m;
  if (ctx->plan_counts && cnt == 0) return;
  constexpr int EP = next_pow2_c(E);
  // (column, rank) fits 32 bits as column << log2(EP) | rank below 2^(32 - log2(EP)) columns
  const bool key32 = (uint64_t)B.n <= (1ull << (32 - __builtin_ctz(EP))) && ctx->prm.esc_sort != 1;
  if (key32) launch_numc_esc_k<E, V, true>(ctx, A, B, rp_c, col_c, val_c, cap, bin, cnt);
  else launch_numc_esc_k<EP, V, false>(ctx, A, B, rp_c, col_c, val_c, cap, bin, cnt);
}

template <int T, typename V, bool PB>
static int launch_numc_dense_pb(ams_ctx_t* ctx, const Csr& A, const Csr& B, int W, int64_t nall,
                                const int64_t* rp_c, int32_t* col_c, V* val_c, int64_t cap, const Probs& pr) {
  Ctrl* c = dctrl(ctx);
  const size_t dyn = dense_lds<T>(W, sizeof(V));
  if (dyn + dense_static_lds<T>() > ctx->lds_max)
    return fail(ctx, AMS_ERR_ARG, "dense_width %d needs %zu B of LDS", W, dyn + dense_static_lds<T>());
  HIPCHK(hipFuncSetAttribute((const void*)k_numc_dense<T, V, PB>, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)dyn));
  int occ;
  occupancy(k_numc_dense<T, V, PB>, T, dyn, &occ);
  const int g = grid_for(ctx, occ, nall, ctx->plan_counts);
  hipLaunchKernelGGL((k_numc_dense<T, V, PB>), dim3(g), dim3(T), dyn, ctx->stream, A, B.m,
                     (const int64_t*)ctx->bptr.p, (const uint16_t*)ctx->bcol.p, (const V*)ctx->bval.p,
                     (const DenseTask*)ctx->dtask.p, (const int64_t*)&c->ntask[0], W, rp_c, col_c, val_c, cap, pr,
                     &c->ticket, &c->err);
  HIPCHK(hipGetLastError());
  return AMS_OK;
}

// The kept row bitmaps serve when one word a thread covers a chunk.
template <int T, typename V>
static int launch_numc_dense(ams_ctx_t* ctx, const Csr& A, const Csr& B, int W, int64_t nall, const int64_t* rp_c,
                             int32_t* col_c, V* val_c, int64_t cap, const Probs& pr) {
  if (pr.rowbits && (W + 31) / 32 <= T)
    return launch_numc_dense_pb<T, V, true>(ctx, A, B, W, nall, rp_c, col_c, val_c, cap, pr);
  return launch_numc_dense_pb<T, V, false>(ctx, A, B, W, nall, rp_c, col_c, val_c, cap, pr);
}

// Light morsels go to the wave kernel when light_max > 0 and a wave's chunk
// bitmap fits its share of the LDS.
static bool light_on(const ams_ctx_t* ctx, int W, int vbytes) {
  const bool pb = ctx->plan_pr.rowbits != nullptr;
  return ctx->prm.light_max > 0 && (W + 31) / 32 <= kLightMaxWords &&
         light_block_lds(W, vbytes, pb) <= ctx->lds_max;
}

// C entries a light morsel may have: up to kLightS at ranks from the kept
// row bitmaps, up to kLightMaxCnt (a table two-thirds full) in hash tables.
static int light_limit(const ams_ctx_t* ctx, int W, int vbytes) {
  if (!light_on(ctx, W, vbytes)) return 0;
  return std::min(ctx->prm.light_max, ctx->plan_pr.rowbits ? kLightCap : kLightMaxCnt);
}

template <typename V, bool PB>
static int launch_numc_light_pb(ams_ctx_t* ctx, const Csr& A, const Csr& B, int W, const int64_t* rp_c,
                                int32_t* col_c, V* val_c, int64_t cap) {
  Ctrl* c = dctrl(ctx);
  const size_t dyn = light_block_lds(W, sizeof(V), PB);
  HIPCHK(hipFuncSetAttribute((const void*)k_numc_light<V, PB>, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)dyn));
  int occ;
  occupancy(k_numc_light<V, PB>, kLightT, dyn, &occ);
  // persistent waves: each walks the light morsels with the grid's stride
  const int g = std::max(1, occ) * ctx->num_cu;
  hipLaunchKernelGGL((k_numc_light<V, PB>), dim3(g), dim3(kLightT), dyn, ctx->stream, A, B.m,
                     (const int64_t*)ctx->bptr.p, (const uint16_t*)ctx->bcol.p, (const V*)ctx->bval.p,
                     (const DenseTask*)ctx->dtask.p, (const int64_t*)c->ntask, W, rp_c, col_c, val_c, cap,
                     ctx->plan_pr, &c->err);
  HIPCHK(hipGetLastError());
  return AMS_OK;
}

// With the symbolic phase's row bitmaps the light morsels add at their
// columns' ranks; without them, in per-wave hash tables.
template <typename V>
static int launch_numc_light(ams_ctx_t* ctx, const Csr& A, const Csr& B, int W, const int64_t* rp_c,
                             int32_t* col_c, V* val_c, int64_t cap) {
  if (ctx->plan_pr.rowbits)
    return launch_numc_light_pb<V, true>(ctx, A, B, W, rp_c, col_c, val_c, cap);
  return launch_numc_light_pb<V, false>(ctx, A, B, W, rp_c, col_c, val_c, cap);
}

template <typename V>
static int run_numeric_t(ams_ctx_t* ctx, const ams_dcsr_t* a, const ams_dcsr_t* b, const int64_t* rp_c,
                         int32_t* col_c, V* val_c, int64_t cap) {
  const Csr A = to_csr(a), B = to_csr(b);
  const Probs& pr = ctx->plan_pr;
  Ctrl* c = dctrl(ctx);
  const int W = ctx->plan_W, R = ctx->plan_R;
  const bool dense = !ctx->plan_counts || ctx->cnt_num[NB_DENSE] > 0;
  const bool on = ctx->plan_counts;
  if (dense && A.m > 0) {
    if (!ctx->bblk_ready) {
      Phase ph(ctx, AMS_PH_BBLK, on);
      CHK(build_bblk<V>(ctx, B, pr, W, R, false));
    }
    const int64_t hmax = ctx->plan_counts ? (int64_t)ctx->cnt_sym[SB_BITMAP] : A.m;
    const int64_t nall = hmax * R;
    {
      // compacted morsel descriptors, chunk-major
      Phase ph(ctx, AMS_PH_BBLK, on);
      CHK(grow(ctx, ctx->dflag, sizeof(int64_t) * (nall + 1)));
      CHK(grow(ctx, ctx->dtask, sizeof(DenseTask) * std::max<int64_t>(nall, 1)));
      int64_t* flag = (int64_t*)ctx->dflag.p;
      CHK(fill_i64(ctx, flag, nall + 1, 0));
      const int gf = (int)std::max<int64_t>(1, std::min<int64_t>((nall + 255) / 256, (int64_t)ctx->num_cu * 16));
      hipLaunchKernelGGL(k_dense_flags, dim3(gf), dim3(256), 0, ctx->stream, (const int32_t*)ctx->lsym.p,
                         (const int32_t*)c->cnt_sym, (int)SB_BITMAP, (const uint8_t*)ctx->nbin.p,
                         (const int32_t*)ctx->coff.p, R, A.rp, light_limit(ctx, W, sizeof(V)), flag);
      CHK(scan_i64(ctx, flag, flag, nall + 1));
      hipLaunchKernelGGL(k_dense_fill, dim3(gf), dim3(256), 0, ctx->stream, A, pr, W, (const int32_t*)ctx->lsym.p,
                         (const int32_t*)c->cnt_sym, (int)SB_BITMAP, (const int32_t*)ctx->coff.p, R, rp_c,
                         (const int64_t*)flag, (DenseTask*)ctx->dtask.p, c->ntask);
      HIPCHK(hipGetLastError());
    }
    {
      Phase ph(ctx, AMS_PH_NUMC_DENSE, on);
      CHK(fill_i64(ctx, (int64_t*)&c->ticket, 1, 0));
      const int rc = dense_block(W, sizeof(V)) == 512
                         ? launch_numc_dense<512, V>(ctx, A, B, W, nall, rp_c, col_c, val_c, cap, pr)
                         : launch_numc_dense<1024, V>(ctx, A, B, W, nall, rp_c, col_c, val_c, cap, pr);
      CHK(rc);
    }
    if (light_on(ctx, W, sizeof(V))) {
      Phase phl(ctx, AMS_PH_NUMC_LIGHT, on);
      CHK(launch_numc_light<V>(ctx, A, B, W, rp_c, col_c, val_c, cap));
    }
  }
  if (A.m > 0) {
    Phase ph(ctx, AMS_PH_NUMC_HASH, on);
    launch_numc_hash<64, V>(ctx, A, B, rp_c, col_c, val_c, cap, 0);
    launch_numc_hash<128, V>(ctx, A, B, rp_c, col_c, val_c, cap, 1);
    launch_numc_hash<256, V>(ctx, A, B, rp_c, col_c, val_c, cap, 2);
    launch_numc_hash<512, V>(ctx, A, B, rp_c, col_c, val_c, cap, 3);
    launch_numc_hash<1024, V>(ctx, A, B, rp_c, col_c, val_c, cap, 4);
    launch_numc_hash<2048, V>(ctx, A, B, rp_c, col_c, val_c, cap, 5);
    launch_numc_hash<4096, V>(ctx, A, B, rp_c, col_c, val_c, cap, 6);
    launch_numc_hash<8192, V>(ctx, A, B, rp_c, col_c, val_c, cap, 7);
  }
  if (A.m > 0 && ctx->plan_counts && ctx->cnt_num[NB_GHASH] > 0) {
    Phase ph(ctx, AMS_PH_NUMC_GHASH, on);
    CHK(grow(ctx, ctx->gnum, (size_t)ctx->ghash_num_bytes));
    CHK(fill_i64(ctx, (int64_t*)&c->ghash_cursor, 1, 0));
    constexpr int TG = 256;
    int occ;
    occupancy(k_numc_ghash<TG, V>, TG, 0, &occ);
    const int g = grid_for(ctx, occ, ctx->cnt_num[NB_GHASH], true);
    const int col_bits = B.n > 1 ? 64 - __builtin_clzll((unsigned long long)(B.n - 1)) : 1;
    hipLaunchKernelGGL((k_numc_ghash<TG, V>), dim3(g), dim3(TG), 0, ctx->stream, A, B, (const int32_t*)ctx->lnum.p,
                       (const int32_t*)c->cnt_num, (int)NB_GHASH, rp_c, col_c, val_c, cap,
                       (unsigned char*)ctx->gnum.p, &c->ghash_cursor, col_bits, &c->err);
    HIPCHK(hipGetLastError());
  }
  if (A.m > 0) {
    Phase ph(ctx, AMS_PH_NUMC_ESC, on);
    launch_numc_esc<64, V>(ctx, A, B, rp_c, col_c, val_c, cap, kNumHashBins + 0);
    launch_numc_esc<128, V>(ctx, A, B, rp_c, col_c, val_c, cap, kNumHashBins + 1);
    launch_numc_esc<256, V>(ctx, A, B, rp_c, col_c, val_c, cap, kNumHashBins + 2);
    launch_numc_esc<512, V>(ctx, A, B, rp_c, col_c, val_c, cap, kNumHashBins + 3);
    launch_numc_esc<1024, V>(ctx, A, B, rp_c, col_c, val_c, cap, kNumHashBins + 4);
    launch_numc_esc<1536, V>(ctx, A, B, rp_c, col_c, val_c, cap, kNumHashBins + 5);
    launch_numc_esc<2048, V>(ctx, A, B, rp_c, col_c, val_c, cap, kNumHashBins + 6);
    launch_numc_esc<3072, V>(ctx, A, B, rp_c, col_c, val_c, cap, kNumHashBins + 7);
    launch_numc_esc<4096, V>(ctx, A, B, rp_c, col_c, val_c, cap, kNumHashBins + 8);
    HIPCHK(hipGetLastError());
  }
  return AMS_OK;
}

static int run_numeric(ams_ctx_t* ctx, const ams_dcsr_t* a, const ams_dcsr_t* b, int dtype, const int64_t* rp_c,
                       int32_t* col_c, void* val_c, int64_t cap) {
  if (dtype == AMS_F64) return run_numeric_t<double>(ctx, a, b, rp_c, col_c, (double*)val_c, cap);
  return run_numeric_t<float>(ctx, a, b, rp_c, col_c, (float*)val_c, cap);
}

static int check_pair(ams_ctx_t* ctx, const ams_dcsr_t* a, const ams_dcsr_t* b, int dtype) {
  if (!ctx) return AMS_ERR_ARG;
  CHK(check_csr(ctx, a, "A"));
  CHK(check_csr(ctx, b, "B"));
  if (a->num_col != b->num_row)
    return fail(ctx, AMS_ERR_ARG, "A is %lld x %lld but B has %lld rows", (long long)a->num_row,
                (long long)a->num_col, (long long)b->num_row);
  if (dtype != AMS_F64 && dtype != AMS_F32) return fail(ctx, AMS_ERR_ARG, "unknown dtype %d", dtype);
  return AMS_OK;
}

static Probs single_probs(const ams_dcsr_t* b) {
  Probs pr{};
  pr.n = 1;
  pr.width0 = b->num_col;
  pr.max_width = b->num_col;
  return pr;
}

// ------------------------------------------------------------- backward

template <int S, typename V>
static void launch_bwd_hash(ams_ctx_t* ctx, const Csr& A, const Csr& B, const Csr& G, V* dA, V* dB, int bin) {
  constexpr int T = hash_threads<S>();
  const int64_t cnt = ctx->cnt_num[bin];
  if (cnt == 0) return;
  int occ;
  occupancy(k_bwd_hash<S, T, V>, T, 0, &occ);
  const int g = grid_for(ctx, occ, cnt, true);
  Ctrl* c = dctrl(ctx);
  hipLaunchKernelGGL((k_bwd_hash<S, T, V>), dim3(g), dim3(T), 0, ctx->stream, A, B, G, (const int32_t*)ctx->lnum.p,
                     (const int32_t*)c->cnt_num, bin, dA, dB, &c->err);
}

template <typename V>
static int masked_bwd_t(ams_ctx_t* ctx, const ams_dcsr_t* a, const ams_dcsr_t* b, const ams_dcsr_t* g, V* dA,
                        V* dB) {
  const Csr A = to_csr(a), B = to_csr(b), G = to_csr(g);
  const int64_t m = A.m;
  ctx->plan_valid = false;  // the routing buffers are reused below
  CHK(grow(ctx, ctx->ctrl, sizeof(Ctrl)));
  CHK(grow(ctx, ctx->nbin, std::max<int64_t>(m, 1)));
  CHK(grow(ctx, ctx->lnum, sizeof(int32_t) * std::max<int64_t>(m, 1)));
  Ctrl* c = dctrl(ctx);
  CHK(fill_i32(ctx, (int32_t*)c, sizeof(Ctrl) / 4, 0));
  if (sizeof(V) == 8) {
    CHK(fill_i64(ctx, (int64_t*)dA, A.nnz, 0));
    CHK(fill_i64(ctx, (int64_t*)dB, B.nnz, 0));
  } else {
    CHK(fill_i32(ctx, (int32_t*)dA, A.nnz, 0));
    CHK(fill_i32(ctx, (int32_t*)dB, B.nnz, 0));
  }
  if (m == 0) return AMS_OK;
  const int gc = (int)std::min<int64_t>((m + 255) / 256, (int64_t)ctx->num_cu * 8);
  hipLaunchKernelGGL(k_classify_bwd, dim3(gc), dim3(256), 0, ctx->stream, m, G.rp, (int64_t)ctx->prm.num_hash_max,
                     (uint8_t*)ctx->nbin.p, c->cnt_num);
  hipLaunchKernelGGL(k_scatter_bins, dim3(gc), dim3(256), 0, ctx->stream, m, (const uint8_t*)ctx->nbin.p, kNumBins,
                     (const int32_t*)c->cnt_num, c->cur_num, (int32_t*)ctx->lnum.p);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(ctx->h_ctrl, c, sizeof(Ctrl), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  std::memcpy(ctx->cnt_num, ctx->h_ctrl->cnt_num, sizeof(ctx->cnt_num));
  launch_bwd_hash<64, V>(ctx, A, B, G, dA, dB, 0);
  launch_bwd_hash<128, V>(ctx, A, B, G, dA, dB, 1);
  launch_bwd_hash<256, V>(ctx, A, B, G, dA, dB, 2);
  launch_bwd_hash<512, V>(ctx, A, B, G, dA, dB, 3);
  launch_bwd_hash<1024, V>(ctx, A, B, G, dA, dB, 4);
  launch_bwd_hash<2048, V>(ctx, A, B, G, dA, dB, 5);
  launch_bwd_hash<4096, V>(ctx, A, B, G, dA, dB, 6);
  launch_bwd_hash<8192, V>(ctx, A, B, G, dA, dB, 7);
  HIPCHK(hipGetLastError());
  if (ctx->cnt_num[NB_DENSE] > 0) {
    const Probs pr = single_probs(b);
    int W, R, Wb;
    dense_geometry(ctx, sizeof(V) == 8 ? AMS_F64 : AMS_F32, pr.max_width, W, R, Wb);
    // The backward builds its own blocked copy of B, so its chunks may be
    // narrower than the forward's: at least two blocks a CU (the FP64
    // forward chunk, 67.6 KB of G values next to this kernel's 14 KB of
    // static LDS, left room for one: 164 ms against 73 ms in FP32 on G500
    // s18 ef16).
    for (int occ = 0;;) {
      occupancy(k_bwd_dense<kDenseT, V>, kDenseT, (size_t)W * sizeof(V), &occ);
      if (occ >= AMS_BWD_MIN_BLOCKS || W <= 1024) break;
      W -= 256;
    }
    R = (int)((pr.max_width + W - 1) / W);
    if (R < 1) R = 1;
    CHK(build_bblk<V>(ctx, B, pr, W, R, true));
    ctx->bblk_ready = false;
    const size_t dyn = (size_t)W * sizeof(V);
    HIPCHK(hipFuncSetAttribute((const void*)k_bwd_dense<kDenseT, V>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)dyn));
    CHK(fill_i64(ctx, (int64_t*)&c->ticket, 1, 0));
    CHK(grow(ctx, ctx->bgrad, sizeof(V) * std::max<int64_t>(B.nnz, 1)));
    V* dBb = (V*)ctx->bgrad.p;
    if (sizeof(V) == 8) CHK(fill_i64(ctx, (int64_t*)dBb, B.nnz, 0));
    else CHK(fill_i32(ctx, (int32_t*)dBb, B.nnz, 0));
    // G's chunk boundaries of every dense row (the forward's dense descriptors are not kept)
    const int64_t nsplit = (int64_t)ctx->cnt_num[NB_DENSE] * (R + 1);
    CHK(grow(ctx, ctx->dflag, sizeof(int64_t) * std::max<int64_t>(nsplit, 1)));
    const int gs = (int)std::max<int64_t>(1, std::min<int64_t>((nsplit + 255) / 256, (int64_t)ctx->num_cu * 16));
    hipLaunchKernelGGL(k_bwd_gsplit, dim3(gs), dim3(256), 0, ctx->stream, G, (const int32_t*)ctx->lnum.p,
                       (const int32_t*)c->cnt_num, (int)NB_DENSE, R, W, pr, (int64_t*)ctx->dflag.p);
    int occ;
    occupancy(k_bwd_dense<kDenseT, V>, kDenseT, dyn, &occ);
    const int g = grid_for(ctx, occ, (int64_t)ctx->cnt_num[NB_DENSE] * R, true);
    hipLaunchKernelGGL((k_bwd_dense<kDenseT, V>), dim3(g), dim3(kDenseT), dyn, ctx->stream, A, G, B.m,
                       (const int64_t*)ctx->bptr.p, (const uint16_t*)ctx->bcol.p, (const V*)ctx->bval.p,
                       (const int32_t*)ctx->lnum.p, (const int32_t*)c->cnt_num,
                       (int)NB_DENSE, R, W, pr, (const int64_t*)ctx->dflag.p, &c->ticket, dA, dBb);
    const int gu = (int)std::max<int64_t>(1, std::min<int64_t>((B.nnz + 255) / 256, (int64_t)ctx->num_cu * 16));
    hipLaunchKernelGGL((k_bwd_unpermute<V>), dim3(gu), dim3(256), 0, ctx->stream, (const int64_t*)ctx->bperm.p,
                       (const V*)dBb, dB, B.nnz);
    HIPCHK(hipGetLastError());
  }
  HIPCHK(hipMemcpyAsync(&ctx->h_ctrl->err, &c->err, 4, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  if (ctx->h_ctrl->err) return fail(ctx, AMS_ERR_STATE, "backward kernels raised error flags 0x%x", ctx->h_ctrl->err);
  return AMS_OK;
}

// ----------------------------------------------------------------- C ABI

extern "C" {

void ams_default_params(ams_params_t* p) {
  // LDS tables up to two-thirds full: rows of up to 5461 products (distinct
  // columns) take an 8192-slot table.  Against 2730 (a third full): G500
  // s19 ef32 391.2 -> 388.7 ms, PaRMAT s19 ef32 382.8 -> 380.0 ms, ER s20
  // ef16 and the join-aggregate product unchanged, 2^24-column rows of
  // about 4096 products 26.5 -> 8.6 ms (an LDS table instead of a global one)
  p->sym_hash_max = 5461;
  p->num_hash_max = 5461;
  p->esc_max = 4096;
  p->acc_thres = 2.0;
  p->dense_width = 8448;
  p->bitmap_bits = 1 << 20;
  p->bitmap_ratio = 1.0;
  p->dense_density = 0.05;
  p->mode = AMS_MODE_HYBRID;
  p->wgs_per_cu = 0;
  p->esc_sort = 0;
  p->sym_blocked = 0;
  // rows too long for an LDS table with fewer products than a quarter of
  // their range's words: 2^24-column rows of about 4096 products (tables
  // then a third full) 764 -> 26 ms against dense chunks, G500 s19 ef32
  // unchanged (391.2 against 391.5 ms); 1.0 sends G500 rows that dense
  // chunks do 2.8 times faster to the global hash (497 ms), and every wide
  // row of the join-aggregate product goes 35 times slower there (740
  // against 21 ms)
  p->ghash_ratio = 0.25;
  // light morsels: at most 512 C entries at ranks from the kept row
  // bitmaps (341 in hash tables); G500 s19 ef32 dense route 243.2 ms at 341,
  // 237.7 ms at 512
  p->light_max = kLightCap;
  p->row_bitmaps = 1;
}

int ams_ctx_create(ams_ctx_t** out, void* stream) {
  if (!out) return AMS_ERR_ARG;
  ams_ctx_t* ctx = new ams_ctx_t();
  ctx->stream = (hipStream_t)stream;
  ams_default_params(&ctx->prm);
  if (hipGetDevice(&ctx->device) != hipSuccess) {
    (void)hipGetLastError();
    delete ctx;
    return AMS_ERR_HIP;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, ctx->device) == hipSuccess) {
    ctx->num_cu = prop.multiProcessorCount;
    ctx->lds_max = std::max(prop.sharedMemPerBlock, prop.sharedMemPerBlockOptin);
    ctx->mem_total = prop.totalGlobalMem;
  }
  if (hipHostMalloc((void**)&ctx->h_ctrl, sizeof(Ctrl), hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&ctx->h_pin, 64 * sizeof(int64_t), hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    ams_ctx_destroy(ctx);
    return AMS_ERR_ALLOC;
  }
  // this GPU's fixed scratch (ams_spgemm_fixed) and the context's own control
  // block, so that ams_spgemm_fixed reports its flags without allocating
  void* fs = nullptr;
  if (fixed_scratch_address(ctx->device, &fs) != hipSuccess || grow(ctx, ctx->ctrl, sizeof(Ctrl)) != AMS_OK ||
      fill_i32(ctx, (int32_t*)ctx->ctrl.p, sizeof(Ctrl) / 4, 0) != AMS_OK) {
    (void)hipGetLastError();
    ams_ctx_destroy(ctx);
    return AMS_ERR_HIP;
  }
  ctx->fixed_base = (unsigned char*)fs;
  *out = ctx;
  return AMS_OK;
}

int ams_ctx_destroy(ams_ctx_t* ctx) {
  if (!ctx) return AMS_OK;
  (void)hipStreamSynchronize(ctx->stream);
  Buf* bufs[] = {&ctx->row_flp, &ctx->sbin, &ctx->nbin, &ctx->lsym, &ctx->lnum, &ctx->ctrl, &ctx->coff,
                 &ctx->bptr, &ctx->bcol, &ctx->bval, &ctx->bperm, &ctx->tsum, &ctx->pool_rp, &ctx->pool_col,
                 &ctx->pool_val, &ctx->probs, &ctx->scratch_a, &ctx->dflag, &ctx->dtask, &ctx->ccnt,
                 &ctx->prow, &ctx->pbrow, &ctx->lrows, &ctx->hflag, &ctx->hrow, &ctx->hpiece, &ctx->gbm,
                 &ctx->gtab, &ctx->gnum, &ctx->bgrad, &ctx->rowbits};
  for (Buf* b : bufs)
    if (b->p) (void)hipFree(b->p);
  for (auto& r : ctx->recs) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  for (hipEvent_t e : ctx->ev_free) (void)hipEventDestroy(e);
  if (ctx->h_ctrl) (void)hipHostFree(ctx->h_ctrl);
  if (ctx->h_pin) (void)hipHostFree(ctx->h_pin);
  delete ctx;
  return AMS_OK;
}

int ams_set_stream(ams_ctx_t* ctx, void* stream) {
  if (!ctx) return AMS_ERR_ARG;
  ctx->stream = (hipStream_t)stream;
  return AMS_OK;
}

const char* ams_last_error(const ams_ctx_t* ctx) { return ctx ? ctx->err.c_str() : "no context"; }

int ams_set_params(ams_ctx_t* ctx, const ams_params_t* p) {
  if (!ctx || !p) return AMS_ERR_ARG;
  if (p->sym_hash_max < 1 || p->sym_hash_max > (kMaxHashSlots * 2) / 3)
    return fail(ctx, AMS_ERR_ARG, "sym_hash_max must be in [1, %d]", (kMaxHashSlots * 2) / 3);
  if (p->num_hash_max < 1 || p->num_hash_max > (kMaxHashSlots * 2) / 3)
    return fail(ctx, AMS_ERR_ARG, "num_hash_max must be in [1, %d]", (kMaxHashSlots * 2) / 3);
  if (p->esc_max < 0 || p->esc_max > kMaxEsc) return fail(ctx, AMS_ERR_ARG, "esc_max must be in [0, %d]", kMaxEsc);
  if (p->dense_width < 64 || p->dense_width > 65536 || p->dense_width % 64)
    return fail(ctx, AMS_ERR_ARG, "dense_width must be a multiple of 64 in [64, 65536]");
  if (p->bitmap_bits < p->dense_width) return fail(ctx, AMS_ERR_ARG, "bitmap_bits must be >= dense_width");
  if (p->mode < 0 || p->mode > 2) return fail(ctx, AMS_ERR_ARG, "unknown mode %d", p->mode);
  if (p->esc_sort < 0 || p->esc_sort > 4) return fail(ctx, AMS_ERR_ARG, "esc_sort must be in [0, 4]");
  if (p->sym_blocked < 0 || p->sym_blocked > 1) return fail(ctx, AMS_ERR_ARG, "sym_blocked must be 0 or 1");
  if (!(p->ghash_ratio >= 0.0)) return fail(ctx, AMS_ERR_ARG, "ghash_ratio must be >= 0");
  if (p->row_bitmaps < 0 || p->row_bitmaps > 1) return fail(ctx, AMS_ERR_ARG, "row_bitmaps must be 0 or 1");
  if (p->light_max < 0 || p->light_max > std::max(kLightCap, kLightMaxCnt))
    return fail(ctx, AMS_ERR_ARG, "light_max must be in [0, %d]", std::max(kLightCap, kLightMaxCnt));
  ctx->prm = *p;
  return AMS_OK;
}

int ams_get_params(const ams_ctx_t* ctx, ams_params_t* p) {
  if (!ctx || !p) return AMS_ERR_ARG;
  *p = ctx->prm;
  return AMS_OK;
}

int ams_get_stats(const ams_ctx_t* ctx, ams_stats_t* s) {
  if (!ctx || !s) return AMS_ERR_ARG;
  *s = ctx->stats;
  return AMS_OK;
}

int ams_row_flops(ams_ctx_t* ctx, const ams_dcsr_t* a, const ams_dcsr_t* b, int64_t* row_flp, int64_t* total) {
  CHK(check_pair(ctx, a, b, AMS_F64));
  CHK(grow(ctx, ctx->ctrl, sizeof(Ctrl)));
  Ctrl* c = dctrl(ctx);
  CHK(fill_i64(ctx, (int64_t*)&c->flops, 1, 0));
  const Csr A = to_csr(a), B = to_csr(b);
  CHK(launch_row_flops(ctx, A, B, row_flp, &c->flops));
  if (total) {
    HIPCHK(hipMemcpyAsync(ctx->h_pin, &c->flops, 8, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    *total = ctx->h_pin[0];
  }
  return AMS_OK;
}

int ams_scan_i64(ams_ctx_t* ctx, const int64_t* in, int64_t* out, int64_t n) {
  if (!ctx || n < 0 || (n > 0 && (!in || !out))) return fail(ctx, AMS_ERR_ARG, "bad scan arguments");
  // n outputs from in[0..n-1): the last input is not read.
  if (n == 0) return AMS_OK;
  if (in != out) {
    HIPCHK(hipMemcpyAsync(out, in, sizeof(int64_t) * (n - 1), hipMemcpyDeviceToDevice, ctx->stream));
  }
  CHK(fill_i64(ctx, out + n - 1, 1, 0));
  return scan_i64(ctx, out, out, n);
}

int ams_spgemm_symbolic(ams_ctx_t* ctx, const ams_dcsr_t* a, const ams_dcsr_t* b, int dtype, int64_t* rp_c,
                        int64_t* nnz) {
  CHK(check_pair(ctx, a, b, dtype));
  if (!rp_c) return fail(ctx, AMS_ERR_ARG, "row_ptr_c is NULL");
  return run_symbolic(ctx, a, b, dtype, single_probs(b), rp_c, true, nnz);
}

// The routing and column-chunk plan of the last symbolic call describes
// exactly these operands: same sizes, same row-pointer and column arrays
// (values may differ).
static bool plan_matches(const ams_ctx_t* ctx, const ams_dcsr_t* a, const ams_dcsr_t* b, int dtype) {
  return ctx->plan_valid && ctx->plan_m == a->num_row && ctx->plan_mB == b->num_row && ctx->plan_dtype == dtype &&
         ctx->plan_nnzA == a->num_nnz && ctx->plan_nnzB == b->num_nnz && ctx->plan_nB == b->num_col &&
         ctx->plan_arp == a->row_ptr && ctx->plan_acol == a->col && ctx->plan_brp == b->row_ptr &&
         ctx->plan_bcol == b->col;
}

// Reads the kernels' error flag word after the stream drains.
static int check_flags(ams_ctx_t* ctx, const char* what) {
  HIPCHK(hipMemcpyAsync(&ctx->h_ctrl->err, &dctrl(ctx)->err, 4, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  if (ctx->h_ctrl->err) return fail(ctx, AMS_ERR_STATE, "%s kernels raised error flags 0x%x", what, ctx->h_ctrl->err);
  return AMS_OK;
}

int ams_spgemm_numeric(ams_ctx_t* ctx, const ams_dcsr_t* a, const ams_dcsr_t* b, int dtype, const int64_t* rp_c,
                       int32_t* col_c, void* val_c) {
  CHK(check_pair(ctx, a, b, dtype));
  if (!plan_matches(ctx, a, b, dtype))
    return fail(ctx, AMS_ERR_STATE, "numeric needs the symbolic phase of the same A and B (same arrays) first");
  if (!rp_c || (ctx->stats.nnz_c > 0 && (!col_c || !val_c)))
    return fail(ctx, AMS_ERR_ARG, "output buffers are NULL");
  const ams_params_t &p0 = ctx->plan_prm, &p1 = ctx->prm;
  if (p0.sym_hash_max != p1.sym_hash_max || p0.dense_width != p1.dense_width || p0.bitmap_bits != p1.bitmap_bits ||
      p0.bitmap_ratio != p1.bitmap_ratio || p0.sym_blocked != p1.sym_blocked)
    return fail(ctx, AMS_ERR_STATE, "symbolic parameters changed since the symbolic phase; run it again");
  if (p0.num_hash_max != p1.num_hash_max || p0.esc_max != p1.esc_max || p0.acc_thres != p1.acc_thres ||
      p0.dense_density != p1.dense_density || p0.mode != p1.mode)
    CHK(reroute_numeric(ctx, a, rp_c));
  CHK(run_numeric(ctx, a, b, dtype, rp_c, col_c, val_c, INT64_MAX));
  return check_flags(ctx, "numeric");
}

int ams_get_error_flags(ams_ctx_t* ctx, int32_t* flags) {
  if (!ctx || !flags) return AMS_ERR_ARG;
  *flags = 0;
  if (!ctx->ctrl.p) return AMS_OK;
  HIPCHK(hipMemcpyAsync(&ctx->h_ctrl->err, &dctrl(ctx)->err, 4, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  *flags = ctx->h_ctrl->err;
  return AMS_OK;
}

// Releases the kept row bitmaps of the current plan; its numeric phase then
// sets the dense chunks' bitmaps per product.
static void drop_rowbits(ams_ctx_t* ctx) {
  (void)hipStreamSynchronize(ctx->stream);
  if (ctx->rowbits.p) (void)hipFree(ctx->rowbits.p);
  (void)hipGetLastError();
  ctx->rowbits = Buf{};
  ++ctx->gen;
  ctx->plan_pr.rowbits = nullptr;
  ctx->plan_pr.bwords = 0;
}

// C's pool of nnz entries; when it does not fit beside the kept row
// bitmaps, they are released and the pool is tried once more.
static int grow_c_pool(ams_ctx_t* ctx, int64_t nnz, size_t vb) {
  for (int attempt = 0;; ++attempt) {
    int rc = grow(ctx, ctx->pool_col, sizeof(int32_t) * std::max<int64_t>(nnz, 1));
    if (rc == AMS_OK) rc = grow(ctx, ctx->pool_val, vb * std::max<int64_t>(nnz, 1));
    if (rc != AMS_ERR_ALLOC || attempt > 0 || !ctx->plan_pr.rowbits) return rc;
    drop_rowbits(ctx);
  }
}

int ams_spgemm(ams_ctx_t* ctx, const ams_dcsr_t* a, const ams_dcsr_t* b, int dtype, ams_dcsr_t* c_out) {
  CHK(check_pair(ctx, a, b, dtype));
  if (!c_out) return fail(ctx, AMS_ERR_ARG, "c_out is NULL");
  CHK(grow(ctx, ctx->pool_rp, sizeof(int64_t) * (a->num_row + 1)));
  int64_t* rp = (int64_t*)ctx->pool_rp.p;
  int64_t nnz = 0;
  CHK(run_symbolic(ctx, a, b, dtype, single_probs(b), rp, true, &nnz));
  const size_t vb = dtype == AMS_F64 ? 8 : 4;
  CHK(grow_c_pool(ctx, nnz, vb));
  CHK(run_numeric(ctx, a, b, dtype, rp, (int32_t*)ctx->pool_col.p, ctx->pool_val.p, INT64_MAX));
  CHK(check_flags(ctx, "numeric"));
  c_out->num_row = a->num_row;
  c_out->num_col = b->num_col;
  c_out->num_nnz = nnz;
  c_out->row_ptr = rp;
  c_out->col = (const int32_t*)ctx->pool_col.p;
  c_out->val = ctx->pool_val.p;
  return AMS_OK;
}

int ams_verify_result(ams_ctx_t* ctx, const ams_dcsr_t* c, int dtype, int64_t* nnz, int64_t* sum_col,
                      double* sum_val, int64_t* unsorted) {
  if (!ctx) return AMS_ERR_ARG;
  CHK(check_csr(ctx, c, "C"));
  CHK(grow(ctx, ctx->ctrl, sizeof(Ctrl)));
  Ctrl* d = dctrl(ctx);
  unsigned long long* out = &d->sums[20];  // [20] col sum, [21] unsorted rows, [22] value sum
  CHK(fill_i64(ctx, (int64_t*)out, 3, 0));
  const Csr C = to_csr(c);
  if (C.m > 0) {
    const int g = (int)std::min<int64_t>((C.m + 3) / 4, (int64_t)ctx->num_cu * 16);
    if (dtype == AMS_F64)
      hipLaunchKernelGGL(k_verify<double>, dim3(g), dim3(256), 0, ctx->stream, C, out, (double*)(out + 2));
    else
      hipLaunchKernelGGL(k_verify<float>, dim3(g), dim3(256), 0, ctx->stream, C, out, (double*)(out + 2));
    HIPCHK(hipGetLastError());
  }
  HIPCHK(hipMemcpyAsync(ctx->h_pin, out, 24, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipMemcpyAsync(ctx->h_pin + 3, c->row_ptr + c->num_row, 8, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  if (sum_col) *sum_col = ctx->h_pin[0];
  if (unsorted) *unsorted = ctx->h_pin[1];
  if (sum_val) std::memcpy(sum_val, &ctx->h_pin[2], 8);
  if (nnz) *nnz = ctx->h_pin[3];
  return AMS_OK;
}

int ams_enable_timing(ams_ctx_t* ctx, int on) {
  if (!ctx) return AMS_ERR_ARG;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  for (auto& r : ctx->recs) {
    ctx->ev_free.push_back(r.a);
    ctx->ev_free.push_back(r.b);
  }
  ctx->recs.clear();
  for (int i = 0; i < AMS_PH_COUNT; ++i) {
    ctx->acc_ms[i] = 0;
    ctx->acc_n[i] = 0;
  }
  ctx->timing = on != 0;
  return AMS_OK;
}

int ams_get_timings(ams_ctx_t* ctx, double* ms, int64_t* launches) {
  if (!ctx) return AMS_ERR_ARG;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  for (auto& r : ctx->recs) {
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, r.a, r.b));
    ctx->acc_ms[r.ph] += t;
    ctx->acc_n[r.ph] += 1;
    ctx->ev_free.push_back(r.a);
    ctx->ev_free.push_back(r.b);
  }
  ctx->recs.clear();
  for (int i = 0; i < AMS_PH_COUNT; ++i) {
    if (ms) ms[i] = ctx->acc_ms[i];
    if (launches) launches[i] = ctx->acc_n[i];
  }
  return AMS_OK;
}

int ams_dev_alloc(void** ptr, int64_t bytes) {
  if (!ptr || bytes < 0) return AMS_ERR_ARG;
  if (hipMalloc(ptr, bytes > 0 ? (size_t)bytes : 8) != hipSuccess) {
    (void)hipGetLastError();
    *ptr = nullptr;
    return AMS_ERR_ALLOC;
  }
  return AMS_OK;
}

int ams_dev_free(void* ptr) {
  if (ptr && hipFree(ptr) != hipSuccess) {
    (void)hipGetLastError();
    return AMS_ERR_HIP;
  }
  return AMS_OK;
}

int ams_memcpy_h2d(ams_ctx_t* ctx, void* dst, const void* src, int64_t bytes) {
  if (!ctx) return AMS_ERR_ARG;
  if (bytes <= 0) return AMS_OK;
  HIPCHK(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return AMS_OK;
}

int ams_memcpy_d2h(ams_ctx_t* ctx, void* dst, const void* src, int64_t bytes) {
  if (!ctx) return AMS_ERR_ARG;
  if (bytes <= 0) return AMS_OK;
  HIPCHK(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return AMS_OK;
}

int ams_sync(ams_ctx_t* ctx) {
  if (!ctx) return AMS_ERR_ARG;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return AMS_OK;
}

int ams_i64_to_i32(ams_ctx_t* ctx, const int64_t* in, int32_t* out, int64_t n) {
  if (!ctx) return AMS_ERR_ARG;
  if (n <= 0) return AMS_OK;
  const int g = (int)std::min<int64_t>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(k_i64_to_i32, dim3(g), dim3(256), 0, ctx->stream, in, out, n);
  HIPCHK(hipGetLastError());
  return AMS_OK;
}

int ams_i32_to_i64(ams_ctx_t* ctx, const int32_t* in, int64_t* out, int64_t n) {
  if (!ctx) return AMS_ERR_ARG;
  if (n <= 0) return AMS_OK;
  const int g = (int)std::min<int64_t>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(k_i32_to_i64, dim3(g), dim3(256), 0, ctx->stream, in, out, n);
  HIPCHK(hipGetLastError());
  return AMS_OK;
}

int ams_batched_plan(ams_ctx_t* ctx, const ams_dcsr_t* a, const ams_dcsr_t* b, int dtype, int32_t nprob,
                     const int64_t* arow, const int64_t* brow, const int64_t* ccol, int64_t* cap) {
  CHK(check_pair(ctx, a, b, dtype));
  ctx->batch_ready = false;
  if (nprob < 1 || !arow || !brow || !ccol || !cap) return fail(ctx, AMS_ERR_ARG, "bad problem table");
  if (arow[0] != 0 || brow[0] != 0 || ccol[0] != 0 || arow[nprob] != a->num_row || brow[nprob] != a->num_col ||
      ccol[nprob] != b->num_col)
    return fail(ctx, AMS_ERR_ARG, "problem offsets do not span the stacked matrices");
  int64_t maxw = 0;
  for (int i = 0; i < nprob; ++i) {
    if (arow[i + 1] < arow[i] || brow[i + 1] < brow[i] || ccol[i + 1] < ccol[i])
      return fail(ctx, AMS_ERR_ARG, "problem offsets must be non-decreasing");
    maxw = std::max(maxw, ccol[i + 1] - ccol[i]);
  }
  Probs pr{};
  pr.n = nprob;
  pr.width0 = b->num_col;
  pr.max_width = nprob == 1 ? b->num_col : maxw;
  if (nprob > 1) {
    CHK(grow(ctx, ctx->probs, sizeof(int64_t) * 3 * (nprob + 1)));
    int64_t* d = (int64_t*)ctx->probs.p;
    std::vector<int64_t> h(3 * (size_t)(nprob + 1));
    std::memcpy(h.data(), arow, sizeof(int64_t) * (nprob + 1));
    std::memcpy(h.data() + (nprob + 1), brow, sizeof(int64_t) * (nprob + 1));
    std::memcpy(h.data() + 2 * (nprob + 1), ccol, sizeof(int64_t) * (nprob + 1));
    HIPCHK(hipMemcpyAsync(d, h.data(), sizeof(int64_t) * h.size(), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    pr.arow = d;
    pr.brow = d + (nprob + 1);
    pr.ccol = d + 2 * (nprob + 1);
    // per-row problem tables: kernels look a row's problem up with one load
    CHK(grow(ctx, ctx->prow, sizeof(int2) * std::max<int64_t>(a->num_row, 1)));
    CHK(grow(ctx, ctx->pbrow, sizeof(int32_t) * std::max<int64_t>(b->num_row, 1)));
    const int64_t mm = std::max<int64_t>(a->num_row, b->num_row);
    const int gp = (int)std::max<int64_t>(1, std::min<int64_t>((mm + 255) / 256, (int64_t)ctx->num_cu * 16));
    hipLaunchKernelGGL(k_prob_rows, dim3(gp), dim3(256), 0, ctx->stream, pr, a->num_row, b->num_row,
                       (int2*)ctx->prow.p, (int32_t*)ctx->pbrow.p);
    HIPCHK(hipGetLastError());
    pr.arow_cols = (const int2*)ctx->prow.p;
    pr.brow_lo = (const int32_t*)ctx->pbrow.p;
  }
  const int64_t m = a->num_row;
  CHK(grow(ctx, ctx->pool_rp, sizeof(int64_t) * (m + 1)));
  int64_t nnz = 0;
  CHK(run_symbolic(ctx, a, b, dtype, pr, (int64_t*)ctx->pool_rp.p, true, &nnz));
  // size every buffer for the unsynchronised run: all rows may be heavy
  int W, R, Wb;
  dense_geometry(ctx, dtype, pr.max_width, W, R, Wb);
  const int64_t ld = b->num_row + 1;
  CHK(grow(ctx, ctx->coff, sizeof(int32_t) * std::max<int64_t>(m, 1) * (R + 1)));
  CHK(grow(ctx, ctx->bptr, sizeof(int64_t) * (size_t)R * ld));
  CHK(grow(ctx, ctx->bcol, sizeof(uint16_t) * std::max<int64_t>(b->num_nnz, 1)));
  CHK(grow(ctx, ctx->bval, (dtype == AMS_F64 ? 8 : 4) * std::max<int64_t>(b->num_nnz, 1)));
  CHK(grow(ctx, ctx->ccnt, sizeof(int32_t) * std::max<int64_t>(m, 1) * R));
  CHK(grow(ctx, ctx->dflag, sizeof(int64_t) * (std::max<int64_t>(m, 1) * R + 1)));
  CHK(grow(ctx, ctx->dtask, sizeof(DenseTask) * std::max<int64_t>(m, 1) * R));
  CHK(grow(ctx, ctx->lrows, sizeof(int32_t) * std::max<int64_t>(b->num_row, 1)));
  const int64_t nmax = std::max<int64_t>(std::max<int64_t>(m + 1, (int64_t)R * ld), std::max<int64_t>(m, 1) * R + 1);
  CHK(grow(ctx, ctx->tsum, sizeof(int64_t) * ((nmax + kScanTile - 1) / kScanTile)));
  if (heavy_words(pr) > 0) CHK(grow_heavy(ctx, pr, m));
  ++ctx->gen;  // the problem tables were rewritten: graphs of earlier plans are stale
  ctx->batch_pr = pr;
  ctx->batch_m = m;
  ctx->batch_mB = b->num_row;
  ctx->batch_nnzA = a->num_nnz;
  ctx->batch_nnzB = b->num_nnz;
  ctx->batch_dtype = dtype;
  ctx->batch_ready = true;
  *cap = nnz;
  return AMS_OK;
}

int ams_batched_run(ams_ctx_t* ctx, const ams_dcsr_t* a, const ams_dcsr_t* b, int dtype, int64_t* c_rp,
                    int32_t* c_col, void* c_val, int64_t cap, int64_t* nnz_dev) {
  CHK(check_pair(ctx, a, b, dtype));
  if (!ctx->batch_ready || ctx->batch_m != a->num_row || ctx->batch_mB != b->num_row ||
      ctx->batch_nnzA != a->num_nnz || ctx->batch_nnzB != b->num_nnz || ctx->batch_dtype != dtype)
    return fail(ctx, AMS_ERR_STATE, "ams_batched_run needs ams_batched_plan of the same inputs");
  if (!c_rp || !nnz_dev || (cap > 0 && (!c_col || !c_val))) return fail(ctx, AMS_ERR_ARG, "NULL output buffer");
  CHK(run_symbolic(ctx, a, b, dtype, ctx->batch_pr, c_rp, false, nullptr));
  CHK(run_numeric(ctx, a, b, dtype, c_rp, c_col, c_val, cap));
  hipLaunchKernelGGL(k_copy_i64, dim3(1), dim3(64), 0, ctx->stream, (const int64_t*)(c_rp + a->num_row), nnz_dev);
  HIPCHK(hipGetLastError());
  return AMS_OK;
}

// The context's scratch buffers, swapped for slices of the fixed scratch for
// the duration of one ams_spgemm_fixed call and restored afterwards.
struct FixedSwap {
  ams_ctx_t* ctx;
  static constexpr int kN = 29;
  Buf* bufs[kN];
  Buf saved[kN];
  explicit FixedSwap(ams_ctx_t* c) : ctx(c) {
    Buf* all[kN] = {&c->row_flp, &c->sbin, &c->nbin, &c->lsym, &c->lnum, &c->ctrl, &c->coff,
                    &c->bptr, &c->bcol, &c->bval, &c->bperm, &c->tsum, &c->pool_rp, &c->pool_col,
                    &c->pool_val, &c->probs, &c->scratch_a, &c->dflag, &c->dtask, &c->ccnt,
                    &c->prow, &c->pbrow, &c->lrows, &c->hflag, &c->hrow, &c->hpiece, &c->gbm,
                    &c->gtab, &c->gnum};
    for (int i = 0; i < kN; ++i) {
      bufs[i] = all[i];
      saved[i] = *all[i];
      *all[i] = Buf{};  // unused buffers stay empty: any use fails instead of touching the context's memory
    }
    c->fixed = true;
  }
  ~FixedSwap() {
    for (int i = 0; i < kN; ++i) *bufs[i] = saved[i];
    ctx->fixed = false;
    // the plan of this call lived in the fixed scratch
    ctx->plan_valid = false;
    ctx->plan_counts = false;
    ctx->bblk_ready = false;
  }
};

int ams_spgemm_fixed(ams_ctx_t* ctx, const ams_dcsr_t* a, const ams_dcsr_t* b, int dtype, int64_t* c_rp,
                     int32_t* c_col, void* c_val, int64_t cap, int64_t* nnz_dev) {
  CHK(check_pair(ctx, a, b, dtype));
  if (!c_rp || !nnz_dev || cap < 0 || (cap > 0 && (!c_col || !c_val)))
    return fail(ctx, AMS_ERR_ARG, "NULL output buffer or negative capacity");
  if (!ctx->fixed_base) return fail(ctx, AMS_ERR_STATE, "no fixed scratch address on this context's GPU");
  int dev = -1;
  HIPCHK(hipGetDevice(&dev));
  if (dev != ctx->device) return fail(ctx, AMS_ERR_STATE, "the context belongs to GPU %d, GPU %d is current",
                                      ctx->device, dev);
  if (ctx->capturing) return fail(ctx, AMS_ERR_STATE, "a batched capture is in progress on this context");
  const Probs pr = single_probs(b);
  int W, R, Wb;
  dense_geometry(ctx, dtype, pr.max_width, W, R, Wb);
  const int64_t m1 = std::max<int64_t>(a->num_row, 1), mB = b->num_row;
  const int64_t nnzB1 = std::max<int64_t>(b->num_nnz, 1);
  const size_t vb = dtype == AMS_F64 ? 8 : 4;
  // every scan of the unsynchronised run: row pointers of C, the blocked
  // copy's chunk pointers, the dense morsel flags
  const int64_t nscan = std::max<int64_t>(std::max<int64_t>(m1 + 1, (int64_t)R * (mB + 1)), m1 * R + 1);
  // the buffers the unsynchronised run grows (every row may be a wide row)
  struct Part { Buf* buf; size_t bytes; };
  const Part parts[] = {
      {&ctx->ctrl, sizeof(Ctrl)},
      {&ctx->row_flp, sizeof(int64_t) * (size_t)m1},
      {&ctx->sbin, (size_t)m1},
      {&ctx->nbin, (size_t)m1},
      {&ctx->lsym, sizeof(int32_t) * (size_t)m1},
      {&ctx->lnum, sizeof(int32_t) * (size_t)m1},
      {&ctx->coff, sizeof(int32_t) * (size_t)m1 * (R + 1)},
      {&ctx->ccnt, ctx->prm.sym_blocked ? sizeof(int32_t) * (size_t)m1 * R : 0},
      {&ctx->bptr, sizeof(int64_t) * (size_t)R * (mB + 1)},
      {&ctx->bcol, sizeof(uint16_t) * (size_t)nnzB1},
      {&ctx->bval, vb * (size_t)nnzB1},
      {&ctx->lrows, sizeof(int32_t) * (size_t)std::max<int64_t>(mB, 1)},
      {&ctx->dflag, sizeof(int64_t) * (size_t)(m1 * R + 1)},
      {&ctx->dtask, sizeof(DenseTask) * (size_t)(m1 * R)},
      {&ctx->tsum, sizeof(int64_t) * (size_t)((nscan + kScanTile - 1) / kScanTile)},
  };
  size_t need = 0;
  for (const Part& q : parts) need += (q.bytes + 255) & ~(size_t)255;
  if (need > kFixedScratchBytes)
    return fail(ctx, AMS_ERR_LIMIT, "this product needs %zu bytes of scratch; the fixed scratch holds %zu "
                "(use ams_spgemm or ams_spgemm_symbolic/numeric, which allocate)", need, kFixedScratchBytes);
  Ctrl* own = dctrl(ctx);  // the context's control block receives the flags
  int rc;
  {
    FixedSwap swap(ctx);
    size_t off = 0;
    for (const Part& q : parts) {
      if (q.bytes == 0) continue;
      *q.buf = Buf{ctx->fixed_base + off, q.bytes};
      off += (q.bytes + 255) & ~(size_t)255;
    }
    rc = run_symbolic(ctx, a, b, dtype, pr, c_rp, false, nullptr);
    if (rc == AMS_OK) rc = run_numeric(ctx, a, b, dtype, c_rp, c_col, c_val, cap);
    if (rc == AMS_OK) {
      hipLaunchKernelGGL(k_copy_i64, dim3(1), dim3(64), 0, ctx->stream, (const int64_t*)(c_rp + a->num_row),
                         nnz_dev);
      hipLaunchKernelGGL(k_copy_i32, dim3(1), dim3(64), 0, ctx->stream, (const int32_t*)&dctrl(ctx)->err,
                         &own->err);
      const hipError_t e = hipGetLastError();
      if (e != hipSuccess) rc = fail(ctx, AMS_ERR_HIP, "launch: %s", hipGetErrorString(e));
    }
  }
  return rc;
}

int ams_batched_capture(ams_ctx_t* ctx, const ams_dcsr_t* a, const ams_dcsr_t* b, int dtype, int64_t* c_rp,
                        int32_t* c_col, void* c_val, int64_t cap, int64_t* nnz_dev, void** exec) {
  if (!ctx || !exec) return AMS_ERR_ARG;
  if (!ctx->stream) return fail(ctx, AMS_ERR_ARG, "graph capture needs a non-null stream");
  const bool timing = ctx->timing;
  ctx->timing = false;  // no phase events inside a graph
  ctx->capturing = true;
  hipError_t e0 = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed);
  if (e0 != hipSuccess) {
    ctx->capturing = false;
    ctx->timing = timing;
    return fail(ctx, AMS_ERR_HIP, "hipStreamBeginCapture: %s", hipGetErrorString(e0));
  }
  const int r = ams_batched_run(ctx, a, b, dtype, c_rp, c_col, c_val, cap, nnz_dev);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
  ctx->capturing = false;
  ctx->timing = timing;
  if (r != AMS_OK) {
    if (g) (void)hipGraphDestroy(g);
    return r;
  }
  if (e != hipSuccess) return fail(ctx, AMS_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
  hipGraphExec_t ge = nullptr;
  const hipError_t e2 = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e2 != hipSuccess) return fail(ctx, AMS_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e2));
  AmsGraph* h = new AmsGraph{ge, ctx, ctx->gen};
  *exec = (void*)h;
  return AMS_OK;
}

int ams_graph_launch(ams_ctx_t* ctx, void* exec) {
  if (!ctx || !exec) return AMS_ERR_ARG;
  const AmsGraph* h = (const AmsGraph*)exec;
  if (h->ctx != ctx) return fail(ctx, AMS_ERR_ARG, "the graph was captured on another context");
  if (h->gen != ctx->gen)
    return fail(ctx, AMS_ERR_STATE, "the graph is stale: the context's scratch moved or was re-planned since its "
                "capture; capture it again");
  HIPCHK(hipGraphLaunch(h->exec, ctx->stream));
  return AMS_OK;
}

int ams_graph_destroy(void* exec) {
  if (!exec) return AMS_OK;
  AmsGraph* h = (AmsGraph*)exec;
  const hipError_t e = hipGraphExecDestroy(h->exec);
  delete h;
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return AMS_ERR_HIP;
  }
  return AMS_OK;
}

// Coordinate search over the routing thresholds and LDS geometry, timing the
// whole one-shot product with events on the context stream.
int ams_autotune(ams_ctx_t* ctx, const ams_dcsr_t* a, const ams_dcsr_t* b, int dtype, int reps,
                 ams_params_t* best_out, double* best_ms_out) {
  CHK(check_pair(ctx, a, b, dtype));
  if (reps < 1) reps = 1;
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  auto timed = [&](const ams_params_t& p, double* ms) -> int {
    CHK(ams_set_params(ctx, &p));
    ams_dcsr_t c;
    CHK(ams_spgemm(ctx, a, b, dtype, &c));
    HIPCHK(hipEventRecord(e0, ctx->stream));
    for (int i = 0; i < reps; ++i) CHK(ams_spgemm(ctx, a, b, dtype, &c));
    HIPCHK(hipEventRecord(e1, ctx->stream));
    HIPCHK(hipEventSynchronize(e1));
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, e0, e1));
    *ms = (double)t / reps;
    return AMS_OK;
  };
  ams_params_t best = ctx->prm;
  double best_ms = 0;
  int r = timed(best, &best_ms);
  if (r != AMS_OK) {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return r;
  }
  const int hash_pairs[][2] = {{682, 682}, {1365, 1365}, {2730, 2730}, {1365, 2730}, {2730, 5461}, {5461, 5461}};
  const int widths[] = {4288, 8448, 12800, 17344, 24576};
  const double thres[] = {0.0, 1.25, 2.0, 4.0};
  const double ratios[] = {0.25, 1.0, 4.0};
  const double dens[] = {0.02, 0.05, 0.2, 1.0};
  const double ghash[] = {0.0, 0.25, 1.0};
  auto consider = [&](ams_params_t p) {
    double ms;
    if (timed(p, &ms) == AMS_OK && ms < best_ms) {
      best_ms = ms;
      best = p;
    }
  };
  for (auto& hp : hash_pairs) { ams_params_t p = best; p.sym_hash_max = hp[0]; p.num_hash_max = hp[1]; consider(p); }
  for (int w : widths) { ams_params_t p = best; p.dense_width = w; if (p.bitmap_bits < w) p.bitmap_bits = w; consider(p); }
  for (double t : thres) { ams_params_t p = best; p.acc_thres = t; consider(p); }
  for (double q : ratios) { ams_params_t p = best; p.bitmap_ratio = q; consider(p); }
  for (double d : dens) { ams_params_t p = best; p.dense_density = d; consider(p); }
  for (int e = 0; e < 5; ++e) { ams_params_t p = best; p.esc_sort = e; consider(p); }
  for (int e = 0; e < 2; ++e) { ams_params_t p = best; p.sym_blocked = e; consider(p); }
  for (double g : ghash) { ams_params_t p = best; p.ghash_ratio = g; consider(p); }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  ctx->err.clear();
  CHK(ams_set_params(ctx, &best));
  if (best_out) *best_out = best;
  if (best_ms_out) *best_ms_out = best_ms;
  return AMS_OK;
}

int ams_spgemm_masked_bwd(ams_ctx_t* ctx, const ams_dcsr_t* a, const ams_dcsr_t* b, const ams_dcsr_t* g, int dtype,
                          void* da, void* db) {
  CHK(check_pair(ctx, a, b, dtype));
  CHK(check_csr(ctx, g, "G"));
  if (g->num_row != a->num_row || g->num_col != b->num_col)
    return fail(ctx, AMS_ERR_ARG, "G must be num_row(A) x num_col(B)");
  if ((a->num_nnz > 0 && !da) || (b->num_nnz > 0 && !db)) return fail(ctx, AMS_ERR_ARG, "NULL gradient buffer");
  if (dtype == AMS_F64) return masked_bwd_t<double>(ctx, a, b, g, (double*)da, (double*)db);
  return masked_bwd_t<float>(ctx, a, b, g, (float*)da, (float*)db);
}

}  // extern "C"
