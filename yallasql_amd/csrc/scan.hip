// scan.hip — single-pass inclusive scans and stream compaction.
//
// One workgroup of 256 threads (4 waves) owns a tile of 4096 rows; a tile
// takes its id from an atomic counter, computes its local aggregate, and
// finds its exclusive prefix by decoupled look-back over the predecessors'
// status words (wave 0 checks 64 predecessors per step).  Each row of input
// is read once and each output written once: the reference's chained
// Brent-Kung scan with a serial carry (prefix_sum_kernel.cu:7-103), its
// separate mask, map and per-column scatter passes (filter_operator.hpp:31-98)
// become one pass.
#include <atomic>
#include <mutex>

#include "expr_dev.hpp"
#include "scan_dev.hpp"

namespace ysql {

constexpr int kScanBlock = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanBlock * kScanItems;  // 4096

struct ScanWs {
  uint32_t *ctr;
  uint64_t *status;
};

static size_t scan_ws_size(uint64_t n) {
  WsCounter c;
  c.take<uint32_t>(64);
  c.take<uint64_t>(ceil_div(n, kScanTile) + 1);
  return c.off;
}

static bool scan_ws_carve(void *ws, size_t bytes, uint64_t n, ScanWs *out) {
  WsCarver c(ws, bytes);
  out->ctr = c.take<uint32_t>(64);
  out->status = c.take<uint64_t>(ceil_div(n, kScanTile) + 1);
  return out->ctr && out->status;
}

__global__ void k_scan_reset(uint32_t *ctr, uint64_t *status, uint32_t ntiles) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) ctr[0] = 0;
  for (; i < ntiles; i += gridDim.x * blockDim.x) status[i] = 0;
}

static int scan_reset(const ScanWs &w, uint32_t ntiles, hipStream_t s) {
  uint32_t g = (uint32_t)ceil_div(ntiles, kScanBlock);
  if (g > 1024) g = 1024;
  if (g < 1) g = 1;
  k_scan_reset<<<g, kScanBlock, 0, s>>>(w.ctr, w.status, ntiles);
  YSQL_LAUNCH_CHECK();
  return YSQL_OK;
}

__device__ __forceinline__ uint32_t take_tile(uint32_t *ctr, uint32_t *s_tile) {
  if (threadIdx.x == 0) *s_tile = atomicAdd(ctr, 1u);
  __syncthreads();
  return *s_tile;
}

// ---------------------------------------------------------------------------
// Inclusive scan of uint32 values or of a byte mask (nonzero = 1).
// Rows go HBM -> LDS striped (coalesced 16-byte loads), each thread then
// scans 16 consecutive rows from LDS (padded by one word per 16 to avoid
// bank conflicts), and results go back through LDS to coalesced stores.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int pad16(int e) { return e + (e >> 4); }

template <bool MASK>
__global__ __launch_bounds__(kScanBlock) void k_scan(const void *in_, uint32_t *out, uint64_t n, uint32_t *ctr,
                                                     uint64_t *status) {
  __shared__ uint32_t s_val[kScanTile + kScanTile / 16];
  __shared__ uint64_t s_red[kScanBlock / kWave + 1];
  __shared__ uint32_t s_tile;
  __shared__ uint64_t s_prefix;
  const uint32_t tile = take_tile(ctr, &s_tile);
  const uint64_t base = (uint64_t)tile * kScanTile;
  const int t = threadIdx.x;
  const uint64_t rows = (n - base) < (uint64_t)kScanTile ? (n - base) : (uint64_t)kScanTile;

  if (MASK) {
    const uint8_t *in = (const uint8_t *)in_;
    // 16 mask bytes per thread: one 16-byte load, already in blocked order.
    uint32_t v[16];
    if (rows == kScanTile && ((uintptr_t)(in + base) & 15) == 0) {
      uint4 q = *(const uint4 *)(in + base + t * 16);
      uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
      for (int e = 0; e < 16; e++) v[e] = ((w[e >> 2] >> (8 * (e & 3))) & 0xFF) != 0;
    } else {
#pragma unroll
      for (int e = 0; e < 16; e++) {
        uint64_t r = (uint64_t)t * 16 + e;
        v[e] = (r < rows) ? (in[base + r] != 0) : 0;
      }
    }
    uint32_t run = 0;
#pragma unroll
    for (int e = 0; e < 16; e++) {
      run += v[e];
      v[e] = run;
    }
    uint64_t total;
    uint64_t excl = block_exclusive_sum<kScanBlock, uint64_t>(run, s_red, &total);
    if (t < kWave) {
      uint64_t p = lookback_exclusive(status, tile, total);
      if (t == 0) s_prefix = p;
    }
    __syncthreads();
    const uint32_t add = (uint32_t)(s_prefix + excl);
#pragma unroll
    for (int e = 0; e < 16; e++) s_val[pad16(t * 16 + e)] = v[e] + add;
    __syncthreads();
  } else {
    const uint32_t *in = (const uint32_t *)in_;
    // striped loads: round j covers rows [j*1024, j*1024+1024), 4 per thread
#pragma unroll
    for (int j = 0; j < kScanItems / 4; j++) {
      const int r0 = j * 1024 + t * 4;
      if (base + r0 + 4 <= n && ((uintptr_t)(in + base) & 15) == 0) {
        uint4 q = *(const uint4 *)(in + base + r0);
        s_val[pad16(r0 + 0)] = q.x;
        s_val[pad16(r0 + 1)] = q.y;
        s_val[pad16(r0 + 2)] = q.z;
        s_val[pad16(r0 + 3)] = q.w;
      } else {
#pragma unroll
        for (int e = 0; e < 4; e++) s_val[pad16(r0 + e)] = (base + r0 + e < n) ? in[base + r0 + e] : 0u;
      }
    }
    __syncthreads();
    uint32_t v[16];
    uint32_t run = 0;
#pragma unroll
    for (int e = 0; e < 16; e++) {
      run += s_val[pad16(t * 16 + e)];
      v[e] = run;
    }
    // u32 prefix sums wrap modulo 2^32 like the reference's uint32 scan
    uint64_t total;
    uint64_t excl = block_exclusive_sum<kScanBlock, uint64_t>(run, s_red, &total);
    if (t < kWave) {
      uint64_t p = lookback_exclusive(status, tile, total & 0xFFFFFFFFull);
      if (t == 0) s_prefix = p;
    }
    __syncthreads();
    const uint32_t add = (uint32_t)(s_prefix + excl);
#pragma unroll
    for (int e = 0; e < 16; e++) s_val[pad16(t * 16 + e)] = v[e] + add;
    __syncthreads();
  }
  // striped coalesced stores
#pragma unroll
  for (int j = 0; j < kScanItems / 4; j++) {
    const int r0 = j * 1024 + t * 4;
    if (base + r0 + 4 <= n && ((uintptr_t)(out + base) & 15) == 0) {
      uint4 q;
      q.x = s_val[pad16(r0 + 0)];
      q.y = s_val[pad16(r0 + 1)];
      q.z = s_val[pad16(r0 + 2)];
      q.w = s_val[pad16(r0 + 3)];
      *(uint4 *)(out + base + r0) = q;
    } else {
#pragma unroll
      for (int e = 0; e < 4; e++)
        if (base + r0 + e < n) out[base + r0 + e] = s_val[pad16(r0 + e)];
    }
  }
}

// ---------------------------------------------------------------------------
// Stream compaction of up to 8 columns under a mask or a fused predicate.
// Each thread owns 16 consecutive rows of the tile; the selected rows' tile
// offsets are listed in LDS in order (s_src), and every column is then
// copied with one output row per thread per step: coalesced stores, and
// loads that stay inside the tile's cache lines.
// ---------------------------------------------------------------------------
struct ColArgs {
  ysql_col c[YSQL_MAX_COLS];
};

constexpr int kCmpRange = 100;  // internal: lo <= pred < hi

template <int MODE, typename PT, int CMP>
__global__ __launch_bounds__(kScanBlock) void k_compact(const uint8_t *mask, const PT *pred, const PT *scalar, ColArgs cols,
                                                        int ncols, uint64_t n, uint64_t cap, uint64_t *count_out,
                                                        uint32_t *ctr, uint64_t *status, uint32_t ntiles) {
  __shared__ CompactSmem<kScanBlock, (MODE == 0 || sizeof(PT) != 4) ? 0 : 4> sm;
  __shared__ uint32_t s_tile;
  const uint32_t tile = take_tile(ctr, &s_tile);
  const PT s = MODE >= 1 ? scalar[0] : PT();
  const PT hi = MODE == 2 ? scalar[1] : PT();
  compact_tile<kScanBlock, MODE, PT, CMP>(sm, tile, mask, pred, s, hi, cols.c, ncols, n, cap, 0, count_out, status,
                                          ntiles);
}

// ---------------------------------------------------------------------------
// Filter by a fused expression program (expr_dev.hpp): the program runs over
// the tile in striped steps (thread t: rows 256 R j + R t .. + R-1, so each
// load instruction of a wave reads one contiguous range), its BOOL result
// goes to LDS as one byte per row, and each thread then takes the flags of
// its 16 consecutive rows for the stable compaction of compact_scatter.
// stage_in >= 0 (32-bit programs): that input column is also an output
// column (stage_src); its values, loaded for the program anyway, are kept in
// the LDS stage and copied out from there instead of read again after the
// tile's look-back.
// ---------------------------------------------------------------------------
template <typename S>
__global__ __launch_bounds__(kScanBlock) void k_xfilter(XProg p, ColArgs cols, int ncols, uint64_t n,
                                                        uint64_t *count_out, uint32_t *ctr, uint64_t *status,
                                                        uint32_t ntiles, int stage_in, const void *stage_src) {
  constexpr int R = sizeof(S) == 8 ? 1 : 4;  // rows per thread per step (64-bit slots: the file is twice as wide)
  constexpr int VB = sizeof(S) == 4 ? 4 : 0;
  __shared__ CompactSmem<kScanBlock, VB> sm;
  __shared__ uint32_t s_sel[kScanTile / 4];
  __shared__ uint32_t s_tile;
  const uint32_t tile = take_tile(ctr, &s_tile);
  const uint64_t base = (uint64_t)tile * kScanTile;
  const int t = threadIdx.x;
  XRegs<S, R> reg;
  reg.clear();
  S acc[R] = {};
  uint32_t nacc = 0;
  uint8_t *sel = reinterpret_cast<uint8_t *>(s_sel);
#pragma unroll
  for (int j = 0; j < kScanItems / R; j++) {
    const uint32_t r0 = j * (kScanBlock * R) + t * R;
    const bool fast = p.aligned && base + (uint64_t)(j + 1) * (kScanBlock * R) <= n;  // wave-uniform
    xload_inputs<S, R, false>(p, base + r0, n, fast, reg);
    if constexpr (VB == 4) {
      if (stage_in >= 0) {  // wave-uniform
        S sv[R];
        xget(reg, stage_in, sv);
#pragma unroll
        for (int e = 0; e < R; e++) sm.pv[r0 + e] = (uint32_t)sv[e];
      }
    }
    xrun<S, R, false>(p, reg, acc, nacc);
    S pr[R];
    if (p.pred == kXAcc) {
#pragma unroll
      for (int e = 0; e < R; e++) pr[e] = acc[e];
    } else {
      xget(reg, p.pred, pr);
    }
    if constexpr (R == 4) {
      s_sel[r0 / 4] = (uint32_t)(pr[0] != 0) | (uint32_t)(pr[1] != 0) << 8 | (uint32_t)(pr[2] != 0) << 16 |
                      (uint32_t)(pr[3] != 0) << 24;
    } else {
#pragma unroll
      for (int e = 0; e < R; e++) sel[r0 + e] = (uint8_t)(pr[e] != 0);
    }
  }
  __syncthreads();
  const uint4 q = reinterpret_cast<const uint4 *>(s_sel)[t];
  const uint32_t w[4] = {q.x, q.y, q.z, q.w};
  uint32_t flags = 0;
#pragma unroll
  for (int e = 0; e < 16; e++) flags |= (uint32_t)(((w[e >> 2] >> (8 * (e & 3))) & 0xFF) != 0) << e;
  compact_scatter<kScanBlock, VB, uint32_t>(sm, tile, flags, stage_in >= 0 ? (const uint32_t *)stage_src : nullptr,
                                            cols.c, ncols, n, ~0ull, 0, count_out, status, ntiles);
}

static bool cols_ok(const ysql_col *cols, int ncols, bool need_ptrs) {
  if (ncols < 0 || ncols > YSQL_MAX_COLS || (ncols > 0 && !cols)) return false;
  for (int i = 0; i < ncols; i++) {
    uint32_t eb = cols[i].elem_bytes;
    if (!(eb == 1 || eb == 2 || eb == 4 || eb == 8 || eb == YSQL_STRING_BYTES)) return false;
    if (need_ptrs && (!cols[i].src || !cols[i].dst)) return false;
    if (cols[i].src && cols[i].src == cols[i].dst) return false;  // the copy reads and writes distinct columns
  }
  return true;
}

// Writes 0 to *count_out when there are no rows (no tile would).
__global__ void k_store_zero(uint64_t *p) { *p = 0; }

template <typename PT>
static void launch_pred(int cmp, dim3 g, hipStream_t s, const PT *pred, const PT *scalar, const ColArgs &ca, int ncols,
                        uint64_t n, uint64_t cap, uint64_t *count_out, const ScanWs &w, uint32_t ntiles) {
  if (cmp == kCmpRange) {
    k_compact<2, PT, YSQL_LT><<<g, kScanBlock, 0, s>>>(nullptr, pred, scalar, ca, ncols, n, cap, count_out, w.ctr,
                                                       w.status, ntiles);
    return;
  }
#define PRED_CASE(C)                                                                                              \
  case C:                                                                                                         \
    k_compact<1, PT, C><<<g, kScanBlock, 0, s>>>(nullptr, pred, scalar, ca, ncols, n, cap, count_out, w.ctr,      \
                                                 w.status, ntiles);                                               \
    break;
  switch (cmp) {
    PRED_CASE(YSQL_LT)
    PRED_CASE(YSQL_LE)
    PRED_CASE(YSQL_GT)
    PRED_CASE(YSQL_GE)
    PRED_CASE(YSQL_EQ)
    PRED_CASE(YSQL_NE)
  }
#undef PRED_CASE
}

static int compact_impl(int mode, int dtype, int cmp, const uint8_t *mask, const void *pred, const void *scalar,
                        const ysql_col *cols, int ncols, uint64_t n, uint64_t cap, uint64_t *count_out, void *ws,
                        size_t ws_bytes, hipStream_t s) {
  if (!cols_ok(cols, ncols, n > 0)) return YSQL_EINVAL;
  if (n >= (1ull << 40)) return YSQL_EINVAL;
  if (n == 0) {
    if (count_out) {
      k_store_zero<<<1, 1, 0, s>>>(count_out);
      YSQL_LAUNCH_CHECK();
    }
    return YSQL_OK;
  }
  if (mode == 0 && !mask) return YSQL_EINVAL;
  if (mode == 1 && (!pred || !scalar || ((cmp < YSQL_LT || cmp > YSQL_NE) && cmp != kCmpRange))) return YSQL_EINVAL;
  ScanWs w;
  if (!ws || ws_bytes < scan_ws_size(n) || !scan_ws_carve(ws, ws_bytes, n, &w)) return YSQL_EWORKSPACE;
  const uint32_t ntiles = (uint32_t)ceil_div(n, kScanTile);
  int rc = scan_reset(w, ntiles, s);
  if (rc) return rc;
  ColArgs ca;
  memset(&ca, 0, sizeof(ca));
  for (int i = 0; i < ncols; i++) ca.c[i] = cols[i];
  dim3 g(ntiles);
  if (mode == 0) {
    k_compact<0, uint8_t, 0><<<g, kScanBlock, 0, s>>>(mask, nullptr, nullptr, ca, ncols, n, cap, count_out, w.ctr,
                                                      w.status, ntiles);
  } else {
    switch (dtype) {
      case YSQL_INT: launch_pred<int32_t>(cmp, g, s, (const int32_t *)pred, (const int32_t *)scalar, ca, ncols, n, cap, count_out, w, ntiles); break;
      case YSQL_FLOAT: launch_pred<float>(cmp, g, s, (const float *)pred, (const float *)scalar, ca, ncols, n, cap, count_out, w, ntiles); break;
      case YSQL_DATETIME: launch_pred<int64_t>(cmp, g, s, (const int64_t *)pred, (const int64_t *)scalar, ca, ncols, n, cap, count_out, w, ntiles); break;
      default: return YSQL_EINVAL;
    }
  }
  YSQL_LAUNCH_CHECK();
  return YSQL_OK;
}

// ---------------------------------------------------------------------------
// Scratch for ysql_filter_compact_noalloc: the tile counter and look-back
// status words of a compaction of up to kNoAllocMaxRows rows, in a static
// device array (each GPU the library runs on holds its own copy until the
// process ends).  Host code never passes `scratch` itself to a kernel: its
// device address comes from hipGetSymbolAddress, resolved once per device on
// that device's first call and kept in g_scratch_addr.
// ---------------------------------------------------------------------------
constexpr uint64_t kNoAllocMaxRows = 1ull << 32;
// the counter words, then one 8-byte status word per 4096-row tile (+1), each piece 256-byte aligned
constexpr size_t kNoAllocScratchBytes = 256 + (8 * (kNoAllocMaxRows / kScanTile + 1) + 255) / 256 * 256;  // ~8 MiB
constexpr int kNoAllocMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

static std::atomic<void *> g_scratch_addr[kNoAllocMaxDevices];
static std::mutex g_scratch_mu;

// The current device's address of `scratch`, or nullptr (error text set).
static void *noalloc_scratch() {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kNoAllocMaxDevices) {
    set_last_error("ysql_filter_compact_noalloc: no usable current device");
    return nullptr;
  }
  void *p = g_scratch_addr[dev].load(std::memory_order_acquire);
  if (p) return p;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  p = g_scratch_addr[dev].load(std::memory_order_relaxed);
  if (p) return p;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    set_last_error(e != hipSuccess ? hipGetErrorString(e) : "hipGetSymbolAddress returned no address");
    return nullptr;
  }
  g_scratch_addr[dev].store(p, std::memory_order_release);
  return p;
}

}  // namespace ysql

using namespace ysql;

extern "C" {

int ysql_scan_ws_bytes(uint64_t n, size_t *bytes) {
  if (!bytes) return YSQL_EINVAL;
  *bytes = scan_ws_size(n);
  return YSQL_OK;
}

static int scan_common(bool mask, const void *in, uint32_t *out, uint64_t n, void *ws, size_t ws_bytes,
                       ysql_stream_t stream) {
  if (n == 0) return YSQL_OK;
  if (!in || !out || n >= (1ull << 40)) return YSQL_EINVAL;
  ScanWs w;
  if (!ws || ws_bytes < scan_ws_size(n) || !scan_ws_carve(ws, ws_bytes, n, &w)) return YSQL_EWORKSPACE;
  hipStream_t s = as_stream(stream);
  const uint32_t ntiles = (uint32_t)ceil_div(n, kScanTile);
  int rc = scan_reset(w, ntiles, s);
  if (rc) return rc;
  if (mask)
    k_scan<true><<<ntiles, kScanBlock, 0, s>>>(in, out, n, w.ctr, w.status);
  else
    k_scan<false><<<ntiles, kScanBlock, 0, s>>>(in, out, n, w.ctr, w.status);
  YSQL_LAUNCH_CHECK();
  return YSQL_OK;
}

int ysql_scan_u32_inclusive(const uint32_t *in, uint32_t *out, uint64_t n, void *ws, size_t ws_bytes,
                            ysql_stream_t stream) {
  return scan_common(false, in, out, n, ws, ws_bytes, stream);
}

int ysql_scan_mask_inclusive(const uint8_t *mask, uint32_t *out, uint64_t n, void *ws, size_t ws_bytes,
                             ysql_stream_t stream) {
  return scan_common(true, mask, out, n, ws, ws_bytes, stream);
}

int ysql_compact_ws_bytes(uint64_t n, size_t *bytes) {
  if (!bytes) return YSQL_EINVAL;
  *bytes = scan_ws_size(n);
  return YSQL_OK;
}

int ysql_compact_cols(const uint8_t *mask, const ysql_col *cols, int ncols, uint64_t n, uint64_t *count_out, void *ws,
                      size_t ws_bytes, ysql_stream_t stream) {
  return compact_impl(0, 0, 0, mask, nullptr, nullptr, cols, ncols, n, ~0ull, count_out, ws, ws_bytes,
                      as_stream(stream));
}

int ysql_filter_compact(int dtype, int cmp, const void *pred_col, const void *scalar, const ysql_col *cols, int ncols,
                        uint64_t n, uint64_t *count_out, void *ws, size_t ws_bytes, ysql_stream_t stream) {
  if (cmp < YSQL_LT || cmp > YSQL_NE) return YSQL_EINVAL;
  return compact_impl(1, dtype, cmp, nullptr, pred_col, scalar, cols, ncols, n, ~0ull, count_out, ws, ws_bytes,
                      as_stream(stream));
}

int ysql_filter_compact_noalloc(int dtype, int cmp, const void *pred_col, const void *scalar, const ysql_col *cols,
                                int ncols, uint64_t n, uint64_t *count_out, ysql_stream_t stream) {
  if (cmp < YSQL_LT || cmp > YSQL_NE) return YSQL_EINVAL;
  if (n > kNoAllocMaxRows) {
    set_last_error("ysql_filter_compact_noalloc: more rows than the static scratch holds (2^32)");
    return YSQL_EWORKSPACE;
  }
  void *ws = nullptr;
  if (n > 0 && !(ws = noalloc_scratch())) return YSQL_EHIP;
  // the tile counter and status words are cleared by a kernel inside
  // compact_impl (scan_reset) before the compaction reads them
  return compact_impl(1, dtype, cmp, nullptr, pred_col, scalar, cols, ncols, n, ~0ull, count_out, ws,
                      kNoAllocScratchBytes, as_stream(stream));
}

int ysql_filter_compact_range(int dtype, const void *pred_col, const void *bounds, const ysql_col *cols, int ncols,
                              uint64_t n, uint64_t cap, uint64_t *count_out, void *ws, size_t ws_bytes,
                              ysql_stream_t stream) {
  return compact_impl(1, dtype, kCmpRange, nullptr, pred_col, bounds, cols, ncols, n, cap, count_out, ws, ws_bytes,
                      as_stream(stream));
}

int ysql_expr_filter_compact(const ysql_xinsn *prog, int ninsn, int pred_reg, const ysql_xinput *inputs, int ninputs,
                             const ysql_col *cols, int ncols, uint64_t n, uint64_t *count_out, void *ws,
                             size_t ws_bytes, ysql_stream_t stream) {
  XProg p;
  bool wide = false;
  int8_t ty[kXRegs];
  if (pred_reg < 0 || pred_reg >= kXRegs) return YSQL_EINVAL;
  int rc = xprog_prepare(prog, ninsn, inputs, ninputs, n, 0u, pred_reg, &p, &wide, ty);
  if (rc) return rc;
  p.nout = 0;
  if (!cols_ok(cols, ncols, n > 0)) return YSQL_EINVAL;
  if (n >= (1ull << 40)) return YSQL_EINVAL;
  hipStream_t s = as_stream(stream);
  if (n == 0) {
    if (count_out) {
      k_store_zero<<<1, 1, 0, s>>>(count_out);
      YSQL_LAUNCH_CHECK();
    }
    return YSQL_OK;
  }
  ScanWs w;
  if (!ws || ws_bytes < scan_ws_size(n) || !scan_ws_carve(ws, ws_bytes, n, &w)) return YSQL_EWORKSPACE;
  const uint32_t ntiles = (uint32_t)ceil_div(n, kScanTile);
  rc = scan_reset(w, ntiles, s);
  if (rc) return rc;
  ColArgs ca;
  memset(&ca, 0, sizeof(ca));
  for (int i = 0; i < ncols; i++) ca.c[i] = cols[i];
  // a 4-byte input column that is also an output column: staged in LDS
  int stage_in = -1;
  for (int c = 0; c < p.nin && stage_in < 0 && !wide; c++) {
    if (p.in[c].dtype != YSQL_INT && p.in[c].dtype != YSQL_FLOAT) continue;
    for (int i = 0; i < ncols; i++)
      if (cols[i].src == p.in[c].data && cols[i].elem_bytes == 4) stage_in = c;
  }
  const void *stage_src = stage_in >= 0 ? p.in[stage_in].data : nullptr;
  if (wide)
    k_xfilter<uint64_t><<<ntiles, kScanBlock, 0, s>>>(p, ca, ncols, n, count_out, w.ctr, w.status, ntiles, -1,
                                                      nullptr);
  else
    k_xfilter<uint32_t><<<ntiles, kScanBlock, 0, s>>>(p, ca, ncols, n, count_out, w.ctr, w.status, ntiles, stage_in,
                                                      stage_src);
  YSQL_LAUNCH_CHECK();
  return YSQL_OK;
}

}  // extern "C"
