// radix.hip — stable LSD radix sort with 8-bit digits (onesweep form).
//
// Reference: launch_radix_sort (radix_sort_kernel.cu:139-170) runs, per
// 4-bit digit, a block-local split sort of 256 keys (4 one-bit splits of
// Hillis-Steele scans, :8-103), a chained histogram scan and a scatter
// (:106-136): 3 launches and 2 full read+write passes of key+index per digit,
// 8 digits for 32-bit keys, 16 for 64-bit.
//
// Here (gfx950, wave64):
//   1. k_radix_hist: one read of the keys builds the 256-bin histogram of
//      every digit at once (LDS atomics, one global atomic per bin per block).
//   2. One k_onesweep launch per 8-bit digit (radix_dev.hpp): one read and
//      one write of key+payload per digit, tiles placed through decoupled
//      look-back.
#include "radix_dev.hpp"

namespace ysql {

template <typename K>
__global__ __launch_bounds__(256) void k_radix_hist(const K *keys, uint64_t n, int begin_bit, int passes,
                                                    uint32_t desc, uint32_t *hist) {
  __shared__ uint32_t s_h[8 * kBins];
  const uint64_t chunk = ceil_div(ceil_div(n, gridDim.x), 64) * 64;
  const uint64_t begin = (uint64_t)blockIdx.x * chunk;
  const uint64_t end = begin + chunk < n ? begin + chunk : n;
  if (begin >= end) return;  // whole blocks leave together
  hist_block<K>(keys, begin, end, begin_bit, passes, desc, s_h, hist);
}

// Zeroes [p, p + words) of 32-bit words.
__global__ void k_zero_u32(uint32_t *p, uint64_t words) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

// A whole sort of n <= one tile in one workgroup.
template <typename K, int BLOCK, int KPT>
__global__ __launch_bounds__(BLOCK) void k_sort_small(const K *kin, const uint32_t *vin, uint32_t val_base, K *kout,
                                                      uint32_t *vout, uint32_t n, int begin_bit, int passes,
                                                      uint32_t desc) {
  __shared__ OnesweepSmem<K, BLOCK, KPT> sm;
  small_sort_block<K, BLOCK, KPT>(sm, kin, vin, val_base, kout, vout, n, begin_bit, passes, desc);
}

template <typename K>
__global__ void k_copy_pairs(const K *ks, const uint32_t *vs, K *kd, uint32_t *vd, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    kd[i] = ks[i];
    vd[i] = vs[i];
  }
}

// One digit pass; tiles are taken in launch order from *ctr.  The digit
// counts are read and the wave counters cleared while the ticket is taken.
template <typename K, int BLOCK, int KPT>
__global__ __launch_bounds__(BLOCK) void k_onesweep(const K *__restrict__ kin, const uint32_t *__restrict__ vin,
                                                    uint32_t val_base, K *__restrict__ kout,
                                                    uint32_t *__restrict__ vout, uint32_t n, int shift, uint32_t desc,
                                                    const uint32_t *__restrict__ hist, uint32_t *status,
                                                    uint32_t *status_next, uint32_t *ctr) {
  __shared__ OnesweepSmem<K, BLOCK, KPT> sm;
  __shared__ uint32_t s_tile;
  const int t = threadIdx.x;
  const uint32_t h = t < kBins ? hist[t] : 0u;
  for (int i = t; i < (BLOCK / kWave) * kBins; i += BLOCK) (&sm.wh[0][0])[i] = 0;
  if (t == 0) s_tile = atomicAdd(ctr, 1u);
  __syncthreads();
  onesweep_tile<K, BLOCK, KPT>(sm, s_tile, kin, vin, val_base, kout, vout, n, shift, desc, h, status, status_next);
}

// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// Host driver
// ---------------------------------------------------------------------------
// Tile shapes of the digit pass (threads x keys per thread).  Variant 0
// (8192-key tiles) measured best on 2^25
// keys; the smaller tiles give mid-size sorts more workgroups.  A sort uses
// one variant for all its passes (the status areas are laid out per tile).
constexpr int kRadixVariants = 5;
template <int V>
struct RadixTile;
template <>
struct RadixTile<0> {
  static constexpr int BLOCK = 512, KPT = 16;
};
template <>
struct RadixTile<1> {
  static constexpr int BLOCK = 256, KPT = 16;
};
template <>
struct RadixTile<2> {
  static constexpr int BLOCK = 256, KPT = 8;
};
template <>
struct RadixTile<3> {
  static constexpr int BLOCK = 512, KPT = 8;
};
template <>
struct RadixTile<4> {
  static constexpr int BLOCK = 256, KPT = 12;
};
constexpr int kSmallSortBlock = 512, kSmallSortKpt = 16;  // one-workgroup sorts up to 8192 keys
constexpr uint64_t kMinTile = 256 * 8;                     // smallest variant tile: sizes the status areas

// Default variant for an n-key sort (both key widths).
// (measured per size bucket on MI355X, both key widths: 2048-key tiles up to
// 2^19 keys, 8192-key tiles above)
static inline int radix_default_variant(uint64_t n) { return n < (1ull << 19) ? 2 : 0; }

static inline uint64_t variant_tile(int v) {
  switch (v) {
    case 0: return RadixTile<0>::BLOCK * RadixTile<0>::KPT;
    case 1: return RadixTile<1>::BLOCK * RadixTile<1>::KPT;
    case 2: return RadixTile<2>::BLOCK * RadixTile<2>::KPT;
    case 3: return RadixTile<3>::BLOCK * RadixTile<3>::KPT;
    default: return RadixTile<4>::BLOCK * RadixTile<4>::KPT;
  }
}

struct RadixWs {
  void *kalt;
  uint32_t *valt;
  uint32_t *hist;  // 8 * 256
  uint32_t *ctr;   // 8, padded
  uint32_t *st0;   // ntiles * 256 (ntiles of the smallest tile)
  uint32_t *st1;   // ntiles * 256
};

template <typename K>
static size_t radix_ws_layout(uint64_t n, void *ws, size_t bytes, RadixWs *out) {
  const uint64_t ntiles = ceil_div(n, kMinTile);
  if (out) {
    WsCarver c(ws, bytes);
    out->kalt = c.take<K>(n);
    out->valt = c.take<uint32_t>(n);
    out->hist = c.take<uint32_t>(8 * kBins + 64);  // hist followed by the pass counters
    out->ctr = out->hist ? out->hist + 8 * kBins : nullptr;
    out->st0 = c.take<uint32_t>(ntiles * kBins);
    out->st1 = c.take<uint32_t>(ntiles * kBins);
    return (out->kalt && out->valt && out->hist && out->st0 && out->st1) ? c.off : 0;
  }
  WsCounter c;
  c.take<K>(n);
  c.take<uint32_t>(n);
  c.take<uint32_t>(8 * kBins + 64);
  c.take<uint32_t>(ntiles * kBins);
  c.take<uint32_t>(ntiles * kBins);
  return c.off;
}

template <typename K>
static int radix_check(const void *keys_in, uint64_t n, int begin_bit, int end_bit) {
  constexpr int KBITS = 8 * sizeof(K);
  if (!keys_in) return YSQL_EINVAL;
  if (n >= (1ull << 30)) return YSQL_EINVAL;
  if (begin_bit < 0 || end_bit > KBITS || begin_bit >= end_bit) return YSQL_EINVAL;
  return YSQL_OK;
}

static inline int radix_passes(int begin_bit, int end_bit) { return (end_bit - begin_bit + 7) / 8; }

// Zeroes the histograms, tile counters and first status area, then builds
// the histograms of every digit from keys_in.
template <typename K>
static int radix_begin_t(int desc, const K *keys_in, uint64_t n, int begin_bit, int end_bit, int variant,
                         const RadixWs &w, hipStream_t s) {
  const uint32_t ntiles = (uint32_t)ceil_div(n, variant_tile(variant));
  // the histograms, pass counters and first status area are adjacent in
  // the workspace: one launch clears them
  const uint64_t zwords = (uint64_t)(w.st0 - w.hist) + (uint64_t)ntiles * kBins;
  uint64_t zg = ceil_div(zwords, 256 * 8);
  if (zg > 2048) zg = 2048;
  if (zg < 1) zg = 1;
  k_zero_u32<<<(unsigned)zg, 256, 0, s>>>(w.hist, zwords);
  YSQL_LAUNCH_CHECK();
  uint64_t g = ceil_div(n, 256 * 16);
  const uint64_t cap = (uint64_t)grid_cap() * 4;
  if (g > cap) g = cap;
  if (g < 1) g = 1;
  k_radix_hist<K><<<(unsigned)g, 256, 0, s>>>(keys_in, n, begin_bit, radix_passes(begin_bit, end_bit),
                                              desc ? 0xFFu : 0u, w.hist);
  YSQL_LAUNCH_CHECK();
  return YSQL_OK;
}

// One digit pass p of `passes`: (kin, vin or val_base + row) -> (kout, vout).
template <typename K, int V>
static void radix_pass_v(int desc, int p, int passes, const K *kin, const uint32_t *vin, uint32_t val_base, K *kout,
                         uint32_t *vout, uint64_t n, int begin_bit, const RadixWs &w, hipStream_t s) {
  constexpr int BLOCK = RadixTile<V>::BLOCK, KPT = RadixTile<V>::KPT, TILE = BLOCK * KPT;
  const uint32_t ntiles = (uint32_t)ceil_div(n, TILE);
  uint32_t *st = (p & 1) ? w.st1 : w.st0;
  uint32_t *snext = (p + 1 < passes) ? ((p & 1) ? w.st0 : w.st1) : nullptr;
  k_onesweep<K, BLOCK, KPT><<<ntiles, BLOCK, 0, s>>>(kin, vin, val_base, kout, vout, (uint32_t)n,
                                                            begin_bit + 8 * p, desc ? 0xFFu : 0u, w.hist + p * kBins,
                                                            st, snext, w.ctr + p);
}

template <typename K>
static int radix_pass_t(int desc, int p, int passes, const K *kin, const uint32_t *vin, uint32_t val_base, K *kout,
                        uint32_t *vout, uint64_t n, int begin_bit, int variant, const RadixWs &w, hipStream_t s) {
  switch (variant) {
    case 0: radix_pass_v<K, 0>(desc, p, passes, kin, vin, val_base, kout, vout, n, begin_bit, w, s); break;
    case 1: radix_pass_v<K, 1>(desc, p, passes, kin, vin, val_base, kout, vout, n, begin_bit, w, s); break;
    case 2: radix_pass_v<K, 2>(desc, p, passes, kin, vin, val_base, kout, vout, n, begin_bit, w, s); break;
    case 3: radix_pass_v<K, 3>(desc, p, passes, kin, vin, val_base, kout, vout, n, begin_bit, w, s); break;
    case 4: radix_pass_v<K, 4>(desc, p, passes, kin, vin, val_base, kout, vout, n, begin_bit, w, s); break;
    default: return YSQL_EINVAL;
  }
  YSQL_LAUNCH_CHECK();
  return YSQL_OK;
}

// Full sort keys_in/vals_in -> keys_out/vals_out.  keys_in may equal
// keys_out (in place); vals_in == nullptr gives payload val_base + row.
template <typename K>
static int radix_sort_t(int desc, const K *kin, const uint32_t *vin, uint32_t vbase, K *kout, uint32_t *vout,
                        uint64_t n, int begin_bit, int end_bit, int variant, void *ws, size_t ws_bytes,
                        hipStream_t s) {
  if (n == 0) return YSQL_OK;
  int rc = radix_check<K>(kin, n, begin_bit, end_bit);
  if (rc) return rc;
  if (!kout || !vout) return YSQL_EINVAL;
  if (variant < 0 || variant >= kRadixVariants) return YSQL_EINVAL;
  RadixWs w;
  if (!ws || ws_bytes < radix_ws_layout<K>(n, nullptr, 0, nullptr) || !radix_ws_layout<K>(n, ws, ws_bytes, &w))
    return YSQL_EWORKSPACE;
  const int P = radix_passes(begin_bit, end_bit);
  if (n <= (uint64_t)kSmallSortBlock * kSmallSortKpt) {  // one tile: the whole sort in one workgroup
    k_sort_small<K, kSmallSortBlock, kSmallSortKpt><<<1, kSmallSortBlock, 0, s>>>(kin, vin, vbase, kout, vout,
                                                                                (uint32_t)n, begin_bit, P,
                                                                                desc ? 0xFFu : 0u);
    YSQL_LAUNCH_CHECK();
    return YSQL_OK;
  }
  rc = radix_begin_t<K>(desc, kin, n, begin_bit, end_bit, variant, w, s);
  if (rc) return rc;
  K *kalt = (K *)w.kalt;
  uint32_t *valt = w.valt;
  // With an even pass count, or distinct in/out buffers, the passes can
  // alternate so that the last one lands in the output.  An odd count in
  // place runs through the workspace and copies back at the end.
  const bool aliased = ((const void *)kin == (const void *)kout) || (vin && vin == vout);
  const bool copy_back = (P & 1) && aliased;
  const K *ksrc = kin;
  const uint32_t *vsrc = vin;
  for (int p = 0; p < P; p++) {
    const bool to_out = copy_back ? (p & 1) : (((P - 1 - p) & 1) == 0);
    K *kd = to_out ? kout : kalt;
    uint32_t *vd = to_out ? vout : valt;
    rc = radix_pass_t<K>(desc, p, P, ksrc, vsrc, vbase, kd, vd, n, begin_bit, variant, w, s);
    if (rc) return rc;
    ksrc = kd;
    vsrc = vd;
  }
  if (copy_back) {
    uint64_t g = ceil_div(n, 256 * 4);
    const uint64_t cap = (uint64_t)grid_cap() * 4;
    if (g > cap) g = cap;
    k_copy_pairs<K><<<(unsigned)g, 256, 0, s>>>(kalt, valt, kout, vout, n);
    YSQL_LAUNCH_CHECK();
  }
  return YSQL_OK;
}

}  // namespace ysql

using namespace ysql;

extern "C" {

int ysql_radix_sort_ws_bytes(int key_dtype, uint64_t n, size_t *bytes) {
  if (!bytes) return YSQL_EINVAL;
  if (key_dtype == YSQL_UKEY32) *bytes = radix_ws_layout<uint32_t>(n, nullptr, 0, nullptr);
  else if (key_dtype == YSQL_UKEY64) *bytes = radix_ws_layout<uint64_t>(n, nullptr, 0, nullptr);
  else return YSQL_EINVAL;
  return YSQL_OK;
}

int ysql_radix_sort_variant(int key_dtype, int variant, int descending, const void *keys_in, const uint32_t *vals_in,
                            uint32_t val_base, void *keys_out, uint32_t *vals_out, uint64_t n, int begin_bit,
                            int end_bit, void *ws, size_t ws_bytes, ysql_stream_t stream) {
  hipStream_t s = as_stream(stream);
  if (variant < 0) variant = radix_default_variant(n);
  if (key_dtype == YSQL_UKEY32)
    return radix_sort_t<uint32_t>(descending, (const uint32_t *)keys_in, vals_in, val_base, (uint32_t *)keys_out,
                                  vals_out, n, begin_bit, end_bit, variant, ws, ws_bytes, s);
  if (key_dtype == YSQL_UKEY64)
    return radix_sort_t<uint64_t>(descending, (const uint64_t *)keys_in, vals_in, val_base, (uint64_t *)keys_out,
                                  vals_out, n, begin_bit, end_bit, variant, ws, ws_bytes, s);
  return YSQL_EINVAL;
}

int ysql_radix_variant_count(void) { return kRadixVariants; }

int ysql_radix_sort_ex(int key_dtype, int descending, const void *keys_in, const uint32_t *vals_in, uint32_t val_base,
                       void *keys_out, uint32_t *vals_out, uint64_t n, int begin_bit, int end_bit, void *ws,
                       size_t ws_bytes, ysql_stream_t stream) {
  return ysql_radix_sort_variant(key_dtype, -1, descending, keys_in, vals_in, val_base, keys_out, vals_out, n,
                                 begin_bit, end_bit, ws, ws_bytes, stream);
}

int ysql_radix_argsort(int key_dtype, int descending, void *keys_inout, uint32_t *perm_out, uint64_t n, int begin_bit,
                       int end_bit, void *ws, size_t ws_bytes, ysql_stream_t stream) {
  if (n > 0 && !perm_out) return YSQL_EINVAL;
  return ysql_radix_sort_ex(key_dtype, descending, keys_inout, nullptr, 0, keys_inout, perm_out, n, begin_bit, end_bit,
                            ws, ws_bytes, stream);
}

int ysql_radix_sort_pairs(int key_dtype, int descending, void *keys_inout, uint32_t *vals_inout, uint64_t n,
                          int begin_bit, int end_bit, void *ws, size_t ws_bytes, ysql_stream_t stream) {
  if (n > 0 && !vals_inout) return YSQL_EINVAL;
  return ysql_radix_sort_ex(key_dtype, descending, keys_inout, vals_inout, 0, keys_inout, vals_inout, n, begin_bit,
                            end_bit, ws, ws_bytes, stream);
}

int ysql_radix_begin(int key_dtype, int descending, const void *keys_in, uint64_t n, int begin_bit, int end_bit,
                     void *ws, size_t ws_bytes, ysql_stream_t stream) {
  if (n == 0) return YSQL_OK;
  hipStream_t s = as_stream(stream);
  RadixWs w;
  if (key_dtype == YSQL_UKEY32) {
    int rc = radix_check<uint32_t>(keys_in, n, begin_bit, end_bit);
    if (rc) return rc;
    if (!ws || ws_bytes < radix_ws_layout<uint32_t>(n, nullptr, 0, nullptr) ||
        !radix_ws_layout<uint32_t>(n, ws, ws_bytes, &w))
      return YSQL_EWORKSPACE;
    return radix_begin_t<uint32_t>(descending, (const uint32_t *)keys_in, n, begin_bit, end_bit,
                                   radix_default_variant(n), w, s);
  }
  if (key_dtype == YSQL_UKEY64) {
    int rc = radix_check<uint64_t>(keys_in, n, begin_bit, end_bit);
    if (rc) return rc;
    if (!ws || ws_bytes < radix_ws_layout<uint64_t>(n, nullptr, 0, nullptr) ||
        !radix_ws_layout<uint64_t>(n, ws, ws_bytes, &w))
      return YSQL_EWORKSPACE;
    return radix_begin_t<uint64_t>(descending, (const uint64_t *)keys_in, n, begin_bit, end_bit,
                                   radix_default_variant(n), w, s);
  }
  return YSQL_EINVAL;
}

int ysql_radix_pass(int key_dtype, int descending, int pass, const void *keys_in, const uint32_t *vals_in,
                    uint32_t val_base, void *keys_out, uint32_t *vals_out, uint64_t n, int begin_bit, int end_bit,
                    void *ws, size_t ws_bytes, ysql_stream_t stream) {
  if (n == 0) return YSQL_OK;
  hipStream_t s = as_stream(stream);
  // each pass scatters into buffers distinct from what other tiles still read
  if (!keys_out || !vals_out || keys_in == keys_out || (vals_in && vals_in == vals_out)) return YSQL_EINVAL;
  const int P = radix_passes(begin_bit, end_bit);
  if (pass < 0 || pass >= P) return YSQL_EINVAL;
  RadixWs w;
  if (key_dtype == YSQL_UKEY32) {
    int rc = radix_check<uint32_t>(keys_in, n, begin_bit, end_bit);
    if (rc) return rc;
    if (!ws || ws_bytes < radix_ws_layout<uint32_t>(n, nullptr, 0, nullptr) ||
        !radix_ws_layout<uint32_t>(n, ws, ws_bytes, &w))
      return YSQL_EWORKSPACE;
    return radix_pass_t<uint32_t>(descending, pass, P, (const uint32_t *)keys_in, vals_in, val_base,
                                  (uint32_t *)keys_out, vals_out, n, begin_bit, radix_default_variant(n), w, s);
  }
  if (key_dtype == YSQL_UKEY64) {
    int rc = radix_check<uint64_t>(keys_in, n, begin_bit, end_bit);
    if (rc) return rc;
    if (!ws || ws_bytes < radix_ws_layout<uint64_t>(n, nullptr, 0, nullptr) ||
        !radix_ws_layout<uint64_t>(n, ws, ws_bytes, &w))
      return YSQL_EWORKSPACE;
    return radix_pass_t<uint64_t>(descending, pass, P, (const uint64_t *)keys_in, vals_in, val_base,
                                  (uint64_t *)keys_out, vals_out, n, begin_bit, radix_default_variant(n), w, s);
  }
  return YSQL_EINVAL;
}

}  // extern "C"
