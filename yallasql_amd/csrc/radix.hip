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
//   2. One k_onesweep launch per 8-bit digit: a workgroup takes a tile of
//      keys, ranks each key inside its wave with 8 ballots (lanes with the same
//      digit are found by matching the 8 digit bits; mbcnt gives the rank), sums
//      per-wave digit counts into tile counts, publishes them for decoupled
//      look-back (one thread per digit walks back over predecessor tiles),
//      reorders the tile in LDS by digit and writes each digit's run to its
//      global position.  One read and one write of key+payload per digit.
// Stability: keys are ranked in (wave, item, lane) order, which is the
// input order inside the tile, and tiles are placed in tile order.
#include "common.hpp"

namespace ysql {

constexpr int kBins = 256;
constexpr uint32_t kRFlagAgg = 1u << 30;
constexpr uint32_t kRFlagInc = 2u << 30;
constexpr uint32_t kRValMask = (1u << 30) - 1;

__device__ __forceinline__ uint32_t rstatus_load(const uint32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void rstatus_store(uint32_t *p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------------------
// Histograms of all digit positions in one pass.
// ---------------------------------------------------------------------------
// Keys narrower than their type (small integers, Zipf ranks) put most of
// their high digits in one bin; 64 lanes adding to one LDS word serialise.
// Each thread therefore counts the bin of digit value `desc` (the encoded
// zero digit) in registers and only adds the other bins to LDS.
template <typename K>
__device__ __forceinline__ void hist_key(K k, int begin_bit, int passes, uint32_t desc, uint32_t *s_h,
                                         uint32_t (&zc)[8]) {
#pragma unroll
  for (int p = 0; p < 8; p++) {
    if (p < passes) {
      const uint32_t d = ((uint32_t)(k >> (begin_bit + 8 * p)) & 0xFFu) ^ desc;
      if (d == desc) zc[p]++;
      else atomicAdd(&s_h[p * kBins + d], 1u);
    }
  }
}

template <typename K>
__global__ __launch_bounds__(256) void k_radix_hist(const K *keys, uint64_t n, int begin_bit, int passes,
                                                    uint32_t desc, uint32_t *hist) {
  __shared__ uint32_t s_h[8 * kBins];
  for (int i = threadIdx.x; i < passes * kBins; i += blockDim.x) s_h[i] = 0;
  __syncthreads();
  uint32_t zc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  constexpr int PER = 16 / sizeof(K);  // keys per 16-byte load
  const uint64_t nv = ((uintptr_t)keys & 15) ? 0 : n / PER;
  const uint4 *kv = reinterpret_cast<const uint4 *>(keys);
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nv; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint4 q = kv[i];
    K ks[PER];
    memcpy(ks, &q, 16);
#pragma unroll
    for (int e = 0; e < PER; e++) hist_key<K>(ks[e], begin_bit, passes, desc, s_h, zc);
  }
  for (uint64_t i = nv * PER + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    hist_key<K>(keys[i], begin_bit, passes, desc, s_h, zc);
#pragma unroll
  for (int p = 0; p < 8; p++) {
    if (p < passes) {
      const uint32_t z = wave_sum(zc[p]);
      if (lane_id() == 0 && z) atomicAdd(&s_h[p * kBins + desc], z);
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < passes * kBins; i += blockDim.x) {
    uint32_t v = s_h[i];
    if (v) atomicAdd(&hist[i], v);
  }
}

// Zeroes [p, p + words) of 32-bit words.
__global__ void k_zero_u32(uint32_t *p, uint64_t words) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

template <typename K>
__global__ void k_copy_pairs(const K *ks, const uint32_t *vs, K *kd, uint32_t *vd, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    kd[i] = ks[i];
    vd[i] = vs[i];
  }
}

// ---------------------------------------------------------------------------
// One digit pass.
//   kin/vin -> kout/vout; vin == nullptr means the payload is val_base + row.
//   hist: the 256 counts of this digit over all keys.
//   status: ntiles * 256 words of this pass, zeroed; status_next: the same
//   area for the following pass, zeroed here tile by tile (it was last used
//   two passes ago, whose kernel has finished).
// ---------------------------------------------------------------------------
template <typename K, int BLOCK, int KPT, bool EARLY>
__global__ __launch_bounds__(BLOCK) void k_onesweep(const K *__restrict__ kin, const uint32_t *__restrict__ vin,
                                                    uint32_t val_base, K *__restrict__ kout,
                                                    uint32_t *__restrict__ vout, uint32_t n, int shift, uint32_t desc,
                                                    const uint32_t *__restrict__ hist,
                                                    uint32_t *status, uint32_t *status_next, uint32_t *ctr) {
  static_assert(BLOCK >= kBins, "one thread per digit");
  constexpr int W = BLOCK / kWave;
  constexpr int TILE = BLOCK * KPT;
  __shared__ K s_k[TILE];
  __shared__ uint32_t s_wh[W][kBins];
  __shared__ uint32_t s_dstart[kBins];
  __shared__ uint32_t s_glob[kBins];
  __shared__ uint32_t s_red[W + 1];
  __shared__ uint32_t s_tile;

  const int t = threadIdx.x;
  const int lane = lane_id();
  const int wave = t / kWave;
  if (t == 0) s_tile = atomicAdd(ctr, 1u);
  for (int i = t; i < W * kBins; i += BLOCK) (&s_wh[0][0])[i] = 0;
  __syncthreads();
  const uint32_t tile = s_tile;
  const uint32_t base = tile * (uint32_t)TILE;
  if (status_next && t < kBins) status_next[tile * kBins + t] = 0;

  // ---- load: wave w owns rows [base + w*KPT*64, ...), item i lane l -> row
  K key[KPT];
  uint32_t val[KPT];
  const uint32_t wbase = base + (uint32_t)wave * KPT * kWave;
#pragma unroll
  for (int i = 0; i < KPT; i++) {
    const uint32_t idx = wbase + i * kWave + lane;
    const bool valid = idx < n;
    key[i] = valid ? kin[idx] : (K)0;
    val[i] = vin ? (valid ? vin[idx] : 0u) : val_base + idx;
  }

  // ---- rank inside the wave, in (item, lane) order
  uint32_t rank[KPT];
#pragma unroll
  for (int i = 0; i < KPT; i++) {
    const uint32_t idx = wbase + i * kWave + lane;
    const bool valid = idx < n;
    const uint32_t d = ((uint32_t)(key[i] >> shift) & 0xFFu) ^ desc;
    uint64_t m = ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; b++) {
      const bool bit = (d >> b) & 1u;
      const uint64_t bal = ballot(bit);
      m &= bit ? bal : ~bal;
    }
    const uint32_t lower = popc_below(m);
    uint32_t pre = 0;
    if (valid) pre = s_wh[wave][d];
    rank[i] = pre + lower;
    if (valid && lower == 0) s_wh[wave][d] = pre + (uint32_t)__popcll(m);
  }
  __syncthreads();

  // ---- per-digit tile counts, wave offsets, publication
  uint32_t cnt = 0;
  if (t < kBins) {
    uint32_t run = 0;
#pragma unroll
    for (int w = 0; w < W; w++) {
      uint32_t c = s_wh[w][t];
      s_wh[w][t] = run;
      run += c;
    }
    cnt = run;
    rstatus_store(&status[tile * kBins + t], (tile == 0 ? kRFlagInc : kRFlagAgg) | cnt);
  }
  // Decoupled look-back, one thread per digit.  EARLY walks back right
  // after publishing the aggregate, so this tile's inclusive prefix (which
  // later tiles wait for) appears as soon as possible.
  uint32_t excl = 0;
  auto lookback = [&]() {
    if (t < kBins && tile > 0) {
      int64_t k = (int64_t)tile - 1;
      while (true) {
        const uint32_t s = rstatus_load(&status[(uint32_t)k * kBins + t]);
        const uint32_t f = s & ~kRValMask;
        if (f == 0) continue;
        excl += s & kRValMask;
        if (f == kRFlagInc) break;
        k--;
      }
      rstatus_store(&status[tile * kBins + t], kRFlagInc | (excl + cnt));
    }
  };
  if constexpr (EARLY) lookback();
  uint32_t tot;
  const uint32_t dstart = block_exclusive_sum<BLOCK, uint32_t>(t < kBins ? cnt : 0u, s_red, &tot);
  const uint32_t h = t < kBins ? hist[t] : 0u;
  const uint32_t gstart = block_exclusive_sum<BLOCK, uint32_t>(h, s_red, &tot);
  if (t < kBins) s_dstart[t] = dstart;
  __syncthreads();

  // ---- reorder the tile's keys in LDS by digit (payloads follow later
  // through the same LDS area, which keeps the block at 36 KB of LDS)
  uint32_t pos[KPT];
#pragma unroll
  for (int i = 0; i < KPT; i++) {
    const uint32_t idx = wbase + i * kWave + lane;
    pos[i] = 0xFFFFFFFFu;
    if (idx < n) {
      const uint32_t d = ((uint32_t)(key[i] >> shift) & 0xFFu) ^ desc;
      pos[i] = s_dstart[d] + s_wh[wave][d] + rank[i];
      s_k[pos[i]] = key[i];
    }
  }

  if constexpr (!EARLY) lookback();
  if (t < kBins) s_glob[t] = gstart + excl - dstart;
  __syncthreads();

  // ---- write each digit's run of keys to its global position
  const uint32_t valid_cnt = (n - base) < (uint32_t)TILE ? (n - base) : (uint32_t)TILE;
  uint32_t opos[KPT];
#pragma unroll
  for (int q = 0; q < KPT; q++) {
    const uint32_t j = t + q * BLOCK;
    opos[q] = 0;
    if (j < valid_cnt) {
      const K kk = s_k[j];
      const uint32_t d = ((uint32_t)(kk >> shift) & 0xFFu) ^ desc;
      opos[q] = s_glob[d] + j;
      kout[opos[q]] = kk;
    }
  }
  __syncthreads();
  // ---- then the payloads, through the same LDS area
  uint32_t *s_v = reinterpret_cast<uint32_t *>(s_k);
#pragma unroll
  for (int i = 0; i < KPT; i++)
    if (pos[i] != 0xFFFFFFFFu) s_v[pos[i]] = val[i];
  __syncthreads();
#pragma unroll
  for (int q = 0; q < KPT; q++) {
    const uint32_t j = t + q * BLOCK;
    if (j < valid_cnt) vout[opos[q]] = s_v[j];
  }
}

// ---------------------------------------------------------------------------
// Host driver
// ---------------------------------------------------------------------------
template <typename K>
struct RadixTile;
template <>
struct RadixTile<uint32_t> {
  static constexpr int BLOCK = 512, KPT = 16;
  static constexpr bool EARLY = false;
};
template <>
struct RadixTile<uint64_t> {
  // 8192-key tiles, look-back after the LDS reorder: measured best of
  // 256..1024 threads x 8..32 keys on 2^25 keys (216 us per pass on MI355X).
  static constexpr int BLOCK = 512, KPT = 16;
  static constexpr bool EARLY = false;
};

struct RadixWs {
  void *kalt;
  uint32_t *valt;
  uint32_t *hist;  // 8 * 256
  uint32_t *ctr;   // 8, padded
  uint32_t *st0;   // ntiles * 256
  uint32_t *st1;   // ntiles * 256
};

template <typename K>
static size_t radix_ws_layout(uint64_t n, void *ws, size_t bytes, RadixWs *out) {
  constexpr int TILE = RadixTile<K>::BLOCK * RadixTile<K>::KPT;
  const uint64_t ntiles = ceil_div(n, TILE);
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
static int radix_begin_t(int desc, const K *keys_in, uint64_t n, int begin_bit, int end_bit, const RadixWs &w,
                         hipStream_t s) {
  constexpr int TILE = RadixTile<K>::BLOCK * RadixTile<K>::KPT;
  const uint32_t ntiles = (uint32_t)ceil_div(n, TILE);
  k_zero_u32<<<64, 256, 0, s>>>(w.hist, 8 * kBins + 64);
  YSQL_LAUNCH_CHECK();
  uint64_t zg = ceil_div((uint64_t)ntiles * kBins, 256 * 8);
  if (zg > 2048) zg = 2048;
  if (zg < 1) zg = 1;
  k_zero_u32<<<(unsigned)zg, 256, 0, s>>>(w.st0, (uint64_t)ntiles * kBins);
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
template <typename K>
static int radix_pass_t(int desc, int p, int passes, const K *kin, const uint32_t *vin, uint32_t val_base, K *kout,
                        uint32_t *vout, uint64_t n, int begin_bit, const RadixWs &w, hipStream_t s) {
  constexpr int BLOCK = RadixTile<K>::BLOCK, KPT = RadixTile<K>::KPT, TILE = BLOCK * KPT;
  const uint32_t ntiles = (uint32_t)ceil_div(n, TILE);
  uint32_t *st = (p & 1) ? w.st1 : w.st0;
  uint32_t *snext = (p + 1 < passes) ? ((p & 1) ? w.st0 : w.st1) : nullptr;
  k_onesweep<K, BLOCK, KPT, RadixTile<K>::EARLY><<<ntiles, BLOCK, 0, s>>>(kin, vin, val_base, kout, vout, (uint32_t)n, begin_bit + 8 * p,
                                                     desc ? 0xFFu : 0u, w.hist + p * kBins, st, snext, w.ctr + p);
  YSQL_LAUNCH_CHECK();
  return YSQL_OK;
}

// Full sort keys_in/vals_in -> keys_out/vals_out.  keys_in may equal
// keys_out (in place); vals_in == nullptr gives payload val_base + row.
template <typename K>
static int radix_sort_t(int desc, const K *kin, const uint32_t *vin, uint32_t vbase, K *kout, uint32_t *vout,
                        uint64_t n, int begin_bit, int end_bit, void *ws, size_t ws_bytes, hipStream_t s) {
  if (n == 0) return YSQL_OK;
  int rc = radix_check<K>(kin, n, begin_bit, end_bit);
  if (rc) return rc;
  if (!kout || !vout) return YSQL_EINVAL;
  RadixWs w;
  if (!ws || ws_bytes < radix_ws_layout<K>(n, nullptr, 0, nullptr) || !radix_ws_layout<K>(n, ws, ws_bytes, &w))
    return YSQL_EWORKSPACE;
  const int P = radix_passes(begin_bit, end_bit);
  rc = radix_begin_t<K>(desc, kin, n, begin_bit, end_bit, w, s);
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
    rc = radix_pass_t<K>(desc, p, P, ksrc, vsrc, vbase, kd, vd, n, begin_bit, w, s);
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

int ysql_radix_sort_ex(int key_dtype, int descending, const void *keys_in, const uint32_t *vals_in, uint32_t val_base,
                       void *keys_out, uint32_t *vals_out, uint64_t n, int begin_bit, int end_bit, void *ws,
                       size_t ws_bytes, ysql_stream_t stream) {
  hipStream_t s = as_stream(stream);
  if (key_dtype == YSQL_UKEY32)
    return radix_sort_t<uint32_t>(descending, (const uint32_t *)keys_in, vals_in, val_base, (uint32_t *)keys_out,
                                  vals_out, n, begin_bit, end_bit, ws, ws_bytes, s);
  if (key_dtype == YSQL_UKEY64)
    return radix_sort_t<uint64_t>(descending, (const uint64_t *)keys_in, vals_in, val_base, (uint64_t *)keys_out,
                                  vals_out, n, begin_bit, end_bit, ws, ws_bytes, s);
  return YSQL_EINVAL;
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
    return radix_begin_t<uint32_t>(descending, (const uint32_t *)keys_in, n, begin_bit, end_bit, w, s);
  }
  if (key_dtype == YSQL_UKEY64) {
    int rc = radix_check<uint64_t>(keys_in, n, begin_bit, end_bit);
    if (rc) return rc;
    if (!ws || ws_bytes < radix_ws_layout<uint64_t>(n, nullptr, 0, nullptr) ||
        !radix_ws_layout<uint64_t>(n, ws, ws_bytes, &w))
      return YSQL_EWORKSPACE;
    return radix_begin_t<uint64_t>(descending, (const uint64_t *)keys_in, n, begin_bit, end_bit, w, s);
  }
  return YSQL_EINVAL;
}

int ysql_radix_pass(int key_dtype, int descending, int pass, const void *keys_in, const uint32_t *vals_in,
                    uint32_t val_base, void *keys_out, uint32_t *vals_out, uint64_t n, int begin_bit, int end_bit,
                    void *ws, size_t ws_bytes, ysql_stream_t stream) {
  if (n == 0) return YSQL_OK;
  hipStream_t s = as_stream(stream);
  if (!keys_out || !vals_out || keys_in == keys_out) return YSQL_EINVAL;
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
                                  (uint32_t *)keys_out, vals_out, n, begin_bit, w, s);
  }
  if (key_dtype == YSQL_UKEY64) {
    int rc = radix_check<uint64_t>(keys_in, n, begin_bit, end_bit);
    if (rc) return rc;
    if (!ws || ws_bytes < radix_ws_layout<uint64_t>(n, nullptr, 0, nullptr) ||
        !radix_ws_layout<uint64_t>(n, ws, ws_bytes, &w))
      return YSQL_EWORKSPACE;
    return radix_pass_t<uint64_t>(descending, pass, P, (const uint64_t *)keys_in, vals_in, val_base,
                                  (uint64_t *)keys_out, vals_out, n, begin_bit, w, s);
  }
  return YSQL_EINVAL;
}

}  // extern "C"
