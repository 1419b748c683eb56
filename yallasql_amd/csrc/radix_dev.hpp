// radix_dev.hpp — device-side building blocks of the 8-bit LSD radix sort
// (used by the kernels of radix.hip).
//
//   hist_block:    one workgroup adds the 256-bin histograms of every digit
//                  position of keys[begin, end) into a global histogram.
//   onesweep_tile: one workgroup performs one digit pass over one tile:
//                  wave-level ranking with 8 ballots (lanes with the same
//                  digit match on all 8 digit bits; mbcnt gives the rank), per
//                  digit tile counts published for decoupled look-back (one
//                  thread per digit walks back over predecessor tiles), the
//                  tile reordered by digit in LDS, then each digit's run of
//                  keys and payloads written to its global position.
//   small_sort_block: a whole sort of one tile's worth of keys in one
//                  workgroup (every pass through LDS).
// Stability: keys are ranked in (wave, item, lane) order, which is the input
// order inside the tile, and tiles are placed in tile order.
#pragma once

#include "common.hpp"

namespace ysql {

constexpr int kBins = 256;
constexpr uint32_t kRFlagAgg = 1u << 30;
constexpr uint32_t kRFlagInc = 2u << 30;
constexpr uint32_t kRValMask = (1u << 30) - 1;
constexpr int kLookbackBatch = 8;  // predecessor tiles read per look-back step

__device__ __forceinline__ uint32_t rstatus_load(const uint32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void rstatus_store(uint32_t *p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Lanes of the wave whose 8-bit digit equals this lane's, among the lanes
// with `valid` set: the AND over the digit's 8 bits of the ballot of lanes
// sharing that bit.  Each bit's lane value is a sign-extended 0 / ~0, so a
// half of the mask is and-not(ballot ^ s); the compiler folds the eight
// steps into or3 / bitop3 chains (about 45 VALU per key instead of ~80 for
// a select per bit).
__device__ __forceinline__ uint64_t match_digit8(uint32_t d, bool valid) {
  const uint64_t v = __builtin_amdgcn_ballot_w64(valid);
  uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
#pragma unroll
  for (int b = 0; b < 8; b++) {
    const uint32_t s = (uint32_t)(((int32_t)(d << (31 - b))) >> 31);
    const uint64_t bal = __builtin_amdgcn_ballot_w64(s != 0u);
    lo &= ~((uint32_t)bal ^ s);
    hi &= ~((uint32_t)(bal >> 32) ^ s);
  }
  return ((uint64_t)hi << 32) | lo;
}

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

// All threads of the block: the histograms of `passes` digit positions of
// keys[begin, end) added into s_h (passes * 256 words of LDS, zeroed by the
// caller).  Ends with a barrier.
template <typename K>
__device__ __forceinline__ void hist_range_lds(const K *keys, uint64_t begin, uint64_t end, int begin_bit, int passes,
                                               uint32_t desc, uint32_t *s_h) {
  uint32_t zc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  constexpr int PER = 16 / sizeof(K);  // keys per 16-byte load
  uint64_t i = begin;
  // head until 16-byte alignment, then vector loads, then the tail
  while (i < end && ((uintptr_t)(keys + i) & 15)) {
    if (i - begin == threadIdx.x) hist_key<K>(keys[i], begin_bit, passes, desc, s_h, zc);
    i++;
  }
  const uint64_t vstart = i;
  const uint64_t nv = (end - vstart) / PER;
  const uint4 *kv = reinterpret_cast<const uint4 *>(keys + vstart);
  // kHistUnroll loads of a thread are issued before any key is counted, so
  // that enough bytes are in flight per wave
  constexpr int kHistUnroll = 4;
  uint64_t v = threadIdx.x;
  for (; v + (kHistUnroll - 1) * (uint64_t)blockDim.x < nv; v += kHistUnroll * (uint64_t)blockDim.x) {
    uint4 q[kHistUnroll];
#pragma unroll
    for (int u = 0; u < kHistUnroll; u++) q[u] = kv[v + u * (uint64_t)blockDim.x];
#pragma unroll
    for (int u = 0; u < kHistUnroll; u++) {
      K ks[PER];
      memcpy(ks, &q[u], 16);
#pragma unroll
      for (int e = 0; e < PER; e++) hist_key<K>(ks[e], begin_bit, passes, desc, s_h, zc);
    }
  }
  for (; v < nv; v += blockDim.x) {
    const uint4 q = kv[v];
    K ks[PER];
    memcpy(ks, &q, 16);
#pragma unroll
    for (int e = 0; e < PER; e++) hist_key<K>(ks[e], begin_bit, passes, desc, s_h, zc);
  }
  for (uint64_t r = vstart + nv * PER + threadIdx.x; r < end; r += blockDim.x)
    hist_key<K>(keys[r], begin_bit, passes, desc, s_h, zc);
#pragma unroll
  for (int p = 0; p < 8; p++) {
    if (p < passes) {
      const uint32_t z = wave_sum(zc[p]);
      if (lane_id() == 0 && z) atomicAdd(&s_h[p * kBins + desc], z);
    }
  }
  __syncthreads();
}

// All threads of the block: histograms of keys[begin, end) (s_h: 8 * 256
// words of LDS) added into hist (passes * 256 words).  Contains barriers.
template <typename K>
__device__ __forceinline__ void hist_block(const K *keys, uint64_t begin, uint64_t end, int begin_bit, int passes,
                                           uint32_t desc, uint32_t *s_h, uint32_t *hist) {
  for (int i = threadIdx.x; i < passes * kBins; i += blockDim.x) s_h[i] = 0;
  __syncthreads();
  hist_range_lds<K>(keys, begin, end, begin_bit, passes, desc, s_h);
  for (int b = threadIdx.x; b < passes * kBins; b += blockDim.x) {
    const uint32_t v = s_h[b];
    if (v) atomicAdd(&hist[b], v);
  }
  __syncthreads();
}

// A tile's keys and payloads are reordered through one 32-bit LDS stage:
// 64-bit keys go through it as two 32-bit halves (low, then high), so the
// stage is TILE words for every key width.  Half the LDS of a 64-bit stage
// lets more workgroups share a CU, and 32-bit scattered LDS stores conflict
// less than 64-bit ones: on MI355X a 2,048-key tile pass over 2^25 u64 keys
// runs in about half the time of the same tile with a 64-bit stage.
template <typename K, int BLOCK, int KPT>
struct OnesweepSmem {
  static constexpr int W = BLOCK / kWave;
  static constexpr int TILE = BLOCK * KPT;
  uint32_t k[TILE];  // key halves, then payloads, of the tile in digit order
  uint32_t wh[W][kBins];
  uint32_t dstart[kBins];
  uint32_t glob[kBins];
  uint32_t red[W + 1];
};

// Moves key[i] to LDS position pos[i] (0xFFFFFFFF: no key) and returns in
// out[q] the key at LDS position rd(q) (0 where rd(q) >= lim).  `between`
// runs after the first stores, before the first barrier (work that overlaps
// the LDS traffic).  Every thread of the block calls it; the caller
// synchronises before it reuses the stage.
template <typename K, int KPT, typename RdIdx, typename Between>
__device__ __forceinline__ void lds_exchange(uint32_t *sk, const K (&key)[KPT], const uint32_t (&pos)[KPT],
                                             K (&out)[KPT], RdIdx rd, uint32_t lim, Between between) {
#pragma unroll
  for (int i = 0; i < KPT; i++)
    if (pos[i] != 0xFFFFFFFFu) sk[pos[i]] = (uint32_t)key[i];
  between();
  __syncthreads();
  uint32_t lo[KPT];
#pragma unroll
  for (int q = 0; q < KPT; q++) {
    const uint32_t j = rd(q);
    lo[q] = j < lim ? sk[j] : 0u;
  }
  if constexpr (sizeof(K) == 8) {
    __syncthreads();
#pragma unroll
    for (int i = 0; i < KPT; i++)
      if (pos[i] != 0xFFFFFFFFu) sk[pos[i]] = (uint32_t)((uint64_t)key[i] >> 32);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < KPT; q++) {
      const uint32_t j = rd(q);
      out[q] = j < lim ? (K)(((uint64_t)sk[j] << 32) | lo[q]) : (K)0;
    }
  } else {
#pragma unroll
    for (int q = 0; q < KPT; q++) out[q] = (K)lo[q];
  }
}

// Makes v opaque to the optimiser at this point (no instruction): the value
// stays in one register instead of being re-derived later from operands
// that would then have to stay live.
template <typename T>
__device__ __forceinline__ void pin_vgpr(T &v) {
  __asm__ volatile("" : "+v"(v));
}

// One digit pass over tile `tile` of a sequence of n keys.
//   kin/vin -> kout/vout; vin == nullptr means the payload is val_base + row.
//   h: for thread t < 256, the count of digit t over all n keys (read by the
//   caller before the tile is known).  The caller zeroes sm.wh and
//   synchronises before the call.
//   status: ntiles * 256 words of this pass (zeroed before the pass);
//   status_next: the same area for the following pass, zeroed here tile by
//   tile (it was last read two passes ago, which has completed).
// Register budget: a thread holds its KPT keys (as 32-bit halves), their
// packed digits and one rank/position word each; the payloads are loaded
// only once the keys' low halves are staged, so they are in flight during
// the look-back instead of occupying registers through the ranking.  Keys
// are written as they are read back from the stage.
// Every thread of the block calls it; contains barriers.
template <typename K, int BLOCK, int KPT>
__device__ __forceinline__ void onesweep_tile(OnesweepSmem<K, BLOCK, KPT> &sm, uint32_t tile,
                                              const K *__restrict__ kin, const uint32_t *__restrict__ vin,
                                              uint32_t val_base, K *__restrict__ kout, uint32_t *__restrict__ vout,
                                              uint32_t n, int shift, uint32_t desc, uint32_t h, uint32_t *status,
                                              uint32_t *status_next) {
  static_assert(BLOCK >= kBins, "one thread per digit");
  static_assert(KPT % 4 == 0, "digits are packed four per word");
  constexpr bool WIDE = sizeof(K) == 8;
  constexpr int W = BLOCK / kWave;
  constexpr int TILE = BLOCK * KPT;
  const int t = threadIdx.x;
  const int lane = lane_id();
  const int wave = t / kWave;
  const uint32_t base = tile * (uint32_t)TILE;
  if (status_next && t < kBins) status_next[tile * kBins + t] = 0;

  // ---- load: wave w owns rows [base + w*KPT*64, ...), item i lane l -> row
  const uint32_t wbase = base + (uint32_t)wave * KPT * kWave;
  uint32_t klo[KPT], khi[KPT];
  uint32_t dg[KPT / 4];  // the items' digits, four per word
  {
    K key[KPT];
#pragma unroll
    for (int i = 0; i < KPT; i++) {
      const uint32_t idx = wbase + i * kWave + lane;
      key[i] = idx < n ? kin[idx] : (K)0;
    }
#pragma unroll
    for (int i = 0; i < KPT / 4; i++) dg[i] = 0;
#pragma unroll
    for (int i = 0; i < KPT; i++) {
      dg[i / 4] |= ((((uint32_t)(key[i] >> shift)) & 0xFFu) ^ desc) << (8 * (i % 4));
      klo[i] = (uint32_t)key[i];
      khi[i] = WIDE ? (uint32_t)((uint64_t)key[i] >> 32) : 0u;
      pin_vgpr(klo[i]);
      if constexpr (WIDE) pin_vgpr(khi[i]);
    }
  }

  // ---- rank inside the wave, in (item, lane) order
  uint32_t pos[KPT];  // the rank inside the wave, then the position in the tile
#pragma unroll
  for (int i = 0; i < KPT; i++) {
    const uint32_t idx = wbase + i * kWave + lane;
    const bool valid = idx < n;
    const uint32_t d = (dg[i / 4] >> (8 * (i % 4))) & 0xFFu;
    const uint64_t m = match_digit8(d, valid);
    const uint32_t lower = popc_below(m);
    uint32_t pre = 0;
    if (valid) pre = sm.wh[wave][d];
    pos[i] = pre + lower;
    pin_vgpr(pos[i]);
    if (valid && lower == 0) sm.wh[wave][d] = pre + (uint32_t)__popcll(m);
  }
  __syncthreads();

  // ---- per-digit tile counts, wave offsets, publication
  uint32_t cnt = 0;
  if (t < kBins) {
    uint32_t run = 0;
#pragma unroll
    for (int w = 0; w < W; w++) {
      const uint32_t c = sm.wh[w][t];
      sm.wh[w][t] = run;
      run += c;
    }
    cnt = run;
    rstatus_store(&status[tile * kBins + t], (tile == 0 ? kRFlagInc : kRFlagAgg) | cnt);
  }
  uint32_t tot;
  const uint32_t dstart = block_exclusive_sum<BLOCK, uint32_t>(t < kBins ? cnt : 0u, sm.red, &tot);
  const uint32_t gstart = block_exclusive_sum<BLOCK, uint32_t>(t < kBins ? h : 0u, sm.red, &tot);
  if (t < kBins) sm.dstart[t] = dstart;
  __syncthreads();

  // ---- positions in the tile's digit order; low halves into the stage
#pragma unroll
  for (int i = 0; i < KPT; i++) {
    const uint32_t idx = wbase + i * kWave + lane;
    if (idx < n) {
      uint32_t w4 = dg[i / 4];
      pin_vgpr(w4);  // re-derive the LDS addresses rather than keep them from the ranking
      const uint32_t d = (w4 >> (8 * (i % 4))) & 0xFFu;
      pos[i] += sm.dstart[d] + sm.wh[wave][d];
    } else {
      pos[i] = 0xFFFFFFFFu;
    }
  }
#pragma unroll
  for (int i = 0; i < KPT; i++)
    if (pos[i] != 0xFFFFFFFFu) sm.k[pos[i]] = klo[i];
  // payloads: in flight during the look-back
  uint32_t val[KPT];
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int i = 0; i < KPT; i++) {
    const uint32_t idx = wbase + i * kWave + lane;
    val[i] = vin ? (idx < n ? vin[idx] : 0u) : val_base + idx;
  }
  __builtin_amdgcn_sched_barrier(0);

  // ---- decoupled look-back, one thread per digit, kLookbackBatch
  // predecessors per step (their status loads are in flight together).  A
  // step counts only when every predecessor up to the first inclusive one in
  // it has published; otherwise it is read again.
  if (t < kBins) {
    uint32_t excl = 0;
    if (tile > 0) {
      int64_t k = (int64_t)tile - 1;
      while (true) {
        uint32_t st[kLookbackBatch];
#pragma unroll
        for (int b = 0; b < kLookbackBatch; b++)  // tile 0 is inclusive, so k - b < 0 is never needed
          st[b] = k - b >= 0 ? rstatus_load(&status[(uint32_t)(k - b) * kBins + t]) : kRFlagInc;
        bool ready = true, inc = false;
        uint32_t add = 0;
#pragma unroll
        for (int b = 0; b < kLookbackBatch; b++) {
          if (!inc) {
            const uint32_t f = st[b] & ~kRValMask;
            ready = ready && f != 0;
            add += st[b] & kRValMask;
            inc = f == kRFlagInc;
          }
        }
        if (!ready) continue;
        excl += add;
        if (inc) break;
        k -= kLookbackBatch;
      }
      rstatus_store(&status[tile * kBins + t], kRFlagInc | (excl + cnt));
    }
    sm.glob[t] = gstart + excl - dstart;
  }
  __syncthreads();

  // ---- keys back from the stage in digit order, each written to its
  // global position as soon as it is whole
  const uint32_t valid_cnt = (n - base) < (uint32_t)TILE ? (n - base) : (uint32_t)TILE;
  uint32_t opos[KPT];
  if constexpr (WIDE) {
    uint32_t lo[KPT];
#pragma unroll
    for (int q = 0; q < KPT; q++) {
      const uint32_t j = t + q * BLOCK;
      lo[q] = j < valid_cnt ? sm.k[j] : 0u;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < KPT; i++)
      if (pos[i] != 0xFFFFFFFFu) sm.k[pos[i]] = khi[i];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < KPT; q++) {
      const uint32_t j = t + q * BLOCK;
      opos[q] = 0;
      if (j < valid_cnt) {
        const K key = (K)(((uint64_t)sm.k[j] << 32) | lo[q]);
        const uint32_t d = ((uint32_t)(key >> shift) & 0xFFu) ^ desc;
        opos[q] = sm.glob[d] + j;
        kout[opos[q]] = key;
      }
    }
  } else {
#pragma unroll
    for (int q = 0; q < KPT; q++) {
      const uint32_t j = t + q * BLOCK;
      opos[q] = 0;
      if (j < valid_cnt) {
        const K key = (K)sm.k[j];
        const uint32_t d = ((uint32_t)(key >> shift) & 0xFFu) ^ desc;
        opos[q] = sm.glob[d] + j;
        kout[opos[q]] = key;
      }
    }
  }
  __syncthreads();
  // ---- then the payloads, through the same LDS area
#pragma unroll
  for (int i = 0; i < KPT; i++)
    if (pos[i] != 0xFFFFFFFFu) sm.k[pos[i]] = val[i];
  __syncthreads();
#pragma unroll
  for (int q = 0; q < KPT; q++) {
    const uint32_t j = t + q * BLOCK;
    if (j < valid_cnt) vout[opos[q]] = sm.k[j];
  }
}

// Whole sort of n <= BLOCK * KPT keys in one workgroup: every digit pass
// ranks the keys held in registers (the onesweep wave ranking), offsets each
// digit by the block's exclusive digit prefix (no histogram, no look-back),
// scatters keys and then payloads through LDS and reloads them in the same
// per-thread order.  Stable, so the permutation equals the multi-pass one.
// kin may equal kout and vin vout (everything is read before the first
// store).  Every thread of the block calls it; contains barriers.
template <typename K, int BLOCK, int KPT>
__device__ __forceinline__ void small_sort_block(OnesweepSmem<K, BLOCK, KPT> &sm, const K *kin, const uint32_t *vin,
                                                 uint32_t val_base, K *kout, uint32_t *vout, uint32_t n,
                                                 int begin_bit, int passes, uint32_t desc) {
  static_assert(BLOCK >= kBins, "one thread per digit");
  constexpr int W = BLOCK / kWave;
  const int t = threadIdx.x;
  const int lane = lane_id();
  const int wave = t / kWave;
  const uint32_t wbase = (uint32_t)wave * KPT * kWave;
  K key[KPT];
  uint32_t val[KPT];
#pragma unroll
  for (int i = 0; i < KPT; i++) {
    const uint32_t idx = wbase + i * kWave + lane;
    const bool valid = idx < n;
    key[i] = valid ? kin[idx] : (K)0;
    val[i] = vin ? (valid ? vin[idx] : 0u) : val_base + idx;
  }
  uint32_t *s_v = reinterpret_cast<uint32_t *>(sm.k);
  for (int p = 0; p < passes; p++) {
    const int shift = begin_bit + 8 * p;
    for (int i = t; i < W * kBins; i += BLOCK) (&sm.wh[0][0])[i] = 0;
    __syncthreads();
    uint32_t rank[KPT];
#pragma unroll
    for (int i = 0; i < KPT; i++) {
      const uint32_t idx = wbase + i * kWave + lane;
      const bool valid = idx < n;
      const uint32_t d = ((uint32_t)(key[i] >> shift) & 0xFFu) ^ desc;
      const uint64_t m = match_digit8(d, valid);
      const uint32_t lower = popc_below(m);
      uint32_t pre = 0;
      if (valid) pre = sm.wh[wave][d];
      rank[i] = pre + lower;
      if (valid && lower == 0) sm.wh[wave][d] = pre + (uint32_t)__popcll(m);
    }
    __syncthreads();
    uint32_t cnt = 0;
    if (t < kBins) {
      uint32_t run = 0;
#pragma unroll
      for (int w = 0; w < W; w++) {
        const uint32_t c = sm.wh[w][t];
        sm.wh[w][t] = run;
        run += c;
      }
      cnt = run;
    }
    uint32_t tot;
    const uint32_t dstart = block_exclusive_sum<BLOCK, uint32_t>(t < kBins ? cnt : 0u, sm.red, &tot);
    if (t < kBins) sm.dstart[t] = dstart;
    __syncthreads();
    uint32_t pos[KPT];
#pragma unroll
    for (int i = 0; i < KPT; i++) {
      const uint32_t idx = wbase + i * kWave + lane;
      pos[i] = 0xFFFFFFFFu;
      if (idx < n) {
        const uint32_t d = ((uint32_t)(key[i] >> shift) & 0xFFu) ^ desc;
        pos[i] = sm.dstart[d] + sm.wh[wave][d] + rank[i];
      }
    }
    {
      K nk[KPT];
      // rows are read back in the same (wave, item, lane) order they were
      // loaded in: the keys of the next pass, in this pass's order
      lds_exchange<K, KPT>(
          sm.k, key, pos, nk, [&](int i) { return wbase + (uint32_t)(i * kWave + lane); }, n, []() {});
#pragma unroll
      for (int i = 0; i < KPT; i++) key[i] = nk[i];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < KPT; i++)
      if (pos[i] != 0xFFFFFFFFu) s_v[pos[i]] = val[i];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < KPT; i++) {
      const uint32_t idx = wbase + i * kWave + lane;
      if (idx < n) val[i] = s_v[idx];
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < KPT; i++) {
    const uint32_t idx = wbase + i * kWave + lane;
    if (idx < n) {
      kout[idx] = key[i];
      vout[idx] = val[i];
    }
  }
}

// One digit pass over tiles [tile_begin, tile_end) of a sequence of n keys,
// in tile order, by one workgroup: each tile is ranked and reordered in LDS
// like a onesweep tile, but its digit offsets come from running per-digit
// counters `gofs` (256 words of LDS, the global start of each digit's next
// key, advanced tile by tile) instead of a look-back.  src/vsrc -> dst/vdst;
// vsrc == nullptr means the payload is val_base + row.  Every thread of the
// block calls it; contains barriers.
template <typename K, int BLOCK, int KPT>
__device__ __forceinline__ void sort_pass_range(OnesweepSmem<K, BLOCK, KPT> &os, uint32_t *gofs, const K *src,
                                                const uint32_t *vsrc, uint32_t val_base, K *dst, uint32_t *vdst,
                                                uint32_t n, uint32_t tile_begin, uint32_t tile_end, int shift,
                                                uint32_t desc) {
  constexpr int W = BLOCK / kWave;
  constexpr uint32_t TILE = BLOCK * KPT;
  const int t = threadIdx.x;
  const int lane = lane_id();
  const int wave = t / kWave;
  uint32_t *s_v = reinterpret_cast<uint32_t *>(os.k);
  for (uint32_t tile = tile_begin; tile < tile_end; tile++) {
    for (int i = t; i < W * kBins; i += BLOCK) (&os.wh[0][0])[i] = 0;
    __syncthreads();
    const uint32_t base = tile * TILE;
    const uint32_t wbase = base + (uint32_t)wave * KPT * kWave;
    K key[KPT];
    uint32_t val[KPT], rank[KPT];
#pragma unroll
    for (int i = 0; i < KPT; i++) {
      const uint32_t idx = wbase + i * kWave + lane;
      const bool valid = idx < n;
      key[i] = valid ? src[idx] : (K)0;
      val[i] = vsrc ? (valid ? vsrc[idx] : 0u) : val_base + idx;
    }
#pragma unroll
    for (int i = 0; i < KPT; i++) {
      const uint32_t idx = wbase + i * kWave + lane;
      const bool valid = idx < n;
      const uint32_t d = ((uint32_t)(key[i] >> shift) & 0xFFu) ^ desc;
      const uint64_t m = match_digit8(d, valid);
      const uint32_t lower = popc_below(m);
      uint32_t pre = 0;
      if (valid) pre = os.wh[wave][d];
      rank[i] = pre + lower;
      if (valid && lower == 0) os.wh[wave][d] = pre + (uint32_t)__popcll(m);
    }
    __syncthreads();
    uint32_t cnt = 0;
    if (t < kBins) {
      uint32_t run = 0;
#pragma unroll
      for (int w = 0; w < W; w++) {
        const uint32_t c = os.wh[w][t];
        os.wh[w][t] = run;
        run += c;
      }
      cnt = run;
    }
    uint32_t tot;
    const uint32_t dstart = block_exclusive_sum<BLOCK, uint32_t>(t < kBins ? cnt : 0u, os.red, &tot);
    if (t < kBins) {
      os.dstart[t] = dstart;
      os.glob[t] = gofs[t] - dstart;  // wraps like the onesweep form; + j brings it back
    }
    __syncthreads();
    uint32_t pos[KPT];
#pragma unroll
    for (int i = 0; i < KPT; i++) {
      const uint32_t idx = wbase + i * kWave + lane;
      pos[i] = 0xFFFFFFFFu;
      if (idx < n) {
        const uint32_t d = ((uint32_t)(key[i] >> shift) & 0xFFu) ^ desc;
        pos[i] = os.dstart[d] + os.wh[wave][d] + rank[i];
      }
    }
    const uint32_t valid_cnt = (n - base) < TILE ? (n - base) : TILE;
    K okey[KPT];
    lds_exchange<K, KPT>(
        os.k, key, pos, okey, [&](int q) { return (uint32_t)(t + q * BLOCK); }, valid_cnt, []() {});
    uint32_t opos[KPT];
#pragma unroll
    for (int q = 0; q < KPT; q++) {
      const uint32_t j = t + q * BLOCK;
      opos[q] = 0;
      if (j < valid_cnt) {
        const uint32_t d = ((uint32_t)(okey[q] >> shift) & 0xFFu) ^ desc;
        opos[q] = os.glob[d] + j;
        dst[opos[q]] = okey[q];
      }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < KPT; i++)
      if (pos[i] != 0xFFFFFFFFu) s_v[pos[i]] = val[i];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < KPT; q++) {
      const uint32_t j = t + q * BLOCK;
      if (j < valid_cnt) vdst[opos[q]] = s_v[j];
    }
    if (t < kBins) gofs[t] += cnt;
    __syncthreads();
  }
}

// The 256 counts of digit (key >> shift) & 0xFF (^ desc) over keys[begin,
// end), written (not added) to out[0..255].  s_h: 256 words of LDS.  Every
// thread of the block calls it; contains barriers.
template <typename K>
__device__ __forceinline__ void digit_counts_range(const K *keys, uint64_t begin, uint64_t end, int shift,
                                                   uint32_t desc, uint32_t *s_h, uint32_t *out) {
  for (int i = threadIdx.x; i < kBins; i += blockDim.x) s_h[i] = 0;
  __syncthreads();
  hist_range_lds<K>(keys, begin, end, shift, 1, desc, s_h);
  for (int b = threadIdx.x; b < kBins; b += blockDim.x) out[b] = s_h[b];
  __syncthreads();
}

// One workgroup sorts a whole job of n keys (segmented batches: one job per
// workgroup, no communication between workgroups).  Up to one tile it is
// small_sort_block; beyond, it builds every digit's histogram in LDS with one
// read, then runs the digit passes over the job's tiles in order
// (sort_pass_range, offsets from running per-digit counters in LDS).
// Buffers alternate kin -> alt -> ... so that the last pass lands in
// kout/vout (kin must not alias kout).  Stable: the same permutation as the
// multi-workgroup sort.  Every thread of the block calls it.
template <typename K, int BLOCK, int KPT>
struct SegSortSmem {
  OnesweepSmem<K, BLOCK, KPT> os;
  uint32_t h[8 * kBins];  // per-digit-position histograms, then running offsets
};

// Source and destination of digit pass p of `passes` that alternate
// kin -> alt -> ... and end in out (pass 0 reads kin with row payloads).
template <typename K>
struct PassBufs {
  const K *src;
  const uint32_t *vsrc;
  K *dst;
  uint32_t *vdst;
};
template <typename K>
__device__ __forceinline__ PassBufs<K> pass_bufs(int p, int passes, const K *kin, K *kout, uint32_t *vout, K *kalt,
                                                 uint32_t *valt) {
  const bool to_out = ((passes - 1 - p) & 1) == 0;
  PassBufs<K> b;
  b.src = p == 0 ? kin : (((passes - p) & 1) == 0 ? kout : kalt);
  b.vsrc = p == 0 ? nullptr : (((passes - p) & 1) == 0 ? vout : valt);
  b.dst = to_out ? kout : kalt;
  b.vdst = to_out ? vout : valt;
  return b;
}

template <typename K, int BLOCK, int KPT>
__device__ __forceinline__ void seg_sort_block(SegSortSmem<K, BLOCK, KPT> &sm, const K *kin, uint32_t val_base,
                                               K *kout, uint32_t *vout, K *kalt, uint32_t *valt, uint32_t n,
                                               int begin_bit, int passes, uint32_t desc) {
  constexpr uint32_t TILE = BLOCK * KPT;
  if (n <= TILE) {
    small_sort_block<K, BLOCK, KPT>(sm.os, kin, nullptr, val_base, kout, vout, n, begin_bit, passes, desc);
    return;
  }
  const int t = threadIdx.x;
  const int lane = lane_id();
  // ---- histograms of every digit position, one read of the keys
  for (int i = t; i < passes * kBins; i += BLOCK) sm.h[i] = 0;
  __syncthreads();
  {
    uint32_t zc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (uint32_t i = t; i < n; i += BLOCK) hist_key<K>(kin[i], begin_bit, passes, desc, sm.h, zc);
#pragma unroll
    for (int p = 0; p < 8; p++) {
      if (p < passes) {
        const uint32_t z = wave_sum(zc[p]);
        if (lane == 0 && z) atomicAdd(&sm.h[p * kBins + desc], z);
      }
    }
  }
  __syncthreads();
  // ---- running offsets: exclusive prefix of each pass's histogram
  for (int p = 0; p < passes; p++) {
    uint32_t tot;
    const uint32_t c = t < kBins ? sm.h[p * kBins + t] : 0u;
    const uint32_t ex = block_exclusive_sum<BLOCK, uint32_t>(c, sm.os.red, &tot);
    if (t < kBins) sm.h[p * kBins + t] = ex;
    __syncthreads();
  }
  const uint32_t ntiles = (n + TILE - 1) / TILE;
  for (int p = 0; p < passes; p++) {
    const PassBufs<K> b = pass_bufs<K>(p, passes, kin, kout, vout, kalt, valt);
    sort_pass_range<K, BLOCK, KPT>(sm.os, sm.h + p * kBins, b.src, b.vsrc, val_base, b.dst, b.vdst, n, 0, ntiles,
                                   begin_bit + 8 * p, desc);
    // the next pass reads what this one wrote (other threads' stores, which
    // may have been cached before they were rewritten)
    __threadfence();
    __syncthreads();
  }
}

}  // namespace ysql
