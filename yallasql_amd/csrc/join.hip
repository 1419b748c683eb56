// join.hip — inner joins producing pair lists, and pair-list intersection.
//
// Reference: outer_join_batches (comparison_operators_kernel.cu:5-22) tests
// every (l, r) of a 32x32-thread 2-D grid and appends matches at an atomic
// slot of an 8 MiB buffer with no bound check; and_join_batches (:24-54)
// intersects two pair lists in an O(P1*P2) grid.
//
// Here every join writes its pairs ordered by (l, r) through a decoupled
// look-back scan of per-row match counts, never past `cap`, and reports the
// full count:
//   EQ on INT/FLOAT/DATETIME: the right side is sorted by key with the radix
//     sort (stable, so equal keys stay in r order), each distinct key's run
//     (start, count) goes into an open-addressing hash table, and each left
//     row probes it once.
//   other ops and strings: a nested loop over right chunks staged in LDS,
//     counted once and then written.
#include <math.h>

#include <algorithm>

#include "common.hpp"

namespace ysql {

constexpr int kJoinBlock = 256;

// ---- key canonicalisation for hashing: equal under '==' <=> equal bits
template <typename T>
struct JoinKey;
template <>
struct JoinKey<int32_t> {
  typedef uint32_t U;
  static __device__ __forceinline__ bool ok(int32_t) { return true; }
  static __device__ __forceinline__ U bits(int32_t v) { return (uint32_t)v; }
};
template <>
struct JoinKey<float> {
  typedef uint32_t U;
  static __device__ __forceinline__ bool ok(float v) { return !isnan(v); }
  static __device__ __forceinline__ U bits(float v) { return v == 0.0f ? 0u : __float_as_uint(v); }
};
template <>
struct JoinKey<int64_t> {
  typedef uint64_t U;
  static __device__ __forceinline__ bool ok(int64_t) { return true; }
  static __device__ __forceinline__ U bits(int64_t v) { return (uint64_t)v; }
};

// 32-bit hash of a 64-bit key: the high word is folded into the low one
// with a multiply, the result is spread by a golden-ratio multiply, and the
// high bits are folded back into the low ones, so both the low bits (the
// slot) and the high bits (the probe's membership bitmap) depend on every
// key bit.  Two 32-bit multiplies; build and probe use the same function.
__device__ __forceinline__ uint32_t hash_u64(uint64_t x) {
  uint32_t h = ((uint32_t)x ^ ((uint32_t)(x >> 32) * 0x85ebca6bu)) * 0x9e3779b1u;
  h ^= h >> 15;
  return h;
}

struct HashSlot {
  uint64_t key;
  uint32_t start;
  uint32_t count;
};

template <typename T>
__global__ void k_join_keys(const T *right, uint64_t R, typename JoinKey<T>::U *keys, uint32_t *nan_count) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < R; i += (uint64_t)gridDim.x * blockDim.x) {
    const T v = right[i];
    // NaN keys never match: map them to the maximum key so they sort last
    // and are excluded from the table.
    keys[i] = JoinKey<T>::ok(v) ? JoinKey<T>::bits(v) : (typename JoinKey<T>::U)(~(typename JoinKey<T>::U)0);
    if (!JoinKey<T>::ok(v)) atomicAdd(nan_count, 1u);
  }
}

template <typename U>
__device__ __forceinline__ uint64_t upper_bound_u(const U *a, uint64_t n, U v, uint64_t lo) {
  uint64_t hi = n;
  while (lo < hi) {
    uint64_t mid = (lo + hi) >> 1;
    if (a[mid] <= v) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// One thread per run start of the sorted keys inserts (key, start, count).
// Keys whose rows are NaN (counted in nan_count, sorted last) are skipped.
// When bloom is given, each key also sets two bits of a 2^lgb-bit membership
// filter.  Split filter (staged in LDS by the probe): bits h >> (32 - lgb)
// and (h * 0x9e3779b1) >> (32 - lgb) of its hash h.  Blocked filter (read
// from global memory, one word per row): both bits in word h >> (37 - lgb),
// at bit positions g >> 27 and (g >> 22) & 31 of g = h * 0x9e3779b1.
__device__ __forceinline__ uint32_t bloom_word_bits(uint32_t h) {
  const uint32_t g = h * 0x9e3779b1u;
  return (1u << (g >> 27)) | (1u << ((g >> 22) & 31u));
}

template <typename U>
__global__ void k_join_build(const U *skeys, uint64_t R, const uint32_t *nan_count, HashSlot *table,
                             uint32_t *used, uint32_t mask, uint32_t *bloom, int lgb, bool blocked) {
  const uint64_t valid = R - *nan_count;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < valid; i += (uint64_t)gridDim.x * blockDim.x) {
    const U k = skeys[i];
    if (i > 0 && skeys[i - 1] == k) continue;
    const uint64_t end = upper_bound_u<U>(skeys, valid, k, i + 1);
    const uint32_t hk = hash_u64((uint64_t)k);
    if (bloom && blocked) {
      atomicOr(&bloom[hk >> (37 - lgb)], bloom_word_bits(hk));
    } else if (bloom) {
      const uint32_t b1 = hk >> (32 - lgb), b2 = (hk * 0x9e3779b1u) >> (32 - lgb);
      atomicOr(&bloom[b1 >> 5], 1u << (b1 & 31));
      atomicOr(&bloom[b2 >> 5], 1u << (b2 & 31));
    }
    uint32_t h = hk & mask;
    while (atomicCAS(&used[h], 0u, 1u) != 0u) h = (h + 1) & mask;
    table[h].key = (uint64_t)k;
    table[h].start = (uint32_t)i;
    table[h].count = (uint32_t)(end - i);
  }
}

__device__ __forceinline__ uint32_t take_tile_j(uint32_t *ctr, uint32_t *s_tile) {
  if (threadIdx.x == 0) *s_tile = atomicAdd(ctr, 1u);
  __syncthreads();
  return *s_tile;
}

// Probes: blocks of kProbeTile left rows, kProbeRpt rows per thread (the
// layout is described at probe_row below).
constexpr int kProbeRpt = 16;
constexpr uint64_t kProbeTile = (uint64_t)kJoinBlock * kProbeRpt;

// EQ probe in three launches, so that no workgroup ever waits on another
// (a decoupled look-back per tile measured slow here: each tile walked back
// over the ~1,000 tiles in flight, one cross-die status read per 64):
//   1. count pass (k_join_probe_lds / k_join_probe_bloom): persistent
//      workgroups stream steps of kProbeTile rows and store, per thread, a
//      hit mask of its 16 rows and, per wave, the pairs its rows produce;
//   2. k_probe_block_totals, k_probe_offsets: prefix of those counts in row
//      order, and the total;
//   3. k_join_write: only waves with hits look their rows up again and
//      write their pairs, ordered by (l, r).
// k_join_probe_lds serves hash tables of up to kLdsSlots slots (right sides
// of up to kLdsSlots / 2 rows): each workgroup stages the table in LDS
// once.  LDS holds each slot's start << 16 | count (0: empty; start and
// count are below 2^11 here) and a 64K-bit two-probe membership filter (a
// Bloom filter over two 16-bit functions of the hash), which rejects all
// but about 0.4% of non-matching rows with two LDS reads; the rest probe in
// rounds, every pending row of a thread advancing one slot per round so
// that their reads overlap, comparing the full key in the global table
// (64 KB, cache-resident).
constexpr uint32_t kLdsSlots = 4096;
constexpr uint32_t kLdsBits = 65536;

// Probe layout: in each block of kProbeTile rows, wave w owns the
// kWaveProbeRows rows from w * kWaveProbeRows; a lane holds kProbeRpt of
// them as groups of R = 16 / key bytes consecutive rows, group i of lane l
// at i * 64 * R + l * R, so that each 16-byte load instruction of a wave
// reads 1 KB contiguous.
constexpr uint64_t kWaveProbeRows = (uint64_t)kWave * kProbeRpt;

template <typename T>
__device__ __forceinline__ uint64_t probe_row(uint64_t wbase, int lane, int e) {
  constexpr int R = 16 / sizeof(T);
  return wbase + (uint64_t)(e / R) * (kWave * R) + (uint64_t)lane * R + (uint64_t)(e % R);
}

template <typename T>
__device__ __forceinline__ void probe_load(const T *left, uint64_t L, uint64_t wbase, int lane, T (&v)[kProbeRpt]) {
  constexpr int R = 16 / sizeof(T), G = kProbeRpt / R;
  if (wbase + kWaveProbeRows <= L && ((uintptr_t)(left + wbase) & 15) == 0) {
    uint4 q[G];
#pragma unroll
    for (int i = 0; i < G; i++)
      q[i] = *reinterpret_cast<const uint4 *>(left + wbase + (uint64_t)i * kWave * R + (uint64_t)lane * R);
    memcpy(&v[0], &q[0], sizeof(q));
  } else {
#pragma unroll
    for (int e = 0; e < kProbeRpt; e++) {
      const uint64_t r = probe_row<T>(wbase, lane, e);
      v[e] = r < L ? left[r] : T();
    }
  }
}

// Writes the pairs of one wave's hit rows (bit e of hm: row probe_row(e)),
// the first at position pos, looking each up again in the global table.
// Pairs come out in row order: groups in order and, within a group, lanes
// in order (a wave scan of the group's counts).  Every lane of the wave
// calls it.
template <typename T>
__device__ __forceinline__ void probe_write(const T *left, uint64_t wbase, int lane, uint32_t hm, uint64_t pos,
                                            const HashSlot *table, const uint32_t *used, uint32_t mask,
                                            const uint32_t *perm, uint32_t *pairs, uint64_t cap) {
  constexpr int R = 16 / sizeof(T), G = kProbeRpt / R;
#pragma unroll 1
  for (int i = 0; i < G; i++) {
    const uint32_t gm = (hm >> (i * R)) & ((1u << R) - 1u);
    if (ballot(gm != 0u) == 0) continue;  // wave-uniform
    uint32_t st[R], cn[R];
    uint64_t ci = 0;
#pragma unroll
    for (int j = 0; j < R; j++) {
      st[j] = 0;
      cn[j] = 0;
      if ((gm >> j) & 1u) {
        const uint64_t key = (uint64_t)JoinKey<T>::bits(left[probe_row<T>(wbase, lane, i * R + j)]);
        uint32_t slot = hash_u64(key) & mask;
        while (used[slot] && table[slot].key != key) slot = (slot + 1) & mask;
        st[j] = table[slot].start;
        cn[j] = table[slot].count;
        ci += cn[j];
      }
    }
    const uint64_t inc = wave_inclusive_sum<uint64_t>(ci);
    uint64_t p = pos + inc - ci;
#pragma unroll
    for (int j = 0; j < R; j++) {
      const uint32_t row = (uint32_t)probe_row<T>(wbase, lane, i * R + j);
      for (uint32_t q = 0; q < cn[j]; q++, p++) {
        if (p < cap) {
          pairs[2 * p] = row;
          pairs[2 * p + 1] = perm[st[j] + q];
        }
      }
    }
    pos += shfl_t<uint64_t>(inc, kWave - 1);
  }
}

// Count pass output of one step (kProbeTile rows): hits[step * kJoinBlock +
// thread] is the thread's hit mask (bit e: row probe_row(e) has matches),
// wcnt[step * kProbeWaves + wave] the number of pairs of the wave's rows.
constexpr int kProbeWaves = kJoinBlock / kWave;

__device__ __forceinline__ void probe_publish(uint16_t *hits, uint64_t *wcnt, uint64_t step, int wave, int lane,
                                              uint32_t hm, uint64_t cnt) {
  hits[step * kJoinBlock + (uint64_t)wave * kWave + lane] = (uint16_t)hm;
  const uint64_t ws = wave_sum<uint64_t>(cnt);
  if (lane == 0) wcnt[step * kProbeWaves + wave] = ws;
}

// Prefix of the steps' pair counts in row order, in two launches of one
// workgroup per kProbeScanBlock steps: block totals, then each block's
// exclusive scan plus the totals of the blocks before it (at most
// kProbeScanBlock blocks: left sides below 2^32 rows).
constexpr int kProbeScanBlock = 1024;

__device__ __forceinline__ uint64_t step_pairs(const uint64_t *wcnt, uint64_t nsteps, uint64_t st) {
  uint64_t v = 0;
  if (st < nsteps) {
#pragma unroll
    for (int w = 0; w < kProbeWaves; w++) v += wcnt[st * kProbeWaves + w];
  }
  return v;
}

__global__ __launch_bounds__(kProbeScanBlock) void k_probe_block_totals(const uint64_t *wcnt, uint64_t nsteps,
                                                                        uint64_t *btot) {
  __shared__ uint64_t s_red[kProbeScanBlock / kWave + 1];
  const uint64_t v = step_pairs(wcnt, nsteps, (uint64_t)blockIdx.x * kProbeScanBlock + threadIdx.x);
  uint64_t total;
  (void)block_exclusive_sum<kProbeScanBlock, uint64_t>(v, s_red, &total);
  if (threadIdx.x == 0) btot[blockIdx.x] = total;
}

__global__ __launch_bounds__(kProbeScanBlock) void k_probe_offsets(const uint64_t *wcnt, uint64_t nsteps,
                                                                   const uint64_t *btot, uint64_t *step_off,
                                                                   uint64_t *count_out) {
  __shared__ uint64_t s_red[kProbeScanBlock / kWave + 1];
  uint64_t prefix;
  (void)block_exclusive_sum<kProbeScanBlock, uint64_t>(threadIdx.x < blockIdx.x ? btot[threadIdx.x] : 0ull, s_red,
                                                       &prefix);
  const uint64_t st = (uint64_t)blockIdx.x * kProbeScanBlock + threadIdx.x;
  uint64_t tot;
  const uint64_t ex = block_exclusive_sum<kProbeScanBlock, uint64_t>(step_pairs(wcnt, nsteps, st), s_red, &tot);
  if (st < nsteps) step_off[st] = prefix + ex;
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) *count_out = prefix + tot;
}

// Write pass, one workgroup per step: every wave whose rows have hits looks
// them up again in the global table and writes their pairs from its offset
// (the step's prefix plus the earlier waves of the step).  No barriers.
template <typename T>
__global__ __launch_bounds__(kJoinBlock) void k_join_write(const T *left, const HashSlot *table, const uint32_t *used,
                                                           uint32_t mask, const uint32_t *perm, const uint16_t *hits,
                                                           const uint64_t *wcnt, const uint64_t *step_off,
                                                           uint32_t *pairs, uint64_t cap) {
  const int lane = lane_id(), wave = threadIdx.x / kWave;
  const uint64_t step = blockIdx.x;
  const uint32_t hm = hits[step * kJoinBlock + threadIdx.x];
  if (ballot(hm != 0u) == 0) return;  // wave-uniform
  uint64_t pos = step_off[step];
  for (int w = 0; w < wave; w++) pos += wcnt[step * kProbeWaves + w];
  probe_write<T>(left, step * kProbeTile + (uint64_t)wave * kWaveProbeRows, lane, hm, pos, table, used, mask, perm,
                 pairs, cap);
}

template <typename T>
__global__ __launch_bounds__(kJoinBlock) void k_join_probe_lds(const T *left, uint64_t L, const HashSlot *table,
                                                               const uint32_t *used, uint32_t mask, uint16_t *hits,
                                                               uint64_t *wcnt, uint64_t nsteps) {
  __shared__ uint32_t s_sc[kLdsSlots];
  __shared__ uint32_t s_bits[kLdsBits / 32];
  const int lane = lane_id(), wave = threadIdx.x / kWave;
  for (uint32_t i = threadIdx.x; i < kLdsBits / 32; i += kJoinBlock) s_bits[i] = 0;
  __syncthreads();
  for (uint32_t i = threadIdx.x; i <= mask; i += kJoinBlock) {
    const bool u = used[i] != 0;
    const HashSlot e = u ? table[i] : HashSlot{0, 0, 0};
    s_sc[i] = u ? (e.start << 16) | e.count : 0u;
    if (u) {
      const uint32_t h = hash_u64(e.key);
      const uint32_t b1 = h >> 16, b2 = (h * 0x9e3779b1u) >> 16;
      atomicOr(&s_bits[b1 >> 5], 1u << (b1 & 31));
      atomicOr(&s_bits[b2 >> 5], 1u << (b2 & 31));
    }
  }
  __syncthreads();  // LDS setup complete
  for (uint64_t step = blockIdx.x; step < nsteps; step += gridDim.x) {
    {
      const uint64_t wbase = step * kProbeTile + (uint64_t)wave * kWaveProbeRows;
      T v[kProbeRpt];
      probe_load<T>(left, L, wbase, lane, v);
      uint32_t slot[kProbeRpt];
      uint32_t pend = 0;  // bit e: row e still probing
#pragma unroll
      for (int e = 0; e < kProbeRpt; e++) {
        slot[e] = 0;
        if (probe_row<T>(wbase, lane, e) < L && JoinKey<T>::ok(v[e])) {
          const uint32_t h = hash_u64((uint64_t)JoinKey<T>::bits(v[e]));
          const uint32_t b1 = h >> 16, b2 = (h * 0x9e3779b1u) >> 16;
          if ((s_bits[b1 >> 5] >> (b1 & 31)) & (s_bits[b2 >> 5] >> (b2 & 31)) & 1u) {
            slot[e] = h & mask;
            pend |= 1u << e;
          }
        }
      }
      uint32_t cnt = 0, hm = 0;
      while (pend) {
#pragma unroll
        for (int e = 0; e < kProbeRpt; e++) {
          if ((pend >> e) & 1u) {
            const uint32_t sc = s_sc[slot[e]];
            const uint64_t key = sc ? table[slot[e]].key : 0ull;
            if (sc == 0u) {
              pend &= ~(1u << e);
            } else if (key == (uint64_t)JoinKey<T>::bits(v[e])) {
              cnt += sc & 0xffffu;
              hm |= 1u << e;
              pend &= ~(1u << e);
            } else {
              slot[e] = (slot[e] + 1) & mask;
            }
          }
        }
      }
      probe_publish(hits, wcnt, step, wave, lane, hm, cnt);
    }
  }
}

// Same probe for hash tables too large for LDS.  The build kernel also
// fills a membership filter, so most non-matching rows are rejected without
// reading the table: LGB > 0, a split filter of 2^LGB bits that the
// persistent workgroups copy into LDS once (two LDS reads per row); LGB ==
// 0, a blocked filter of 2^lgb bits in global memory (L2-resident, one word
// read per row, the words of a thread's 16 rows in flight together).  The
// rows that pass probe the global table in rounds (every pending row of a
// thread advances one slot per round, so their reads overlap).  Count pass
// of the three-launch probe, like k_join_probe_lds.
template <typename T, int LGB>
__global__ __launch_bounds__(kJoinBlock) void k_join_probe_bloom(const T *left, uint64_t L, const HashSlot *table,
                                                                 const uint32_t *used, uint32_t mask,
                                                                 const uint32_t *bloom, int lgb, uint16_t *hits,
                                                                 uint64_t *wcnt, uint64_t nsteps) {
  constexpr uint32_t kWords = LGB ? 1u << (LGB - 5) : 1u;
  __shared__ uint32_t s_bits[kWords];
  const int lane = lane_id(), wave = threadIdx.x / kWave;
  if constexpr (LGB != 0)
    for (uint32_t i = threadIdx.x; i < kWords; i += kJoinBlock) s_bits[i] = bloom[i];
  __syncthreads();  // LDS setup complete
  for (uint64_t step = blockIdx.x; step < nsteps; step += gridDim.x) {
    {
      const uint64_t wbase = step * kProbeTile + (uint64_t)wave * kWaveProbeRows;
      T v[kProbeRpt];
      probe_load<T>(left, L, wbase, lane, v);
      uint32_t slot[kProbeRpt];
      uint32_t pend = 0;  // bit e: row e still probing
      if constexpr (LGB != 0) {
#pragma unroll
        for (int e = 0; e < kProbeRpt; e++) {
          slot[e] = 0;
          if (probe_row<T>(wbase, lane, e) < L && JoinKey<T>::ok(v[e])) {
            const uint32_t h = hash_u64((uint64_t)JoinKey<T>::bits(v[e]));
            const uint32_t b1 = h >> (32 - LGB), b2 = (h * 0x9e3779b1u) >> (32 - LGB);
            if ((s_bits[b1 >> 5] >> (b1 & 31)) & (s_bits[b2 >> 5] >> (b2 & 31)) & 1u) {
              slot[e] = h & mask;
              pend |= 1u << e;
            }
          }
        }
      } else {
        uint32_t wd[kProbeRpt];
#pragma unroll
        for (int e = 0; e < kProbeRpt; e++) {  // every row's filter word is in flight before any test
          slot[e] = hash_u64((uint64_t)JoinKey<T>::bits(v[e]));
          wd[e] = (probe_row<T>(wbase, lane, e) < L && JoinKey<T>::ok(v[e])) ? bloom[slot[e] >> (37 - lgb)] : 0u;
        }
#pragma unroll
        for (int e = 0; e < kProbeRpt; e++) {
          const uint32_t m = bloom_word_bits(slot[e]);
          if ((wd[e] & m) == m) pend |= 1u << e;
          slot[e] &= mask;
        }
      }
      uint64_t cnt = 0;
      uint32_t hm = 0;
      while (pend) {
#pragma unroll
        for (int e = 0; e < kProbeRpt; e++) {
          if ((pend >> e) & 1u) {
            const uint32_t u = used[slot[e]];
            const HashSlot hs = table[slot[e]];
            if (u == 0u) {
              pend &= ~(1u << e);
            } else if (hs.key == (uint64_t)JoinKey<T>::bits(v[e])) {
              cnt += hs.count;
              hm |= 1u << e;
              pend &= ~(1u << e);
            } else {
              slot[e] = (slot[e] + 1) & mask;
            }
          }
        }
      }
      probe_publish(hits, wcnt, step, wave, lane, hm, cnt);
    }
  }
}

// ---- nested loop (θ-joins, strings): one left row per thread, right rows
// staged in LDS in chunks; PASS 0 counts, PASS 1 writes.
constexpr int kNlChunk = 1024;

template <typename T, int CMP>
__device__ __forceinline__ bool nl_cmp(const T &a, const T &b) {
  return cmp_static<CMP>(a, b);
}

template <typename T, int CMP>
__global__ __launch_bounds__(kJoinBlock) void k_join_nl(const T *left, uint64_t L, const T *right, uint64_t R,
                                                        uint32_t *pairs, uint64_t cap, uint64_t *count_out,
                                                        uint32_t *ctr, uint64_t *status, uint32_t ntiles) {
  __shared__ T s_r[kNlChunk];
  __shared__ uint64_t s_red[kJoinBlock / kWave + 1];
  __shared__ uint32_t s_tile;
  __shared__ uint64_t s_prefix;
  const uint32_t tile = take_tile_j(ctr, &s_tile);
  const uint64_t l = (uint64_t)tile * kJoinBlock + threadIdx.x;
  const bool lv = l < L;
  const T a = lv ? left[l] : T();
  uint64_t mine = 0;
  for (uint64_t r0 = 0; r0 < R; r0 += kNlChunk) {
    const int cnt = (int)((R - r0) < (uint64_t)kNlChunk ? (R - r0) : (uint64_t)kNlChunk);
    __syncthreads();
    for (int x = threadIdx.x; x < cnt; x += kJoinBlock) s_r[x] = right[r0 + x];
    __syncthreads();
    if (lv)
      for (int x = 0; x < cnt; x++) mine += nl_cmp<T, CMP>(a, s_r[x]);
  }
  uint64_t total;
  const uint64_t excl = block_exclusive_sum<kJoinBlock, uint64_t>(mine, s_red, &total);
  if (threadIdx.x < kWave) {
    const uint64_t p = lookback_exclusive(status, tile, total);
    if (threadIdx.x == 0) s_prefix = p;
  }
  __syncthreads();
  uint64_t pos = s_prefix + excl;
  for (uint64_t r0 = 0; r0 < R; r0 += kNlChunk) {
    const int cnt = (int)((R - r0) < (uint64_t)kNlChunk ? (R - r0) : (uint64_t)kNlChunk);
    __syncthreads();
    for (int x = threadIdx.x; x < cnt; x += kJoinBlock) s_r[x] = right[r0 + x];
    __syncthreads();
    if (lv && mine)
      for (int x = 0; x < cnt; x++)
        if (nl_cmp<T, CMP>(a, s_r[x])) {
          if (pos < cap) {
            pairs[2 * pos] = (uint32_t)l;
            pairs[2 * pos + 1] = (uint32_t)(r0 + x);
          }
          pos++;
        }
  }
  if (tile == ntiles - 1 && threadIdx.x == 0) *count_out = s_prefix + total;
}

// Strings: global reads, no staging (256-byte rows).
__global__ __launch_bounds__(kJoinBlock) void k_join_nl_str(const char *left, uint64_t L, const char *right,
                                                            uint64_t R, int cmp, uint32_t *pairs, uint64_t cap,
                                                            uint64_t *count_out, uint32_t *ctr, uint64_t *status,
                                                            uint32_t ntiles) {
  __shared__ uint64_t s_red[kJoinBlock / kWave + 1];
  __shared__ uint32_t s_tile;
  __shared__ uint64_t s_prefix;
  const uint32_t tile = take_tile_j(ctr, &s_tile);
  const uint64_t l = (uint64_t)tile * kJoinBlock + threadIdx.x;
  const bool lv = l < L;
  const char *a = left + (lv ? l : 0) * YSQL_STRING_BYTES;
  uint64_t mine = 0;
  if (lv)
    for (uint64_t r = 0; r < R; r++) mine += cmp_apply(cmp, str_cmp(a, right + r * YSQL_STRING_BYTES), 0);
  uint64_t total;
  const uint64_t excl = block_exclusive_sum<kJoinBlock, uint64_t>(mine, s_red, &total);
  if (threadIdx.x < kWave) {
    const uint64_t p = lookback_exclusive(status, tile, total);
    if (threadIdx.x == 0) s_prefix = p;
  }
  __syncthreads();
  uint64_t pos = s_prefix + excl;
  if (lv && mine)
    for (uint64_t r = 0; r < R; r++)
      if (cmp_apply(cmp, str_cmp(a, right + r * YSQL_STRING_BYTES), 0)) {
        if (pos < cap) {
          pairs[2 * pos] = (uint32_t)l;
          pairs[2 * pos + 1] = (uint32_t)r;
        }
        pos++;
      }
  if (tile == ntiles - 1 && threadIdx.x == 0) *count_out = s_prefix + total;
}

// ---- pair-list intersection: p2 sorted as (l << 32 | r), p1 entries binary
// searched, kept entries compacted in p1 order.
__global__ void k_pairs_key(const uint32_t *p, uint64_t n, uint64_t *keys) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    keys[i] = ((uint64_t)p[2 * i] << 32) | p[2 * i + 1];
}

__global__ __launch_bounds__(kJoinBlock) void k_pairs_probe(const uint32_t *p1, uint64_t n1, const uint64_t *skeys,
                                                            uint64_t n2, uint32_t *out, uint64_t cap,
                                                            uint64_t *count_out, uint32_t *ctr, uint64_t *status,
                                                            uint32_t ntiles) {
  __shared__ uint64_t s_red[kJoinBlock / kWave + 1];
  __shared__ uint32_t s_tile;
  __shared__ uint64_t s_prefix;
  const uint32_t tile = take_tile_j(ctr, &s_tile);
  const uint64_t i = (uint64_t)tile * kJoinBlock + threadIdx.x;
  bool hit = false;
  uint64_t k = 0;
  if (i < n1) {
    k = ((uint64_t)p1[2 * i] << 32) | p1[2 * i + 1];
    uint64_t lo = 0, hi = n2;
    while (lo < hi) {
      uint64_t mid = (lo + hi) >> 1;
      if (skeys[mid] < k) lo = mid + 1;
      else hi = mid;
    }
    hit = lo < n2 && skeys[lo] == k;
  }
  uint64_t total;
  const uint64_t excl = block_exclusive_sum<kJoinBlock, uint64_t>(hit ? 1 : 0, s_red, &total);
  if (threadIdx.x < kWave) {
    const uint64_t p = lookback_exclusive(status, tile, total);
    if (threadIdx.x == 0) s_prefix = p;
  }
  __syncthreads();
  const uint64_t pos = s_prefix + excl;
  if (hit && pos < cap) {
    out[2 * pos] = (uint32_t)(k >> 32);
    out[2 * pos + 1] = (uint32_t)k;
  }
  if (tile == ntiles - 1 && threadIdx.x == 0) *count_out = s_prefix + total;
}

__global__ void k_join_reset(uint32_t *ctr, uint64_t *status, uint32_t ntiles, uint32_t *used, uint32_t nused,
                             uint32_t *bloom = nullptr, uint32_t nbloom = 0) {
  const uint32_t i0 = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t stride = gridDim.x * blockDim.x;
  if (i0 < 64) ctr[i0] = 0;
  for (uint32_t i = i0; i < ntiles; i += stride) status[i] = 0;
  for (uint32_t i = i0; i < nused; i += stride) used[i] = 0;
  for (uint32_t i = i0; i < nbloom; i += stride) bloom[i] = 0;
}

__global__ void k_store_zero_u64(uint64_t *p) { *p = 0; }

// ---- workspace layout
struct JoinWs {
  uint32_t *ctr;      // [0] tile counter (nested loop), [1] NaN count of the right keys (hash)
  uint64_t *status;   // ntiles + 1: look-back status (nested loop) or step offsets (hash probe)
  void *keys;         // R keys (u32 or u64)
  uint32_t *perm;     // R
  HashSlot *table;    // cap
  uint32_t *used;     // cap
  uint32_t *bloom;    // kBloomMaxWords (membership filter of the global-table count pass)
  uint16_t *hits;     // probe steps * kJoinBlock hit masks
  uint64_t *wcnt;     // probe steps * kProbeWaves pair counts
  uint64_t *btot;     // kProbeScanBlock block totals of the step scan
  void *sort_ws;
  size_t sort_ws_bytes;
};

// Filter of the global-table probe: a split filter of 2^17 bits in LDS up
// to 8,192 right rows (16 bits per row), else a blocked filter in global
// memory of 2^20..2^24 bits, 16 bits per row up to the cap.  (A 2^19-bit LDS
// filter, 64 KB, measured slower than the blocked filter from 8,192 rows:
// 1.39 / 2.61 ms against 1.33 / 1.71 ms at 20,000 / 50,000 rows.)
constexpr int kBloomLdsLgb = 17;
constexpr int kBloomMaxLgb = 24;
constexpr uint32_t kBloomMaxWords = (1u << kBloomMaxLgb) / 32;
static int bloom_lgb(uint64_t R) {
  if (16 * R <= (1u << kBloomLdsLgb)) return kBloomLdsLgb;
  int lg = 20;
  while (lg < kBloomMaxLgb && (1ull << lg) < 16 * R) lg++;
  return lg;
}

static uint32_t hash_cap(uint64_t R) {
  uint64_t c = 64;
  while (c < 2 * R) c <<= 1;
  return (uint32_t)c;
}

static bool is_hash_eq(int dtype, int cmp) {
  return cmp == YSQL_EQ && (dtype == YSQL_INT || dtype == YSQL_FLOAT || dtype == YSQL_DATETIME);
}

static size_t join_ws_layout(int dtype, int cmp, uint64_t L, uint64_t R, void *ws, size_t bytes, JoinWs *out) {
  const bool hash = is_hash_eq(dtype, cmp);
  const uint64_t tile_rows = hash ? kProbeTile : kJoinBlock;
  const uint64_t ntiles = ceil_div(L, tile_rows) + 1;
  const int kdt = dtype == YSQL_DATETIME ? YSQL_UKEY64 : YSQL_UKEY32;
  size_t sort_bytes = 0;
  if (hash) ysql_radix_sort_ws_bytes(kdt, R, &sort_bytes);
  const size_t kb = dtype == YSQL_DATETIME ? 8 : 4;
  const uint32_t cap = hash ? hash_cap(R) : 1;
  if (!out) {
    WsCounter c;
    c.take<uint32_t>(64);
    c.take<uint64_t>(ntiles);
    if (hash) {
      c.take<char>(R * kb + 8);
      c.take<uint32_t>(R + 1);
      c.take<HashSlot>(cap);
      c.take<uint32_t>(cap);
      c.take<uint32_t>(kBloomMaxWords);
      c.take<uint16_t>((ntiles - 1) * kJoinBlock);
      c.take<uint64_t>((ntiles - 1) * kProbeWaves);
      c.take<uint64_t>(kProbeScanBlock);
      c.take<char>(sort_bytes);
    }
    return c.off;
  }
  WsCarver c(ws, bytes);
  out->ctr = c.take<uint32_t>(64);
  out->status = c.take<uint64_t>(ntiles);
  bool ok = out->ctr && out->status;
  if (hash) {
    out->keys = c.take<char>(R * kb + 8);
    out->perm = c.take<uint32_t>(R + 1);
    out->table = c.take<HashSlot>(cap);
    out->used = c.take<uint32_t>(cap);
    out->bloom = c.take<uint32_t>(kBloomMaxWords);
    out->hits = c.take<uint16_t>((ntiles - 1) * kJoinBlock);
    out->wcnt = c.take<uint64_t>((ntiles - 1) * kProbeWaves);
    out->btot = c.take<uint64_t>(kProbeScanBlock);
    out->sort_ws = c.take<char>(sort_bytes);
    out->sort_ws_bytes = sort_bytes;
    ok = ok && out->keys && out->perm && out->table && out->used && out->bloom && out->hits && out->wcnt && out->btot &&
         out->sort_ws;
  }
  return ok ? c.off : 0;
}

// Persistent grid of a count-pass kernel: enough workgroups to fill the
// chip at its occupancy, at most one per step.
template <typename K>
static uint32_t probe_grid(K kernel, uint64_t nsteps) {
  int b = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, kernel, kJoinBlock, 0) != hipSuccess || b < 1) b = 1;
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(nsteps, (uint64_t)b * grid_cap()));
}

template <typename T>
static int join_hash(const T *left, uint64_t L, const T *right, uint64_t R, uint32_t *pairs, uint64_t cap,
                     uint64_t *count_out, const JoinWs &w, hipStream_t s) {
  typedef typename JoinKey<T>::U U;
  const uint32_t tcap = hash_cap(R);
  const uint64_t nsteps = ceil_div(L, kProbeTile);
  const bool filtered = tcap > kLdsSlots;
  const int lgb = bloom_lgb(R);
  k_join_reset<<<256, 256, 0, s>>>(w.ctr, w.status, 0, w.used, tcap, w.bloom, filtered ? (1u << lgb) / 32 : 0u);
  YSQL_LAUNCH_CHECK();
  uint64_t g = ceil_div(R, 256);
  if (g > 4096) g = 4096;
  if (g < 1) g = 1;
  if (R > 0) {
    k_join_keys<T><<<(unsigned)g, 256, 0, s>>>(right, R, (U *)w.keys, w.ctr + 1);
    YSQL_LAUNCH_CHECK();
    const int kdt = sizeof(U) == 8 ? YSQL_UKEY64 : YSQL_UKEY32;
    int rc = ysql_radix_argsort(kdt, 0, w.keys, w.perm, R, 0, 8 * sizeof(U), w.sort_ws, w.sort_ws_bytes, s);
    if (rc) return rc;
    k_join_build<U><<<(unsigned)g, 256, 0, s>>>((const U *)w.keys, R, w.ctr + 1, w.table, w.used, tcap - 1,
                                                filtered ? w.bloom : nullptr, lgb, lgb > kBloomLdsLgb);
    YSQL_LAUNCH_CHECK();
  }
  // 1. count pass: hit masks and per-wave pair counts of every step
  if (!filtered) {
    static const uint32_t gl = probe_grid(k_join_probe_lds<T>, ~0ull);
    k_join_probe_lds<T><<<(uint32_t)std::min<uint64_t>(gl, nsteps), kJoinBlock, 0, s>>>(
        left, L, w.table, w.used, tcap - 1, w.hits, w.wcnt, nsteps);
  } else if (lgb == kBloomLdsLgb) {
    static const uint32_t gb = probe_grid(k_join_probe_bloom<T, kBloomLdsLgb>, ~0ull);
    k_join_probe_bloom<T, kBloomLdsLgb><<<(uint32_t)std::min<uint64_t>(gb, nsteps), kJoinBlock, 0, s>>>(
        left, L, w.table, w.used, tcap - 1, w.bloom, lgb, w.hits, w.wcnt, nsteps);
  } else {
    static const uint32_t gg = probe_grid(k_join_probe_bloom<T, 0>, ~0ull);
    k_join_probe_bloom<T, 0><<<(uint32_t)std::min<uint64_t>(gg, nsteps), kJoinBlock, 0, s>>>(
        left, L, w.table, w.used, tcap - 1, w.bloom, lgb, w.hits, w.wcnt, nsteps);
  }
  YSQL_LAUNCH_CHECK();
  // 2. step offsets and the total; 3. pairs of the steps with hits
  const uint32_t nblk = (uint32_t)ceil_div(nsteps, (uint64_t)kProbeScanBlock);  // <= kProbeScanBlock
  k_probe_block_totals<<<nblk, kProbeScanBlock, 0, s>>>(w.wcnt, nsteps, w.btot);
  YSQL_LAUNCH_CHECK();
  k_probe_offsets<<<nblk, kProbeScanBlock, 0, s>>>(w.wcnt, nsteps, w.btot, w.status, count_out);
  YSQL_LAUNCH_CHECK();
  k_join_write<T><<<(uint32_t)nsteps, kJoinBlock, 0, s>>>(left, w.table, w.used, tcap - 1, w.perm, w.hits, w.wcnt,
                                                          w.status, pairs, cap);
  YSQL_LAUNCH_CHECK();
  return YSQL_OK;
}

template <typename T>
static int join_nl(int cmp, const T *left, uint64_t L, const T *right, uint64_t R, uint32_t *pairs, uint64_t cap,
                   uint64_t *count_out, const JoinWs &w, hipStream_t s) {
  const uint32_t ntiles = (uint32_t)ceil_div(L, kJoinBlock);
  k_join_reset<<<64, 256, 0, s>>>(w.ctr, w.status, ntiles + 1, nullptr, 0);
  YSQL_LAUNCH_CHECK();
#define NL_CASE(C)                                                                                                \
  case C:                                                                                                         \
    k_join_nl<T, C><<<ntiles, kJoinBlock, 0, s>>>(left, L, right, R, pairs, cap, count_out, w.ctr, w.status,     \
                                                  ntiles);                                                        \
    break;
  switch (cmp) {
    NL_CASE(YSQL_LT)
    NL_CASE(YSQL_LE)
    NL_CASE(YSQL_GT)
    NL_CASE(YSQL_GE)
    NL_CASE(YSQL_EQ)
    NL_CASE(YSQL_NE)
    default: return YSQL_EINVAL;
  }
#undef NL_CASE
  YSQL_LAUNCH_CHECK();
  return YSQL_OK;
}

}  // namespace ysql

using namespace ysql;

extern "C" {

int ysql_join_ws_bytes(int dtype, int cmp, uint64_t L, uint64_t R, size_t *bytes) {
  if (!bytes) return YSQL_EINVAL;
  *bytes = join_ws_layout(dtype, cmp, L, R, nullptr, 0, nullptr);
  return YSQL_OK;
}

int ysql_join_pairs(int dtype, int cmp, const void *left, uint64_t L, const void *right, uint64_t R,
                    uint32_t *pairs_out, uint64_t cap, uint64_t *count_out, void *ws, size_t ws_bytes,
                    ysql_stream_t stream) {
  hipStream_t s = as_stream(stream);
  if (!count_out || cmp < YSQL_LT || cmp > YSQL_NE) return YSQL_EINVAL;
  if (L >= (1ull << 32) || R >= (1ull << 30)) return YSQL_EINVAL;
  if (cap > 0 && !pairs_out) return YSQL_EINVAL;
  if (L == 0 || R == 0) {
    k_store_zero_u64<<<1, 1, 0, s>>>(count_out);
    YSQL_LAUNCH_CHECK();
    return YSQL_OK;
  }
  if (!left || !right) return YSQL_EINVAL;
  JoinWs w;
  memset(&w, 0, sizeof(w));
  if (!ws || ws_bytes < join_ws_layout(dtype, cmp, L, R, nullptr, 0, nullptr) ||
      !join_ws_layout(dtype, cmp, L, R, ws, ws_bytes, &w))
    return YSQL_EWORKSPACE;
  if (is_hash_eq(dtype, cmp)) {
    switch (dtype) {
      case YSQL_INT: return join_hash<int32_t>((const int32_t *)left, L, (const int32_t *)right, R, pairs_out, cap, count_out, w, s);
      case YSQL_FLOAT: return join_hash<float>((const float *)left, L, (const float *)right, R, pairs_out, cap, count_out, w, s);
      case YSQL_DATETIME: return join_hash<int64_t>((const int64_t *)left, L, (const int64_t *)right, R, pairs_out, cap, count_out, w, s);
    }
    return YSQL_EINVAL;
  }
  switch (dtype) {
    case YSQL_INT: return join_nl<int32_t>(cmp, (const int32_t *)left, L, (const int32_t *)right, R, pairs_out, cap, count_out, w, s);
    case YSQL_FLOAT: return join_nl<float>(cmp, (const float *)left, L, (const float *)right, R, pairs_out, cap, count_out, w, s);
    case YSQL_DATETIME: return join_nl<int64_t>(cmp, (const int64_t *)left, L, (const int64_t *)right, R, pairs_out, cap, count_out, w, s);
    case YSQL_STRING: {
      const uint32_t ntiles = (uint32_t)ceil_div(L, kJoinBlock);
      k_join_reset<<<64, 256, 0, s>>>(w.ctr, w.status, ntiles + 1, nullptr, 0);
      YSQL_LAUNCH_CHECK();
      k_join_nl_str<<<ntiles, kJoinBlock, 0, s>>>((const char *)left, L, (const char *)right, R, cmp, pairs_out, cap,
                                                   count_out, w.ctr, w.status, ntiles);
      YSQL_LAUNCH_CHECK();
      return YSQL_OK;
    }
  }
  return YSQL_EINVAL;
}

int ysql_pairs_intersect_ws_bytes(uint64_t n1, uint64_t n2, size_t *bytes) {
  if (!bytes) return YSQL_EINVAL;
  size_t sb = 0;
  ysql_radix_sort_ws_bytes(YSQL_UKEY64, n2, &sb);
  WsCounter c;
  c.take<uint32_t>(64);
  c.take<uint64_t>(ceil_div(n1, kJoinBlock) + 1);
  c.take<uint64_t>(n2 + 1);
  c.take<uint32_t>(n2 + 1);
  c.take<char>(sb);
  *bytes = c.off;
  return YSQL_OK;
}

int ysql_pairs_intersect(const uint32_t *p1, uint64_t n1, const uint32_t *p2, uint64_t n2, uint32_t *pairs_out,
                         uint64_t cap, uint64_t *count_out, void *ws, size_t ws_bytes, ysql_stream_t stream) {
  hipStream_t s = as_stream(stream);
  if (!count_out) return YSQL_EINVAL;
  if (n1 >= (1ull << 32) || n2 >= (1ull << 30)) return YSQL_EINVAL;
  if (n1 == 0 || n2 == 0) {
    k_store_zero_u64<<<1, 1, 0, s>>>(count_out);
    YSQL_LAUNCH_CHECK();
    return YSQL_OK;
  }
  if (!p1 || !p2 || (cap > 0 && !pairs_out)) return YSQL_EINVAL;
  size_t need = 0, sb = 0;
  ysql_pairs_intersect_ws_bytes(n1, n2, &need);
  ysql_radix_sort_ws_bytes(YSQL_UKEY64, n2, &sb);
  if (!ws || ws_bytes < need) return YSQL_EWORKSPACE;
  WsCarver c(ws, ws_bytes);
  uint32_t *ctr = c.take<uint32_t>(64);
  const uint32_t ntiles = (uint32_t)ceil_div(n1, kJoinBlock);
  uint64_t *status = c.take<uint64_t>(ntiles + 1);
  uint64_t *keys = c.take<uint64_t>(n2 + 1);
  uint32_t *perm = c.take<uint32_t>(n2 + 1);
  void *sws = c.take<char>(sb);
  if (!ctr || !status || !keys || !perm || !sws) return YSQL_EWORKSPACE;
  k_join_reset<<<64, 256, 0, s>>>(ctr, status, ntiles + 1, nullptr, 0);
  YSQL_LAUNCH_CHECK();
  uint64_t g = ceil_div(n2, 256);
  if (g > 4096) g = 4096;
  k_pairs_key<<<(unsigned)g, 256, 0, s>>>(p2, n2, keys);
  YSQL_LAUNCH_CHECK();
  int rc = ysql_radix_argsort(YSQL_UKEY64, 0, keys, perm, n2, 0, 64, sws, sb, s);
  if (rc) return rc;
  k_pairs_probe<<<ntiles, kJoinBlock, 0, s>>>(p1, n1, keys, n2, pairs_out, cap, count_out, ctr, status, ntiles);
  YSQL_LAUNCH_CHECK();
  return YSQL_OK;
}

}  // extern "C"
