// common.hpp — shared device and host helpers for the gfx950 kernels.
//
// Wave64 primitives (ballot / mbcnt ranks, shuffles), the decoupled
// look-back used by every single-pass scan on this path, dtype traits, and
// the host-side error plumbing of the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include "ysql_kernels.h"

namespace ysql {

constexpr int kWave = 64;

// ---------------------------------------------------------------------------
// Host-side status plumbing
// ---------------------------------------------------------------------------
void set_last_error(const char *msg);

#define YSQL_HIP_TRY(expr)                          \
  do {                                              \
    hipError_t _e = (expr);                         \
    if (_e != hipSuccess) {                         \
      ::ysql::set_last_error(hipGetErrorString(_e)); \
      return YSQL_EHIP;                             \
    }                                               \
  } while (0)

// Checks the launch that was just issued.
#define YSQL_LAUNCH_CHECK() YSQL_HIP_TRY(hipGetLastError())

static inline hipStream_t as_stream(ysql_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

__host__ __device__ static inline uint64_t ceil_div(uint64_t a, uint64_t b) { return (a + b - 1) / b; }
static inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// Number of workgroups that keeps every CU busy for a streaming kernel
// (256 CUs on MI355X; a few groups per CU hide HBM latency).
int grid_cap();

// Bump allocator over a caller-supplied workspace (256-byte aligned pieces).
struct WsCarver {
  char *base;
  size_t size;
  size_t off = 0;
  WsCarver(void *b, size_t s) : base(static_cast<char *>(b)), size(s) {}
  template <typename T>
  T *take(size_t count) {
    size_t bytes = align_up(count * sizeof(T), 256);
    if (off + bytes > size) return nullptr;
    T *p = reinterpret_cast<T *>(base + off);
    off += bytes;
    return p;
  }
};
// Same arithmetic, used by the *_ws_bytes() queries.
struct WsCounter {
  size_t off = 0;
  template <typename T>
  void take(size_t count) { off += align_up(count * sizeof(T), 256); }
};

// Fills `n` 64-bit words with zero on `stream` (a kernel, not a memset node,
// so the sequence stays correct under graph capture).
int zero_words(void *p, size_t bytes, hipStream_t stream);

// ---------------------------------------------------------------------------
// Device helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ int lane_id() { return __lane_id(); }

// Bits of `mask` held by lanes below the calling lane.
__device__ __forceinline__ uint32_t popc_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

template <typename T>
__device__ __forceinline__ T shfl_xor_t(T v, int m) {
  return __shfl_xor(v, m, kWave);
}
template <typename T>
__device__ __forceinline__ T shfl_up_t(T v, int d) {
  return __shfl_up(v, d, kWave);
}
template <typename T>
__device__ __forceinline__ T shfl_t(T v, int src) {
  return __shfl(v, src, kWave);
}

// Inclusive wave-wide prefix sum (Kogge-Stone over 64 lanes).
template <typename T>
__device__ __forceinline__ T wave_inclusive_sum(T v) {
  const int lane = lane_id();
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    T o = shfl_up_t(v, d);
    if (lane >= d) v += o;
  }
  return v;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += shfl_xor_t(v, m);
  return v;
}

// Block-wide exclusive prefix sum over one value per thread.  `scratch`
// holds at least BLOCK/64 + 1 values.  Returns the exclusive prefix and sets
// *total.  Contains barriers: every thread of the block must call it.
template <int BLOCK, typename T>
__device__ __forceinline__ T block_exclusive_sum(T v, T *scratch, T *total) {
  constexpr int W = BLOCK / kWave;
  const int lane = lane_id();
  const int wave = threadIdx.x / kWave;
  T inc = wave_inclusive_sum(v);
  if (lane == kWave - 1) scratch[wave] = inc;
  __syncthreads();
  T wave_off = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < W; w++) {
    T s = scratch[w];
    if (w < wave) wave_off += s;
    tot += s;
  }
  __syncthreads();
  *total = tot;
  return wave_off + inc - v;
}

// ---------------------------------------------------------------------------
// Decoupled look-back over 64-bit tile status words:
//   bits 63..62 = flag (0 = not ready, 1 = aggregate, 2 = inclusive prefix)
//   bits 61..0  = value
// Tiles take their id from an atomic counter in launch order, so every
// predecessor a tile waits on is already running and never waits on a later
// tile: the wait always ends.
// ---------------------------------------------------------------------------
constexpr uint64_t kFlagAgg = 1ull << 62;
constexpr uint64_t kFlagInc = 2ull << 62;
constexpr uint64_t kFlagMask = 3ull << 62;
constexpr uint64_t kValMask = (1ull << 62) - 1;

__device__ __forceinline__ uint64_t status_load(const uint64_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void status_store(uint64_t *p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Called by all 64 lanes of ONE wave.  Publishes the tile aggregate, walks
// back over predecessors 64 at a time, publishes the inclusive prefix and
// returns the exclusive prefix (same value in every lane).
__device__ __forceinline__ uint64_t lookback_exclusive(uint64_t *status, uint32_t tile, uint64_t agg) {
  const int lane = lane_id();
  if (tile == 0) {
    if (lane == 0) status_store(&status[0], kFlagInc | agg);
    return 0;
  }
  if (lane == 0) status_store(&status[tile], kFlagAgg | agg);
  uint64_t excl = 0;
  int64_t base = (int64_t)tile - 1;
  while (true) {
    const int64_t idx = base - lane;
    uint64_t s;
    uint64_t inc_mask, zero_mask;
    do {
      s = idx >= 0 ? status_load(&status[idx]) : kFlagInc;
      inc_mask = ballot((s & kFlagMask) == kFlagInc);
      zero_mask = ballot((s & kFlagMask) == 0);
      // lanes that matter: up to and including the first inclusive one
      // (lowest_bit << 1) - 1 wraps to all ones when the lowest bit is 63
      const uint64_t upto = inc_mask ? ((inc_mask & (~inc_mask + 1)) << 1) - 1 : ~0ull;
      zero_mask &= upto;
    } while (zero_mask);
    const int first_inc = inc_mask ? __builtin_ctzll(inc_mask) : kWave;
    uint64_t v = (lane <= first_inc) ? (s & kValMask) : 0;
    excl += wave_sum(v);
    if (inc_mask) break;
    base -= kWave;
  }
  if (lane == 0) status_store(&status[tile], kFlagInc | (excl + agg));
  return excl;
}

// ---------------------------------------------------------------------------
// dtype traits
// ---------------------------------------------------------------------------
static inline int dtype_bytes(int dtype) {
  switch (dtype) {
    case YSQL_INT: case YSQL_FLOAT: case YSQL_UKEY32: return 4;
    case YSQL_DATETIME: case YSQL_UKEY64: return 8;
    case YSQL_STRING: return YSQL_STRING_BYTES;
    case YSQL_BOOL: return 1;
  }
  return 0;
}

struct Str256 {
  char data[YSQL_STRING_BYTES];
};

// strcmp over fixed 256-byte NUL-padded strings (string_kernel.hpp:46-56).
__device__ __forceinline__ int str_cmp(const char *a, const char *b) {
  int i = 0;
  while (i < YSQL_STRING_BYTES && a[i] && a[i] == b[i]) i++;
  if (i >= YSQL_STRING_BYTES) return 0;
  return (int)(unsigned char)a[i] - (int)(unsigned char)b[i];
}

template <typename T>
__device__ __forceinline__ bool cmp_apply(int cmp, T a, T b) {
  switch (cmp) {
    case YSQL_LT: return a < b;
    case YSQL_LE: return a <= b;
    case YSQL_GT: return a > b;
    case YSQL_GE: return a >= b;
    case YSQL_EQ: return a == b;
    default: return a != b;
  }
}

template <int CMP, typename T>
__device__ __forceinline__ bool cmp_static(T a, T b) {
  if constexpr (CMP == YSQL_LT) return a < b;
  else if constexpr (CMP == YSQL_LE) return a <= b;
  else if constexpr (CMP == YSQL_GT) return a > b;
  else if constexpr (CMP == YSQL_GE) return a >= b;
  else if constexpr (CMP == YSQL_EQ) return a == b;
  else return a != b;
}

}  // namespace ysql
