// reduce.hip — MIN/MAX/SUM/AVG/COUNT reductions, grouped aggregation and
// the fused star join + grouped SUM.
//
// Reference: reduction_sum_wrap (reduction_kernel.cu:78-121) reduces 512
// values per block with a 32-lane shuffle tree and folds each block into the
// accumulator with one global atomic (float SUM through a double atomicAdd,
// so the result depends on block timing).
//
// Here the input is cut into canonical tiles of 65536 rows.  A tile's partial
// is computed in a fixed order (each thread folds its rows in order, a wave64
// butterfly, then the 4 waves in order), tile partials are stored, and the
// last workgroup to finish folds them in a fixed order and then folds the
// total into the caller's accumulator.  The result is a function of the input
// alone: the same float SUM comes out of every run and every launch shape.
#include <float.h>
#include <algorithm>
#include <math.h>

#include "common.hpp"

namespace ysql {

constexpr int kRedBlock = 256;
constexpr uint64_t kRedTile = 65536;

// Accumulator kinds
//   KI: int64 in .i (INT, DATETIME: MIN/MAX/SUM/AVG)
//   KF: float in .f (FLOAT MIN/MAX)
//   KD: double in .d (FLOAT SUM/AVG)
//   KC: count only (COUNT, COUNT_STAR)
enum { KI = 0, KF = 1, KD = 2, KC = 3 };

__host__ __device__ static inline int acc_kind(int dtype, int agg) {
  if (agg == YSQL_AGG_COUNT || agg == YSQL_AGG_COUNT_STAR) return KC;
  if (dtype == YSQL_FLOAT) return (agg == YSQL_AGG_SUM || agg == YSQL_AGG_AVG) ? KD : KF;
  return KI;
}

// AGG reduced to its combine operator: 0 = sum, 1 = min, 2 = max
__host__ __device__ static inline int agg_op(int agg) {
  return agg == YSQL_AGG_MIN ? 1 : (agg == YSQL_AGG_MAX ? 2 : 0);
}

template <int KIND, int OP>
__device__ __forceinline__ void acc_identity(ysql_acc &a, int dtype) {
  a.i = 0;
  a.d = 0.0;
  a.f = 0.f;
  a.pad = 0;
  a.count = 0;
  if (KIND == KI) {
    if (OP == 1) a.i = dtype == YSQL_INT ? (int64_t)INT32_MAX : INT64_MAX;
    if (OP == 2) a.i = dtype == YSQL_INT ? (int64_t)INT32_MIN : INT64_MIN;
  } else if (KIND == KF) {
    if (OP == 1) a.f = __builtin_inff();
    if (OP == 2) a.f = -__builtin_inff();
  }
}

template <int KIND, int OP>
__device__ __forceinline__ void acc_combine(ysql_acc &a, const ysql_acc &b) {
  a.count += b.count;
  if (KIND == KI) {
    if (OP == 0) a.i = (int64_t)((uint64_t)a.i + (uint64_t)b.i);
    else if (OP == 1) a.i = b.i < a.i ? b.i : a.i;
    else a.i = b.i > a.i ? b.i : a.i;
  } else if (KIND == KF) {
    if (OP == 1) a.f = fminf(a.f, b.f);
    else a.f = fmaxf(a.f, b.f);
  } else if (KIND == KD) {
    a.d += b.d;
  }
}

template <typename T, int KIND, int OP>
__device__ __forceinline__ void acc_fold(ysql_acc &a, T v) {
  a.count += 1;
  if (KIND == KI) {
    const int64_t x = (int64_t)v;
    if (OP == 0) a.i = (int64_t)((uint64_t)a.i + (uint64_t)x);
    else if (OP == 1) a.i = x < a.i ? x : a.i;
    else a.i = x > a.i ? x : a.i;
  } else if (KIND == KF) {
    if (OP == 1) a.f = fminf(a.f, (float)v);
    else a.f = fmaxf(a.f, (float)v);
  } else if (KIND == KD) {
    a.d += (double)v;
  }
}

template <int KIND, int OP>
__device__ __forceinline__ ysql_acc acc_shfl_xor(const ysql_acc &a, int m) {
  ysql_acc b;
  b.count = shfl_xor_t(a.count, m);
  b.i = (KIND == KI) ? shfl_xor_t(a.i, m) : 0;
  b.d = (KIND == KD) ? shfl_xor_t(a.d, m) : 0.0;
  b.f = (KIND == KF) ? shfl_xor_t(a.f, m) : 0.f;
  b.pad = 0;
  return b;
}

// Fixed-shape block reduction; result valid in thread 0.
template <int KIND, int OP>
__device__ __forceinline__ ysql_acc block_reduce_acc(ysql_acc v, ysql_acc *s_w) {
  constexpr int W = kRedBlock / kWave;
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    ysql_acc o = acc_shfl_xor<KIND, OP>(v, m);
    acc_combine<KIND, OP>(v, o);
  }
  const int wave = threadIdx.x / kWave;
  if (lane_id() == 0) s_w[wave] = v;
  __syncthreads();
  ysql_acc r = s_w[0];
#pragma unroll
  for (int w = 1; w < W; w++) acc_combine<KIND, OP>(r, s_w[w]);
  __syncthreads();
  return r;
}

// Folds the rows [base, end) of one canonical tile, in a fixed order:
// rounds of 1024 rows, thread t owning rows t*4 .. t*4+3 of each round.
template <typename T, int KIND, int OP, bool STAR>
__device__ __forceinline__ void fold_tile(ysql_acc &p, const T *in, const uint8_t *nul, uint64_t base, uint64_t end) {
  const int t = threadIdx.x;
  for (uint64_t r = base; r < end; r += 1024) {
    const uint64_t i0 = r + (uint64_t)t * 4;
    if (STAR || (KIND == KC && !nul)) {  // every row counts; values unread
      uint64_t c = 0;
      for (int e = 0; e < 4; e++) c += (i0 + e < end);
      p.count += c;
      continue;
    }
    if (KIND == KC) {  // COUNT: rows whose null byte is 0; values unread
      uint64_t c = 0;
      for (int e = 0; e < 4; e++) c += (i0 + e < end) && !nul[i0 + e];
      p.count += c;
      continue;
    }
    T v[4];
    uint8_t nn[4] = {0, 0, 0, 0};
    if (i0 + 4 <= end && ((uintptr_t)(in + i0) & 15) == 0) {
      if (sizeof(T) == 4) {
        uint4 q = *(const uint4 *)(in + i0);
        memcpy(v, &q, 16);
      } else {
        uint4 q[2];
        q[0] = ((const uint4 *)(in + i0))[0];
        q[1] = ((const uint4 *)(in + i0))[1];
        memcpy(v, q, 32);
      }
      if (nul) {
        if (((uintptr_t)(nul + i0) & 3) == 0) {
          uint32_t w = *(const uint32_t *)(nul + i0);
          memcpy(nn, &w, 4);
        } else {
          for (int e = 0; e < 4; e++) nn[e] = nul[i0 + e];
        }
      }
#pragma unroll
      for (int e = 0; e < 4; e++)
        if (!nn[e]) acc_fold<T, KIND, OP>(p, v[e]);
    } else {
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const uint64_t i = i0 + e;
        if (i < end && !(nul && nul[i])) acc_fold<T, KIND, OP>(p, in[i]);
      }
    }
  }
}

// Tile partials of one aggregate; k_reduce_finish combines them.  The
// combine runs as a second launch on the stream rather than in the last
// workgroup to finish: a last-finisher hand-off needs an agent-scope release
// per workgroup, which on this multi-XCD device writes back the L2 each time
// (measured: 0.07 to 0.28 ms per 200M-row call, growing with the grid).
template <typename T, int KIND, int OP, bool STAR>
__global__ __launch_bounds__(kRedBlock) void k_reduce(const T *in, const uint8_t *nul, uint64_t n, int dtype,
                                                      ysql_acc *partials, uint32_t ntiles) {
  __shared__ ysql_acc s_w[kRedBlock / kWave];
  for (uint32_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    ysql_acc p;
    acc_identity<KIND, OP>(p, dtype);
    const uint64_t base = (uint64_t)tile * kRedTile;
    const uint64_t end = base + kRedTile < n ? base + kRedTile : n;
    fold_tile<T, KIND, OP, STAR>(p, in, nul, base, end);
    ysql_acc tot = block_reduce_acc<KIND, OP>(p, s_w);
    if (threadIdx.x == 0) partials[tile] = tot;
  }
}

// One workgroup: the tile partials in a fixed order, folded into *acc.
template <int KIND, int OP>
__global__ __launch_bounds__(kRedBlock) void k_reduce_finish(int dtype, ysql_acc *acc, const ysql_acc *partials,
                                                             uint32_t ntiles) {
  __shared__ ysql_acc s_w[kRedBlock / kWave];
  ysql_acc p;
  acc_identity<KIND, OP>(p, dtype);
  for (uint32_t x = threadIdx.x; x < ntiles; x += kRedBlock) {
    ysql_acc q = partials[x];
    acc_combine<KIND, OP>(p, q);
  }
  ysql_acc tot = block_reduce_acc<KIND, OP>(p, s_w);
  if (threadIdx.x == 0) {
    ysql_acc a = *acc;
    acc_combine<KIND, OP>(a, tot);
    *acc = a;
  }
}

__global__ void k_acc_init(ysql_acc *acc, int nacc, int dtype, int agg) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nacc) return;
  ysql_acc a;
  const int kind = acc_kind(dtype, agg), op = agg_op(agg);
  a.i = 0;
  a.d = 0.0;
  a.f = 0.f;
  a.pad = 0;
  a.count = 0;
  if (kind == KI && op == 1) a.i = dtype == YSQL_INT ? (int64_t)INT32_MAX : INT64_MAX;
  if (kind == KI && op == 2) a.i = dtype == YSQL_INT ? (int64_t)INT32_MIN : INT64_MIN;
  if (kind == KF && op == 1) a.f = __builtin_inff();
  if (kind == KF && op == 2) a.f = -__builtin_inff();
  acc[i] = a;
}

__global__ void k_finalize(int dtype, int agg, const ysql_acc *acc, void *out) {
  const ysql_acc a = *acc;
  switch (agg) {
    case YSQL_AGG_MIN:
    case YSQL_AGG_MAX:
      if (dtype == YSQL_INT) *(int32_t *)out = (int32_t)a.i;
      else if (dtype == YSQL_DATETIME) *(int64_t *)out = a.i;
      else *(float *)out = a.f;
      break;
    case YSQL_AGG_SUM:
      if (dtype == YSQL_FLOAT) *(float *)out = (float)a.d;
      else *(int64_t *)out = a.i;
      break;
    case YSQL_AGG_AVG: {
      const double s = dtype == YSQL_FLOAT ? a.d : (double)a.i;
      *(float *)out = (float)(s / (double)a.count);
      break;
    }
    default: *(int64_t *)out = a.count; break;
  }
}

// The reference keeps one plain typed accumulator per aggregate
// (aggregate.hpp:201-205): INT -> int32, FLOAT MIN/MAX -> float, FLOAT
// SUM/AVG -> double, DATETIME -> int64, COUNT and AVG's counter -> int32.
__global__ void k_acc_from_value(int dtype, int agg, const void *v, ysql_acc *acc) {
  ysql_acc a;
  a.i = 0;
  a.d = 0.0;
  a.f = 0.f;
  a.pad = 0;
  a.count = 0;
  if (agg == YSQL_AGG_COUNT || agg == YSQL_AGG_COUNT_STAR) a.count = *(const int32_t *)v;
  else if (dtype == YSQL_INT) a.i = *(const int32_t *)v;
  else if (dtype == YSQL_DATETIME) a.i = *(const int64_t *)v;
  else if (agg == YSQL_AGG_SUM || agg == YSQL_AGG_AVG) a.d = *(const double *)v;
  else a.f = *(const float *)v;
  *acc = a;
}

// Inverse of k_acc_from_value.  An INT sum keeps its low 32 bits: the
// reference's int32 accumulator wraps the same way (aggregate.hpp:122-124).
__global__ void k_acc_to_value(int dtype, int agg, const ysql_acc *acc, void *v) {
  const ysql_acc a = *acc;
  if (agg == YSQL_AGG_COUNT || agg == YSQL_AGG_COUNT_STAR) *(int32_t *)v = (int32_t)(uint32_t)(uint64_t)a.count;
  else if (dtype == YSQL_INT) *(int32_t *)v = (int32_t)(uint32_t)(uint64_t)a.i;
  else if (dtype == YSQL_DATETIME) *(int64_t *)v = a.i;
  else if (agg == YSQL_AGG_SUM || agg == YSQL_AGG_AVG) *(double *)v = a.d;
  else *(float *)v = a.f;
}

// div_avg_kernel (reduction_kernel.cu:5-10): avg = (double)sum / (double)count.
__global__ void k_div_avg(int dtype, const int32_t *counter, const void *sum, float *avg) {
  const double s = dtype == YSQL_INT ? (double)*(const int32_t *)sum
                                     : (dtype == YSQL_DATETIME ? (double)*(const int64_t *)sum : *(const double *)sum);
  *avg = (float)(s / (double)*counter);
}

__global__ void k_double_to_float(const double *in, float *out) { *out = (float)*in; }

static size_t reduce_ws_size(uint64_t n) {
  WsCounter c;
  c.take<ysql_acc>(ceil_div(n, kRedTile) + 1);
  return c.off;
}

template <typename T, int KIND, int OP, bool STAR>
static int reduce_launch(const void *in, const uint8_t *nul, uint64_t n, int dtype, ysql_acc *acc,
                         ysql_acc *partials, hipStream_t s) {
  const uint32_t ntiles = (uint32_t)ceil_div(n, kRedTile);
  uint32_t g = ntiles;
  const uint32_t cap = (uint32_t)grid_cap() * 4;
  if (g > cap) g = cap;
  k_reduce<T, KIND, OP, STAR><<<g, kRedBlock, 0, s>>>((const T *)in, nul, n, dtype, partials, ntiles);
  YSQL_LAUNCH_CHECK();
  k_reduce_finish<KIND, OP><<<1, kRedBlock, 0, s>>>(dtype, acc, partials, ntiles);
  YSQL_LAUNCH_CHECK();
  return YSQL_OK;
}

template <typename T>
static int reduce_typed(int dtype, int agg, const void *in, const uint8_t *nul, uint64_t n, ysql_acc *acc,
                        ysql_acc *partials, hipStream_t s) {
  const int kind = acc_kind(dtype, agg), op = agg_op(agg);
  const bool star = agg == YSQL_AGG_COUNT_STAR;
  if (star) return reduce_launch<T, KC, 0, true>(in, nullptr, n, dtype, acc, partials, s);
  if (kind == KC) return reduce_launch<T, KC, 0, false>(in, nul, n, dtype, acc, partials, s);
  if (kind == KI) {
    if (op == 0) return reduce_launch<T, KI, 0, false>(in, nul, n, dtype, acc, partials, s);
    if (op == 1) return reduce_launch<T, KI, 1, false>(in, nul, n, dtype, acc, partials, s);
    return reduce_launch<T, KI, 2, false>(in, nul, n, dtype, acc, partials, s);
  }
  if constexpr (std::is_same<T, float>::value) {
    if (kind == KD) return reduce_launch<T, KD, 0, false>(in, nul, n, dtype, acc, partials, s);
    if (op == 1) return reduce_launch<T, KF, 1, false>(in, nul, n, dtype, acc, partials, s);
    return reduce_launch<T, KF, 2, false>(in, nul, n, dtype, acc, partials, s);
  }
  return YSQL_EINVAL;
}

// ---------------------------------------------------------------------------
// GROUP BY: LDS-privatised per-group accumulators (8-byte value + 32-bit
// count per group) when they fit, folded into the caller's accumulators with
// one global atomic per touched group per workgroup.
// ---------------------------------------------------------------------------
constexpr uint32_t kGroupLdsMax = 4096;

// float <-> order-preserving uint32 (so float MIN/MAX become integer atomics)
__device__ __forceinline__ uint32_t f2ord(float f) {
  uint32_t b = __float_as_uint(f);
  return (b & 0x80000000u) ? ~b : (b ^ 0x80000000u);
}
__device__ __forceinline__ float ord2f(uint32_t k) {
  uint32_t b = (k & 0x80000000u) ? (k ^ 0x80000000u) : ~k;
  return __uint_as_float(b);
}

__device__ __forceinline__ void atomic_fmin_global(float *p, float v) {
  unsigned int *a = (unsigned int *)p;
  unsigned int old = __hip_atomic_load(a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  while (true) {
    float cur = __uint_as_float(old);
    float nv = fminf(cur, v);
    if (__float_as_uint(nv) == old) break;
    unsigned int prev = atomicCAS(a, old, __float_as_uint(nv));
    if (prev == old) break;
    old = prev;
  }
}
__device__ __forceinline__ void atomic_fmax_global(float *p, float v) {
  unsigned int *a = (unsigned int *)p;
  unsigned int old = __hip_atomic_load(a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  while (true) {
    float cur = __uint_as_float(old);
    float nv = fmaxf(cur, v);
    if (__float_as_uint(nv) == old) break;
    unsigned int prev = atomicCAS(a, old, __float_as_uint(nv));
    if (prev == old) break;
    old = prev;
  }
}

template <int KIND, int OP>
__device__ __forceinline__ void global_fold(ysql_acc *a, uint64_t val_bits, uint64_t cnt) {
  if (cnt == 0) return;
  atomicAdd((unsigned long long *)&a->count, (unsigned long long)cnt);
  if (KIND == KI) {
    const long long v = (long long)val_bits;
    if (OP == 0) atomicAdd((unsigned long long *)&a->i, (unsigned long long)v);
    else if (OP == 1) atomicMin((long long *)&a->i, v);
    else atomicMax((long long *)&a->i, v);
  } else if (KIND == KD) {
    atomicAdd(&a->d, __longlong_as_double((long long)val_bits));
  } else if (KIND == KF) {
    const float f = ord2f((uint32_t)val_bits);
    if (OP == 1) atomic_fmin_global(&a->f, f);
    else atomic_fmax_global(&a->f, f);
  }
}

template <typename T, int KIND, int OP>
__device__ __forceinline__ uint64_t slot_identity(int dtype) {
  if (KIND == KI) {
    if (OP == 1) return (uint64_t)(dtype == YSQL_INT ? (int64_t)INT32_MAX : INT64_MAX);
    if (OP == 2) return (uint64_t)(dtype == YSQL_INT ? (int64_t)INT32_MIN : INT64_MIN);
    return 0;
  }
  if (KIND == KF) return OP == 1 ? (uint64_t)f2ord(__builtin_inff()) : (uint64_t)f2ord(-__builtin_inff());
  return 0;  // KD: bits of +0.0
}

template <typename T, int KIND, int OP>
__device__ __forceinline__ void lds_fold(unsigned long long *slot, T v) {
  if (KIND == KI) {
    const long long x = (long long)v;
    if (OP == 0) atomicAdd(slot, (unsigned long long)x);
    else if (OP == 1) atomicMin((long long *)slot, x);
    else atomicMax((long long *)slot, x);
  } else if (KIND == KD) {
    atomicAdd((double *)slot, (double)v);
  } else if (KIND == KF) {
    const unsigned long long k = f2ord((float)v);
    if (OP == 1) atomicMin(slot, k);
    else atomicMax(slot, k);
  }
}

template <typename T, int KIND, int OP>
__global__ __launch_bounds__(512) void k_group_lds(const uint32_t *gid, uint32_t ng, const T *in, const uint8_t *nul,
                                                   uint64_t n, int dtype, bool star, ysql_acc *acc) {
  __shared__ unsigned long long s_val[kGroupLdsMax];
  __shared__ uint32_t s_cnt[kGroupLdsMax];
  const uint64_t ident = slot_identity<T, KIND, OP>(dtype);
  for (uint32_t g = threadIdx.x; g < ng; g += blockDim.x) {
    s_val[g] = ident;
    s_cnt[g] = 0;
  }
  __syncthreads();
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t g = gid[i];
    if (g >= ng) continue;
    if (!star && nul && nul[i]) continue;
    atomicAdd(&s_cnt[g], 1u);
    if (KIND != KC) lds_fold<T, KIND, OP>(&s_val[g], in[i]);
  }
  __syncthreads();
  for (uint32_t g = threadIdx.x; g < ng; g += blockDim.x) global_fold<KIND, OP>(&acc[g], s_val[g], s_cnt[g]);
}

template <typename T, int KIND, int OP>
__global__ __launch_bounds__(256) void k_group_global(const uint32_t *gid, uint32_t ng, const T *in,
                                                      const uint8_t *nul, uint64_t n, bool star, ysql_acc *acc) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t g = gid[i];
    if (g >= ng) continue;
    if (!star && nul && nul[i]) continue;
    uint64_t bits = 0;
    if (KIND == KI) bits = (uint64_t)(int64_t)in[i];
    else if (KIND == KD) bits = (uint64_t)__double_as_longlong((double)in[i]);
    else if (KIND == KF) bits = f2ord((float)in[i]);
    global_fold<KIND, OP>(&acc[g], bits, 1);
  }
}

template <typename T, int KIND, int OP>
static int group_launch(const uint32_t *gid, uint32_t ng, const void *in, const uint8_t *nul, uint64_t n, int dtype,
                        bool star, ysql_acc *acc, hipStream_t s) {
  uint64_t g = ceil_div(n, 512 * 8);
  const uint64_t cap = (uint64_t)grid_cap() * 2;
  if (g > cap) g = cap;
  if (g < 1) g = 1;
  if (ng <= kGroupLdsMax)
    k_group_lds<T, KIND, OP><<<(unsigned)g, 512, 0, s>>>(gid, ng, (const T *)in, nul, n, dtype, star, acc);
  else
    k_group_global<T, KIND, OP><<<(unsigned)(g * 2), 256, 0, s>>>(gid, ng, (const T *)in, nul, n, star, acc);
  YSQL_LAUNCH_CHECK();
  return YSQL_OK;
}

template <typename T>
static int group_typed(int dtype, int agg, const uint32_t *gid, uint32_t ng, const void *in, const uint8_t *nul,
                       uint64_t n, ysql_acc *acc, hipStream_t s) {
  const int kind = acc_kind(dtype, agg), op = agg_op(agg);
  const bool star = agg == YSQL_AGG_COUNT_STAR;
  if (kind == KC) return group_launch<T, KC, 0>(gid, ng, in, nul, n, dtype, star, acc, s);
  if (kind == KI) {
    if (op == 0) return group_launch<T, KI, 0>(gid, ng, in, nul, n, dtype, star, acc, s);
    if (op == 1) return group_launch<T, KI, 1>(gid, ng, in, nul, n, dtype, star, acc, s);
    return group_launch<T, KI, 2>(gid, ng, in, nul, n, dtype, star, acc, s);
  }
  if constexpr (std::is_same<T, float>::value) {
    if (kind == KD) return group_launch<T, KD, 0>(gid, ng, in, nul, n, dtype, star, acc, s);
    if (op == 1) return group_launch<T, KF, 1>(gid, ng, in, nul, n, dtype, star, acc, s);
    return group_launch<T, KF, 2>(gid, ng, in, nul, n, dtype, star, acc, s);
  }
  return YSQL_EINVAL;
}

// ---------------------------------------------------------------------------
// Star join + grouped SUM(m1 * m2): one pass over the fact table.  The
// dimension columns (pass flags and group attributes) are small and stay in
// L2; the fact foreign keys are read in dimension order and a row stops at
// its first failing dimension, so the measures are read only for rows that
// qualify.
// ---------------------------------------------------------------------------
struct StarArgs {
  ysql_star_dim d[YSQL_MAX_DIMS];
  uint32_t stride[YSQL_MAX_DIMS];
};

constexpr uint32_t kStarLdsMax = 8192;

// 4 consecutive int32 values from row r (rows >= n read 0): one 128-bit
// load when whole and aligned.
__device__ __forceinline__ void star_load4p(const int32_t *p, uint64_t r, uint64_t n, int32_t *v) {
  if (r + 4 <= n && ((uintptr_t)(p + r) & 15) == 0) {
    const int4 q = *reinterpret_cast<const int4 *>(p + r);
    v[0] = q.x;
    v[1] = q.y;
    v[2] = q.z;
    v[3] = q.w;
  } else {
#pragma unroll
    for (int e = 0; e < 4; e++) v[e] = r + e < n ? p[r + e] : 0;
  }
}

// The values of rows r..r+3 whose flag is set (0 elsewhere): a load per
// such row only, so a column is read only in the lines holding such rows.
__device__ __forceinline__ void star_load_sel(const int32_t *p, uint64_t r, const bool *sel, int32_t *v) {
#pragma unroll
  for (int e = 0; e < 4; e++) v[e] = sel[e] ? p[r + e] : 0;
}

// Each thread takes G groups of 4 consecutive fact rows per step (group j at
// row r + j * S, S = 4 * threads in the grid, so each group's 128-bit loads
// are contiguous across the grid), so that a step's chain of dependent
// latencies (column loads, filter lookups, group lookups) covers 4G rows.
// With LATE (some dimension has a filter), only the filtered dimensions'
// foreign keys are read up front; their filter bytes are looked up (the
// rows' gathers of a dimension together), and only the rows that pass every
// filter read the other foreign keys and the two measures: at the join's
// selectivity most lines of those columns are never fetched.  Without
// filters every column is read up front.  Then the group values of the
// qualifying rows are looked up, and the rows fold into per-block LDS group
// arrays (LDS) or straight into the global ones.  The rows that qualify, and
// so every sum and count, are the same either way.
template <bool LDS, bool LATE>
__global__ __launch_bounds__(1024) void k_star(StarArgs a, int nd, const int32_t *__restrict__ m1,
                                               const int32_t *__restrict__ m2, uint64_t n, int64_t *sums,
                                               uint64_t *counts, uint32_t ng) {
  constexpr int G = 2;  // 8 rows: their keys, measures and group values fit the registers of a 1,024-thread block
  constexpr int RPT = 4 * G;
  __shared__ unsigned long long s_sum[LDS ? kStarLdsMax : 1];
  __shared__ uint32_t s_cnt[LDS ? kStarLdsMax : 1];
  if (LDS) {
    for (uint32_t g = threadIdx.x; g < ng; g += blockDim.x) {
      s_sum[g] = 0;
      s_cnt[g] = 0;
    }
    __syncthreads();
  }
  const uint64_t S = (uint64_t)gridDim.x * blockDim.x * 4;
  for (uint64_t r = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * 4; r < n; r += S * G) {
    int32_t f[YSQL_MAX_DIMS][RPT], x1[RPT], x2[RPT];
    bool ok[RPT];
#pragma unroll
    for (int e = 0; e < RPT; e++) ok[e] = r + (uint64_t)(e / 4) * S + (e % 4) < n;
#pragma unroll
    for (int d = 0; d < YSQL_MAX_DIMS; d++)
      if (d < nd && (!LATE || a.d[d].pass))
#pragma unroll
        for (int jg = 0; jg < G; jg++) star_load4p(a.d[d].fk, r + jg * S, n, &f[d][4 * jg]);
    if constexpr (!LATE) {
#pragma unroll
      for (int jg = 0; jg < G; jg++) {
        star_load4p(m1, r + jg * S, n, &x1[4 * jg]);
        star_load4p(m2, r + jg * S, n, &x2[4 * jg]);
      }
    }
    // in-range checks of the keys read so far, then all gathers
#pragma unroll
    for (int d = 0; d < YSQL_MAX_DIMS; d++)
      if (d < nd && (!LATE || a.d[d].pass)) {
#pragma unroll
        for (int e = 0; e < RPT; e++) ok[e] = ok[e] && f[d][e] >= 0 && (uint32_t)f[d][e] < a.d[d].rows;
      }
    // every dimension's filter first, then the group lookups of the rows
    // that pass them all; the rows' gathers of one dimension go together
#pragma unroll
    for (int d = 0; d < YSQL_MAX_DIMS; d++) {
      if (d >= nd || !a.d[d].pass) continue;
      uint8_t pv[RPT];
#pragma unroll
      for (int e = 0; e < RPT; e++) pv[e] = ok[e] ? a.d[d].pass[f[d][e]] : (uint8_t)0;
#pragma unroll
      for (int e = 0; e < RPT; e++) ok[e] = ok[e] && pv[e] != 0;
    }
    if constexpr (LATE) {  // the rest of the row, for the rows still qualifying
#pragma unroll
      for (int d = 0; d < YSQL_MAX_DIMS; d++)
        if (d < nd && !a.d[d].pass) {
#pragma unroll
          for (int jg = 0; jg < G; jg++) star_load_sel(a.d[d].fk, r + jg * S, &ok[4 * jg], &f[d][4 * jg]);
#pragma unroll
          for (int e = 0; e < RPT; e++) ok[e] = ok[e] && f[d][e] >= 0 && (uint32_t)f[d][e] < a.d[d].rows;
        }
#pragma unroll
      for (int jg = 0; jg < G; jg++) {
        star_load_sel(m1, r + jg * S, &ok[4 * jg], &x1[4 * jg]);
        star_load_sel(m2, r + jg * S, &ok[4 * jg], &x2[4 * jg]);
      }
    }
    uint32_t g[RPT];
#pragma unroll
    for (int e = 0; e < RPT; e++) g[e] = 0;
#pragma unroll
    for (int d = 0; d < YSQL_MAX_DIMS; d++) {
      if (d >= nd || !a.d[d].group) continue;
      int32_t gv[RPT];
#pragma unroll
      for (int e = 0; e < RPT; e++) gv[e] = ok[e] ? a.d[d].group[f[d][e]] : a.d[d].group_base;
#pragma unroll
      for (int e = 0; e < RPT; e++) {
        const int64_t v = (int64_t)gv[e] - a.d[d].group_base;
        ok[e] = ok[e] && v >= 0 && v < (int64_t)a.d[d].group_card;
        g[e] += (uint32_t)v * a.stride[d];
      }
    }
#pragma unroll
    for (int e = 0; e < RPT; e++) {
      const bool q = ok[e];
      if (!q) continue;
      const long long prod = (long long)x1[e] * (long long)x2[e];
      if (LDS) {
        atomicAdd(&s_sum[g[e]], (unsigned long long)prod);
        atomicAdd(&s_cnt[g[e]], 1u);
      } else {
        atomicAdd((unsigned long long *)&sums[g[e]], (unsigned long long)prod);
        atomicAdd((unsigned long long *)&counts[g[e]], 1ull);
      }
    }
  }
  if (LDS) {
    __syncthreads();
    for (uint32_t g = threadIdx.x; g < ng; g += blockDim.x) {
      if (s_cnt[g]) {
        atomicAdd((unsigned long long *)&sums[g], s_sum[g]);
        atomicAdd((unsigned long long *)&counts[g], (unsigned long long)s_cnt[g]);
      }
    }
  }
}

// Several star-join queries sharing one read of the facts (the serving
// loop's batching): they differ only in their dimension filters, so bit q
// of a dimension's pass byte is query q's filter.  A row's group value is
// the same for every query; the row folds into the group arrays of each
// query whose bit survives (sums[q * ng + g], counts[q * ng + g]).
// Integer folds commute, so every query's result equals its lone run.
__global__ __launch_bounds__(1024) void k_star_multi(StarArgs a, int nd, uint32_t qmask,
                                                     const int32_t *__restrict__ m1, const int32_t *__restrict__ m2,
                                                     uint64_t n, int64_t *sums, uint64_t *counts, uint32_t ng) {
  constexpr int G = 2;  // groups of 4 rows per thread and step, as in k_star
  constexpr int RPT = 4 * G;
  const uint64_t S = (uint64_t)gridDim.x * blockDim.x * 4;
  for (uint64_t r = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * 4; r < n; r += S * G) {
    int32_t f[YSQL_MAX_DIMS][RPT], x1[RPT], x2[RPT];
#pragma unroll
    for (int d = 0; d < YSQL_MAX_DIMS; d++)
      if (d < nd)
#pragma unroll
        for (int jg = 0; jg < G; jg++) star_load4p(a.d[d].fk, r + jg * S, n, &f[d][4 * jg]);
#pragma unroll
    for (int jg = 0; jg < G; jg++) {
      star_load4p(m1, r + jg * S, n, &x1[4 * jg]);
      star_load4p(m2, r + jg * S, n, &x2[4 * jg]);
    }
    uint32_t mk[RPT];
#pragma unroll
    for (int e = 0; e < RPT; e++) mk[e] = r + (uint64_t)(e / 4) * S + (e % 4) < n ? qmask : 0u;
#pragma unroll
    for (int d = 0; d < YSQL_MAX_DIMS; d++)
      if (d < nd) {
#pragma unroll
        for (int e = 0; e < RPT; e++)
          if (!(f[d][e] >= 0 && (uint32_t)f[d][e] < a.d[d].rows)) mk[e] = 0;
      }
#pragma unroll
    for (int d = 0; d < YSQL_MAX_DIMS; d++) {
      if (d >= nd || !a.d[d].pass) continue;
      uint32_t pv[RPT];
#pragma unroll
      for (int e = 0; e < RPT; e++) pv[e] = mk[e] ? (uint32_t)a.d[d].pass[f[d][e]] : 0u;
#pragma unroll
      for (int e = 0; e < RPT; e++) mk[e] &= pv[e];
    }
    uint32_t g[RPT];
#pragma unroll
    for (int e = 0; e < RPT; e++) g[e] = 0;
#pragma unroll
    for (int d = 0; d < YSQL_MAX_DIMS; d++) {
      if (d >= nd || !a.d[d].group) continue;
      int32_t gv[RPT];
#pragma unroll
      for (int e = 0; e < RPT; e++) gv[e] = mk[e] ? a.d[d].group[f[d][e]] : a.d[d].group_base;
#pragma unroll
      for (int e = 0; e < RPT; e++) {
        const int64_t v = (int64_t)gv[e] - a.d[d].group_base;
        if (!(v >= 0 && v < (int64_t)a.d[d].group_card)) mk[e] = 0;
        g[e] += (uint32_t)v * a.stride[d];
      }
    }
#pragma unroll
    for (int e = 0; e < RPT; e++) {
      uint32_t m = mk[e];
      if (!m) continue;
      const long long prod = (long long)x1[e] * (long long)x2[e];
      while (m) {
        const uint32_t q = (uint32_t)__builtin_ctz(m);
        m &= m - 1;
        atomicAdd((unsigned long long *)&sums[(uint64_t)q * ng + g[e]], (unsigned long long)prod);
        atomicAdd((unsigned long long *)&counts[(uint64_t)q * ng + g[e]], 1ull);
      }
    }
  }
}

struct GidArgs {
  const int32_t *col[4];
  int32_t base[4];
  uint32_t card[4];
  uint32_t stride[4];
};

__global__ void k_group_ids(GidArgs a, int nc, uint64_t n, uint32_t *gid) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t g = 0;
    bool ok = true;
#pragma unroll
    for (int c = 0; c < 4; c++) {
      if (c < nc) {
        const int64_t v = (int64_t)a.col[c][i] - a.base[c];
        if (v < 0 || v >= (int64_t)a.card[c]) ok = false;
        g += (uint32_t)v * a.stride[c];
      }
    }
    gid[i] = ok ? g : 0xFFFFFFFFu;
  }
}

}  // namespace ysql

using namespace ysql;

extern "C" {

int ysql_group_ids(const int32_t *const *cols, const int32_t *base, const uint32_t *card, int ncols, uint64_t n,
                   uint32_t *gid_out, ysql_stream_t stream) {
  if (ncols < 1 || ncols > 4 || !cols || !base || !card) return YSQL_EINVAL;
  if (n == 0) return YSQL_OK;
  if (!gid_out) return YSQL_EINVAL;
  GidArgs a;
  memset(&a, 0, sizeof(a));
  uint64_t stride = 1;
  for (int c = ncols - 1; c >= 0; c--) {
    if (!cols[c] || card[c] == 0) return YSQL_EINVAL;
    a.col[c] = cols[c];
    a.base[c] = base[c];
    a.card[c] = card[c];
    a.stride[c] = (uint32_t)stride;
    stride *= card[c];
    if (stride >= 0xFFFFFFFFull) return YSQL_EINVAL;
  }
  uint64_t g = ceil_div(n, 256 * 4);
  const uint64_t cap = (uint64_t)grid_cap() * 4;
  if (g > cap) g = cap;
  k_group_ids<<<(unsigned)g, 256, 0, as_stream(stream)>>>(a, ncols, n, gid_out);
  YSQL_LAUNCH_CHECK();
  return YSQL_OK;
}

int ysql_acc_init(int dtype, int agg, ysql_acc *acc, int nacc, ysql_stream_t stream) {
  if (nacc <= 0) return YSQL_OK;
  if (!acc || agg < YSQL_AGG_MAX || agg > YSQL_AGG_COUNT_STAR) return YSQL_EINVAL;
  k_acc_init<<<(nacc + 255) / 256, 256, 0, as_stream(stream)>>>(acc, nacc, dtype, agg);
  YSQL_LAUNCH_CHECK();
  return YSQL_OK;
}

int ysql_reduce_ws_bytes(uint64_t n, size_t *bytes) {
  if (!bytes) return YSQL_EINVAL;
  *bytes = reduce_ws_size(n);
  return YSQL_OK;
}

int ysql_reduce(int dtype, int agg, const void *in, const uint8_t *null_opt, uint64_t n, ysql_acc *acc_inout, void *ws,
                size_t ws_bytes, ysql_stream_t stream) {
  if (!acc_inout || agg < YSQL_AGG_MAX || agg > YSQL_AGG_COUNT_STAR) return YSQL_EINVAL;
  if (n == 0) return YSQL_OK;
  if (!in && agg != YSQL_AGG_COUNT_STAR && agg != YSQL_AGG_COUNT) return YSQL_EINVAL;
  if (n >= (1ull << 40)) return YSQL_EINVAL;
  WsCarver c(ws, ws_bytes);
  ysql_acc *partials = c.take<ysql_acc>(ceil_div(n, kRedTile) + 1);
  if (!ws || !partials) return YSQL_EWORKSPACE;
  hipStream_t s = as_stream(stream);
  if (agg == YSQL_AGG_COUNT || agg == YSQL_AGG_COUNT_STAR) {
    // counts read only the null mask (reduction_count_notnull reads only isnull)
    return reduce_typed<int32_t>(YSQL_INT, agg, nullptr, null_opt, n, acc_inout, partials, s);
  }
  switch (dtype) {
    case YSQL_INT: return reduce_typed<int32_t>(dtype, agg, in, null_opt, n, acc_inout, partials, s);
    case YSQL_FLOAT: return reduce_typed<float>(dtype, agg, in, null_opt, n, acc_inout, partials, s);
    case YSQL_DATETIME: return reduce_typed<int64_t>(dtype, agg, in, null_opt, n, acc_inout, partials, s);
  }
  return YSQL_EINVAL;
}

int ysql_agg_finalize(int dtype, int agg, const ysql_acc *acc, void *out, ysql_stream_t stream) {
  if (!acc || !out || agg < YSQL_AGG_MAX || agg > YSQL_AGG_COUNT_STAR) return YSQL_EINVAL;
  k_finalize<<<1, 1, 0, as_stream(stream)>>>(dtype, agg, acc, out);
  YSQL_LAUNCH_CHECK();
  return YSQL_OK;
}

static bool acc_value_ok(int dtype, int agg) {
  return agg >= YSQL_AGG_MAX && agg <= YSQL_AGG_COUNT_STAR &&
         (dtype == YSQL_INT || dtype == YSQL_FLOAT || dtype == YSQL_DATETIME);
}

int ysql_acc_from_value(int dtype, int agg, const void *value, ysql_acc *acc, ysql_stream_t stream) {
  if (!value || !acc || !acc_value_ok(dtype, agg)) return YSQL_EINVAL;
  k_acc_from_value<<<1, 1, 0, as_stream(stream)>>>(dtype, agg, value, acc);
  YSQL_LAUNCH_CHECK();
  return YSQL_OK;
}

int ysql_acc_to_value(int dtype, int agg, const ysql_acc *acc, void *value, ysql_stream_t stream) {
  if (!value || !acc || !acc_value_ok(dtype, agg)) return YSQL_EINVAL;
  k_acc_to_value<<<1, 1, 0, as_stream(stream)>>>(dtype, agg, acc, value);
  YSQL_LAUNCH_CHECK();
  return YSQL_OK;
}

int ysql_div_avg(int dtype, const int32_t *counter, const void *sum, float *avg, ysql_stream_t stream) {
  if (!counter || !sum || !avg || !acc_value_ok(dtype, YSQL_AGG_AVG)) return YSQL_EINVAL;
  k_div_avg<<<1, 1, 0, as_stream(stream)>>>(dtype, counter, sum, avg);
  YSQL_LAUNCH_CHECK();
  return YSQL_OK;
}

int ysql_double_to_float(const double *in, float *out, ysql_stream_t stream) {
  if (!in || !out) return YSQL_EINVAL;
  k_double_to_float<<<1, 1, 0, as_stream(stream)>>>(in, out);
  YSQL_LAUNCH_CHECK();
  return YSQL_OK;
}

int ysql_group_reduce(int dtype, int agg, const uint32_t *group_id, uint32_t n_groups, const void *in,
                      const uint8_t *null_opt, uint64_t n, ysql_acc *acc_inout, ysql_stream_t stream) {
  if (n == 0 || n_groups == 0) return YSQL_OK;
  if (!group_id || !acc_inout || agg < YSQL_AGG_MAX || agg > YSQL_AGG_COUNT_STAR) return YSQL_EINVAL;
  if (!in && agg != YSQL_AGG_COUNT_STAR) return YSQL_EINVAL;
  hipStream_t s = as_stream(stream);
  const void *src = in ? in : (const void *)group_id;
  switch (dtype) {
    case YSQL_INT: return group_typed<int32_t>(dtype, agg, group_id, n_groups, src, null_opt, n, acc_inout, s);
    case YSQL_FLOAT: return group_typed<float>(dtype, agg, group_id, n_groups, src, null_opt, n, acc_inout, s);
    case YSQL_DATETIME: return group_typed<int64_t>(dtype, agg, group_id, n_groups, src, null_opt, n, acc_inout, s);
  }
  return YSQL_EINVAL;
}

static int star_args(const ysql_star_dim *dims, int ndims, uint32_t n_groups, StarArgs *out) {
  StarArgs &a = *out;
  memset(&a, 0, sizeof(a));
  uint64_t stride = 1;
  for (int d = ndims - 1; d >= 0; d--) {
    a.d[d] = dims[d];
    if (!dims[d].fk) return YSQL_EINVAL;
    a.stride[d] = (uint32_t)stride;
    if (dims[d].group) {
      if (dims[d].group_card == 0) return YSQL_EINVAL;
      stride *= dims[d].group_card;
    }
  }
  return stride > n_groups ? YSQL_EINVAL : YSQL_OK;
}

// Whether some dimension filters the rows (then the other columns are read
// late, for the rows that pass).
static bool star_has_filter(const ysql_star_dim *dims, int ndims) {
  for (int d = 0; d < ndims; d++)
    if (dims[d].pass) return true;
  return false;
}

int ysql_star_join_group_sum_multi(const ysql_star_dim *dims, int ndims, int nq, const int32_t *m1, const int32_t *m2,
                                   uint64_t n, int64_t *sums, uint64_t *counts, uint32_t n_groups,
                                   ysql_stream_t stream) {
  if (ndims < 1 || ndims > YSQL_MAX_DIMS || !dims || !m1 || !m2 || !sums || !counts) return YSQL_EINVAL;
  if (n_groups == 0 || n_groups > YSQL_MAX_GROUPS || nq < 1 || nq > 8) return YSQL_EINVAL;
  StarArgs a;
  if (int rc = star_args(dims, ndims, n_groups, &a)) return rc;
  if (n == 0) return YSQL_OK;
  hipStream_t s = as_stream(stream);
  uint64_t g = ceil_div(n, 1024 * 16);
  const uint64_t cap = (uint64_t)grid_cap() * 2;
  if (g > cap) g = cap;
  if (g < 1) g = 1;
  k_star_multi<<<(unsigned)g, 1024, 0, s>>>(a, ndims, (1u << nq) - 1u, m1, m2, n, sums, counts, n_groups);
  YSQL_LAUNCH_CHECK();
  return YSQL_OK;
}

int ysql_star_join_group_sum(const ysql_star_dim *dims, int ndims, const int32_t *m1, const int32_t *m2, uint64_t n,
                             int64_t *sums, uint64_t *counts, uint32_t n_groups, ysql_stream_t stream) {
  if (ndims < 1 || ndims > YSQL_MAX_DIMS || !dims || !m1 || !m2 || !sums || !counts) return YSQL_EINVAL;
  if (n_groups == 0 || n_groups > YSQL_MAX_GROUPS) return YSQL_EINVAL;
  StarArgs a;
  if (int rc = star_args(dims, ndims, n_groups, &a)) return rc;
  if (n == 0) return YSQL_OK;
  hipStream_t s = as_stream(stream);
  uint64_t g = ceil_div(n, 1024 * 16);
  const uint64_t cap = (uint64_t)grid_cap() * (n_groups <= kStarLdsMax ? 1 : 2);
  if (g > cap) g = cap;
  if (g < 1) g = 1;
  const bool late = star_has_filter(dims, ndims);
  if (n_groups <= kStarLdsMax) {
    if (late) k_star<true, true><<<(unsigned)g, 1024, 0, s>>>(a, ndims, m1, m2, n, sums, counts, n_groups);
    else k_star<true, false><<<(unsigned)g, 1024, 0, s>>>(a, ndims, m1, m2, n, sums, counts, n_groups);
  } else {
    if (late) k_star<false, true><<<(unsigned)g, 1024, 0, s>>>(a, ndims, m1, m2, n, sums, counts, n_groups);
    else k_star<false, false><<<(unsigned)g, 1024, 0, s>>>(a, ndims, m1, m2, n, sums, counts, n_groups);
  }
  YSQL_LAUNCH_CHECK();
  return YSQL_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Filtered multi-aggregate (agg_dev.hpp), per-call form.
// ---------------------------------------------------------------------------
#include "agg_dev.hpp"

namespace ysql {

// QMAX = 1 for single queries (fewer registers, higher occupancy), 4 for
// shared scans; a query's fold sequence is the same in both.  The
// comparison is the range test of fa2_range (exact for every operator).
template <typename PT, int QMAX>
__global__ __launch_bounds__(kFaBlock) void k_filter_agg(const PT *pred, const PT *scalars, int cmp, int nq, FaArgs a,
                                                         int ns, uint64_t n, uint64_t tile_rows, ysql_acc *partials,
                                                         uint32_t ntiles) {
  __shared__ ysql_acc s_w[kFaBlock / kWave];
  const FaSelRange<PT, QMAX> sel = fa_sel_range<PT, QMAX>(cmp, scalars, nq);
  for (uint32_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x)
    filter_agg_tile<PT, FaSelRange<PT, QMAX>, kFaBlock, QMAX>(tile, tile_rows, pred, sel, nq, a.spec, ns, n, partials,
                                                             s_w);
}

template <typename K>
static uint32_t fa_blocks_per_cu(K kernel) {
  int b = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, kernel, kFaBlock, 0) != hipSuccess || b < 1) b = 2;
  return (uint32_t)b;
}

// One workgroup per (query, spec): its tile partials combined in a fixed
// order into its output (a second launch on the stream, for the reason
// given at k_reduce).
__global__ __launch_bounds__(kFaBlock) void k_fa_finish(FaArgs a, int ns, int nq, uint32_t ntiles,
                                                        const ysql_acc *partials, int overwrite) {
  __shared__ ysql_acc s_w[kFaBlock / kWave];
  const int q = (int)blockIdx.x / ns, sp = (int)blockIdx.x % ns;
  filter_agg_finish_one<kFaBlock>(a.spec, ns, nq, q, sp, ntiles, partials, a.out, overwrite != 0, s_w);
}

static int fa_finish_launch(const FaArgs &a, int ns, int nq, uint32_t ntiles, const ysql_acc *partials, int overwrite,
                            hipStream_t s) {
  k_fa_finish<<<(unsigned)(nq * ns), kFaBlock, 0, s>>>(a, ns, nq, ntiles, partials, overwrite);
  YSQL_LAUNCH_CHECK();
  return YSQL_OK;
}

// The two-column form (agg_dev.hpp, filter_agg2_tile): slot 1 holds W1-byte
// elements (0: no column besides the predicate).
struct Fa2Args {
  int code[kFaMaxSpecs];
  const void *col1;
  int cmp;
  int aligned;
};

template <typename PT, int W1, int QMAX>
__global__ __launch_bounds__(kFaBlock) void k_filter_agg2(const PT *pred, const PT *scalars, int nq, FaArgs a,
                                                          Fa2Args b, int ns, uint64_t n, uint64_t tile_rows,
                                                          ysql_acc *partials, uint32_t ntiles) {
  __shared__ ysql_acc s_w[kFaBlock / kWave];
  const FaSelRange<PT, QMAX> sel = fa_sel_range<PT, QMAX>(b.cmp, scalars, nq);
  for (uint32_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x)
    filter_agg2_tile<PT, FaSelRange<PT, QMAX>, W1, kFaBlock, QMAX>(tile, tile_rows, pred, b.col1, sel, nq, a.spec,
                                                                  b.code, ns, n, b.aligned != 0, partials, s_w);
}

template <typename PT, int W1, int QMAX>
static int filter_agg2_launch(const void *pred, const void *scalars, int nq, const FaArgs &a, const Fa2Args &b,
                              int ns, uint64_t n, ysql_acc *partials, int overwrite,
                              hipStream_t s) {
  const uint64_t tile_rows = fa_tile_rows(n);
  const uint32_t ntiles = (uint32_t)ceil_div(n, tile_rows);
  static const uint32_t per_cu = fa_blocks_per_cu(k_filter_agg2<PT, W1, QMAX>);
  const uint32_t g = std::min<uint32_t>(ntiles, per_cu * (uint32_t)grid_cap());
  k_filter_agg2<PT, W1, QMAX><<<g, kFaBlock, 0, s>>>((const PT *)pred, (const PT *)scalars, nq, a, b, ns, n,
                                                      tile_rows, partials, ntiles);
  YSQL_LAUNCH_CHECK();
  return fa_finish_launch(a, ns, nq, ntiles, partials, overwrite, s);
}

// Common shapes compiled whole (agg_dev.hpp, filter_agg3_tile): the spec
// codes (fa2_code: opcode * 2 + slot) are template arguments.
// Waves per SIMD the compiled shapes are built for: the register budget at
// which each compiles without spilling (measured per shape; one more wave
// spills the four-query forms).
template <typename PT, int QMAX>
struct Fa3Waves {
  static constexpr int value = QMAX == 1 ? 7 : (sizeof(PT) == 4 ? 5 : 4);
};

template <typename PT, int W1, int QMAX, int C0, int C1, int C2, int C3>
__global__ __launch_bounds__(kFaBlock) __attribute__((amdgpu_waves_per_eu(Fa3Waves<PT, QMAX>::value)))
void k_filter_agg3(const PT *pred, const PT *scalars, int nq, FaArgs a,
                                                          Fa2Args b, int ns, uint64_t n, uint64_t tile_rows,
                                                          ysql_acc *partials, uint32_t ntiles) {
  __shared__ ysql_acc s_w[kFaBlock / kWave];
  const FaSelRange<PT, QMAX> sel = fa_sel_range<PT, QMAX>(b.cmp, scalars, nq);
  for (uint32_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x)
    filter_agg3_tile<PT, FaSelRange<PT, QMAX>, W1, kFaBlock, QMAX, C0, C1, C2, C3>(
        tile, tile_rows, pred, b.col1, sel, nq, a.spec, ns, n, b.aligned != 0, partials, s_w);
}

template <typename PT, int W1, int QMAX, int C0, int C1, int C2, int C3>
static int filter_agg3_launch(const void *pred, const void *scalars, int nq, const FaArgs &a, const Fa2Args &b,
                              int ns, uint64_t n, ysql_acc *partials, int overwrite,
                              hipStream_t s) {
  const uint64_t tile_rows = fa_tile_rows(n);
  const uint32_t ntiles = (uint32_t)ceil_div(n, tile_rows);
  static const uint32_t per_cu = fa_blocks_per_cu(k_filter_agg3<PT, W1, QMAX, C0, C1, C2, C3>);
  const uint32_t g = std::min<uint32_t>(ntiles, per_cu * (uint32_t)grid_cap());
  k_filter_agg3<PT, W1, QMAX, C0, C1, C2, C3><<<g, kFaBlock, 0, s>>>(
      (const PT *)pred, (const PT *)scalars, nq, a, b, ns, n, tile_rows, partials, ntiles);
  YSQL_LAUNCH_CHECK();
  return fa_finish_launch(a, ns, nq, ntiles, partials, overwrite, s);
}

// The shapes compiled whole: COUNT(*), SUM(x), MIN(y), MAX(y) with the
// predicate a float column and y an int64 column (the summary of a filtered
// range), and with the predicate an int64 key whose MIN/MAX are taken and x
// an int32 column (the serving loop's filtered aggregate).  Returns
// YSQL_EUNSUPPORTED for any other shape.
static int filter_agg3_try(int pred_dtype, int w1, int nq, const void *pred, const void *scalars, const FaArgs &a,
                           const Fa2Args &b, int ns, uint64_t n, ysql_acc *partials, int overwrite,
                           hipStream_t s) {
  if (ns != 4) return YSQL_EUNSUPPORTED;
  const int *c = b.code;
#define FA3_SHAPE(DT, PT, W, C0, C1, C2, C3)                                                                   \
  if (pred_dtype == DT && w1 == W && c[0] == C0 && c[1] == C1 && c[2] == C2 && c[3] == C3)                   \
    return nq == 1 ? filter_agg3_launch<PT, W, 1, C0, C1, C2, C3>(pred, scalars, nq, a, b, ns, n, partials, \
                                                                 overwrite, s)                                     \
                   : filter_agg3_launch<PT, W, kFaMaxQueries, C0, C1, C2, C3>(pred, scalars, nq, a, b, ns, n,        \
                                                                             partials, overwrite, s);
  // COUNT(*), SUM(float predicate), MIN(int64 y), MAX(int64 y)
  FA3_SHAPE(YSQL_FLOAT, float, 8, 0, FA_OPC_D_SUM * 2, FA_OPC_I64_MIN * 2 + 1, FA_OPC_I64_MAX * 2 + 1)
  // COUNT(*), SUM(int32 x), MIN(int64 predicate), MAX(int64 predicate)
  FA3_SHAPE(YSQL_DATETIME, int64_t, 4, 0, FA_OPC_I32_SUM * 2 + 1, FA_OPC_I64_MIN * 2, FA_OPC_I64_MAX * 2)
#undef FA3_SHAPE
  return YSQL_EUNSUPPORTED;
}

template <typename PT, int W1>
static int filter_agg2_dispatch(const void *pred, const void *scalars, int nq, const FaArgs &a, const Fa2Args &b,
                                int ns, uint64_t n, ysql_acc *partials, int overwrite,
                                hipStream_t s) {
  return nq == 1 ? filter_agg2_launch<PT, W1, 1>(pred, scalars, nq, a, b, ns, n, partials, overwrite, s)
                 : filter_agg2_launch<PT, W1, kFaMaxQueries>(pred, scalars, nq, a, b, ns, n, partials,
                                                             overwrite, s);
}

// Identities into the outputs (an empty input with overwrite).
__global__ void k_fa_identity(FaArgs a, int ns, int nq) {
  const int k = threadIdx.x;
  if (k < ns * nq) {
    ysql_acc v;
    fa_identity(v, a.spec[k % ns]);
    *a.out[k] = v;
  }
}


template <typename PT, int QMAX>
static int filter_agg_launch(int cmp, const void *pred, const void *scalars, int nq, const FaArgs &a, int ns,
                             uint64_t n, ysql_acc *partials, int overwrite, hipStream_t s) {
  if (cmp < YSQL_LT || cmp > YSQL_NE) return YSQL_EINVAL;
  const uint64_t tile_rows = fa_tile_rows(n);
  const uint32_t ntiles = (uint32_t)ceil_div(n, tile_rows);
  // grid: the workgroups that fit at once (tiles are taken grid-stride)
  static const uint32_t per_cu = fa_blocks_per_cu(k_filter_agg<PT, QMAX>);
  const uint32_t g = std::min<uint32_t>(ntiles, per_cu * (uint32_t)grid_cap());
  k_filter_agg<PT, QMAX><<<g, kFaBlock, 0, s>>>((const PT *)pred, (const PT *)scalars, cmp, nq, a, ns, n, tile_rows,
                                                 partials, ntiles);
  YSQL_LAUNCH_CHECK();
  return fa_finish_launch(a, ns, nq, ntiles, partials, overwrite, s);
}

static size_t filter_agg_ws_size(uint64_t n, int nq, int ns) {
  WsCounter c;
  c.take<ysql_acc>(ceil_div(n, fa_tile_rows(n)) * (uint64_t)nq * ns + 1);
  return c.off;
}

// nq queries; specs[q * nspecs + s]: every query has the same columns,
// dtypes and aggregates (only the accumulators differ).
static int filter_agg_impl(int pred_dtype, int cmp, const void *pred_col, const void *scalars, int nq,
                           const ysql_agg_spec *specs, int nspecs, uint64_t n, int flags, void *ws, size_t ws_bytes,
                           hipStream_t s) {
  if (nspecs < 1 || nspecs > kFaMaxSpecs || nq < 1 || nq > kFaMaxQueries || !specs || cmp < YSQL_LT ||
      cmp > YSQL_NE)
    return YSQL_EINVAL;
  if (flags & ~(YSQL_FA_OVERWRITE | YSQL_FA_GENERIC)) return YSQL_EINVAL;
  const int overwrite = (flags & YSQL_FA_OVERWRITE) ? 1 : 0;
  FaArgs a;
  memset(&a, 0, sizeof(a));
  for (int k = 0; k < nspecs; k++) {
    const ysql_agg_spec &x = specs[k];
    if (x.agg < YSQL_AGG_MAX || x.agg > YSQL_AGG_COUNT_STAR) return YSQL_EINVAL;
    const bool counting = x.agg == YSQL_AGG_COUNT || x.agg == YSQL_AGG_COUNT_STAR;
    if (!counting && x.dtype != YSQL_INT && x.dtype != YSQL_FLOAT && x.dtype != YSQL_DATETIME) return YSQL_EINVAL;
    if (!counting && !x.col && n > 0) return YSQL_EINVAL;  // an empty input reads no column
    a.spec[k].kind = fa_kind(x.dtype, x.agg);
    a.spec[k].op = fa_op(x.agg);
    a.spec[k].dtype = x.dtype;
    a.spec[k].col = x.col;
    // a spec over the predicate column, or over the column of an earlier
    // spec of the same width, reuses those loaded values
    a.spec[k].src = FA_OWN;
    if (x.col && x.col == pred_col && x.dtype == pred_dtype) a.spec[k].src = FA_PRED;
    for (int t = 0; t < k && a.spec[k].src == FA_OWN && !counting; t++)
      if (a.spec[t].src == FA_OWN && a.spec[t].kind != FA_KC && a.spec[t].col == x.col &&
          fa_wide(a.spec[t]) == fa_wide(a.spec[k]))
        a.spec[k].src = FA_SPEC0 + t;  // same column, same width: same words
  }
  for (int q = 0; q < nq; q++) {
    for (int k = 0; k < nspecs; k++) {
      const ysql_agg_spec &x = specs[q * nspecs + k];
      const ysql_agg_spec &x0 = specs[k];
      if (!x.acc || x.dtype != x0.dtype || x.agg != x0.agg || x.col != x0.col) return YSQL_EINVAL;
      a.out[q * nspecs + k] = x.acc;
    }
  }
  if (n == 0) {
    if (overwrite) {
      k_fa_identity<<<1, kWave, 0, s>>>(a, nspecs, nq);
      YSQL_LAUNCH_CHECK();
    }
    return YSQL_OK;
  }
  if (!pred_col || !scalars) return YSQL_EINVAL;
  if (n >= (1ull << 40)) return YSQL_EINVAL;
  WsCarver c(ws, ws_bytes);
  ysql_acc *partials = c.take<ysql_acc>(ceil_div(n, fa_tile_rows(n)) * (uint64_t)nq * nspecs + 1);
  if (!ws || !partials) return YSQL_EWORKSPACE;
  // the two-column form, when every spec reads the predicate column, one
  // other column or nothing
  if (!(flags & YSQL_FA_GENERIC)) {
    Fa2Args b;
    memset(&b, 0, sizeof(b));
    b.cmp = cmp;
    int w1 = 0;
    bool ok = true;
    const int pb = dtype_bytes(pred_dtype);
    for (int k = 0; k < nspecs && ok; k++) {
      const FaSpec &x = a.spec[k];
      const int opc = fa_opcode(x);
      const int wb = fa_opc_wide(opc) ? 8 : 4;
      if (opc == FA_OPC_NONE) {
        b.code[k] = fa2_code(opc, 0);
      } else if (x.src == FA_PRED && wb == pb) {
        b.code[k] = fa2_code(opc, 0);
      } else if (b.col1 == nullptr || (b.col1 == x.col && w1 == wb)) {
        b.col1 = x.col;
        w1 = wb;
        b.code[k] = fa2_code(opc, 1);
      } else {
        ok = false;
      }
    }
    if (ok) {
      b.aligned = ((uintptr_t)pred_col & (2 * pb - 1)) == 0 && (w1 == 0 || ((uintptr_t)b.col1 & (2 * w1 - 1)) == 0);
      const int rc3 = filter_agg3_try(pred_dtype, w1, nq, pred_col, scalars, a, b, nspecs, n, partials,
                                      overwrite, s);
      if (rc3 != YSQL_EUNSUPPORTED) return rc3;
#define FA2_DT(PT)                                                                                                 \
  return w1 == 8   ? filter_agg2_dispatch<PT, 8>(pred_col, scalars, nq, a, b, nspecs, n, partials, overwrite, s) \
         : w1 == 4 ? filter_agg2_dispatch<PT, 4>(pred_col, scalars, nq, a, b, nspecs, n, partials, overwrite, s) \
                   : filter_agg2_dispatch<PT, 0>(pred_col, scalars, nq, a, b, nspecs, n, partials, overwrite, s);
      switch (pred_dtype) {
        case YSQL_INT: FA2_DT(int32_t)
        case YSQL_FLOAT: FA2_DT(float)
        case YSQL_DATETIME: FA2_DT(int64_t)
      }
#undef FA2_DT
      return YSQL_EINVAL;
    }
  }
#define FA_DT(PT)                                                                                              \
  return nq == 1 ? filter_agg_launch<PT, 1>(cmp, pred_col, scalars, nq, a, nspecs, n, partials, overwrite, s) \
                 : filter_agg_launch<PT, kFaMaxQueries>(cmp, pred_col, scalars, nq, a, nspecs, n, partials,    \
                                                        overwrite, s);
  switch (pred_dtype) {
    case YSQL_INT: FA_DT(int32_t)
    case YSQL_FLOAT: FA_DT(float)
    case YSQL_DATETIME: FA_DT(int64_t)
  }
#undef FA_DT
  return YSQL_EINVAL;
}

}  // namespace ysql

extern "C" {

int ysql_filter_aggregate_ws_bytes(uint64_t n, int nspecs, size_t *bytes) {
  if (!bytes || nspecs < 1 || nspecs > kFaMaxSpecs) return YSQL_EINVAL;
  *bytes = filter_agg_ws_size(n, 1, nspecs);
  return YSQL_OK;
}

int ysql_filter_aggregate_ex(int pred_dtype, int cmp, const void *pred_col, const void *scalar,
                             const ysql_agg_spec *specs, int nspecs, uint64_t n, int flags, void *ws, size_t ws_bytes,
                             ysql_stream_t stream) {
  return filter_agg_impl(pred_dtype, cmp, pred_col, scalar, 1, specs, nspecs, n, flags, ws, ws_bytes,
                         as_stream(stream));
}

int ysql_filter_aggregate(int pred_dtype, int cmp, const void *pred_col, const void *scalar, const ysql_agg_spec *specs,
                          int nspecs, uint64_t n, void *ws, size_t ws_bytes, ysql_stream_t stream) {
  return filter_agg_impl(pred_dtype, cmp, pred_col, scalar, 1, specs, nspecs, n, 0, ws, ws_bytes, as_stream(stream));
}

int ysql_filter_aggregate_multi_ws_bytes(uint64_t n, int nqueries, int nspecs, size_t *bytes) {
  if (!bytes || nspecs < 1 || nspecs > kFaMaxSpecs || nqueries < 1 || nqueries > kFaMaxQueries) return YSQL_EINVAL;
  *bytes = filter_agg_ws_size(n, nqueries, nspecs);
  return YSQL_OK;
}

int ysql_filter_aggregate_multi(int pred_dtype, int cmp, const void *pred_col, const void *scalars, int nqueries,
                                const ysql_agg_spec *specs, int nspecs, uint64_t n, int flags, void *ws,
                                size_t ws_bytes, ysql_stream_t stream) {
  return filter_agg_impl(pred_dtype, cmp, pred_col, scalars, nqueries, specs, nspecs, n, flags, ws, ws_bytes,
                         as_stream(stream));
}

}  // extern "C"
