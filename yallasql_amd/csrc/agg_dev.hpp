// agg_dev.hpp — filtered multi-aggregate over canonical tiles, shared by
// the per-call kernel (reduce.hip, ysql_filter_aggregate) and the batched
// executor.
//
// SELECT agg_1(c_1), ..., agg_k(c_k) WHERE pred <cmp> scalar in one read of
// the columns.  The input is cut into canonical tiles of 65536 rows; a
// tile's partial is folded in a fixed order (thread t folds rows t, t+256,
// ... of the tile, then a fixed wave64 butterfly, then the 4 waves in order)
// and the partials are combined in a fixed order, so float sums are a
// function of the input alone — the plain per-call form and the batched form
// return the same bits.
#pragma once

#include "common.hpp"

namespace ysql {

constexpr uint64_t kFaTile = 65536;
constexpr int kFaBlock = 256;
constexpr int kFaMaxSpecs = 4;

// Accumulator kinds: int64 in .i, float in .f (MIN/MAX), double in .d (SUM),
// count only.
enum { FA_KI = 0, FA_KF = 1, FA_KD = 2, FA_KC = 3 };

struct FaSpec {
  int kind, op, dtype, pad;  // op: 0 sum, 1 min, 2 max
  const void *col;
};

__host__ __device__ inline int fa_kind(int dtype, int agg) {
  if (agg == YSQL_AGG_COUNT || agg == YSQL_AGG_COUNT_STAR) return FA_KC;
  if (dtype == YSQL_FLOAT) return (agg == YSQL_AGG_SUM || agg == YSQL_AGG_AVG) ? FA_KD : FA_KF;
  return FA_KI;
}
__host__ __device__ inline int fa_op(int agg) { return agg == YSQL_AGG_MIN ? 1 : (agg == YSQL_AGG_MAX ? 2 : 0); }

__device__ __forceinline__ void fa_identity(ysql_acc &a, const FaSpec &s) {
  a.i = 0;
  a.d = 0.0;
  a.f = 0.f;
  a.pad = 0;
  a.count = 0;
  if (s.kind == FA_KI) {
    if (s.op == 1) a.i = s.dtype == YSQL_INT ? (int64_t)INT32_MAX : INT64_MAX;
    if (s.op == 2) a.i = s.dtype == YSQL_INT ? (int64_t)INT32_MIN : INT64_MIN;
  } else if (s.kind == FA_KF) {
    a.f = s.op == 1 ? __builtin_inff() : -__builtin_inff();
  }
}

__device__ __forceinline__ void fa_combine(ysql_acc &a, const ysql_acc &b, const FaSpec &s) {
  a.count += b.count;
  if (s.kind == FA_KI) {
    if (s.op == 0) a.i = (int64_t)((uint64_t)a.i + (uint64_t)b.i);
    else if (s.op == 1) a.i = b.i < a.i ? b.i : a.i;
    else a.i = b.i > a.i ? b.i : a.i;
  } else if (s.kind == FA_KF) {
    a.f = s.op == 1 ? fminf(a.f, b.f) : fmaxf(a.f, b.f);
  } else if (s.kind == FA_KD) {
    a.d += b.d;
  }
}

__device__ __forceinline__ void fa_fold_row(ysql_acc &a, const FaSpec &s, uint64_t row) {
  a.count += 1;
  if (s.kind == FA_KI) {
    const int64_t x = s.dtype == YSQL_INT ? (int64_t)((const int32_t *)s.col)[row] : ((const int64_t *)s.col)[row];
    if (s.op == 0) a.i = (int64_t)((uint64_t)a.i + (uint64_t)x);
    else if (s.op == 1) a.i = x < a.i ? x : a.i;
    else a.i = x > a.i ? x : a.i;
  } else if (s.kind == FA_KF) {
    const float x = ((const float *)s.col)[row];
    a.f = s.op == 1 ? fminf(a.f, x) : fmaxf(a.f, x);
  } else if (s.kind == FA_KD) {
    a.d += (double)((const float *)s.col)[row];
  }
}

// Fixed-shape block reduction (BLOCK threads); result in every thread.
template <int BLOCK>
__device__ __forceinline__ ysql_acc fa_block_reduce(ysql_acc v, const FaSpec &s, ysql_acc *s_w) {
  constexpr int W = BLOCK / kWave;
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    ysql_acc o;
    o.count = shfl_xor_t(v.count, m);
    o.i = shfl_xor_t(v.i, m);
    o.d = shfl_xor_t(v.d, m);
    o.f = shfl_xor_t(v.f, m);
    o.pad = 0;
    fa_combine(v, o, s);
  }
  if (lane_id() == 0) s_w[threadIdx.x / kWave] = v;
  __syncthreads();
  ysql_acc r = s_w[0];
#pragma unroll
  for (int w = 1; w < W; w++) fa_combine(r, s_w[w], s);
  __syncthreads();
  return r;
}

// Partials of canonical tile `tile`: partials[tile * ns + s].
template <typename PT, int CMP, int BLOCK>
__device__ __forceinline__ void filter_agg_tile(uint32_t tile, const PT *pred, PT scalar, const FaSpec *specs, int ns,
                                                uint64_t n, ysql_acc *partials, ysql_acc *s_w) {
  const uint64_t base = (uint64_t)tile * kFaTile;
  const uint64_t end = base + kFaTile < n ? base + kFaTile : n;
  ysql_acc acc[kFaMaxSpecs];
  FaSpec sp[kFaMaxSpecs];
#pragma unroll
  for (int s = 0; s < kFaMaxSpecs; s++) {
    if (s < ns) {
      sp[s] = specs[s];
      fa_identity(acc[s], sp[s]);
    }
  }
  for (uint64_t row = base + threadIdx.x; row < end; row += BLOCK) {
    if (!cmp_static<CMP>(pred[row], scalar)) continue;
#pragma unroll
    for (int s = 0; s < kFaMaxSpecs; s++)
      if (s < ns) fa_fold_row(acc[s], sp[s], row);
  }
#pragma unroll
  for (int s = 0; s < kFaMaxSpecs; s++) {
    if (s < ns) {
      const ysql_acc tot = fa_block_reduce<BLOCK>(acc[s], sp[s], s_w);
      if (threadIdx.x == 0) partials[(uint64_t)tile * ns + s] = tot;
    }
  }
}

// Combines the ntiles partials in a fixed order and folds the totals into
// out[s] (each an accumulator that persists across calls).
template <int BLOCK>
__device__ __forceinline__ void filter_agg_finish(const FaSpec *specs, int ns, uint32_t ntiles,
                                                  const ysql_acc *partials, ysql_acc *const *out, ysql_acc *s_w) {
  for (int s = 0; s < ns; s++) {
    const FaSpec sp = specs[s];
    ysql_acc p;
    fa_identity(p, sp);
    for (uint32_t x = threadIdx.x; x < ntiles; x += BLOCK) {
      const ysql_acc q = partials[(uint64_t)x * ns + s];
      fa_combine(p, q, sp);
    }
    const ysql_acc tot = fa_block_reduce<BLOCK>(p, sp, s_w);
    if (threadIdx.x == 0) {
      ysql_acc a = *out[s];
      fa_combine(a, tot, sp);
      *out[s] = a;
    }
  }
}

}  // namespace ysql
