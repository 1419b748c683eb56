// agg_dev.hpp — filtered multi-aggregate over canonical tiles (the
// kernels behind ysql_filter_aggregate*, reduce.hip).
//
// SELECT agg_1(c_1), ..., agg_k(c_k) WHERE pred <cmp> scalar in one read of
// the columns.  The input is cut into canonical tiles (fa_tile_rows(n) rows); a
// tile's partial is folded in a fixed order (thread t folds its runs of
// kFaRpt rows in row order, then a fixed wave64 butterfly, then the 4 waves
// in order) and the partials are combined in a fixed order, so float sums are a
// function of the input alone, whatever the grid and whichever caller
// launches the tiles.
#pragma once

#include "common.hpp"

namespace ysql {

// Canonical tile rows for an input of n rows: a power of two in
// [4096, 32768] aiming at about 1024 tiles, so a small input still spreads
// over many workgroups and a large one amortises each tile's reductions.
// It depends on n alone, so the fold order (and a float SUM's bits) is a
// function of the input.
__host__ __device__ inline uint64_t fa_tile_rows(uint64_t n) {
  uint64_t t = 4096;
  while (t < 32768 && t * 1024 < n) t <<= 1;
  return t;
}
constexpr int kFaBlock = 256;
constexpr int kFaMaxSpecs = 4;
constexpr int kFaMaxQueries = 4;  // queries sharing one scan
constexpr int kFaRpt = 8;  // rows per thread per step

// Accumulator kinds: int64 in .i, float in .f (MIN/MAX), double in .d (SUM),
// count only.
enum { FA_KI = 0, FA_KF = 1, FA_KD = 2, FA_KC = 3 };

struct FaSpec {
  int kind, op, dtype;  // op: 0 sum, 1 min, 2 max
  int src;              // where the values come from: FA_OWN (load col), FA_PRED (the
                        // predicate column), or FA_SPEC0 + t (the column of spec t < this one)
  const void *col;
};
enum { FA_OWN = 0, FA_PRED = 1, FA_SPEC0 = 2 };

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

// Fixed-shape block reduction (BLOCK threads); result in every thread.
// Only the count and the field of the spec's kind travel.
template <int BLOCK>
__device__ __forceinline__ ysql_acc fa_block_reduce(ysql_acc v, const FaSpec &s, ysql_acc *s_w) {
  constexpr int W = BLOCK / kWave;
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    ysql_acc o = v;
    o.count = shfl_xor_t(v.count, m);
    if (s.kind == FA_KI) o.i = shfl_xor_t(v.i, m);
    else if (s.kind == FA_KD) o.d = shfl_xor_t(v.d, m);
    else if (s.kind == FA_KF) o.f = shfl_xor_t(v.f, m);
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

// kFaRpt consecutive values from row r0 (rows >= end read as T()): 128-bit
// loads when the run is whole and 16-byte aligned.
template <typename T>
__device__ __forceinline__ void fa_load_rows(const T *p, uint64_t r0, uint64_t end, T (&v)[kFaRpt]) {
  if (r0 + kFaRpt <= end && ((uintptr_t)(p + r0) & 15) == 0) {
    constexpr int NQ = kFaRpt * sizeof(T) / 16;
    uint4 q[NQ];
    const uint4 *src = reinterpret_cast<const uint4 *>(p + r0);
#pragma unroll
    for (int i = 0; i < NQ; i++) q[i] = src[i];
    memcpy(&v[0], &q[0], sizeof(q));
  } else {
#pragma unroll
    for (int e = 0; e < kFaRpt; e++) v[e] = (r0 + e < end) ? p[r0 + e] : T();
  }
}

// Element width of a spec's column: 8 bytes for int64, else 4.
__host__ __device__ inline bool fa_wide(const FaSpec &s) { return s.kind == FA_KI && s.dtype != YSQL_INT; }

// The kFaRpt values of a spec's column at rows r0.. as 32-bit words (element e
// in w[e], or in w[2e], w[2e+1] when 8 bytes wide); rows >= end read 0.
__device__ __forceinline__ void fa_load_words(const FaSpec &s, uint64_t r0, uint64_t end, uint32_t (&w)[2 * kFaRpt]) {
  const bool full = r0 + kFaRpt <= end;
  if (fa_wide(s)) {
    const uint64_t *p = (const uint64_t *)s.col + r0;
    if (full && ((uintptr_t)p & 15) == 0) {
      const uint4 *q = reinterpret_cast<const uint4 *>(p);
#pragma unroll
      for (int i = 0; i < kFaRpt / 2; i++) {
        const uint4 x = q[i];
        w[4 * i] = x.x;
        w[4 * i + 1] = x.y;
        w[4 * i + 2] = x.z;
        w[4 * i + 3] = x.w;
      }
    } else {
#pragma unroll
      for (int e = 0; e < kFaRpt; e++) {
        const uint64_t x = r0 + e < end ? p[e] : 0ull;
        w[2 * e] = (uint32_t)x;
        w[2 * e + 1] = (uint32_t)(x >> 32);
      }
    }
  } else {
    const uint32_t *p = (const uint32_t *)s.col + r0;
    if (full && ((uintptr_t)p & 15) == 0) {
      const uint4 *q = reinterpret_cast<const uint4 *>(p);
#pragma unroll
      for (int i = 0; i < kFaRpt / 4; i++) {
        const uint4 x = q[i];
        w[4 * i] = x.x;
        w[4 * i + 1] = x.y;
        w[4 * i + 2] = x.z;
        w[4 * i + 3] = x.w;
      }
    } else {
#pragma unroll
      for (int e = 0; e < kFaRpt; e++) w[e] = r0 + e < end ? p[e] : 0u;
    }
  }
}

// Per-thread running value of one spec, one 64-bit word: the int64 (KI),
// the double's bits (KD) or the float's bits (KF).  The row count is kept
// once per query: without null masks every spec of a query folds the same
// rows.
__device__ __forceinline__ uint64_t fa_val_identity(const FaSpec &s) {
  ysql_acc a;
  fa_identity(a, s);
  if (s.kind == FA_KI) return (uint64_t)a.i;
  if (s.kind == FA_KF) return (uint64_t)__float_as_uint(a.f);
  return (uint64_t)__double_as_longlong(a.d);
}

__device__ __forceinline__ ysql_acc fa_val_to_acc(uint64_t v, uint32_t count, const FaSpec &s) {
  ysql_acc a;
  fa_identity(a, s);
  if (s.kind == FA_KI) a.i = (int64_t)v;
  else if (s.kind == FA_KF) a.f = __uint_as_float((uint32_t)v);
  else if (s.kind == FA_KD) a.d = __longlong_as_double((long long)v);
  a.count = count;
  return a;
}

// Folds the selected rows (bit e of `flags` = row r0 + e) of one spec, in
// row order, from its loaded words.  Unselected rows fold the operator's
// identity (0 for SUM, the current value for MIN/MAX), which leaves the
// value bit-for-bit as if they were skipped.  One short body per
// accumulator kind keeps the kernel small enough for the instruction cache.
__device__ __forceinline__ void fa_fold_words(uint64_t &val, const FaSpec &s, const uint32_t (&w)[2 * kFaRpt],
                                              uint32_t flags) {
  if (s.kind == FA_KI) {
    const bool wide = s.dtype != YSQL_INT;
    int64_t x[kFaRpt];
#pragma unroll
    for (int e = 0; e < kFaRpt; e++)
      x[e] = wide ? (int64_t)((uint64_t)w[2 * e] | ((uint64_t)w[2 * e + 1] << 32)) : (int64_t)(int32_t)w[e];
    if (s.op == 0) {
      uint64_t sum = val;
#pragma unroll
      for (int e = 0; e < kFaRpt; e++) sum += ((flags >> e) & 1u) ? (uint64_t)x[e] : 0ull;
      val = sum;
    } else {
      const bool mn = s.op == 1;
      int64_t m = (int64_t)val;
#pragma unroll
      for (int e = 0; e < kFaRpt; e++) {
        const bool better = mn ? x[e] < m : x[e] > m;
        m = (((flags >> e) & 1u) && better) ? x[e] : m;
      }
      val = (uint64_t)m;
    }
  } else if (s.kind == FA_KF) {
    const bool mn = s.op == 1;
    float m = __uint_as_float((uint32_t)val);
#pragma unroll
    for (int e = 0; e < kFaRpt; e++) {
      const float x = __uint_as_float(w[e]);
      const float r = mn ? fminf(m, x) : fmaxf(m, x);
      m = ((flags >> e) & 1u) ? r : m;
    }
    val = (uint64_t)__float_as_uint(m);
  } else if (s.kind == FA_KD) {
    double d = __longlong_as_double((long long)val);
#pragma unroll
    for (int e = 0; e < kFaRpt; e++) d += ((flags >> e) & 1u) ? (double)__uint_as_float(w[e]) : 0.0;
    val = (uint64_t)__double_as_longlong(d);
  }
}

// Partials of canonical tile `tile` for nq <= QMAX queries that share the
// predicate column, the comparison and the spec columns and differ in the
// constant: partials[(tile * nq + q) * ns + s].  Thread t owns rows
// base + t*kFaRpt .. of every step of BLOCK*kFaRpt rows.  Each step first
// issues every load (the predicate and each spec column that is not the
// predicate column, 128-bit loads), then evaluates each query's predicate
// and folds its rows in row order.  A query's fold sequence does not depend
// on the other queries of the launch, so its result is the same alone or
// batched.
template <typename PT, int CMP, int BLOCK, int QMAX>
__device__ __forceinline__ void filter_agg_tile(uint32_t tile, uint64_t tile_rows, const PT *pred,
                                                const PT (&scalar)[QMAX], int nq, const FaSpec *specs, int ns,
                                                uint64_t n, ysql_acc *partials, ysql_acc *s_w) {
  const uint64_t base = (uint64_t)tile * tile_rows;
  const uint64_t end = base + tile_rows < n ? base + tile_rows : n;
  uint64_t val[QMAX][kFaMaxSpecs];
  uint32_t cnt[QMAX];
  FaSpec sp[kFaMaxSpecs];
#pragma unroll
  for (int s = 0; s < kFaMaxSpecs; s++) {
    if (s < ns) {
      sp[s] = specs[s];
#pragma unroll
      for (int q = 0; q < QMAX; q++) val[q][s] = fa_val_identity(sp[s]);
    }
  }
#pragma unroll
  for (int q = 0; q < QMAX; q++) cnt[q] = 0;
  for (uint64_t r0 = base + (uint64_t)threadIdx.x * kFaRpt; r0 < end; r0 += (uint64_t)BLOCK * kFaRpt) {
    PT pv[kFaRpt];
    fa_load_rows<PT>(pred, r0, end, pv);
    uint32_t w[kFaMaxSpecs][2 * kFaRpt];
#pragma unroll
    for (int s = 0; s < kFaMaxSpecs; s++)
      if (s < ns && sp[s].kind != FA_KC && sp[s].src == FA_OWN) fa_load_words(sp[s], r0, end, w[s]);
    // specs over the predicate column or an earlier spec's column take a
    // copy of those words (register moves, no second load)
    {
      uint32_t pw[2 * kFaRpt];
      memcpy(pw, pv, sizeof(pv));
#pragma unroll
      for (int s = 0; s < kFaMaxSpecs; s++) {
        if (s >= ns) continue;
        const int src = sp[s].src;
        if (src == FA_PRED) {
#pragma unroll
          for (int i = 0; i < 2 * kFaRpt; i++) w[s][i] = i < (int)(sizeof(pv) / 4) ? pw[i] : 0u;
        }
#pragma unroll
        for (int t = 0; t < s; t++)
          if (src == FA_SPEC0 + t) {
#pragma unroll
            for (int i = 0; i < 2 * kFaRpt; i++) w[s][i] = w[t][i];
          }
      }
    }
#pragma unroll
    for (int q = 0; q < QMAX; q++) {
      if (q >= nq) break;
      uint32_t flags = 0;
#pragma unroll
      for (int e = 0; e < kFaRpt; e++) flags |= (uint32_t)(r0 + e < end && cmp_static<CMP>(pv[e], scalar[q])) << e;
      cnt[q] += __builtin_popcount(flags);
#pragma unroll
      for (int s = 0; s < kFaMaxSpecs; s++)
        if (s < ns) fa_fold_words(val[q][s], sp[s], w[s], flags);
    }
  }
#pragma unroll
  for (int q = 0; q < QMAX; q++) {
    if (q >= nq) break;
#pragma unroll
    for (int s = 0; s < kFaMaxSpecs; s++) {
      if (s < ns) {
        const ysql_acc tot = fa_block_reduce<BLOCK>(fa_val_to_acc(val[q][s], cnt[q], sp[s]), sp[s], s_w);
        if (threadIdx.x == 0) partials[((uint64_t)tile * nq + q) * ns + s] = tot;
      }
    }
  }
}

// Combines the ntiles partials of every (query, spec) in a fixed order; folds
// the totals into out[q * ns + s] (accumulators that persist across calls),
// or with `overwrite` stores them there.
template <int BLOCK>
__device__ __forceinline__ void filter_agg_finish(const FaSpec *specs, int ns, int nq, uint32_t ntiles,
                                                  const ysql_acc *partials, ysql_acc *const *out, bool overwrite,
                                                  ysql_acc *s_w) {
  for (int q = 0; q < nq; q++) {
    for (int s = 0; s < ns; s++) {
      const FaSpec sp = specs[s];
      ysql_acc p;
      fa_identity(p, sp);
      for (uint32_t x = threadIdx.x; x < ntiles; x += BLOCK) {
        const ysql_acc v = partials[((uint64_t)x * nq + q) * ns + s];
        fa_combine(p, v, sp);
      }
      const ysql_acc tot = fa_block_reduce<BLOCK>(p, sp, s_w);
      if (threadIdx.x == 0) {
        ysql_acc a;
        if (overwrite) fa_identity(a, sp);
        else a = *out[q * ns + s];
        fa_combine(a, tot, sp);
        *out[q * ns + s] = a;
      }
    }
  }
}

}  // namespace ysql
