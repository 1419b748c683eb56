// agg_dev.hpp — filtered multi-aggregate over canonical tiles (the
// kernels behind ysql_filter_aggregate*, reduce.hip).
//
// SELECT agg_1(c_1), ..., agg_k(c_k) WHERE pred <cmp> scalar in one read of
// the columns.  The input is cut into canonical tiles (fa_tile_rows(n) rows); a
// tile's partial is folded in a fixed order (each thread folds its rows in a
// fixed order, then a fixed wave64 butterfly, then the 4 waves in order) and the partials are combined in a fixed order, so float sums are a
// function of the input alone, whatever the grid and whichever caller
// launches the tiles.
#pragma once

#include <limits>
#include <type_traits>

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
constexpr int kFaRpt = 8;   // rows per thread per step
constexpr int kFaPair = 2;  // consecutive rows per lane per load (dense 8/16-byte loads)

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
// Element width of a spec's column: 8 bytes for int64, else 4.
__host__ __device__ inline bool fa_wide(const FaSpec &s) { return s.kind == FA_KI && s.dtype != YSQL_INT; }

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

// Fixed wave64 butterfly over a spec's accumulators; result in every lane.
// Only the count and the field of the spec's kind travel.
__device__ __forceinline__ ysql_acc fa_wave_reduce(ysql_acc v, const FaSpec &s) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    ysql_acc o = v;
    o.count = shfl_xor_t(v.count, m);
    if (s.kind == FA_KI) o.i = shfl_xor_t(v.i, m);
    else if (s.kind == FA_KD) o.d = shfl_xor_t(v.d, m);
    else if (s.kind == FA_KF) o.f = shfl_xor_t(v.f, m);
    fa_combine(v, o, s);
  }
  return v;
}

// Fixed-shape block reduction (BLOCK threads): the wave butterfly, then the
// waves in order.  Result in every thread.
template <int BLOCK>
__device__ __forceinline__ ysql_acc fa_block_reduce(ysql_acc v, const FaSpec &s, ysql_acc *s_w) {
  constexpr int W = BLOCK / kWave;
  v = fa_wave_reduce(v, s);
  if (lane_id() == 0) s_w[threadIdx.x / kWave] = v;
  __syncthreads();
  ysql_acc r = s_w[0];
#pragma unroll
  for (int w = 1; w < W; w++) fa_combine(r, s_w[w], s);
  __syncthreads();
  return r;
}

// Rows of a tile.  The tile is walked in steps of BLOCK*kFaRpt rows; in
// every step, virtual thread v < BLOCK owns rows step + u*BLOCK*2 + 2v +
// {0, 1} (u < kFaRpt/2), so at every load the 64 lanes of a wave read one
// contiguous range.  A virtual thread folds its rows in that order, steps
// in order, into one running value per (query, spec); a tile's partial is
// the wave64 butterfly over each group of 64 virtual threads followed by the
// four groups in order.  A query's fold sequence does not depend on the
// other queries or specs of the launch, so its result is a function of n and
// its own inputs alone: the same alone or sharing a scan, whichever kernel
// runs the tile.

// Fold bodies, one per (accumulator kind, operator, column width).
enum {
  FA_OPC_NONE = 0,  // COUNT: the per-query row count carries it
  FA_OPC_I32_SUM,
  FA_OPC_I32_MIN,
  FA_OPC_I32_MAX,
  FA_OPC_I64_SUM,
  FA_OPC_I64_MIN,
  FA_OPC_I64_MAX,
  FA_OPC_F_MIN,
  FA_OPC_F_MAX,
  FA_OPC_D_SUM
};

__host__ __device__ inline int fa_opcode(const FaSpec &s) {
  if (s.kind == FA_KI) {
    const int base = s.dtype == YSQL_INT ? FA_OPC_I32_SUM : FA_OPC_I64_SUM;
    return base + s.op;
  }
  if (s.kind == FA_KF) return s.op == 1 ? FA_OPC_F_MIN : FA_OPC_F_MAX;
  if (s.kind == FA_KD) return FA_OPC_D_SUM;
  return FA_OPC_NONE;
}
constexpr __host__ __device__ inline bool fa_opc_wide(int opc) { return opc >= FA_OPC_I64_SUM && opc <= FA_OPC_I64_MAX; }

// Column element type of an opcode (4 bytes for COUNT, which loads none).
template <int OPC>
using fa_col_t = typename std::conditional<(OPC >= FA_OPC_I64_SUM && OPC <= FA_OPC_I64_MAX), uint64_t,
                                           uint32_t>::type;

// Folds one value into a running value; sel false leaves it bit-for-bit
// unchanged (a double sum that starts at +0.0 never becomes -0.0, so adding
// +0.0 is exact).
template <int OPC, typename VT>
__device__ __forceinline__ void fa_fold1(uint64_t &val, VT raw, bool sel) {
  if constexpr (OPC == FA_OPC_I32_SUM || OPC == FA_OPC_I64_SUM) {
    const uint64_t x = OPC == FA_OPC_I64_SUM ? (uint64_t)raw : (uint64_t)(int64_t)(int32_t)(uint32_t)raw;
    val += sel ? x : 0ull;
  } else if constexpr (OPC == FA_OPC_I32_MIN || OPC == FA_OPC_I32_MAX || OPC == FA_OPC_I64_MIN ||
                       OPC == FA_OPC_I64_MAX) {
    constexpr bool wide = OPC == FA_OPC_I64_MIN || OPC == FA_OPC_I64_MAX;
    constexpr bool mn = OPC == FA_OPC_I32_MIN || OPC == FA_OPC_I64_MIN;
    const int64_t x = wide ? (int64_t)raw : (int64_t)(int32_t)(uint32_t)raw;
    const int64_t m = (int64_t)val;
    val = (sel && (mn ? x < m : x > m)) ? (uint64_t)x : val;
  } else if constexpr (OPC == FA_OPC_F_MIN || OPC == FA_OPC_F_MAX) {
    const float x = __uint_as_float((uint32_t)raw), m = __uint_as_float((uint32_t)val);
    const float r = OPC == FA_OPC_F_MIN ? fminf(m, x) : fmaxf(m, x);
    val = sel ? (uint64_t)__float_as_uint(r) : val;
  } else if constexpr (OPC == FA_OPC_D_SUM) {
    double d = __longlong_as_double((long long)val);
    d += (double)(sel ? __uint_as_float((uint32_t)raw) : 0.f);
    val = (uint64_t)__double_as_longlong(d);
  }
}

// Per-thread running value of one spec, one 64-bit word: the int64 (KI),
// the double's bits (KD) or the float's bits (KF).
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

// `x <cmp> c` as `(lo <= x && x <= hi) != neg`, so one kernel serves every
// comparison: strict bounds become inclusive at the neighbouring value (the
// next float toward the open side, c -/+ 1 for integers), NE is EQ negated,
// and an impossible range (x < -inf, x > INT_MAX, any ordered comparison
// with NaN) is lo > hi.  Exact for every x: NaN fails both tests, as it
// fails every ordered comparison, and passes NE by the negation.
template <typename PT>
struct Fa2Range {
  PT lo, hi;
};
template <typename PT>
__device__ __forceinline__ void fa2_empty(Fa2Range<PT> &r) {
  if constexpr (std::is_same<PT, float>::value) {
    r.lo = __builtin_inff();
    r.hi = -__builtin_inff();
  } else {
    r.lo = 1;
    r.hi = 0;
  }
}
// The neighbour of c toward -inf (down) or +inf; false when there is none.
template <typename PT>
__device__ __forceinline__ bool fa2_step_value(PT c, bool down, PT *out) {
  if constexpr (std::is_same<PT, float>::value) {
    if (isnan(c) || (down && c == -__builtin_inff()) || (!down && c == __builtin_inff())) return false;
    const uint32_t b = __float_as_uint(c);
    uint32_t nb;
    if (c == 0.f) nb = down ? 0x80000001u : 0x00000001u;  // the smallest subnormals
    else if ((b >> 31) == 0) nb = down ? b - 1 : b + 1;    // positive: bits grow with the value
    else nb = down ? b + 1 : b - 1;                        // negative: bits grow with the magnitude
    *out = __uint_as_float(nb);
    return true;
  } else {
    if (down && c == std::numeric_limits<PT>::min()) return false;
    if (!down && c == std::numeric_limits<PT>::max()) return false;
    *out = down ? (PT)(c - 1) : (PT)(c + 1);
    return true;
  }
}
template <typename PT>
__device__ __forceinline__ Fa2Range<PT> fa2_range(int cmp, PT c) {
  Fa2Range<PT> r;
  PT lowest, highest;
  if constexpr (std::is_same<PT, float>::value) {
    lowest = -__builtin_inff();
    highest = __builtin_inff();
    if (isnan(c)) {
      fa2_empty(r);
      return r;
    }
  } else {
    lowest = std::numeric_limits<PT>::min();
    highest = std::numeric_limits<PT>::max();
  }
  PT v;
  switch (cmp) {
    case YSQL_LT:
      r.lo = lowest;
      if (fa2_step_value(c, true, &v)) r.hi = v;
      else fa2_empty(r);
      break;
    case YSQL_LE: r.lo = lowest; r.hi = c; break;
    case YSQL_GT:
      r.hi = highest;
      if (fa2_step_value(c, false, &v)) r.lo = v;
      else fa2_empty(r);
      break;
    case YSQL_GE: r.lo = c; r.hi = highest; break;
    default: r.lo = c; r.hi = c; break;  // EQ, and NE negated
  }
  return r;
}

// Row selection of the filtered aggregates, by query: a comparison fixed at
// compile time (FaSelCmp) or any comparison as a range test (FaSelRange).
template <typename PT, int CMP, int QMAX>
struct FaSelCmp {
  PT c[QMAX];
  __device__ __forceinline__ bool operator()(int q, PT x) const { return cmp_static<CMP>(x, c[q]); }
};
template <typename PT, int QMAX>
struct FaSelRange {
  Fa2Range<PT> rg[QMAX];
  bool neg;
  __device__ __forceinline__ bool operator()(int q, PT x) const { return (rg[q].lo <= x && x <= rg[q].hi) != neg; }
};
template <typename PT, int QMAX>
__device__ __forceinline__ FaSelRange<PT, QMAX> fa_sel_range(int cmp, const PT *scalars, int nq) {
  FaSelRange<PT, QMAX> r;
#pragma unroll
  for (int q = 0; q < QMAX; q++) r.rg[q] = fa2_range<PT>(cmp, q < nq ? scalars[q] : PT());
  r.neg = cmp == YSQL_NE;
  return r;
}

// ---------------------------------------------------------------------------
// The tile kernel: each thread folds every spec of every query (up to QMAX
// sharing the scan) for its virtual thread, kFaRpt rows per step.
// ---------------------------------------------------------------------------

// The kFaRpt values of a column at virtual thread r0's rows: pair u at
// r0 + u * stride.  FULL: every row < end and the column aligned to two
// elements (one vector load per pair); otherwise guarded, rows >= end read 0.
template <bool FULL, typename T>
__device__ __forceinline__ void fa_load_rows(const T *p, uint64_t r0, uint64_t stride, uint64_t end,
                                             T (&v)[kFaRpt]) {
#pragma unroll
  for (int u = 0; u < kFaRpt / kFaPair; u++) {
    const uint64_t r = r0 + u * stride;
    if constexpr (FULL) {
      if constexpr (sizeof(T) == 8) {
        const uint4 q = *static_cast<const uint4 *>(__builtin_assume_aligned(p + r, 16));
        const uint64_t x = (uint64_t)q.x | ((uint64_t)q.y << 32), y = (uint64_t)q.z | ((uint64_t)q.w << 32);
        memcpy(&v[2 * u], &x, 8);
        memcpy(&v[2 * u + 1], &y, 8);
      } else {
        const uint2 q = *static_cast<const uint2 *>(__builtin_assume_aligned(p + r, 8));
        memcpy(&v[2 * u], &q.x, 4);
        memcpy(&v[2 * u + 1], &q.y, 4);
      }
    } else {
      v[2 * u] = r < end ? p[r] : T();
      v[2 * u + 1] = r + 1 < end ? p[r + 1] : T();
    }
  }
}

// A spec's kFaRpt values as words: row k in w[k], or w[2k], w[2k + 1] when wide.
template <bool FULL>
__device__ __forceinline__ void fa_load_words(const void *col, bool wide, uint64_t r0, uint64_t stride, uint64_t end,
                                              uint32_t (&w)[2 * kFaRpt]) {
  if (wide) {
    uint64_t v[kFaRpt];
    fa_load_rows<FULL, uint64_t>((const uint64_t *)col, r0, stride, end, v);
#pragma unroll
    for (int k = 0; k < kFaRpt; k++) {
      w[2 * k] = (uint32_t)v[k];
      w[2 * k + 1] = (uint32_t)(v[k] >> 32);
    }
  } else {
    uint32_t v[kFaRpt];
    fa_load_rows<FULL, uint32_t>((const uint32_t *)col, r0, stride, end, v);
#pragma unroll
    for (int k = 0; k < kFaRpt; k++) w[k] = v[k];
  }
}

template <int OPC>
__device__ __forceinline__ void fa_fold_words_op(uint64_t &val, const uint32_t (&w)[2 * kFaRpt],
                                                 const bool (&sel)[kFaRpt]) {
#pragma unroll
  for (int k = 0; k < kFaRpt; k++) {
    if constexpr (fa_opc_wide(OPC))
      fa_fold1<OPC>(val, (uint64_t)w[2 * k] | ((uint64_t)w[2 * k + 1] << 32), sel[k]);
    else
      fa_fold1<OPC>(val, w[k], sel[k]);
  }
}

// One body per opcode, chosen by a wave-uniform switch: no body evaluates
// another's operations.
__device__ __forceinline__ void fa_fold_words(uint64_t &val, int opc, const uint32_t (&w)[2 * kFaRpt],
                                              const bool (&sel)[kFaRpt]) {
  switch (opc) {
    case FA_OPC_I32_SUM: fa_fold_words_op<FA_OPC_I32_SUM>(val, w, sel); break;
    case FA_OPC_I32_MIN: fa_fold_words_op<FA_OPC_I32_MIN>(val, w, sel); break;
    case FA_OPC_I32_MAX: fa_fold_words_op<FA_OPC_I32_MAX>(val, w, sel); break;
    case FA_OPC_I64_SUM: fa_fold_words_op<FA_OPC_I64_SUM>(val, w, sel); break;
    case FA_OPC_I64_MIN: fa_fold_words_op<FA_OPC_I64_MIN>(val, w, sel); break;
    case FA_OPC_I64_MAX: fa_fold_words_op<FA_OPC_I64_MAX>(val, w, sel); break;
    case FA_OPC_F_MIN: fa_fold_words_op<FA_OPC_F_MIN>(val, w, sel); break;
    case FA_OPC_F_MAX: fa_fold_words_op<FA_OPC_F_MAX>(val, w, sel); break;
    case FA_OPC_D_SUM: fa_fold_words_op<FA_OPC_D_SUM>(val, w, sel); break;
    default: break;
  }
}

// One step of BLOCK*kFaRpt rows from s0: issues every load (the predicate
// and each spec column that is not the predicate column or an earlier
// spec's column), then evaluates each query's predicate and folds the
// thread's rows in row order.
template <bool FULL, bool SPECS, typename PT, typename SEL, int BLOCK, int QMAX>
__device__ __forceinline__ void fa_step(uint64_t s0, uint64_t end, const PT *pred, const SEL &selq, int nq,
                                        const FaSpec (&sp)[kFaMaxSpecs], const int (&opc)[kFaMaxSpecs], int ns,
                                        uint64_t (&val)[QMAX][kFaMaxSpecs], uint32_t (&cnt)[QMAX]) {
  constexpr uint64_t stride = (uint64_t)BLOCK * kFaPair;
  const uint64_t r0 = s0 + (uint64_t)threadIdx.x * kFaPair;
  PT pv[kFaRpt];
  fa_load_rows<FULL, PT>(pred, r0, stride, end, pv);
  uint32_t w[kFaMaxSpecs][2 * kFaRpt];
  if constexpr (SPECS) {
#pragma unroll
    for (int s = 0; s < kFaMaxSpecs; s++)
      if (s < ns && opc[s] != FA_OPC_NONE && sp[s].src == FA_OWN)
        fa_load_words<FULL>(sp[s].col, fa_opc_wide(opc[s]), r0, stride, end, w[s]);
  }
  // specs over the predicate column or an earlier spec's column take a copy
  // of those words (register moves, no second load)
  if constexpr (SPECS) {
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
    bool sel[kFaRpt];
#pragma unroll
    for (int k = 0; k < kFaRpt; k++) {
      sel[k] = selq(q, pv[k]);
      if constexpr (!FULL) sel[k] = sel[k] && r0 + (uint64_t)(k >> 1) * stride + (k & 1) < end;
      cnt[q] += sel[k] ? 1u : 0u;
    }
    if constexpr (SPECS) {
#pragma unroll
      for (int s = 0; s < kFaMaxSpecs; s++)
        if (s < ns) fa_fold_words(val[q][s], opc[s], w[s], sel);
    }
  }
}

// Partials of canonical tile `tile` for nq <= QMAX queries that share the
// predicate column, the comparison and the spec columns and differ in the
// constant: partials[(tile * nq + q) * ns + s].  Whole steps over aligned
// columns take unguarded vector loads; the rest (the last, partial step of
// the input, or columns not aligned to two elements) the guarded path, with
// the same rows per thread and the same fold order.
template <typename PT, typename SEL, int BLOCK, int QMAX>
__device__ __forceinline__ void filter_agg_tile(uint32_t tile, uint64_t tile_rows, const PT *pred,
                                                const SEL &selq, int nq, const FaSpec *specs, int ns,
                                                uint64_t n, ysql_acc *partials, ysql_acc *s_w) {
  const uint64_t base = (uint64_t)tile * tile_rows;
  const uint64_t end = base + tile_rows < n ? base + tile_rows : n;
  uint64_t val[QMAX][kFaMaxSpecs];
  uint32_t cnt[QMAX];
  FaSpec sp[kFaMaxSpecs];
  int opc[kFaMaxSpecs];
  bool aligned = ((uintptr_t)pred & (2 * sizeof(PT) - 1)) == 0;
  bool any = false;  // any spec other than COUNT
#pragma unroll
  for (int s = 0; s < kFaMaxSpecs; s++) {
    opc[s] = FA_OPC_NONE;
    if (s < ns) {
      sp[s] = specs[s];
      opc[s] = fa_opcode(sp[s]);
      any = any || opc[s] != FA_OPC_NONE;
      if (opc[s] != FA_OPC_NONE && sp[s].src == FA_OWN)
        aligned = aligned && ((uintptr_t)sp[s].col & (fa_opc_wide(opc[s]) ? 15 : 7)) == 0;
#pragma unroll
      for (int q = 0; q < QMAX; q++) val[q][s] = fa_val_identity(sp[s]);
    }
  }
#pragma unroll
  for (int q = 0; q < QMAX; q++) cnt[q] = 0;
  constexpr uint64_t step = (uint64_t)BLOCK * kFaRpt;
  uint64_t s0 = base;
  // COUNT-only queries: a predicate-only loop
  if (aligned && any)
    for (; s0 + step <= end; s0 += step)
      fa_step<true, true, PT, SEL, BLOCK, QMAX>(s0, end, pred, selq, nq, sp, opc, ns, val, cnt);
  if (aligned && !any)
    for (; s0 + step <= end; s0 += step)
      fa_step<true, false, PT, SEL, BLOCK, QMAX>(s0, end, pred, selq, nq, sp, opc, ns, val, cnt);
  for (; s0 < end; s0 += step)
    fa_step<false, true, PT, SEL, BLOCK, QMAX>(s0, end, pred, selq, nq, sp, opc, ns, val, cnt);
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

// ---------------------------------------------------------------------------
// Two-column form: every spec reads the predicate column (slot 0), one other
// column (slot 1, elements of W1 bytes; W1 = 0: none) or nothing (COUNT).
// The step loads both columns once, evaluates the comparison as a range test
// (fa2_range) and folds each spec with the static body of its (opcode, slot)
// pair.  The kernel stays a few hundred instructions
// long, where the generic tile above, which must move any spec's words into
// place at run time, runs to thousands and misses the instruction cache.
// Rows per thread, fold order and partial layout are those of
// filter_agg_tile, so every result is bit-identical to it.
// ---------------------------------------------------------------------------
__host__ __device__ inline int fa2_code(int opc, int slot) { return opc * 2 + slot; }


// One spec's fold over the step's rows from slot 0 (pw) or slot 1 (cw).
template <int PB, int W1>
__device__ __forceinline__ void fa2_fold(uint64_t &val, int code, const uint32_t (&pw)[2 * kFaRpt],
                                         const uint32_t (&cw)[2 * kFaRpt], const bool (&sel)[kFaRpt]) {
  switch (code) {
#define YSQL_FA2_OPC(OPC)                                                                       \
  case OPC * 2:                                                                                 \
    if constexpr ((fa_opc_wide(OPC) ? 8 : 4) == PB) fa_fold_words_op<OPC>(val, pw, sel);        \
    break;                                                                                      \
  case OPC * 2 + 1:                                                                             \
    if constexpr ((fa_opc_wide(OPC) ? 8 : 4) == W1) fa_fold_words_op<OPC>(val, cw, sel);        \
    break;
    YSQL_FA2_OPC(FA_OPC_I32_SUM)
    YSQL_FA2_OPC(FA_OPC_I32_MIN)
    YSQL_FA2_OPC(FA_OPC_I32_MAX)
    YSQL_FA2_OPC(FA_OPC_I64_SUM)
    YSQL_FA2_OPC(FA_OPC_I64_MIN)
    YSQL_FA2_OPC(FA_OPC_I64_MAX)
    YSQL_FA2_OPC(FA_OPC_F_MIN)
    YSQL_FA2_OPC(FA_OPC_F_MAX)
    YSQL_FA2_OPC(FA_OPC_D_SUM)
#undef YSQL_FA2_OPC
    default: break;  // COUNT: the per-query row count carries it
  }
}

template <bool FULL, typename PT, typename SEL, int W1, int BLOCK, int QMAX>
__device__ __forceinline__ void fa2_step(uint64_t s0, uint64_t end, const PT *pred, const void *col1,
                                         const SEL &selq, int nq, const int (&code)[kFaMaxSpecs], int ns,
                                         uint64_t (&val)[QMAX][kFaMaxSpecs], uint32_t (&cnt)[QMAX]) {
  constexpr uint64_t stride = (uint64_t)BLOCK * kFaPair;
  const uint64_t r0 = s0 + (uint64_t)threadIdx.x * kFaPair;
  PT pv[kFaRpt];
  fa_load_rows<FULL, PT>(pred, r0, stride, end, pv);
  uint32_t cw[2 * kFaRpt] = {};
  if constexpr (W1 == 8) {
    uint64_t v[kFaRpt];
    fa_load_rows<FULL, uint64_t>((const uint64_t *)col1, r0, stride, end, v);
    memcpy(cw, v, sizeof(v));
  } else if constexpr (W1 == 4) {
    uint32_t v[kFaRpt];
    fa_load_rows<FULL, uint32_t>((const uint32_t *)col1, r0, stride, end, v);
    memcpy(cw, v, sizeof(v));
  }
  uint32_t pw[2 * kFaRpt] = {};
  memcpy(pw, pv, sizeof(pv));
#pragma unroll
  for (int q = 0; q < QMAX; q++) {
    if (q >= nq) break;
    bool sel[kFaRpt];
#pragma unroll
    for (int k = 0; k < kFaRpt; k++) {
      sel[k] = selq(q, pv[k]);
      if constexpr (!FULL) sel[k] = sel[k] && r0 + (uint64_t)(k >> 1) * stride + (k & 1) < end;
      cnt[q] += sel[k] ? 1u : 0u;
    }
#pragma unroll
    for (int s = 0; s < kFaMaxSpecs; s++)
      if (s < ns) fa2_fold<(int)sizeof(PT), W1>(val[q][s], code[s], pw, cw, sel);
  }
}

// The same step with every spec's (opcode, slot) pair fixed at compile time
// (code C < 0: no spec there): no dispatch is left in the loop.  For the
// common aggregate shapes (reduce.hip lists them); same rows, order and bits.
template <int PB, int W1, int C>
__device__ __forceinline__ void fa3_fold(uint64_t &val, const uint32_t (&pw)[2 * kFaRpt],
                                         const uint32_t (&cw)[2 * kFaRpt], const bool (&sel)[kFaRpt]) {
  if constexpr (C >= 2) {  // C / 2 = opcode, C % 2 = slot; opcode 0 (COUNT) folds nothing
    constexpr int OPC = C / 2;
    if constexpr (C % 2 == 0) {
      static_assert((fa_opc_wide(OPC) ? 8 : 4) == PB, "slot 0 width");
      fa_fold_words_op<OPC>(val, pw, sel);
    } else {
      static_assert((fa_opc_wide(OPC) ? 8 : 4) == W1, "slot 1 width");
      fa_fold_words_op<OPC>(val, cw, sel);
    }
  }
}

template <bool FULL, typename PT, typename SEL, int W1, int BLOCK, int QMAX, int C0, int C1, int C2, int C3>
__device__ __forceinline__ void fa3_step(uint64_t s0, uint64_t end, const PT *pred, const void *col1,
                                         const SEL &selq, int nq, uint64_t (&val)[QMAX][kFaMaxSpecs],
                                         uint32_t (&cnt)[QMAX]) {
  constexpr uint64_t stride = (uint64_t)BLOCK * kFaPair;
  const uint64_t r0 = s0 + (uint64_t)threadIdx.x * kFaPair;
  PT pv[kFaRpt];
  fa_load_rows<FULL, PT>(pred, r0, stride, end, pv);
  uint32_t cw[2 * kFaRpt] = {};
  if constexpr (W1 == 8) {
    uint64_t v[kFaRpt];
    fa_load_rows<FULL, uint64_t>((const uint64_t *)col1, r0, stride, end, v);
    memcpy(cw, v, sizeof(v));
  } else if constexpr (W1 == 4) {
    uint32_t v[kFaRpt];
    fa_load_rows<FULL, uint32_t>((const uint32_t *)col1, r0, stride, end, v);
    memcpy(cw, v, sizeof(v));
  }
  uint32_t pw[2 * kFaRpt] = {};
  memcpy(pw, pv, sizeof(pv));
#pragma unroll
  for (int q = 0; q < QMAX; q++) {
    if (q >= nq) break;
    bool sel[kFaRpt];
#pragma unroll
    for (int k = 0; k < kFaRpt; k++) {
      sel[k] = selq(q, pv[k]);
      if constexpr (!FULL) sel[k] = sel[k] && r0 + (uint64_t)(k >> 1) * stride + (k & 1) < end;
      cnt[q] += sel[k] ? 1u : 0u;
    }
    fa3_fold<(int)sizeof(PT), W1, C0>(val[q][0], pw, cw, sel);
    fa3_fold<(int)sizeof(PT), W1, C1>(val[q][1], pw, cw, sel);
    fa3_fold<(int)sizeof(PT), W1, C2>(val[q][2], pw, cw, sel);
    fa3_fold<(int)sizeof(PT), W1, C3>(val[q][3], pw, cw, sel);
  }
}

template <typename PT, typename SEL, int W1, int BLOCK, int QMAX, int C0, int C1, int C2, int C3>
__device__ __forceinline__ void filter_agg3_tile(uint32_t tile, uint64_t tile_rows, const PT *pred,
                                                 const void *col1, const SEL &selq, int nq, const FaSpec *specs,
                                                 int ns, uint64_t n, bool aligned, ysql_acc *partials,
                                                 ysql_acc *s_w) {
  const uint64_t base = (uint64_t)tile * tile_rows;
  const uint64_t end = base + tile_rows < n ? base + tile_rows : n;
  uint64_t val[QMAX][kFaMaxSpecs];
  uint32_t cnt[QMAX];
#pragma unroll
  for (int s = 0; s < kFaMaxSpecs; s++) {
    const uint64_t id = s < ns ? fa_val_identity(specs[s]) : 0ull;
#pragma unroll
    for (int q = 0; q < QMAX; q++) val[q][s] = id;
  }
#pragma unroll
  for (int q = 0; q < QMAX; q++) cnt[q] = 0;
  constexpr uint64_t step = (uint64_t)BLOCK * kFaRpt;
  uint64_t s0 = base;
  if (aligned)
    for (; s0 + step <= end; s0 += step)
      fa3_step<true, PT, SEL, W1, BLOCK, QMAX, C0, C1, C2, C3>(s0, end, pred, col1, selq, nq, val, cnt);
  for (; s0 < end; s0 += step)
    fa3_step<false, PT, SEL, W1, BLOCK, QMAX, C0, C1, C2, C3>(s0, end, pred, col1, selq, nq, val, cnt);
#pragma unroll
  for (int q = 0; q < QMAX; q++) {
    if (q >= nq) break;
#pragma unroll
    for (int s = 0; s < kFaMaxSpecs; s++) {
      if (s < ns) {
        const ysql_acc tot = fa_block_reduce<BLOCK>(fa_val_to_acc(val[q][s], cnt[q], specs[s]), specs[s], s_w);
        if (threadIdx.x == 0) partials[((uint64_t)tile * nq + q) * ns + s] = tot;
      }
    }
  }
}

// filter_agg_tile for the two-column form; `code` from fa2_code per spec.
template <typename PT, typename SEL, int W1, int BLOCK, int QMAX>
__device__ __forceinline__ void filter_agg2_tile(uint32_t tile, uint64_t tile_rows, const PT *pred,
                                                 const void *col1, const SEL &selq, int nq,
                                                 const FaSpec *specs, const int (&code)[kFaMaxSpecs], int ns,
                                                 uint64_t n, bool aligned, ysql_acc *partials, ysql_acc *s_w) {
  const uint64_t base = (uint64_t)tile * tile_rows;
  const uint64_t end = base + tile_rows < n ? base + tile_rows : n;
  uint64_t val[QMAX][kFaMaxSpecs];
  uint32_t cnt[QMAX];
#pragma unroll
  for (int s = 0; s < kFaMaxSpecs; s++) {
    if (s < ns) {
      const uint64_t id = fa_val_identity(specs[s]);
#pragma unroll
      for (int q = 0; q < QMAX; q++) val[q][s] = id;
    }
  }
#pragma unroll
  for (int q = 0; q < QMAX; q++) cnt[q] = 0;
  constexpr uint64_t step = (uint64_t)BLOCK * kFaRpt;
  uint64_t s0 = base;
  if (aligned)
    for (; s0 + step <= end; s0 += step)
      fa2_step<true, PT, SEL, W1, BLOCK, QMAX>(s0, end, pred, col1, selq, nq, code, ns, val, cnt);
  for (; s0 < end; s0 += step)
    fa2_step<false, PT, SEL, W1, BLOCK, QMAX>(s0, end, pred, col1, selq, nq, code, ns, val, cnt);
#pragma unroll
  for (int q = 0; q < QMAX; q++) {
    if (q >= nq) break;
#pragma unroll
    for (int s = 0; s < kFaMaxSpecs; s++) {
      if (s < ns) {
        const ysql_acc tot = fa_block_reduce<BLOCK>(fa_val_to_acc(val[q][s], cnt[q], specs[s]), specs[s], s_w);
        if (threadIdx.x == 0) partials[((uint64_t)tile * nq + q) * ns + s] = tot;
      }
    }
  }
}

// Kernel arguments of a filtered multi-aggregate launch.
struct FaArgs {
  FaSpec spec[kFaMaxSpecs];
  ysql_acc *out[kFaMaxQueries * kFaMaxSpecs];  // [query][spec]
};

// Combines the ntiles partials of (query q, spec s) in a fixed order and
// folds the total into out[q * ns + s] (accumulators that persist across
// calls), or with `overwrite` stores it there.
template <int BLOCK>
__device__ __forceinline__ void filter_agg_finish_one(const FaSpec *specs, int ns, int nq, int q, int s,
                                                      uint32_t ntiles, const ysql_acc *partials, ysql_acc *const *out,
                                                      bool overwrite, ysql_acc *s_w) {
  constexpr int B8 = 8;  // partials loaded together per thread (their combine order is unchanged)
  const FaSpec sp = specs[s];
  ysql_acc p;
  fa_identity(p, sp);
  for (uint32_t x0 = threadIdx.x; x0 < ntiles; x0 += BLOCK * B8) {
    ysql_acc v[B8];
#pragma unroll
    for (int j = 0; j < B8; j++) {
      const uint32_t x = x0 + (uint32_t)j * BLOCK;
      if (x < ntiles) v[j] = partials[((uint64_t)x * nq + q) * ns + s];
    }
#pragma unroll
    for (int j = 0; j < B8; j++)
      if (x0 + (uint32_t)j * BLOCK < ntiles) fa_combine(p, v[j], sp);
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

// Every (query, spec) by one workgroup, one after another.
template <int BLOCK>
__device__ __forceinline__ void filter_agg_finish(const FaSpec *specs, int ns, int nq, uint32_t ntiles,
                                                  const ysql_acc *partials, ysql_acc *const *out, bool overwrite,
                                                  ysql_acc *s_w) {
  for (int q = 0; q < nq; q++)
    for (int s = 0; s < ns; s++) filter_agg_finish_one<BLOCK>(specs, ns, nq, q, s, ntiles, partials, out, overwrite, s_w);
}

}  // namespace ysql
