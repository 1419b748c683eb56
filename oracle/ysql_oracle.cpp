/*
 * ysql_oracle.cpp — CPU restatement of the YallaSQL operator hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle that the
 * parity tests, the smoke check and the CPU-baseline leg of the benchmark
 * compare the HIP kernels against.  It is never linked into, called by or
 * shipped with the product library.
 *
 * Every function restates one piece of the reference's semantics and cites
 * the reference file:line it follows (paths relative to the reference repo).
 * Where the reference has a defect on this path, the oracle states the
 * intended rule and the divergence is named in the comment.
 *
 * Pinning (what checks this oracle beyond restating it) is done in
 * tests/test_oracle.py: brute-force enumeration, closed forms, invariants and
 * library routines (numpy) chosen so that a dropped term, a wrong sign or a
 * swapped operand fails at least one of them.
 *
 * All pointers are host pointers.  Build: oracle/Makefile.
 */
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <numeric>
#include <thread>
#include <vector>

enum { O_INT = 0, O_FLOAT = 1, O_DATETIME = 2, O_STRING = 3, O_BOOL = 4, O_UKEY32 = 16, O_UKEY64 = 17 };
enum { O_LT = 0, O_LE = 1, O_GT = 2, O_GE = 3, O_EQ = 4, O_NE = 5 };
enum { O_ADD = 0, O_SUB = 1, O_MUL = 2, O_DIV = 3, O_REM = 4 };
enum { A_MAX = 0, A_MIN = 1, A_SUM = 2, A_AVG = 3, A_COUNT = 4, A_COUNT_STAR = 5 };

/* Same layout as ysql_acc in include/ysql_kernels.h. */
typedef struct orc_acc {
  int64_t i;
  double d;
  float f;
  uint32_t pad;
  int64_t count;
} orc_acc;

static const int STR_BYTES = 256;

/* ------------------------------------------------------------------ */
/* A1: sort-key normalisation — radix_sort_utils.cu:4-37;
 * bound_ref_sort.hpp:14-64.  int32 -> x ^ 0x80000000; float ->
 * sign ? ~bits : bits ^ 0x80000000.  int64 gets a sign flip (the reference
 * used raw bits, bound_ref_sort.hpp:53-55; identical for x >= 0).
 * descending complements the key (DescOp, radix_sort_kernel.hpp:38-45,
 * complements every digit = complements the key). */
extern "C" int orc_sortkey_encode(int dtype, int desc, const void *in, void *out, uint64_t n) {
  if (dtype == O_INT || dtype == O_FLOAT) {
    const uint32_t *x = (const uint32_t *)in;
    uint32_t *y = (uint32_t *)out;
    for (uint64_t i = 0; i < n; i++) {
      uint32_t b = x[i];
      uint32_t k = (dtype == O_INT) ? (b ^ 0x80000000u) : ((b & 0x80000000u) ? ~b : (b ^ 0x80000000u));
      y[i] = desc ? ~k : k;
    }
    return 0;
  }
  if (dtype == O_DATETIME) {
    const uint64_t *x = (const uint64_t *)in;
    uint64_t *y = (uint64_t *)out;
    for (uint64_t i = 0; i < n; i++) {
      uint64_t k = x[i] ^ 0x8000000000000000ull;
      y[i] = desc ? ~k : k;
    }
    return 0;
  }
  return -1;
}

/* A2-A6: the reference's LSD radix argsort is stable (split sort per bit,
 * radix_sort_kernel.cu:8-136); its result is the stable ascending order of
 * the (possibly complemented) key: perm[sorted_pos] = original row,
 * radix_sort_kernel.cu:139-170.  std::stable_sort states exactly that. */
extern "C" int orc_stable_argsort(int key_dtype, int desc, const void *keys, uint32_t *perm, uint64_t n) {
  for (uint64_t i = 0; i < n; i++) perm[i] = (uint32_t)i;
  if (key_dtype == O_UKEY32) {
    const uint32_t *k = (const uint32_t *)keys;
    if (desc)
      std::stable_sort(perm, perm + n, [&](uint32_t a, uint32_t b) { return k[a] > k[b]; });
    else
      std::stable_sort(perm, perm + n, [&](uint32_t a, uint32_t b) { return k[a] < k[b]; });
    return 0;
  }
  if (key_dtype == O_UKEY64) {
    const uint64_t *k = (const uint64_t *)keys;
    if (desc)
      std::stable_sort(perm, perm + n, [&](uint32_t a, uint32_t b) { return k[a] > k[b]; });
    else
      std::stable_sort(perm, perm + n, [&](uint32_t a, uint32_t b) { return k[a] < k[b]; });
    return 0;
  }
  return -1;
}

/* Typed "a <= b" of the merge ops (merge_batchs.hpp:20-34): AscSortMergeOp
 * is_less_or_equal = a <= b, DescSortMergeOp = a >= b. */
static int str_cmp(const char *a, const char *b);

static inline bool merge_le(int dtype, int desc, const void *A, uint64_t i, const void *B, uint64_t j) {
  switch (dtype) {
    case O_INT: {
      int32_t a = ((const int32_t *)A)[i], b = ((const int32_t *)B)[j];
      return desc ? a >= b : a <= b;
    }
    case O_FLOAT: {
      float a = ((const float *)A)[i], b = ((const float *)B)[j];
      return desc ? a >= b : a <= b;
    }
    case O_DATETIME: {
      int64_t a = ((const int64_t *)A)[i], b = ((const int64_t *)B)[j];
      return desc ? a >= b : a <= b;
    }
    case O_UKEY32: {
      uint32_t a = ((const uint32_t *)A)[i], b = ((const uint32_t *)B)[j];
      return desc ? a >= b : a <= b;
    }
    case O_UKEY64: {
      uint64_t a = ((const uint64_t *)A)[i], b = ((const uint64_t *)B)[j];
      return desc ? a >= b : a <= b;
    }
    case O_STRING: { /* AscSortStrOp / DescSortStrOp, merge_batchs.hpp:36-48 */
      int c = str_cmp((const char *)A + i * STR_BYTES, (const char *)B + j * STR_BYTES);
      return desc ? c >= 0 : c <= 0;
    }
  }
  return true;
}

static inline int dtype_bytes(int dtype) {
  switch (dtype) {
    case O_INT: case O_FLOAT: case O_UKEY32: return 4;
    case O_DATETIME: case O_UKEY64: return 8;
    case O_STRING: return STR_BYTES;
    case O_BOOL: return 1;
  }
  return 0;
}

/* A8-A9: the sequential merge of merge_sorted_array_kernel_v1
 * (merge_batchs.cu:45-75): take A while A[i] <= B[j] (A first on ties),
 * outputs (C, newIdxs, tableIdxs) for the first k_max outputs, and
 * lasti = number of A elements among them (merge_batchs.cu:38-40).  The
 * reference's parallel co-rank is wrong on ties (merge_batchs.cu:22); this
 * restates the intended rule of its sequential loop. */
extern "C" int orc_merge(int dtype, int desc, const void *A, uint64_t m, const void *B, uint64_t n,
              uint64_t k_max, void *C, uint32_t *src_idx, uint8_t *src_is_a, int64_t *consumed_a) {
  int eb = dtype_bytes(dtype);
  uint64_t i = 0, j = 0, k = 0, kmax = std::min<uint64_t>(k_max, m + n);
  for (; k < kmax; k++) {
    bool take_a = (j >= n) || (i < m && merge_le(dtype, desc, A, i, B, j));
    if (take_a) {
      if (C) memcpy((char *)C + k * eb, (const char *)A + i * eb, eb);
      src_idx[k] = (uint32_t)i;
      src_is_a[k] = 1;
      i++;
    } else {
      if (C) memcpy((char *)C + k * eb, (const char *)B + j * eb, eb);
      src_idx[k] = (uint32_t)j;
      src_is_a[k] = 0;
      j++;
    }
  }
  *consumed_a = (int64_t)i;
  return 0;
}

/* ORDER BY over key columns, strings included (OrderOperator,
 * order_operator.hpp:196-265; AscOpString / DescOpString,
 * radix_sort_kernel.hpp:45-73, ordering strings as strcmp_device does,
 * string_kernel.hpp:46-56; the reference sorted by orders[0] only,
 * order_operator.hpp:277, and the multi-column rule is the intended one):
 * a stable sort of row indices comparing the columns in order, numeric
 * columns by their sort-key encoding (orc_sortkey_encode without the DESC
 * complement), strings by str_cmp, a descending column's result negated. */
extern "C" int orc_sort_rows(int nkeys, const int *dtypes, const int *desc, const void *const *cols, uint64_t n,
                             uint32_t *perm) {
  std::vector<std::vector<uint64_t>> enc((size_t)nkeys);
  for (int c = 0; c < nkeys; c++) {
    if (dtypes[c] == O_STRING) continue;
    if (dtypes[c] != O_INT && dtypes[c] != O_FLOAT && dtypes[c] != O_DATETIME) return -1;
    enc[c].resize(n);
    if (dtypes[c] == O_DATETIME) {
      orc_sortkey_encode(dtypes[c], 0, cols[c], enc[c].data(), n);
    } else {
      std::vector<uint32_t> k(n);
      orc_sortkey_encode(dtypes[c], 0, cols[c], k.data(), n);
      for (uint64_t i = 0; i < n; i++) enc[c][i] = k[i];
    }
  }
  for (uint64_t i = 0; i < n; i++) perm[i] = (uint32_t)i;
  std::stable_sort(perm, perm + n, [&](uint32_t x, uint32_t y) {
    for (int c = 0; c < nkeys; c++) {
      int r;
      if (dtypes[c] == O_STRING) {
        r = str_cmp((const char *)cols[c] + (uint64_t)x * STR_BYTES, (const char *)cols[c] + (uint64_t)y * STR_BYTES);
      } else {
        r = enc[c][x] < enc[c][y] ? -1 : (enc[c][x] > enc[c][y] ? 1 : 0);
      }
      if (r) return desc[c] ? r > 0 : r < 0;
    }
    return false;
  });
  return 0;
}

/* A4 / A13: inclusive scans (prefix_sum_kernel.cu:7-156). */
extern "C" int orc_scan_u32_inclusive(const uint32_t *in, uint32_t *out, uint64_t n) {
  uint32_t s = 0;
  for (uint64_t i = 0; i < n; i++) { s += in[i]; out[i] = s; }
  return 0;
}
extern "C" int orc_scan_mask_inclusive(const uint8_t *mask, uint32_t *out, uint64_t n) {
  uint32_t s = 0;
  for (uint64_t i = 0; i < n; i++) { s += mask[i] ? 1u : 0u; out[i] = s; }
  return 0;
}

/* strcmp_device, string_kernel.hpp:46-56. */
static int str_cmp(const char *a, const char *b) {
  int i = 0;
  while (i < STR_BYTES && a[i] && a[i] == b[i]) i++;
  if (i >= STR_BYTES) return 0;
  return (int)(unsigned char)a[i] - (int)(unsigned char)b[i];
}

template <typename T>
static inline bool cmp_apply(int cmp, T a, T b) {
  switch (cmp) {
    case O_LT: return a < b;   /* LEOperator  */
    case O_LE: return a <= b;  /* LEQOperator */
    case O_GT: return a > b;   /* GEOperator  */
    case O_GE: return a >= b;  /* GEQOperator */
    case O_EQ: return a == b;
    case O_NE: return a != b;
  }
  return false;
}

static inline bool cmp_at(int dtype, int cmp, const void *l, uint64_t li, const void *r, uint64_t ri) {
  switch (dtype) {
    case O_INT: return cmp_apply(cmp, ((const int32_t *)l)[li], ((const int32_t *)r)[ri]);
    case O_FLOAT: return cmp_apply(cmp, ((const float *)l)[li], ((const float *)r)[ri]);
    case O_DATETIME: return cmp_apply(cmp, ((const int64_t *)l)[li], ((const int64_t *)r)[ri]);
    case O_STRING:
      return cmp_apply(cmp, str_cmp((const char *)l + li * STR_BYTES, (const char *)r + ri * STR_BYTES), 0);
    case O_BOOL: return cmp_apply(cmp, ((const uint8_t *)l)[li] != 0, ((const uint8_t *)r)[ri] != 0);
  }
  return false;
}

/* A12: apply_batches / apply_batches_scalar_rhs / _lhs
 * (comparison_operators_kernel.cu:57-112): res[i] = op(lhs[i|0], rhs[i|0])
 * XOR isneg, operand order op(lhs, rhs) (comparison.hpp:68-73 passes
 * (rhs, lhs) to a launcher that applies op(lhs, rhs)). */
extern "C" int orc_compare(int dtype, int cmp, const void *lhs, int lhs_scalar, const void *rhs, int rhs_scalar,
                int negate, uint8_t *mask, uint64_t n) {
  for (uint64_t i = 0; i < n; i++) {
    bool v = cmp_at(dtype, cmp, lhs, lhs_scalar ? 0 : i, rhs, rhs_scalar ? 0 : i);
    mask[i] = (uint8_t)(negate ? !v : v);
  }
  return 0;
}

/* ANDOperator / OROperator (comparison_operators_kernel.hpp:78-86). */
extern "C" int orc_mask_logic(int op, const uint8_t *a, int a_scalar, const uint8_t *b, int b_scalar, uint8_t *out,
                   uint64_t n) {
  for (uint64_t i = 0; i < n; i++) {
    bool x = a[a_scalar ? 0 : i] != 0, y = b[b_scalar ? 0 : i] != 0;
    out[i] = (uint8_t)(op == 0 ? (x && y) : (x || y));
  }
  return 0;
}

/* A15: BinaryOperator functors (binary_operators_kernel.hpp:27-51):
 * a+b, a-b, a*b, b!=0 ? a/b : 0, b!=0 ? a%b : 0.  Integer add/sub/mul wrap
 * (two's complement, as the device does). */
template <typename T> struct wrap_type { typedef T type; };
template <> struct wrap_type<int32_t> { typedef uint32_t type; };
template <> struct wrap_type<int64_t> { typedef uint64_t type; };

template <typename T>
static inline T arith_apply(int op, T a, T b) {
  typedef typename wrap_type<T>::type U;
  switch (op) {
    case O_ADD: return (T)((U)a + (U)b);
    case O_SUB: return (T)((U)a - (U)b);
    case O_MUL: return (T)((U)a * (U)b);
    case O_DIV:
      if (b == 0) return 0;
      if (std::is_integral<T>::value && b == (T)-1) return (T)((U)0 - (U)a);
      return a / b;
  }
  return 0;
}
template <typename T>
static inline T rem_apply(T a, T b) {
  if (b == 0) return 0;
  if (b == (T)-1) return 0;
  return a % b;
}

extern "C" int orc_arith(int dtype, int op, const void *lhs, int lhs_scalar, const void *rhs, int rhs_scalar, void *out,
              uint64_t n) {
  for (uint64_t i = 0; i < n; i++) {
    uint64_t li = lhs_scalar ? 0 : i, ri = rhs_scalar ? 0 : i;
    switch (dtype) {
      case O_INT: {
        int32_t a = ((const int32_t *)lhs)[li], b = ((const int32_t *)rhs)[ri];
        ((int32_t *)out)[i] = op == O_REM ? rem_apply(a, b) : arith_apply(op, a, b);
        break;
      }
      case O_DATETIME: {
        int64_t a = ((const int64_t *)lhs)[li], b = ((const int64_t *)rhs)[ri];
        ((int64_t *)out)[i] = op == O_REM ? rem_apply(a, b) : arith_apply(op, a, b);
        break;
      }
      case O_FLOAT: {
        if (op == O_REM) return -1; /* RemOperator is not instantiated for float (.cu:71-74) */
        float a = ((const float *)lhs)[li], b = ((const float *)rhs)[ri];
        ((float *)out)[i] = arith_apply(op, a, b);
        break;
      }
      default: return -1;
    }
  }
  return 0;
}

/* numerical_cast_kernel (cast_operators_kernel.cu:5-17): static_cast.
 * float -> int conversions of out-of-range values are undefined in C++; the
 * oracle saturates (and maps NaN to 0) and the device path does the same. */
template <typename D>
static D cast_from_double(double v) {
  if (std::is_integral<D>::value) {
    if (std::isnan(v)) return 0;
    double lo = (double)std::numeric_limits<D>::min(), hi = (double)std::numeric_limits<D>::max();
    if (v <= lo) return std::numeric_limits<D>::min();
    if (v >= hi) return std::numeric_limits<D>::max();
    return (D)v;
  }
  return (D)v;
}
extern "C" int orc_cast(int src, int dst, const void *in, void *out, uint64_t n) {
  for (uint64_t i = 0; i < n; i++) {
    switch (src) {
      case O_INT: {
        int32_t v = ((const int32_t *)in)[i];
        if (dst == O_FLOAT) ((float *)out)[i] = (float)v;
        else if (dst == O_DATETIME) ((int64_t *)out)[i] = v;
        else if (dst == O_INT) ((int32_t *)out)[i] = v;
        else return -1;
        break;
      }
      case O_FLOAT: {
        float v = ((const float *)in)[i];
        if (dst == O_INT) ((int32_t *)out)[i] = cast_from_double<int32_t>(v);
        else if (dst == O_DATETIME) ((int64_t *)out)[i] = cast_from_double<int64_t>(v);
        else if (dst == O_FLOAT) ((float *)out)[i] = v;
        else return -1;
        break;
      }
      case O_DATETIME: {
        int64_t v = ((const int64_t *)in)[i];
        if (dst == O_INT) ((int32_t *)out)[i] = (int32_t)(uint32_t)(uint64_t)v;
        else if (dst == O_FLOAT) ((float *)out)[i] = (float)v;
        else if (dst == O_DATETIME) ((int64_t *)out)[i] = v;
        else return -1;
        break;
      }
      case O_BOOL: {
        uint8_t v = ((const uint8_t *)in)[i] != 0;
        if (dst == O_INT) ((int32_t *)out)[i] = v; /* bool -> unsigned int (.cu:33) */
        else return -1;
        break;
      }
      default: return -1;
    }
  }
  return 0;
}

/* A14: move_rows_filter_kernel (move_rows_kernel.cu:32-49) with the
 * stable order of the mask scan: selected rows keep their relative order.
 * NULL rows are written with their null byte (the reference left a hole,
 * move_rows_kernel.cu:45 — divergent).  Returns the count. */
extern "C" uint64_t orc_compact(uint32_t eb, const void *src, const uint8_t *src_null, const uint8_t *mask, uint64_t n,
                     void *dst, uint8_t *dst_null) {
  uint64_t o = 0;
  for (uint64_t i = 0; i < n; i++) {
    if (!mask[i]) continue;
    if (dst) memcpy((char *)dst + o * eb, (const char *)src + i * eb, eb);
    if (dst_null) dst_null[o] = src_null ? src_null[i] : 0;
    o++;
  }
  return o;
}

/* A7: move_one_col (move_rows_kernel.cu:51-70). */
extern "C" int orc_gather(uint32_t eb, const void *src, const uint8_t *src_null, const uint32_t *idx, uint64_t n, void *dst,
               uint8_t *dst_null) {
  for (uint64_t i = 0; i < n; i++) {
    memcpy((char *)dst + i * eb, (const char *)src + (uint64_t)idx[i] * eb, eb);
    if (dst_null) dst_null[i] = src_null ? src_null[idx[i]] : 0;
  }
  return 0;
}
/* A10: move_one_col_merge (move_rows_kernel.cu:72-105). */
extern "C" int orc_gather2(uint32_t eb, const void *a, const uint8_t *a_null, const void *b, const uint8_t *b_null,
                const uint32_t *idx, const uint8_t *is_a, uint64_t n, int reverse, void *dst, uint8_t *dst_null) {
  for (uint64_t i = 0; i < n; i++) {
    uint64_t d = reverse ? n - 1 - i : i;
    const void *s = is_a[i] ? a : b;
    const uint8_t *sn = is_a[i] ? a_null : b_null;
    memcpy((char *)dst + d * eb, (const char *)s + (uint64_t)idx[i] * eb, eb);
    if (dst_null) dst_null[d] = sn ? sn[idx[i]] : 0;
  }
  return 0;
}
/* A20: move_rows_join_kernel (move_rows_kernel.cu:7-30). */
extern "C" int orc_gather_pairs(uint32_t eb, const void *src, const uint8_t *src_null, const uint32_t *pairs, int side,
                     uint64_t n, void *dst, uint8_t *dst_null) {
  for (uint64_t i = 0; i < n; i++) {
    uint32_t r = pairs[2 * i + (side ? 1 : 0)];
    memcpy((char *)dst + i * eb, (const char *)src + (uint64_t)r * eb, eb);
    if (dst_null) dst_null[i] = src_null ? src_null[r] : 0;
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* A16-A17: aggregates.
 * Identities: aggregate.hpp:208-272 with the defects fixed: MAX(float)
 * starts at -inf (the reference used FLT_MIN, aggregate.hpp:235); float SUM
 * starts at exactly 0 (the reference seeded from an unassigned member,
 * aggregate.hpp:261-263).  MIN/MAX ops: a <= b ? a : b / a >= b ? a : b
 * (reduction_kernel.hpp:24-32) with NaN ignored as fminf/fmaxf do in the
 * float atomics (reduction_kernel.cu:56-75).  SUM(int32) accumulates in
 * int64 (the reference wrapped in int32, aggregate.hpp:122-124).
 * Float SUM uses a compensated (Neumaier) long-double sum: the exact value to
 * within far less than the 1e-6 relative tolerance of the tests.  COUNT =
 * number of non-NULL rows (reduction_count_notnull, reduction_kernel.cu:124-166);
 * COUNT(*) = all rows (aggregate.hpp:56-58). */
extern "C" int orc_acc_init(int dtype, int agg, orc_acc *acc) {
  memset(acc, 0, sizeof(*acc));
  if (agg == A_MIN) {
    if (dtype == O_INT) acc->i = std::numeric_limits<int32_t>::max();
    else if (dtype == O_DATETIME) acc->i = std::numeric_limits<int64_t>::max();
    else acc->f = std::numeric_limits<float>::infinity();
  } else if (agg == A_MAX) {
    if (dtype == O_INT) acc->i = std::numeric_limits<int32_t>::min();
    else if (dtype == O_DATETIME) acc->i = std::numeric_limits<int64_t>::min();
    else acc->f = -std::numeric_limits<float>::infinity();
  }
  return 0;
}

extern "C" int orc_reduce(int dtype, int agg, const void *in, const uint8_t *nul, uint64_t n, orc_acc *acc) {
  if (agg == A_COUNT_STAR) { acc->count += (int64_t)n; return 0; }
  long double s = 0, c = 0;
  for (uint64_t i = 0; i < n; i++) {
    if (nul && nul[i]) continue;
    acc->count++;
    if (agg == A_COUNT) continue;
    if (dtype == O_INT || dtype == O_DATETIME) {
      int64_t v = dtype == O_INT ? (int64_t)((const int32_t *)in)[i] : ((const int64_t *)in)[i];
      if (agg == A_MIN) acc->i = std::min(acc->i, v);
      else if (agg == A_MAX) acc->i = std::max(acc->i, v);
      else acc->i = (int64_t)((uint64_t)acc->i + (uint64_t)v);
    } else if (dtype == O_FLOAT) {
      float v = ((const float *)in)[i];
      if (agg == A_MIN) acc->f = std::fmin(acc->f, v);
      else if (agg == A_MAX) acc->f = std::fmax(acc->f, v);
      else {
        long double t = s + (long double)v;
        if (std::fabs((double)s) >= std::fabs((double)v)) c += (s - t) + v;
        else c += ((long double)v - t) + s;
        s = t;
      }
    } else {
      return -1;
    }
  }
  if (dtype == O_FLOAT && (agg == A_SUM || agg == A_AVG)) acc->d = (double)((long double)acc->d + s + c);
  return 0;
}

/* getAggregate (aggregate.hpp:149-192), div_avg_kernel (reduction_kernel.cu:5-10),
 * convert_double_to_float_kernel (:17-22). */
extern "C" int orc_agg_finalize(int dtype, int agg, const orc_acc *acc, void *out) {
  switch (agg) {
    case A_MIN: case A_MAX:
      if (dtype == O_INT) *(int32_t *)out = (int32_t)acc->i;
      else if (dtype == O_DATETIME) *(int64_t *)out = acc->i;
      else *(float *)out = acc->f;
      return 0;
    case A_SUM:
      if (dtype == O_FLOAT) *(float *)out = (float)acc->d;
      else *(int64_t *)out = acc->i;
      return 0;
    case A_AVG: {
      double s = dtype == O_FLOAT ? acc->d : (double)acc->i;
      *(float *)out = (float)(s / (double)acc->count);
      return 0;
    }
    case A_COUNT: case A_COUNT_STAR:
      *(int64_t *)out = acc->count;
      return 0;
  }
  return -1;
}

/* GROUP BY: no reference function (aggregate_operator.hpp:92-103 ignores
 * groups).  Per-group fold with the same rules as orc_reduce. */
extern "C" int orc_group_reduce(int dtype, int agg, const uint32_t *gid, uint32_t ng, const void *in, const uint8_t *nul,
                     uint64_t n, orc_acc *accs) {
  int eb = dtype_bytes(dtype);
  for (uint64_t i = 0; i < n; i++) {
    uint32_t g = gid[i];
    if (g >= ng) continue;
    int rc = orc_reduce(dtype, agg, (const char *)in + i * eb, nul ? nul + i : nullptr, 1, &accs[g]);
    if (rc) return rc;
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* A18: outer_join_batches (comparison_operators_kernel.cu:5-22): every
 * (l, r) with op(lhs[l], rhs[r]) (launch_outer_join_operators receives
 * (rhs, lhs) and applies op(lhs[lidx], rhs[ridx])).  The reference emits the
 * pairs in atomic order; the oracle emits them ordered by (l, r).  Writes at
 * most cap pairs, returns the full count. */
extern "C" uint64_t orc_join_pairs(int dtype, int cmp, const void *left, uint64_t L, const void *right, uint64_t R,
                        uint32_t *pairs, uint64_t cap) {
  uint64_t c = 0;
  for (uint64_t l = 0; l < L; l++)
    for (uint64_t r = 0; r < R; r++)
      if (cmp_at(dtype, cmp, left, l, right, r)) {
        if (c < cap) { pairs[2 * c] = (uint32_t)l; pairs[2 * c + 1] = (uint32_t)r; }
        c++;
      }
  return c;
}

/* A19: and_join_batches (comparison_operators_kernel.cu:24-54): a p1 entry
 * is kept (once, guarded by the bitset over its index) if the same (l, r)
 * occurs in p2.  Kept in p1 order. */
extern "C" uint64_t orc_pairs_intersect(const uint32_t *p1, uint64_t n1, const uint32_t *p2, uint64_t n2, uint32_t *out,
                             uint64_t cap) {
  std::vector<uint64_t> s(n2);
  for (uint64_t i = 0; i < n2; i++) s[i] = ((uint64_t)p2[2 * i] << 32) | p2[2 * i + 1];
  std::sort(s.begin(), s.end());
  uint64_t c = 0;
  for (uint64_t i = 0; i < n1; i++) {
    uint64_t k = ((uint64_t)p1[2 * i] << 32) | p1[2 * i + 1];
    if (std::binary_search(s.begin(), s.end(), k)) {
      if (c < cap) { out[2 * c] = p1[2 * i]; out[2 * c + 1] = p1[2 * i + 1]; }
      c++;
    }
  }
  return c;
}

/* A21: cross_product_kernel (cross_product_kernel.cu:7-27):
 * out[l*R + r] = isleft ? col[l] : col[r]. */
extern "C" int orc_cross_expand(uint32_t eb, const void *src, const uint8_t *src_null, uint64_t L, uint64_t R, int is_left,
                     void *dst, uint8_t *dst_null) {
  for (uint64_t l = 0; l < L; l++)
    for (uint64_t r = 0; r < R; r++) {
      uint64_t row = l * R + r, s = is_left ? l : r;
      memcpy((char *)dst + row * eb, (const char *)src + s * eb, eb);
      if (dst_null) dst_null[row] = src_null ? src_null[s] : 0;
    }
  return 0;
}

/* Star join + grouped SUM: the composition of A18 equi-joins on dense
 * primary keys (pair set {(i, fk[i]) : fk[i] in [0, rows)}), the dimension
 * filters, and a per-group int64 SUM(m1 * m2). */
typedef struct orc_star_dim {
  const int32_t *fk;
  uint32_t rows;
  uint32_t group_card;
  const uint8_t *pass;
  const int32_t *group;
  int32_t group_base;
  uint32_t pad;
} orc_star_dim;

extern "C" int orc_star_join_group_sum(const orc_star_dim *dims, int ndims, const int32_t *m1, const int32_t *m2, uint64_t n,
                            int64_t *sums, uint64_t *counts, uint32_t n_groups) {
  uint64_t stride[8];
  uint64_t s = 1;
  for (int d = ndims - 1; d >= 0; d--) {
    stride[d] = s;
    if (dims[d].group) s *= dims[d].group_card;
  }
  if (s > n_groups) return -1;
  for (uint64_t i = 0; i < n; i++) {
    uint64_t g = 0;
    bool ok = true;
    for (int d = 0; d < ndims && ok; d++) {
      int32_t f = dims[d].fk[i];
      if (f < 0 || (uint32_t)f >= dims[d].rows) { ok = false; break; }
      if (dims[d].pass && !dims[d].pass[f]) { ok = false; break; }
      if (dims[d].group) {
        int64_t v = (int64_t)dims[d].group[f] - dims[d].group_base;
        if (v < 0 || v >= (int64_t)dims[d].group_card) { ok = false; break; }
        g += (uint64_t)v * stride[d];
      }
    }
    if (!ok) continue;
    sums[g] = (int64_t)((uint64_t)sums[g] + (uint64_t)((int64_t)m1[i] * (int64_t)m2[i]));
    counts[g]++;
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* CPU baseline of the out-of-core sort (multi-run LSD radix + merges),
 * the same pipeline as the device path: each run of `run_rows` rows is
 * stably sorted by an 8-bit LSD radix sort over 64-bit keys carrying the
 * global row index, then runs are merged pairwise (stable, A first),
 * level by level as OrderOperator::merge_all_sorted (order_operator.hpp:160-193).
 * Threads sort runs in parallel and merge pairs in parallel.
 * Output: out_keys / out_vals.  Returns 0. */
static void lsd_radix_u64(uint64_t *k, uint32_t *v, uint64_t n, uint64_t *tk, uint32_t *tv, int passes) {
  uint64_t *sk = k, *dk = tk;
  uint32_t *sv = v, *dv = tv;
  for (int p = 0; p < passes; p++) {
    int sh = 8 * p;
    uint64_t cnt[257] = {0};
    for (uint64_t i = 0; i < n; i++) cnt[((sk[i] >> sh) & 255) + 1]++;
    for (int d = 0; d < 256; d++) cnt[d + 1] += cnt[d];
    for (uint64_t i = 0; i < n; i++) {
      uint64_t pos = cnt[(sk[i] >> sh) & 255]++;
      dk[pos] = sk[i];
      dv[pos] = sv[i];
    }
    std::swap(sk, dk);
    std::swap(sv, dv);
  }
  if (sk != k) {
    memcpy(k, sk, n * 8);
    memcpy(v, sv, n * 4);
  }
}

extern "C" int orc_cpu_multirun_sort_u64(const uint64_t *keys, uint64_t n, uint64_t run_rows, int threads,
                              uint64_t *out_keys, uint32_t *out_vals) {
  if (threads < 1) threads = 1;
  uint64_t nruns = (n + run_rows - 1) / run_rows;
  std::vector<uint64_t> k(keys, keys + n), tk(n);
  std::vector<uint32_t> v(n), tv(n);
  for (uint64_t i = 0; i < n; i++) v[i] = (uint32_t)i;
  uint64_t mx = 0;
  for (uint64_t i = 0; i < n; i++) mx |= keys[i];
  int passes = 0;
  while (passes < 8 && (mx >> (8 * passes))) passes++;
  /* run generation */
  {
    std::atomic<uint64_t> next{0};
    std::vector<std::thread> th;
    for (int t = 0; t < threads; t++)
      th.emplace_back([&]() {
        for (uint64_t r; (r = next.fetch_add(1)) < nruns;) {
          uint64_t b = r * run_rows, e = std::min(n, b + run_rows);
          lsd_radix_u64(&k[b], &v[b], e - b, &tk[b], &tv[b], passes);
        }
      });
    for (auto &t : th) t.join();
  }
  /* pairwise merge levels */
  std::vector<std::pair<uint64_t, uint64_t>> runs;
  for (uint64_t r = 0; r < nruns; r++) runs.push_back({r * run_rows, std::min(n, (r + 1) * run_rows)});
  uint64_t *src_k = k.data(), *dst_k = tk.data();
  uint32_t *src_v = v.data(), *dst_v = tv.data();
  while (runs.size() > 1) {
    std::vector<std::pair<uint64_t, uint64_t>> out;
    std::vector<std::thread> th;
    for (size_t i = 0; i < runs.size(); i += 2) {
      if (i + 1 == runs.size()) {
        auto rr = runs[i];
        th.emplace_back([=]() {
          memcpy(dst_k + rr.first, src_k + rr.first, (rr.second - rr.first) * 8);
          memcpy(dst_v + rr.first, src_v + rr.first, (rr.second - rr.first) * 4);
        });
        out.push_back(rr);
        continue;
      }
      auto a = runs[i], b = runs[i + 1];
      th.emplace_back([=]() {
        uint64_t x = a.first, y = b.first, o = a.first;
        while (x < a.second && y < b.second) {
          if (src_k[x] <= src_k[y]) { dst_k[o] = src_k[x]; dst_v[o++] = src_v[x++]; }
          else { dst_k[o] = src_k[y]; dst_v[o++] = src_v[y++]; }
        }
        while (x < a.second) { dst_k[o] = src_k[x]; dst_v[o++] = src_v[x++]; }
        while (y < b.second) { dst_k[o] = src_k[y]; dst_v[o++] = src_v[y++]; }
      });
      out.push_back({a.first, b.second});
      if ((int)th.size() >= threads) {
        for (auto &t : th) t.join();
        th.clear();
      }
    }
    for (auto &t : th) t.join();
    runs = out;
    std::swap(src_k, dst_k);
    std::swap(src_v, dst_v);
  }
  memcpy(out_keys, src_k, n * 8);
  memcpy(out_vals, src_v, n * 4);
  return 0;
}

extern "C" const char *orc_version(void) { return "ysql-oracle 1"; }

