// elementwise.hip — key normalisation, predicates, mask logic, arithmetic,
// casts and the row-movement kernels (gathers, mask compaction, cross
// product).  All are streaming kernels bound by HBM bandwidth: grid-stride
// loops over 64-bit row indices, one wave reading 64 consecutive rows.
#include <math.h>

#include <limits>

#include "common.hpp"
#include "ops_dev.hpp"

namespace ysql {

constexpr int kEwBlock = 256;

static inline dim3 ew_grid(uint64_t n) {
  uint64_t g = ceil_div(n, (uint64_t)kEwBlock * 4);
  uint64_t cap = (uint64_t)grid_cap() * 4;
  if (g > cap) g = cap;
  if (g < 1) g = 1;
  return dim3((unsigned)g);
}

#define GRID_STRIDE(i, n) \
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += (uint64_t)gridDim.x * blockDim.x)

// ---------------------------------------------------------------------------
// zero fill (used instead of hipMemsetAsync so callers may capture graphs)
// ---------------------------------------------------------------------------
__global__ void k_zero_u64(uint64_t *p, uint64_t n) {
  GRID_STRIDE(i, n) p[i] = 0;
}
__global__ void k_zero_u8(uint8_t *p, uint64_t n) {
  GRID_STRIDE(i, n) p[i] = 0;
}

int zero_words(void *p, size_t bytes, hipStream_t stream) {
  if (bytes == 0) return YSQL_OK;
  if (((uintptr_t)p & 7) == 0 && (bytes & 7) == 0) {
    k_zero_u64<<<ew_grid(bytes / 8), kEwBlock, 0, stream>>>((uint64_t *)p, bytes / 8);
  } else {
    k_zero_u8<<<ew_grid(bytes), kEwBlock, 0, stream>>>((uint8_t *)p, bytes);
  }
  YSQL_LAUNCH_CHECK();
  return YSQL_OK;
}

// ---------------------------------------------------------------------------
// A1: sort-key normalisation (radix_sort_utils.cu:4-37)
// ---------------------------------------------------------------------------
__global__ void k_encode32(const uint32_t *in, uint32_t *out, uint64_t n, int is_float, uint32_t flip) {
  GRID_STRIDE(i, n) {
    uint32_t b = in[i];
    uint32_t k = is_float ? ((b & 0x80000000u) ? ~b : (b ^ 0x80000000u)) : (b ^ 0x80000000u);
    out[i] = k ^ flip;
  }
}
__global__ void k_decode32(const uint32_t *in, uint32_t *out, uint64_t n, int is_float, uint32_t flip) {
  GRID_STRIDE(i, n) {
    uint32_t k = in[i] ^ flip;
    out[i] = is_float ? ((k & 0x80000000u) ? (k ^ 0x80000000u) : ~k) : (k ^ 0x80000000u);
  }
}
__global__ void k_xor64(const uint64_t *in, uint64_t *out, uint64_t n, uint64_t x) {
  GRID_STRIDE(i, n) out[i] = in[i] ^ x;
}

__global__ void k_pack32(const uint32_t *in, uint64_t *key, uint64_t n, int is_float, uint32_t flip) {
  GRID_STRIDE(i, n) {
    uint32_t b = in[i];
    uint32_t k = is_float ? ((b & 0x80000000u) ? ~b : (b ^ 0x80000000u)) : (b ^ 0x80000000u);
    key[i] = (key[i] << 32) | (uint64_t)(k ^ flip);
  }
}

static int encode_impl(int dtype, int desc, const void *in, void *out, uint64_t n, hipStream_t s, bool decode) {
  if (n == 0) return YSQL_OK;
  if (!in || !out) return YSQL_EINVAL;
  if (dtype == YSQL_INT || dtype == YSQL_FLOAT) {
    uint32_t flip = desc ? 0xFFFFFFFFu : 0u;
    if (decode)
      k_decode32<<<ew_grid(n), kEwBlock, 0, s>>>((const uint32_t *)in, (uint32_t *)out, n, dtype == YSQL_FLOAT, flip);
    else
      k_encode32<<<ew_grid(n), kEwBlock, 0, s>>>((const uint32_t *)in, (uint32_t *)out, n, dtype == YSQL_FLOAT, flip);
  } else if (dtype == YSQL_DATETIME) {
    uint64_t x = 0x8000000000000000ull ^ (desc ? ~0ull : 0ull);
    k_xor64<<<ew_grid(n), kEwBlock, 0, s>>>((const uint64_t *)in, (uint64_t *)out, n, x);
  } else {
    return YSQL_EINVAL;
  }
  YSQL_LAUNCH_CHECK();
  return YSQL_OK;
}

// ---------------------------------------------------------------------------
// A12: predicates (comparison_operators_kernel.cu:57-112), op(lhs, rhs)
// ---------------------------------------------------------------------------
template <typename T, int CMP>
__global__ void k_compare(const T *lhs, int ls, const T *rhs, int rs, uint8_t neg, uint8_t *mask, uint64_t n) {
  const T l0 = ls ? lhs[0] : T();
  const T r0 = rs ? rhs[0] : T();
  GRID_STRIDE(i, n) {
    T a = ls ? l0 : lhs[i];
    T b = rs ? r0 : rhs[i];
    mask[i] = (uint8_t)(cmp_static<CMP>(a, b) ^ (neg != 0));
  }
}
__global__ void k_compare_str(const char *lhs, int ls, const char *rhs, int rs, int cmp, uint8_t neg, uint8_t *mask,
                              uint64_t n) {
  GRID_STRIDE(i, n) {
    const char *a = lhs + (ls ? 0 : i * YSQL_STRING_BYTES);
    const char *b = rhs + (rs ? 0 : i * YSQL_STRING_BYTES);
    mask[i] = (uint8_t)(cmp_apply(cmp, str_cmp(a, b), 0) ^ (neg != 0));
  }
}

template <typename T>
static void launch_compare_t(int cmp, const void *l, int ls, const void *r, int rs, int neg, uint8_t *m, uint64_t n,
                             hipStream_t s) {
  dim3 g = ew_grid(n);
  const T *L = (const T *)l;
  const T *R = (const T *)r;
  switch (cmp) {
    case YSQL_LT: k_compare<T, YSQL_LT><<<g, kEwBlock, 0, s>>>(L, ls, R, rs, neg, m, n); break;
    case YSQL_LE: k_compare<T, YSQL_LE><<<g, kEwBlock, 0, s>>>(L, ls, R, rs, neg, m, n); break;
    case YSQL_GT: k_compare<T, YSQL_GT><<<g, kEwBlock, 0, s>>>(L, ls, R, rs, neg, m, n); break;
    case YSQL_GE: k_compare<T, YSQL_GE><<<g, kEwBlock, 0, s>>>(L, ls, R, rs, neg, m, n); break;
    case YSQL_EQ: k_compare<T, YSQL_EQ><<<g, kEwBlock, 0, s>>>(L, ls, R, rs, neg, m, n); break;
    default: k_compare<T, YSQL_NE><<<g, kEwBlock, 0, s>>>(L, ls, R, rs, neg, m, n); break;
  }
}

__global__ void k_mask_logic(const uint8_t *a, int as, const uint8_t *b, int bs, int op, uint8_t *out, uint64_t n) {
  GRID_STRIDE(i, n) {
    bool x = a[as ? 0 : i] != 0, y = b[bs ? 0 : i] != 0;
    out[i] = (uint8_t)(op == YSQL_AND ? (x && y) : (x || y));
  }
}

// ---------------------------------------------------------------------------
// A15: arithmetic (binary_operators_kernel.hpp:27-51).  Integer add/sub/mul
// wrap; DIV/REM by 0 give 0; INT_MIN / -1 wraps and INT_MIN % -1 is 0.
// ---------------------------------------------------------------------------
template <typename T, int OP>
__global__ void k_arith(const T *lhs, int ls, const T *rhs, int rs, T *out, uint64_t n) {
  const T l0 = ls ? lhs[0] : T();
  const T r0 = rs ? rhs[0] : T();
  GRID_STRIDE(i, n) {
    T a = ls ? l0 : lhs[i];
    T b = rs ? r0 : rhs[i];
    out[i] = arith_op<T, OP>(a, b);
  }
}

template <typename T>
static int launch_arith_t(int op, const void *l, int ls, const void *r, int rs, void *o, uint64_t n, hipStream_t s) {
  dim3 g = ew_grid(n);
  const T *L = (const T *)l;
  const T *R = (const T *)r;
  T *O = (T *)o;
  switch (op) {
    case YSQL_ADD: k_arith<T, YSQL_ADD><<<g, kEwBlock, 0, s>>>(L, ls, R, rs, O, n); break;
    case YSQL_SUB: k_arith<T, YSQL_SUB><<<g, kEwBlock, 0, s>>>(L, ls, R, rs, O, n); break;
    case YSQL_MUL: k_arith<T, YSQL_MUL><<<g, kEwBlock, 0, s>>>(L, ls, R, rs, O, n); break;
    case YSQL_DIV: k_arith<T, YSQL_DIV><<<g, kEwBlock, 0, s>>>(L, ls, R, rs, O, n); break;
    case YSQL_REM:
      if constexpr (std::numeric_limits<T>::is_integer) {
        k_arith<T, YSQL_REM><<<g, kEwBlock, 0, s>>>(L, ls, R, rs, O, n);
        break;
      } else {
        return YSQL_EUNSUPPORTED;
      }
    default: return YSQL_EINVAL;
  }
  YSQL_LAUNCH_CHECK();
  return YSQL_OK;
}

// ---------------------------------------------------------------------------
// A15: casts (cast_operators_kernel.cu:5-17).  float -> integer saturates
// and maps NaN to 0 (what v_cvt_i32_f32 does; made explicit for int64).
// ---------------------------------------------------------------------------
template <typename S, typename D>
__global__ void k_cast(const S *in, D *out, uint64_t n) {
  GRID_STRIDE(i, n) {
    if constexpr (std::is_same<S, float>::value && std::numeric_limits<D>::is_integer)
      out[i] = sat_from_float<D>(in[i]);
    else if constexpr (std::is_same<S, uint8_t>::value)
      out[i] = (D)(in[i] != 0);
    else
      out[i] = (D)in[i];
  }
}

// ---------------------------------------------------------------------------
// A7 / A10 / A14 / A20 / A21: row movement
// ---------------------------------------------------------------------------
template <typename E>
__global__ void k_gather(const E *src, const uint8_t *sn, const uint32_t *idx, uint64_t n, E *dst, uint8_t *dn) {
  GRID_STRIDE(i, n) {
    uint32_t s = idx[i];
    dst[i] = src[s];
    if (dn) dn[i] = sn ? sn[s] : 0;
  }
}
template <typename E>
__global__ void k_gather2(const E *a, const uint8_t *an, const E *b, const uint8_t *bn, const uint32_t *idx,
                          const uint8_t *is_a, uint64_t n, int reverse, E *dst, uint8_t *dn) {
  GRID_STRIDE(i, n) {
    uint64_t d = reverse ? n - 1 - i : i;
    uint32_t s = idx[i];
    bool fa = is_a[i] != 0;
    dst[d] = fa ? a[s] : b[s];
    if (dn) {
      const uint8_t *nn = fa ? an : bn;
      dn[d] = nn ? nn[s] : 0;
    }
  }
}
template <typename E>
__global__ void k_gather_pairs(const E *src, const uint8_t *sn, const uint32_t *pairs, int side, uint64_t n, E *dst,
                               uint8_t *dn) {
  GRID_STRIDE(i, n) {
    uint32_t s = pairs[2 * i + side];
    dst[i] = src[s];
    if (dn) dn[i] = sn ? sn[s] : 0;
  }
}
template <typename E>
__global__ void k_compact_map(const E *src, const uint8_t *sn, const uint8_t *mask, const uint32_t *map, uint64_t n,
                              E *dst, uint8_t *dn) {
  GRID_STRIDE(i, n) {
    if (mask[i]) {
      uint32_t o = map[i] - 1;
      dst[o] = src[i];
      if (dn) dn[o] = sn ? sn[i] : 0;
    }
  }
}
template <typename E>
__global__ void k_cross(const E *src, const uint8_t *sn, uint64_t L, uint64_t R, int is_left, E *dst, uint8_t *dn) {
  const uint64_t total = L * R;
  GRID_STRIDE(row, total) {
    uint64_t l = row / R;
    uint64_t s = is_left ? l : row - l * R;
    dst[row] = src[s];
    if (dn) dn[row] = sn ? sn[s] : 0;
  }
}

// Dispatch over element width: 1, 2, 4, 8, 16 or 256 bytes.
#define DISPATCH_ELEM(eb, E, ...)                    \
  switch (eb) {                                      \
    case 1: { typedef uint8_t E; __VA_ARGS__; break; }  \
    case 2: { typedef uint16_t E; __VA_ARGS__; break; } \
    case 4: { typedef uint32_t E; __VA_ARGS__; break; } \
    case 8: { typedef uint64_t E; __VA_ARGS__; break; } \
    case 16: { typedef uint4 E; __VA_ARGS__; break; }   \
    case 256: { typedef Str256 E; __VA_ARGS__; break; } \
    default: return YSQL_EINVAL;                     \
  }

}  // namespace ysql

using namespace ysql;

extern "C" {

int ysql_sortkey_encode(int dtype, int descending, const void *in, void *out, uint64_t n, ysql_stream_t stream) {
  return encode_impl(dtype, descending, in, out, n, as_stream(stream), false);
}
int ysql_sortkey_decode(int dtype, int descending, const void *in, void *out, uint64_t n, ysql_stream_t stream) {
  return encode_impl(dtype, descending, in, out, n, as_stream(stream), true);
}
int ysql_sortkey_pack(int dtype, int descending, const void *in, uint64_t n, uint64_t *key_inout,
                      ysql_stream_t stream) {
  if (n == 0) return YSQL_OK;
  if (!in || !key_inout) return YSQL_EINVAL;
  hipStream_t s = as_stream(stream);
  if (dtype == YSQL_INT || dtype == YSQL_FLOAT) {
    k_pack32<<<ew_grid(n), kEwBlock, 0, s>>>((const uint32_t *)in, key_inout, n, dtype == YSQL_FLOAT,
                                             descending ? 0xFFFFFFFFu : 0u);
  } else if (dtype == YSQL_DATETIME) {
    k_xor64<<<ew_grid(n), kEwBlock, 0, s>>>((const uint64_t *)in, key_inout, n,
                                            0x8000000000000000ull ^ (descending ? ~0ull : 0ull));
  } else {
    return YSQL_EINVAL;
  }
  YSQL_LAUNCH_CHECK();
  return YSQL_OK;
}

int ysql_compare(int dtype, int cmp, const void *lhs, int lhs_is_scalar, const void *rhs, int rhs_is_scalar,
                 int negate, uint8_t *mask_out, uint64_t n, ysql_stream_t stream) {
  if (n == 0) return YSQL_OK;
  if (!lhs || !rhs || !mask_out || cmp < YSQL_LT || cmp > YSQL_NE) return YSQL_EINVAL;
  hipStream_t s = as_stream(stream);
  switch (dtype) {
    case YSQL_INT: launch_compare_t<int32_t>(cmp, lhs, lhs_is_scalar, rhs, rhs_is_scalar, negate, mask_out, n, s); break;
    case YSQL_FLOAT: launch_compare_t<float>(cmp, lhs, lhs_is_scalar, rhs, rhs_is_scalar, negate, mask_out, n, s); break;
    case YSQL_DATETIME: launch_compare_t<int64_t>(cmp, lhs, lhs_is_scalar, rhs, rhs_is_scalar, negate, mask_out, n, s); break;
    case YSQL_BOOL: launch_compare_t<uint8_t>(cmp, lhs, lhs_is_scalar, rhs, rhs_is_scalar, negate, mask_out, n, s); break;
    case YSQL_STRING:
      k_compare_str<<<ew_grid(n), kEwBlock, 0, s>>>((const char *)lhs, lhs_is_scalar, (const char *)rhs, rhs_is_scalar,
                                                   cmp, (uint8_t)negate, mask_out, n);
      break;
    default: return YSQL_EINVAL;
  }
  YSQL_LAUNCH_CHECK();
  return YSQL_OK;
}

int ysql_mask_logic(int op, const uint8_t *a, int a_is_scalar, const uint8_t *b, int b_is_scalar, uint8_t *out,
                    uint64_t n, ysql_stream_t stream) {
  if (n == 0) return YSQL_OK;
  if (!a || !b || !out || (op != YSQL_AND && op != YSQL_OR)) return YSQL_EINVAL;
  k_mask_logic<<<ew_grid(n), kEwBlock, 0, as_stream(stream)>>>(a, a_is_scalar, b, b_is_scalar, op, out, n);
  YSQL_LAUNCH_CHECK();
  return YSQL_OK;
}

int ysql_arith(int dtype, int op, const void *lhs, int lhs_is_scalar, const void *rhs, int rhs_is_scalar, void *out,
               uint64_t n, ysql_stream_t stream) {
  if (n == 0) return YSQL_OK;
  if (!lhs || !rhs || !out) return YSQL_EINVAL;
  hipStream_t s = as_stream(stream);
  switch (dtype) {
    case YSQL_INT: return launch_arith_t<int32_t>(op, lhs, lhs_is_scalar, rhs, rhs_is_scalar, out, n, s);
    case YSQL_FLOAT: return launch_arith_t<float>(op, lhs, lhs_is_scalar, rhs, rhs_is_scalar, out, n, s);
    case YSQL_DATETIME: return launch_arith_t<int64_t>(op, lhs, lhs_is_scalar, rhs, rhs_is_scalar, out, n, s);
  }
  return YSQL_EINVAL;
}

int ysql_cast(int src_dtype, int dst_dtype, const void *in, void *out, uint64_t n, ysql_stream_t stream) {
  if (n == 0) return YSQL_OK;
  if (!in || !out) return YSQL_EINVAL;
  hipStream_t s = as_stream(stream);
  dim3 g = ew_grid(n);
#define CAST_CASE(SD, DD, S, D) \
  if (src_dtype == SD && dst_dtype == DD) { k_cast<S, D><<<g, kEwBlock, 0, s>>>((const S *)in, (D *)out, n); YSQL_LAUNCH_CHECK(); return YSQL_OK; }
  CAST_CASE(YSQL_INT, YSQL_FLOAT, int32_t, float)
  CAST_CASE(YSQL_INT, YSQL_DATETIME, int32_t, int64_t)
  CAST_CASE(YSQL_INT, YSQL_INT, int32_t, int32_t)
  CAST_CASE(YSQL_FLOAT, YSQL_INT, float, int32_t)
  CAST_CASE(YSQL_FLOAT, YSQL_DATETIME, float, int64_t)
  CAST_CASE(YSQL_FLOAT, YSQL_FLOAT, float, float)
  CAST_CASE(YSQL_DATETIME, YSQL_INT, int64_t, int32_t)
  CAST_CASE(YSQL_DATETIME, YSQL_FLOAT, int64_t, float)
  CAST_CASE(YSQL_DATETIME, YSQL_DATETIME, int64_t, int64_t)
  CAST_CASE(YSQL_BOOL, YSQL_INT, uint8_t, int32_t)
#undef CAST_CASE
  return YSQL_EINVAL;
}

int ysql_gather(uint32_t elem_bytes, const void *src, const uint8_t *src_null, const uint32_t *idx, uint64_t n,
                void *dst, uint8_t *dst_null, ysql_stream_t stream) {
  if (n == 0) return YSQL_OK;
  if (!src || !idx || !dst) return YSQL_EINVAL;
  hipStream_t s = as_stream(stream);
  DISPATCH_ELEM(elem_bytes, E,
                k_gather<E><<<ew_grid(n), kEwBlock, 0, s>>>((const E *)src, src_null, idx, n, (E *)dst, dst_null));
  YSQL_LAUNCH_CHECK();
  return YSQL_OK;
}

int ysql_gather2(uint32_t elem_bytes, const void *a, const uint8_t *a_null, const void *b, const uint8_t *b_null,
                 const uint32_t *idx, const uint8_t *is_a, uint64_t n, int reverse, void *dst, uint8_t *dst_null,
                 ysql_stream_t stream) {
  if (n == 0) return YSQL_OK;
  if (!a || !b || !idx || !is_a || !dst) return YSQL_EINVAL;
  hipStream_t s = as_stream(stream);
  DISPATCH_ELEM(elem_bytes, E,
                k_gather2<E><<<ew_grid(n), kEwBlock, 0, s>>>((const E *)a, a_null, (const E *)b, b_null, idx, is_a, n,
                                                            reverse, (E *)dst, dst_null));
  YSQL_LAUNCH_CHECK();
  return YSQL_OK;
}

int ysql_gather_pairs(uint32_t elem_bytes, const void *src, const uint8_t *src_null, const uint32_t *pairs, int side,
                      uint64_t n, void *dst, uint8_t *dst_null, ysql_stream_t stream) {
  if (n == 0) return YSQL_OK;
  if (!src || !pairs || !dst || (side != 0 && side != 1)) return YSQL_EINVAL;
  hipStream_t s = as_stream(stream);
  DISPATCH_ELEM(elem_bytes, E,
                k_gather_pairs<E><<<ew_grid(n), kEwBlock, 0, s>>>((const E *)src, src_null, pairs, side, n, (E *)dst,
                                                                 dst_null));
  YSQL_LAUNCH_CHECK();
  return YSQL_OK;
}

int ysql_compact(uint32_t elem_bytes, const void *src, const uint8_t *src_null, const uint8_t *mask,
                 const uint32_t *map, uint64_t n, void *dst, uint8_t *dst_null, ysql_stream_t stream) {
  if (n == 0) return YSQL_OK;
  if (!src || !mask || !map || !dst) return YSQL_EINVAL;
  hipStream_t s = as_stream(stream);
  DISPATCH_ELEM(elem_bytes, E,
                k_compact_map<E><<<ew_grid(n), kEwBlock, 0, s>>>((const E *)src, src_null, mask, map, n, (E *)dst,
                                                                dst_null));
  YSQL_LAUNCH_CHECK();
  return YSQL_OK;
}

int ysql_cross_expand(uint32_t elem_bytes, const void *src, const uint8_t *src_null, uint64_t L, uint64_t R,
                      int is_left, void *dst, uint8_t *dst_null, ysql_stream_t stream) {
  if (L == 0 || R == 0) return YSQL_OK;
  if (!src || !dst) return YSQL_EINVAL;
  if (R != 0 && L > UINT64_MAX / R) return YSQL_EINVAL;
  hipStream_t s = as_stream(stream);
  DISPATCH_ELEM(elem_bytes, E,
                k_cross<E><<<ew_grid(L * R), kEwBlock, 0, s>>>((const E *)src, src_null, L, R, is_left, (E *)dst,
                                                              dst_null));
  YSQL_LAUNCH_CHECK();
  return YSQL_OK;
}

}  // extern "C"
