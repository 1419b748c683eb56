// strings.hip — STRING (256-byte, NUL-padded) sort keys and merges.
//
// Reference: OrderOperator sorts STRING keys with AscOpString /
// DescOpString (radix_sort_kernel.hpp:45-73, instantiated at
// radix_sort_kernel.cu:174,178, called at order_operator.hpp:220,250): a
// 4-bit LSD radix sort over all 256 bytes, 512 passes whatever the strings
// hold.  It merges sorted runs with AscSortStrOp / DescSortStrOp
// (merge_batchs.hpp:36-48), strcmp_device (string_kernel.hpp:46-56).
//
// Here the order is strcmp's (unsigned bytes, up to the first NUL, at most
// 256 bytes), stable:
//   * ysql_string_sort: one pass finds every string's length and the
//     longest; then an LSD sort over 8-byte chunks, last chunk first, only
//     as many chunks as the longest string needs.  Chunk c of a string is
//     its bytes [8c, 8c+8) with the bytes at or past its length zeroed,
//     byte-swapped so that a uint64 compare is a byte compare (complemented
//     for DESC); each chunk is one stable 64-bit radix sort carrying the
//     permutation.  Masking past the length makes the order strcmp's even
//     where bytes after the NUL are not zero.
//   * merges compare 8 bytes at a time: the first byte that differs or is
//     NUL in the left string decides, as in strcmp_device.
#include <algorithm>
#include <utility>

#include "common.hpp"

namespace ysql {

constexpr uint64_t kLo7 = 0x7f7f7f7f7f7f7f7full, kHi = 0x8080808080808080ull, kOnes = 0x0101010101010101ull;

// strcmp_device (string_kernel.hpp:46-56) on 8-byte words of two 8-byte
// aligned 256-byte strings.  Per word: bit 7 of each byte that differs,
// and of each zero byte of `a` (exact up to the first zero byte, which is
// all that is used); the lowest such byte is where strcmp stops.
__device__ __forceinline__ int str_cmp_words(const char *a, const char *b) {
  const uint64_t *pa = reinterpret_cast<const uint64_t *>(a);
  const uint64_t *pb = reinterpret_cast<const uint64_t *>(b);
  for (int w = 0; w < YSQL_STRING_BYTES / 8; w++) {
    const uint64_t x = pa[w], y = pb[w];
    const uint64_t d = x ^ y;
    const uint64_t differs = (((d & kLo7) + kLo7) | d) & kHi;
    const uint64_t zero = (x - kOnes) & ~x & kHi;
    const uint64_t stop = differs | zero;
    if (stop) {
      const int sh = __builtin_ctzll(stop) - 7;
      return (int)((x >> sh) & 0xff) - (int)((y >> sh) & 0xff);
    }
  }
  return 0;
}

// Length up to the first NUL (256 if none), from 8-byte words.
__device__ __forceinline__ uint32_t str_len_words(const char *s) {
  const uint64_t *p = reinterpret_cast<const uint64_t *>(s);
  for (int w = 0; w < YSQL_STRING_BYTES / 8; w++) {
    const uint64_t x = p[w];
    const uint64_t zero = (x - kOnes) & ~x & kHi;
    if (zero) return (uint32_t)(w * 8 + (__builtin_ctzll(zero) >> 3));
  }
  return YSQL_STRING_BYTES;
}

constexpr int kStrBlock = 256;

__global__ __launch_bounds__(kStrBlock) void k_str_len(const char *__restrict__ strs, uint64_t n,
                                                       uint16_t *__restrict__ len, uint32_t *__restrict__ max_len) {
  __shared__ uint32_t s_max;
  if (threadIdx.x == 0) s_max = 0;
  __syncthreads();
  uint32_t m = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * kStrBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kStrBlock) {
    const uint32_t l = str_len_words(strs + i * YSQL_STRING_BYTES);
    len[i] = (uint16_t)l;
    m = l > m ? l : m;
  }
  for (int o = 32; o > 0; o >>= 1) {
    const uint32_t v = __shfl_xor(m, o);
    m = v > m ? v : m;
  }
  if (lane_id() == 0) atomicMax(&s_max, m);
  __syncthreads();
  if (threadIdx.x == 0 && s_max) atomicMax(max_len, s_max);
}

// keys[i] = chunk c of string perm[i], big-endian, bytes at or past the
// string's length zeroed, xor flip (all ones for DESC).
__global__ __launch_bounds__(kStrBlock) void k_str_chunk_key(const char *__restrict__ strs,
                                                             const uint32_t *__restrict__ perm,
                                                             const uint16_t *__restrict__ len, uint64_t n, int c,
                                                             uint64_t flip, uint64_t *__restrict__ keys) {
  const uint64_t i = (uint64_t)blockIdx.x * kStrBlock + threadIdx.x;
  if (i >= n) return;
  const uint32_t row = perm[i];
  const int l = (int)len[row] - 8 * c;  // bytes of this chunk inside the string
  uint64_t k = 0;
  if (l > 0) {
    uint64_t x = reinterpret_cast<const uint64_t *>(strs + (uint64_t)row * YSQL_STRING_BYTES)[c];
    if (l < 8) x &= (1ull << (8 * l)) - 1;
    k = __builtin_bswap64(x);
  }
  keys[i] = k ^ flip;
}

__global__ void k_iota_u32(uint32_t *p, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = (uint32_t)i;
}

__global__ void k_copy_u32(const uint32_t *__restrict__ src, uint32_t *__restrict__ dst, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

// ------------------------------------------------------------------ merges
// a <= b in the merge order of one key column.
struct KeyCol {
  int dtype, desc;
  const char *a, *b;
};

// Order-preserving unsigned key of a numeric value (ysql_sortkey_encode's
// rule, without the DESC complement).
__device__ __forceinline__ uint64_t enc_key(int dtype, const char *base, uint64_t i) {
  if (dtype == YSQL_DATETIME) return reinterpret_cast<const uint64_t *>(base)[i] ^ (1ull << 63);
  const uint32_t x = reinterpret_cast<const uint32_t *>(base)[i];
  if (dtype == YSQL_FLOAT) return (x & 0x80000000u) ? ~x : (x ^ 0x80000000u);
  return x ^ 0x80000000u;
}

// Sign of (A row i) vs (B row j) over the key columns, each column's
// direction applied.
template <int NK>
__device__ __forceinline__ int row_cmp(const KeyCol *k, int nk, uint64_t i, uint64_t j) {
  for (int c = 0; c < (NK > 0 ? NK : nk); c++) {
    int r;
    if (k[c].dtype == YSQL_STRING) {
      r = str_cmp_words(k[c].a + i * YSQL_STRING_BYTES, k[c].b + j * YSQL_STRING_BYTES);
    } else {
      const uint64_t x = enc_key(k[c].dtype, k[c].a, i), y = enc_key(k[c].dtype, k[c].b, j);
      r = x < y ? -1 : (x > y ? 1 : 0);
    }
    if (r) return k[c].desc ? -r : r;
  }
  return 0;
}

struct RowKeys {
  KeyCol col[YSQL_MAX_COLS];
  int nk;
};

constexpr int kRowIpt = 8;

// Stable merge of two sorted runs by their key rows, ties from A first:
// thread t emits outputs [8t, 8t + 8): co-rank i = min{i : not (A[i] <=
// B[k-1-i])} by binary search, then a sequential merge.  The thread that
// emits the last output stores the number of A rows consumed.
template <int NK>
__global__ __launch_bounds__(kStrBlock) void k_row_merge(RowKeys keys, uint64_t m, uint64_t n, uint64_t kmax,
                                                         uint32_t *__restrict__ src_idx,
                                                         uint8_t *__restrict__ src_is_a, char *__restrict__ c_out,
                                                         int64_t *__restrict__ consumed_a) {
  const uint64_t d0 = ((uint64_t)blockIdx.x * kStrBlock + threadIdx.x) * kRowIpt;
  if (d0 >= kmax) return;
  const uint64_t d1 = d0 + kRowIpt < kmax ? d0 + kRowIpt : kmax;
  const KeyCol *k = keys.col;
  uint64_t lo = d0 > n ? d0 - n : 0, hi = d0 < m ? d0 : m;
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    if (row_cmp<NK>(k, keys.nk, mid, d0 - 1 - mid) <= 0) lo = mid + 1;
    else hi = mid;
  }
  uint64_t i = lo, j = d0 - lo;
  for (uint64_t d = d0; d < d1; d++) {
    const bool take_a = j >= n || (i < m && row_cmp<NK>(k, keys.nk, i, j) <= 0);
    const uint64_t src = take_a ? i : j;
    src_idx[d] = (uint32_t)src;
    src_is_a[d] = take_a ? 1 : 0;
    if (c_out) {  // the key string itself (one STRING key column)
      const uint4 *s = reinterpret_cast<const uint4 *>((take_a ? k[0].a : k[0].b) + src * YSQL_STRING_BYTES);
      uint4 *o = reinterpret_cast<uint4 *>(c_out + d * YSQL_STRING_BYTES);
#pragma unroll
      for (int w = 0; w < YSQL_STRING_BYTES / 16; w++) o[w] = s[w];
    }
    if (take_a) i++;
    else j++;
  }
  if (d1 == kmax && consumed_a) *consumed_a = (int64_t)i;
}

static int row_merge(const RowKeys &keys, uint64_t m, uint64_t n, uint64_t kmax, uint32_t *src_idx, uint8_t *src_is_a,
                     char *c_out, int64_t *consumed_a, hipStream_t s) {
  const unsigned grid = (unsigned)ceil_div(ceil_div(kmax, kRowIpt), kStrBlock);
  if (keys.nk == 1)
    k_row_merge<1><<<grid, kStrBlock, 0, s>>>(keys, m, n, kmax, src_idx, src_is_a, c_out, consumed_a);
  else
    k_row_merge<0><<<grid, kStrBlock, 0, s>>>(keys, m, n, kmax, src_idx, src_is_a, c_out, consumed_a);
  YSQL_LAUNCH_CHECK();
  return YSQL_OK;
}

__global__ void k_store_i64_s(int64_t *p, int64_t v) { *p = v; }

static bool aligned8(const void *p) { return (reinterpret_cast<uintptr_t>(p) & 7) == 0; }

static size_t string_sort_ws(uint64_t n, size_t *radix_ws) {
  size_t r = 0;
  if (ysql_radix_sort_ws_bytes(YSQL_UKEY64, n, &r) != YSQL_OK) return 0;
  if (radix_ws) *radix_ws = r;
  WsCounter c;
  c.take<char>(r);
  c.take<uint64_t>(n);  // chunk keys
  c.take<uint64_t>(n);  // sorted chunk keys (discarded)
  c.take<uint32_t>(n);  // the other permutation buffer
  c.take<uint16_t>(n);  // lengths
  c.take<uint32_t>(64);  // longest length
  return c.off;
}

}  // namespace ysql

using namespace ysql;

extern "C" {

int ysql_string_sort_ws_bytes(uint64_t n, size_t *bytes) {
  if (!bytes || n >= (1ull << 30)) return YSQL_EINVAL;
  *bytes = string_sort_ws(n, nullptr);
  return *bytes ? YSQL_OK : YSQL_EINVAL;
}

int ysql_string_sort(int descending, const void *strs, uint32_t *perm_inout, uint64_t n, void *ws, size_t ws_bytes,
                     ysql_stream_t stream) {
  if (n >= (1ull << 30)) return YSQL_EINVAL;
  if (n <= 1) return YSQL_OK;
  if (!strs || !perm_inout || !aligned8(strs)) return YSQL_EINVAL;
  size_t rws = 0;
  const size_t need = string_sort_ws(n, &rws);
  if (!ws || ws_bytes < need) return YSQL_EWORKSPACE;
  hipStream_t s = as_stream(stream);
  WsCarver c(ws, ws_bytes);
  char *radix_ws = c.take<char>(rws);
  uint64_t *keys = c.take<uint64_t>(n), *keys_sorted = c.take<uint64_t>(n);
  uint32_t *alt = c.take<uint32_t>(n);
  uint16_t *len = c.take<uint16_t>(n);
  uint32_t *max_len = c.take<uint32_t>(64);
  const char *sp = static_cast<const char *>(strs);
  int rc = zero_words(max_len, 8, s);
  if (rc) return rc;
  const unsigned g = (unsigned)std::min<uint64_t>(ceil_div(n, kStrBlock), (uint64_t)grid_cap());
  k_str_len<<<g, kStrBlock, 0, s>>>(sp, n, len, max_len);
  YSQL_LAUNCH_CHECK();
  uint32_t longest = 0;
  YSQL_HIP_TRY(hipMemcpyAsync(&longest, max_len, 4, hipMemcpyDeviceToHost, s));
  YSQL_HIP_TRY(hipStreamSynchronize(s));  // the number of chunk passes
  const int chunks = (int)ceil_div(longest, 8);
  const uint64_t flip = descending ? ~0ull : 0ull;
  uint32_t *cur = perm_inout, *nxt = alt;
  const unsigned gk = (unsigned)ceil_div(n, kStrBlock);
  for (int ch = chunks - 1; ch >= 0; ch--) {
    k_str_chunk_key<<<gk, kStrBlock, 0, s>>>(sp, cur, len, n, ch, flip, keys);
    YSQL_LAUNCH_CHECK();
    rc = ysql_radix_sort_ex(YSQL_UKEY64, 0, keys, cur, 0, keys_sorted, nxt, n, 0, 64, radix_ws, rws, stream);
    if (rc) return rc;
    std::swap(cur, nxt);
  }
  if (cur != perm_inout) {
    k_copy_u32<<<gk, kStrBlock, 0, s>>>(cur, perm_inout, n);
    YSQL_LAUNCH_CHECK();
  }
  return YSQL_OK;
}

int ysql_string_argsort(int descending, const void *strs, uint32_t *perm_out, uint64_t n, void *ws, size_t ws_bytes,
                        ysql_stream_t stream) {
  if (n >= (1ull << 30) || (n && !perm_out)) return YSQL_EINVAL;
  if (n == 0) return YSQL_OK;
  if (!strs || !aligned8(strs)) return YSQL_EINVAL;
  size_t need = 0;
  if (n > 1 && (ysql_string_sort_ws_bytes(n, &need) != YSQL_OK || !ws || ws_bytes < need)) return YSQL_EWORKSPACE;
  k_iota_u32<<<(unsigned)ceil_div(n, kStrBlock), kStrBlock, 0, as_stream(stream)>>>(perm_out, n);
  YSQL_LAUNCH_CHECK();
  return ysql_string_sort(descending, strs, perm_out, n, ws, ws_bytes, stream);
}

int ysql_merge_rows(const ysql_sortcol *keys, int nkeys, uint64_t m, uint64_t n, uint64_t k_max, uint32_t *src_idx,
                    uint8_t *src_is_a, void *keys_out, int64_t *consumed_a_out, ysql_stream_t stream) {
  if (!keys || nkeys < 1 || nkeys > YSQL_MAX_COLS) return YSQL_EINVAL;
  if (m >= (1ull << 32) || n >= (1ull << 32)) return YSQL_EINVAL;
  RowKeys rk;
  rk.nk = nkeys;
  for (int c = 0; c < nkeys; c++) {
    const int t = keys[c].dtype;
    if (t != YSQL_INT && t != YSQL_FLOAT && t != YSQL_DATETIME && t != YSQL_STRING) return YSQL_EINVAL;
    if ((m && !keys[c].a) || (n && !keys[c].b)) return YSQL_EINVAL;
    const size_t al = t == YSQL_STRING ? 8 : (size_t)dtype_bytes(t);
    if ((reinterpret_cast<uintptr_t>(keys[c].a) | reinterpret_cast<uintptr_t>(keys[c].b)) % al) return YSQL_EINVAL;
    rk.col[c] = KeyCol{t, keys[c].descending ? 1 : 0, static_cast<const char *>(keys[c].a),
                       static_cast<const char *>(keys[c].b)};
  }
  if (keys_out && (nkeys != 1 || rk.col[0].dtype != YSQL_STRING || (reinterpret_cast<uintptr_t>(keys_out) & 15) ||
                   (reinterpret_cast<uintptr_t>(rk.col[0].a) & 15) || (reinterpret_cast<uintptr_t>(rk.col[0].b) & 15)))
    return YSQL_EINVAL;
  hipStream_t s = as_stream(stream);
  const uint64_t kmax = k_max < m + n ? k_max : m + n;
  if (kmax == 0) {
    if (consumed_a_out) {
      k_store_i64_s<<<1, 1, 0, s>>>(consumed_a_out, 0);
      YSQL_LAUNCH_CHECK();
    }
    return YSQL_OK;
  }
  if (!src_idx || !src_is_a) return YSQL_EINVAL;
  return row_merge(rk, m, n, kmax, src_idx, src_is_a, static_cast<char *>(keys_out), consumed_a_out, s);
}

}  // extern "C"

// ysql_merge_path for YSQL_STRING (merge.hip dispatches here).
namespace ysql {
int merge_path_strings(int descending, const void *A, uint64_t m, const void *B, uint64_t n, uint64_t kmax, void *C_opt,
                       uint32_t *src_idx, uint8_t *src_is_a, int64_t *consumed_a_out, hipStream_t s) {
  if ((m && !aligned8(A)) || (n && !aligned8(B))) return YSQL_EINVAL;
  if (C_opt && ((reinterpret_cast<uintptr_t>(C_opt) | reinterpret_cast<uintptr_t>(A) |
                 reinterpret_cast<uintptr_t>(B)) & 15))
    return YSQL_EINVAL;
  RowKeys rk;
  rk.nk = 1;
  rk.col[0] = KeyCol{YSQL_STRING, descending ? 1 : 0, static_cast<const char *>(A), static_cast<const char *>(B)};
  return row_merge(rk, m, n, kmax, src_idx, src_is_a, static_cast<char *>(C_opt), consumed_a_out, s);
}
}  // namespace ysql
