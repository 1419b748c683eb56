// merge.hip — merge path of two sorted runs.
//
// Reference: merge_sorted_array_kernel_v1 (merge_batchs.cu:32-77) gives
// every thread 6 outputs and a private co-rank binary search in global
// memory, with a strict comparison (:22) that disagrees with the A-first tie
// rule of its sequential loop (:61), so equal keys spanning both runs are
// duplicated and lost.  Its LDS-tiled variant is commented out (:78-141).
//
// Here: k_merge_partition finds, for every tile of 2048 outputs, the
// co-rank (number of A elements among the first k outputs) with the A-first
// rule: i = min{i : not (A[i] <= B[k-1-i])}.  k_merge then stages the tile's
// A and B ranges in LDS with coalesced loads, each thread finds its own
// co-rank inside LDS and merges 8 outputs sequentially (A first on ties), and
// the merged tile leaves through LDS with coalesced stores.
#include "common.hpp"

namespace ysql {

constexpr int kMergeBlock = 256;
constexpr int kMergeIpt = 8;
constexpr int kMergeTile = kMergeBlock * kMergeIpt;  // 2048

// a <= b in merge order (AscSortMergeOp / DescSortMergeOp is_less_or_equal,
// merge_batchs.hpp:20-34).
template <typename T, bool DESC>
__device__ __forceinline__ bool mle(T a, T b) {
  return DESC ? (a >= b) : (a <= b);
}

template <typename T, bool DESC>
__device__ __forceinline__ uint64_t corank(const T *A, uint64_t m, const T *B, uint64_t n, uint64_t k) {
  uint64_t lo = k > n ? k - n : 0;
  uint64_t hi = k < m ? k : m;
  while (lo < hi) {
    uint64_t mid = (lo + hi) >> 1;
    if (mle<T, DESC>(A[mid], B[k - 1 - mid])) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// One thread per partition point.  (A wave-wide 65-way search needs fewer
// dependent steps but touches 64 cache lines per step; measured slower.)
template <typename T, bool DESC>
__global__ __launch_bounds__(256) void k_merge_partition(const T *A, uint64_t m, const T *B, uint64_t n, uint64_t kmax,
                                                         uint64_t nparts, uint64_t *part, int64_t *consumed_a) {
  const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p > nparts) return;
  uint64_t diag = p * kMergeTile;
  if (diag > kmax) diag = kmax;
  const uint64_t i = corank<T, DESC>(A, m, B, n, diag);
  part[p] = i;
  if (p == nparts && consumed_a) *consumed_a = (int64_t)i;
}

// MODE 0: (key, payload) pairs -> out_keys/out_vals.
// MODE 1: reference outputs -> src_idx, src_is_a, optional keys C.
template <typename T, bool DESC, int MODE>
__global__ __launch_bounds__(kMergeBlock) void k_merge(const T *__restrict__ A, const uint32_t *__restrict__ av,
                                                       uint64_t m, const T *__restrict__ B,
                                                       const uint32_t *__restrict__ bv, uint64_t n, uint64_t kmax,
                                                       const uint64_t *__restrict__ part, T *__restrict__ outk,
                                                       uint32_t *__restrict__ outv, uint8_t *__restrict__ out_is_a) {
  __shared__ T s_k[kMergeTile];
  __shared__ uint32_t s_v[kMergeTile];
  __shared__ uint8_t s_a[MODE == 1 ? kMergeTile : 1];
  const int t = threadIdx.x;
  const uint64_t tile = blockIdx.x;
  const uint64_t k0 = tile * kMergeTile;
  const uint64_t k1 = (k0 + kMergeTile < kmax) ? k0 + kMergeTile : kmax;
  const uint64_t i0 = part[tile], i1 = part[tile + 1];
  const uint64_t j0 = k0 - i0, j1 = k1 - i1;
  const int na = (int)(i1 - i0), nb = (int)(j1 - j0), nt = na + nb;

  for (int x = t; x < nt; x += kMergeBlock) {
    if (x < na) {
      s_k[x] = A[i0 + x];
      if (MODE == 0) s_v[x] = av[i0 + x];
    } else {
      s_k[x] = B[j0 + (x - na)];
      if (MODE == 0) s_v[x] = bv[j0 + (x - na)];
    }
  }
  __syncthreads();

  const T *sA = s_k;
  const T *sB = s_k + na;
  int d = t * kMergeIpt;
  if (d > nt) d = nt;
  int lo = d > nb ? d - nb : 0;
  int hi = d < na ? d : na;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (mle<T, DESC>(sA[mid], sB[d - 1 - mid])) lo = mid + 1;
    else hi = mid;
  }
  int i = lo, j = d - lo;
  T ok[kMergeIpt];
  uint32_t ov[kMergeIpt];
  uint8_t oa[kMergeIpt];
#pragma unroll
  for (int e = 0; e < kMergeIpt; e++) {
    if (d + e < nt) {
      const bool take_a = (j >= nb) || (i < na && mle<T, DESC>(sA[i], sB[j]));
      if (take_a) {
        ok[e] = sA[i];
        ov[e] = MODE == 0 ? s_v[i] : (uint32_t)(i0 + i);
        oa[e] = 1;
        i++;
      } else {
        ok[e] = sB[j];
        ov[e] = MODE == 0 ? s_v[na + j] : (uint32_t)(j0 + j);
        oa[e] = 0;
        j++;
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int e = 0; e < kMergeIpt; e++) {
    if (d + e < nt) {
      s_k[d + e] = ok[e];
      s_v[d + e] = ov[e];
      if (MODE == 1) s_a[d + e] = oa[e];
    }
  }
  __syncthreads();
  for (int x = t; x < nt; x += kMergeBlock) {
    if (outk) outk[k0 + x] = s_k[x];
    outv[k0 + x] = s_v[x];
    if (MODE == 1) out_is_a[k0 + x] = s_a[x];
  }
}

static size_t merge_ws_size(uint64_t kmax) {
  WsCounter c;
  c.take<uint64_t>(ceil_div(kmax, kMergeTile) + 2);
  return c.off;
}

template <typename T, bool DESC, int MODE>
static int merge_launch(const T *A, const uint32_t *av, uint64_t m, const T *B, const uint32_t *bv, uint64_t n,
                        uint64_t kmax, T *outk, uint32_t *outv, uint8_t *out_is_a, int64_t *consumed_a, void *ws,
                        size_t ws_bytes, hipStream_t s) {
  const uint64_t nparts = ceil_div(kmax, kMergeTile);
  WsCarver c(ws, ws_bytes);
  uint64_t *part = c.take<uint64_t>(nparts + 2);
  if (!part) return YSQL_EWORKSPACE;
  k_merge_partition<T, DESC><<<(unsigned)ceil_div(nparts + 1, 256), 256, 0, s>>>(A, m, B, n, kmax, nparts, part,
                                                                                 consumed_a);
  YSQL_LAUNCH_CHECK();
  if (nparts > 0) {
    k_merge<T, DESC, MODE><<<(unsigned)nparts, kMergeBlock, 0, s>>>(A, av, m, B, bv, n, kmax, part, outk, outv,
                                                                    out_is_a);
    YSQL_LAUNCH_CHECK();
  }
  return YSQL_OK;
}

template <typename T, int MODE>
static int merge_dir(int desc, const T *A, const uint32_t *av, uint64_t m, const T *B, const uint32_t *bv, uint64_t n,
                     uint64_t kmax, T *outk, uint32_t *outv, uint8_t *out_is_a, int64_t *consumed_a, void *ws,
                     size_t ws_bytes, hipStream_t s) {
  if (desc)
    return merge_launch<T, true, MODE>(A, av, m, B, bv, n, kmax, outk, outv, out_is_a, consumed_a, ws, ws_bytes, s);
  return merge_launch<T, false, MODE>(A, av, m, B, bv, n, kmax, outk, outv, out_is_a, consumed_a, ws, ws_bytes, s);
}

__global__ void k_store_i64(int64_t *p, int64_t v) { *p = v; }

}  // namespace ysql

using namespace ysql;

extern "C" {

int ysql_merge_ws_bytes(uint64_t m, uint64_t n, uint64_t k_max, size_t *bytes) {
  if (!bytes) return YSQL_EINVAL;
  uint64_t kmax = k_max < m + n ? k_max : m + n;
  *bytes = merge_ws_size(kmax);
  return YSQL_OK;
}

int ysql_merge_path(int dtype, int descending, const void *A, uint64_t m, const void *B, uint64_t n, uint64_t k_max,
                    void *C_opt, uint32_t *src_idx, uint8_t *src_is_a, int64_t *consumed_a_out, void *ws,
                    size_t ws_bytes, ysql_stream_t stream) {
  hipStream_t s = as_stream(stream);
  const uint64_t kmax = k_max < m + n ? k_max : m + n;
  if (m >= (1ull << 32) || n >= (1ull << 32)) return YSQL_EINVAL;
  if ((m && !A) || (n && !B)) return YSQL_EINVAL;
  if (kmax == 0) {
    if (consumed_a_out) {
      k_store_i64<<<1, 1, 0, s>>>(consumed_a_out, 0);
      YSQL_LAUNCH_CHECK();
    }
    return YSQL_OK;
  }
  if (!src_idx || !src_is_a) return YSQL_EINVAL;
  if (!ws || ws_bytes < merge_ws_size(kmax)) return YSQL_EWORKSPACE;
  switch (dtype) {
    case YSQL_INT:
      return merge_dir<int32_t, 1>(descending, (const int32_t *)A, nullptr, m, (const int32_t *)B, nullptr, n, kmax,
                                   (int32_t *)C_opt, src_idx, src_is_a, consumed_a_out, ws, ws_bytes, s);
    case YSQL_FLOAT:
      return merge_dir<float, 1>(descending, (const float *)A, nullptr, m, (const float *)B, nullptr, n, kmax,
                                 (float *)C_opt, src_idx, src_is_a, consumed_a_out, ws, ws_bytes, s);
    case YSQL_DATETIME:
      return merge_dir<int64_t, 1>(descending, (const int64_t *)A, nullptr, m, (const int64_t *)B, nullptr, n, kmax,
                                   (int64_t *)C_opt, src_idx, src_is_a, consumed_a_out, ws, ws_bytes, s);
    case YSQL_UKEY32:
      return merge_dir<uint32_t, 1>(descending, (const uint32_t *)A, nullptr, m, (const uint32_t *)B, nullptr, n, kmax,
                                    (uint32_t *)C_opt, src_idx, src_is_a, consumed_a_out, ws, ws_bytes, s);
    case YSQL_UKEY64:
      return merge_dir<uint64_t, 1>(descending, (const uint64_t *)A, nullptr, m, (const uint64_t *)B, nullptr, n, kmax,
                                    (uint64_t *)C_opt, src_idx, src_is_a, consumed_a_out, ws, ws_bytes, s);
  }
  return YSQL_EINVAL;
}

int ysql_merge_pairs(int key_dtype, int descending, const void *a_keys, const uint32_t *a_vals, uint64_t m,
                     const void *b_keys, const uint32_t *b_vals, uint64_t n, void *out_keys, uint32_t *out_vals,
                     void *ws, size_t ws_bytes, ysql_stream_t stream) {
  hipStream_t s = as_stream(stream);
  const uint64_t kmax = m + n;
  if (kmax == 0) return YSQL_OK;
  if (m >= (1ull << 32) || n >= (1ull << 32)) return YSQL_EINVAL;
  if ((m && (!a_keys || !a_vals)) || (n && (!b_keys || !b_vals)) || !out_keys || !out_vals) return YSQL_EINVAL;
  if (!ws || ws_bytes < merge_ws_size(kmax)) return YSQL_EWORKSPACE;
  if (key_dtype == YSQL_UKEY32)
    return merge_dir<uint32_t, 0>(descending, (const uint32_t *)a_keys, a_vals, m, (const uint32_t *)b_keys, b_vals, n,
                                  kmax, (uint32_t *)out_keys, out_vals, nullptr, nullptr, ws, ws_bytes, s);
  if (key_dtype == YSQL_UKEY64)
    return merge_dir<uint64_t, 0>(descending, (const uint64_t *)a_keys, a_vals, m, (const uint64_t *)b_keys, b_vals, n,
                                  kmax, (uint64_t *)out_keys, out_vals, nullptr, nullptr, ws, ws_bytes, s);
  return YSQL_EINVAL;
}

}  // extern "C"
