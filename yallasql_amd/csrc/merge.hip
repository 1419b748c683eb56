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
#include <cstring>

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

// STRING merge path (strings.hip).
int merge_path_strings(int descending, const void *A, uint64_t m, const void *B, uint64_t n, uint64_t kmax, void *C_opt,
                       uint32_t *src_idx, uint8_t *src_is_a, int64_t *consumed_a_out, hipStream_t s);

// ---------------------------------------------------------------------------
// k-way merge of R <= 16 sorted runs in one pass over the data (the
// reference merges runs pairwise, level by level: every row moves
// ceil(log2 R) times, order_operator.hpp:160-193).  Merged order: ascending
// key, ties by run, then by position in the run (for runs cut in row order
// and sorted stably, one global stable sort).
//   1. Every run's element at positions S-1, 2S-1, ... is a sample (key, its
//      global position); the samples are radix-sorted by key (stable, so
//      ties stay in run-then-position order, the merged order).
//   2. Every G-th sorted sample j is a partition boundary: for every run r,
//      how many of r's elements come before or at the sample in merged order
//      (r before the sample's run: keys <= key_j; after it: keys < key_j;
//      its own run: position + 1).  These splits are monotone in j.  If k_r
//      of the G samples of a partition belong to run r, r contributes fewer
//      than (k_r + 1) * S elements (S consecutive elements of a run hold one
//      of its samples), so a partition holds fewer than (G + R) * S =
//      kRunPart elements and about G * S on average.
//   3. One workgroup per partition stages its R ranges in LDS, merges them
//      there pairwise, level by level (merge path per thread, earlier run
//      first on ties), and writes the partition with coalesced stores.
// Keys and payloads move through HBM once each way.
// ---------------------------------------------------------------------------
constexpr int kMaxRuns = 16;
constexpr int kRunPart = 4096;  // elements per partition, at most
constexpr int kRunBlock = 512;
constexpr int kRunIpt = kRunPart / kRunBlock;  // 8

constexpr uint32_t kRunStride = 128;  // S

struct RunArgs {
  uint64_t off[kMaxRuns + 1];  // run r = [off[r], off[r + 1]) of the input
  int nruns;
  uint32_t stride;  // S
  uint32_t group;   // G = kRunPart / S - R: samples per partition
  uint32_t pad;
};

template <typename K>
__global__ void k_run_samples(const K *__restrict__ keys, RunArgs a, K *__restrict__ skey, uint32_t *__restrict__ spos) {
  // samples of run r occupy [sbase_r, sbase_r + (len_r / S)) in run order
  const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t sb = 0;
  for (int r = 0; r < a.nruns; r++) {
    const uint64_t m = (a.off[r + 1] - a.off[r]) / a.stride;
    if (tid < sb + m) {
      const uint64_t p = a.off[r] + (tid - sb + 1) * a.stride - 1;
      skey[tid] = keys[p];
      spos[tid] = (uint32_t)p;
      return;
    }
    sb += m;
  }
}

// Number of keys of run[0, len) that come before key v: < v (strict), or <= v.
template <typename K, bool LE>
__device__ __forceinline__ uint64_t run_bound(const K *run, uint64_t len, K v) {
  uint64_t lo = 0, hi = len;
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    const bool before = LE ? !(v < run[mid]) : (run[mid] < v);
    if (before) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// One thread per (partition boundary b, run r): split[b * R + r], the
// boundary being sorted sample (b + 1) * G - 1.
template <typename K>
__global__ void k_run_splits(const K *__restrict__ keys, RunArgs a, const K *__restrict__ skey,
                             const uint32_t *__restrict__ spos, uint64_t nbounds, uint32_t *__restrict__ split) {
  const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t b = tid / (uint64_t)a.nruns;
  const int r = (int)(tid % (uint64_t)a.nruns);
  if (b >= nbounds) return;
  const uint64_t j = (b + 1) * a.group - 1;
  const K v = skey[j];
  const uint64_t p = spos[j];
  int rj = 0;
  while (rj + 1 < a.nruns && p >= a.off[rj + 1]) rj++;
  const uint64_t len = a.off[r + 1] - a.off[r];
  uint64_t c;
  if (r == rj) c = p - a.off[r] + 1;
  else if (r < rj) c = run_bound<K, true>(keys + a.off[r], len, v);
  else c = run_bound<K, false>(keys + a.off[r], len, v);
  split[b * a.nruns + r] = (uint32_t)c;
}

template <typename K>
struct RunSmem {
  K k[kRunPart];
  uint32_t v[kRunPart];
  uint32_t lo[kMaxRuns], len[kMaxRuns], at[kMaxRuns + 1];
};

template <typename K>
__global__ __launch_bounds__(kRunBlock) void k_run_merge(const K *__restrict__ keys, const uint32_t *__restrict__ vals,
                                                         RunArgs a, const uint32_t *__restrict__ split,
                                                         uint64_t nbounds, K *__restrict__ outk,
                                                         uint32_t *__restrict__ outv) {
  __shared__ RunSmem<K> sm;
  const int t = threadIdx.x;
  const int R = a.nruns;
  const uint64_t j = blockIdx.x;  // partition j: after boundary j-1, up to and including boundary j
  if (t < R) {
    const uint64_t len_r = a.off[t + 1] - a.off[t];
    const uint32_t lo = j == 0 ? 0u : split[(j - 1) * R + t];
    const uint32_t hi = j == nbounds ? (uint32_t)len_r : split[j * R + t];
    sm.lo[t] = lo;
    sm.len[t] = hi - lo;
  }
  __syncthreads();
  if (t == 0) {
    uint32_t acc = 0;
    for (int r = 0; r < R; r++) {
      sm.at[r] = acc;
      acc += sm.len[r];
    }
    sm.at[R] = acc;
  }
  __syncthreads();
  const uint32_t total = sm.at[R];
  // output position of the partition: every run's elements before it
  uint64_t base = 0;
  for (int r = 0; r < R; r++) base += sm.lo[r];
  // ---- stage the R ranges (each run's range contiguous in LDS)
  {
    // every load of the thread in flight before its first LDS store
    K stk[kRunIpt];
    uint32_t stv[kRunIpt];
#pragma unroll
    for (int q = 0; q < kRunIpt; q++) {
      const uint32_t x = t + q * kRunBlock;
      if (x < total) {
        int r = 0;
        while (x >= sm.at[r + 1]) r++;
        const uint64_t g = a.off[r] + sm.lo[r] + (x - sm.at[r]);
        stk[q] = keys[g];
        stv[q] = vals[g];
      }
    }
#pragma unroll
    for (int q = 0; q < kRunIpt; q++) {
      const uint32_t x = t + q * kRunBlock;
      if (x < total) {
        sm.k[x] = stk[q];
        sm.v[x] = stv[q];
      }
    }
  }
  __syncthreads();
  // ---- merge the staged runs in LDS, adjacent spans pairwise, level by
  // level (runs 2m and 2m+1 at width 1, then pairs of pairs, ...): a merged
  // pair occupies the span its two parts held, so one buffer serves every
  // level.  Thread t produces the outputs [o0, o1): the co-rank of its first
  // output in its pair (the number of left elements among the first d
  // outputs, left first on ties), then a sequential merge; a thread whose
  // outputs run into the next pair starts that pair's co-rank afresh.
  const uint32_t ipt = (total + kRunBlock - 1) / kRunBlock;
  const uint32_t o0 = t * ipt < total ? t * ipt : total;
  const uint32_t o1 = o0 + ipt < total ? o0 + ipt : total;
#pragma unroll 1
  for (int w = 1; w < R; w <<= 1) {
    K ok[kRunIpt];
    uint32_t ov[kRunIpt];
    uint32_t s1 = 0, s2 = 0, ia = 0, ib = 0;
    K ka = 0, kb = 0;  // the keys at ia and ib, kept in registers
#pragma unroll
    for (int q = 0; q < kRunIpt; q++) {
      const uint32_t o = o0 + q;
      if (o < o1) {
        if (o >= s2) {
          int m = 0;
          while (sm.at[min(R, (2 * m + 2) * w)] <= o) m++;
          const uint32_t s0 = sm.at[min(R, 2 * m * w)];
          s1 = sm.at[min(R, (2 * m + 1) * w)];
          s2 = sm.at[min(R, (2 * m + 2) * w)];
          const uint32_t d = o - s0, na = s1 - s0, nb = s2 - s1;
          uint32_t lo = d > nb ? d - nb : 0u, hi = d < na ? d : na;
          while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (!(sm.k[s1 + d - 1 - mid] < sm.k[s0 + mid])) lo = mid + 1;  // A[mid] <= B[d-1-mid]
            else hi = mid;
          }
          ia = s0 + lo;
          ib = s1 + (d - lo);
          ka = ia < s1 ? sm.k[ia] : (K)0;
          kb = ib < s2 ? sm.k[ib] : (K)0;
        }
        const bool take_a = ib >= s2 || (ia < s1 && !(kb < ka));
        if (take_a) {
          ok[q] = ka;
          ov[q] = sm.v[ia];
          ia++;
          if (ia < s1) ka = sm.k[ia];
        } else {
          ok[q] = kb;
          ov[q] = sm.v[ib];
          ib++;
          if (ib < s2) kb = sm.k[ib];
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kRunIpt; q++)
      if (o0 + q < o1) {
        sm.k[o0 + q] = ok[q];
        sm.v[o0 + q] = ov[q];
      }
    __syncthreads();
  }
  for (uint32_t x = t; x < total; x += kRunBlock) {
    outk[base + x] = sm.k[x];
    outv[base + x] = sm.v[x];
  }
}

struct RunWs {
  void *skey, *skey2;
  uint32_t *spos, *spos2, *split;
  void *sort_ws;
  size_t sort_ws_bytes;
};

template <typename K>
static size_t runs_ws_layout(uint64_t nsamples, int nruns, void *ws, size_t bytes, RunWs *out) {
  size_t sort_bytes = 0;
  ysql_radix_sort_ws_bytes(sizeof(K) == 8 ? YSQL_UKEY64 : YSQL_UKEY32, nsamples, &sort_bytes);
  const uint64_t m = nsamples ? nsamples : 1;
  if (out) {
    WsCarver c(ws, bytes);
    out->skey = c.take<K>(m);
    out->skey2 = c.take<K>(m);
    out->spos = c.take<uint32_t>(m);
    out->spos2 = c.take<uint32_t>(m);
    out->split = c.take<uint32_t>(m * (uint64_t)nruns);  // >= boundaries * R
    out->sort_ws = c.take<uint8_t>(sort_bytes);
    out->sort_ws_bytes = sort_bytes;
    return (out->skey && out->skey2 && out->spos && out->spos2 && out->split && out->sort_ws) ? c.off : 0;
  }
  WsCounter c;
  c.take<K>(m);
  c.take<K>(m);
  c.take<uint32_t>(m);
  c.take<uint32_t>(m);
  c.take<uint32_t>(m * (uint64_t)nruns);
  c.take<uint8_t>(sort_bytes);
  return c.off;
}

static bool runs_args(const uint64_t *off, int nruns, RunArgs *a, uint64_t *n, uint64_t *nsamples) {
  if (!off || nruns < 1 || nruns > kMaxRuns) return false;
  memset(a, 0, sizeof(*a));
  a->nruns = nruns;
  a->stride = kRunStride;
  a->group = (uint32_t)kRunPart / kRunStride - (uint32_t)nruns;
  uint64_t ns = 0;
  for (int r = 0; r <= nruns; r++) {
    a->off[r] = off[r] - off[0];
    if (r > 0 && off[r] < off[r - 1]) return false;
  }
  for (int r = 0; r < nruns; r++) ns += (a->off[r + 1] - a->off[r]) / a->stride;
  *n = a->off[nruns];
  *nsamples = ns;
  return *n < (1ull << 32);
}

template <typename K>
static int merge_runs_t(const K *keys, const uint32_t *vals, const RunArgs &a, uint64_t n, uint64_t nsamples,
                        K *outk, uint32_t *outv, void *ws, size_t ws_bytes, hipStream_t s) {
  RunWs w;
  if (!ws || ws_bytes < runs_ws_layout<K>(nsamples, a.nruns, nullptr, 0, nullptr) ||
      !runs_ws_layout<K>(nsamples, a.nruns, ws, ws_bytes, &w))
    return YSQL_EWORKSPACE;
  if (nsamples) {
    k_run_samples<K><<<(unsigned)ceil_div(nsamples, 256), 256, 0, s>>>(keys, a, (K *)w.skey, w.spos);
    YSQL_LAUNCH_CHECK();
    const int rc = ysql_radix_sort_ex(sizeof(K) == 8 ? YSQL_UKEY64 : YSQL_UKEY32, 0, w.skey, w.spos, 0, w.skey2,
                                      w.spos2, nsamples, 0, 8 * (int)sizeof(K), w.sort_ws, w.sort_ws_bytes, s);
    if (rc) return rc;
  }
  const uint64_t nbounds = nsamples / a.group;
  if (nbounds) {
    const uint64_t pairs = nbounds * (uint64_t)a.nruns;
    k_run_splits<K><<<(unsigned)ceil_div(pairs, 256), 256, 0, s>>>(keys, a, (const K *)w.skey2, w.spos2, nbounds,
                                                                   w.split);
    YSQL_LAUNCH_CHECK();
  }
  k_run_merge<K><<<(unsigned)(nbounds + 1), kRunBlock, 0, s>>>(keys, vals, a, w.split, nbounds, outk, outv);
  YSQL_LAUNCH_CHECK();
  return YSQL_OK;
}

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
    case YSQL_STRING:
      return merge_path_strings(descending, A, m, B, n, kmax, C_opt, src_idx, src_is_a, consumed_a_out, s);
  }
  return YSQL_EINVAL;
}

int ysql_merge_runs_ws_bytes(int key_dtype, const uint64_t *run_offsets, int nruns, size_t *bytes) {
  RunArgs a;
  uint64_t n = 0, ns = 0;
  if (!bytes || !runs_args(run_offsets, nruns, &a, &n, &ns)) return YSQL_EINVAL;
  if (key_dtype == YSQL_UKEY32) *bytes = runs_ws_layout<uint32_t>(ns, nruns, nullptr, 0, nullptr);
  else if (key_dtype == YSQL_UKEY64) *bytes = runs_ws_layout<uint64_t>(ns, nruns, nullptr, 0, nullptr);
  else return YSQL_EINVAL;
  return YSQL_OK;
}

int ysql_merge_runs(int key_dtype, const void *keys, const uint32_t *vals, const uint64_t *run_offsets, int nruns,
                    void *keys_out, uint32_t *vals_out, void *ws, size_t ws_bytes, ysql_stream_t stream) {
  RunArgs a;
  uint64_t n = 0, ns = 0;
  if (!runs_args(run_offsets, nruns, &a, &n, &ns)) return YSQL_EINVAL;
  if (n == 0) return YSQL_OK;
  if (!keys || !vals || !keys_out || !vals_out) return YSQL_EINVAL;
  // the merge writes while partitions still read: distinct buffers only
  if (keys == keys_out || vals == vals_out) return YSQL_EINVAL;
  hipStream_t s = as_stream(stream);
  const uint64_t o0 = run_offsets[0];
  if (key_dtype == YSQL_UKEY32)
    return merge_runs_t<uint32_t>((const uint32_t *)keys + o0, vals + o0, a, n, ns, (uint32_t *)keys_out, vals_out, ws,
                                  ws_bytes, s);
  if (key_dtype == YSQL_UKEY64)
    return merge_runs_t<uint64_t>((const uint64_t *)keys + o0, vals + o0, a, n, ns, (uint64_t *)keys_out, vals_out, ws,
                                  ws_bytes, s);
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
