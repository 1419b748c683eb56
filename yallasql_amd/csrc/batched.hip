// batched.hip — thousands of ragged sort / filter / aggregate jobs replayed
// as one captured hipGraph.
//
// ysql_batch_create assigns the jobs to K independent chains (largest job
// first onto the least-loaded chain; a job's cost is the HBM bytes it moves
// plus a fixed per-launch cost), gives every chain its own workspace and
// stream, and captures each chain's per-call launches between one fork event
// and one join event into a single hipGraph.  ysql_batch_run replays it: one
// host call for the whole job set, no per-launch host cost, and the chains'
// kernels overlap on the device, so small jobs no longer run one at a time
// on an otherwise idle GPU.
//
// Jobs up to a size limit per kind (the tuning's seg_*_max) do not get a
// chain: each runs in ONE workgroup of a segmented kernel (one launch per
// kind and key width for all of them), which walks the job's tiles in order
// with running offsets in LDS and needs no communication between
// workgroups.  Jobs are dealt to workgroups largest first.
//
// ysql_batch_run_plain issues the per-call entry points one job after
// another on one stream (the operator-at-a-time pattern of the reference's
// engine).  Every path produces the same bits: sorts and compactions are
// stable (unique results), and the segmented aggregate folds the same
// canonical tiles with the same device functions as the per-call kernel.
#include <algorithm>
#include <numeric>
#include <vector>

#include "agg_dev.hpp"
#include "radix_dev.hpp"
#include "scan_dev.hpp"

namespace ysql {

// ---------------------------------------------------------------- segmented
constexpr int kSegSortBlock = 256, kSegSortKpt = 16;  // 4096-key tiles
constexpr int kSegFiltBlock = 256;                    // 4096-row compaction tiles

struct SegSort {
  const void *keys;
  void *keys_out;
  uint32_t *perm_out;
  void *kalt;
  uint32_t *valt;
  uint32_t n, pad;
};

struct SegFilter {
  const void *keys;
  const float *vals;
  void *keys_out;
  float *vals_out;
  uint64_t *count_out;
  float thr;
  uint32_t n, kb, pad;
};

struct SegAgg {
  const void *keys;
  const float *vals;
  ysql_acc *accs_out;
  ysql_acc *partials;
  float thr;
  uint32_t n;
  int key64, pad;
};

template <typename K>
__global__ __launch_bounds__(kSegSortBlock) void k_seg_sort(const SegSort *jobs) {
  __shared__ SegSortSmem<K, kSegSortBlock, kSegSortKpt> sm;
  const SegSort j = jobs[blockIdx.x];
  seg_sort_block<K, kSegSortBlock, kSegSortKpt>(sm, (const K *)j.keys, 0u, (K *)j.keys_out, j.perm_out, (K *)j.kalt,
                                               j.valt, j.n, 0, (int)sizeof(K), 0u);
}

__global__ __launch_bounds__(kSegFiltBlock) void k_seg_filter(const SegFilter *jobs) {
  __shared__ CompactSmem<kSegFiltBlock, 4> sm;
  const SegFilter j = jobs[blockIdx.x];
  if (j.n == 0) {
    if (threadIdx.x == 0) *j.count_out = 0;
    return;
  }
  const ysql_col cols[2] = {{j.keys, nullptr, j.keys_out, nullptr, j.kb, 0}, {j.vals, nullptr, j.vals_out, nullptr, 4, 0}};
  constexpr uint32_t TILE = kSegFiltBlock * 16;
  const uint32_t ntiles = (j.n + TILE - 1) / TILE;
  uint64_t running = 0;
  for (uint32_t tile = 0; tile < ntiles; tile++)
    running += compact_tile<kSegFiltBlock, 1, float, YSQL_LT>(sm, tile, nullptr, j.vals, j.thr, 0.f, cols, 2, j.n,
                                                               ~0ull, running, j.count_out, nullptr, ntiles);
}

// SUM(vals), MIN(keys), MAX(keys) WHERE vals < thr: the specs
// ysql_filter_aggregate_ex builds for the same job.
__device__ __forceinline__ void seg_agg_specs(const SegAgg &j, FaSpec (&sp)[3]) {
  const int kdt = j.key64 ? YSQL_DATETIME : YSQL_INT;
  sp[0] = FaSpec{FA_KD, 0, YSQL_FLOAT, FA_PRED, j.vals};
  sp[1] = FaSpec{FA_KI, 1, kdt, FA_OWN, j.keys};
  sp[2] = FaSpec{FA_KI, 2, kdt, FA_SPEC0 + 1, j.keys};
}

__global__ __launch_bounds__(kFaBlock) void k_seg_agg(const SegAgg *jobs) {
  __shared__ ysql_acc s_w[kFaBlock / kWave];
  const SegAgg j = jobs[blockIdx.x];
  FaSpec sp[3];
  seg_agg_specs(j, sp);
  ysql_acc *const out[3] = {j.accs_out, j.accs_out + 1, j.accs_out + 2};
  if (j.n == 0) {
    if (threadIdx.x < 3) {
      ysql_acc a;
      fa_identity(a, sp[threadIdx.x]);
      *out[threadIdx.x] = a;
    }
    return;
  }
  const uint64_t tile_rows = fa_tile_rows(j.n);
  const uint32_t ntiles = (uint32_t)ceil_div(j.n, tile_rows);
  const float sc[1] = {j.thr};
  for (uint32_t tile = 0; tile < ntiles; tile++)
    filter_agg_tile<float, YSQL_LT, kFaBlock, 1>(tile, tile_rows, j.vals, sc, 1, sp, 3, j.n, j.partials, s_w);
  __threadfence();
  __syncthreads();
  filter_agg_finish<kFaBlock>(sp, 3, 1, ntiles, j.partials, out, true, s_w);
}

constexpr int kMaxChains = 64;

// Bytes a job moves through HBM, plus a fixed cost per launch (in bytes at
// HBM rate), for the longest-first chain assignment.
static double job_cost(const ysql_job &x) {
  const double kb = x.key_dtype == YSQL_UKEY64 ? 8.0 : 4.0;
  const double n = (double)x.n;
  constexpr double kLaunch = 40e3;  // ~5 us of launch gap at ~8 TB/s
  switch (x.kind) {
    case YSQL_JOB_SORT: {
      const double passes = kb;  // 8-bit digits
      const double launches = x.n <= 8192 ? 1.0 : passes + 2.0;  // one tile sorts in one launch
      return n * kb + passes * 2.0 * n * (kb + 4.0) + launches * kLaunch;
    }
    case YSQL_JOB_FILTER: return n * (kb + 4.0) * 1.5 + 2.0 * kLaunch;
    default: return n * (kb + 4.0) + 2.0 * kLaunch;
  }
}

static size_t job_ws_bytes(const ysql_job &x) {
  size_t w = 0;
  if (x.kind == YSQL_JOB_SORT) ysql_radix_sort_ws_bytes(x.key_dtype, x.n, &w);
  else if (x.kind == YSQL_JOB_FILTER) ysql_compact_ws_bytes(x.n, &w);
  else ysql_filter_aggregate_ws_bytes(x.n, 3, &w);
  return std::max<size_t>(w, 256);
}

static int sort_variant(const ysql_batch_tuning &t, const ysql_job &x) {
  int b = 0;
  while (b < 31 && (2ull << b) <= x.n) b++;  // floor(log2(n))
  return t.sort_variant[x.key_dtype == YSQL_UKEY64 ? 1 : 0][b];
}

// The per-call launches of one job on `stream` (workspace `ws`).
static int issue_job(const ysql_job &x, const ysql_batch_tuning &t, const float *thr_dev, void *ws, size_t ws_bytes,
                     ysql_stream_t stream) {
  const int kdt = x.key_dtype == YSQL_UKEY64 ? YSQL_DATETIME : YSQL_INT;
  if (x.kind == YSQL_JOB_SORT) {
    if (x.n == 0) return YSQL_OK;
    return ysql_radix_sort_variant(x.key_dtype, sort_variant(t, x), 0, x.keys, nullptr, 0, x.keys_out, x.perm_out,
                                   x.n, 0, x.key_dtype == YSQL_UKEY64 ? 64 : 32, ws, ws_bytes, stream);
  }
  if (x.kind == YSQL_JOB_FILTER) {
    const uint32_t kb = x.key_dtype == YSQL_UKEY64 ? 8 : 4;
    ysql_col cols[2] = {{x.keys, nullptr, x.keys_out, nullptr, kb, 0}, {x.vals, nullptr, x.vals_out, nullptr, 4, 0}};
    return ysql_filter_compact(YSQL_FLOAT, YSQL_LT, x.vals, thr_dev, cols, 2, x.n, x.count_out, ws, ws_bytes, stream);
  }
  ysql_agg_spec specs[3] = {{YSQL_FLOAT, YSQL_AGG_SUM, x.vals, x.accs_out},
                            {kdt, YSQL_AGG_MIN, x.keys, x.accs_out + 1},
                            {kdt, YSQL_AGG_MAX, x.keys, x.accs_out + 2}};
  return ysql_filter_aggregate_ex(YSQL_FLOAT, YSQL_LT, x.vals, thr_dev, specs, 3, x.n, YSQL_FA_OVERWRITE, ws, ws_bytes,
                                  stream);
}

}  // namespace ysql

using namespace ysql;

struct ysql_batch {
  std::vector<ysql_job> jobs;
  ysql_batch_tuning tune;
  float *d_thr = nullptr;  // thresholds as device scalars
  void *plain_ws = nullptr;
  size_t plain_ws_bytes = 0;
  int nchains = 0;
  std::vector<std::vector<int>> chain_jobs;
  std::vector<size_t> chain_ws_off, chain_ws_len;
  void *chain_ws = nullptr;
  size_t chain_ws_bytes = 0;
  // segmented jobs: device tables (largest first) and their workspace
  std::vector<SegSort> seg_sort32, seg_sort64;
  std::vector<SegFilter> seg_filter;
  std::vector<SegAgg> seg_agg;
  void *seg_dev = nullptr;  // the four tables, then per-job workspace
  size_t seg_bytes = 0;
  std::vector<hipStream_t> streams;  // streams[0] is the capture stream
  std::vector<hipEvent_t> events;    // fork, then one join per chain and per segmented launch
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

extern "C" {

void ysql_batch_tuning_default(ysql_batch_tuning *t) {
  if (!t) return;
  memset(t, 0, sizeof(*t));
  t->chains = 32;
  // Measured on 4,096 log-uniform jobs of 1K..4M rows: one-workgroup jobs
  // win up to these sizes because the many jobs fill the CUs side by side
  // (a batch of a few large jobs is better off with lower limits).
  t->seg_sort_max = 1 << 19;
  t->seg_filter_max = 1 << 21;
  t->seg_agg_max = 1 << 22;
  memset(t->sort_variant, -1, sizeof(t->sort_variant));
}

void ysql_batch_destroy(ysql_batch *b) {
  if (!b) return;
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  for (auto e : b->events) (void)hipEventDestroy(e);
  for (auto s : b->streams) (void)hipStreamDestroy(s);
  (void)hipFree(b->d_thr);
  (void)hipFree(b->plain_ws);
  (void)hipFree(b->chain_ws);
  (void)hipFree(b->seg_dev);
  delete b;
}

int ysql_batch_info(const ysql_batch *b, uint64_t *tasks, uint32_t *grid, uint64_t *ws_bytes) {
  if (!b) return YSQL_EINVAL;
  if (tasks) *tasks = b->jobs.size();
  if (grid) *grid = (uint32_t)b->nchains;
  if (ws_bytes) *ws_bytes = b->chain_ws_bytes + b->plain_ws_bytes + b->seg_bytes;
  return YSQL_OK;
}

int ysql_batch_create(const ysql_job *jobs, int njobs, const ysql_batch_tuning *tuning, ysql_batch **out) {
  if (!out || njobs < 0 || (njobs > 0 && !jobs)) return YSQL_EINVAL;
  *out = nullptr;
  ysql_batch_tuning tune;
  ysql_batch_tuning_default(&tune);
  if (tuning) tune = *tuning;
  tune.chains = std::max(1, std::min(tune.chains, kMaxChains));
  for (auto &row : tune.sort_variant)
    for (auto &v : row)
      if (v >= ysql_radix_variant_count()) return YSQL_EINVAL;
  for (int q = 0; q < njobs; q++) {
    const ysql_job &x = jobs[q];
    if (x.kind < YSQL_JOB_SORT || x.kind > YSQL_JOB_AGG) return YSQL_EINVAL;
    if (x.key_dtype != YSQL_UKEY32 && x.key_dtype != YSQL_UKEY64) return YSQL_EINVAL;
    if (x.n >= (1ull << 30)) return YSQL_EINVAL;
    if (x.kind == YSQL_JOB_FILTER && !x.count_out) return YSQL_EINVAL;
    if (x.kind == YSQL_JOB_AGG && !x.accs_out) return YSQL_EINVAL;
    if (x.n == 0) continue;
    if (!x.keys) return YSQL_EINVAL;
    if (x.kind == YSQL_JOB_SORT && (!x.keys_out || !x.perm_out)) return YSQL_EINVAL;
    if (x.kind == YSQL_JOB_FILTER && (!x.vals || !x.keys_out || !x.vals_out)) return YSQL_EINVAL;
    if (x.kind == YSQL_JOB_AGG && !x.vals) return YSQL_EINVAL;
  }
  auto *b = new ysql_batch();
  b->jobs.assign(jobs, jobs + njobs);
  b->tune = tune;
  auto fail = [&](hipError_t e) {
    set_last_error(hipGetErrorString(e));
    ysql_batch_destroy(b);
    return YSQL_EHIP;
  };
  hipError_t e;

  // ---- device thresholds and the plain loop's (largest) workspace
  std::vector<float> thr(std::max(1, njobs));
  std::vector<size_t> wsb(njobs);
  size_t pws = 256;
  for (int q = 0; q < njobs; q++) {
    thr[q] = b->jobs[q].threshold;
    wsb[q] = job_ws_bytes(b->jobs[q]);
    pws = std::max(pws, wsb[q]);
  }
  b->plain_ws_bytes = pws;
  if ((e = hipMalloc(&b->plain_ws, pws)) != hipSuccess) return fail(e);
  if ((e = hipMalloc(&b->d_thr, thr.size() * 4)) != hipSuccess) return fail(e);
  if ((e = hipMemcpy(b->d_thr, thr.data(), thr.size() * 4, hipMemcpyHostToDevice)) != hipSuccess) return fail(e);

  // ---- segmented jobs (one workgroup each) and their workspace
  std::vector<int> chained;
  {
    std::vector<int> segs;
    for (int q = 0; q < njobs; q++) {
      const ysql_job &x = b->jobs[q];
      const int lim = x.kind == YSQL_JOB_SORT ? tune.seg_sort_max
                                               : (x.kind == YSQL_JOB_FILTER ? tune.seg_filter_max : tune.seg_agg_max);
      if (x.n <= (uint64_t)std::max(lim, 0) && lim > 0) segs.push_back(q);
      else chained.push_back(q);
    }
    std::stable_sort(segs.begin(), segs.end(), [&](int a, int c) { return b->jobs[a].n > b->jobs[c].n; });
    // workspace offsets (256-byte pieces), after the tables
    size_t wso = 0;
    std::vector<size_t> wkalt(njobs, 0), wvalt(njobs, 0), wpart(njobs, 0);
    for (int q : segs) {
      const ysql_job &x = b->jobs[q];
      const size_t kb = x.key_dtype == YSQL_UKEY64 ? 8 : 4;
      if (x.kind == YSQL_JOB_SORT) {
        wkalt[q] = wso;
        wso += align_up(std::max<size_t>(x.n * kb, 1), 256);
        wvalt[q] = wso;
        wso += align_up(std::max<size_t>(x.n * 4, 1), 256);
      } else if (x.kind == YSQL_JOB_AGG) {
        wpart[q] = wso;
        wso += align_up((ceil_div(x.n, fa_tile_rows(x.n)) * 3 + 1) * sizeof(ysql_acc), 256);
      }
    }
    for (int q : segs) {
      const ysql_job &x = b->jobs[q];
      if (x.kind == YSQL_JOB_SORT) {
        SegSort j{x.keys, x.keys_out, x.perm_out, (void *)wkalt[q], (uint32_t *)wvalt[q], (uint32_t)x.n, 0};
        (x.key_dtype == YSQL_UKEY64 ? b->seg_sort64 : b->seg_sort32).push_back(j);
      } else if (x.kind == YSQL_JOB_FILTER) {
        b->seg_filter.push_back(SegFilter{x.keys, x.vals, x.keys_out, x.vals_out, x.count_out, x.threshold,
                                          (uint32_t)x.n, x.key_dtype == YSQL_UKEY64 ? 8u : 4u, 0});
      } else {
        b->seg_agg.push_back(SegAgg{x.keys, x.vals, x.accs_out, (ysql_acc *)wpart[q], x.threshold, (uint32_t)x.n,
                                    x.key_dtype == YSQL_UKEY64 ? 1 : 0, 0});
      }
    }
    const size_t t0 = align_up(b->seg_sort32.size() * sizeof(SegSort), 256);
    const size_t t1 = align_up(b->seg_sort64.size() * sizeof(SegSort), 256);
    const size_t t2 = align_up(b->seg_filter.size() * sizeof(SegFilter), 256);
    const size_t t3 = align_up(b->seg_agg.size() * sizeof(SegAgg), 256);
    const size_t tables = t0 + t1 + t2 + t3;
    b->seg_bytes = tables + wso;
    if ((e = hipMalloc(&b->seg_dev, std::max<size_t>(b->seg_bytes, 256))) != hipSuccess) return fail(e);
    char *base = (char *)b->seg_dev;
    char *wbase = base + tables;
    for (auto &j : b->seg_sort32) {
      j.kalt = wbase + (size_t)j.kalt;
      j.valt = (uint32_t *)(wbase + (size_t)j.valt);
    }
    for (auto &j : b->seg_sort64) {
      j.kalt = wbase + (size_t)j.kalt;
      j.valt = (uint32_t *)(wbase + (size_t)j.valt);
    }
    for (auto &j : b->seg_agg) j.partials = (ysql_acc *)(wbase + (size_t)j.partials);
    size_t off = 0;
    auto put = [&](const void *src, size_t bytes, size_t span) {
      if (bytes && (e = hipMemcpy(base + off, src, bytes, hipMemcpyHostToDevice)) != hipSuccess) return false;
      off += span;
      return true;
    };
    if (!put(b->seg_sort32.data(), b->seg_sort32.size() * sizeof(SegSort), t0) ||
        !put(b->seg_sort64.data(), b->seg_sort64.size() * sizeof(SegSort), t1) ||
        !put(b->seg_filter.data(), b->seg_filter.size() * sizeof(SegFilter), t2) ||
        !put(b->seg_agg.data(), b->seg_agg.size() * sizeof(SegAgg), t3))
      return fail(e);
  }

  // ---- chains: largest job first onto the least-loaded chain
  const int K = std::max(1, std::min(tune.chains, std::max(1, (int)chained.size())));
  b->nchains = K;
  b->chain_jobs.assign(K, {});
  std::vector<int> order = chained;
  std::vector<double> cost(njobs);
  for (int q = 0; q < njobs; q++) cost[q] = job_cost(b->jobs[q]);
  std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return cost[a] > cost[c]; });
  std::vector<double> load(K, 0.0);
  for (int q : order) {
    const int c = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    load[c] += cost[q];
    b->chain_jobs[c].push_back(q);
  }
  b->chain_ws_off.assign(K, 0);
  b->chain_ws_len.assign(K, 256);
  size_t off = 0;
  for (int c = 0; c < K; c++) {
    for (int q : b->chain_jobs[c]) b->chain_ws_len[c] = std::max(b->chain_ws_len[c], wsb[q]);
    b->chain_ws_off[c] = off;
    off += align_up(b->chain_ws_len[c], 256);
  }
  b->chain_ws_bytes = off;
  if ((e = hipMalloc(&b->chain_ws, std::max<size_t>(off, 256))) != hipSuccess) return fail(e);

  // ---- capture: fork -> K chains + the segmented launches -> join
  constexpr int kSegLaunches = 4;
  b->streams.assign(K + 1 + kSegLaunches, nullptr);
  for (auto &s : b->streams)
    if ((e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking)) != hipSuccess) return fail(e);
  b->events.assign(K + 1 + kSegLaunches, nullptr);
  for (auto &ev : b->events)
    if ((e = hipEventCreateWithFlags(&ev, hipEventDisableTiming)) != hipSuccess) return fail(e);
  hipStream_t cap = b->streams[0];
  if ((e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal)) != hipSuccess) return fail(e);
  int rc = YSQL_OK;
  bool ok = hipEventRecord(b->events[0], cap) == hipSuccess;
  for (int c = 0; c < K && ok && !rc; c++) {
    hipStream_t s = b->streams[c + 1];
    ok = hipStreamWaitEvent(s, b->events[0], 0) == hipSuccess;
    char *ws = (char *)b->chain_ws + b->chain_ws_off[c];
    for (size_t i = 0; ok && !rc && i < b->chain_jobs[c].size(); i++) {
      const int q = b->chain_jobs[c][i];
      rc = issue_job(b->jobs[q], b->tune, b->d_thr + q, ws, b->chain_ws_len[c], (ysql_stream_t)s);
    }
    ok = ok && hipEventRecord(b->events[c + 1], s) == hipSuccess;
    ok = ok && hipStreamWaitEvent(cap, b->events[c + 1], 0) == hipSuccess;
  }
  {
    const char *base = (const char *)b->seg_dev;
    const size_t t0 = align_up(b->seg_sort32.size() * sizeof(SegSort), 256);
    const size_t t1 = align_up(b->seg_sort64.size() * sizeof(SegSort), 256);
    const size_t t2 = align_up(b->seg_filter.size() * sizeof(SegFilter), 256);
    const size_t counts[kSegLaunches] = {b->seg_sort32.size(), b->seg_sort64.size(), b->seg_filter.size(),
                                         b->seg_agg.size()};
    for (int k = 0; k < kSegLaunches && ok && !rc; k++) {
      if (!counts[k]) continue;
      hipStream_t s = b->streams[K + 1 + k];
      ok = hipStreamWaitEvent(s, b->events[0], 0) == hipSuccess;
      const unsigned g = (unsigned)counts[k];
      if (k == 0) k_seg_sort<uint32_t><<<g, kSegSortBlock, 0, s>>>((const SegSort *)base);
      if (k == 1) k_seg_sort<uint64_t><<<g, kSegSortBlock, 0, s>>>((const SegSort *)(base + t0));
      if (k == 2) k_seg_filter<<<g, kSegFiltBlock, 0, s>>>((const SegFilter *)(base + t0 + t1));
      if (k == 3) k_seg_agg<<<g, kFaBlock, 0, s>>>((const SegAgg *)(base + t0 + t1 + t2));
      ok = ok && hipGetLastError() == hipSuccess;
      ok = ok && hipEventRecord(b->events[K + 1 + k], s) == hipSuccess;
      ok = ok && hipStreamWaitEvent(cap, b->events[K + 1 + k], 0) == hipSuccess;
    }
  }
  hipGraph_t g = nullptr;
  e = hipStreamEndCapture(cap, &g);
  b->graph = g;
  if (rc) {
    ysql_batch_destroy(b);
    return rc;
  }
  if (!ok || e != hipSuccess) return fail(e != hipSuccess ? e : hipErrorUnknown);
  if ((e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0)) != hipSuccess) return fail(e);
  *out = b;
  return YSQL_OK;
}

int ysql_batch_run(ysql_batch *b, ysql_stream_t stream) {
  if (!b) return YSQL_EINVAL;
  YSQL_HIP_TRY(hipGraphLaunch(b->exec, as_stream(stream)));
  return YSQL_OK;
}

int ysql_batch_run_plain(ysql_batch *b, ysql_stream_t stream) {
  if (!b) return YSQL_EINVAL;
  for (size_t q = 0; q < b->jobs.size(); q++) {
    const int rc = issue_job(b->jobs[q], b->tune, b->d_thr + q, b->plain_ws, b->plain_ws_bytes, stream);
    if (rc) return rc;
  }
  return YSQL_OK;
}

}  // extern "C"
