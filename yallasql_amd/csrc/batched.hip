// batched.hip — thousands of ragged sort / filter / aggregate jobs in one
// call.  Three forms, all giving the same bits:
//
// ysql_batch_run: ONE persistent kernel launch (k_persist, below) whose
// workgroups take whole jobs and chunk tasks from device-side job lists and
// a phase queue; no workgroup ever waits for another.
//
// ysql_batch_run_graph: a captured hipGraph.  Jobs up to a size limit per
// kind (the tuning's seg_*_max) each run in ONE workgroup of a segmented
// kernel (one launch per kind and key width for all of them), which walks
// the job's tiles in order with running offsets in LDS.  Larger jobs are
// assigned to K chains (largest first onto the least-loaded chain; a job's
// cost is the HBM bytes it moves plus a fixed per-launch cost), each with
// its own workspace and stream, and their per-call launches are captured
// between one fork event and one join event.
//
// ysql_batch_run_plain issues the per-call entry points one job after
// another on one stream (the operator-at-a-time pattern of the reference's
// engine).  Sorts and compactions are stable (unique results), and every
// aggregate path folds the same canonical tiles with the same device
// functions as the per-call kernel, so all three forms agree bit for bit.
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

// One filter job by one workgroup: its tiles compacted in order with a
// running prefix (no look-back).  Every thread of the block calls it.
__device__ __forceinline__ void seg_filter_job(CompactSmem<kSegFiltBlock, 4> &sm, const SegFilter &j) {
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

__global__ __launch_bounds__(kSegFiltBlock) void k_seg_filter(const SegFilter *jobs) {
  __shared__ CompactSmem<kSegFiltBlock, 4> sm;
  seg_filter_job(sm, jobs[blockIdx.x]);
}

// SUM(vals), MIN(keys), MAX(keys) WHERE vals < thr: the specs
// ysql_filter_aggregate_ex builds for the same job.
__device__ __forceinline__ void seg_agg_specs(const SegAgg &j, FaSpec (&sp)[3]) {
  const int kdt = j.key64 ? YSQL_DATETIME : YSQL_INT;
  sp[0] = FaSpec{FA_KD, 0, YSQL_FLOAT, FA_PRED, j.vals};
  sp[1] = FaSpec{FA_KI, 1, kdt, FA_OWN, j.keys};
  sp[2] = FaSpec{FA_KI, 2, kdt, FA_SPEC0 + 1, j.keys};
}

// One aggregate job by one workgroup: its canonical tiles folded in order,
// then combined like the per-call kernel's last block.
__device__ __forceinline__ void seg_agg_job(ysql_acc *s_w, const SegAgg &j) {
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
  const FaSelCmp<float, YSQL_LT, 1> sel = {{j.thr}};
  for (uint32_t tile = 0; tile < ntiles; tile++)
    filter_agg_tile<float, FaSelCmp<float, YSQL_LT, 1>, kFaBlock, 1>(tile, tile_rows, j.vals, sel, 1, sp, 3, j.n,
                                                                     j.partials, s_w);
  __threadfence();
  __syncthreads();
  filter_agg_finish<kFaBlock>(sp, 3, 1, ntiles, j.partials, out, true, s_w);
}

__global__ __launch_bounds__(kFaBlock) void k_seg_agg(const SegAgg *jobs) {
  __shared__ ysql_acc s_w[kFaBlock / kWave];
  seg_agg_job(s_w, jobs[blockIdx.x]);
}

// ------------------------------------------------------------- persistent
// One persistent launch for the whole job set.  Workgroups loop taking
// tasks from two device-side sources until every job is done:
//   * two job lists (cursors advanced with atomics): split sorts, then the
//     whole jobs largest first;
//   * a phase queue: entries (job, phase, ntasks) with an atomic task
//     counter; a workgroup takes one task (chunk) of the oldest entry that
//     still has untaken tasks.
// A whole job (filter, aggregate, or a sort up to the split size) is run by
// the workgroup that takes it, with the segmented bodies above.  A split
// sort runs as phases over chunks of `chunk` rows: U_p (each chunk's counts
// of digit p) then D_p (each chunk's tiles scattered in order from running
// per-digit offsets); the workgroup that completes the last task of a phase
// (the tasks' done counter) computes what the next phase needs (after U_p:
// every chunk's digit offsets, an exclusive scan over chunks) and pushes the
// next phase's entry.  No workgroup ever waits for another: a task reads
// only data whose producers completed before its entry was pushed.  A worker
// that finds nothing to take re-checks for a bounded number of rounds (a
// running task may push new work) and leaves; leaving is always safe,
// because the worker that completes a phase pushes its successor and then
// keeps taking tasks itself.  Release fences before every completion count
// and an acquire fence at the start of every task carry the data between
// workgroups.
constexpr int kPB = 256;            // threads per worker (4 waves)
constexpr uint32_t kPTile = kPB * 16;  // sort tile of the split phases (as the segmented sort)

enum { PJ_SORT = 0, PJ_SORT_SPLIT = 1, PJ_FILTER = 2, PJ_AGG = 3, PJ_FILTER_SPLIT = 4, PJ_AGG_SPLIT = 5 };
enum { PT_NONE = 0, PT_JOB = 1, PT_PHASE = 2, PT_EXIT = 3 };
// control words, 64 words (256 bytes) apart
enum { PC_CUR_SPLIT = 0, PC_CUR_WHOLE = 64, PC_HEAD = 128, PC_TAIL = 192, PC_DONE = 256, PC_ERR = 320, PC_WORDS = 384 };

struct PJob {
  int type, key64;
  uint32_t n, nchunks;
  const void *keys;
  void *keys_out;
  uint32_t *perm_out;
  void *kalt;
  uint32_t *valt;
  uint32_t *counts;   // split sort: nchunks * 256 words (counts, then offsets)
  uint64_t *fcounts;  // split filter: nchunks selected-row counts, then offsets
  uint32_t agg_tiles; // split aggregate: canonical tiles per task
  uint32_t pad;
  SegFilter f;
  SegAgg a;
};

struct PEntry {
  uint64_t desc;  // job << 32 | phase << 24 | ntasks; 0 while not yet pushed
  uint32_t taken, done;
};

struct PArgs {
  const PJob *jobs;
  const uint32_t *split_order, *whole_order;
  uint32_t nsplit, nwhole, total;  // total: jobs with rows
  PEntry *entries;
  uint32_t cap_entries;
  uint32_t *ctl;
  uint32_t chunk;          // rows per chunk, a multiple of kPTile
  uint32_t jobs_first_every;  // every k-th worker takes whole jobs before phases (0: none)
  uint32_t idle_limit;     // empty rounds before a worker leaves
};

union PSmem {
  SegSortSmem<uint64_t, kPB, 16> s64;
  SegSortSmem<uint32_t, kPB, 16> s32;
  CompactSmem<kSegFiltBlock, 4> cs;
  ysql_acc fa[kFaBlock / kWave];
};

__device__ __forceinline__ uint32_t pld(const uint32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t pld64(const uint64_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Cross-workgroup hand-offs.  The workgroup barrier first makes every
// thread's stores complete, then one thread's agent-scope fence writes the
// workgroup's data back (release) or drops stale cached lines (acquire):
// the cache operations act for the whole compute unit and L2, so one wave
// issuing them is enough.
__device__ __forceinline__ void wg_release() {
  __syncthreads();
  if (threadIdx.x == 0) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
}
__device__ __forceinline__ void wg_acquire() {
  if (threadIdx.x == 0) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  __syncthreads();
}

// Thread 0: the next index of a job list, or false.
__device__ __forceinline__ bool ptake_job(uint32_t *cur, uint32_t count, const uint32_t *order, uint32_t *job) {
  if (pld(cur) >= count) return false;
  const uint32_t i = atomicAdd(cur, 1u);
  if (i >= count) return false;
  *job = order[i];
  return true;
}

// Thread 0: one task of the oldest phase entry with untaken tasks, or false.
__device__ __forceinline__ bool ptake_phase(const PArgs &a, uint32_t *entry, uint32_t *task) {
  uint32_t tail = pld(&a.ctl[PC_TAIL]);
  if (tail > a.cap_entries) tail = a.cap_entries;
  uint32_t h = pld(&a.ctl[PC_HEAD]);
  for (uint32_t e = h; e < tail; e++) {
    PEntry *E = &a.entries[e];
    const uint64_t d = pld64(&E->desc);
    if (d == 0) continue;  // reserved, not yet written
    const uint32_t nt = (uint32_t)d & 0xFFFFFFu;
    if (pld(&E->taken) < nt) {
      const uint32_t t = atomicAdd(&E->taken, 1u);
      if (t < nt) {
        *entry = e;
        *task = t;
        return true;
      }
    }
    if (e == h) {  // every task of the head entry is taken: move the head on
      atomicCAS(&a.ctl[PC_HEAD], h, h + 1);
      h++;
    }
  }
  return false;
}

// Thread 0 (after the workgroup's writes are fenced): appends a phase entry.
__device__ __forceinline__ void ppush(const PArgs &a, uint32_t job, uint32_t phase, uint32_t ntasks) {
  const uint32_t e = atomicAdd(&a.ctl[PC_TAIL], 1u);
  if (e >= a.cap_entries) {  // cannot happen with the capacity ysql_batch_create sizes
    atomicOr(&a.ctl[PC_ERR], 1u);
    return;
  }
  const uint64_t d = ((uint64_t)job << 32) | ((uint64_t)phase << 24) | ntasks;
  __hip_atomic_store(&a.entries[e].desc, d, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void pjob_done(const PArgs &a) {
  wg_release();
  if (threadIdx.x == 0) atomicAdd(&a.ctl[PC_DONE], 1u);
}

// After U_p: counts[c][d] -> the global start of digit d's keys of chunk c
// (digit totals' exclusive scan plus the counts of earlier chunks).  Thread
// d walks the chunks 16 at a time with all 16 loads issued together.
__device__ __forceinline__ void psplit_offsets(uint32_t *counts, uint32_t nch, uint32_t *red) {
  constexpr uint32_t B = 16;
  const int t = threadIdx.x;
  uint32_t tot = 0;
  if (t < kBins) {
    for (uint32_t c = 0; c < nch; c += B) {
      uint32_t v[B];
#pragma unroll
      for (uint32_t u = 0; u < B; u++) v[u] = c + u < nch ? counts[(c + u) * kBins + t] : 0u;
#pragma unroll
      for (uint32_t u = 0; u < B; u++) tot += v[u];
    }
  }
  uint32_t all;
  uint32_t run = block_exclusive_sum<kPB, uint32_t>(t < kBins ? tot : 0u, red, &all);
  if (t < kBins) {
    for (uint32_t c = 0; c < nch; c += B) {
      uint32_t v[B];
#pragma unroll
      for (uint32_t u = 0; u < B; u++) v[u] = c + u < nch ? counts[(c + u) * kBins + t] : 0u;
#pragma unroll
      for (uint32_t u = 0; u < B; u++) {
        if (c + u < nch) counts[(c + u) * kBins + t] = run;
        run += v[u];
      }
    }
  }
}

template <typename K>
__device__ __noinline__ void psplit_task(const PArgs &a, SegSortSmem<K, kPB, 16> &sm, const PJob &j,
                                            uint32_t job, uint32_t phase, uint32_t entry, uint32_t c,
                                            uint32_t *s_flag) {
  const int passes = (int)sizeof(K);
  const int p = (int)(phase >> 1);
  const bool scatter = phase & 1u;
  const PassBufs<K> b = pass_bufs<K>(p, passes, (const K *)j.keys, (K *)j.keys_out, j.perm_out, (K *)j.kalt, j.valt);
  const uint64_t r0 = (uint64_t)c * a.chunk;
  const uint64_t r1 = r0 + a.chunk < j.n ? r0 + a.chunk : j.n;
  uint32_t *cnt = j.counts + (uint64_t)c * kBins;
  if (!scatter) {
    digit_counts_range<K>(b.src, r0, r1, 8 * p, 0u, sm.h, cnt);
  } else {
    if (threadIdx.x < kBins) sm.h[threadIdx.x] = cnt[threadIdx.x];
    __syncthreads();
    sort_pass_range<K, kPB, 16>(sm.os, sm.h, b.src, b.vsrc, 0u, b.dst, b.vdst, j.n, (uint32_t)(r0 / kPTile),
                                (uint32_t)ceil_div(r1, (uint64_t)kPTile), 8 * p, 0u);
  }
  // completion: the last task of the phase starts the next one
  wg_release();
  if (threadIdx.x == 0) *s_flag = atomicAdd(&a.entries[entry].done, 1u) == j.nchunks - 1 ? 1u : 0u;
  __syncthreads();
  if (!*s_flag) return;
  wg_acquire();  // every chunk's output of this phase
  if (!scatter) {
    psplit_offsets(j.counts, j.nchunks, sm.os.red);
    wg_release();
    if (threadIdx.x == 0) ppush(a, job, phase + 1, j.nchunks);
  } else if (p + 1 < passes) {
    if (threadIdx.x == 0) ppush(a, job, phase + 1, j.nchunks);
  } else {
    pjob_done(a);
  }
}

// Split filter: phase 0 counts each chunk's selected rows (the compaction's
// own loads and comparison); its last finisher turns the counts into
// offsets (exclusive scan over chunks); phase 1 compacts each chunk's tiles
// in order from its offset (the last tile writes *count_out).
__device__ __noinline__ void pfilter_task(const PArgs &a, CompactSmem<kSegFiltBlock, 4> &sm, const PJob &j,
                                          uint32_t job, uint32_t phase, uint32_t entry, uint32_t c,
                                          uint32_t *s_flag) {
  const SegFilter &f = j.f;
  constexpr uint32_t TILE = kSegFiltBlock * 16;
  const uint32_t ntiles = (f.n + TILE - 1) / TILE;
  const uint32_t tpc = a.chunk / TILE;
  const uint32_t t0 = c * tpc, t1 = t0 + tpc < ntiles ? t0 + tpc : ntiles;
  if (phase == 0) {
    uint32_t cnt = 0;
    for (uint32_t tile = t0; tile < t1; tile++) {
      const uint64_t base = (uint64_t)tile * TILE;
      const uint64_t rows = (f.n - base) < (uint64_t)TILE ? (f.n - base) : (uint64_t)TILE;
      float v[16];
      load16<float>(f.vals, base, rows, threadIdx.x, v);
#pragma unroll
      for (int e = 0; e < 16; e++) cnt += ((uint64_t)threadIdx.x * 16 + e < rows && cmp_static<YSQL_LT>(v[e], f.thr));
    }
    uint64_t tot;
    block_exclusive_sum<kSegFiltBlock, uint64_t>(cnt, sm.red, &tot);
    if (threadIdx.x == 0) j.fcounts[c] = tot;
  } else {
    const ysql_col cols[2] = {{f.keys, nullptr, f.keys_out, nullptr, f.kb, 0},
                              {f.vals, nullptr, f.vals_out, nullptr, 4, 0}};
    uint64_t running = j.fcounts[c];
    for (uint32_t tile = t0; tile < t1; tile++)
      running += compact_tile<kSegFiltBlock, 1, float, YSQL_LT>(sm, tile, nullptr, f.vals, f.thr, 0.f, cols, 2, f.n,
                                                                 ~0ull, running, f.count_out, nullptr, ntiles);
  }
  wg_release();
  if (threadIdx.x == 0) *s_flag = atomicAdd(&a.entries[entry].done, 1u) == j.nchunks - 1 ? 1u : 0u;
  __syncthreads();
  if (!*s_flag) return;
  wg_acquire();
  if (phase == 0) {
    uint64_t carry = 0;
    for (uint32_t b = 0; b < j.nchunks; b += kSegFiltBlock) {
      const uint32_t i = b + threadIdx.x;
      const uint64_t v = i < j.nchunks ? j.fcounts[i] : 0;
      uint64_t tot;
      const uint64_t ex = block_exclusive_sum<kSegFiltBlock, uint64_t>(v, sm.red, &tot);
      if (i < j.nchunks) j.fcounts[i] = carry + ex;
      carry += tot;
    }
    wg_release();
    if (threadIdx.x == 0) ppush(a, job, 1u, j.nchunks);
  } else {
    pjob_done(a);
  }
}

// Split aggregate: each task folds its run of canonical tiles into the
// per-tile partials; the last finisher combines every tile's partials in
// tile order, exactly as the one-workgroup and per-call forms do.
__device__ __noinline__ void pagg_task(const PArgs &a, ysql_acc *s_w, const PJob &j, uint32_t entry, uint32_t c,
                                       uint32_t *s_flag) {
  const SegAgg &g = j.a;
  FaSpec sp[3];
  seg_agg_specs(g, sp);
  const uint64_t tile_rows = fa_tile_rows(g.n);
  const uint32_t ntiles = (uint32_t)ceil_div((uint64_t)g.n, tile_rows);
  const uint32_t t0 = c * j.agg_tiles, t1 = t0 + j.agg_tiles < ntiles ? t0 + j.agg_tiles : ntiles;
  const FaSelCmp<float, YSQL_LT, 1> sel = {{g.thr}};
  for (uint32_t tile = t0; tile < t1; tile++)
    filter_agg_tile<float, FaSelCmp<float, YSQL_LT, 1>, kFaBlock, 1>(tile, tile_rows, g.vals, sel, 1, sp, 3, g.n,
                                                                     g.partials, s_w);
  wg_release();
  if (threadIdx.x == 0) *s_flag = atomicAdd(&a.entries[entry].done, 1u) == j.nchunks - 1 ? 1u : 0u;
  __syncthreads();
  if (!*s_flag) return;
  wg_acquire();
  ysql_acc *const out[3] = {g.accs_out, g.accs_out + 1, g.accs_out + 2};
  filter_agg_finish<kFaBlock>(sp, 3, 1, ntiles, g.partials, out, true, s_w);
  pjob_done(a);
}

// The task bodies stay separate functions: each keeps its own register
// allocation, so the kernel needs the largest body's registers, not more.
template <typename K>
__device__ __noinline__ void pwhole_sort(SegSortSmem<K, kPB, 16> &sm, const PJob &j) {
  seg_sort_block<K, kPB, 16>(sm, (const K *)j.keys, 0u, (K *)j.keys_out, j.perm_out, (K *)j.kalt, j.valt, j.n, 0,
                             (int)sizeof(K), 0u);
}
__device__ __noinline__ void pwhole_filter(CompactSmem<kSegFiltBlock, 4> &sm, const SegFilter &f) {
  seg_filter_job(sm, f);
}
__device__ __noinline__ void pwhole_agg(ysql_acc *s_w, const SegAgg &g) { seg_agg_job(s_w, g); }

__global__ __launch_bounds__(kPB, 3) void k_persist(const PArgs *__restrict__ pa) {
  const PArgs &a = *pa;
  __shared__ PSmem sm;
  __shared__ uint32_t s_kind, s_x, s_y, s_flag;
  const bool jobs_first = a.jobs_first_every && blockIdx.x % a.jobs_first_every == 0;
  uint32_t idle = 0;
  while (true) {
    if (threadIdx.x == 0) {
      uint32_t x = 0, y = 0, k = PT_NONE;
      if (jobs_first && ptake_job(&a.ctl[PC_CUR_WHOLE], a.nwhole, a.whole_order, &x)) k = PT_JOB;
      else if (ptake_phase(a, &x, &y)) k = PT_PHASE;
      else if (ptake_job(&a.ctl[PC_CUR_SPLIT], a.nsplit, a.split_order, &x)) k = PT_JOB;
      else if (ptake_job(&a.ctl[PC_CUR_WHOLE], a.nwhole, a.whole_order, &x)) k = PT_JOB;
      else if (pld(&a.ctl[PC_DONE]) >= a.total) k = PT_EXIT;
      s_kind = k;
      s_x = x;
      s_y = y;
    }
    __syncthreads();
    const uint32_t kind = s_kind, x = s_x, y = s_y;
    __syncthreads();
    if (kind == PT_EXIT) return;
    if (kind == PT_NONE) {
      if (++idle > a.idle_limit) return;
      __builtin_amdgcn_s_sleep(32);
      continue;
    }
    idle = 0;
    wg_acquire();  // see what the producers of this task's inputs wrote
    if (kind == PT_JOB) {
      const PJob &j = a.jobs[x];
      if (j.type == PJ_SORT_SPLIT || j.type == PJ_FILTER_SPLIT || j.type == PJ_AGG_SPLIT) {
        if (threadIdx.x == 0) ppush(a, x, 0u, j.nchunks);
        __syncthreads();
        continue;
      }
      if (j.type == PJ_SORT) {
        if (j.key64) pwhole_sort<uint64_t>(sm.s64, j);
        else pwhole_sort<uint32_t>(sm.s32, j);
      } else if (j.type == PJ_FILTER) {
        pwhole_filter(sm.cs, j.f);
      } else {
        pwhole_agg(sm.fa, j.a);
      }
      pjob_done(a);
    } else {
      const uint64_t d = pld64(&a.entries[x].desc);
      const uint32_t job = (uint32_t)(d >> 32), phase = (uint32_t)(d >> 24) & 0xFFu;
      const PJob &j = a.jobs[job];
      if (j.type == PJ_FILTER_SPLIT) pfilter_task(a, sm.cs, j, job, phase, x, y, &s_flag);
      else if (j.type == PJ_AGG_SPLIT) pagg_task(a, sm.fa, j, x, y, &s_flag);
      else if (j.key64) psplit_task<uint64_t>(a, sm.s64, j, job, phase, x, y, &s_flag);
      else psplit_task<uint32_t>(a, sm.s32, j, job, phase, x, y, &s_flag);
    }
    __syncthreads();
  }
}

// Per run, before k_persist: control words and phase entries zeroed, and
// the outputs of jobs without rows set (filter count 0, aggregate identities).
__global__ void k_persist_reset(uint32_t *ctl, PEntry *entries, uint32_t cap, const PJob *jobs, uint32_t njobs) {
  const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = tid; i < PC_WORDS; i += stride) ctl[i] = 0;
  for (uint64_t i = tid; i < cap; i += stride) entries[i] = PEntry{0, 0, 0};
  for (uint64_t q = tid; q < njobs; q += stride) {
    const PJob &j = jobs[q];
    if (j.n) continue;
    if (j.type == PJ_FILTER || j.type == PJ_FILTER_SPLIT) {
      *j.f.count_out = 0;
    } else if (j.type == PJ_AGG || j.type == PJ_AGG_SPLIT) {
      FaSpec sp[3];
      seg_agg_specs(j.a, sp);
      for (int s = 0; s < 3; s++) {
        ysql_acc acc;
        fa_identity(acc, sp[s]);
        j.a.accs_out[s] = acc;
      }
    }
  }
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
  // persistent executor: one buffer holding the job table, the two job
  // lists, the control words, the phase entries and every job's workspace
  void *p_dev = nullptr;
  size_t p_bytes = 0;
  PArgs p_args{};
  const PArgs *p_args_dev = nullptr;  // the same, in device memory (the kernel's argument)
  uint32_t p_workers = 0;
  hipStream_t p_stream = nullptr;
  hipGraph_t p_graph = nullptr;
  hipGraphExec_t p_exec = nullptr;
};

// Builds the persistent executor's tables and workspace (one allocation)
// and captures reset + persistent launch into b->p_graph.
static int persist_setup(ysql_batch *b) {
  const ysql_batch_tuning &t = b->tune;
  const int njobs = (int)b->jobs.size();
  const uint32_t chunk = (uint32_t)t.persist_chunk;
  const uint64_t split_min = (uint64_t)std::max(t.persist_split_min, 0);
  std::vector<PJob> pj(njobs);
  std::vector<uint32_t> split_order, whole_order;
  std::vector<size_t> o_kalt(njobs, 0), o_valt(njobs, 0), o_cnt(njobs, 0), o_part(njobs, 0);
  size_t wso = 0;  // workspace offsets, 256-byte pieces
  auto take = [&](size_t bytes) {
    const size_t o = wso;
    wso += align_up(std::max<size_t>(bytes, 1), 256);
    return o;
  };
  uint64_t cap = 0;
  uint32_t total = 0;
  for (int q = 0; q < njobs; q++) {
    const ysql_job &x = b->jobs[q];
    PJob &j = pj[q];
    memset(&j, 0, sizeof(j));
    j.key64 = x.key_dtype == YSQL_UKEY64;
    j.n = (uint32_t)x.n;
    j.keys = x.keys;
    const size_t kb = j.key64 ? 8 : 4;
    if (x.kind == YSQL_JOB_SORT) {
      j.type = x.n > split_min ? PJ_SORT_SPLIT : PJ_SORT;
      j.keys_out = x.keys_out;
      j.perm_out = x.perm_out;
      if (x.n) {
        o_kalt[q] = take(x.n * kb);
        o_valt[q] = take(x.n * 4);
      }
      if (j.type == PJ_SORT_SPLIT) {
        j.nchunks = (uint32_t)ceil_div(x.n, (uint64_t)chunk);
        o_cnt[q] = take((size_t)j.nchunks * kBins * 4);
        cap += 2 * kb;  // U_p and D_p for every digit pass
      }
    } else if (x.kind == YSQL_JOB_FILTER) {
      j.type = x.n > (uint64_t)std::max(t.persist_split_filter_min, 0) ? PJ_FILTER_SPLIT : PJ_FILTER;
      j.f = SegFilter{x.keys, x.vals, x.keys_out, x.vals_out, x.count_out, x.threshold, (uint32_t)x.n, (uint32_t)kb,
                      0};
      if (j.type == PJ_FILTER_SPLIT) {
        j.nchunks = (uint32_t)ceil_div(x.n, (uint64_t)chunk);
        o_cnt[q] = take((size_t)j.nchunks * 8);
        cap += 2;
      }
    } else {
      j.type = x.n > (uint64_t)std::max(t.persist_split_agg_min, 0) ? PJ_AGG_SPLIT : PJ_AGG;
      if (x.n) o_part[q] = take((ceil_div(x.n, fa_tile_rows(x.n)) * 3 + 1) * sizeof(ysql_acc));
      j.a = SegAgg{x.keys, x.vals, x.accs_out, nullptr, x.threshold, (uint32_t)x.n, j.key64 ? 1 : 0, 0};
      if (j.type == PJ_AGG_SPLIT) {
        const uint64_t tr = fa_tile_rows(x.n);
        j.agg_tiles = (uint32_t)std::max<uint64_t>(1, chunk / tr);
        j.nchunks = (uint32_t)ceil_div(ceil_div(x.n, tr), (uint64_t)j.agg_tiles);
        cap += 1;
      }
    }
    if (x.n) {
      const bool split = j.type == PJ_SORT_SPLIT || j.type == PJ_FILTER_SPLIT || j.type == PJ_AGG_SPLIT;
      (split ? split_order : whole_order).push_back((uint32_t)q);
      total++;
    }
  }
  std::stable_sort(split_order.begin(), split_order.end(),
                   [&](uint32_t a, uint32_t c) { return job_cost(b->jobs[a]) > job_cost(b->jobs[c]); });
  std::stable_sort(whole_order.begin(), whole_order.end(),
                   [&](uint32_t a, uint32_t c) { return job_cost(b->jobs[a]) > job_cost(b->jobs[c]); });
  cap = std::max<uint64_t>(cap, 1);
  // layout: jobs | split list | whole list | control words | entries | workspace
  const size_t sz_jobs = align_up(std::max<size_t>(pj.size(), 1) * sizeof(PJob), 256);
  const size_t sz_split = align_up(std::max<size_t>(split_order.size(), 1) * 4, 256);
  const size_t sz_whole = align_up(std::max<size_t>(whole_order.size(), 1) * 4, 256);
  const size_t sz_ctl = align_up(PC_WORDS * 4, 256);
  const size_t sz_ent = align_up(cap * sizeof(PEntry), 256);
  const size_t sz_args = align_up(sizeof(PArgs), 256);
  const size_t head = sz_jobs + sz_split + sz_whole + sz_ctl + sz_ent + sz_args;
  b->p_bytes = head + wso;
  hipError_t e;
  if ((e = hipMalloc(&b->p_dev, b->p_bytes)) != hipSuccess) {
    set_last_error(hipGetErrorString(e));
    return YSQL_EHIP;
  }
  char *base = (char *)b->p_dev;
  char *ws = base + head;
  for (int q = 0; q < njobs; q++) {
    PJob &j = pj[q];
    if (!j.n) continue;
    if (j.type == PJ_SORT || j.type == PJ_SORT_SPLIT) {
      j.kalt = ws + o_kalt[q];
      j.valt = (uint32_t *)(ws + o_valt[q]);
      if (j.type == PJ_SORT_SPLIT) j.counts = (uint32_t *)(ws + o_cnt[q]);
    } else if (j.type == PJ_AGG || j.type == PJ_AGG_SPLIT) {
      j.a.partials = (ysql_acc *)(ws + o_part[q]);
    } else if (j.type == PJ_FILTER_SPLIT) {
      j.fcounts = (uint64_t *)(ws + o_cnt[q]);
    }
  }
  PArgs &a = b->p_args;
  a.jobs = (const PJob *)base;
  a.split_order = (const uint32_t *)(base + sz_jobs);
  a.whole_order = (const uint32_t *)(base + sz_jobs + sz_split);
  a.nsplit = (uint32_t)split_order.size();
  a.nwhole = (uint32_t)whole_order.size();
  a.total = total;
  a.ctl = (uint32_t *)(base + sz_jobs + sz_split + sz_whole);
  a.entries = (PEntry *)(base + sz_jobs + sz_split + sz_whole + sz_ctl);
  a.cap_entries = (uint32_t)cap;
  a.chunk = chunk;
  a.jobs_first_every = (uint32_t)std::max(t.persist_jobs_first_every, 0);
  a.idle_limit = (uint32_t)std::max(t.persist_idle_limit, 1);
  b->p_args_dev = (const PArgs *)(base + sz_jobs + sz_split + sz_whole + sz_ctl + sz_ent);
  if ((e = hipMemcpy((void *)b->p_args_dev, &a, sizeof(PArgs), hipMemcpyHostToDevice)) != hipSuccess ||
      (njobs && (e = hipMemcpy(base, pj.data(), pj.size() * sizeof(PJob), hipMemcpyHostToDevice)) != hipSuccess) ||
      (!split_order.empty() && (e = hipMemcpy(base + sz_jobs, split_order.data(), split_order.size() * 4,
                                              hipMemcpyHostToDevice)) != hipSuccess) ||
      (!whole_order.empty() && (e = hipMemcpy(base + sz_jobs + sz_split, whole_order.data(), whole_order.size() * 4,
                                              hipMemcpyHostToDevice)) != hipSuccess)) {
    set_last_error(hipGetErrorString(e));
    return YSQL_EHIP;
  }
  // workers: as many as can be resident at once (none of them ever waits
  // for another, so more would only queue behind the resident ones)
  int per_cu = t.persist_workers_per_cu;
  if (per_cu <= 0) {
    if ((e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_persist, kPB, 0)) != hipSuccess) {
      set_last_error(hipGetErrorString(e));
      return YSQL_EHIP;
    }
  }
  per_cu = std::max(1, per_cu);
  b->p_workers = (uint32_t)(per_cu * grid_cap());
  if ((e = hipStreamCreateWithFlags(&b->p_stream, hipStreamNonBlocking)) != hipSuccess ||
      (e = hipStreamBeginCapture(b->p_stream, hipStreamCaptureModeThreadLocal)) != hipSuccess) {
    set_last_error(hipGetErrorString(e));
    return YSQL_EHIP;
  }
  {
    const uint64_t words = std::max<uint64_t>(std::max<uint64_t>(cap, PC_WORDS), (uint64_t)njobs);
    const unsigned g = (unsigned)std::min<uint64_t>(std::max<uint64_t>(ceil_div(words, 256), 1), 1024);
    k_persist_reset<<<g, 256, 0, b->p_stream>>>(a.ctl, a.entries, a.cap_entries, a.jobs, (uint32_t)njobs);
    if (total) k_persist<<<b->p_workers, kPB, 0, b->p_stream>>>(b->p_args_dev);
  }
  hipGraph_t g = nullptr;
  e = hipStreamEndCapture(b->p_stream, &g);
  b->p_graph = g;
  if (e != hipSuccess || (e = hipGraphInstantiate(&b->p_exec, b->p_graph, nullptr, nullptr, 0)) != hipSuccess) {
    set_last_error(hipGetErrorString(e));
    return YSQL_EHIP;
  }
  return YSQL_OK;
}

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
  t->persist_split_min = 1 << 19;
  t->persist_chunk = 1 << 18;
  t->persist_jobs_first_every = 4;
  t->persist_idle_limit = 1 << 15;
  t->persist_split_filter_min = 1 << 20;
  t->persist_split_agg_min = 1 << 20;
  t->persist_workers_per_cu = 0;
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
  if (b->p_exec) (void)hipGraphExecDestroy(b->p_exec);
  if (b->p_graph) (void)hipGraphDestroy(b->p_graph);
  if (b->p_stream) (void)hipStreamDestroy(b->p_stream);
  (void)hipFree(b->p_dev);
  delete b;
}

int ysql_batch_info(const ysql_batch *b, uint64_t *tasks, uint32_t *grid, uint64_t *ws_bytes) {
  if (!b) return YSQL_EINVAL;
  if (tasks) *tasks = b->jobs.size();
  if (grid) *grid = (uint32_t)b->nchains;
  if (ws_bytes) *ws_bytes = b->chain_ws_bytes + b->plain_ws_bytes + b->seg_bytes + b->p_bytes;
  return YSQL_OK;
}

int ysql_batch_create(const ysql_job *jobs, int njobs, const ysql_batch_tuning *tuning, ysql_batch **out) {
  if (!out || njobs < 0 || (njobs > 0 && !jobs)) return YSQL_EINVAL;
  *out = nullptr;
  ysql_batch_tuning tune;
  ysql_batch_tuning_default(&tune);
  if (tuning) tune = *tuning;
  tune.chains = std::max(1, std::min(tune.chains, kMaxChains));
  // split-sort chunks are whole tiles of the segmented sort
  tune.persist_chunk = (int)std::min<uint64_t>(
      align_up((uint64_t)std::max(tune.persist_chunk, (int)kPTile), (uint64_t)kPTile), 1ull << 28);
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
  if ((rc = persist_setup(b)) != YSQL_OK) {
    ysql_batch_destroy(b);
    return rc;
  }
  *out = b;
  return YSQL_OK;
}

int ysql_batch_run(ysql_batch *b, ysql_stream_t stream) {
  if (!b) return YSQL_EINVAL;
  YSQL_HIP_TRY(hipGraphLaunch(b->p_exec, as_stream(stream)));
  return YSQL_OK;
}

int ysql_batch_run_graph(ysql_batch *b, ysql_stream_t stream) {
  if (!b) return YSQL_EINVAL;
  YSQL_HIP_TRY(hipGraphLaunch(b->exec, as_stream(stream)));
  return YSQL_OK;
}

int ysql_batch_check(ysql_batch *b, ysql_stream_t stream, uint32_t *jobs_done, uint32_t *jobs_total) {
  if (!b || !jobs_done || !jobs_total) return YSQL_EINVAL;
  uint32_t w[2] = {0, 0};
  hipStream_t s = as_stream(stream);
  YSQL_HIP_TRY(hipMemcpyAsync(&w[0], b->p_args.ctl + PC_DONE, 4, hipMemcpyDeviceToHost, s));
  YSQL_HIP_TRY(hipMemcpyAsync(&w[1], b->p_args.ctl + PC_ERR, 4, hipMemcpyDeviceToHost, s));
  YSQL_HIP_TRY(hipStreamSynchronize(s));
  *jobs_done = w[0];
  *jobs_total = b->p_args.total;
  if (w[1]) {
    set_last_error("ysql_batch_check: the phase queue overflowed");
    return YSQL_EOVERFLOW;
  }
  return YSQL_OK;
}

int ysql_batch_workers(const ysql_batch *b, uint32_t *workers) {
  if (!b || !workers) return YSQL_EINVAL;
  *workers = b->p_workers;
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
