/*
 * cpu_baseline.cpp — multi-threaded CPU versions of the hot path, timed by
 * bench.py's CPU-baseline leg beside the GPU figures.
 *
 * TEST INFRASTRUCTURE ONLY (linked into oracle/liboracle.so with the
 * oracle).  Nothing in the product library calls these.  The reference has
 * no CPU executor; these follow the reference semantics the oracle
 * restates, with the algorithms a good CPU implementation would use:
 *   sort       per-thread stable LSD radix (8-bit digits) over each run,
 *              runs merged pairwise, each merge split over the threads by
 *              merge-path co-ranks (stable, earlier run first on ties) —
 *              the multi-run pipeline of OrderOperator
 *              (order_operator.hpp:89-193);
 *   filter     chunked predicate, per-thread counts, exclusive prefix,
 *              compaction (FilterOperator, filter_operator.hpp:31-98);
 *   reductions per-thread partials combined in thread order, float SUM in
 *              double (AggregateExpression, aggregate.hpp:82-173);
 *   EQ join    hash table of the right keys, probe split over the threads,
 *              pairs per thread concatenated in thread order;
 *   star join  direct-index join: dimension payloads indexed by the foreign
 *              key, per-thread private group arrays summed at the end.
 * Every function takes the thread count and splits work into contiguous
 * ranges, one per thread.
 */
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace {

template <typename F>
void parallel(int threads, F &&f) {
  if (threads <= 1) {
    f(0);
    return;
  }
  std::vector<std::thread> th;
  th.reserve((size_t)threads - 1);
  for (int t = 1; t < threads; t++) th.emplace_back(f, t);
  f(0);
  for (auto &x : th) x.join();
}

inline void range_of(uint64_t n, int threads, int t, uint64_t *b, uint64_t *e) {
  const uint64_t per = (n + (uint64_t)threads - 1) / (uint64_t)threads;
  *b = std::min(n, (uint64_t)t * per);
  *e = std::min(n, *b + per);
}

// Stable LSD radix sort of (k, v) over `passes` 8-bit digits, the threads
// splitting every pass: per-thread digit counts, a prefix over (digit,
// thread), then each thread scatters its range in order.
void radix_pairs(uint64_t *k, uint32_t *v, uint64_t n, uint64_t *tk, uint32_t *tv, int passes, int threads) {
  std::vector<uint64_t> cnt((size_t)threads * 256);
  uint64_t *sk = k, *dk = tk;
  uint32_t *sv = v, *dv = tv;
  for (int p = 0; p < passes; p++) {
    const int sh = 8 * p;
    std::fill(cnt.begin(), cnt.end(), 0);
    parallel(threads, [&](int t) {
      uint64_t b, e;
      range_of(n, threads, t, &b, &e);
      uint64_t *c = &cnt[(size_t)t * 256];
      for (uint64_t i = b; i < e; i++) c[(sk[i] >> sh) & 255]++;
    });
    uint64_t run = 0;
    for (int d = 0; d < 256; d++)
      for (int t = 0; t < threads; t++) {
        const uint64_t x = cnt[(size_t)t * 256 + d];
        cnt[(size_t)t * 256 + d] = run;
        run += x;
      }
    parallel(threads, [&](int t) {
      uint64_t b, e;
      range_of(n, threads, t, &b, &e);
      uint64_t *c = &cnt[(size_t)t * 256];
      for (uint64_t i = b; i < e; i++) {
        const uint64_t pos = c[(sk[i] >> sh) & 255]++;
        dk[pos] = sk[i];
        dv[pos] = sv[i];
      }
    });
    std::swap(sk, dk);
    std::swap(sv, dv);
  }
  if (sk != k) {
    memcpy(k, sk, n * 8);
    memcpy(v, sv, n * 4);
  }
}

// Number of A elements among the first d outputs of the stable merge (A
// first on ties).
uint64_t corank(const uint64_t *a, uint64_t m, const uint64_t *b, uint64_t n, uint64_t d) {
  uint64_t lo = d > n ? d - n : 0, hi = d < m ? d : m;
  while (lo < hi) {
    const uint64_t mid = (lo + hi) / 2;
    if (a[mid] <= b[d - 1 - mid]) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

void merge_parallel(const uint64_t *ak, const uint32_t *av, uint64_t m, const uint64_t *bk, const uint32_t *bv,
                    uint64_t n, uint64_t *ok, uint32_t *ov, int threads) {
  parallel(threads, [&](int t) {
    uint64_t d0, d1;
    range_of(m + n, threads, t, &d0, &d1);
    uint64_t i = corank(ak, m, bk, n, d0), j = d0 - i;
    for (uint64_t d = d0; d < d1; d++) {
      if (j >= n || (i < m && ak[i] <= bk[j])) {
        ok[d] = ak[i];
        ov[d] = av[i++];
      } else {
        ok[d] = bk[j];
        ov[d] = bv[j++];
      }
    }
  });
}

}  // namespace

extern "C" {

unsigned cpb_hardware_concurrency(void) { return std::thread::hardware_concurrency(); }

// Multi-run sort of n keys (row index payloads), runs of run_rows rows.
int cpb_multirun_sort_u64(const uint64_t *keys, uint64_t n, uint64_t run_rows, int threads, uint64_t *out_keys,
                          uint32_t *out_vals) {
  if (threads < 1) threads = 1;
  std::vector<uint64_t> k(keys, keys + n), tk(n);
  std::vector<uint32_t> v(n), tv(n);
  parallel(threads, [&](int t) {
    uint64_t b, e;
    range_of(n, threads, t, &b, &e);
    for (uint64_t i = b; i < e; i++) v[i] = (uint32_t)i;
  });
  uint64_t mx = 0;
  for (uint64_t i = 0; i < n; i++) mx |= keys[i];
  int passes = 0;
  while (passes < 8 && (mx >> (8 * passes))) passes++;
  std::vector<std::pair<uint64_t, uint64_t>> runs;
  for (uint64_t b = 0; b < n; b += run_rows) runs.push_back({b, std::min(n, b + run_rows)});
  for (auto &r : runs) radix_pairs(&k[r.first], &v[r.first], r.second - r.first, &tk[r.first], &tv[r.first], passes,
                                   threads);
  uint64_t *sk = k.data(), *dk = tk.data();
  uint32_t *sv = v.data(), *dv = tv.data();
  while (runs.size() > 1) {
    std::vector<std::pair<uint64_t, uint64_t>> nxt;
    for (size_t i = 0; i < runs.size(); i += 2) {
      if (i + 1 == runs.size()) {
        const auto r = runs[i];
        memcpy(dk + r.first, sk + r.first, (r.second - r.first) * 8);
        memcpy(dv + r.first, sv + r.first, (r.second - r.first) * 4);
        nxt.push_back(r);
        continue;
      }
      const auto a = runs[i], b = runs[i + 1];
      merge_parallel(sk + a.first, sv + a.first, a.second - a.first, sk + b.first, sv + b.first, b.second - b.first,
                     dk + a.first, dv + a.first, threads);
      nxt.push_back({a.first, b.second});
    }
    runs = nxt;
    std::swap(sk, dk);
    std::swap(sv, dv);
  }
  memcpy(out_keys, sk, n * 8);
  memcpy(out_vals, sv, n * 4);
  return 0;
}

// WHERE vals < thr: COUNT(*), SUM(vals) in double, MIN(keys), MAX(keys).
int cpb_filter_aggregate(const float *vals, const int64_t *keys, uint64_t n, float thr, int threads, uint64_t *count,
                         double *sum, int64_t *mn, int64_t *mx) {
  if (threads < 1) threads = 1;
  std::vector<uint64_t> c((size_t)threads, 0);
  std::vector<double> s((size_t)threads, 0.0);
  std::vector<int64_t> lo((size_t)threads, INT64_MAX), hi((size_t)threads, INT64_MIN);
  parallel(threads, [&](int t) {
    uint64_t b, e;
    range_of(n, threads, t, &b, &e);
    uint64_t cc = 0;
    double ss = 0;
    int64_t l = INT64_MAX, h = INT64_MIN;
    for (uint64_t i = b; i < e; i++)
      if (vals[i] < thr) {
        cc++;
        ss += vals[i];
        l = std::min(l, keys[i]);
        h = std::max(h, keys[i]);
      }
    c[t] = cc;
    s[t] = ss;
    lo[t] = l;
    hi[t] = h;
  });
  *count = 0;
  *sum = 0;
  *mn = INT64_MAX;
  *mx = INT64_MIN;
  for (int t = 0; t < threads; t++) {  // thread order
    *count += c[t];
    *sum += s[t];
    *mn = std::min(*mn, lo[t]);
    *mx = std::max(*mx, hi[t]);
  }
  return 0;
}

// WHERE vals < thr compacting (keys, vals); returns the count.
uint64_t cpb_filter_compact(const float *vals, const int64_t *keys, uint64_t n, float thr, int threads,
                            int64_t *out_keys, float *out_vals) {
  if (threads < 1) threads = 1;
  std::vector<uint64_t> c((size_t)threads + 1, 0);
  parallel(threads, [&](int t) {
    uint64_t b, e, cc = 0;
    range_of(n, threads, t, &b, &e);
    for (uint64_t i = b; i < e; i++) cc += vals[i] < thr;
    c[(size_t)t + 1] = cc;
  });
  for (int t = 0; t < threads; t++) c[(size_t)t + 1] += c[t];
  parallel(threads, [&](int t) {
    uint64_t b, e, o = c[t];
    range_of(n, threads, t, &b, &e);
    for (uint64_t i = b; i < e; i++)
      if (vals[i] < thr) {
        out_keys[o] = keys[i];
        out_vals[o++] = vals[i];
      }
  });
  return c[(size_t)threads];
}

// Inner EQ join: (l, r) pairs of left[l] == right[r], left-major (right
// matches of one left row in right order); writes at most cap pairs and
// returns the count.
uint64_t cpb_join_eq(const int64_t *left, uint64_t L, const int64_t *right, uint64_t R, int threads,
                     uint32_t *pairs, uint64_t cap) {
  if (threads < 1) threads = 1;
  uint64_t nb = 1;
  while (nb < 2 * R) nb <<= 1;
  // chained hash: head per bucket, next per right row (rows inserted in
  // reverse so each chain lists its rows in ascending order)
  std::vector<int64_t> head(nb, -1), next(R, -1);
  for (uint64_t r = R; r-- > 0;) {
    const uint64_t h = ((uint64_t)right[r] * 0x9E3779B97F4A7C15ull) >> 1;
    const uint64_t bkt = h & (nb - 1);
    next[r] = head[bkt];
    head[bkt] = (int64_t)r;
  }
  std::vector<std::vector<uint32_t>> part((size_t)threads);
  parallel(threads, [&](int t) {
    uint64_t b, e;
    range_of(L, threads, t, &b, &e);
    auto &out = part[t];
    for (uint64_t l = b; l < e; l++) {
      const uint64_t h = ((uint64_t)left[l] * 0x9E3779B97F4A7C15ull) >> 1;
      for (int64_t r = head[h & (nb - 1)]; r >= 0; r = next[r])
        if (right[r] == left[l]) {
          out.push_back((uint32_t)l);
          out.push_back((uint32_t)r);
        }
    }
  });
  uint64_t total = 0;
  for (auto &p : part)
    for (size_t i = 0; i + 1 < p.size(); i += 2, total++)
      if (total < cap) {
        pairs[2 * total] = p[i];
        pairs[2 * total + 1] = p[i + 1];
      }
  return total;
}

typedef struct cpb_dim {   /* the layout of the oracle's orc_star_dim */
  const int32_t *fk;       /* fact foreign keys into this dimension */
  uint32_t rows;           /* dimension rows */
  uint32_t group_card;     /* group values [group_base, group_base + group_card) */
  const uint8_t *pass;     /* per dimension row: passes the filter (NULL: all) */
  const int32_t *group;    /* per dimension row: group value (NULL: not a GROUP BY dim) */
  int32_t group_base;
  uint32_t pad;
} cpb_dim;

// SUM(int64(m1) * m2), COUNT(*) over the facts whose foreign keys all hit a
// dimension row that passes its filter, grouped by the GROUP BY dimensions'
// values (row-major over the dimensions in order).
int cpb_star_join_group_sum(const cpb_dim *dims, int ndims, const int32_t *m1, const int32_t *m2, uint64_t n,
                            int threads, int64_t *sums, uint64_t *counts, uint32_t n_groups) {
  if (threads < 1) threads = 1;
  uint64_t stride[8], s = 1;
  for (int d = ndims - 1; d >= 0; d--) {
    stride[d] = s;
    if (dims[d].group) s *= dims[d].group_card;
  }
  if (s > n_groups || ndims > 8) return -1;
  std::vector<std::vector<int64_t>> ps((size_t)threads, std::vector<int64_t>(n_groups, 0));
  std::vector<std::vector<uint64_t>> pc((size_t)threads, std::vector<uint64_t>(n_groups, 0));
  parallel(threads, [&](int t) {
    uint64_t b, e;
    range_of(n, threads, t, &b, &e);
    int64_t *S = ps[t].data();
    uint64_t *Cn = pc[t].data();
    for (uint64_t i = b; i < e; i++) {
      uint64_t g = 0;
      bool ok = true;
      for (int d = 0; d < ndims; d++) {
        const int32_t f = dims[d].fk[i];
        if (f < 0 || (uint32_t)f >= dims[d].rows || (dims[d].pass && !dims[d].pass[f])) {
          ok = false;
          break;
        }
        if (dims[d].group) {
          const int64_t v = (int64_t)dims[d].group[f] - dims[d].group_base;
          if (v < 0 || v >= (int64_t)dims[d].group_card) {
            ok = false;
            break;
          }
          g += (uint64_t)v * stride[d];
        }
      }
      if (!ok) continue;
      S[g] = (int64_t)((uint64_t)S[g] + (uint64_t)((int64_t)m1[i] * (int64_t)m2[i]));
      Cn[g]++;
    }
  });
  for (uint32_t g = 0; g < n_groups; g++) {
    int64_t a = 0;
    uint64_t c = 0;
    for (int t = 0; t < threads; t++) {
      a = (int64_t)((uint64_t)a + (uint64_t)ps[t][g]);
      c += pc[t][g];
    }
    sums[g] = a;
    counts[g] = c;
  }
  return 0;
}

// Config 1: WHERE price > 500 ORDER BY price DESC, id ASC; SELECT id, price,
// 2 * id.  Filter as cpb_filter_compact, then the composite key
// (~price key << 32 | id key) sorted by the parallel LSD radix sort.
// Returns the selected row count.
uint64_t cpb_config1(const int32_t *ids, const float *price, uint64_t n, int threads, int32_t *out_id,
                     float *out_price, int32_t *out_2id) {
  if (threads < 1) threads = 1;
  std::vector<uint64_t> c((size_t)threads + 1, 0);
  parallel(threads, [&](int t) {
    uint64_t b, e, cc = 0;
    range_of(n, threads, t, &b, &e);
    for (uint64_t i = b; i < e; i++) cc += price[i] > 500.0f;
    c[(size_t)t + 1] = cc;
  });
  for (int t = 0; t < threads; t++) c[(size_t)t + 1] += c[t];
  const uint64_t m = c[(size_t)threads];
  std::vector<uint64_t> key(m), tk(m);
  std::vector<uint32_t> row(m), tv(m);
  parallel(threads, [&](int t) {
    uint64_t b, e, o = c[t];
    range_of(n, threads, t, &b, &e);
    for (uint64_t i = b; i < e; i++)
      if (price[i] > 500.0f) {
        uint32_t pb;
        memcpy(&pb, &price[i], 4);
        const uint32_t pk = ~((pb & 0x80000000u) ? ~pb : (pb ^ 0x80000000u));  // DESC
        const uint32_t ik = (uint32_t)ids[i] ^ 0x80000000u;
        key[o] = ((uint64_t)pk << 32) | ik;
        row[o++] = (uint32_t)i;
      }
  });
  radix_pairs(key.data(), row.data(), m, tk.data(), tv.data(), 8, threads);
  parallel(threads, [&](int t) {
    uint64_t b, e;
    range_of(m, threads, t, &b, &e);
    for (uint64_t i = b; i < e; i++) {
      const uint32_t r = row[i];
      out_id[i] = ids[r];
      out_price[i] = price[r];
      out_2id[i] = (int32_t)(2u * (uint32_t)ids[r]);
    }
  });
  return m;
}

}  // extern "C"
