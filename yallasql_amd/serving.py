"""Serving loop: concurrent queries against HBM-resident tables.

No reference counterpart (the reference runs one query at a time from its
CLI).  Queries arrive over time (a Poisson process in the benchmark); the
dispatcher admits them as execution slots and device memory free up,
batches operators across the queries in flight, and keeps every
intermediate on the device; only final results come back to the host.

Query kinds over the table T(key int64, payload int32 = row id) and the
star schema of the star-join workload (a fact table F with four foreign keys
and two measures; dimensions date, supplier, customer, part):
  FILTER_AGG  SELECT COUNT(*), SUM(payload), MIN(key), MAX(key)
              FROM T WHERE key < c
  TOPK        SELECT key, payload FROM T WHERE key >= lo
              ORDER BY key LIMIT 1000        (ties in row order)
  JOIN_GROUP  SELECT year, brand, SUM(extendedprice * discount), COUNT(*)
              FROM F JOIN date JOIN supplier JOIN customer JOIN part
              WHERE customer.region = r AND part.category = c
              GROUP BY year, brand           (7,000 groups)

Operator batching across queries: up to 4 FILTER_AGG queries in flight
share one scan of the table (ysql_filter_aggregate_multi, one read of the
columns for all of them), and up to 8 JOIN_GROUP queries share one read of
the fact table (ysql_star_join_group_sum_multi, each query's filters as one
bit of the dimension filter bytes); TOPK pipelines run concurrently on
their own streams.  A query's result does not depend on what it was batched
with: the shared scan folds each query in the same order as a lone scan,
and every other operator runs per query.

Admission: each unit of work (a shared scan, or one TOPK / JOIN_GROUP
pipeline) reserves from the HBM budget the device memory it allocates,
computed from the same sizes and workspace queries its operators use
(each allocation rounded up to the device allocator's 512-byte blocks), and
waits (in arrival order) while the budget or every slot is taken.  A TOPK
whose key window needs a third widening (64 times the estimate) allocates
past its reservation.
"""
import collections
import contextlib
import math
import threading
import time

import numpy as np
import torch

from . import _lib as L
from . import ops
from .datagen import multirun_keys, star_schema

FILTER_AGG, TOPK, JOIN_GROUP = 0, 1, 2
KIND_NAMES = {FILTER_AGG: "filter_agg", TOPK: "topk", JOIN_GROUP: "join_group"}
TOPK_K = 1000
TOPK_CAPS = (16 * TOPK_K, 128 * TOPK_K)  # compaction capacity of the first two window widths
STAR_GROUPS = 7000  # 7 years x 1,000 brands
STAR_REGIONS, STAR_CATEGORIES = 7, 25
SHARED_SCAN_MAX = 4
SHARED_STAR_MAX = 8
I64_MAX = 2**63 - 1


class Table:
    """T(key, payload) and the star schema, resident in HBM."""

    def __init__(self, n, device, seed=1, star_rows=None, sample=1 << 20):
        self.n, self.device = n, device
        self.key = multirun_keys(n, device, seed=seed)
        self.payload = torch.arange(n, dtype=torch.int32, device=device)
        g = torch.Generator(device=device)
        g.manual_seed(seed + 100)
        # sorted key sample: statistics for choosing TOPK windows
        idx = torch.randint(0, n, (min(sample, n),), generator=g, device=device)
        self.sample = np.sort(self.key[idx].cpu().numpy())
        self.star_rows = n if star_rows is None else star_rows
        self.star = star_schema(self.star_rows, device, seed=seed + 7, attrs=True)

    def host(self):
        """Host copies (tests compare against numpy and the oracle)."""
        dims, m1, m2, attrs = self.star
        hd = [{k: (v.cpu().numpy() if hasattr(v, "cpu") else v) for k, v in d.items()} for d in dims]
        star = (hd, m1.cpu().numpy(), m2.cpu().numpy(), {k: v.cpu().numpy() for k, v in attrs.items()})
        return self.key.cpu().numpy(), self.payload.cpu().numpy(), star

    def topk_window(self, lo, widen=1):
        """[lo, hi) expected to hold about 4 * TOPK_K * widen rows, by the
        key sample; hi > lo (a key that fills many sample slots gets a window
        up to the next distinct sampled key)."""
        s = self.sample
        i = int(np.searchsorted(s, lo, side="left"))
        m = int(math.ceil(4 * TOPK_K * widen * len(s) / self.n)) + 1
        if i + m >= len(s):
            return lo, I64_MAX
        hi = int(s[i + m])
        if hi <= lo:
            j = int(np.searchsorted(s, lo, side="right"))
            hi = int(s[j]) if j < len(s) else I64_MAX
        return lo, hi


class Query:
    __slots__ = ("qid", "kind", "param", "arrival", "start", "done", "result")

    def __init__(self, qid, kind, param, arrival=0.0):
        self.qid, self.kind, self.param, self.arrival = qid, kind, param, arrival
        self.start = self.done = None
        self.result = None


def make_queries(table, count, rate_qps=None, seed=0, mix=(0.5, 0.3, 0.2)):
    """Random queries (kind by `mix`, random constants); arrival times of a
    Poisson process at rate_qps (all at t=0 when None)."""
    r = np.random.default_rng(seed)
    kinds = r.choice(3, size=count, p=mix)
    t = np.cumsum(r.exponential(1.0 / rate_qps, count)) if rate_qps else np.zeros(count)
    s = table.sample
    out = []
    for q in range(count):
        k = int(kinds[q])
        if k == FILTER_AGG:
            param = int(s[r.integers(0, len(s))])  # key < c: a quantile of the keys
        elif k == TOPK:
            param = int(s[r.integers(0, len(s))])  # key >= lo
        else:
            param = (int(r.integers(0, STAR_REGIONS)), int(r.integers(0, STAR_CATEGORIES)))  # region, category
        out.append(Query(q, k, param, float(t[q])))
    return out


# ---------------------------------------------------------------------------
# Units of work (each runs on one worker's stream)
# ---------------------------------------------------------------------------
def _fa_columns(table):
    """COUNT(*), SUM(payload), MIN(key), MAX(key)."""
    return [(L.INT, L.AGG_COUNT_STAR, None), (L.INT, L.AGG_SUM, table.payload), (L.DATETIME, L.AGG_MIN, table.key),
            (L.DATETIME, L.AGG_MAX, table.key)]


def run_filter_agg(table, queries):
    """One shared scan for up to 4 FILTER_AGG queries."""
    dev = table.device
    cols = _fa_columns(table)
    accs = [[torch.empty(32, dtype=torch.uint8, device=dev) for _ in cols] for _ in queries]
    ops.filter_aggregate_multi(table.key, L.LT, [q.param for q in queries], cols, accs, overwrite=True)
    host = torch.stack([torch.cat(a) for a in accs]).cpu().numpy()
    for q, row in zip(queries, host):
        a = [L.Acc.from_buffer_copy(row[32 * k:32 * k + 32].tobytes()) for k in range(len(cols))]
        q.result = (int(a[0].count), int(a[1].i), int(a[2].i) if a[0].count else None,
                    int(a[3].i) if a[0].count else None)
    return table.n


def _range(table, lo, hi, cap):
    """The first min(c, cap) rows in row order with lo <= key < hi (keys,
    payloads), and c."""
    (k, p), c = ops.filter_compact_range(table.key, lo, hi, [table.key, table.payload], cap=cap, retry=False)
    return k, p, c


def _first_sorted(k, p, m):
    """The first m of (k, p) stably sorted by key (non-negative int64 keys:
    their raw bits order like the values)."""
    perm = ops.radix_argsort(k.clone().view(torch.int64))
    take = perm[:m]
    return ops.gather(k, take)[0], ops.gather(p, take)[0]


def run_topk(table, q):
    """ORDER BY key LIMIT 1000 over key >= lo.  The key sample picks a
    window [lo, hi) of about 4,000 rows; a window with too few rows is
    widened 8 times, then to the top of the key range.  A window holding
    more rows than the compaction capacity (a hot key) is not compacted
    whole: count-only scans bisect for the key s at which the count from lo
    first reaches 1,000; every row below s is taken and sorted, then the
    first rows of key s in row order."""
    lo, K = q.param, TOPK_K
    scans = 0
    hi = I64_MAX
    dense = False
    for cap, widen in ((TOPK_CAPS[0], 1), (TOPK_CAPS[1], 8), (TOPK_CAPS[1], None)):
        lo_, hi = table.topk_window(lo, widen) if widen else (lo, I64_MAX)
        k, p, c = _range(table, lo_, hi, cap)
        scans += 1
        if c > cap:
            dense = True
            break
        if c >= K or hi == I64_MAX:
            if c:
                keys, pays = _first_sorted(k, p, K)
                q.result = (keys.cpu().numpy(), pays.cpu().numpy())
            else:
                q.result = (np.zeros(0, np.int64), np.zeros(0, np.int32))
            return table.n * scans
    assert dense
    del k, p
    L_, H = lo, hi  # count(lo, L_) < K <= count(lo, H)
    while H - L_ > 1:
        mid = L_ + (H - L_) // 2
        _, _, cm = _range(table, lo, mid, 1)
        scans += 1
        if cm >= K:
            H = mid
        else:
            L_ = mid
    s_ = L_
    k1, p1, c1 = _range(table, lo, s_, K)  # c1 < K: every row below s
    k2, p2, c2 = _range(table, s_, s_ + 1, K - c1)  # the first rows of key s
    scans += 2
    if c1:
        k1, p1 = _first_sorted(k1, p1, c1)
    q.result = (torch.cat([k1, k2]).cpu().numpy(), torch.cat([p1, p2]).cpu().numpy())
    return table.n * scans


def join_group_dims(table, params):
    """The star join's dimension list for queries with these (region,
    category) constants: one query's 0/1 filters, or for several queries
    filter bytes whose bit q is query q's."""
    dims, _, _, attrs = table.star
    d = [dict(x) for x in dims]
    if len(params) == 1:
        d[2]["pass"] = (attrs["region"] == params[0][0]).to(torch.uint8)
        d[3]["pass"] = (attrs["cat"] == params[0][1]).to(torch.uint8)
        return d
    dev = table.device
    weights = torch.tensor([1 << q for q in range(len(params))], dtype=torch.uint8, device=dev)
    for j, (name, k) in ((2, ("region", 0)), (3, ("cat", 1))):
        consts = torch.tensor([p[k] for p in params], dtype=torch.int32, device=dev)
        hit = attrs[name].unsqueeze(1) == consts.unsqueeze(0)
        d[j]["pass"] = (hit.to(torch.uint8) * weights).sum(dim=1, dtype=torch.uint8)
    return d


def run_join_group(table, queries):
    """Up to SHARED_STAR_MAX star-join queries in one read of the facts."""
    _, m1, m2, _ = table.star
    dims = join_group_dims(table, [q.param for q in queries])
    if len(queries) == 1:
        sums, counts = ops.star_join_group_sum(dims, m1, m2, STAR_GROUPS)
        sums, counts = sums.view(1, -1), counts.view(1, -1)
    else:
        sums, counts = ops.star_join_group_sum_multi(dims, m1, m2, STAR_GROUPS, len(queries))
    hs, hc = sums.cpu().numpy(), counts.cpu().numpy()
    for q, a, b in zip(queries, hs, hc):
        q.result = (a, b)
    return table.star_rows


def _blk(nbytes):
    """A device allocation as the caching allocator accounts it (512-byte blocks)."""
    return max(512, (int(nbytes) + 511) // 512 * 512)


def unit_bytes(table, kind, nq=1):
    """Device memory a unit allocates, reserved from the budget at admission:
    the sum of its operators' buffers and workspaces."""
    n = table.n
    if kind == FILTER_AGG:
        ws = L.ws_bytes("ysql_filter_aggregate_multi_ws_bytes", n, nq, 4)
        # constants, workspace, 4 accumulators per query, the per-query
        # concatenation and the stacked copy read back to the host
        return _blk(8 * nq) + _blk(max(ws, 256)) + nq * 4 * _blk(32) + nq * _blk(128) + _blk(128 * nq)
    if kind == TOPK:
        cws = _blk(max(L.ws_bytes("ysql_compact_ws_bytes", n), 256))

        def compaction(cap):  # bounds, two output columns, count, workspace
            return _blk(16) + _blk(8 * cap) + _blk(4 * cap) + _blk(8) + cws

        def first_sorted(c, m):  # key copy, permutation, sort workspace, two gathers
            rws = L.ws_bytes("ysql_radix_sort_ws_bytes", L.UKEY64, c)
            return _blk(8 * c) + _blk(4 * c) + _blk(max(rws, 256)) + _blk(8 * m) + _blk(m) + _blk(4 * m) + _blk(m)

        K = TOPK_K
        # two window widths alive at once (the wider window's rows sorted),
        # or the hot-key path: count-only scans, then rows below s and of s,
        # the sort of the former and the concatenation of the result
        windows = 2 * compaction(TOPK_CAPS[1]) + first_sorted(TOPK_CAPS[1], K)
        hot = compaction(1) + 2 * compaction(K) + first_sorted(K, K) + _blk(8 * K) + _blk(4 * K)
        return max(windows, hot)
    # JOIN_GROUP: the two dimension filters, sums and counts.  One query:
    # a bool then a uint8 per filtered dimension row; nq queries: the bit
    # weights, and per dimension the constants, the row x query comparison,
    # its uint8 copy, the weighted copy and the summed filter byte.
    rows = (30000, 200000)
    if nq == 1:
        filters = sum(2 * _blk(r) for r in rows)
    else:
        filters = _blk(nq) + sum(_blk(4 * nq) + 3 * _blk(r * nq) + _blk(r) for r in rows)
    return filters + 2 * _blk(8 * STAR_GROUPS * nq)


# ---------------------------------------------------------------------------
# Dispatcher and workers
# ---------------------------------------------------------------------------
class Server:
    """Dispatcher (the calling thread) and a pool of `slots` workers, each
    with its own stream.  Call close() when done."""

    def __init__(self, table, slots=8, hbm_budget=200 << 30, shared_scan=True, runners=None):
        import queue

        self.table = table
        # unit runners by query kind: fn(table, queries) -> table rows read
        self.runners = runners or {FILTER_AGG: run_filter_agg, TOPK: lambda t, qs: run_topk(t, qs[0]),
                                   JOIN_GROUP: run_join_group}
        self.error = None  # the first exception a unit raised; serve() re-raises it
        self.slots = slots
        self.budget = hbm_budget
        self.shared_scan = shared_scan
        self.cv = threading.Condition()
        self.free_bytes = hbm_budget
        self.idle = list(range(slots))
        self.outstanding = 0
        self.rows_queried = 0  # table rows covered, summed over queries
        self.rows_scanned = 0  # table rows read (a shared scan reads once)
        self.inboxes = [queue.Queue() for _ in range(slots)]
        self.threads = [threading.Thread(target=self._worker, args=(k,), daemon=True) for k in range(slots)]
        for th in self.threads:
            th.start()

    def close(self):
        for ib in self.inboxes:
            ib.put(None)
        for th in self.threads:
            th.join()

    def _worker(self, k):
        dev = self.table.device
        if dev.type == "cuda":
            torch.cuda.set_device(dev)
            stream = torch.cuda.Stream(device=dev)
            ctx = torch.cuda.stream(stream)
        else:  # host-only runners (tests of the dispatcher itself)
            stream, ctx = None, contextlib.nullcontext()
        with ctx:
            while True:
                item = self.inboxes[k].get()
                if item is None:
                    return
                kind, qs, nbytes = item
                t0 = time.perf_counter()
                for q in qs:
                    q.start = t0
                scanned = 0
                try:
                    scanned = self.runners[kind](self.table, qs)
                    if stream is not None:
                        stream.synchronize()
                except BaseException as e:  # kept for serve(); the slot and bytes are still returned
                    with self.cv:
                        if self.error is None:
                            self.error = e
                finally:
                    t1 = time.perf_counter()
                    for q in qs:
                        q.done = t1
                    with self.cv:
                        self.free_bytes += nbytes
                        self.outstanding -= 1
                        self.rows_scanned += scanned
                        self.rows_queried += self.table.n * len(qs)
                        self.idle.append(k)
                        self.cv.notify_all()

    def _next_unit(self, pending):
        head = pending[0]
        if head.kind == FILTER_AGG and self.shared_scan:
            return head.kind, [q for q in pending if q.kind == FILTER_AGG][:SHARED_SCAN_MAX]
        if head.kind == JOIN_GROUP and self.shared_scan:
            return head.kind, [q for q in pending if q.kind == JOIN_GROUP][:SHARED_STAR_MAX]
        return head.kind, [head]

    def serve(self, queries, realtime=True):
        """Runs the queries; arrival times are seconds after the start when
        `realtime`, else everything arrives at once.  Sets each query's
        arrival/start/done (seconds after the start) and result; returns the
        wall time."""
        arrivals = collections.deque(sorted(queries, key=lambda q: q.arrival))
        pending = collections.deque()
        self.rows_queried = self.rows_scanned = 0
        t0 = time.perf_counter()
        while True:
            now = time.perf_counter() - t0
            while arrivals and (not realtime or arrivals[0].arrival <= now):
                q = arrivals.popleft()
                if not realtime:
                    q.arrival = now
                pending.append(q)
            with self.cv:
                if self.error is not None:
                    # a unit failed: admit nothing more, let the units in
                    # flight finish, then raise the failure
                    while self.outstanding:
                        self.cv.wait()
                    err, self.error = self.error, None
                    raise RuntimeError("a serving unit failed") from err
                while pending and self.idle:
                    kind, group = self._next_unit(pending)
                    nbytes = unit_bytes(self.table, kind, len(group))
                    if nbytes > self.free_bytes and self.outstanding:
                        break  # admission waits for memory; arrival order is kept
                    self.free_bytes -= nbytes
                    self.outstanding += 1
                    k = self.idle.pop()
                    for q in group:
                        pending.remove(q)
                    self.inboxes[k].put((kind, group, nbytes))
                if not arrivals and not pending and not self.outstanding:
                    break
                wait = 0.002
                if arrivals:
                    wait = max(0.0, min(wait, arrivals[0].arrival - (time.perf_counter() - t0)))
                if wait > 0:
                    self.cv.wait(timeout=wait)
        wall = time.perf_counter() - t0
        for q in queries:
            q.start -= t0
            q.done -= t0
        return wall

    def run_alone(self, queries):
        """Every query by itself, one after another (no shared scans, nothing
        else in flight)."""
        saved = self.shared_scan
        self.shared_scan = False
        try:
            for q in queries:
                self.serve([q], realtime=False)
        finally:
            self.shared_scan = saved


def latency_report(queries, wall, server):
    lat = np.array([q.done - q.arrival for q in queries]) * 1e3
    rep = {"queries": len(queries), "wall_s": round(wall, 3), "qps": round(len(queries) / wall, 1),
           "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
           "rows_per_s": round(server.rows_queried / wall, 1),
           "scanned_rows_per_s": round(server.rows_scanned / wall, 1)}
    for k, name in KIND_NAMES.items():
        lk = np.array([q.done - q.arrival for q in queries if q.kind == k]) * 1e3
        if lk.size:
            rep[name] = {"n": int(lk.size), "p50_ms": round(float(np.percentile(lk, 50)), 3),
                         "p99_ms": round(float(np.percentile(lk, 99)), 3)}
    return rep
