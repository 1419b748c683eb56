"""Serving loop: concurrent queries against HBM-resident tables.

No reference counterpart (the reference runs one query at a time from its
CLI).  Queries arrive over time (a Poisson process in the benchmark); the
dispatcher admits them as execution slots and device memory free up,
batches operators across the queries in flight, and keeps every
intermediate on the device; only final results come back to the host.

Query kinds over the table T(key int64, payload int32 = row id) and a
dimension D(dkey int64, grp int32):
  FILTER_AGG  SELECT COUNT(*), SUM(payload), MIN(key), MAX(key)
              FROM T WHERE key < c
  TOPK        SELECT key, payload FROM T WHERE key >= lo
              ORDER BY key LIMIT 1000        (ties in row order)
  JOIN_GROUP  SELECT grp, SUM(payload), COUNT(*) FROM T JOIN D ON key = dkey
              WHERE payload < c GROUP BY grp

Operator batching across queries: up to 4 FILTER_AGG queries in flight
share one scan of the table (ysql_filter_aggregate_multi, one read of the
columns for all of them); TOPK and JOIN_GROUP pipelines run concurrently on
their own streams.  A query's result does not depend on what it was batched
with: the shared scan folds each query in the same order as a lone scan,
and every other operator runs per query.

Admission: each unit of work (a shared scan, or one TOPK / JOIN_GROUP
pipeline) reserves an estimate of its device memory from the HBM budget and
waits (in arrival order) while the budget or every slot is taken.
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
from .datagen import multirun_keys

FILTER_AGG, TOPK, JOIN_GROUP = 0, 1, 2
KIND_NAMES = {FILTER_AGG: "filter_agg", TOPK: "topk", JOIN_GROUP: "join_group"}
TOPK_K = 1000
SHARED_SCAN_MAX = 4
I64_MAX = 2**63 - 1


class Table:
    """T(key, payload) resident in HBM, and the join dimension D."""

    def __init__(self, n, device, seed=1, dim_rows=2000, groups=64, sample=1 << 20):
        self.n, self.device, self.groups = n, device, groups
        self.key = multirun_keys(n, device, seed=seed)
        self.payload = torch.arange(n, dtype=torch.int32, device=device)
        g = torch.Generator(device=device)
        g.manual_seed(seed + 100)
        # sorted key sample: statistics for choosing TOPK windows
        idx = torch.randint(0, n, (min(sample, n),), generator=g, device=device)
        self.sample = np.sort(self.key[idx].cpu().numpy())
        # D: distinct ranks of the Zipf half away from the head (each matches
        # a few hundred rows or fewer), random groups
        r = np.random.default_rng(seed)
        dk = r.choice(np.arange(1000, 10_000_000, dtype=np.int64), size=dim_rows, replace=False)
        self.dkey = torch.from_numpy(dk).to(device)
        self.dgrp = torch.from_numpy(r.integers(0, groups, dim_rows).astype(np.int32)).to(device)

    def host(self):
        """Host copies (tests compare against numpy)."""
        return (self.key.cpu().numpy(), self.payload.cpu().numpy(), self.dkey.cpu().numpy(),
                self.dgrp.cpu().numpy())

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
            param = int(r.integers(0, table.n + 1))  # payload < c
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


def run_topk(table, q):
    lo = q.param
    widen, scans = 1, 0
    while True:
        scans += 1
        lo_, hi = table.topk_window(lo, widen)
        (k, p), c = ops.filter_compact_range(table.key, lo_, hi, [table.key, table.payload], cap=16 * TOPK_K * widen)
        if c >= TOPK_K or hi == I64_MAX:
            break
        widen *= 8
    if c:
        kk = k.clone()  # non-negative int64: the raw bits order like the values
        perm = ops.radix_argsort(kk.view(torch.int64))
        take = perm[:TOPK_K]
        keys, _ = ops.gather(k, take)
        pays, _ = ops.gather(p, take)
        q.result = (keys.cpu().numpy(), pays.cpu().numpy())
    else:
        q.result = (np.zeros(0, np.int64), np.zeros(0, np.int32))
    return table.n * scans


def run_join_group(table, q):
    dev = table.device
    pairs, c = ops.join_pairs(L.EQ, table.key, table.dkey, cap=1 << 16)
    G = table.groups
    sums = ops.acc_new(L.INT, L.AGG_SUM, nacc=G, device=dev)
    cnts = ops.acc_new(L.INT, L.AGG_COUNT, nacc=G, device=dev)
    if c:
        pay, _ = ops.gather(table.payload, pairs[:, 0].contiguous())
        grp, _ = ops.gather(table.dgrp, pairs[:, 1].contiguous())
        cst = torch.tensor([q.param], dtype=torch.int32, device=dev)
        skip = ops.compare(L.GE, pay, cst, rhs_scalar=True)  # payload >= c: not selected
        ops.group_reduce(L.INT, L.AGG_SUM, grp, G, pay, sums, null=skip)
        ops.group_reduce(L.INT, L.AGG_COUNT, grp, G, pay, cnts, null=skip)
    s = ops.acc_read(sums)
    n = ops.acc_read(cnts)
    q.result = (np.array([a["i"] for a in s], np.int64), np.array([a["count"] for a in n], np.int64))
    return table.n


def unit_bytes(table, kind, nq=1):
    """Device memory a unit reserves from the budget (an upper estimate)."""
    if kind == FILTER_AGG:
        return (table.n // 32768 + 2) * 32 * 4 * nq + (1 << 20)
    if kind == TOPK:
        return 16 * TOPK_K * 8 * 12 * 3 + (8 << 20)
    return (1 << 16) * 8 * 3 + (16 << 20)


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
                                   JOIN_GROUP: lambda t, qs: run_join_group(t, qs[0])}
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
