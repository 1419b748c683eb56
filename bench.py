"""Benchmark: out-of-core multi-run sort on one MI355X.

Workload (one step): 200,000,000 rows; key int64 drawn 50% uniform over
[0, 2^63) and 50% Zipf(s=1.1) over 10^7 distinct values; payload = uint32
row index.  The table is cut into 256 MB key batches of 2^25 rows (6 runs:
5 x 33,554,432 + 32,227,840); each run is radix-sorted with 8-bit digits
(8 passes over the 63 significant bits), then the runs are merged in one
k-way merge (ysql_merge_runs: every row moves once; the reference's
OrderOperator merges pairwise, level by level, order_operator.hpp:160-193,
which bench.py --pairwise-merge measures).  Inputs are resident in HBM
before timing.

Prints one JSON line (rows/s of the whole job) with the roofline of the
dominant kernel (the onesweep digit pass), measured with HIP events around
each pass launch inside the timed steps, and a CPU baseline of the same
pipeline (the oracle's multi-run sort) on a bounded sample on the host.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N_ROWS = 200_000_000
RUN_ROWS = 1 << 25
ZIPF_N = 10_000_000
ZIPF_S = 1.1
SEED = 1
HBM_PEAK_GBS = 8000.0
KEY_BITS = 63


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--rows", type=int, default=N_ROWS)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-verify", action="store_true")
    ap.add_argument("--calib-copy", action="store_true",
                    help="run one plain 1.6 GB device copy first (calibrates HBM byte counters in a profile)")
    ap.add_argument("--no-ops", action="store_true", help="skip the per-op roofline lines (filter, filter+aggregate)")
    ap.add_argument("--no-batched", action="store_true", help="skip the 4,096-job batched vs plain comparison")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving loop")
    ap.add_argument("--no-configs", action="store_true", help="skip the config-1 and config-3 lines")
    ap.add_argument("--serving-queries", type=int, default=10_000, help="queries per rate, at most")
    ap.add_argument("--serving-seconds", type=float, default=60.0, help="arrival window per rate, at most")
    ap.add_argument("--pairwise-merge", action="store_true",
                    help="merge the runs pairwise, level by level, instead of one k-way merge")
    return ap.parse_args()


def gen_keys(n, dev):
    """Config-2 keys on the GPU from a seeded torch generator."""
    from yallasql_amd.datagen import multirun_keys

    return multirun_keys(n, dev, seed=SEED)


CPU_SAMPLE_ROWS = 50_000_000  # rows of the 200M-row table the CPU op baselines run on
CPU_JOIN_ROWS = 20_000_000
CPU_STAR_ROWS = 60_000_000


def _cpu_threads():
    """(std::thread::hardware_concurrency(), threads used): all the host's
    hardware threads, capped by OMP_NUM_THREADS where it is set (the CPU share
    a GPU gets on a shared box)."""
    from oracle import pyoracle as O

    hc = O.cpb_hardware_concurrency()
    cap = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or hc
    return hc, max(1, min(hc, cap))


def _cpu_model():
    import subprocess

    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except Exception:
        pass
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def _cpu_median(fn):
    """2 untimed warm-up runs, then the median of 5 timed runs (seconds)."""
    for _ in range(2):
        fn()
    ts = []
    for _ in range(5):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    ts.sort()
    return ts[2]


def cpu_baseline(inp):
    """The oracle's multi-threaded CPU versions (oracle/cpu_baseline.cpp,
    -O3 -march=native) on bounded samples of the same inputs, at 1 thread
    and at all permitted threads; 2 warm-ups + median of 5 each."""
    from oracle import pyoracle as O

    hc, T = _cpu_threads()
    ops = {}

    def both(name, rows, fn, sample, rows_1=None, fn_1=None):
        t1 = _cpu_median(lambda: (fn_1 or fn)(1))
        tT = _cpu_median(lambda: fn(T))
        ops[name] = {"rows_per_s_1_thread": (rows_1 or rows) / t1, "rows_per_s": rows / tT, "threads": T,
                     "ms_1_thread": t1 * 1e3, "ms": tT * 1e3, "sample": sample}

    k2 = inp["config2_keys"]  # 2 runs of RUN_ROWS
    k1 = k2[:RUN_ROWS // 2]  # 2 runs of RUN_ROWS / 4 for the 1-thread figure
    both("config2_multirun_sort", k2.size, lambda t: O.cpb_multirun_sort(k2, RUN_ROWS, t),
         f"all threads: first {k2.size:,} keys of the config-2 stream as 2 runs of {RUN_ROWS:,} + 1 merge; "
         f"1 thread: first {k1.size:,} keys as 2 runs of {RUN_ROWS // 4:,} + 1 merge",
         rows_1=k1.size, fn_1=lambda t: O.cpb_multirun_sort(k1, RUN_ROWS // 4, t))
    if "vals" in inp:
        v, k = inp["vals"], inp["keys"]
        both("filter_aggregate", v.size, lambda t: O.cpb_filter_aggregate(v, k, 0.5, t),
             f"first {v.size:,} rows of the op table: WHERE f32 < 0.5 -> COUNT, SUM, MIN/MAX(int64)")
        both("filter_compact", v.size, lambda t: O.cpb_filter_compact(v, k, 0.5, t),
             f"first {v.size:,} rows: WHERE f32 < 0.5 compacting (int64 key, f32)")
        for R, right in inp["join_right"].items():
            left = inp["join_left"]
            both(f"join_eq_R{R}", left.size, lambda t: O.cpb_join_eq(left, right, t, 1 << 24),
                 f"first {left.size:,} left int64 keys x {R:,} right keys: hash build + parallel probe + pairs")
    if "config1" in inp:
        ids, price = inp["config1"]
        both("config1_filter_sort_project", ids.size, lambda t: O.cpb_config1(ids, price, t),
             f"all {ids.size:,} rows: WHERE price > 500, composite-key radix sort (price DESC, id), 2*id")
    if "star" in inp:
        dims, m1, m2 = inp["star"]
        both("config3_star_join_group_by", m1.size, lambda t: O.cpb_star_join_group_sum(dims, m1, m2, 7000, t),
             f"first {m1.size:,} of the 600M fact rows: direct-index join, per-thread private group arrays")
    head = ops["config2_multirun_sort"]
    return {
        "value": head["rows_per_s"],
        "unit": "rows/s",
        "cores": T,
        "kind": "port",
        "sample": head["sample"],
        "value_1_thread": head["rows_per_s_1_thread"],
        "hardware_concurrency": hc,
        "cpu_model": _cpu_model(),
        "runs": "2 warm-up + median of 5",
        "ops": ops,
    }


def _timed_ms(fn, reps=10):
    """Median of `reps` runs of fn() on the current stream, HIP events."""
    import torch

    ts = []
    for _ in range(reps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        fn()
        b.record()
        b.synchronize()
        ts.append(a.elapsed_time(b))
    ts.sort()
    return ts[len(ts) // 2]


def op_lines(keys, cpu_inputs=None):
    """Filter (compaction) and fused filter+aggregate over the 200M-row table:
    rows/s and fraction of HBM peak by their algorithmic bytes.  With
    `cpu_inputs`, host copies of the samples the CPU baseline runs on."""
    import torch

    from yallasql_amd import _lib as L
    from yallasql_amd import ops

    n = keys.numel()
    dev = keys.device
    g = torch.Generator(device=dev)
    g.manual_seed(SEED + 7)
    vals = torch.rand(n, generator=g, device=dev)
    out = {}
    if cpu_inputs is not None:
        S = min(n, CPU_SAMPLE_ROWS)
        cpu_inputs["vals"] = vals[:S].cpu().numpy()
        cpu_inputs["keys"] = keys[:S].cpu().numpy()
        cpu_inputs["join_left"] = keys[:min(n, CPU_JOIN_ROWS)].cpu().numpy()
        cpu_inputs["join_right"] = {}
    # WHERE vals < 0.5: COUNT, SUM(vals) (double), MIN(key), MAX(key) -- reads 12 B/row
    cols = [(L.INT, L.AGG_COUNT_STAR, None), (L.FLOAT, L.AGG_SUM, vals), (L.DATETIME, L.AGG_MIN, keys),
            (L.DATETIME, L.AGG_MAX, keys)]
    accs = [[torch.empty(32, dtype=torch.uint8, device=dev) for _ in cols]]
    run = lambda: ops.filter_aggregate_multi(vals, L.LT, [0.5], cols, accs)  # noqa: E731
    run()
    ms = _timed_ms(run)
    b = n * 12
    out["filter_aggregate"] = {"rows_per_s": n / (ms * 1e-3), "ms": ms, "algorithmic_GB": b / 1e9,
                               "GB_s": b / (ms * 1e-3) / 1e9, "frac": b / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                               "what": "WHERE f32 < c -> COUNT, SUM(f32), MIN/MAX(int64) over 200M rows"}
    # four queries sharing one scan (the serving loop's batching)
    accs4 = [[torch.empty(32, dtype=torch.uint8, device=dev) for _ in cols] for _ in range(4)]
    run4 = lambda: ops.filter_aggregate_multi(vals, L.LT, [0.2, 0.4, 0.6, 0.8], cols, accs4)  # noqa: E731
    run4()
    ms4 = _timed_ms(run4)
    out["filter_aggregate_shared_x4"] = {"queries_rows_per_s": 4 * n / (ms4 * 1e-3), "ms": ms4,
                                         "GB_s": b / (ms4 * 1e-3) / 1e9, "frac": b / (ms4 * 1e-3) / 1e9 / HBM_PEAK_GBS}
    # WHERE vals < 0.5 compacting (key, vals): N*(4+8) read + M*(8+4) written
    cnt = torch.zeros(1, dtype=torch.int64, device=dev)
    kout, vout = torch.empty_like(keys), torch.empty_like(vals)
    col = (L.Col * 2)()
    col[0] = L.Col(keys.data_ptr(), None, kout.data_ptr(), None, 8, 0)
    col[1] = L.Col(vals.data_ptr(), None, vout.data_ptr(), None, 4, 0)
    sc = torch.tensor([0.5], device=dev)
    ws = ops.workspace(L.ws_bytes("ysql_compact_ws_bytes", n), dev)
    runf = lambda: L.call("ysql_filter_compact", L.FLOAT, L.LT, ops._p(vals), ops._p(sc), col, 2, n,  # noqa: E731
                          ops._p(cnt), ops._p(ws), ws.numel(), ops._s())
    runf()
    msf = _timed_ms(runf)
    m = int(cnt.item())
    bf = n * 12 + m * 12
    out["filter_compact"] = {"rows_per_s": n / (msf * 1e-3), "ms": msf, "selected": m,
                             "GB_s": bf / (msf * 1e-3) / 1e9, "frac": bf / (msf * 1e-3) / 1e9 / HBM_PEAK_GBS,
                             "what": "WHERE f32 < 0.5 compacting (int64 key, f32) of 200M rows"}
    del vals, kout, vout, ws
    # inner EQ join of the 200M int64 keys with a dimension of R int64 keys in
    # [1000, 1e7): hash build of the right side + probe, (l, r) pairs written
    for R in (2_000, 500_000):
        right = torch.randint(1000, 10_000_000, (R,), generator=g, device=dev, dtype=torch.int64)
        if cpu_inputs is not None:
            cpu_inputs["join_right"][R] = right.cpu().numpy()
        jws = ops.workspace(L.ws_bytes("ysql_join_ws_bytes", L.DATETIME, L.EQ, n, R), dev)
        jcap = 1 << 24
        pairs = torch.empty((jcap, 2), dtype=torch.int32, device=dev)
        jcnt = torch.zeros(1, dtype=torch.int64, device=dev)
        runj = lambda: L.call("ysql_join_pairs", L.DATETIME, L.EQ, ops._p(keys), n, ops._p(right), R,  # noqa: E731
                              ops._p(pairs), jcap, ops._p(jcnt), ops._p(jws), jws.numel(), ops._s())
        runj()
        msj = _timed_ms(runj)
        npairs = int(jcnt.item())
        assert npairs <= jcap
        bj = n * 8 + R * 8 + npairs * 8
        out[f"join_eq_R{R}"] = {"rows_per_s": n / (msj * 1e-3), "ms": msj, "pairs": npairs,
                                "GB_s": bj / (msj * 1e-3) / 1e9, "frac": bj / (msj * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                "what": f"inner EQ join, 200M int64 left keys x {R:,} right keys: build + "
                                        "filtered probe + pairs"}
        del right, jws, pairs
    return out


def config1_line(cpu_inputs=None):
    """Config 1: WHERE price > 500 ORDER BY price DESC, id ASC; SELECT id, price,
    2*id over 1,048,576 rows in one batch, through the engine (result copied
    to the host), median wall time of 10 queries."""
    import numpy as np

    from yallasql_amd import engine as E

    n = 1 << 20
    r = np.random.default_rng(0)
    ids = r.permutation(n).astype(np.int32)
    price = np.minimum((1 + 999 * r.random(n)).astype(np.float32), np.float32(999.9999))
    cat = r.integers(0, 64, n).astype(np.int32)
    ctx = E.Context()
    t = ctx.table("t", [("id", ids), ("price", price), ("category", cat)])
    if cpu_inputs is not None:
        cpu_inputs["config1"] = (ids, price)

    def query():
        p = ctx.scan(t, n).filter(E.cmp(2, E.ref(1), E.f32(500.0))).order([1, 0], [1, 0])
        return p.project([E.ref(0), E.ref(1), E.arith(2, E.i32(2), E.ref(0))]).execute()

    cols, _, _ = query()
    ts = []
    for _ in range(10):
        t0 = time.perf_counter()
        query()
        ts.append(time.perf_counter() - t0)
    ts.sort()
    ms = ts[len(ts) // 2] * 1e3
    m = int(cols[0].shape[0])
    return {"rows": n, "selected": m, "ms": ms, "rows_per_s": n / (ms * 1e-3),
            "what": "engine scan -> filter -> order (price desc, id) -> projection, result on the host"}


def csv_load_line(n=2_000_000):
    """CSV file -> device column batches (the scan path's feeder): a
    config-1-shaped table (id INT, price FLOAT, qty INT) written to a temporary
    file, then read batch by batch through pinned staging with each batch's
    copies overlapping the next batch's parse."""
    import os
    import tempfile

    import numpy as np
    import pandas as pd
    import torch

    from yallasql_amd import loader

    r = np.random.default_rng(SEED)
    df = pd.DataFrame({"id": np.arange(n, dtype=np.int32),
                       "price": np.round(1 + 999 * r.random(n), 2).astype(np.float32),
                       "qty": r.integers(0, 100, n, dtype=np.int32)})
    fd, path = tempfile.mkstemp(suffix=".csv")
    os.close(fd)
    try:
        df.to_csv(path, index=False)
        size = os.path.getsize(path)
        names, types = ["id", "price", "qty"], [0, 1, 0]
        dev = torch.device("cuda:0")

        def load(device_path):
            rd = loader.CsvReader(path, names, types)
            s = torch.cuda.Stream(device=dev)
            rows = 0
            t0 = time.perf_counter()
            while True:
                b = rd.next_device(dev, stream=s) if device_path else rd.next_host()
                if b is None:
                    break
                rows += int(b[0][0].shape[0])
            s.synchronize()
            dt = time.perf_counter() - t0
            rd.close()
            return rows, dt

        load(True)
        rows_h, t_host = load(False)
        rows_d, t_dev = load(True)
        assert rows_h == rows_d == n
        return {"rows": n, "csv_MB": size / 1e6, "s": t_dev, "rows_per_s": n / t_dev, "MB_s": size / t_dev / 1e6,
                "parse_only_s": t_host,
                "what": "CSV (id INT, price FLOAT, qty INT) -> device batches of 349,440 rows: file mapped, "
                        "record starts found in one SSE2 pass, up to 16 host threads converting row ranges, pinned "
                        "double-buffered staging, H2D copies on a side stream"}
    finally:
        os.unlink(path)


def config3_line(cpu_inputs=None):
    """Config 3: star join of a 600M-row fact table with four dimensions
    (2,556 / 2,000 / 30,000 / 200,000 rows; region = 0 and category = 0
    filters), SUM(extendedprice * discount) GROUP BY year x brand (7,000
    groups), one fused kernel; HBM roofline by the fact columns read."""
    import torch

    from yallasql_amd import ops
    from yallasql_amd.datagen import star_schema

    dev = torch.device("cuda:0")
    n = 600_000_000
    dims, m1, m2 = star_schema(n, dev, seed=2)
    sums = torch.zeros(7000, dtype=torch.int64, device=dev)
    counts = torch.zeros(7000, dtype=torch.int64, device=dev)

    def run():
        sums.zero_()
        counts.zero_()
        ops.star_join_group_sum(dims, m1, m2, 7000, sums=sums, counts=counts)

    run()
    ms = _timed_ms(run)
    if cpu_inputs is not None:
        S = min(n, CPU_STAR_ROWS)
        hd = []
        for d in dims:
            h = {k: (v.cpu().numpy() if hasattr(v, "cpu") else v) for k, v in d.items()}
            h["fk"] = h["fk"][:S]
            hd.append(h)
        cpu_inputs["star"] = (hd, m1[:S].cpu().numpy(), m2[:S].cpu().numpy())
    b = n * 24  # four int32 foreign keys + two int32 measures per fact row
    out = {"fact_rows": n, "ms": ms, "rows_per_s": n / (ms * 1e-3), "qualifying_rows": int(counts.sum().item()),
           "GB_s": b / (ms * 1e-3) / 1e9, "frac": b / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
           "what": "fused 4-FK star join + GROUP BY year x brand SUM(int64(price)*discount)"}
    del dims, m1, m2
    torch.cuda.empty_cache()
    return out


def batched_lines():
    """4,096 ragged sort/filter/aggregate jobs (log-uniform 1,024..4,194,304
    rows): the per-job launch loop against the persistent launch (one kernel,
    device work queue, captured with its reset in a hipGraph) and against the
    launch-chain graph.  Outputs compared on the device, buffer by buffer."""
    import torch

    from yallasql_amd.autotune import load_sort_variants
    from yallasql_amd.batched import JobSet, config4_jobs

    dev = torch.device("cuda:0")
    sizes, kinds, key64, thr = config4_jobs()
    js = JobSet(sizes, kinds, key64, thr, dev, seed=4).create(chains=32, sort_variants=load_sort_variants())
    js.run_batched()
    js.run_graph()
    js.run_plain()
    torch.cuda.synchronize()

    def wall(fn, reps=3):
        best = None
        for _ in range(reps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            best = dt if best is None else min(best, dt)
        return best

    t_plain = wall(js.run_plain)
    t_graph = wall(js.run_graph)
    t_pers = wall(js.run_batched)
    outs = (js.o32, js.o64, js.ov, js.perm, js.counts, js.accs)
    js.clear_outputs()
    js.run_plain()
    torch.cuda.synchronize()
    ref = [t.clone() for t in outs]
    js.clear_outputs()
    js.run_batched()
    done, total = js.check()
    same = done == total and all(torch.equal(a, b) for a, b in zip(ref, outs))
    js.clear_outputs()
    js.run_graph()
    torch.cuda.synchronize()
    same_graph = all(torch.equal(a, b) for a, b in zip(ref, outs))
    del ref
    alg = js.algorithmic_bytes()
    out = {"jobs": len(sizes), "rows": js.total_rows, "plain_s": t_plain, "batched_s": t_pers,
           "speedup": t_plain / t_pers, "identical_outputs": bool(same),
           "batched_GB_s": alg / t_pers / 1e9, "batched_frac": alg / t_pers / 1e9 / HBM_PEAK_GBS,
           "workers": js.workers(), "jobs_completed": done,
           "graph_s": t_graph, "graph_speedup": t_plain / t_graph, "graph_identical_outputs": bool(same_graph),
           "what": "plain: per-job launches on one stream; batched: ONE persistent launch (workgroups take whole "
                   "jobs and split-sort chunk tasks from device job lists and a phase queue), replayed as a "
                   "hipGraph with its reset kernel; graph: launch chains beside segmented kernels"}
    js.close()
    del js
    torch.cuda.empty_cache()
    return out


def serving_lines(max_queries, max_seconds):
    """Serving loop over the 200M-row table and the 600M-row star schema at
    50/200/800 queries/s: each rate runs min(max_queries, rate *
    max_seconds) Poisson arrivals."""
    import torch

    from yallasql_amd.serving import Server, Table, latency_report, make_queries

    dev = torch.device("cuda:0")
    table = Table(N_ROWS, dev, seed=SEED, star_rows=600_000_000)
    srv = Server(table, slots=8, hbm_budget=200 << 30)
    out = {}
    try:
        srv.serve(make_queries(table, 24, rate_qps=None, seed=1), realtime=False)  # warm-up
        for rate in (50, 200, 800):
            qs = make_queries(table, int(min(max_queries, rate * max_seconds)), rate_qps=rate, seed=4 + rate)
            wall = srv.serve(qs)
            out[f"{rate}qps"] = latency_report(qs, wall, srv)
            print(f"serving {rate} qps: {len(qs)} queries done", file=sys.stderr, flush=True)
    finally:
        srv.close()
    out["what"] = ("Poisson arrivals; 50% WHERE key < c -> COUNT/SUM/MIN/MAX over 200M rows (up to 4 share a scan), "
                   "30% ORDER BY key LIMIT 1000 above a random key, 20% the star join + GROUP BY year, brand over "
                   "600M fact rows with random region/category constants; 8 slots, 200 GB HBM budget reserved by "
                   "each unit's own allocation sizes; latency = arrival to result on the host")
    del table
    torch.cuda.empty_cache()
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU figures are not produced",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import torch

    dev = torch.device("cuda:0")
    keys = gen_keys(args.rows, dev)
    if args.calib_copy:
        torch.empty_like(keys).copy_(keys)
    torch.cuda.synchronize()
    from yallasql_amd.multirun import MultiRunSort, verify_global_sort

    job = MultiRunSort(keys, RUN_ROWS, key_bits=KEY_BITS, merge="pairwise" if args.pairwise_merge else "kway")
    for _ in range(args.warmup):
        job.step()
    torch.cuda.synchronize()
    job.record = True
    t0 = time.perf_counter()
    for _ in range(args.steps):
        job.step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    job.record = False
    ms = (t1 - t0) * 1e3 / args.steps
    rows_per_s = args.rows * args.steps / (t1 - t0)

    pass_ms = [a.elapsed_time(b) for a, b, _ in job.pass_events]
    pass_bytes = [2 * r * 12 for _, _, r in job.pass_events]
    pass_gbs = sum(pass_bytes) / (sum(pass_ms) * 1e-3) / 1e9
    merge_ms = [a.elapsed_time(b) for a, b, _ in job.merge_events]
    merge_bytes = [2 * r * 12 for _, _, r in job.merge_events]
    merge_gbs = sum(merge_bytes) / (sum(merge_ms) * 1e-3) / 1e9 if merge_ms else None
    alg = job.algorithmic_bytes()

    if not args.no_verify:
        ko, vo = job.result()
        verify_global_sort(keys, ko, vo)

    # attainable HBM bandwidth: a plain 1.6 GB device copy (read + write)
    dst = torch.empty_like(keys)
    copy_ms = []
    for _ in range(3):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        dst.copy_(keys)
        b.record()
        torch.cuda.synchronize()
        copy_ms.append(a.elapsed_time(b))
    copy_gbs = 2 * keys.numel() * 8 / (min(copy_ms) * 1e-3) / 1e9
    del dst

    traffic = None
    tf = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tf):
        traffic = json.load(open(tf)).get("onesweep_u64_bytes_per_launch")

    cpu_inputs = None
    if not args.no_cpu:
        cpu_inputs = {"config2_keys": keys[: 2 * RUN_ROWS].cpu().numpy().view("uint64")}

    print(f"sort: {ms:.2f} ms per step", file=sys.stderr, flush=True)
    extra_ops = None if args.no_ops else op_lines(keys, cpu_inputs)
    print("op lines done", file=sys.stderr, flush=True)
    nruns = len(job.runs)
    del job
    del keys
    torch.cuda.empty_cache()
    configs = None if args.no_configs else {"config1_filter_sort_project": config1_line(cpu_inputs),
                                            "config3_star_join_group_by": config3_line(cpu_inputs),
                                            "a23_csv_load": csv_load_line()}
    print("config lines done", file=sys.stderr, flush=True)
    cpu = None if cpu_inputs is None else cpu_baseline(cpu_inputs)
    del cpu_inputs
    print("cpu baseline done", file=sys.stderr, flush=True)
    batched = None if args.no_batched else batched_lines()
    print("batched lines done", file=sys.stderr, flush=True)
    serving = None if args.no_serving else serving_lines(args.serving_queries, args.serving_seconds)

    out = {
        "metric": "rows/s & %HBM roofline per op; batched vs plain speedup; p99 query latency",
        "value": rows_per_s,
        "unit": "rows/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic (seeded torch generator on the GPU: 50% uniform [0,2^63), 50% Zipf(1.1) over 1e7)",
        "config": {
            "workload": "out-of-core multi-run sort: 200M int64 keys + uint32 row index, 2^25-row runs "
                        "radix-sorted with 8-bit digits, then merged ("
                        + ("pairwise merge path, level by level" if args.pairwise_merge else "one k-way merge") + ")",
            "rows": args.rows,
            "run_rows": RUN_ROWS,
            "runs": nruns,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": pass_gbs,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": pass_gbs / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_source": None if traffic is None else
            "profiles/onesweep_u64_hbm_traffic.md: rocprofv3 FETCH_SIZE + WRITE_SIZE of one digit pass of a "
            "2^25-row run, each scaled by the counter's reading for a device copy of known size (not measured "
            "in this run)",
            "kernel": "k_onesweep<uint64,512,16> (one 8-bit digit pass of a 2^25-row run)",
            "bytes_per_launch": 2 * RUN_ROWS * 12,
            "avg_launch_ms": sum(pass_ms) / len(pass_ms),
        },
        "ops": {
            "onesweep_pass": {"GB_s": pass_gbs, "frac": pass_gbs / HBM_PEAK_GBS,
                              "avg_ms": sum(pass_ms) / len(pass_ms), "launches": len(pass_ms)},
            "merge": None if merge_gbs is None else {"GB_s": merge_gbs, "frac": merge_gbs / HBM_PEAK_GBS,
                                                     "avg_ms": sum(merge_ms) / len(merge_ms),
                                                     "calls_per_step": len(merge_ms) // max(1, args.steps),
                                                     "kind": "pairwise" if args.pairwise_merge else "k-way"},
            "whole_job": {"algorithmic_GB": alg / 1e9, "GB_s": alg / (ms * 1e-3) / 1e9,
                          "frac": alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS},
            "device_copy": {"GB_s": copy_gbs, "frac": copy_gbs / HBM_PEAK_GBS,
                            "what": "torch copy_ of 1.6 GB (read + write), best of 3: attainable HBM rate"},
            **(extra_ops or {}),
        },
        "configs": configs,
        "batched": batched,
        "serving": serving,
        "cpu_baseline": cpu,
        "verified": not args.no_verify,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
