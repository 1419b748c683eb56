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
    ap.add_argument("--no-stream", action="store_true", help="skip config 2 streamed from pinned host memory")
    ap.add_argument("--serving-queries", type=int, default=10_000, help="queries per rate, at most")
    ap.add_argument("--serving-seconds", type=float, default=60.0, help="arrival window per rate, at most")
    ap.add_argument("--config1-only", action="store_true",
                    help="run only the config-1 engine query (for a kernel trace of that query alone)")
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
    out.update(fusion_lines(keys, vals, g))
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


def fusion_lines(keys, vals, g):
    """Expression fusion over the 200M-row table: a two-comparison predicate
    compacting (key, f32), and a three-column projection, each as one fused
    program against the same tree run node by node (one kernel and one
    intermediate column per node, as the reference evaluates it)."""
    import struct

    import torch

    from yallasql_amd import _lib as L
    from yallasql_amd import ops

    n = keys.numel()
    dev = keys.device
    cat = torch.randint(0, 64, (n,), generator=g, device=dev, dtype=torch.int32)
    X = ops.xinsn
    fbits = lambda x: struct.unpack("<I", struct.pack("<f", x))[0]  # noqa: E731
    out = {}
    # WHERE f32 > 0.25 AND cat < 32, compacting (key, f32)
    inp = (L.XInput * 2)(L.XInput(vals.data_ptr(), None, L.FLOAT, 0), L.XInput(cat.data_ptr(), None, L.INT, 0))
    prog = [X(L.XOP_CONST, 2, type=L.FLOAT, imm=fbits(0.25)), X(L.XOP_CMP, 2, a=0, b=2, type=L.FLOAT, sub=L.GT),
            X(L.XOP_CONST, 3, type=L.INT, imm=32), X(L.XOP_CMP, 3, a=1, b=3, type=L.INT, sub=L.LT),
            X(L.XOP_LOGIC, 2, a=2, b=3, sub=L.AND)]
    xp = (L.XInsn * len(prog))(*prog)
    kout, vout = torch.empty_like(keys), torch.empty_like(vals)
    col = (L.Col * 2)()
    col[0] = L.Col(keys.data_ptr(), None, kout.data_ptr(), None, 8, 0)
    col[1] = L.Col(vals.data_ptr(), None, vout.data_ptr(), None, 4, 0)
    cnt = torch.zeros(1, dtype=torch.int64, device=dev)
    ws = ops.workspace(L.ws_bytes("ysql_compact_ws_bytes", n), dev)
    fused = lambda: L.call("ysql_expr_filter_compact", xp, len(prog), 2, inp, 2, col, 2, n, ops._p(cnt),  # noqa: E731
                           ops._p(ws), ws.numel(), ops._s())
    c025 = torch.tensor([0.25], device=dev)
    c32 = torch.tensor([32], dtype=torch.int32, device=dev)
    m1, m2, m3 = (torch.empty(n, dtype=torch.uint8, device=dev) for _ in range(3))

    def nodes():
        L.call("ysql_compare", L.FLOAT, L.GT, ops._p(vals), 0, ops._p(c025), 1, 0, ops._p(m1), n, ops._s())
        L.call("ysql_compare", L.INT, L.LT, ops._p(cat), 0, ops._p(c32), 1, 0, ops._p(m2), n, ops._s())
        L.call("ysql_mask_logic", L.AND, ops._p(m1), 0, ops._p(m2), 0, ops._p(m3), n, ops._s())
        L.call("ysql_compact_cols", ops._p(m3), col, 2, n, ops._p(cnt), ops._p(ws), ws.numel(), ops._s())

    fused()
    mf = int(cnt.item())
    ms_f = _timed_ms(fused)
    nodes()
    assert int(cnt.item()) == mf
    ms_n = _timed_ms(nodes)
    b = n * (4 + 4 + 8) + mf * 12
    out["fused_filter_tree"] = {
        "rows_per_s": n / (ms_f * 1e-3), "ms": ms_f, "node_by_node_ms": ms_n, "speedup": ms_n / ms_f,
        "selected": mf, "GB_s": b / (ms_f * 1e-3) / 1e9, "frac": b / (ms_f * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "what": "WHERE f32 > 0.25 AND i32 < 32 compacting (int64 key, f32) of 200M rows: one fused program vs "
                "2 compares + AND + mask compaction"}
    del m1, m2, m3, kout, vout
    # SELECT 2*cat + 1, f32*1.5 - 0.5, CAST(f32*100 AS INT) + cat
    prog = [X(L.XOP_CONST, 2, type=L.INT, imm=2), X(L.XOP_ARITH, 2, a=2, b=1, type=L.INT, sub=L.MUL),
            X(L.XOP_CONST, 3, type=L.INT, imm=1), X(L.XOP_ARITH, 2, a=2, b=3, type=L.INT, sub=L.ADD),
            X(L.XOP_CONST, 3, type=L.FLOAT, imm=fbits(1.5)), X(L.XOP_ARITH, 3, a=0, b=3, type=L.FLOAT, sub=L.MUL),
            X(L.XOP_CONST, 4, type=L.FLOAT, imm=fbits(0.5)), X(L.XOP_ARITH, 3, a=3, b=4, type=L.FLOAT, sub=L.SUB),
            X(L.XOP_CONST, 4, type=L.FLOAT, imm=fbits(100.0)), X(L.XOP_ARITH, 4, a=0, b=4, type=L.FLOAT, sub=L.MUL),
            X(L.XOP_CAST, 4, a=4, type=L.FLOAT, sub=L.INT), X(L.XOP_ARITH, 4, a=4, b=1, type=L.INT, sub=L.ADD)]
    xp = (L.XInsn * len(prog))(*prog)
    o1, o3 = torch.empty_like(cat), torch.empty_like(cat)
    o2 = torch.empty_like(vals)
    xo = (L.XOutput * 3)(L.XOutput(o1.data_ptr(), None, 2, L.INT), L.XOutput(o2.data_ptr(), None, 3, L.FLOAT),
                         L.XOutput(o3.data_ptr(), None, 4, L.INT))
    fusedp = lambda: L.call("ysql_expr_eval", xp, len(prog), inp, 2, xo, 3, n, ops._s())  # noqa: E731
    t1, t2 = torch.empty_like(cat), torch.empty_like(vals)
    k2, k1, k15, k05, k100 = (torch.tensor([v], dtype=dt, device=dev) for v, dt in
                              ((2, torch.int32), (1, torch.int32), (1.5, torch.float32), (0.5, torch.float32),
                               (100.0, torch.float32)))

    def nodesp():
        s = ops._s()
        L.call("ysql_arith", L.INT, L.MUL, ops._p(k2), 1, ops._p(cat), 0, ops._p(t1), n, s)
        L.call("ysql_arith", L.INT, L.ADD, ops._p(t1), 0, ops._p(k1), 1, ops._p(o1), n, s)
        L.call("ysql_arith", L.FLOAT, L.MUL, ops._p(vals), 0, ops._p(k15), 1, ops._p(t2), n, s)
        L.call("ysql_arith", L.FLOAT, L.SUB, ops._p(t2), 0, ops._p(k05), 1, ops._p(o2), n, s)
        L.call("ysql_arith", L.FLOAT, L.MUL, ops._p(vals), 0, ops._p(k100), 1, ops._p(t2), n, s)
        L.call("ysql_cast", L.FLOAT, L.INT, ops._p(t2), ops._p(t1), n, s)
        L.call("ysql_arith", L.INT, L.ADD, ops._p(t1), 0, ops._p(cat), 0, ops._p(o3), n, s)

    fusedp()
    ref = [o1.clone(), o2.clone(), o3.clone()]
    ms_pf = _timed_ms(fusedp)
    nodesp()
    assert all(torch.equal(a, b) for a, b in zip(ref, (o1, o2, o3)))
    ms_pn = _timed_ms(nodesp)
    bp = n * (4 + 4) + n * 12
    out["fused_projection_tree"] = {
        "rows_per_s": n / (ms_pf * 1e-3), "ms": ms_pf, "node_by_node_ms": ms_pn, "speedup": ms_pn / ms_pf,
        "GB_s": bp / (ms_pf * 1e-3) / 1e9, "frac": bp / (ms_pf * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "what": "SELECT 2*i32+1, f32*1.5-0.5, CAST(f32*100 AS INT)+i32 over 200M rows: one fused program vs 7 "
                "node kernels"}
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


def csv_scale_line(target_bytes=2 << 30):
    """CSV at scale: a file of about 2 GB (a 1M-row block of the config-1
    shape written repeatedly) read into device batches two ways, every batch
    consumed by a fused filter+aggregate kernel on the caller's stream:
    the reference's sizing (4 MiB batches, copies on the caller's stream) and
    scale mode (16M-row batches, 3 pinned staging areas, copies on the
    reader's own stream overlapping the kernels).  Also the pinned
    host-to-device copy rate of 1 GiB: the PCIe roofline of this path."""
    import os
    import tempfile

    import numpy as np
    import pandas as pd
    import torch

    from yallasql_amd import _lib as L
    from yallasql_amd import loader, ops

    dev = torch.device("cuda:0")
    blk = 1 << 20
    r = np.random.default_rng(SEED + 5)
    body = pd.DataFrame({"id": np.arange(blk, dtype=np.int32),
                         "price": np.round(1 + 999 * r.random(blk), 2).astype(np.float32),
                         "qty": r.integers(0, 100, blk, dtype=np.int32)}).to_csv(index=False, header=False)
    body = body.encode()
    reps = max(1, target_bytes // len(body))
    fd, path = tempfile.mkstemp(suffix=".csv")
    os.close(fd)
    try:
        with open(path, "wb") as f:
            f.write(b"id,price,qty\n")
            for _ in range(reps):
                f.write(body)
        size = os.path.getsize(path)
        n = reps * blk
        names, types = ["id", "price", "qty"], [L.INT, L.FLOAT, L.INT]

        def run(scale):
            rd = (loader.CsvReader(path, names, types, batch_rows_=1 << 24, slots=3, async_copies=True) if scale
                  else loader.CsvReader(path, names, types))
            s = torch.cuda.Stream(device=dev)
            accs = [[torch.empty(32, dtype=torch.uint8, device=dev) for _ in range(2)]]
            rows, first = 0, True
            t0 = time.perf_counter()
            with torch.cuda.stream(s):
                while True:
                    b = rd.next_async(stream=s) if scale else rd.next_device(dev, stream=s)
                    if b is None:
                        break
                    cols = [(L.INT, L.AGG_COUNT_STAR, None), (L.INT, L.AGG_SUM, b[0][2])]
                    ops.filter_aggregate_multi(b[0][1], L.LT, [500.0], cols, accs, overwrite=first)
                    first = False
                    rows += int(b[0][0].shape[0])
            s.synchronize()
            dt = time.perf_counter() - t0
            rd.close()
            return rows, dt, [ops.acc_read(a) for a in accs[0]]

        run(True)
        rows_r, t_ref, acc_r = run(False)
        rows_s, t_scale, acc_s = run(True)
        assert rows_r == rows_s == n and acc_r == acc_s
        # PCIe roofline of the path: pinned host -> device, 1 GiB, best of 3
        hb = torch.empty(1 << 30, dtype=torch.uint8, pin_memory=True)
        db = torch.empty(1 << 30, dtype=torch.uint8, device=dev)
        h2d_ms = _timed_ms(lambda: db.copy_(hb, non_blocking=True), reps=3)
        pcie = (1 << 30) / (h2d_ms * 1e-3) / 1e9
        del hb, db
        col_bytes = n * (4 + 4 + 4 + 3)  # three columns and their null bytes cross PCIe
        return {"rows": n, "csv_GB": size / 1e9, "reference_sizing_s": t_ref, "scale_mode_s": t_scale,
                "speedup": t_ref / t_scale, "rows_per_s": n / t_scale, "file_GB_s": size / t_scale / 1e9,
                "h2d_GB": col_bytes / 1e9, "pcie_h2d_GB_s": pcie,
                "pcie_frac": col_bytes / t_scale / 1e9 / pcie,
                "what": "CSV (id INT, price FLOAT, qty INT) of ~2 GB -> device, each batch folded by a fused "
                        "WHERE price < 500 -> COUNT, SUM(qty) kernel: 4 MiB batches with copies on the caller's "
                        "stream vs 16M-row batches, 3 pinned staging areas, copies on the reader's own stream"}
    finally:
        os.unlink(path)


def config2_streamed_line(job, keys, steps):
    """Config 2 out of core: the 200M keys start in pinned host memory; each
    2^25-row run is copied on a copy stream while the previous run sorts,
    then the runs are merged.  Also the bare host-to-device copy of the keys
    (the PCIe bound of this form)."""
    import torch

    from yallasql_amd.multirun import verify_global_sort

    n = keys.numel()
    hk = torch.empty(n, dtype=torch.int64, pin_memory=True)
    hk.copy_(keys)
    tmp = torch.empty_like(keys)
    h2d_ms = _timed_ms(lambda: tmp.copy_(hk, non_blocking=True), reps=3)
    del tmp
    cs = torch.cuda.Stream()
    job.step_streamed(hk, cs)
    torch.cuda.synchronize()
    ts = []
    for _ in range(steps):
        t0 = time.perf_counter()
        job.step_streamed(hk, cs)
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
    ko, vo = job.result()
    verify_global_sort(keys, ko, vo)
    ts.sort()
    ms = ts[len(ts) // 2] * 1e3
    del hk
    return {"rows": n, "ms": ms, "rows_per_s": n / (ms * 1e-3), "h2d_ms": h2d_ms,
            "h2d_GB_s": n * 8 / (h2d_ms * 1e-3) / 1e9, "pcie_frac": (n * 8 / (ms * 1e-3)) / (n * 8 / (h2d_ms * 1e-3)),
            "verified": True,
            "what": "keys in pinned host memory (1.6 GB; the row-index payload is made on the device): runs copied "
                    "on a copy stream while the previous run sorts, then one k-way merge; median of the steps"}


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
    # The values the query needs: the filtered dimensions' foreign keys of every row, then the other
    # foreign keys and the two measures of the qualifying rows only (the kernel reads those columns
    # for the rows that pass the filters).
    nfilt = sum(1 for d in dims if d.get("pass") is not None)
    nq = int(counts.sum().item())
    b = n * 4 * nfilt + nq * (4 * (len(dims) - nfilt) + 8)
    out = {"fact_rows": n, "ms": ms, "rows_per_s": n / (ms * 1e-3), "qualifying_rows": nq,
           "GB_s": b / (ms * 1e-3) / 1e9, "frac": b / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, "algorithmic_GB": b / 1e9,
           "all_columns_GB": n * 24 / 1e9,
           "what": "fused 4-FK star join + GROUP BY year x brand SUM(int64(price)*discount); algorithmic bytes: "
                   "the filtered dimensions' keys of every row, the other keys and measures of qualifying rows"}
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

    if args.config1_only:
        print(json.dumps({"config1_filter_sort_project": config1_line()}), flush=True)
        return
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

    streamed = None
    if not args.no_stream and not args.pairwise_merge:
        streamed = config2_streamed_line(job, keys, args.steps)
    print(f"sort: {ms:.2f} ms per step", file=sys.stderr, flush=True)
    extra_ops = None if args.no_ops else op_lines(keys, cpu_inputs)
    print("op lines done", file=sys.stderr, flush=True)
    nruns = len(job.runs)
    del job
    del keys
    torch.cuda.empty_cache()
    configs = None if args.no_configs else {"config1_filter_sort_project": config1_line(cpu_inputs),
                                            "config3_star_join_group_by": config3_line(cpu_inputs),
                                            "a23_csv_load": csv_load_line(),
                                            "a23_csv_scale": csv_scale_line()}
    if configs is not None and streamed is not None:
        configs["config2_streamed_from_host"] = streamed
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
