"""Benchmark: out-of-core multi-run sort on one MI355X.

Workload (one step): 200,000,000 rows; key int64 drawn 50% uniform over
[0, 2^63) and 50% Zipf(s=1.1) over 10^7 distinct values; payload = uint32
row index.  The table is cut into 256 MB key batches of 2^25 rows (6 runs:
5 x 33,554,432 + 32,227,840); each run is radix-sorted with 8-bit digits
(8 passes over the 63 significant bits), then the runs are merged pairwise
with merge path, level by level, as the reference's OrderOperator does
(order_operator.hpp:160-193).  Inputs are resident in HBM before timing.

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
    return ap.parse_args()


def gen_keys(n, dev):
    """Config-2 keys on the GPU from a seeded torch generator."""
    from yallasql_amd.datagen import multirun_keys

    return multirun_keys(n, dev, seed=SEED)


class MultiRunSort:
    """Runs of RUN_ROWS rows radix-sorted stepwise, then merged pairwise."""

    def __init__(self, keys_in, run_rows):
        import torch

        from yallasql_amd import _lib as L
        from yallasql_amd import ops

        self.L, self.ops, self.torch = L, ops, torch
        dev = keys_in.device
        self.n = n = keys_in.numel()
        self.keys_in = keys_in
        self.runs = [(o, min(run_rows, n - o)) for o in range(0, n, run_rows)]
        self.bk = [torch.empty(n, dtype=torch.int64, device=dev) for _ in range(3)]
        self.bv = [torch.empty(n, dtype=torch.int32, device=dev) for _ in range(3)]
        self.alt_k = torch.empty(run_rows, dtype=torch.int64, device=dev)
        self.alt_v = torch.empty(run_rows, dtype=torch.int32, device=dev)
        self.sort_ws = ops.workspace(L.ws_bytes("ysql_radix_sort_ws_bytes", L.UKEY64, run_rows), dev)
        self.merge_ws = ops.workspace(L.ws_bytes("ysql_merge_ws_bytes", n, n, n), dev)
        self.passes = (KEY_BITS + 7) // 8
        self.pass_events = []   # (start, end, rows) per onesweep launch
        self.merge_events = []  # (start, end, rows) per merge
        self.record = False
        self.final = None

    def _ev(self):
        e = self.torch.cuda.Event(enable_timing=True)
        e.record()
        return e

    def sort_run(self, off, rows):
        L, ops, p = self.L, self.ops, self.ops._p
        s = ops._s()
        kin = self.keys_in[off:off + rows]
        ko, vo = self.bk[0][off:off + rows], self.bv[0][off:off + rows]
        ka, va = self.alt_k[:rows], self.alt_v[:rows]
        ws = self.sort_ws
        L.call("ysql_radix_begin", L.UKEY64, 0, p(kin), rows, 0, KEY_BITS, p(ws), ws.numel(), s)
        src_k, src_v = kin, None
        P = self.passes
        for q in range(P):
            to_out = ((P - 1 - q) & 1) == 0
            dk, dv = (ko, vo) if to_out else (ka, va)
            e0 = self._ev() if self.record else None
            L.call("ysql_radix_pass", L.UKEY64, 0, q, p(src_k), p(src_v), off, p(dk), p(dv), rows, 0, KEY_BITS,
                   p(ws), ws.numel(), s)
            if self.record:
                self.pass_events.append((e0, self._ev(), rows))
            src_k, src_v = dk, dv
        return (0, off, rows)

    def merge(self, a, b):
        (ba, oa, na), (bb, ob, nb) = a, b
        out = min({0, 1, 2} - {ba, bb})
        off = min(oa, ob)
        ok, ov = self.bk[out][off:off + na + nb], self.bv[out][off:off + na + nb]
        e0 = self._ev() if self.record else None
        self.ops.merge_pairs(self.bk[ba][oa:oa + na], self.bv[ba][oa:oa + na], self.bk[bb][ob:ob + nb],
                             self.bv[bb][ob:ob + nb], out_keys=ok, out_vals=ov, ws=self.merge_ws)
        if self.record:
            self.merge_events.append((e0, self._ev(), na + nb))
        return (out, off, na + nb)

    def step(self):
        runs = [self.sort_run(o, r) for o, r in self.runs]
        while len(runs) > 1:
            nxt = [self.merge(runs[i], runs[i + 1]) for i in range(0, len(runs) - 1, 2)]
            if len(runs) % 2:
                nxt.append(runs[-1])
            runs = nxt
        self.final = runs[0]

    def result(self):
        b, o, n = self.final
        return self.bk[b][o:o + n], self.bv[b][o:o + n]

    def algorithmic_bytes(self):
        """Sort: P * 2 * n * (8 + 4) per run; merge: 2 * (m + n) * 12 per merge."""
        b = sum(self.passes * 2 * r * 12 for _, r in self.runs)
        sizes = [r for _, r in self.runs]
        while len(sizes) > 1:
            nxt = []
            for i in range(0, len(sizes) - 1, 2):
                b += 2 * (sizes[i] + sizes[i + 1]) * 12
                nxt.append(sizes[i] + sizes[i + 1])
            if len(sizes) % 2:
                nxt.append(sizes[-1])
            sizes = nxt
        return b


def verify(keys_in, keys_out, vals_out):
    """Sorted, stable, a permutation of the rows, and keys match their rows:
    together these equal the oracle's global stable sort exactly."""
    import torch

    n = keys_in.numel()
    assert keys_out.numel() == n
    d = keys_out[1:] - keys_out[:-1]
    assert bool((d >= 0).all()), "keys not sorted"
    ties = d == 0
    assert bool((vals_out[1:][ties] > vals_out[:-1][ties]).all()), "ties out of input order"
    cnt = torch.zeros(n, dtype=torch.int32, device=keys_in.device)
    cnt.index_add_(0, vals_out.long(), torch.ones_like(vals_out))
    assert bool((cnt == 1).all()), "payload is not a permutation"
    del cnt
    assert bool((keys_in[vals_out.long()] == keys_out).all()), "keys do not match their rows"


def cpu_baseline(keys_host, run_rows):
    """Oracle multi-run sort (same pipeline) on the first two runs, all host threads."""
    from oracle import pyoracle as O

    threads = min(16, os.cpu_count() or 1)
    sample = keys_host[: 2 * run_rows]
    best = None
    for _ in range(2):
        t0 = time.perf_counter()
        O.cpu_multirun_sort_u64(sample, run_rows, threads)
        dt = time.perf_counter() - t0
        best = dt if best is None else min(best, dt)
    return {
        "value": sample.size / best,
        "unit": "rows/s",
        "cores": threads,
        "kind": "port",
        "sample": f"first {sample.size} rows (2 runs of {run_rows} + 1 merge) of the same key stream, "
                  f"oracle LSD radix + stable merge, best of 2",
    }


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
    job = MultiRunSort(keys, RUN_ROWS)
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
        verify(keys, ko, vo)

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

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(keys[: 2 * RUN_ROWS].cpu().numpy().view("uint64"), RUN_ROWS)

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
                        "radix-sorted with 8-bit digits then merge-path merged pairwise",
            "rows": args.rows,
            "run_rows": RUN_ROWS,
            "runs": len(job.runs),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": pass_gbs,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": pass_gbs / HBM_PEAK_GBS,
            "traffic": traffic,
            "kernel": "k_onesweep<uint64,512,16> (one 8-bit digit pass of a 2^25-row run)",
            "bytes_per_launch": 2 * RUN_ROWS * 12,
            "avg_launch_ms": sum(pass_ms) / len(pass_ms),
        },
        "ops": {
            "onesweep_pass": {"GB_s": pass_gbs, "frac": pass_gbs / HBM_PEAK_GBS,
                              "avg_ms": sum(pass_ms) / len(pass_ms), "launches": len(pass_ms)},
            "merge_pairs": None if merge_gbs is None else {"GB_s": merge_gbs, "frac": merge_gbs / HBM_PEAK_GBS,
                                                           "avg_ms": sum(merge_ms) / len(merge_ms)},
            "whole_job": {"algorithmic_GB": alg / 1e9, "GB_s": alg / (ms * 1e-3) / 1e9,
                          "frac": alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS},
            "device_copy": {"GB_s": copy_gbs, "frac": copy_gbs / HBM_PEAK_GBS,
                            "what": "torch copy_ of 1.6 GB (read + write), best of 3: attainable HBM rate"},
        },
        "cpu_baseline": cpu,
        "verified": not args.no_verify,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
