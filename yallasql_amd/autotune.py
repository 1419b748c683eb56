"""Autotuner of the sort's digit-pass tile shape, per key width and size
bucket.

For each key width (32, 64) and size bucket floor(log2(n)) it times every
tile variant of ysql_radix_sort_variant on uniform keys of 1.5 * 2^bucket
rows (median of HIP-event timings) and keeps the fastest.  Every variant
returns the same permutation, so the choice changes speed only.  The table
is written as JSON next to this file and fed to the batched executor
(JobSet.create(sort_variants=...)).

    python -m yallasql_amd.autotune            # tune on cuda:0, write tuning/<arch>.json
"""
import ctypes as C
import json
import os
import sys

import torch

from . import _lib as L
from . import ops

HERE = os.path.dirname(os.path.abspath(__file__))
BUCKETS = range(13, 23)  # 8K .. 4M rows; up to 8192 keys sort in one workgroup


def _time_variant(kdt, variant, keys, kout, vout, ws, reps):
    n = keys.numel()
    bits = 64 if kdt == L.UKEY64 else 32

    def run():
        L.call("ysql_radix_sort_variant", kdt, variant, 0, ops._p(keys), None, 0, ops._p(kout), ops._p(vout), n, 0,
               bits, ops._p(ws), ws.numel(), ops._s())

    run()
    ts = []
    for _ in range(reps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        run()
        b.record()
        b.synchronize()
        ts.append(a.elapsed_time(b))
    ts.sort()
    return ts[len(ts) // 2]


def tune_sort_variants(device="cuda:0", buckets=BUCKETS, reps=7, seed=0):
    dev = torch.device(device)
    nvar = L.lib.ysql_radix_variant_count()
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    out = {"arch": torch.cuda.get_device_properties(dev).gcnArchName.split(":")[0], "variants": nvar}
    for width in (32, 64):
        kdt = L.UKEY32 if width == 32 else L.UKEY64
        table = [-1] * 32
        times = {}
        for b in buckets:
            n = int(1.5 * (1 << b))
            if width == 32:
                keys = torch.randint(-2**31, 2**31, (n,), generator=g, device=dev, dtype=torch.int64).to(torch.int32)
            else:
                keys = torch.randint(-2**63, 2**63 - 1, (n,), generator=g, device=dev, dtype=torch.int64)
            kout = torch.empty_like(keys)
            vout = torch.empty(n, dtype=torch.int32, device=dev)
            ws = ops.workspace(L.ws_bytes("ysql_radix_sort_ws_bytes", kdt, n), dev)
            ms = [_time_variant(kdt, v, keys, kout, vout, ws, reps) for v in range(nvar)]
            table[b] = int(min(range(nvar), key=lambda v: ms[v]))
            times[str(b)] = [round(m, 4) for m in ms]
        out[f"u{width}"] = table
        out[f"ms_u{width}"] = times
    return out


def table_path(arch="gfx950"):
    return os.path.join(HERE, "tuning", f"{arch}.json")


def load_sort_variants(arch="gfx950"):
    """{32: [...], 64: [...]} from the saved table, or None."""
    p = table_path(arch)
    if not os.path.exists(p):
        return None
    t = json.load(open(p))
    return {32: t["u32"], 64: t["u64"]}


def main():
    res = tune_sort_variants()
    path = sys.argv[1] if len(sys.argv) > 1 else table_path(res["arch"])
    with open(path, "w") as f:
        json.dump(res, f, indent=1)
    print(json.dumps({"u32": res["u32"], "u64": res["u64"], "path": path}))


if __name__ == "__main__":
    main()
