"""Writes the golden fixtures of this directory from the CPU oracle.

Each fixture is a small seeded input (<= 4,096 rows) with the oracle's
expected output, saved as .npz.  tests/test_golden.py checks that the
oracle still reproduces them (CPU) and that the HIP kernels match them (GPU).
Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import pyoracle as O  # noqa: E402


def cases():
    r = np.random.default_rng(2024)
    out = {}
    # A1-A6: stable argsort of encoded keys, ties and extremes included
    k32 = r.integers(-50, 50, 4096).astype(np.int32)
    k32[:4] = [np.iinfo(np.int32).min, np.iinfo(np.int32).max, 0, -1]
    f32 = r.standard_normal(3000).astype(np.float32)
    f32[:6] = [0.0, -0.0, np.inf, -np.inf, np.nan, 1e-45]
    i64 = r.integers(-2**62, 2**62, 4095).astype(np.int64)
    i64[::7] = 12345
    for name, x in (("int32", k32), ("float32", f32), ("int64", i64)):
        for desc in (0, 1):
            enc = O.sortkey_encode(x, desc=bool(desc))
            out[f"sort_{name}_{'desc' if desc else 'asc'}"] = {"x": x, "perm": O.stable_argsort(enc)}
    # A8-A10: merge of two sorted encoded runs with cross ties
    a = np.sort(r.integers(0, 300, 2000).astype(np.uint32))
    b = np.sort(r.integers(0, 300, 1500).astype(np.uint32))
    mi = O.merge(O.UKEY32, 0, a, b)
    out["merge_u32"] = {"a": a, "b": b, **{f"out{i}": v for i, v in enumerate(mi) if isinstance(v, np.ndarray)}}
    # A12-A14: predicate mask and compaction
    p = r.integers(-100, 100, 4097).astype(np.int32)
    m = O.compare(2, p, np.array([10], np.int32), rhs_scalar=True)  # GT
    vals = r.random(4097).astype(np.float32)
    cv, _ = O.compact(vals, m)
    out["filter_gt10"] = {"p": p, "vals": vals, "mask": m, "out": cv}
    # A16: reductions of int32 (bit-exact)
    x = r.integers(-10**6, 10**6, 4000).astype(np.int32)
    red = []
    for agg in range(6):
        acc = O.acc_new(O.INT, agg)
        O.reduce(O.INT, agg, x, acc)
        red.append([acc[0].i, acc[0].count])
    out["reduce_int32"] = {"x": x, "i_count": np.array(red, np.int64)}
    # A18: equi and theta join pair lists
    l = r.integers(0, 60, 700).astype(np.int32)
    rr = r.integers(0, 60, 300).astype(np.int32)
    out["join_eq"] = {"l": l, "r": rr, "pairs": O.join_pairs(4, l, rr)}
    out["join_lt"] = {"l": l[:200], "r": rr[:100], "pairs": O.join_pairs(0, l[:200], rr[:100])}
    # A-gap: grouped int64 sums
    gid = r.integers(0, 37, 4000).astype(np.uint32)
    v = r.integers(-10**9, 10**9, 4000).astype(np.int64)
    accs = O.acc_new(O.DATETIME, 2, n=37)
    O.group_reduce(O.DATETIME, 2, gid, 37, v, accs)
    out["group_sum_i64"] = {"gid": gid, "v": v, "sums": np.array([a.i for a in accs], np.int64),
                            "counts": np.array([a.count for a in accs], np.int64)}
    return out


def main():
    for name, arrays in cases().items():
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **arrays)
    print("wrote", len(cases()), "fixtures")


if __name__ == "__main__":
    main()
