"""numpy binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, never by the product package.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


if not os.path.exists(_LIB):
    build()
lib = C.CDLL(_LIB)

P = C.c_void_p
U64 = C.c_uint64
I = C.c_int


class Acc(C.Structure):
    _fields_ = [("i", C.c_int64), ("d", C.c_double), ("f", C.c_float), ("pad", C.c_uint32), ("count", C.c_int64)]


class StarDim(C.Structure):
    _fields_ = [
        ("fk", P), ("rows", C.c_uint32), ("group_card", C.c_uint32), ("pass_", P), ("group", P),
        ("group_base", C.c_int32), ("pad", C.c_uint32),
    ]


_SIG = {
    "orc_sortkey_encode": ([I, I, P, P, U64], I),
    "orc_stable_argsort": ([I, I, P, P, U64], I),
    "orc_merge": ([I, I, P, U64, P, U64, U64, P, P, P, P], I),
    "orc_scan_u32_inclusive": ([P, P, U64], I),
    "orc_scan_mask_inclusive": ([P, P, U64], I),
    "orc_compare": ([I, I, P, I, P, I, I, P, U64], I),
    "orc_mask_logic": ([I, P, I, P, I, P, U64], I),
    "orc_arith": ([I, I, P, I, P, I, P, U64], I),
    "orc_cast": ([I, I, P, P, U64], I),
    "orc_compact": ([C.c_uint32, P, P, P, U64, P, P], U64),
    "orc_gather": ([C.c_uint32, P, P, P, U64, P, P], I),
    "orc_gather2": ([C.c_uint32, P, P, P, P, P, P, U64, I, P, P], I),
    "orc_gather_pairs": ([C.c_uint32, P, P, P, I, U64, P, P], I),
    "orc_acc_init": ([I, I, P], I),
    "orc_reduce": ([I, I, P, P, U64, P], I),
    "orc_agg_finalize": ([I, I, P, P], I),
    "orc_group_reduce": ([I, I, P, C.c_uint32, P, P, U64, P], I),
    "orc_join_pairs": ([I, I, P, U64, P, U64, P, U64], U64),
    "orc_pairs_intersect": ([P, U64, P, U64, P, U64], U64),
    "orc_cross_expand": ([C.c_uint32, P, P, U64, U64, I, P, P], I),
    "orc_star_join_group_sum": ([C.POINTER(StarDim), I, P, P, U64, P, P, C.c_uint32], I),
    "orc_cpu_multirun_sort_u64": ([P, U64, U64, I, P, P], I),
    "orc_sort_rows": ([I, C.POINTER(I), C.POINTER(I), C.POINTER(P), U64, P], I),
    "cpb_hardware_concurrency": ([], C.c_uint),
    "cpb_multirun_sort_u64": ([P, U64, U64, I, P, P], I),
    "cpb_filter_aggregate": ([P, P, U64, C.c_float, I, P, P, P, P], I),
    "cpb_filter_compact": ([P, P, U64, C.c_float, I, P, P], U64),
    "cpb_join_eq": ([P, U64, P, U64, I, P, U64], U64),
    "cpb_star_join_group_sum": ([C.POINTER(StarDim), I, P, P, U64, I, P, P, C.c_uint32], I),
    "cpb_config1": ([P, P, U64, I, P, P, P], U64),
}
for _n, (_a, _r) in _SIG.items():
    f = getattr(lib, _n)
    f.argtypes = _a
    f.restype = _r

INT, FLOAT, DATETIME, STRING, BOOL, UKEY32, UKEY64 = 0, 1, 2, 3, 4, 16, 17
_DT = {np.dtype(np.int32): INT, np.dtype(np.float32): FLOAT, np.dtype(np.int64): DATETIME, np.dtype(np.uint8): BOOL}


def _p(a):
    return None if a is None else a.ctypes.data_as(P)


def dt(a):
    if a.dtype == np.uint8 and a.ndim == 2 and a.shape[1] == 256:
        return STRING
    return _DT[a.dtype]


def eb(a):
    return a.itemsize * (a.shape[1] if a.ndim == 2 else 1)


def _chk(rc):
    if rc != 0:
        raise RuntimeError(f"oracle returned {rc}")


def sortkey_encode(x, desc=False):
    x = np.ascontiguousarray(x)
    out = np.empty(x.shape, dtype=np.uint64 if x.itemsize == 8 else np.uint32)
    _chk(lib.orc_sortkey_encode(dt(x), int(desc), _p(x), _p(out), x.size))
    return out


def stable_argsort(keys, desc=False):
    keys = np.ascontiguousarray(keys)
    perm = np.empty(keys.size, dtype=np.uint32)
    _chk(lib.orc_stable_argsort(UKEY64 if keys.itemsize == 8 else UKEY32, int(desc), _p(keys), _p(perm), keys.size))
    return perm


def merge(dtype_c, desc, A, B, k_max=None, want_keys=True):
    A = np.ascontiguousarray(A)
    B = np.ascontiguousarray(B)
    m, n = A.shape[0], B.shape[0]
    kmax = m + n if k_max is None else min(k_max, m + n)
    Cc = np.empty((kmax,) + A.shape[1:], dtype=A.dtype) if want_keys else None
    idx = np.empty(kmax, dtype=np.uint32)
    isa = np.empty(kmax, dtype=np.uint8)
    cons = np.zeros(1, dtype=np.int64)
    _chk(lib.orc_merge(dtype_c, int(desc), _p(A), m, _p(B), n, kmax, _p(Cc), _p(idx), _p(isa), _p(cons)))
    return Cc, idx, isa, int(cons[0])


def scan_u32(x):
    x = np.ascontiguousarray(x, dtype=np.uint32)
    out = np.empty_like(x)
    _chk(lib.orc_scan_u32_inclusive(_p(x), _p(out), x.size))
    return out


def scan_mask(m):
    m = np.ascontiguousarray(m, dtype=np.uint8)
    out = np.empty(m.size, dtype=np.uint32)
    _chk(lib.orc_scan_mask_inclusive(_p(m), _p(out), m.size))
    return out


def compare(cmp, lhs, rhs, lhs_scalar=False, rhs_scalar=False, negate=False, n=None):
    lhs = np.ascontiguousarray(lhs)
    rhs = np.ascontiguousarray(rhs)
    n = n if n is not None else (rhs.shape[0] if lhs_scalar else lhs.shape[0])
    out = np.empty(n, dtype=np.uint8)
    _chk(lib.orc_compare(dt(lhs), cmp, _p(lhs), int(lhs_scalar), _p(rhs), int(rhs_scalar), int(negate), _p(out), n))
    return out


def mask_logic(op, a, b, a_scalar=False, b_scalar=False, n=None):
    a = np.ascontiguousarray(a, dtype=np.uint8)
    b = np.ascontiguousarray(b, dtype=np.uint8)
    n = n if n is not None else (b.size if a_scalar else a.size)
    out = np.empty(n, dtype=np.uint8)
    _chk(lib.orc_mask_logic(op, _p(a), int(a_scalar), _p(b), int(b_scalar), _p(out), n))
    return out


def arith(op, lhs, rhs, lhs_scalar=False, rhs_scalar=False, n=None):
    lhs = np.ascontiguousarray(lhs)
    rhs = np.ascontiguousarray(rhs)
    n = n if n is not None else (rhs.size if lhs_scalar else lhs.size)
    out = np.empty(n, dtype=lhs.dtype)
    _chk(lib.orc_arith(dt(lhs), op, _p(lhs), int(lhs_scalar), _p(rhs), int(rhs_scalar), _p(out), n))
    return out


def cast(x, dst):
    x = np.ascontiguousarray(x)
    out = np.empty(x.shape, dtype=dst)
    _chk(lib.orc_cast(dt(x), _DT[np.dtype(dst)], _p(x), _p(out), x.size))
    return out


def compact(src, mask, src_null=None):
    src = np.ascontiguousarray(src)
    mask = np.ascontiguousarray(mask, dtype=np.uint8)
    dst = np.empty_like(src)
    dn = np.empty(mask.size, dtype=np.uint8)
    c = lib.orc_compact(eb(src), _p(src), _p(src_null), _p(mask), mask.size, _p(dst), _p(dn))
    return dst[:c], dn[:c]


def gather(src, idx, src_null=None):
    src = np.ascontiguousarray(src)
    idx = np.ascontiguousarray(idx, dtype=np.uint32)
    dst = np.empty((idx.size,) + src.shape[1:], dtype=src.dtype)
    dn = np.empty(idx.size, dtype=np.uint8)
    _chk(lib.orc_gather(eb(src), _p(src), _p(src_null), _p(idx), idx.size, _p(dst), _p(dn)))
    return dst, dn


def gather2(a, b, idx, is_a, reverse=False, a_null=None, b_null=None):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    idx = np.ascontiguousarray(idx, dtype=np.uint32)
    is_a = np.ascontiguousarray(is_a, dtype=np.uint8)
    dst = np.empty((idx.size,) + a.shape[1:], dtype=a.dtype)
    dn = np.empty(idx.size, dtype=np.uint8)
    _chk(lib.orc_gather2(eb(a), _p(a), _p(a_null), _p(b), _p(b_null), _p(idx), _p(is_a), idx.size, int(reverse),
                         _p(dst), _p(dn)))
    return dst, dn


def gather_pairs(src, pairs, side, src_null=None):
    src = np.ascontiguousarray(src)
    pairs = np.ascontiguousarray(pairs, dtype=np.uint32)
    n = pairs.shape[0]
    dst = np.empty((n,) + src.shape[1:], dtype=src.dtype)
    dn = np.empty(n, dtype=np.uint8)
    _chk(lib.orc_gather_pairs(eb(src), _p(src), _p(src_null), _p(pairs), side, n, _p(dst), _p(dn)))
    return dst, dn


def acc_new(dtype_c, agg, n=1):
    accs = (Acc * n)()
    for k in range(n):
        _chk(lib.orc_acc_init(dtype_c, agg, C.byref(accs[k])))
    return accs


def reduce(dtype_c, agg, x, acc, null=None):
    x = np.ascontiguousarray(x)
    _chk(lib.orc_reduce(dtype_c, agg, _p(x), _p(null), x.size, C.byref(acc[0])))
    return acc


def agg_finalize(dtype_c, agg, acc):
    out = np.zeros(8, dtype=np.uint8)
    _chk(lib.orc_agg_finalize(dtype_c, agg, C.byref(acc[0]), _p(out)))
    return out


def group_reduce(dtype_c, agg, gid, n_groups, x, accs, null=None):
    x = np.ascontiguousarray(x)
    gid = np.ascontiguousarray(gid, dtype=np.uint32)
    _chk(lib.orc_group_reduce(dtype_c, agg, _p(gid), n_groups, _p(x), _p(null), gid.size, accs))
    return accs


def join_pairs(cmp, left, right):
    left = np.ascontiguousarray(left)
    right = np.ascontiguousarray(right)
    Ln, Rn = left.shape[0], right.shape[0]
    c = lib.orc_join_pairs(dt(left), cmp, _p(left), Ln, _p(right), Rn, None, 0)
    pairs = np.empty((max(c, 1), 2), dtype=np.uint32)
    lib.orc_join_pairs(dt(left), cmp, _p(left), Ln, _p(right), Rn, _p(pairs), c)
    return pairs[:c]


def pairs_intersect(p1, p2):
    p1 = np.ascontiguousarray(p1, dtype=np.uint32)
    p2 = np.ascontiguousarray(p2, dtype=np.uint32)
    out = np.empty((max(p1.shape[0], 1), 2), dtype=np.uint32)
    c = lib.orc_pairs_intersect(_p(p1), p1.shape[0], _p(p2), p2.shape[0], _p(out), p1.shape[0])
    return out[:c]


def cross_expand(src, Ln, Rn, is_left, src_null=None):
    src = np.ascontiguousarray(src)
    dst = np.empty((Ln * Rn,) + src.shape[1:], dtype=src.dtype)
    dn = np.empty(Ln * Rn, dtype=np.uint8)
    _chk(lib.orc_cross_expand(eb(src), _p(src), _p(src_null), Ln, Rn, int(is_left), _p(dst), _p(dn)))
    return dst, dn


def star_join_group_sum(dims, m1, m2, n_groups):
    arr = (StarDim * len(dims))()
    keep = []
    for i, d in enumerate(dims):
        fk = np.ascontiguousarray(d["fk"], dtype=np.int32)
        ps = None if d.get("pass") is None else np.ascontiguousarray(d["pass"], dtype=np.uint8)
        gr = None if d.get("group") is None else np.ascontiguousarray(d["group"], dtype=np.int32)
        keep += [fk, ps, gr]
        arr[i] = StarDim(fk.ctypes.data, d["rows"], d.get("group_card", 0), None if ps is None else ps.ctypes.data,
                         None if gr is None else gr.ctypes.data, d.get("group_base", 0), 0)
    m1 = np.ascontiguousarray(m1, dtype=np.int32)
    m2 = np.ascontiguousarray(m2, dtype=np.int32)
    sums = np.zeros(n_groups, dtype=np.int64)
    counts = np.zeros(n_groups, dtype=np.uint64)
    _chk(lib.orc_star_join_group_sum(arr, len(dims), _p(m1), _p(m2), m1.size, _p(sums), _p(counts), n_groups))
    return sums, counts


def cpu_multirun_sort_u64(keys, run_rows, threads=1):
    keys = np.ascontiguousarray(keys, dtype=np.uint64)
    ok = np.empty_like(keys)
    ov = np.empty(keys.size, dtype=np.uint32)
    _chk(lib.orc_cpu_multirun_sort_u64(_p(keys), keys.size, run_rows, threads, _p(ok), _p(ov)))
    return ok, ov

LT, LE, GT, GE, EQ, NE = 0, 1, 2, 3, 4, 5


def sort_rows(cols, desc):
    """Stable ORDER BY permutation over key columns (int32, float32, int64
    arrays or (n, 256) uint8 string arrays), desc[c] per column."""
    cols = [np.ascontiguousarray(c) for c in cols]
    n = cols[0].shape[0]
    k = len(cols)
    perm = np.empty(n, dtype=np.uint32)
    ptrs = (P * k)(*[c.ctypes.data_as(P) for c in cols])
    _chk(lib.orc_sort_rows(k, (I * k)(*[dt(c) for c in cols]), (I * k)(*[int(d) for d in desc]), ptrs, n, _p(perm)))
    return perm


# ---- multi-threaded CPU baselines (oracle/cpu_baseline.cpp) ---------------
def cpb_hardware_concurrency():
    return int(lib.cpb_hardware_concurrency())


def cpb_multirun_sort(keys, run_rows, threads):
    keys = np.ascontiguousarray(keys, dtype=np.uint64)
    ok = np.empty(keys.size, np.uint64)
    ov = np.empty(keys.size, np.uint32)
    _chk(lib.cpb_multirun_sort_u64(_p(keys), keys.size, run_rows, threads, _p(ok), _p(ov)))
    return ok, ov


def cpb_filter_aggregate(vals, keys, thr, threads):
    c, s_, lo, hi = np.zeros(1, np.uint64), np.zeros(1, np.float64), np.zeros(1, np.int64), np.zeros(1, np.int64)
    _chk(lib.cpb_filter_aggregate(_p(vals), _p(keys), vals.size, thr, threads, _p(c), _p(s_), _p(lo), _p(hi)))
    return int(c[0]), float(s_[0]), int(lo[0]), int(hi[0])


def cpb_filter_compact(vals, keys, thr, threads):
    ok = np.empty(vals.size, np.int64)
    ov = np.empty(vals.size, np.float32)
    m = lib.cpb_filter_compact(_p(vals), _p(keys), vals.size, thr, threads, _p(ok), _p(ov))
    return ok[:m], ov[:m]


def cpb_join_eq(left, right, threads, cap):
    pairs = np.empty((cap, 2), np.uint32)
    m = lib.cpb_join_eq(_p(left), left.size, _p(right), right.size, threads, _p(pairs), cap)
    return pairs[:min(m, cap)], int(m)


def cpb_star_join_group_sum(dims, m1, m2, n_groups, threads):
    arr = (StarDim * len(dims))()
    keep = []
    for i, d in enumerate(dims):
        fk = np.ascontiguousarray(d["fk"], dtype=np.int32)
        ps = None if d.get("pass") is None else np.ascontiguousarray(d["pass"], dtype=np.uint8)
        gr = None if d.get("group") is None else np.ascontiguousarray(d["group"], dtype=np.int32)
        keep += [fk, ps, gr]
        arr[i] = StarDim(fk.ctypes.data, d["rows"], d.get("group_card", 0), None if ps is None else ps.ctypes.data,
                         None if gr is None else gr.ctypes.data, d.get("group_base", 0), 0)
    sums = np.zeros(n_groups, dtype=np.int64)
    counts = np.zeros(n_groups, dtype=np.uint64)
    _chk(lib.cpb_star_join_group_sum(arr, len(dims), _p(m1), _p(m2), m1.size, threads, _p(sums), _p(counts),
                                     n_groups))
    return sums, counts


def cpb_config1(ids, price, threads):
    n = ids.size
    oi, op, o2 = np.empty(n, np.int32), np.empty(n, np.float32), np.empty(n, np.int32)
    m = lib.cpb_config1(_p(ids), _p(price), n, threads, _p(oi), _p(op), _p(o2))
    return oi[:m], op[:m], o2[:m]


# ---------------------------------------------------------------------------
# Expression trees, evaluated node by node as the reference's evaluate()
# chain does (include/engine/operators/expressions/): every node is one call
# of the oracle's per-node function above on its children's results.
#   ("ref", i)                     bound_ref.hpp:30-37: column i (values, nulls)
#   ("i32"|"f32"|"i64", v)         bound_value.hpp:21-48: a 1-element scalar
#   ("cmp", op, l, r)              filters/comparison.hpp:51-215
#   ("logic", op, l, r)            filters/conjunction.hpp:76-113
#   ("arith", op, l, r)            bound_func.hpp:45-167
#   ("cast", dtype, child)         cast.hpp:23-90
# NULL rules: a comparison or arithmetic node takes the nullset of its
# non-constant operand, children[0]'s when it is not a bound value
# (comparison.hpp:213, bound_func.hpp:165); a conjunction its left
# operand's (conjunction.hpp:111); a cast its child's (cast.hpp:87).  A
# constant next to an operand of another numeric type is converted to that
# type first (the reference left this to DuckDB's literal casts).
# Returns (values, is_scalar, nulls or None); masks are uint8.
# ---------------------------------------------------------------------------
_NP_OF = {INT: np.int32, FLOAT: np.float32, DATETIME: np.int64}
_CONST_DT = {"i32": np.int32, "f32": np.float32, "i64": np.int64}


def _tree_type(node, cols):
    k = node[0]
    if k == "ref":
        return np.dtype(cols[node[1]].dtype)
    if k in _CONST_DT:
        return np.dtype(_CONST_DT[k])
    if k in ("cmp", "logic"):
        return np.dtype(np.int32)  # a mask reports INT, as the engine's returnType does
    if k == "arith":
        return _tree_type(node[2], cols)
    return np.dtype(_NP_OF[node[1]])


def _coerce(node, want):
    if node[0] not in _CONST_DT:
        return node
    v = node[1]
    if want == np.dtype(np.float32):
        return ("f32", float(np.float32(v)))
    return ("i64" if want == np.dtype(np.int64) else "i32", int(np.int64(np.float32(v))) if node[0] == "f32" else int(v))


def eval_tree(node, cols, nulls=None):
    nulls = nulls or [None] * len(cols)
    k = node[0]
    if k == "ref":
        return np.ascontiguousarray(cols[node[1]]), False, nulls[node[1]]
    if k in _CONST_DT:
        return np.array([node[1]], dtype=_CONST_DT[k]), True, None
    if k == "cast":
        v, sc, nl = eval_tree(node[2], cols, nulls)
        return cast(v, _NP_OF[node[1]]), sc, nl
    lnode, rnode = node[2], node[3]
    if k in ("cmp", "arith"):
        rnode = _coerce(rnode, _tree_type(lnode, cols))
        lnode = _coerce(lnode, _tree_type(rnode, cols))
    lv, ls, ln = eval_tree(lnode, cols, nulls)
    rv, rs, rn = eval_tree(rnode, cols, nulls)
    n = 1 if (ls and rs) else (rv.shape[0] if ls else lv.shape[0])
    if k == "logic":
        return mask_logic(node[1], lv, rv, ls, rs, n=n), ls and rs, ln
    nl = rn if ls else ln
    if k == "cmp":
        return compare(node[1], lv, rv, ls, rs, n=n), ls and rs, nl
    return arith(node[1], lv, rv, ls, rs, n=n), ls and rs, nl
