"""Pins the CPU oracle (CPU only, no GPU).

The reference ships no tests, fixtures or golden vectors for this path, so
each oracle function is checked against what the mathematics fixes: brute
force over tiny inputs, closed forms, invariants, and independent library
routines (numpy / Python).  Each check is chosen so a plausible mistake
(wrong tie order, swapped operands, dropped term, wrong sign) fails it.
"""
import itertools
import math

import numpy as np
import pytest

from oracle import pyoracle as O

R = np.random.default_rng(20240601)


# ---- A1: key encoding preserves order ---------------------------------------
def test_encode_int_order_and_extremes():
    x = np.array([-2**31, -5, -1, 0, 1, 7, 2**31 - 1], np.int32)
    k = O.sortkey_encode(x)
    assert list(k) == sorted(k) and len(set(k)) == len(k)
    assert k[0] == 0 and k[-1] == 0xFFFFFFFF
    kd = O.sortkey_encode(x, desc=True)
    assert list(kd) == sorted(kd, reverse=True)


def test_encode_float_total_order():
    x = np.array([-np.inf, -1e30, -1.0, -1e-40, -0.0, 0.0, 1e-40, 1.0, 1e30, np.inf], np.float32)
    k = O.sortkey_encode(x)
    assert all(a < b for a, b in zip(k, k[1:]))  # strict: -0 < +0 in the total order
    y = (R.standard_normal(10000) * 100).astype(np.float32)
    assert np.array_equal(np.argsort(O.sortkey_encode(y), kind="stable"), np.argsort(y, kind="stable"))


def test_encode_int64_sign_flip():
    x = np.array([-2**63, -1, 0, 1, 2**63 - 1], np.int64)
    k = O.sortkey_encode(x)
    assert all(a < b for a, b in zip(k, k[1:]))
    assert k[2] == 2**63  # 0 maps to the middle


# ---- A2-A6: stable argsort -------------------------------------------------
def test_argsort_bruteforce_all_small_multisets():
    for n in range(0, 7):
        for vals in itertools.product(range(3), repeat=n):
            keys = np.array(vals, np.uint32)
            for desc in (False, True):
                got = O.stable_argsort(keys, desc)
                exp = sorted(range(n), key=lambda i: (-int(keys[i]) if desc else int(keys[i]), i))
                assert list(got) == exp


def test_argsort_matches_numpy_stable():
    keys = R.integers(0, 1000, 50000).astype(np.uint64)
    assert np.array_equal(O.stable_argsort(keys), np.argsort(keys, kind="stable"))
    # descending with ties kept in input order
    got = O.stable_argsort(keys, desc=True)
    exp = np.lexsort((np.arange(keys.size), -keys.astype(np.int64)))
    assert np.array_equal(got, exp)


# ---- A8-A10: merge = stable sort of the concatenation -----------------------
@pytest.mark.parametrize("m,n", [(0, 0), (7, 1), (1, 7), (50, 60), (1000, 3)])
def test_merge_is_stable_sort_of_concat(m, n):
    A = np.sort(R.integers(0, 5, m)).astype(np.int32)
    B = np.sort(R.integers(0, 5, n)).astype(np.int32)
    if (m, n) == (7, 1):
        A[:], B[:] = 5, 5  # reference co-rank counterexample (merge_batchs.cu:22)
    C, idx, isa, cons = O.merge(O.INT, False, A, B)
    cat = np.concatenate([A, B])
    order = np.argsort(cat, kind="stable")
    assert np.array_equal(C, cat[order])
    src = np.where(isa == 1, idx, idx + m)
    assert np.array_equal(src, order)  # every row once, A first on ties
    assert cons == m


def test_merge_window_consumed_a():
    A = np.arange(0, 20, 2, dtype=np.int32)
    B = np.arange(1, 21, 2, dtype=np.int32)
    _, _, isa, cons = O.merge(O.INT, False, A, B, k_max=7)
    assert cons == int(isa.sum()) == 4  # 0,1,2,3,4,5,6 -> A gives 0,2,4,6


def test_merge_descending():
    A = np.array([9, 5, 5, 1], np.int32)
    B = np.array([7, 5, 2], np.int32)
    C, idx, isa, _ = O.merge(O.INT, True, A, B)
    assert list(C) == [9, 7, 5, 5, 5, 2, 1]
    assert list(isa) == [1, 0, 1, 1, 0, 0, 1]


# ---- A4/A13 ---------------------------------------------------------------
def test_scans_match_cumsum():
    x = R.integers(0, 2**32, 10000, dtype=np.uint64).astype(np.uint32)
    assert np.array_equal(O.scan_u32(x), np.cumsum(x, dtype=np.uint64).astype(np.uint32))
    m = (R.random(10000) < 0.3).astype(np.uint8) * 7
    assert np.array_equal(O.scan_mask(m), np.cumsum(m != 0).astype(np.uint32))


# ---- A12: operand order and ops --------------------------------------------
def test_compare_operand_order_and_ops():
    a = np.array([1, 2, 3], np.int32)
    s = np.array([2], np.int32)
    # LT/LE/GT/GE/EQ/NE = < <= > >= == != with op(lhs, rhs)
    exp = {0: [1, 0, 0], 1: [1, 1, 0], 2: [0, 0, 1], 3: [0, 1, 1], 4: [0, 1, 0], 5: [1, 0, 1]}
    for cmp, e in exp.items():
        assert list(O.compare(cmp, a, s, rhs_scalar=True)) == e
        # swapping the operands mirrors the relation
        assert list(O.compare(cmp, s, a, lhs_scalar=True)) == list(O.compare([2, 3, 0, 1, 4, 5][cmp], a, s, rhs_scalar=True))
    assert list(O.compare(2, a, s, rhs_scalar=True, negate=True)) == [1, 1, 0]
    f = np.array([np.nan, 1.0], np.float32)
    assert list(O.compare(5, f, f)) == [1, 0]  # NaN != NaN


def test_compare_strings_strcmp():
    def S(txt):
        b = np.zeros(256, np.uint8)
        b[: len(txt)] = np.frombuffer(txt.encode(), np.uint8)
        return b

    words = ["", "a", "ab", "abc", "b", "ba"]
    arr = np.stack([S(w) for w in words])
    for i, j in itertools.product(range(len(words)), repeat=2):
        got = O.compare(0, arr[i:i + 1], arr[j:j + 1])[0]
        assert got == (words[i] < words[j])


def test_mask_logic_truth_table():
    a = np.array([0, 0, 1, 1], np.uint8)
    b = np.array([0, 1, 0, 1], np.uint8)
    assert list(O.mask_logic(0, a, b)) == [0, 0, 0, 1]
    assert list(O.mask_logic(1, a, b)) == [0, 1, 1, 1]


# ---- A15 ---------------------------------------------------------------------
def test_arith_matches_numpy_and_zero_rules():
    a = R.integers(-1000, 1000, 1000).astype(np.int32)
    b = R.integers(-5, 5, 1000).astype(np.int32)
    assert np.array_equal(O.arith(0, a, b), a + b)
    assert np.array_equal(O.arith(1, a, b), a - b)
    assert np.array_equal(O.arith(2, a, b), a * b)
    nz = b != 0
    d = O.arith(3, a, b)
    assert np.all(d[~nz] == 0)
    assert np.array_equal(d[nz], np.trunc(a[nz] / b[nz]).astype(np.int32))  # C truncation, not floor
    r = O.arith(4, a, b)
    assert np.all(r[~nz] == 0)
    assert np.array_equal(r[nz], np.fmod(a[nz], b[nz]))
    big = np.array([2**31 - 1], np.int32)
    assert O.arith(0, big, np.array([1], np.int32))[0] == -2**31  # wraps
    s = np.array([2], np.int32)
    assert np.array_equal(O.arith(2, s, a, lhs_scalar=True), 2 * a)
    assert np.array_equal(O.arith(1, s, a, lhs_scalar=True), 2 - a)  # lhs - rhs, not rhs - lhs


def test_cast_matches_numpy_in_range_and_saturates():
    f = np.array([1.5, -1.5, 1e20, -1e20, np.nan], np.float32)
    assert list(O.cast(f, np.int32)) == [1, -1, 2**31 - 1, -2**31, 0]
    i = R.integers(-2**31, 2**31, 1000, dtype=np.int64).astype(np.int32)
    assert np.array_equal(O.cast(i, np.float32), i.astype(np.float32))
    assert np.array_equal(O.cast(i, np.int64), i.astype(np.int64))
    l = R.integers(-2**62, 2**62, 1000, dtype=np.int64)
    assert np.array_equal(O.cast(l, np.int32), l.astype(np.int32))


# ---- A7/A10/A14/A20/A21 --------------------------------------------------------
def test_row_movement_matches_numpy():
    src = R.integers(0, 100, 500).astype(np.int32)
    mask = (R.random(500) < 0.4).astype(np.uint8)
    d, _ = O.compact(src, mask)
    assert np.array_equal(d, src[mask.astype(bool)])
    idx = R.integers(0, 500, 300).astype(np.uint32)
    assert np.array_equal(O.gather(src, idx)[0], src[idx])
    b = R.integers(0, 100, 500).astype(np.int32)
    isa = (R.random(300) < 0.5).astype(np.uint8)
    g2, _ = O.gather2(src, b, idx, isa, reverse=True)
    assert np.array_equal(g2[::-1], np.where(isa == 1, src[idx], b[idx]))
    pairs = R.integers(0, 500, (200, 2)).astype(np.uint32)
    assert np.array_equal(O.gather_pairs(src, pairs, 1)[0], src[pairs[:, 1]])
    x = np.arange(5, dtype=np.int32)
    assert np.array_equal(O.cross_expand(x, 5, 3, True)[0], np.repeat(x, 3))
    assert np.array_equal(O.cross_expand(x[:3], 5, 3, False)[0], np.tile(x[:3], 5))


# ---- A16/A17 -------------------------------------------------------------------
def test_reduce_closed_forms():
    n = 100_000
    x = np.arange(n, dtype=np.int64)
    a = O.acc_new(O.DATETIME, 2)
    O.reduce(O.DATETIME, 2, x, a)
    assert a[0].i == n * (n - 1) // 2 and a[0].count == n
    # int32 SUM does not wrap (accumulates in int64)
    y = np.full(10, 2**31 - 1, np.int32)
    a = O.acc_new(O.INT, 2)
    O.reduce(O.INT, 2, y, a)
    assert a[0].i == 10 * (2**31 - 1)
    # MAX of all-negative floats is negative (identity is -inf, not FLT_MIN)
    f = np.array([-3.0, -2.0, -5.0], np.float32)
    a = O.acc_new(O.FLOAT, 0)
    O.reduce(O.FLOAT, 0, f, a)
    assert a[0].f == -2.0
    # float SUM vs math.fsum, AVG = sum / count, nulls skipped
    g = R.random(20001).astype(np.float32)
    nul = (R.random(20001) < 0.1).astype(np.uint8)
    a = O.acc_new(O.FLOAT, 3)
    O.reduce(O.FLOAT, 3, g, a, null=nul)
    keep = g[nul == 0]
    assert a[0].count == keep.size
    assert abs(a[0].d - math.fsum(keep.astype(np.float64))) <= 1e-9 * a[0].d
    out = O.agg_finalize(O.FLOAT, 3, a)
    assert out[:4].view(np.float32)[0] == np.float32(math.fsum(keep.astype(np.float64)) / keep.size)
    # MIN/MAX int, COUNT, COUNT(*)
    z = R.integers(-10**9, 10**9, 5000).astype(np.int32)
    for agg, exp in ((1, z.min()), (0, z.max())):
        a = O.acc_new(O.INT, agg)
        O.reduce(O.INT, agg, z, a)
        assert a[0].i == exp
    zn = (R.random(5000) < 0.2).astype(np.uint8)
    a = O.acc_new(O.INT, 4)
    O.reduce(O.INT, 4, z, a, null=zn)
    assert a[0].count == int((zn == 0).sum())
    a = O.acc_new(O.INT, 5)
    O.reduce(O.INT, 5, z, a, null=zn)
    assert a[0].count == 5000


def test_group_reduce_matches_numpy():
    gid = R.integers(0, 10, 5000).astype(np.uint32)
    x = R.integers(-100, 100, 5000).astype(np.int64)
    accs = O.acc_new(O.DATETIME, 2, 8)  # ids 8, 9 are skipped
    O.group_reduce(O.DATETIME, 2, gid, 8, x, accs)
    exp = np.zeros(10, np.int64)
    np.add.at(exp, gid, x)
    assert [accs[g].i for g in range(8)] == list(exp[:8])


# ---- A18/A19 -----------------------------------------------------------------------
@pytest.mark.parametrize("cmp", range(6))
def test_join_matches_broadcast(cmp):
    left = R.integers(0, 10, 60).astype(np.int32)
    right = R.integers(0, 10, 40).astype(np.int32)
    ops = [np.less, np.less_equal, np.greater, np.greater_equal, np.equal, np.not_equal]
    l, r = np.nonzero(ops[cmp](left[:, None], right[None, :]))
    got = O.join_pairs(cmp, left, right)
    assert np.array_equal(got, np.stack([l, r], 1).astype(np.uint32))


def test_pairs_intersect_matches_sets():
    p1 = R.integers(0, 6, (300, 2)).astype(np.uint32)
    p2 = R.integers(0, 6, (200, 2)).astype(np.uint32)
    s2 = {tuple(p) for p in p2}
    exp = [tuple(p) for p in p1 if tuple(p) in s2]
    assert [tuple(p) for p in O.pairs_intersect(p1, p2)] == exp


def test_star_join_matches_explicit_joins():
    n = 20000
    fk0 = R.integers(0, 50, n).astype(np.int32)
    fk1 = R.integers(-2, 40, n).astype(np.int32)  # some keys miss the dimension
    g0 = R.integers(10, 13, 50).astype(np.int32)
    pass1 = (R.random(38) < 0.5).astype(np.uint8)
    g1 = R.integers(0, 4, 38).astype(np.int32)
    m1 = R.integers(90000, 10**7, n).astype(np.int32)
    m2 = R.integers(0, 11, n).astype(np.int32)
    dims = [{"fk": fk0, "rows": 50, "group": g0, "group_base": 10, "group_card": 3},
            {"fk": fk1, "rows": 38, "pass": pass1, "group": g1, "group_card": 4}]
    sums, counts = O.star_join_group_sum(dims, m1, m2, 12)
    ok = (fk1 >= 0) & (fk1 < 38)
    ok[ok] &= pass1[fk1[ok]] == 1
    g = (g0[fk0] - 10) * 4 + np.where(ok, g1[np.clip(fk1, 0, 37)], 0)
    es = np.zeros(12, np.int64)
    np.add.at(es, g[ok], m1[ok].astype(np.int64) * m2[ok])
    ec = np.bincount(g[ok], minlength=12)
    assert np.array_equal(sums, es) and np.array_equal(counts.astype(np.int64), ec)
    assert sums.sum() == int((m1[ok].astype(np.int64) * m2[ok]).sum())


# ---- CPU baseline pipeline = one global stable sort -----------------------------
def test_cpu_multirun_sort_is_global_stable_sort():
    keys = R.integers(0, 2**63, 30011, dtype=np.uint64)
    keys[::3] = R.integers(0, 20, keys[::3].size)
    for threads in (1, 4):
        ok, ov = O.cpu_multirun_sort_u64(keys, 4096, threads)
        order = np.argsort(keys, kind="stable")
        assert np.array_equal(ov, order) and np.array_equal(ok, keys[order])


def _strings(r, n, alphabet=b"ab\x80\xffz", maxlen=20):
    """(n, 256) NUL-padded strings over a small alphabet (shared prefixes,
    duplicates, bytes above 0x7f), and their texts."""
    s = np.zeros((n, 256), np.uint8)
    texts = []
    for i in range(n):
        k = int(r.integers(0, maxlen + 1))
        t = bytes(r.choice(list(alphabet), k).tolist()) if k else b""
        s[i, :k] = list(t)
        texts.append(t)
    return s, texts


def test_sort_rows_strings_match_python_sorted():
    """orc_sort_rows on one STRING column equals Python's stable sorted()
    over the bytes up to the first NUL (unsigned lexicographic = strcmp),
    ascending and descending (reverse=True keeps ties in input order)."""
    from oracle import pyoracle as P

    r = np.random.default_rng(71)
    s, texts = _strings(r, 3000)
    s[5, 3:] = 0
    s[5, :3] = list(b"ab\0")[:3]
    s[6, :] = 0x61  # no NUL in 256 bytes: strcmp compares all of them
    s[7, :4] = list(b"ab\0z")  # a byte after the NUL is ignored by strcmp
    texts = [bytes(row).split(b"\0")[0] for row in s]
    for desc in (False, True):
        want = sorted(range(len(texts)), key=lambda i: texts[i], reverse=desc)
        assert list(P.sort_rows([s], [desc])) == want


def test_sort_rows_mixed_keys_match_tuple_sort():
    """Several key columns: string ASC, int DESC, float ASC, datetime DESC
    against Python's stable sorted() on key tuples (a wrong column order,
    direction or tie rule fails)."""
    from oracle import pyoracle as P

    r = np.random.default_rng(72)
    n = 4000
    s, texts = _strings(r, n, alphabet=b"ab", maxlen=3)
    iv = r.integers(-3, 3, n).astype(np.int32)
    fv = r.choice(np.array([-1.5, -0.25, 0.5, 2.0], np.float32), n)
    dv = r.integers(-2, 2, n).astype(np.int64)
    perm = P.sort_rows([s, iv, fv, dv], [0, 1, 0, 1])
    want = sorted(range(n), key=lambda i: (texts[i], -int(iv[i]), float(fv[i]), -int(dv[i])))
    assert list(perm) == want


def test_merge_strings_is_stable_sort_of_concatenation():
    """orc_merge on STRING runs equals the stable sort of A then B (ties
    from A first), checked against Python's sorted()."""
    from oracle import pyoracle as P

    r = np.random.default_rng(73)
    for desc in (False, True):
        a, ta = _strings(r, 700, maxlen=4)
        b, tb = _strings(r, 500, maxlen=4)
        oa = sorted(range(700), key=lambda i: ta[i], reverse=desc)
        ob = sorted(range(500), key=lambda i: tb[i], reverse=desc)
        a, ta = a[oa], [ta[i] for i in oa]
        b, tb = b[ob], [tb[i] for i in ob]
        C_, idx, isa, cons = P.merge(P.STRING, desc, a, b)
        got = [ta[i] if x else tb[i] for i, x in zip(idx, isa)]
        tags = [(0, i) if x else (1, i) for i, x in zip(idx, isa)]
        allt = [(t, (0, i)) for i, t in enumerate(ta)] + [(t, (1, i)) for i, t in enumerate(tb)]
        want = sorted(allt, key=lambda x: x[0], reverse=desc)
        assert got == [w[0] for w in want] and tags == [w[1] for w in want]
        assert cons == 700


def test_cpu_baselines_compute_the_oracle_results():
    """The multi-threaded CPU baselines that bench.py times give the oracle's
    results at 1 and 5 threads (so the CPU figures time correct work)."""
    from oracle import pyoracle as P

    r = np.random.default_rng(81)
    n = 300_001
    keys = np.where(r.random(n) < 0.5, r.integers(0, 2**62, n), r.zipf(1.1, n) % 10**7).astype(np.uint64)
    want_k, want_v = P.cpu_multirun_sort_u64(keys, 70_000, 1)
    vals = r.random(n).astype(np.float32)
    ik = keys.astype(np.int64)
    m = vals < np.float32(0.5)
    right = r.integers(0, 5000, 3000).astype(np.int64)
    left = r.integers(0, 5000, 20_000).astype(np.int64)
    ids = r.permutation(50_000).astype(np.int32)
    price = (1 + 999 * r.random(50_000)).astype(np.float32)
    price[::7] = np.float32(750.5)
    for t in (1, 5):
        k, v = P.cpb_multirun_sort(keys, 70_000, t)
        assert np.array_equal(k, want_k) and np.array_equal(v, want_v)
        c, s_, lo, hi = P.cpb_filter_aggregate(vals, ik, 0.5, t)
        assert c == m.sum() and lo == ik[m].min() and hi == ik[m].max()
        assert abs(s_ - vals[m].astype(np.float64).sum()) <= 1e-9 * s_
        ck, cv = P.cpb_filter_compact(vals, ik, 0.5, t)
        assert np.array_equal(ck, ik[m]) and np.array_equal(cv, vals[m])
        pairs, cnt = P.cpb_join_eq(left, right, t, 1 << 20)
        want = sorted((int(i), int(j)) for i in range(left.size) for j in np.nonzero(right == left[i])[0])
        assert cnt == len(want) and [tuple(p) for p in pairs.tolist()] == want
        oi, op, o2 = P.cpb_config1(ids, price, t)
        sel = price > np.float32(500)
        perm = P.sort_rows([price[sel], ids[sel]], [1, 0])
        assert np.array_equal(oi, ids[sel][perm]) and np.array_equal(op, price[sel][perm])
        assert np.array_equal(o2, 2 * ids[sel][perm])
    dims = [{"fk": r.integers(0, 100, n).astype(np.int32), "rows": 100, "group": r.integers(0, 7, 100).astype(np.int32),
             "group_base": 0, "group_card": 7},
            {"fk": r.integers(0, 50, n).astype(np.int32), "rows": 50, "pass": (r.random(50) < 0.5).astype(np.uint8)}]
    m1 = r.integers(0, 1000, n).astype(np.int32)
    m2 = r.integers(0, 10, n).astype(np.int32)
    ws, wc = P.star_join_group_sum(dims, m1, m2, 7)
    for t in (1, 5):
        s2, c2 = P.cpb_star_join_group_sum(dims, m1, m2, 7, t)
        assert np.array_equal(s2, ws) and np.array_equal(c2, wc)


def test_eval_tree_node_rules():
    """The oracle's node-by-node tree evaluation against numpy written
    directly, and its NULL rule: a comparison or arithmetic node takes the
    NULLs of its non-constant operand (the left one between two columns), a
    conjunction the left operand's, a cast the child's; a constant is
    converted to the other operand's type."""
    r = np.random.default_rng(0)
    n = 1000
    a = r.integers(-20, 20, n).astype(np.int32)
    b = r.integers(-3, 4, n).astype(np.int32)
    f = (r.standard_normal(n) * 10).astype(np.float32)
    na = (r.random(n) < 0.3).astype(np.uint8)
    nb = (r.random(n) < 0.3).astype(np.uint8)
    cols, nulls = [a, b, f], [na, nb, None]
    v, sc, nl = O.eval_tree(("arith", 0, ("i32", 3), ("ref", 0)), cols, nulls)
    assert not sc and np.array_equal(v, a + 3) and np.array_equal(nl, na)
    v, _, nl = O.eval_tree(("arith", 2, ("ref", 1), ("ref", 0)), cols, nulls)
    assert np.array_equal(v, a * b) and np.array_equal(nl, nb)
    v, _, nl = O.eval_tree(("cmp", 0, ("ref", 0), ("f32", 2.75)), cols, nulls)  # 2.75 -> INT 2
    assert np.array_equal(v, (a < 2).astype(np.uint8)) and np.array_equal(nl, na)
    v, _, nl = O.eval_tree(("logic", 1, ("cmp", 2, ("ref", 2), ("f32", 0.0)), ("cmp", 4, ("ref", 1), ("i32", 0))),
                           cols, nulls)
    assert np.array_equal(v, ((f > 0) | (b == 0)).astype(np.uint8)) and nl is None
    v, _, nl = O.eval_tree(("logic", 0, ("cmp", 2, ("ref", 1), ("i32", 0)), ("cmp", 2, ("ref", 2), ("f32", 0.0))),
                           cols, nulls)
    assert np.array_equal(v, ((b > 0) & (f > 0)).astype(np.uint8)) and np.array_equal(nl, nb)
    v, _, nl = O.eval_tree(("cast", 2, ("arith", 3, ("ref", 0), ("ref", 1))), cols, nulls)
    q = np.where(b == 0, 0, np.where(b == -1, -a, np.trunc(a / np.where(b == 0, 1, b)))).astype(np.int64)
    assert v.dtype == np.int64 and np.array_equal(v, q) and np.array_equal(nl, na)
    v, sc, nl = O.eval_tree(("cmp", 0, ("i32", 1), ("i32", 2)), cols, nulls)
    assert sc and v.tolist() == [1] and nl is None
