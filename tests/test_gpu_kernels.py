"""Parity of every C-ABI kernel against the CPU oracle on seeded inputs.

Bit-exact for integer, index and mask results; float SUM/AVG within 1e-6
relative.  Sizes cover empty, single-row, wave (63/64/65) and tile (4095,
4096, 4097) boundaries and a large case.
"""
import numpy as np
import pytest
import torch

from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

SIZES = [0, 1, 63, 64, 65, 4095, 4096, 4097, 100_003]


@pytest.fixture(scope="module")
def ops(dev):
    from yallasql_amd import ops as _ops

    return _ops


def T(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def N(t):
    return t.cpu().numpy()


def u32(t):
    return N(t).view(np.uint32)


def rng(seed):
    return np.random.default_rng(seed)


# ---------------------------------------------------------------- A1
@pytest.mark.parametrize("dtype", [np.int32, np.float32, np.int64])
@pytest.mark.parametrize("desc", [False, True])
def test_sortkey_encode(ops, dev, dtype, desc):
    r = rng(1)
    if dtype == np.float32:
        x = (r.standard_normal(5000) * 1e3).astype(np.float32)
        x[:6] = [0.0, -0.0, np.inf, -np.inf, 1e-40, -1e-40]
    elif dtype == np.int32:
        x = r.integers(-2**31, 2**31, 5000, dtype=np.int64).astype(np.int32)
        x[:2] = [-2**31, 2**31 - 1]
    else:
        x = r.integers(-2**63, 2**63 - 1, 5000, dtype=np.int64)
    got = N(ops.sortkey_encode(T(x, dev), desc))
    exp = O.sortkey_encode(x, desc)
    assert np.array_equal(got.view(exp.dtype), exp)
    back = N(ops.sortkey_decode(T(got, dev), torch.from_numpy(x[:1]).dtype, desc))
    assert np.array_equal(back.view(x.dtype), x)


# ---------------------------------------------------------------- A2-A6
@pytest.mark.parametrize("n", SIZES + [8191, 8192, 8193, 1 << 20])
@pytest.mark.parametrize("bits", [32, 64])
@pytest.mark.parametrize("desc", [False, True])
def test_radix_argsort(ops, dev, n, bits, desc):
    r = rng(n + bits)
    if bits == 32:
        keys = r.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
        if n > 10:
            keys[: n // 3] = keys[0]  # heavy ties: stability
        kt = T(keys.view(np.int32), dev)
    else:
        keys = r.integers(0, 2**63, n, dtype=np.uint64)
        if n > 10:
            keys[n // 2:] = r.integers(0, 50, n - n // 2)
        kt = T(keys.view(np.int64), dev)
    perm = ops.radix_argsort(kt, descending=desc)
    exp = O.stable_argsort(keys, desc)
    assert np.array_equal(u32(perm), exp)
    assert np.array_equal(N(kt).view(keys.dtype), keys[exp])


@pytest.mark.parametrize("n", [8193, 100_003, (1 << 20) + 7])
@pytest.mark.parametrize("bits", [32, 64])
def test_radix_sort_variants_same_permutation(ops, dev, n, bits):
    """Every digit-pass tile shape returns the stable permutation."""
    from yallasql_amd import _lib as L

    r = rng(n + bits + 1)
    keys = r.integers(0, 2**bits - 1, n, dtype=np.uint64)
    keys[: n // 4] = keys[7]  # ties
    kin = T((keys.astype(np.uint32).view(np.int32) if bits == 32 else keys.view(np.int64)), dev)
    kdt = L.UKEY32 if bits == 32 else L.UKEY64
    exp = O.stable_argsort(keys.astype(np.uint32) if bits == 32 else keys)
    ws = ops.workspace(L.ws_bytes("ysql_radix_sort_ws_bytes", kdt, n), dev)
    for v in range(L.lib.ysql_radix_variant_count()):
        kout, vout = torch.empty_like(kin), torch.empty(n, dtype=torch.int32, device=dev)
        L.call("ysql_radix_sort_variant", kdt, v, 0, ops._p(kin), None, 0, ops._p(kout), ops._p(vout), n, 0, bits,
               ops._p(ws), ws.numel(), ops._s())
        assert np.array_equal(N(vout).view(np.uint32), exp), v
    with pytest.raises(L.YsqlError):
        L.call("ysql_radix_sort_variant", kdt, 99, 0, ops._p(kin), None, 0, ops._p(kout), ops._p(vout), n, 0, bits,
               ops._p(ws), ws.numel(), ops._s())


@pytest.mark.parametrize("n", [5_000, 70_001])  # one-workgroup sort, multi-tile sort
def test_radix_sort_pairs_bits(ops, dev, n):
    r = rng(7)
    keys = r.integers(0, 2**20, n, dtype=np.uint64).astype(np.uint32)
    vals = r.integers(0, 2**31, n, dtype=np.int64).astype(np.int32)
    kt, vt = T(keys.view(np.int32), dev), T(vals, dev)
    ops.radix_sort_pairs(kt, vt, begin_bit=0, end_bit=20)
    exp = O.stable_argsort(keys)
    assert np.array_equal(N(vt), vals[exp])
    # 3 passes (odd) lands back in the caller's buffers
    assert np.array_equal(N(kt).view(np.uint32), keys[exp])


@pytest.mark.parametrize("n", [8_000, 123_457])
@pytest.mark.parametrize("end_bit", [64, 63, 24, 20])
def test_radix_sort_ex_out_of_place(ops, dev, end_bit, n):
    r = rng(end_bit)
    keys = r.integers(0, 2**min(end_bit, 63), n, dtype=np.uint64)
    kin = T(keys.view(np.int64), dev)
    kout = torch.empty_like(kin)
    vout = torch.empty(n, dtype=torch.int32, device=dev)
    ops.radix_sort_ex(kin, kout, vout, val_base=1000, end_bit=end_bit)
    exp = O.stable_argsort(keys)
    assert np.array_equal(N(vout), (exp + 1000).astype(np.int32))
    assert np.array_equal(N(kout).view(np.uint64), keys[exp])
    assert np.array_equal(N(kin).view(np.uint64), keys)  # input untouched
    # in place with the same pass count (odd counts go through the workspace)
    ops.radix_sort_ex(kin, kin, vout, end_bit=end_bit)
    assert np.array_equal(N(vout), exp.astype(np.int32))
    assert np.array_equal(N(kin).view(np.uint64), keys[exp])


def test_radix_all_equal_and_sorted(ops, dev):
    n = 50_000
    for keys in (np.full(n, 7, np.uint32), np.arange(n, dtype=np.uint32), np.arange(n, dtype=np.uint32)[::-1].copy()):
        kt = T(keys.view(np.int32), dev)
        perm = ops.radix_argsort(kt)
        assert np.array_equal(u32(perm), O.stable_argsort(keys))


# ---------------------------------------------------------------- A8-A10
@pytest.mark.parametrize("m,n", [(0, 0), (0, 5), (5, 0), (7, 1), (1, 7), (2048, 2048), (3000, 5000), (100_000, 77_777)])
@pytest.mark.parametrize("desc", [False, True])
def test_merge_path(ops, dev, m, n, desc):
    r = rng(m * 31 + n)
    A = np.sort(r.integers(0, 20, m).astype(np.int32))
    B = np.sort(r.integers(0, 20, n).astype(np.int32))
    if desc:
        A, B = A[::-1].copy(), B[::-1].copy()
    if (m, n) in ((7, 1), (1, 7)):  # the co-rank tie counterexample
        A[:] = 5
        B[:] = 5
    for kmax in (m + n, (m + n) // 2):
        c, idx, isa, cons = ops.merge_path(O.INT, desc, T(A, dev), T(B, dev), k_max=kmax)
        ec, eidx, eisa, econs = O.merge(O.INT, desc, A, B, k_max=kmax)
        assert np.array_equal(N(c), ec)
        assert np.array_equal(u32(idx), eidx)
        assert np.array_equal(N(isa), eisa)
        assert int(cons.item()) == econs


@pytest.mark.parametrize("dtype", [np.float32, np.int64])
def test_merge_path_typed(ops, dev, dtype):
    r = rng(3)
    A = np.sort(r.standard_normal(5000).astype(dtype) if dtype == np.float32 else r.integers(-99, 99, 5000))
    B = np.sort(r.standard_normal(3000).astype(dtype) if dtype == np.float32 else r.integers(-99, 99, 3000))
    dc = O.FLOAT if dtype == np.float32 else O.DATETIME
    c, idx, isa, cons = ops.merge_path(dc, False, T(A, dev), T(B, dev))
    ec, eidx, eisa, econs = O.merge(dc, False, A, B)
    assert np.array_equal(u32(idx), eidx) and np.array_equal(N(isa), eisa) and int(cons.item()) == econs


@pytest.mark.parametrize("m,n", [(1, 0), (4097, 1), (100_000, 123_457)])
def test_merge_pairs(ops, dev, m, n):
    r = rng(m + n)
    ak = np.sort(r.integers(0, 1000, m).astype(np.uint64))
    bk = np.sort(r.integers(0, 1000, n).astype(np.uint64))
    av = np.arange(m, dtype=np.int32)
    bv = np.arange(m, m + n, dtype=np.int32)
    ok, ov = ops.merge_pairs(T(ak.view(np.int64), dev), T(av, dev), T(bk.view(np.int64), dev), T(bv, dev))
    allk = np.concatenate([ak, bk])
    exp = O.stable_argsort(allk)
    assert np.array_equal(N(ok).view(np.uint64), allk[exp])
    assert np.array_equal(N(ov), np.concatenate([av, bv])[exp])


# ---------------------------------------------------------------- A4/A13
@pytest.mark.parametrize("n", SIZES)
def test_scans(ops, dev, n):
    r = rng(n)
    mask = (r.random(n) < 0.4).astype(np.uint8) * r.integers(1, 255, n).astype(np.uint8)
    got = ops.scan_mask_inclusive(T(mask, dev))
    assert np.array_equal(u32(got), O.scan_mask(mask))
    x = r.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    got = ops.scan_u32_inclusive(T(x.view(np.int32), dev))
    assert np.array_equal(u32(got), O.scan_u32(x))


# ---------------------------------------------------------------- A14 / fused filter
@pytest.mark.parametrize("n", SIZES + [1 << 20])
def test_compact_cols(ops, dev, n):
    r = rng(n + 5)
    mask = (r.random(n) < 0.5).astype(np.uint8)
    a = r.integers(-1000, 1000, n).astype(np.int32)
    b = r.standard_normal(n).astype(np.float32)
    c = r.integers(0, 2**62, n, dtype=np.int64)
    nul = (r.random(n) < 0.1).astype(np.uint8)
    outs, cnt = ops.compact_cols(T(mask, dev), [T(a, dev), T(b, dev), T(c, dev)], [T(nul, dev), None, None])
    assert cnt == int(mask.sum())
    for (d, dn), src, sn in zip(outs, [a, b, c], [nul, None, None]):
        ed, edn = O.compact(src, mask, sn)
        assert np.array_equal(N(d), ed)
        assert np.array_equal(N(dn), edn)


@pytest.mark.parametrize("n", [0, 1, 4097, 1 << 20])
@pytest.mark.parametrize("cmp", [0, 1, 2, 3, 4, 5])
def test_filter_compact(ops, dev, n, cmp):
    r = rng(n + cmp)
    price = (1 + 999 * r.random(n)).astype(np.float32)
    ids = r.permutation(n).astype(np.int32)
    thr = np.array([500.0], dtype=np.float32)
    if n > 10:
        price[:5] = 500.0
    outs, cnt = ops.filter_compact(cmp, T(price, dev), T(thr, dev), [T(ids, dev), T(price, dev)])
    mask = O.compare(cmp, price, thr, rhs_scalar=True)
    assert cnt == int(mask.sum())
    assert np.array_equal(N(outs[0][0]), O.compact(ids, mask)[0])
    assert np.array_equal(N(outs[1][0]), O.compact(price, mask)[0])


def test_compact_rejects_in_place(ops, dev):
    """A column whose destination is its source is rejected (tiles would
    overwrite rows other tiles still read)."""
    from yallasql_amd import _lib as L
    n = 1000
    x = torch.arange(n, dtype=torch.int32, device=dev)
    m = torch.ones(n, dtype=torch.uint8, device=dev)
    cnt = torch.zeros(1, dtype=torch.int64, device=dev)
    col = (L.Col * 1)()
    col[0] = L.Col(x.data_ptr(), None, x.data_ptr(), None, 4, 0)
    ws = ops.workspace(L.ws_bytes("ysql_compact_ws_bytes", n), dev)
    rc = L.lib.ysql_compact_cols(ops._p(m), col, 1, n, ops._p(cnt), ops._p(ws), ws.numel(), ops._s())
    assert rc == L.EINVAL
    assert torch.equal(x, torch.arange(n, dtype=torch.int32, device=dev))


def test_compact_strings(ops, dev):
    r = rng(9)
    n = 5000
    s = r.integers(0, 128, (n, 256)).astype(np.uint8)
    mask = (r.random(n) < 0.3).astype(np.uint8)
    outs, cnt = ops.compact_cols(T(mask, dev), [T(s, dev)])
    assert np.array_equal(N(outs[0][0]), s[mask.astype(bool)])


def test_compact_map(ops, dev):
    r = rng(11)
    n = 10_000
    mask = (r.random(n) < 0.5).astype(np.uint8)
    x = r.integers(0, 100, n).astype(np.int32)
    mt = T(mask, dev)
    mp = ops.scan_mask_inclusive(mt)
    cnt = int(u32(mp)[-1])
    d, dn = ops.compact(T(x, dev), mt, mp, cnt)
    assert np.array_equal(N(d), O.compact(x, mask)[0])


# ---------------------------------------------------------------- A12 / A15
@pytest.mark.parametrize("dtype", [np.int32, np.float32, np.int64])
@pytest.mark.parametrize("cmp", range(6))
def test_compare(ops, dev, dtype, cmp):
    r = rng(cmp)
    n = 9999
    a = r.integers(-5, 5, n).astype(dtype)
    b = r.integers(-5, 5, n).astype(dtype)
    s = np.array([1], dtype=dtype)
    for neg in (False, True):
        assert np.array_equal(N(ops.compare(cmp, T(a, dev), T(b, dev), negate=neg)), O.compare(cmp, a, b, negate=neg))
        assert np.array_equal(N(ops.compare(cmp, T(a, dev), T(s, dev), rhs_scalar=True)),
                              O.compare(cmp, a, s, rhs_scalar=True))
        assert np.array_equal(N(ops.compare(cmp, T(s, dev), T(b, dev), lhs_scalar=True)),
                              O.compare(cmp, s, b, lhs_scalar=True))


def test_compare_strings(ops, dev):
    r = rng(2)
    n = 3000
    a = np.zeros((n, 256), np.uint8)
    b = np.zeros((n, 256), np.uint8)
    a[:, :3] = r.integers(97, 100, (n, 3))
    b[:, :3] = r.integers(97, 100, (n, 3))
    for cmp in range(6):
        assert np.array_equal(N(ops.compare(cmp, T(a, dev), T(b, dev))), O.compare(cmp, a, b))


def test_mask_logic(ops, dev):
    r = rng(4)
    a = (r.random(5000) < 0.5).astype(np.uint8)
    b = (r.random(5000) < 0.5).astype(np.uint8)
    for op in (0, 1):
        assert np.array_equal(N(ops.mask_logic(op, T(a, dev), T(b, dev))), O.mask_logic(op, a, b))


@pytest.mark.parametrize("dtype", [np.int32, np.int64, np.float32])
@pytest.mark.parametrize("op", range(5))
def test_arith(ops, dev, dtype, op):
    if dtype == np.float32 and op == 4:
        return
    r = rng(op)
    n = 7777
    if dtype == np.float32:
        a = r.standard_normal(n).astype(dtype)
        b = r.standard_normal(n).astype(dtype)
        b[::7] = 0
    else:
        info = np.iinfo(dtype)
        a = r.integers(info.min, info.max, n, dtype=dtype)
        b = r.integers(-3, 3, n).astype(dtype)
        a[:3] = [info.min, info.min, info.max]
        b[:3] = [-1, 0, -1]
    s = np.array([3], dtype=dtype)
    assert np.array_equal(N(ops.arith(op, T(a, dev), T(b, dev))), O.arith(op, a, b))
    assert np.array_equal(N(ops.arith(op, T(s, dev), T(b, dev), lhs_scalar=True)), O.arith(op, s, b, lhs_scalar=True))


def test_cast(ops, dev):
    r = rng(5)
    f = (r.standard_normal(5000) * 1e10).astype(np.float32)
    f[:4] = [np.nan, np.inf, -np.inf, 3e9]
    i = r.integers(-2**31, 2**31, 5000, dtype=np.int64).astype(np.int32)
    l = r.integers(-2**62, 2**62, 5000, dtype=np.int64)
    for x, dst in [(f, np.int32), (f, np.int64), (i, np.float32), (i, np.int64), (l, np.int32), (l, np.float32)]:
        got = N(ops.cast(T(x, dev), torch.from_numpy(np.zeros(1, dst)).dtype))
        assert np.array_equal(got, O.cast(x, dst)), (x.dtype, dst)


# ---------------------------------------------------------------- gathers / cross
def test_gathers(ops, dev):
    r = rng(6)
    n = 20_000
    src = r.integers(0, 1 << 40, n, dtype=np.int64)
    sn = (r.random(n) < 0.2).astype(np.uint8)
    idx = r.integers(0, n, n).astype(np.uint32)
    d, dn = ops.gather(T(src, dev), T(idx.view(np.int32), dev), T(sn, dev))
    ed, edn = O.gather(src, idx, sn)
    assert np.array_equal(N(d), ed) and np.array_equal(N(dn), edn)
    b = r.integers(0, 1 << 40, n, dtype=np.int64)
    isa = (r.random(n) < 0.5).astype(np.uint8)
    for rev in (False, True):
        d, dn = ops.gather2(T(src, dev), T(b, dev), T(idx.view(np.int32), dev), T(isa, dev), reverse=rev)
        ed, _ = O.gather2(src, b, idx, isa, reverse=rev)
        assert np.array_equal(N(d), ed)
    pairs = r.integers(0, n, (n, 2)).astype(np.uint32)
    for side in (0, 1):
        d, _ = ops.gather_pairs(T(src, dev), T(pairs.view(np.int32), dev), side, n)
        assert np.array_equal(N(d), O.gather_pairs(src, pairs, side)[0])


def test_cross_expand(ops, dev):
    r = rng(8)
    left = r.integers(0, 100, 37).astype(np.int32)
    right = r.integers(0, 100, 256).astype(np.int32)
    nul = (r.random(256) < 0.3).astype(np.uint8)
    d, _ = ops.cross_expand(T(left, dev), 37, 256, True)
    assert np.array_equal(N(d), O.cross_expand(left, 37, 256, True)[0])
    d, dn = ops.cross_expand(T(right, dev), 37, 256, False, src_null=T(nul, dev))
    e, en = O.cross_expand(right, 37, 256, False, src_null=nul)
    assert np.array_equal(N(d), e) and np.array_equal(N(dn), en)
    s = r.integers(0, 128, (5, 256)).astype(np.uint8)
    d, _ = ops.cross_expand(T(s, dev), 5, 3, True)
    assert np.array_equal(N(d), np.repeat(s, 3, axis=0))


# ---------------------------------------------------------------- A16/A17
@pytest.mark.parametrize("n", [1, 65, 4097, 65536, 65537, 1 << 20, 3_000_001])
@pytest.mark.parametrize("dtype", [np.int32, np.float32, np.int64])
def test_reduce(ops, dev, n, dtype):
    r = rng(n)
    if dtype == np.float32:
        x = r.random(n).astype(np.float32)
    else:
        x = r.integers(-10**6, 10**6, n).astype(dtype)
    nul = (r.random(n) < 0.05).astype(np.uint8)
    dc = O._DT[np.dtype(dtype)]
    xt, nt = T(x, dev), T(nul, dev)
    for agg in range(6):
        acc = ops.acc_new(dc, agg, device=dev)
        # two batches folded into the same accumulator
        h = n // 2
        ops.reduce(dc, agg, xt[:h], acc, null=nt[:h])
        ops.reduce(dc, agg, xt[h:], acc, null=nt[h:])
        got = ops.acc_read(acc)[0]
        eacc = O.acc_new(dc, agg)
        O.reduce(dc, agg, x, eacc, null=nul if agg != 5 else None)
        e = eacc[0]
        assert got["count"] == e.count
        if agg in (0, 1):
            if dtype == np.float32:
                assert got["f"] == e.f
            else:
                assert got["i"] == e.i
        if agg in (2, 3):
            if dtype == np.float32:
                assert abs(got["d"] - e.d) <= 1e-6 * max(1.0, abs(e.d))
            else:
                assert got["i"] == e.i
        fin = N(ops.agg_finalize(dc, agg, acc))
        efin = O.agg_finalize(dc, agg, eacc)
        if dtype == np.float32 and agg in (2, 3):
            g32, e32 = fin[:4].view(np.float32)[0], efin[:4].view(np.float32)[0]
            assert abs(g32 - e32) <= 1e-6 * max(1.0, abs(e32))
        elif agg == 3:
            assert abs(fin[:4].view(np.float32)[0] - efin[:4].view(np.float32)[0]) <= 1e-6 * abs(efin[:4].view(np.float32)[0]) + 1e-30
        else:
            assert np.array_equal(fin, efin)


def test_reduce_deterministic(ops, dev):
    r = rng(12)
    x = T(r.random(5_000_000).astype(np.float32), dev)
    outs = set()
    for _ in range(3):
        acc = ops.acc_new(O.FLOAT, 2, device=dev)
        ops.reduce(O.FLOAT, 2, x, acc)
        outs.add(ops.acc_read(acc)[0]["d"])
    assert len(outs) == 1


@pytest.mark.parametrize("n", [0, 1, 4095, 4096, 4097, 300_001])
@pytest.mark.parametrize("pdt", [np.int32, np.float32, np.int64])
def test_filter_compact_range(ops, dev, n, pdt):
    """lo <= pred < hi with a capacity smaller than the count (the wrapper
    retries with the reported count), against the oracle's two comparisons
    ANDed and its compaction."""
    r = rng(n + 3)
    if pdt == np.float32:
        pred = r.random(n).astype(np.float32)
        lo, hi = 0.25, 0.5
    else:
        pred = r.integers(-1000, 1000, n).astype(pdt)
        lo, hi = -100, 300
    pay = np.arange(n, dtype=np.int32)
    m1 = O.compare(3, pred, np.array([lo], dtype=pdt), rhs_scalar=True, n=n)  # GE
    m2 = O.compare(0, pred, np.array([hi], dtype=pdt), rhs_scalar=True, n=n)  # LT
    m = O.mask_logic(0, m1, m2, n=n)
    exp_k, _ = O.compact(pred, m)
    exp_p, _ = O.compact(pay, m)
    for cap in (None, 7):
        (k, p), c = ops.filter_compact_range(T(pred, dev), lo, hi, [T(pred, dev), T(pay, dev)], cap=cap)
        assert c == len(exp_k)
        assert np.array_equal(N(k), exp_k) and np.array_equal(N(p), exp_p)


FA_SIZES = [0, 1, 7, 8, 9, 32767, 32768, 32769, 1_000_003]


@pytest.mark.parametrize("n", FA_SIZES)
@pytest.mark.parametrize("pdt", [np.int32, np.float32, np.int64])
def test_filter_aggregate(ops, dev, n, pdt):
    """WHERE pred <cmp> c with SUM(float), MIN(int32), MAX(int64), COUNT,
    SUM(int32) in one pass; against the oracle's compare + reduce."""
    r = rng(n + 7)
    if pdt == np.float32:
        pred = r.random(n + 1).astype(np.float32)
        c = np.float32(0.37)
    else:
        pred = r.integers(-1000, 1000, n + 1).astype(pdt)
        c = pdt(-250)
    f = r.random(n + 1).astype(np.float32)
    i32 = r.integers(-10**9, 10**9, n + 1).astype(np.int32)
    i64 = r.integers(-2**62, 2**62, n + 1).astype(np.int64)
    full = [T(x, dev) for x in (pred, f, i32, i64)]
    # device views at an odd offset (elements 1..n) take the unaligned path
    for off in (0, 1):
        P, F, I, J = (x[off:off + n] for x in (pred, f, i32, i64))
        tp, tf, ti, tj = (x[off:off + n] for x in full)
        for cmp in range(6):
            m = O.compare(cmp, P, np.array([c], dtype=pdt), rhs_scalar=True, n=n).astype(bool)
            specs = [(O.FLOAT, 2, tf), (O.INT, 1, ti), (O.DATETIME, 0, tj), (O.INT, 4, None), (O.INT, 2, ti)]
            pdc = O._DT[np.dtype(pdt)]
            # specs over the predicate column and two specs over one column
            # (the kernel loads such a column once)
            shared = [(pdc, 2, tp), (pdc, 1, tp), (O.DATETIME, 0, tj), (O.DATETIME, 1, tj)]
            # COUNT specs only: the predicate-only loop
            counts = [(O.INT, 5, None), (O.INT, 4, None)]
            for group in (specs[:4], specs[4:], shared, counts):
                accs = [ops.acc_new(dc, agg, device=dev) for dc, agg, _ in group]
                ops.filter_aggregate(tp, cmp, c.item(), [(dc, agg, col, a) for (dc, agg, col), a in zip(group, accs)],
                                     overwrite=(off == 1))
                src = {id(tf): F, id(ti): I, id(tj): J, id(tp): P}
                for (dc, agg, col), a in zip(group, accs):
                    got = ops.acc_read(a)[0]
                    e = O.acc_new(dc, agg)
                    x = src[id(col)][m] if col is not None else np.zeros(int(m.sum()), np.int32)
                    O.reduce(dc, agg, x, e)
                    e = e[0]
                    assert got["count"] == e.count, (cmp, dc, agg)
                    if dc == O.FLOAT and agg == 1:
                        assert got["f"] == e.f, (cmp, agg)
                    elif dc == O.FLOAT and agg == 2:
                        assert abs(got["d"] - e.d) <= 1e-6 * max(1.0, abs(e.d)), (cmp, agg)
                    elif agg != 4:
                        assert got["i"] == e.i, (cmp, dc, agg)


@pytest.mark.parametrize("n", [0, 9, 32769, 2_000_003])
@pytest.mark.parametrize("pdt", [np.int32, np.float32, np.int64])
def test_filter_aggregate_shared_scan_same_bits(ops, dev, n, pdt):
    """Each query of a shared scan returns the bytes it returns alone."""
    r = rng(n + 11)
    if pdt == np.float32:
        pred = r.random(n).astype(np.float32)
        consts = [float(c) for c in r.random(4)]
    else:
        pred = r.integers(-1000, 1000, n).astype(pdt)
        consts = [int(c) for c in r.integers(-1000, 1000, 4)]
    f, i64 = T(r.random(n).astype(np.float32), dev), T(r.integers(-2**62, 2**62, n).astype(np.int64), dev)
    tp = T(pred, dev)
    cols = [(O.FLOAT, 2, f), (O.DATETIME, 1, i64), (O.DATETIME, 0, i64), (O.INT, 5, None)]
    for cmp in (0, 3, 5):
        for nq in (1, 2, 3, 4):
            accs = [[ops.acc_new(dc, agg, device=dev) for dc, agg, _ in cols] for _ in range(nq)]
            ops.filter_aggregate_multi(tp, cmp, consts[:nq], cols, accs)
            for q in range(nq):
                alone = [ops.acc_new(dc, agg, device=dev) for dc, agg, _ in cols]
                ops.filter_aggregate(tp, cmp, consts[q], [(dc, agg, col, a) for (dc, agg, col), a in zip(cols, alone)],
                                     overwrite=True)
                for a, b in zip(accs[q], alone):
                    assert np.array_equal(N(a), N(b)), (cmp, nq, q)


@pytest.mark.parametrize("pdt", [np.int32, np.float32, np.int64])
def test_filter_aggregate_two_column_kernel_equals_generic(ops, dev, pdt):
    """The two-column kernel (a comparison as a range test, static fold
    bodies) returns the bytes of the general tile kernel for every operator
    and the edge constants: NaN, +-inf, +-0, the smallest subnormals,
    INT_MIN / INT_MAX; with predicate values holding the same edges; one
    query and shared scans of four."""
    r = rng(5)
    n = 70_001
    if pdt == np.float32:
        pred = (r.standard_normal(n) * 3).astype(np.float32)
        edges = np.array([np.nan, np.inf, -np.inf, 0.0, -0.0, 1.4e-45, -1.4e-45, 3.0, -3.0], np.float32)
        pred[::97] = edges[r.integers(0, len(edges), len(pred[::97]))]
        consts = [float(x) for x in edges] + [0.5]
    else:
        info = np.iinfo(pdt)
        pred = r.integers(-5, 6, n).astype(pdt)
        pred[::101] = info.min
        pred[::103] = info.max
        consts = [int(info.min), int(info.max), int(info.min) + 1, int(info.max) - 1, 0, 3, -2]
    tp = T(pred, dev)
    other = T(r.integers(-2**40, 2**40, n).astype(np.int64), dev)
    i32 = T(r.integers(-1000, 1000, n).astype(np.int32), dev)
    pdc = O._DT[np.dtype(pdt)]
    shapes = [
        [(O.INT, 5, None), (pdc, 2, tp), (O.DATETIME, 1, other), (O.DATETIME, 0, other)],
        [(O.INT, 4, None), (O.INT, 2, i32), (pdc, 1, tp), (pdc, 0, tp)],
        [(O.INT, 5, None)],
    ]
    for cmp in range(6):
        for c in consts:
            for cols in shapes:
                got, ref = ([ops.acc_new(dc, agg, device=dev) for dc, agg, _ in cols] for _ in range(2))
                ops.filter_aggregate(tp, cmp, c, [(dc, agg, col, a) for (dc, agg, col), a in zip(cols, got)],
                                     overwrite=True)
                ops.filter_aggregate(tp, cmp, c, [(dc, agg, col, a) for (dc, agg, col), a in zip(cols, ref)],
                                     overwrite=True, generic=True)
                for a, b in zip(got, ref):
                    assert np.array_equal(N(a), N(b)), (cmp, c, cols)
                m = O.compare(cmp, pred, np.array([c], dtype=pdt), rhs_scalar=True, n=n).astype(bool)
                assert ops.acc_read(got[0])[0]["count"] == int(m.sum()), (cmp, c)
        cols = shapes[0]
        accs = [[ops.acc_new(dc, agg, device=dev) for dc, agg, _ in cols] for _ in range(4)]
        gacc = [[ops.acc_new(dc, agg, device=dev) for dc, agg, _ in cols] for _ in range(4)]
        ops.filter_aggregate_multi(tp, cmp, consts[:4], cols, accs)
        ops.filter_aggregate_multi(tp, cmp, consts[:4], cols, gacc, generic=True)
        for qa, qb in zip(accs, gacc):
            for a, b in zip(qa, qb):
                assert np.array_equal(N(a), N(b)), (cmp, "shared")


def test_filter_aggregate_fold_vs_overwrite(ops, dev):
    r = rng(3)
    x = r.random(100_000).astype(np.float32)
    tx = T(x, dev)
    acc = ops.acc_new(O.FLOAT, 2, device=dev)
    ops.filter_aggregate(tx, 0, 0.5, [(O.FLOAT, 2, tx, acc)])
    ops.filter_aggregate(tx, 0, 0.5, [(O.FLOAT, 2, tx, acc)])  # folds: twice the sum
    twice = ops.acc_read(acc)[0]
    ops.filter_aggregate(tx, 0, 0.5, [(O.FLOAT, 2, tx, acc)], overwrite=True)
    once = ops.acc_read(acc)[0]
    assert twice["count"] == 2 * once["count"] == 2 * int((x < 0.5).sum())
    assert abs(twice["d"] - 2 * once["d"]) <= 1e-9 * abs(twice["d"])
    # empty input with overwrite stores the identity
    ops.filter_aggregate(tx[:0], 0, 0.5, [(O.FLOAT, 2, tx, acc)], overwrite=True)
    assert ops.acc_read(acc)[0]["count"] == 0 and ops.acc_read(acc)[0]["d"] == 0.0


@pytest.mark.parametrize("ng", [7, 4096, 7000])
def test_group_reduce(ops, dev, ng):
    r = rng(ng)
    n = 300_000
    gid = r.integers(0, ng + 3, n).astype(np.uint32)  # some out-of-range ids skip
    xi = r.integers(-10**6, 10**6, n).astype(np.int64)
    for agg in (0, 1, 2, 4, 5):
        acc = ops.acc_new(O.DATETIME, agg, ng, device=dev)
        ops.group_reduce(O.DATETIME, agg, T(gid.view(np.int32), dev), ng, T(xi, dev), acc)
        got = ops.acc_read(acc)
        eacc = O.acc_new(O.DATETIME, agg, ng)
        O.group_reduce(O.DATETIME, agg, gid, ng, xi, eacc)
        for g in range(ng):
            assert got[g]["count"] == eacc[g].count
            if agg in (0, 1, 2):
                assert got[g]["i"] == eacc[g].i


# ---------------------------------------------------------------- joins
@pytest.mark.parametrize("dtype", [np.int32, np.float32, np.int64])
@pytest.mark.parametrize("cmp", range(6))
def test_join_pairs(ops, dev, dtype, cmp):
    r = rng(cmp + 10)
    L_, R_ = 1500, 700
    left = r.integers(0, 300, L_).astype(dtype)
    right = r.integers(0, 300, R_).astype(dtype)
    pairs, c = ops.join_pairs(cmp, T(left, dev), T(right, dev))
    exp = O.join_pairs(cmp, left, right)
    assert c == exp.shape[0]
    assert np.array_equal(N(pairs).view(np.uint32), exp)


@pytest.mark.parametrize("R_", [2000, 5000])
@pytest.mark.parametrize("L_", [4095, 4096, 4097, 300_001])
@pytest.mark.parametrize("dtype", [np.int32, np.float32, np.int64])
def test_join_eq_hash_probe(ops, dev, L_, dtype, R_):
    """EQ hash join at probe-tile boundaries, on an unaligned left view,
    with duplicate right keys, and (float) NaN and signed zeros; R_ = 2000
    probes a table staged in LDS, R_ = 5000 the table in global memory."""
    r = rng(L_ + 5)
    left = r.integers(0, 5000, L_ + 1).astype(dtype)
    right = r.integers(0, 5000, R_).astype(dtype)
    if dtype == np.float32:
        left[::97] = np.nan
        right[::53] = np.nan
        left[1::89] = -0.0
        right[5] = 0.0
    lt = T(left, dev)[1:]  # unaligned device view of rows 1..L_
    pairs, c = ops.join_pairs(4, lt, T(right, dev), cap=1024)
    exp = O.join_pairs(4, left[1:], right)
    assert c == exp.shape[0]
    assert np.array_equal(N(pairs).view(np.uint32), exp)


@pytest.mark.parametrize("dtype", [np.int32, np.float32, np.int64])
def test_join_eq_large_filter(ops, dev, dtype):
    """EQ hash join whose right side (20,000 rows) takes the smallest
    blocked filter in global memory (2^20 bits): left keys hit, miss and
    land in long probe chains."""
    r = rng(77)
    L_, R_ = 32_769, 20_000
    right = r.integers(0, 60_000, R_).astype(dtype)
    right[:500] = right[500]  # one key with 501 right rows
    left = r.integers(0, 120_000, L_ + 1).astype(dtype)
    left[::1000] = right[500]
    if dtype == np.float32:
        left[::31] = np.nan
        right[::17] = np.nan
    exp = O.join_pairs(4, left[1:], right)
    for lt in (T(left, dev)[1:], T(left[1:], dev)):  # unaligned view, aligned copy
        pairs, c = ops.join_pairs(4, lt, T(right, dev))
        assert c == exp.shape[0] > 10_000
        assert np.array_equal(N(pairs).view(np.uint32), exp)


@pytest.mark.parametrize("dtype", [np.int32, np.float32])
def test_join_eq_blocked_filter(ops, dev, dtype):
    """EQ hash join with 70,000 right rows: a blocked filter in global
    memory of 16 bits per right row (2^21 bits)."""
    r = rng(78)
    L_, R_ = 8_193, 70_000
    right = r.integers(0, 200_000, R_).astype(dtype)
    right[:300] = right[300]
    left = r.integers(0, 400_000, L_ + 1).astype(dtype)
    left[::100] = right[300]
    if dtype == np.float32:
        left[::37] = np.nan
        right[::19] = np.nan
    exp = O.join_pairs(4, left[1:], right)
    for lt in (T(left, dev)[1:], T(left[1:], dev)):  # unaligned view, aligned copy
        pairs, c = ops.join_pairs(4, lt, T(right, dev))
        assert c == exp.shape[0] > 10_000
        assert np.array_equal(N(pairs).view(np.uint32), exp)


@pytest.mark.parametrize("R_", [2_000, 20_000])
def test_join_eq_many_steps(ops, dev, R_):
    """EQ join of 5,000,001 left rows (1,221 probe steps: the step-offset
    scan spans two blocks, the last step is partial).  Expected pairs from a
    stable sort of the right keys and a binary search per left key: an
    independent computation whose order is (l, r) by construction."""
    r = rng(R_ + 9)
    L_ = 5_000_001
    left = r.integers(0, 100_000, L_).astype(np.int64)
    right = r.integers(0, 100_000, R_).astype(np.int64)
    right[:40] = right[40]
    order = np.argsort(right, kind="stable")
    rs = right[order]
    lo = np.searchsorted(rs, left, "left")
    cnt = np.searchsorted(rs, left, "right") - lo
    total = int(cnt.sum())
    first = np.cumsum(cnt) - cnt  # output position of each left row's first pair
    exp = np.empty((total, 2), np.uint32)
    exp[:, 0] = np.repeat(np.arange(L_, dtype=np.uint32), cnt)
    exp[:, 1] = order[np.repeat(lo - first, cnt) + np.arange(total)]
    lt, rt = T(left, dev), T(right, dev)
    pairs, c = ops.join_pairs(4, lt, rt)
    assert c == total > 50_000
    assert np.array_equal(N(pairs).view(np.uint32), exp)
    # capacity of half the pairs: the exact count, and exactly the first half
    from yallasql_amd import _lib as L

    cap = total // 2
    ws = ops.workspace(L.ws_bytes("ysql_join_ws_bytes", L.DATETIME, L.EQ, L_, R_), dev)
    out = torch.full((cap + 64, 2), -1, dtype=torch.int32, device=dev)
    cnt = torch.zeros(1, dtype=torch.int64, device=dev)
    L.call("ysql_join_pairs", L.DATETIME, L.EQ, ops._p(lt), L_, ops._p(rt), R_, ops._p(out), cap, ops._p(cnt),
           ops._p(ws), ws.numel(), ops._s())
    assert int(cnt.item()) == total
    got = N(out).view(np.uint32)
    assert np.array_equal(got[:cap], exp[:cap])
    assert (N(out)[cap:] == -1).all()  # nothing written past the capacity


@pytest.mark.parametrize("R_", [2_000, 20_000])
def test_join_eq_dense_hits_int32(ops, dev, R_):
    """EQ join where almost every left row matches many right rows, with
    multiplicities that differ from row to row (1 to ~60 right rows per key,
    one key with none), int32 keys: every lane of a wave writes several
    pairs, so the write pass's lane order and output offsets are exercised
    on every wave.  R_ = 2,000: table in LDS; 20,000: table and blocked
    filter in global memory.  Expected pairs from a stable sort of the right
    keys and a binary search per left key, in (l, r) order by construction."""
    r = rng(R_ + 21)
    L_ = 100_003 if R_ == 2_000 else 20_003  # about 3M and 6M pairs
    keys = np.arange(64, dtype=np.int32) * 7919 - 250_000
    w = np.arange(1, 65, dtype=np.float64)
    right = r.choice(keys[1:], R_, p=w[1:] / w[1:].sum()).astype(np.int32)  # keys[0] has no right row
    left = r.choice(keys, L_ + 1).astype(np.int32)
    left[::101] = 12345  # some rows match nothing
    order = np.argsort(right, kind="stable")
    rs = right[order]
    lk = left[1:]
    lo = np.searchsorted(rs, lk, "left")
    cnt = np.searchsorted(rs, lk, "right") - lo
    total = int(cnt.sum())
    first = np.cumsum(cnt) - cnt
    exp = np.empty((total, 2), np.uint32)
    exp[:, 0] = np.repeat(np.arange(L_, dtype=np.uint32), cnt)
    exp[:, 1] = order[np.repeat(lo - first, cnt) + np.arange(total)]
    for lt in (T(left, dev)[1:], T(lk, dev)):  # unaligned view, aligned copy
        pairs, c = ops.join_pairs(4, lt, T(right, dev))
        assert c == total > 1_000_000
        assert np.array_equal(N(pairs).view(np.uint32), exp)


def test_join_overflow_reports_count(ops, dev):
    left = np.zeros(100, np.int32)
    right = np.zeros(50, np.int32)
    pairs, c = ops.join_pairs(4, T(left, dev), T(right, dev), cap=10)
    assert c == 5000 and pairs.shape[0] == 5000  # retried with the exact capacity


def test_pairs_intersect(ops, dev):
    r = rng(13)
    p1 = r.integers(0, 40, (3000, 2)).astype(np.uint32)
    p2 = r.integers(0, 40, (2000, 2)).astype(np.uint32)
    got, c = ops.pairs_intersect(T(p1.view(np.int32), dev), T(p2.view(np.int32), dev))
    exp = O.pairs_intersect(p1, p2)
    assert np.array_equal(N(got).view(np.uint32), exp)


@pytest.mark.parametrize("n,off,filters", [(2_000_000, 0, True), (1_000_003, 1, True), (1_000_003, 1, False)])
def test_star_join_group_sum(ops, dev, n, off, filters):
    """off = 1: every fact column is an unaligned device view with a ragged tail.  With filters the
    unfiltered dimensions' keys and the measures are read only for rows that pass them; some of those
    keys are out of their dimension's range (the row joins nothing).  Without filters every column is
    read up front."""
    r = rng(14)
    dims_np = []
    rows = [2556, 2000, 30000, 200000]
    fks = [r.integers(0, R, n).astype(np.int32) for R in rows]
    year = (1992 + (7 * np.arange(2556)) // 2556).astype(np.int32)
    region = r.integers(0, 7, 30000).astype(np.int32)
    cat = r.integers(0, 25, 200000).astype(np.int32)
    brand = r.integers(0, 1000, 200000).astype(np.int32)
    bad = r.random(n) < 0.01
    fks[1][bad] = np.where(r.random(int(bad.sum())) < 0.5, -1, 2000 + 7).astype(np.int32)
    dims_np = [
        {"fk": fks[0], "rows": 2556, "group": year, "group_base": 1992, "group_card": 7},
        {"fk": fks[1], "rows": 2000},
        {"fk": fks[2], "rows": 30000, "pass": (region == 0).astype(np.uint8)},
        {"fk": fks[3], "rows": 200000, "pass": (cat == 0).astype(np.uint8), "group": brand, "group_card": 1000},
    ]
    if not filters:
        for d in dims_np:
            d.pop("pass", None)
    m1 = r.integers(90000, 10000001, n).astype(np.int32)
    m2 = r.integers(0, 11, n).astype(np.int32)
    def fact(v):  # device column, shifted by `off` elements when requested
        if not off:
            return T(v, dev)
        return T(np.concatenate([np.zeros(off, v.dtype), v]), dev)[off:]

    dims_t = []
    for d in dims_np:
        dt_ = {k: ((fact(v) if k == "fk" else T(v, dev)) if isinstance(v, np.ndarray) else v) for k, v in d.items()}
        dims_t.append(dt_)
    sums, counts = ops.star_join_group_sum(dims_t, fact(m1), fact(m2), 7000)
    es, ec = O.star_join_group_sum(dims_np, m1, m2, 7000)
    assert np.array_equal(N(sums), es)
    assert np.array_equal(N(counts).astype(np.uint64), ec)


@pytest.mark.parametrize("n,off", [(2_000_000, 0), (1_000_003, 1)])
def test_star_join_group_sum_multi(ops, dev, n, off):
    """Three star-join queries in one read of the facts (bit q of each pass byte is query q's filter):
    each query's sums and counts equal the oracle's for that query's filters alone.  off = 1: unaligned
    fact columns with a ragged tail; 1% of the supplier keys are out of range."""
    r = rng(15)
    rows = [2556, 2000, 30000, 200000]
    fks = [r.integers(0, R, n).astype(np.int32) for R in rows]
    year = (1992 + (7 * np.arange(2556)) // 2556).astype(np.int32)
    region = r.integers(0, 7, 30000).astype(np.int32)
    cat = r.integers(0, 25, 200000).astype(np.int32)
    brand = r.integers(0, 1000, 200000).astype(np.int32)
    bad = r.random(n) < 0.01
    fks[1][bad] = np.where(r.random(int(bad.sum())) < 0.5, -1, 2000 + 7).astype(np.int32)
    m1 = r.integers(90000, 10000001, n).astype(np.int32)
    m2 = r.integers(0, 11, n).astype(np.int32)
    params = [(0, 0), (3, 7), (6, 24)]
    rp = np.zeros(30000, np.uint8)
    cp = np.zeros(200000, np.uint8)
    for q, (rv, cv) in enumerate(params):
        rp |= ((region == rv).astype(np.uint8) << q).astype(np.uint8)
        cp |= ((cat == cv).astype(np.uint8) << q).astype(np.uint8)

    def fact(v):
        if not off:
            return T(v, dev)
        return T(np.concatenate([np.zeros(off, v.dtype), v]), dev)[off:]

    dims_t = [
        {"fk": fact(fks[0]), "rows": 2556, "group": T(year, dev), "group_base": 1992, "group_card": 7},
        {"fk": fact(fks[1]), "rows": 2000},
        {"fk": fact(fks[2]), "rows": 30000, "pass": T(rp, dev)},
        {"fk": fact(fks[3]), "rows": 200000, "pass": T(cp, dev), "group": T(brand, dev), "group_card": 1000},
    ]
    sums, counts = ops.star_join_group_sum_multi(dims_t, fact(m1), fact(m2), 7000, len(params))
    for q, (rv, cv) in enumerate(params):
        dims_np = [
            {"fk": fks[0], "rows": 2556, "group": year, "group_base": 1992, "group_card": 7},
            {"fk": fks[1], "rows": 2000},
            {"fk": fks[2], "rows": 30000, "pass": (region == rv).astype(np.uint8)},
            {"fk": fks[3], "rows": 200000, "pass": (cat == cv).astype(np.uint8), "group": brand,
             "group_card": 1000},
        ]
        es, ec = O.star_join_group_sum(dims_np, m1, m2, 7000)
        assert np.array_equal(N(sums[q]), es), q
        assert np.array_equal(N(counts[q]).astype(np.uint64), ec), q
