"""include/ysql_kernels.hpp, the reference-shaped launchers, executed: a C++
translation unit (tests/compat/compat_launchers.cpp, built by build())
calls every launcher the way the reference's operators call theirs, and
each result is compared with the oracle."""
import ctypes as C
import math
import os

import numpy as np
import pytest

from oracle import pyoracle as O
from tests.conftest import ROOT

pytestmark = pytest.mark.gpu

LIB = os.path.join(ROOT, "tests", "compat", "libcompat_launchers.so")
P = C.c_void_p


@pytest.fixture(scope="module")
def cl(dev):
    from yallasql_amd import _lib  # noqa: F401  (the kernel library, loaded first)

    if not os.path.exists(LIB):
        pytest.fail(f"{LIB} is not built (run __graft_entry__.build())")
    lib = C.CDLL(LIB)
    for f in ("compat_sortkey", "compat_radix_sort", "compat_prefix_sum", "compat_merge", "compat_move",
              "compat_conditional", "compat_binary", "compat_outer_join", "compat_and_join", "compat_cross",
              "compat_reduce", "compat_finalize", "compat_move_batch", "compat_sort_strings", "compat_merge_strings"):
        getattr(lib, f).restype = C.c_int
    lib.compat_outer_join.argtypes = [C.c_int, P, P, P, P, C.c_uint32, C.c_uint32, C.c_uint64, P]
    return lib


def _t(a, dev):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _p(t):
    return None if t is None else P(t.data_ptr())


def _s():
    import torch

    return P(torch.cuda.current_stream().cuda_stream)


def _h(t):
    import torch

    torch.cuda.synchronize()
    return t.cpu().numpy()


def test_sortkeys_and_radix_sort(cl, dev):
    import torch

    r = np.random.default_rng(1)
    n = 70_001
    ints = r.integers(-2**31, 2**31, n, dtype=np.int64).astype(np.int32)
    ints[:4] = [-2**31, 2**31 - 1, 0, -1]
    fl = r.standard_normal(n).astype(np.float32)
    fl[:4] = [np.inf, -np.inf, -0.0, 0.0]
    for which, x in ((0, ints), (1, fl)):
        keys = torch.empty(n, dtype=torch.int32, device=dev)
        assert cl.compat_sortkey(which, _p(_t(x, dev)), _p(keys), n, _s()) == 0
        exp = O.sortkey_encode(x)
        assert np.array_equal(_h(keys).view(np.uint32), exp)
        back = torch.empty(n, dtype=torch.int32 if which == 0 else torch.float32, device=dev)
        assert cl.compat_sortkey(which + 2, _p(keys), _p(back), n, _s()) == 0
        assert np.array_equal(_h(back).view(np.uint32), x.view(np.uint32))
    for key64 in (0, 1):
        for desc in (0, 1):
            k = r.integers(0, 1000, n, dtype=np.uint64 if key64 else np.uint32)  # heavy ties
            kd = _t(k.view(np.int64 if key64 else np.int32), dev)
            idx = torch.empty(n, dtype=torch.int32, device=dev)
            assert cl.compat_radix_sort(key64, desc, _p(kd), n, _p(idx), _s()) == 0
            exp = O.stable_argsort(k, desc=bool(desc))
            assert np.array_equal(_h(idx).view(np.uint32), exp)
            assert np.array_equal(_h(kd).view(k.dtype), k[exp])


def test_prefix_sums(cl, dev):
    import torch

    r = np.random.default_rng(2)
    n = 100_003
    m = (r.random(n) < 0.3).astype(np.uint8)
    out = torch.empty(n, dtype=torch.int32, device=dev)
    assert cl.compat_prefix_sum(1, _p(_t(m, dev)), _p(out), n, _s()) == 0
    assert np.array_equal(_h(out).view(np.uint32), O.scan_mask(m))
    v = r.integers(0, 1000, n, dtype=np.uint32)
    assert cl.compat_prefix_sum(0, _p(_t(v.view(np.int32), dev)), _p(out), n, _s()) == 0
    assert np.array_equal(_h(out).view(np.uint32), O.scan_u32(v))


@pytest.mark.parametrize("dtype,desc", [(0, 0), (1, 1), (2, 0)])
def test_merge(cl, dev, dtype, desc):
    import torch

    r = np.random.default_rng(3)
    npd = [np.int32, np.float32, np.int64][dtype]
    A = np.sort(r.integers(0, 50, 3001).astype(npd))
    B = np.sort(r.integers(0, 50, 2500).astype(npd))
    if desc:
        A, B = A[::-1].copy(), B[::-1].copy()
    k = 4000
    Cd = torch.empty(k, dtype=_t(A, dev).dtype, device=dev)
    idx = torch.empty(k, dtype=torch.int32, device=dev)
    isa = torch.empty(k, dtype=torch.uint8, device=dev)
    lasti = torch.zeros(1, dtype=torch.int32, device=dev)
    assert cl.compat_merge(dtype, desc, _p(_t(A, dev)), _p(_t(B, dev)), _p(Cd), _p(idx), _p(isa), _p(lasti), k,
                           A.size, B.size, _s()) == 0
    ec, ei, ea, cons = O.merge([O.INT, O.FLOAT, O.DATETIME][dtype], desc, A, B, k)
    assert np.array_equal(_h(Cd), ec) and np.array_equal(_h(idx).view(np.uint32), ei)
    assert np.array_equal(_h(isa), ea) and int(_h(lasti)[0]) == cons


def test_row_movement(cl, dev):
    import torch

    r = np.random.default_rng(4)
    n = 50_000
    src = r.integers(-10**9, 10**9, n).astype(np.int32)
    nul = (r.random(n) < 0.1).astype(np.uint8)
    src_d, nul_d = _t(src, dev), _t(nul, dev)
    res = torch.empty(n, dtype=torch.int32, device=dev)
    rn = torch.empty(n, dtype=torch.uint8, device=dev)
    # gather
    perm = r.permutation(n).astype(np.uint32)
    assert cl.compat_move(0, _p(src_d), None, _p(res), _p(nul_d), None, _p(rn), _p(_t(perm.view(np.int32), dev)),
                          None, 0, n, _s()) == 0
    eg, egn = O.gather(src, perm, nul)
    assert np.array_equal(_h(res), eg) and np.array_equal(_h(rn), egn)
    # merge gather, both directions
    src2 = r.integers(0, 100, n).astype(np.int32)
    nul2 = np.zeros(n, np.uint8)
    isa = (r.random(n) < 0.5).astype(np.uint8)
    for asc in (1, 0):
        assert cl.compat_move(1, _p(src_d), _p(_t(src2, dev)), _p(res), _p(nul_d), _p(_t(nul2, dev)), _p(rn),
                              _p(_t(perm.view(np.int32), dev)), _p(_t(isa, dev)), asc, n, _s()) == 0
        e2, e2n = O.gather2(src, src2, perm, isa, reverse=not asc, a_null=nul, b_null=nul2)
        assert np.array_equal(_h(res), e2) and np.array_equal(_h(rn), e2n)
    # filter compaction from a mask and its inclusive map
    mask = (r.random(n) < 0.4).astype(np.uint8)
    mp = O.scan_mask(mask)
    res.fill_(-1)
    assert cl.compat_move(2, _p(src_d), None, _p(res), _p(nul_d), None, _p(rn), _p(_t(mp.view(np.int32), dev)),
                          _p(_t(mask, dev)), 0, n, _s()) == 0
    ec, ecn = O.compact(src, mask, nul)
    m = ec.size
    assert np.array_equal(_h(res)[:m], ec) and np.array_equal(_h(rn)[:m], ecn)
    # join gather, left and right sides
    pairs = np.stack([r.integers(0, n, 7000), r.integers(0, n, 7000)], axis=1).astype(np.uint32)
    for side in (0, 1):
        assert cl.compat_move(3, _p(src_d), None, _p(res), _p(nul_d), None, _p(rn),
                              _p(_t(pairs.view(np.int32), dev)), None, side, 7000, _s()) == 0
        ej, ejn = O.gather_pairs(src, pairs, side, nul)
        assert np.array_equal(_h(res)[:7000], ej) and np.array_equal(_h(rn)[:7000], ejn)


def test_conditions_arith_cast(cl, dev):
    import torch

    r = np.random.default_rng(5)
    n = 30_000
    out = torch.empty(n, dtype=torch.uint8, device=dev)
    for dtype, npd in ((0, np.int32), (1, np.float32), (2, np.int64)):
        lhs = r.integers(-50, 50, n).astype(npd)
        rhs = r.integers(-50, 50, n).astype(npd)
        for op in range(6):
            for neg in (0, 1):
                # launch_conditional_operators(d_rhs, d_lhs, ...) computes op(lhs, rhs)
                assert cl.compat_conditional(dtype, op, _p(_t(rhs, dev)), _p(_t(lhs, dev)), 0, 0, _p(out), n, neg,
                                             _s()) == 0
                assert np.array_equal(_h(out), O.compare(op, lhs, rhs, negate=bool(neg)))
        sc = np.array([7], dtype=npd)
        assert cl.compat_conditional(dtype, 2, _p(_t(sc, dev)), _p(_t(lhs, dev)), 1, 0, _p(out), n, 0, _s()) == 0
        assert np.array_equal(_h(out), O.compare(2, lhs, sc, rhs_scalar=True))
    a = (r.random(n) < 0.5).astype(np.uint8)
    b = (r.random(n) < 0.5).astype(np.uint8)
    for op in (0, 1):
        assert cl.compat_conditional(3, op, _p(_t(b, dev)), _p(_t(a, dev)), 0, 0, _p(out), n, 0, _s()) == 0
        assert np.array_equal(_h(out), O.mask_logic(op, a, b))
    lhs = r.integers(-2**31, 2**31, n, dtype=np.int64).astype(np.int32)
    rhs = r.integers(-5, 5, n).astype(np.int32)
    lhs[:2], rhs[:2] = [-2**31, 5], [-1, 0]
    res = torch.empty(n, dtype=torch.int32, device=dev)
    resf = torch.empty(n, dtype=torch.float32, device=dev)
    for op in range(5):
        assert cl.compat_binary(op, _p(_t(rhs, dev)), _p(_t(lhs, dev)), 0, 0, _p(res), _p(resf), n, _s()) == 0
        e = O.arith(op, lhs, rhs)
        assert np.array_equal(_h(res), e)
        assert np.array_equal(_h(resf), O.cast(e, np.float32))


def test_joins_and_overflow(cl, dev):
    import torch

    r = np.random.default_rng(6)
    L, R = 5_000, 300  # the LT join yields about 750,000 pairs, under the capacity
    left = r.integers(0, 2000, L).astype(np.int64)
    right = r.integers(0, 2000, R).astype(np.int64)
    cap = 1 << 20
    pairs = torch.empty((cap, 2), dtype=torch.int32, device=dev)
    act = torch.zeros(1, dtype=torch.int32, device=dev)
    for op in (4, 0):  # EQ, LT
        assert cl.compat_outer_join(op, _p(_t(right, dev)), _p(_t(left, dev)), _p(pairs), _p(act), L, R, cap,
                                    _s()) == 0
        e = O.join_pairs(op, left, right)
        c = int(_h(act)[0])
        assert c == e.shape[0] and np.array_equal(_h(pairs)[:c].view(np.uint32), e)
    # a pair buffer smaller than the result: the launcher throws YSQL_EOVERFLOW
    e = O.join_pairs(0, left, right)
    assert cl.compat_outer_join(0, _p(_t(right, dev)), _p(_t(left, dev)), _p(pairs), _p(act), L, R,
                                e.shape[0] - 1, _s()) == -3
    # AND of two pair lists
    p1 = O.join_pairs(4, left, right)
    p2 = O.join_pairs(4, left % 1000, right % 1000)
    outp = torch.empty((max(p1.shape[0], 1), 2), dtype=torch.int32, device=dev)
    assert cl.compat_and_join(_p(_t(p1.view(np.int32), dev)), _p(_t(p2.view(np.int32), dev)), _p(outp), _p(act),
                              p1.shape[0], p2.shape[0], _s()) == 0
    ei = O.pairs_intersect(p1, p2)
    c = int(_h(act)[0])
    assert c == ei.shape[0] and np.array_equal(_h(outp)[:c].view(np.uint32), ei)


def test_cross_products(cl, dev):
    import torch

    r = np.random.default_rng(7)
    lbs, rbs, off = 37, 23, 11
    ints = r.integers(-100, 100, 100).astype(np.int32)
    dt = r.integers(0, 2**40, 100).astype(np.int64)
    strs = r.integers(0, 255, (100, 256)).astype(np.uint8)
    nul = (r.random(100) < 0.2).astype(np.uint8)
    for isleft in (1, 0):
        rows = lbs if isleft else rbs
        for which, typ, src in ((0, 0, ints), (1, 0, ints), (1, 2, dt), (1, 3, strs), (2, 2, dt), (2, 3, strs)):
            o = off if which == 2 else 0
            dst = torch.empty((lbs * rbs,) + src.shape[1:], dtype=_t(src, dev).dtype, device=dev)
            dn = torch.empty(lbs * rbs, dtype=torch.uint8, device=dev)
            assert cl.compat_cross(which, typ, _p(_t(src, dev)), _p(dst), _p(_t(nul, dev)), _p(dn), o, lbs, rbs,
                                   isleft, _s()) == 0
            e, en = O.cross_expand(src[o:o + rows], lbs, rbs, isleft, nul[o:o + rows].copy())
            assert np.array_equal(_h(dst), e) and np.array_equal(_h(dn), en), (which, typ, isleft)


def test_reductions_fold_into_reference_accumulators(cl, dev):
    import torch

    r = np.random.default_rng(8)
    n = 200_001
    nul = (r.random(n) < 0.05).astype(np.uint8)
    keep = nul == 0
    i32 = r.integers(2**29, 2**30, n).astype(np.int32)  # the int32 sum wraps, as the reference's does
    f32 = r.random(n).astype(np.float32)
    i64 = r.integers(-2**60, 2**60, n).astype(np.int64)
    for dtype, x in ((0, i32), (1, f32), (2, i64)):
        xd = _t(x, dev)
        for agg, init in ((0, x.min()), (1, x.max()), (2, 5)):  # MAX, MIN, SUM; a nonzero starting value
            if agg == 2:
                acc_np = np.array([init], dtype=np.float64 if dtype == 1 else x.dtype)
            else:
                acc_np = np.array([init], dtype=x.dtype)
            acc = _t(acc_np, dev)
            assert cl.compat_reduce(dtype, agg, _p(xd), _p(acc), _p(_t(nul, dev)), n, _s()) == 0
            got = _h(acc)[0]
            v = x[keep]
            if agg == 0:
                assert got == max(init, v.max())
            elif agg == 1:
                assert got == min(init, v.min())
            elif dtype == 0:
                assert int(got) == ((int(init) + int(v.astype(np.int64).sum()) + 2**31) % 2**32) - 2**31
            elif dtype == 2:
                assert int(got) == ((int(init) + int(v.astype(object).sum()) + 2**63) % 2**64) - 2**63
            else:
                s = init + math.fsum(v.astype(np.float64))
                assert abs(got - s) <= 1e-6 * abs(s)
    cnt = _t(np.array([3], np.int32), dev)
    assert cl.compat_reduce(0, 4, None, _p(cnt), _p(_t(nul, dev)), n, _s()) == 0
    assert int(_h(cnt)[0]) == 3 + int(keep.sum())
    # AVG and double -> float finalisers
    counter = _t(np.array([7], np.int32), dev)
    avg = torch.empty(1, dtype=torch.float32, device=dev)
    fs = torch.empty(1, dtype=torch.float32, device=dev)
    ds = _t(np.array([1.0 / 3.0]), dev)
    for dtype, sv in ((0, np.array([100], np.int32)), (1, np.array([100.5])), (2, np.array([2**40], np.int64))):
        assert cl.compat_finalize(dtype, _p(counter), _p(_t(sv, dev)), _p(avg), _p(ds), _p(fs)) == 0
        assert _h(avg)[0] == np.float32(float(sv[0]) / 7.0)
        assert _h(fs)[0] == np.float32(1.0 / 3.0)


@pytest.mark.parametrize("merge", [0, 1])
def test_batch_level_gathers(cl, dev, merge):
    import torch

    r = np.random.default_rng(9 + merge)
    na, nb, rows = 4000, 3000, 5000
    mk = lambda k: (r.integers(-10**6, 10**6, k).astype(np.int32), r.integers(0, 2**50, k).astype(np.int64),  # noqa
                    r.integers(0, 255, (k, 256)).astype(np.uint8), [(r.random(k) < 0.1).astype(np.uint8)
                                                                     for _ in range(3)])
    a_cols, b_cols = mk(na), mk(nb)
    keep = []

    def arr(cols):
        d = [_t(c, dev) for c in cols[:3]]
        nl = [_t(x, dev) for x in cols[3]]
        keep.extend(d + nl)
        return (P * 3)(*[P(t.data_ptr()) for t in d]), (P * 3)(*[P(t.data_ptr()) for t in nl])

    ac, an = arr(a_cols)
    bc, bn = arr(b_cols)
    if merge:
        isa = (r.random(rows) < 0.5).astype(np.uint8)
        idx = np.where(isa == 1, r.integers(0, na, rows), r.integers(0, nb, rows)).astype(np.uint32)
    else:
        rows = na
        idx = r.permutation(na).astype(np.uint32)
        isa = np.ones(rows, np.uint8)
    outs = [torch.empty(rows, dtype=torch.int32, device=dev), torch.empty(rows, dtype=torch.int64, device=dev),
            torch.empty((rows, 256), dtype=torch.uint8, device=dev)]
    onul = [torch.empty(rows, dtype=torch.uint8, device=dev) for _ in range(3)]
    oc = (P * 3)(*[P(t.data_ptr()) for t in outs])
    on = (P * 3)(*[P(t.data_ptr()) for t in onul])
    for asc in ((1, 0) if merge else (1,)):
        assert cl.compat_move_batch(merge, ac, an, bc, bn, na, _p(_t(idx.view(np.int32), dev)), _p(_t(isa, dev)),
                                    rows, asc, oc, on, _s()) == 0
        for c in range(3):
            if merge:
                e, en = O.gather2(a_cols[c], b_cols[c], idx, isa, reverse=not asc, a_null=a_cols[3][c],
                                  b_null=b_cols[3][c])
            else:
                e, en = O.gather(a_cols[c], idx, a_cols[3][c])
            assert np.array_equal(_h(outs[c]), e) and np.array_equal(_h(onul[c]), en), (merge, asc, c)


@pytest.mark.parametrize("desc", [0, 1])
def test_string_sort_and_merge(cl, dev, desc):
    """launch_radix_sort<AscOpString / DescOpString, String, ...> sorts the
    strings in place and returns the permutation; launch_merge_sorted_array_kernel
    <String, AscSortStrOp / DescSortStrOp> merges two sorted string runs."""
    import torch

    r = np.random.default_rng(40 + desc)

    def strs(n):
        s = np.zeros((n, 256), np.uint8)
        for i in range(n):
            k = int(r.integers(0, 6))
            s[i, :k] = r.choice(np.frombuffer(b"ab\x90", np.uint8), k)
        return s

    s = strs(5000)
    t = _t(s, dev)
    idx = torch.empty(5000, dtype=torch.int32, device=dev)
    assert cl.compat_sort_strings(desc, _p(t), 5000, _p(idx), _s()) == 0
    want = O.sort_rows([s], [desc])
    assert np.array_equal(_h(idx).view(np.uint32), want) and np.array_equal(_h(t), s[want])
    A, B = s[want][:3000].copy(), strs(2000)
    B = B[O.sort_rows([B], [desc])]
    k = 4500
    Cd = torch.empty((k, 256), dtype=torch.uint8, device=dev)
    ix = torch.empty(k, dtype=torch.int32, device=dev)
    isa = torch.empty(k, dtype=torch.uint8, device=dev)
    lasti = torch.zeros(1, dtype=torch.int32, device=dev)
    assert cl.compat_merge_strings(desc, _p(_t(A, dev)), _p(_t(B, dev)), _p(Cd), _p(ix), _p(isa), _p(lasti), k,
                                   3000, 2000, _s()) == 0
    ec, ei, ea, cons = O.merge(O.STRING, desc, A, B, k)
    assert np.array_equal(_h(Cd), ec) and np.array_equal(_h(ix).view(np.uint32), ei)
    assert np.array_equal(_h(isa), ea) and int(_h(lasti)[0]) == cons
