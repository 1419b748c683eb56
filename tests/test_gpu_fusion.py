"""Expression fusion: predicate and projection trees run as one register
program per batch (ysql_expr_eval / ysql_expr_filter_compact) and must give
the columns of the node-by-node evaluation (the reference's evaluate()
chain, include/engine/operators/expressions/), which the oracle restates
node by node (oracle/pyoracle.py eval_tree).  Integers, masks and NULL bytes
bit-exact; floats bit-exact too (each node is one IEEE operation in both
forms), with NaNs compared as NaNs (the host and the GPU set different NaN
sign bits)."""
import struct

import numpy as np
import pytest

from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

LT, LE, GT, GE, EQ, NE = range(6)
AND, OR = 0, 1
ADD, SUB, MUL, DIV, REM = range(5)
INT, FLOAT, DATETIME, BOOL = 0, 1, 2, 4
NP = {INT: np.int32, FLOAT: np.float32, DATETIME: np.int64}


@pytest.fixture(scope="module")
def E(dev):
    from yallasql_amd import engine

    return engine


@pytest.fixture(scope="module")
def ops(dev):
    from yallasql_amd import ops

    return ops


def same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    if a.dtype != b.dtype or a.shape != b.shape:
        return False
    if a.dtype == np.float32:
        na, nb = np.isnan(a), np.isnan(b)
        return bool(np.array_equal(na, nb) and np.array_equal(a[~na].view(np.uint32), b[~nb].view(np.uint32)))
    return bool(np.array_equal(a, b))


def table_data(n, seed):
    r = np.random.default_rng(seed)
    i1 = r.integers(-50, 50, n).astype(np.int32)
    i2 = r.integers(-5, 6, n).astype(np.int32)
    i2[::7] = 0
    i2[::11] = -1
    i1[::13] = np.iinfo(np.int32).min
    f1 = (r.standard_normal(n) * 100).astype(np.float32)
    f1[::17] = 0.0
    f2 = r.random(n).astype(np.float32) * 3
    d1 = r.integers(-10**12, 10**12, n).astype(np.int64)
    d1[::19] = np.iinfo(np.int64).min
    d2 = r.integers(-3, 4, n).astype(np.int64)
    cols = [i1, i2, f1, f2, d1, d2]
    nulls = [None] * len(cols)
    nulls[0] = (r.random(n) < 0.1).astype(np.uint8)
    nulls[2] = (r.random(n) < 0.2).astype(np.uint8)
    nulls[4] = (r.random(n) < 0.05).astype(np.uint8)
    return cols, nulls


COLS_OF = {INT: [0, 1], FLOAT: [2, 3], DATETIME: [4, 5]}


def gen_value(r, t, depth):
    k = r.integers(0, 10) if depth > 0 else r.integers(0, 3)
    if k <= 1:
        return ("ref", int(r.choice(COLS_OF[t])))
    if k == 2:
        v = int(r.integers(-9, 10))
        return {INT: ("i32", v), FLOAT: ("f32", float(np.float32(v * 0.75))), DATETIME: ("i64", v * 1000003)}[t]
    if k <= 7:
        ops = [ADD, SUB, MUL, DIV] + ([REM] if t != FLOAT else [])
        return ("arith", int(r.choice(ops)), gen_value(r, t, depth - 1), gen_value(r, t, depth - 1))
    src = int(r.choice([x for x in (INT, FLOAT, DATETIME) if x != t]))
    return ("cast", t, gen_value(r, src, depth - 1))


def gen_bool(r, depth):
    if depth > 0 and r.random() < 0.4:
        return ("logic", int(r.integers(0, 2)), gen_bool(r, depth - 1), gen_bool(r, depth - 1))
    t = int(r.choice([INT, FLOAT, DATETIME]))
    return ("cmp", int(r.integers(0, 6)), gen_value(r, t, depth - 1), gen_value(r, t, depth - 1))


def has_ref(node):
    if node[0] == "ref":
        return True
    return any(has_ref(c) for c in node[1:] if isinstance(c, tuple))


def to_engine(E, node):
    k = node[0]
    if k == "ref":
        return E.ref(node[1])
    if k == "i32":
        return E.i32(node[1])
    if k == "f32":
        return E.f32(node[1])
    if k == "i64":
        return E.i64(node[1])
    if k == "cast":
        return E.cast(node[1], to_engine(E, node[2]))
    f = {"cmp": E.cmp, "logic": E.logic, "arith": E.arith}[k]
    return f(node[1], to_engine(E, node[2]), to_engine(E, node[3]))


def oracle_column(node, cols, nulls):
    v, sc, nl = O.eval_tree(node, cols, nulls)
    assert not sc
    if v.dtype == np.uint8:  # a projected mask is an INT column of 0/1
        v = O.cast(v, np.int32)
    return v, (np.zeros(v.shape[0], np.uint8) if nl is None else nl)


def run_projection(E, cols, nulls, trees, batch_rows, fusion):
    ctx = E.Context(fusion=fusion)
    t = ctx.table("t", [(f"c{i}", c) for i, c in enumerate(cols)], nulls=nulls)
    p = ctx.scan(t, batch_rows).project([to_engine(E, x) for x in trees])
    out = p.execute()
    ctx.close()
    return out


# ---------------------------------------------------------------- engine


@pytest.mark.parametrize("seed,batch_rows", [(1, 1 << 16), (2, 1001), (3, 50_000)])
def test_random_projection_trees(E, seed, batch_rows):
    n = 120_007
    cols, nulls = table_data(n, seed)
    r = np.random.default_rng(100 + seed)
    trees = []
    while len(trees) < 12:
        t = gen_value(r, int(r.choice([INT, FLOAT, DATETIME])), 3) if r.random() < 0.7 else gen_bool(r, 3)
        if has_ref(t):
            trees.append(t)
    fused, fn, _ = run_projection(E, cols, nulls, trees, batch_rows, True)
    plain, pn, _ = run_projection(E, cols, nulls, trees, batch_rows, False)
    for i, tree in enumerate(trees):
        ev, en = oracle_column(tree, cols, nulls)
        assert same(fused[i], ev), (i, tree)
        assert same(plain[i], ev), (i, tree)
        assert np.array_equal(fn[i], en), (i, tree)
        assert np.array_equal(pn[i], en), (i, tree)


@pytest.mark.parametrize("seed,batch_rows", [(4, 1 << 15), (5, 999)])
def test_random_filter_trees(E, seed, batch_rows):
    n = 100_003
    cols, nulls = table_data(n, seed)
    strs = np.zeros((n, 256), np.uint8)
    strs[:, 0] = 65 + (np.arange(n) % 26)
    r = np.random.default_rng(200 + seed)
    for _ in range(6):
        tree = gen_bool(r, 3)
        if not has_ref(tree):
            continue
        res = []
        for fusion in (True, False):
            ctx = E.Context(fusion=fusion)
            t = ctx.table("t", [(f"c{i}", c) for i, c in enumerate(cols)] + [("s", strs)], nulls=nulls + [None])
            res.append(ctx.scan(t, batch_rows).filter(to_engine(E, tree)).execute())
            ctx.close()
        mask, _, _ = O.eval_tree(tree, cols, nulls)
        for got, gnul, _ in res:
            for c in range(len(cols)):
                ev, en = O.compact(cols[c], mask, nulls[c])
                assert same(got[c], ev), (tree, c)
                assert np.array_equal(gnul[c], en if nulls[c] is not None else np.zeros(ev.shape[0], np.uint8))
            assert np.array_equal(got[len(cols)], strs[mask.astype(bool)])


def test_config1_shapes_fused_equal_unfused(E):
    """WHERE price > 500 AND category < 32 ... SELECT id, 2*id, price*1.5,
    CAST(price AS INT), id % 7 = 3 over 2^20 rows (the config-1 table)."""
    n = 1 << 20
    r = np.random.default_rng(0)
    ids = r.permutation(n).astype(np.int32)
    price = np.minimum((1 + 999 * r.random(n)).astype(np.float32), np.float32(999.9999))
    cat = r.integers(0, 64, n).astype(np.int32)
    cols = [ids, price, cat]
    pred = ("logic", AND, ("cmp", GT, ("ref", 1), ("f32", 500.0)), ("cmp", LT, ("ref", 2), ("i32", 32)))
    proj = [("ref", 0), ("arith", MUL, ("i32", 2), ("ref", 0)), ("arith", MUL, ("ref", 1), ("f32", 1.5)),
            ("cast", INT, ("ref", 1)), ("cmp", EQ, ("arith", REM, ("ref", 0), ("i32", 7)), ("i32", 3))]
    outs = []
    for fusion in (True, False):
        ctx = E.Context(fusion=fusion)
        t = ctx.table("t", [("id", ids), ("price", price), ("category", cat)])
        p = ctx.scan(t, 1 << 18).filter(to_engine(E, pred)).project([to_engine(E, x) for x in proj])
        outs.append(p.execute())
        ctx.close()
    mask, _, _ = O.eval_tree(pred, cols)
    sel = [O.compact(c, mask)[0] for c in cols]
    for got, _, _ in outs:
        for i, tree in enumerate(proj):
            ev, _ = oracle_column(tree, sel, None)
            assert same(got[i], ev), i


def test_wide_projection_groups_and_register_overflow(E):
    """More than 8 projected trees over more than 8 columns (several fused
    programs), and a right-deep tree needing more than 16 registers (left to
    the node-by-node path): every column still equals the oracle's."""
    n = 50_001
    r = np.random.default_rng(9)
    cols = [r.integers(-100, 100, n).astype(np.int32) for _ in range(10)]
    trees = [("arith", ADD, ("ref", i), ("ref", (i + 1) % 10)) for i in range(10)]
    deep = ("ref", 0)
    for k in range(18):
        deep = ("arith", ADD, ("i32", k), deep)
    trees.append(deep)
    trees.append(("arith", SUB, ("ref", 9), ("i32", 5)))
    got, _, _ = run_projection(E, cols, None, trees, 1 << 14, True)
    for i, tree in enumerate(trees):
        ev, _ = oracle_column(tree, cols, None)
        assert same(got[i], ev), i


def test_constant_coercion_and_null_sides(E):
    """A float constant against an INT column is converted to INT; a
    constant on the left takes the right operand's NULLs; a column-column
    node takes the left column's NULLs; a conjunction its left side's."""
    n = 4099
    cols, nulls = table_data(n, 11)
    trees = [
        ("cmp", LT, ("ref", 0), ("f32", 2.75)),
        ("arith", SUB, ("i32", 7), ("ref", 0)),
        ("arith", ADD, ("ref", 1), ("ref", 0)),
        ("arith", ADD, ("ref", 0), ("ref", 1)),
        ("logic", OR, ("cmp", GT, ("ref", 2), ("f32", 0.0)), ("cmp", LT, ("ref", 0), ("i32", 0))),
        ("cast", DATETIME, ("ref", 2)),
        ("cast", FLOAT, ("ref", 4)),
        ("cast", INT, ("ref", 4)),
        ("arith", DIV, ("ref", 4), ("ref", 5)),
        ("arith", REM, ("ref", 0), ("ref", 1)),
    ]
    got, gn, _ = run_projection(E, cols, nulls, trees, 1000, True)
    for i, tree in enumerate(trees):
        ev, en = oracle_column(tree, cols, nulls)
        assert same(got[i], ev), (i, tree)
        assert np.array_equal(gn[i], en), (i, tree)
    # the NULL-side facts themselves, independent of the oracle
    assert np.array_equal(gn[1], nulls[0]) and np.array_equal(gn[2], np.zeros(n, np.uint8))
    assert np.array_equal(gn[3], nulls[0]) and np.array_equal(gn[4], nulls[2])


# ---------------------------------------------------------------- C ABI


def f32bits(x):
    return struct.unpack("<I", struct.pack("<f", x))[0]


def test_abi_eval_program_with_nulls_and_datetime(ops):
    import torch

    n = 10_007
    cols, nulls = table_data(n, 21)
    dev = torch.device("cuda", 0)
    t = [torch.from_numpy(c).to(dev) for c in (cols[0], cols[2], cols[4])]
    tn = [torch.from_numpy(nulls[0]).to(dev), torch.from_numpy(nulls[2]).to(dev), torch.from_numpy(nulls[4]).to(dev)]
    X = ops.xinsn
    L = ops.L
    prog = [
        X(L.XOP_CONST, 3, type=INT, imm=3),
        X(L.XOP_ARITH, 4, a=0, b=3, type=INT, sub=MUL, nsrc=0),          # r4 = c0 * 3
        X(L.XOP_CAST, 5, a=4, type=INT, sub=DATETIME, nsrc=4),            # r5 = (int64) r4
        X(L.XOP_ARITH, 5, a=5, b=2, type=DATETIME, sub=ADD, nsrc=2),      # r5 = r5 + c2 (nulls of c2)
        X(L.XOP_CONST, 6, type=FLOAT, imm=f32bits(-1.5)),
        X(L.XOP_CMP, 7, a=1, b=6, type=FLOAT, sub=GE, nsrc=1),            # r7 = c1 >= -1.5
        X(L.XOP_NOT, 8, a=7, nsrc=7),                                     # r8 = !r7
    ]
    res = ops.expr_eval(prog, t, [(4, torch.int32, True), (5, torch.int64, True), (8, torch.int32, True)], nulls=tn)
    e4, _, _ = O.eval_tree(("arith", MUL, ("ref", 0), ("i32", 3)), [cols[0]])
    e5, _, _ = O.eval_tree(("arith", ADD, ("cast", DATETIME, ("arith", MUL, ("ref", 0), ("i32", 3))), ("ref", 1)),
                           [cols[0], cols[4]])
    e7 = O.compare(GE, cols[2], np.array([-1.5], np.float32), rhs_scalar=True)
    assert same(res[0][0].cpu().numpy(), e4) and np.array_equal(res[0][1].cpu().numpy(), nulls[0])
    assert same(res[1][0].cpu().numpy(), e5) and np.array_equal(res[1][1].cpu().numpy(), nulls[4])
    assert np.array_equal(res[2][0].cpu().numpy(), (1 - e7).astype(np.int32))
    assert np.array_equal(res[2][1].cpu().numpy(), nulls[2])


def test_abi_filter_program_16_registers(ops):
    """A predicate using all 16 registers: 8 inputs and 8 temporaries."""
    import torch

    n = 70_001
    r = np.random.default_rng(5)
    cols = [r.integers(0, 10, n).astype(np.int32) for _ in range(8)]
    dev = torch.device("cuda", 0)
    t = [torch.from_numpy(c).to(dev) for c in cols]
    X, L = ops.xinsn, ops.L
    prog = [X(L.XOP_CONST, 8 + i, type=INT, imm=5) for i in range(8)]
    prog += [X(L.XOP_CMP, 8 + i, a=i, b=8 + i, type=INT, sub=LT) for i in range(8)]  # r8+i = c_i < 5
    for i in range(1, 8):
        prog.append(X(L.XOP_LOGIC, 8, a=8, b=8 + i, sub=AND if i % 2 else OR))
    outs, cnt = ops.expr_filter_compact(prog, 8, t, [t[0], t[7]])
    m = cols[0] < 5
    for i in range(1, 8):
        m = (m & (cols[i] < 5)) if i % 2 else (m | (cols[i] < 5))
    assert cnt == int(m.sum())
    assert np.array_equal(outs[0][0].cpu().numpy(), cols[0][m])
    assert np.array_equal(outs[1][0].cpu().numpy(), cols[7][m])


def test_abi_empty_input(ops):
    import torch

    dev = torch.device("cuda", 0)
    x = torch.zeros(0, dtype=torch.int32, device=dev)
    X, L = ops.xinsn, ops.L
    prog = [X(L.XOP_CONST, 1, type=INT, imm=0), X(L.XOP_CMP, 2, a=0, b=1, type=INT, sub=GT)]
    outs, cnt = ops.expr_filter_compact(prog, 2, [x], [x])
    assert cnt == 0
