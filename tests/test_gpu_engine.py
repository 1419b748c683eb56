"""Operator-level parity: plans built like the reference's operator trees
(scan -> filter -> order -> projection, aggregate, join, cross product) run
through the HIP kernels and are checked against the CPU oracle."""
import numpy as np
import pytest

from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def E(dev):
    from yallasql_amd import engine

    return engine


def config1_table(n, seed=0):
    r = np.random.default_rng(seed)
    ids = r.permutation(n).astype(np.int32)
    price = (1 + 999 * r.random(n)).astype(np.float32)
    price = np.minimum(price, np.float32(999.9999))
    price[::97] = np.float32(750.5)  # ties on the sort key
    cat = r.integers(0, 64, n).astype(np.int32)
    return ids, price, cat


def oracle_config1(ids, price, by_id):
    mask = O.compare(O.GT, price, np.array([500.0], np.float32), rhs_scalar=True)
    sid, _ = O.compact(ids, mask)
    sp, _ = O.compact(price, mask)
    if by_id:
        key = (O.sortkey_encode(sp, desc=True).astype(np.uint64) << np.uint64(32)) | O.sortkey_encode(sid).astype(np.uint64)
        perm = O.stable_argsort(key)
    else:
        perm = O.stable_argsort(O.sortkey_encode(sp), desc=True)
    return sid[perm], sp[perm]


@pytest.mark.parametrize("batch_rows", [1 << 20, 1 << 18, 100_000])
@pytest.mark.parametrize("by_id", [True, False])
def test_config1_filter_sort_project(E, batch_rows, by_id):
    n = 1 << 20
    ids, price, cat = config1_table(n)
    ctx = E.Context()
    t = ctx.table("t", [("id", ids), ("price", price), ("category", cat)])
    p = ctx.scan(t, batch_rows).filter(E.cmp(2, E.ref(1), E.f32(500.0)))
    p = p.order([1, 0], [1, 0]) if by_id else p.order([1], [1])
    p = p.project([E.ref(0), E.ref(1), E.arith(2, E.i32(2), E.ref(0))])
    cols, nulls, batches = p.execute()
    eid, eprice = oracle_config1(ids, price, by_id)
    assert cols[0].shape[0] == eid.shape[0]
    assert np.array_equal(cols[0], eid)
    assert np.array_equal(cols[1], eprice)
    assert np.array_equal(cols[2], (2 * eid).astype(np.int32))


def test_aggregates_across_batches_with_empty_filter_result(E):
    r = np.random.default_rng(3)
    n = 300_000
    key = r.integers(-10**6, 10**6, n).astype(np.int64)
    val = r.random(n).astype(np.float32)
    val[:100_000] = 2.0  # the first batch is filtered out entirely
    ctx = E.Context()
    t = ctx.table("t", [("key", key), ("val", val)])
    p = ctx.scan(t, 100_000).filter(E.cmp(0, E.ref(1), E.f32(0.4)))
    p = p.aggregate([(4, 1), (2, 1), (1, 0), (0, 0), (3, 1), (5, 0)])
    cols, nulls, batches = p.execute()
    m = val < np.float32(0.4)
    assert cols[0][0] == m.sum()
    assert abs(cols[1][0] - val[m].astype(np.float64).sum()) <= 1e-6 * val[m].sum()
    assert cols[2][0] == key[m].min() and cols[3][0] == key[m].max()
    assert abs(cols[4][0] - val[m].astype(np.float64).mean()) <= 1e-6
    assert cols[5][0] == m.sum()


def test_group_by(E):
    r = np.random.default_rng(4)
    n = 200_000
    year = r.integers(1992, 1999, n).astype(np.int32)
    brand = r.integers(0, 40, n).astype(np.int32)
    x = r.integers(0, 10**6, n).astype(np.int64)
    ctx = E.Context()
    t = ctx.table("t", [("year", year), ("brand", brand), ("x", x)])
    p = ctx.scan(t, 50_000).aggregate([(2, 2), (5, 0)], groups=[(0, 1992, 7), (1, 0, 40)])
    cols, _, _ = p.execute()
    g = (year - 1992) * 40 + brand
    es = np.zeros(280, np.int64)
    np.add.at(es, g, x)
    ec = np.bincount(g, minlength=280)
    live = np.nonzero(ec)[0]
    assert np.array_equal(cols[0], (1992 + live // 40).astype(np.int32))
    assert np.array_equal(cols[1], (live % 40).astype(np.int32))
    assert np.array_equal(cols[2], es[live]) and np.array_equal(cols[3], ec[live])


@pytest.mark.parametrize("op", [4, 0, 5])
def test_join_multi_batch(E, op):
    r = np.random.default_rng(5 + op)
    Ln, Rn = 3000, 1700
    lk = r.integers(0, 500, Ln).astype(np.int32)
    rk = r.integers(0, 500, Rn).astype(np.int32)
    lid = np.arange(Ln, dtype=np.int32)
    rid = np.arange(Rn, dtype=np.int32)
    ctx = E.Context()
    tl = ctx.table("l", [("k", lk), ("id", lid)])
    tr = ctx.table("r", [("k", rk), ("id", rid)])
    p = ctx.scan(tl, 1000).join(ctx.scan(tr, 800), 0, 0, op)
    cols, _, batches = p.execute()
    got = sorted(zip(cols[1].tolist(), cols[3].tolist()))
    exp = O.join_pairs(op, lk, rk)
    assert got == sorted(map(tuple, exp.tolist()))
    assert np.array_equal(cols[0], lk[cols[1]]) and np.array_equal(cols[2], rk[cols[3]])


def test_cross_product(E):
    lid = np.arange(37, dtype=np.int32)
    rid = np.arange(300, dtype=np.int32) * 10
    ctx = E.Context()
    tl = ctx.table("l", [("id", lid)])
    tr = ctx.table("r", [("id", rid)])
    cols, _, _ = ctx.scan(tl, 20).cross(ctx.scan(tr, 256)).execute()
    got = sorted(zip(cols[0].tolist(), cols[1].tolist()))
    assert got == sorted((a, b) for a in lid.tolist() for b in rid.tolist())


def test_projection_arith_cast_and_nulls(E):
    r = np.random.default_rng(6)
    n = 50_000
    a = r.integers(-100, 100, n).astype(np.int32)
    b = r.integers(-3, 3, n).astype(np.int32)
    nul = (r.random(n) < 0.1).astype(np.uint8)
    ctx = E.Context()
    t = ctx.table("t", [("a", a), ("b", b)], nulls=[nul, None])
    p = ctx.scan(t, 20_000).project([E.arith(3, E.ref(0), E.ref(1)), E.cast(1, E.ref(0)),
                                     E.cmp(4, E.ref(1), E.i32(0))])
    cols, nulls, _ = p.execute()
    assert np.array_equal(cols[0], O.arith(3, a, b))
    assert np.array_equal(cols[1], a.astype(np.float32))
    assert np.array_equal(cols[2], (b == 0).astype(np.int32))
    assert np.array_equal(nulls[0], nul)  # nulls of the non-constant operand


def test_filter_keeps_null_bytes_and_strings(E):
    r = np.random.default_rng(7)
    n = 10_000
    k = r.integers(0, 10, n).astype(np.int32)
    s = np.zeros((n, 256), np.uint8)
    s[:, :4] = r.integers(97, 123, (n, 4))
    nul = (r.random(n) < 0.2).astype(np.uint8)
    ctx = E.Context()
    t = ctx.table("t", [("k", k), ("s", s)], nulls=[nul, None])
    cols, nulls, _ = ctx.scan(t, 3000).filter(E.cmp(1, E.ref(0), E.i32(4))).execute()
    m = k <= 4
    assert np.array_equal(cols[0], k[m]) and np.array_equal(cols[1], s[m])
    assert np.array_equal(nulls[0], nul[m])


def test_budget_moves_batches_out_of_hbm(E):
    """A 64 KB HBM budget: ORDER BY's sorted runs (40 KB each) leave HBM for
    host memory as new ones arrive; a merged run larger than the budget stays
    in HBM once everything else has moved out; the result is unchanged."""
    n = 100_000
    r = np.random.default_rng(1)
    a = r.integers(-1000, 1000, n).astype(np.int32)
    ctx = E.Context(hbm_budget=1 << 16)
    t = ctx.table("t", [("a", a)])
    cols, _, _ = ctx.scan(t, 10_000).order([0], [0]).execute()
    assert np.array_equal(cols[0], a[np.argsort(a, kind="stable")])
    st = ctx.cache_stats()
    assert st["evictions"] > 0 and st["restores_from_host"] > 0 and st["spills"] == 0, st


@pytest.mark.parametrize("strings", [False, True])
def test_order_runs_spill_to_host_and_files(E, tmp_path, strings):
    """An ORDER BY over 16 batches under a 3 MB HBM budget and a 2 MB host
    budget: its sorted runs leave HBM for pinned host memory and then for
    spill files, and come back from the files for their merges; the result
    equals the unlimited run and the oracle."""
    n = 1 << 20
    ids, price, cat = config1_table(n, seed=5)
    cols = [("id", ids), ("price", price), ("category", cat)]
    if strings:
        names = np.zeros((n, 256), np.uint8)
        names[:, 0] = 97 + (cat % 3)
        cols = [("id", ids), ("price", price), ("name", names)]

    def run(ctx):
        t = ctx.table("t", cols)
        keys = ([2, 1, 0], [1, 1, 0]) if strings else ([1, 0], [1, 0])
        return ctx.scan(t, 1 << 16).order(*keys).execute()

    big = E.Context()
    want, _, _ = run(big)
    assert big.cache_stats()["evictions"] == 0
    small = E.Context(hbm_budget=3 << 20, host_budget=2 << 20, spill_dir=tmp_path)
    got, _, _ = run(small)
    for a, b in zip(got, want):
        assert np.array_equal(a, b)
    st = small.cache_stats()
    assert st["evictions"] > 0 and st["spills"] > 0 and st["restores_from_file"] > 0, st
    if strings:
        p = O.sort_rows([cols[2][1], price, ids], [1, 1, 0])
    else:
        p = O.sort_rows([price, ids], [1, 0])
    assert np.array_equal(got[0], ids[p]) and np.array_equal(got[1], price[p])
    assert not list(tmp_path.iterdir())  # spill files are removed once read back
