"""STRING sort keys and merges (HIP) against the oracle: the stable
strcmp-order sort (ASC / DESC, shared prefixes, bytes above 0x7f, full-length
rows with no NUL, bytes after the NUL), one LSD step of a multi-column sort,
merge path of string runs, the multi-key row merge, and ORDER BY on string
columns through the engine, over several batches (sorted runs merged)."""
import numpy as np
import pytest
import torch

from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ops(dev):
    from yallasql_amd import ops as _ops

    return _ops


def strings(r, n, maxlen, alphabet=b"ab\x80\xffzq", odd=True):
    """(n, 256) NUL-padded strings; with `odd`, a few rows with no NUL and a
    few with bytes after the NUL (strcmp ignores those)."""
    lens = r.integers(0, maxlen + 1, n)
    codes = np.frombuffer(alphabet, np.uint8)
    s = codes[r.integers(0, len(codes), (n, 256))]
    s[np.arange(256)[None, :] >= lens[:, None]] = 0
    if odd and n > 20:
        s[r.integers(0, n, 3)] = 0x61  # 256 bytes, no NUL
        rows = r.integers(0, n, 5)
        s[rows, 200] = 0x7a  # garbage after the NUL (lens <= maxlen < 200)
    return np.ascontiguousarray(s)


@pytest.mark.parametrize("desc", [False, True])
@pytest.mark.parametrize("n,maxlen", [(1, 5), (2, 0), (7, 3), (100_003, 40), (20_000, 255), (50_000, 9)])
def test_string_argsort_matches_oracle(ops, dev, n, maxlen, desc):
    r = np.random.default_rng(n + maxlen + desc)
    s = strings(r, n, maxlen)
    if n > 50:
        s[10:40] = s[9]  # a run of equal strings: ties keep row order
    t = torch.from_numpy(s).to(dev)
    perm = ops.string_argsort(t, descending=desc).cpu().numpy().astype(np.uint32)
    assert np.array_equal(perm, O.sort_rows([s], [desc]))


def test_all_empty_and_all_equal(ops, dev):
    for fill in (0, 0x61):
        s = np.zeros((5000, 256), np.uint8)
        s[:, :10] = fill
        t = torch.from_numpy(s).to(dev)
        for desc in (False, True):
            assert np.array_equal(ops.string_argsort(t, desc).cpu().numpy(), np.arange(5000))


def test_string_sort_refines_a_permutation(ops, dev):
    """LSD over two columns: sort by the int column, then refine that
    permutation by the string column = ORDER BY (string, int)."""
    r = np.random.default_rng(5)
    n = 60_000
    s = strings(r, n, 3, alphabet=b"abc", odd=False)
    iv = r.integers(-50, 50, n).astype(np.int32)
    for ds, di in ((False, False), (True, False), (False, True)):
        k = ops.sortkey_encode(torch.from_numpy(iv).to(dev), descending=di)
        perm = ops.radix_argsort(k)
        ops.string_sort(torch.from_numpy(s).to(dev), perm, descending=ds)
        assert np.array_equal(perm.cpu().numpy().astype(np.uint32), O.sort_rows([s, iv], [ds, di]))


def _sorted_run(r, n, maxlen, desc):
    s = strings(r, n, maxlen, odd=False)
    return np.ascontiguousarray(s[O.sort_rows([s], [desc])])


@pytest.mark.parametrize("desc", [False, True])
@pytest.mark.parametrize("m,n,kmax", [(3000, 5000, None), (0, 70, None), (64, 0, None), (10_000, 9_999, 12_345)])
def test_merge_path_strings_matches_oracle(ops, dev, m, n, kmax, desc):
    from yallasql_amd import _lib as L

    r = np.random.default_rng(m + n)
    a, b = _sorted_run(r, m, 4, desc), _sorted_run(r, n, 4, desc)
    ta, tb = torch.from_numpy(a).to(dev), torch.from_numpy(b).to(dev)
    C_, idx, isa, cons = ops.merge_path(L.STRING, desc, ta, tb, k_max=kmax)
    oc, oidx, oisa, ocons = O.merge(O.STRING, desc, a, b, k_max=kmax)
    assert np.array_equal(isa.cpu().numpy(), oisa)
    assert np.array_equal(idx.cpu().numpy().astype(np.uint32), oidx)
    assert np.array_equal(C_.cpu().numpy(), oc)
    assert int(cons.item()) == ocons


def test_merge_rows_mixed_keys(ops, dev):
    """Runs sorted by (string ASC, int DESC, float ASC, datetime DESC)
    merged by the key rows = the stable sort of A then B."""
    from yallasql_amd import _lib as L

    r = np.random.default_rng(9)

    def run(n):
        s = strings(r, n, 2, alphabet=b"ab", odd=False)
        iv = r.integers(-3, 3, n).astype(np.int32)
        fv = r.choice(np.array([-0.0, 0.0, -1.5, 2.0, np.inf, -np.inf], np.float32), n)
        dv = r.integers(-2, 2, n).astype(np.int64)
        p = O.sort_rows([s, iv, fv, dv], [0, 1, 0, 1])
        return [np.ascontiguousarray(c[p]) for c in (s, iv, fv, dv)]

    A, B = run(4000), run(3000)
    keys = [(t, d, torch.from_numpy(a).to(dev), torch.from_numpy(b).to(dev))
            for t, d, a, b in zip((L.STRING, L.INT, L.FLOAT, L.DATETIME), (0, 1, 0, 1), A, B)]
    idx, isa, cons, _ = ops.merge_rows(keys, 4000, 3000)
    got = np.where(isa.cpu().numpy() == 1, idx.cpu().numpy(), 4000 + idx.cpu().numpy())
    cat = [np.concatenate([x, y]) for x, y in zip(A, B)]
    assert np.array_equal(got.astype(np.uint32), O.sort_rows(cat, [0, 1, 0, 1]))
    assert int(cons.item()) == 4000


@pytest.mark.parametrize("batch_rows", [1 << 16, 9_000])
def test_engine_order_by_strings(dev, batch_rows):
    """ORDER BY name DESC, id ASC over several batches: each batch sorted
    column by column, the runs merged by their key rows."""
    from yallasql_amd import engine as E

    r = np.random.default_rng(batch_rows)
    n = 50_000
    names = strings(r, n, 6, alphabet=b"abcd", odd=False)
    ids = r.integers(0, 100, n).astype(np.int32)
    price = r.random(n).astype(np.float32)
    ctx = E.Context()
    t = ctx.table("t", [("id", ids), ("name", names), ("price", price)])
    cols, nulls, batches = ctx.scan(t, batch_rows).order([1, 0], [True, False]).execute()
    p = O.sort_rows([names, ids], [1, 0])
    assert np.array_equal(cols[1], names[p])
    assert np.array_equal(cols[0], ids[p]) and np.array_equal(cols[2], price[p])
