"""k-way run merge (ysql_merge_runs) against the oracle: the merged output
of stably sorted runs equals one stable argsort of their concatenation
(ties by run, then position), for 1..16 runs, empty and ragged runs,
heavy ties, uint32 and uint64 keys; the partition bound holds at its edge
(every key equal)."""
import numpy as np
import pytest

from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def _check(dev, keys, sizes):
    import torch

    from yallasql_amd import ops

    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    runs = []
    for r in range(len(sizes)):
        k = keys[offs[r]:offs[r + 1]]
        perm = O.stable_argsort(k)
        runs.append((k[perm], (perm + offs[r]).astype(np.uint32)))
    rk = np.concatenate([x[0] for x in runs]) if runs else keys[:0]
    rv = np.concatenate([x[1] for x in runs]) if runs else np.zeros(0, np.uint32)
    sdt = np.int64 if keys.dtype == np.uint64 else np.int32
    kd = torch.from_numpy(rk.view(sdt).copy()).to(dev)
    vd = torch.from_numpy(rv.view(np.int32).copy()).to(dev)
    ok, ov = ops.merge_runs(kd, vd, [int(o) for o in offs])
    torch.cuda.synchronize()
    exp = O.stable_argsort(keys)
    assert np.array_equal(ov.cpu().numpy().view(np.uint32), exp)
    assert np.array_equal(ok.cpu().numpy().view(keys.dtype), keys[exp])


@pytest.mark.parametrize("nruns", [1, 2, 3, 6, 7, 16])
@pytest.mark.parametrize("key64", [False, True])
def test_merge_runs_random(dev, nruns, key64):
    r = np.random.default_rng(nruns * 10 + key64)
    sizes = [int(x) for x in r.integers(0, 300_000, nruns)]
    if nruns > 2:
        sizes[1] = 0  # an empty run
        sizes[2] = 1
    n = sum(sizes)
    dt = np.uint64 if key64 else np.uint32
    hi = 2**63 if key64 else 2**32
    keys = np.where(r.random(n) < 0.5, r.integers(0, hi, n, dtype=np.uint64),
                    r.integers(0, 1000, n, dtype=np.uint64)).astype(dt)  # half of the keys heavily tied
    print(f"merge_runs: {nruns} runs, key64={key64}, n={n}", flush=True)
    _check(dev, keys, sizes)


@pytest.mark.parametrize("nruns", [2, 5, 16])
def test_merge_runs_all_equal_and_edges(dev, nruns):
    from yallasql_amd import ops  # noqa: F401

    # every key equal: partitions are cut purely by run and position
    sizes = [4096 * 3 + 7] * nruns
    keys = np.full(sum(sizes), 42, np.uint64)
    _check(dev, keys, sizes)
    # sizes at the sample stride's edges
    r = np.random.default_rng(nruns)
    sizes = [255, 256, 257, 511, 512, 513, 4095, 4096][:nruns] + [1000] * max(0, nruns - 8)
    keys = r.integers(0, 50, sum(sizes), dtype=np.uint64).astype(np.uint32)
    _check(dev, keys, sizes)


def test_merge_runs_rejects_bad_arguments(dev):
    import ctypes as C

    import torch

    from yallasql_amd import _lib as L

    k = torch.zeros(10, dtype=torch.int64, device=dev)
    v = torch.zeros(10, dtype=torch.int32, device=dev)
    offs = (C.c_uint64 * 3)(0, 5, 10)
    ws = torch.empty(1 << 20, dtype=torch.uint8, device=dev)
    p = lambda t: C.c_void_p(t.data_ptr())  # noqa: E731
    # in place
    assert L.lib.ysql_merge_runs(L.UKEY64, p(k), p(v), offs, 2, p(k), p(v), p(ws), ws.numel(), None) == L.EINVAL
    # too many runs, offsets going back
    assert L.lib.ysql_merge_runs(L.UKEY64, p(k), p(v), offs, 17, p(k), p(v), p(ws), ws.numel(), None) == L.EINVAL
    bad = (C.c_uint64 * 3)(0, 6, 5)
    assert L.lib.ysql_merge_runs(L.UKEY64, p(k), p(v), bad, 2, p(k), p(v), p(ws), ws.numel(), None) == L.EINVAL
