"""Parity at the benchmark workloads' full sizes on the GPU.

* Out-of-core multi-run sort of int64 keys (50% uniform over [0, 2^63), 50%
  Zipf(1.1) ranks over 10^7 values): the run sort (ysql_radix_begin +
  ysql_radix_pass with the run offset as payload base, runs written into
  slices of shared buffers), then either the pairwise merge levels
  (ysql_merge_pairs) or the k-way merge of all runs (ysql_merge_runs),
  against the oracle's stable argsort at 4 runs of 2^22 rows plus a partial
  one, and at 200M rows in 2^25-row runs by the properties that together
  equal one global stable sort.
* Star join + GROUP BY over the 600M-row fact table against the oracle's
  restatement, every group's sum and count.
* The 4,096 ragged jobs: the persistent launch and the launch-chain graph
  each equal the per-job loop on every output buffer, and a seeded sample of jobs from every kind and size bucket
  (the largest job of each kind included) equals the oracle.
"""
import numpy as np
import pytest

from oracle import pyoracle as O
from tests.test_gpu_batched import check_jobs_against_oracle

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("merge", ["kway", "pairwise"])
@pytest.mark.parametrize("runs,partial", [(4, 123_457), (2, 0)])
def test_multirun_sort_matches_oracle(dev, runs, partial, merge):
    import torch

    from yallasql_amd.datagen import multirun_keys
    from yallasql_amd.multirun import MultiRunSort

    run_rows = 1 << 22
    n = runs * run_rows + partial
    _say(f"multirun sort vs oracle: {n} rows, runs of {run_rows}, {merge} merge")
    keys = multirun_keys(n, dev, seed=1)
    job = MultiRunSort(keys, run_rows, key_bits=63, merge=merge)
    job.step()
    ko, vo = job.result()
    torch.cuda.synchronize()
    kh = keys.cpu().numpy().view(np.uint64)
    exp = O.stable_argsort(kh)
    assert np.array_equal(vo.cpu().numpy().view(np.uint32), exp)
    assert np.array_equal(ko.cpu().numpy().view(np.uint64), kh[exp])
    # the oracle's own multi-run pipeline (run sorts + stable merges) agrees
    ok, ov = O.cpu_multirun_sort_u64(kh, run_rows, 8)
    assert np.array_equal(ov, exp) and np.array_equal(ok, kh[exp])
    # a second step over the same buffers repeats the bits
    job.step()
    ko2, vo2 = job.result()
    assert torch.equal(ko2, ko) and torch.equal(vo2, vo)


@pytest.mark.timeout(600)
@pytest.mark.parametrize("merge", ["kway", "pairwise"])
def test_multirun_sort_full_size(dev, merge):
    import torch

    from yallasql_amd.datagen import multirun_keys
    from yallasql_amd.multirun import MultiRunSort, verify_global_sort

    n, run_rows = 200_000_000, 1 << 25
    _say(f"multirun sort full size: {n} rows, runs of {run_rows}, {merge} merge")
    keys = multirun_keys(n, dev, seed=1)
    job = MultiRunSort(keys, run_rows, key_bits=63, merge=merge)
    assert len(job.runs) == 6 and job.runs[-1][1] == 32_227_840
    job.step()
    ko, vo = job.result()
    verify_global_sort(keys, ko, vo)
    # a sampled window of the output equals the oracle's stable sort of the rows it holds
    lo = n // 2
    w = vo[lo:lo + 1_000_000].long()
    sel = keys[w].cpu().numpy()
    assert np.array_equal(ko[lo:lo + 1_000_000].cpu().numpy(), sel)
    del job, keys, ko, vo
    torch.cuda.empty_cache()


@pytest.mark.timeout(900)
def test_star_join_group_sum_full_size(dev):
    import torch

    from yallasql_amd import ops
    from yallasql_amd.datagen import star_schema

    n = 600_000_000
    _say(f"star join + group by, {n} fact rows")
    dims, m1, m2 = star_schema(n, dev, seed=2)
    sums, counts = ops.star_join_group_sum(dims, m1, m2, 7000)
    torch.cuda.synchronize()
    host_dims = []
    for d in dims:
        h = {"fk": d["fk"].cpu().numpy(), "rows": d["rows"]}
        for k in ("pass", "group"):
            if d.get(k) is not None:
                h[k] = d[k].cpu().numpy()
        for k in ("group_base", "group_card"):
            if k in d:
                h[k] = d[k]
        host_dims.append(h)
    m1h, m2h = m1.cpu().numpy(), m2.cpu().numpy()
    del dims, m1, m2
    torch.cuda.empty_cache()
    es, ec = O.star_join_group_sum(host_dims, m1h, m2h, 7000)
    gs, gc = sums.cpu().numpy(), counts.cpu().numpy().view(np.uint64)
    assert np.array_equal(gc, ec)
    assert np.array_equal(gs, es)
    # about 1/7 * 1/25 of the fact rows qualify, over all 7,000 groups
    q = int(ec.sum())
    assert 0.8 * n / 175 < q < 1.2 * n / 175
    assert (ec > 0).sum() == 7000


@pytest.mark.timeout(900)
def test_batched_4096_jobs(dev):
    import torch

    from yallasql_amd.autotune import load_sort_variants
    from yallasql_amd.batched import JobSet, config4_jobs

    sizes, kinds, key64, thr = config4_jobs()
    _say(f"batched 4,096 jobs: {int(np.sum(sizes))} rows")
    js = JobSet(sizes, kinds, key64, thr, dev, seed=4).create(sort_variants=load_sort_variants())
    try:
        outs = (js.o32, js.o64, js.ov, js.perm, js.counts, js.accs)
        js.clear_outputs()
        js.run_plain()
        torch.cuda.synchronize()
        plain = [t.clone() for t in outs]
        js.clear_outputs()
        js.run_batched()
        done, total = js.check()
        assert done == total == int(np.count_nonzero(sizes))
        for a, b in zip(plain, outs):
            assert torch.equal(a, b)
        js.clear_outputs()
        js.run_graph()
        torch.cuda.synchronize()
        for a, b in zip(plain, outs):
            assert torch.equal(a, b)
        del plain
        # a seeded sample of every kind and size bucket, plus each kind's largest job
        r = np.random.default_rng(12)
        sizes = np.asarray(sizes)
        pick = []
        edges = [1 << 10, 1 << 14, 1 << 18, 1 << 20, (1 << 22) + 1]
        for k in range(3):
            ids = np.nonzero(np.asarray(kinds) == k)[0]
            pick.append(int(ids[np.argmax(sizes[ids])]))
            for lo, hi in zip(edges[:-1], edges[1:]):
                b = ids[(sizes[ids] >= lo) & (sizes[ids] < hi)]
                if b.size:
                    pick += [int(x) for x in r.choice(b, size=min(3, b.size), replace=False)]
        assert max(sizes[pick]) > 4_000_000
        check_jobs_against_oracle(js, pick)
    finally:
        js.close()
