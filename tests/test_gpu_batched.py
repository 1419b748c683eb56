"""Batched executor, launch-chain graph form: the same ragged jobs replayed
as one captured graph of parallel launch chains and run as the per-job
launch loop give identical bits, every job matches the oracle, and graph
replays repeat the bits."""
import math

import numpy as np
import pytest

from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

def make_jobs(seed=0, extra=60, hi=300_000):
    r = np.random.default_rng(seed)
    edge = [0, 1, 100, 4095, 4096, 4097, 6144, 8193, 65535, 65536, 65537, 200_000]
    rand = list(np.floor(np.exp(r.uniform(0, math.log(hi), extra))).astype(int))
    # every edge size once per kind (sort, filter, aggregate), then random sizes
    sizes = [s for s in edge for _ in range(3)] + rand
    kinds = [j % 3 for j in range(len(sizes))]
    key64 = [(j // 3) % 2 == 1 for j in range(len(sizes))]
    thr = r.random(len(sizes)).astype(np.float32)
    return sizes, kinds, key64, thr


@pytest.fixture(scope="module")
def jobset(dev):
    from yallasql_amd.batched import JobSet

    sizes, kinds, key64, thr = make_jobs()
    js = JobSet(sizes, kinds, key64, thr, dev, seed=1).create()
    yield js
    js.close()


def test_graph_batch_equals_plain_bitwise(jobset):
    jobset.clear_outputs()
    jobset.run_plain()
    plain = jobset.snapshot()
    jobset.clear_outputs()
    jobset.run_graph()
    batched = jobset.snapshot()
    for a, b in zip(plain, batched):
        assert np.array_equal(a, b)
    # the captured graph replays to the same bits, with eager work and
    # synchronisations between the replays
    for _ in range(3):
        jobset.clear_outputs()
        jobset.run_plain()
        jobset.snapshot()
        jobset.clear_outputs()
        jobset.run_graph()
        again = jobset.snapshot()
        for a, b in zip(batched, again):
            assert np.array_equal(a, b)


def check_jobs_against_oracle(jobset, jobs=None):
    """Each job (all, or the indices in `jobs`) against the oracle: sorts by
    the stable argsort, filters by boolean indexing, aggregates exactly
    (float SUM within 1e-6 relative of math.fsum)."""
    from yallasql_amd.batched import AGG, FILTER, SORT  # noqa: F401

    for j in (range(len(jobset.sizes)) if jobs is None else jobs):
        keys, vals = jobset.job_inputs(j)
        kind, thr = jobset.kinds[j], np.float32(jobset.thr[j])
        ukeys = keys.view(np.uint64 if keys.itemsize == 8 else np.uint32)
        if kind == SORT:
            ko, perm = jobset.job_outputs(j)
            exp = O.stable_argsort(ukeys)
            assert np.array_equal(perm.view(np.uint32), exp), j
            assert np.array_equal(ko.view(ukeys.dtype), ukeys[exp]), j
        elif kind == FILTER:
            ko, vo, c = jobset.job_outputs(j)
            m = vals < thr
            assert c == int(m.sum()), j
            assert np.array_equal(ko, keys[m]) and np.array_equal(vo, vals[m]), j
        else:
            accs = jobset.job_outputs(j)
            m = vals < thr
            assert accs[0].count == int(m.sum()), j
            s = math.fsum(vals[m].astype(np.float64))
            assert abs(accs[0].d - s) <= 1e-6 * max(1.0, abs(s)), j
            if m.any():
                assert accs[1].i == int(keys[m].min()) and accs[2].i == int(keys[m].max()), j


def test_graph_batch_jobs_match_oracle(jobset):
    jobset.clear_outputs()
    jobset.run_graph()
    check_jobs_against_oracle(jobset)


def test_plain_loop_jobs_match_oracle(jobset):
    jobset.clear_outputs()
    jobset.run_plain()
    check_jobs_against_oracle(jobset)


def test_batch_info(jobset):
    info = jobset.info()
    assert info["jobs"] == len(jobset.sizes) and 1 <= info["chains"] <= 64


@pytest.mark.parametrize("seg_max", [0, 65536, 1 << 22])
def test_segmented_and_chained_paths_same_bits(dev, seg_max):
    """Every job through launch chains only (0), the default split, and one
    workgroup per job (4M): the per-job loop's bits each time."""
    from yallasql_amd.batched import JobSet

    sizes, kinds, key64, thr = make_jobs(seed=8, extra=30, hi=400_000)
    js = JobSet(sizes, kinds, key64, thr, dev, seed=3).create(seg_max=seg_max)
    try:
        js.clear_outputs()
        js.run_plain()
        plain = js.snapshot()
        js.clear_outputs()
        js.run_graph()
        for a, b in zip(plain, js.snapshot()):
            assert np.array_equal(a, b)
        check_jobs_against_oracle(js)
    finally:
        js.close()


def test_single_chain_equals_plain(dev):
    from yallasql_amd.batched import JobSet

    sizes, kinds, key64, thr = make_jobs(seed=5, extra=12, hi=50_000)
    js = JobSet(sizes, kinds, key64, thr, dev, seed=2).create(chains=1)
    try:
        js.clear_outputs()
        js.run_plain()
        plain = js.snapshot()
        js.clear_outputs()
        js.run_graph()
        for a, b in zip(plain, js.snapshot()):
            assert np.array_equal(a, b)
    finally:
        js.close()
