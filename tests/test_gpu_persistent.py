"""Batched executor, persistent form: one launch whose workgroups take whole
jobs from device-side job lists and chunk tasks of split sorts from a phase
queue (split sorts, filters and aggregates).  Every job completes (ysql_batch_check), the outputs equal the
per-job launch loop bit for bit and the oracle, across graph replays and
across the splitting choices (every sort split, none split, one-tile
chunks, workers that take whole jobs first or phase tasks first)."""
import numpy as np
import pytest

from tests.test_gpu_batched import check_jobs_against_oracle, make_jobs

pytestmark = pytest.mark.gpu


def _run_and_compare(js, replays=1):
    js.clear_outputs()
    js.run_plain()
    plain = js.snapshot()
    for _ in range(replays):
        js.clear_outputs()
        js.run_batched()
        done, total = js.check()
        assert done == total, (done, total)
        got = js.snapshot()
        for a, b in zip(plain, got):
            assert np.array_equal(a, b)
    return plain


@pytest.mark.parametrize("tuning", [
    {},
    {"split_min": 0, "chunk": 4096},
    {"split_min": 0, "split_filter_min": 0, "split_agg_min": 0, "chunk": 4096},
    {"split_filter_min": 4096, "split_agg_min": 4096, "chunk": 8192, "jobs_first_every": 2},
    {"split_filter_min": 100_000, "split_agg_min": 50_000, "chunk": 65536},
    {"split_min": 4096, "chunk": 8192, "jobs_first_every": 1},
    {"split_min": 1 << 30, "jobs_first_every": 0},
    {"split_min": 65536, "chunk": 65536, "jobs_first_every": 0, "workers_per_cu": 1},
])
def test_persistent_equals_plain_and_oracle(dev, tuning):
    from yallasql_amd.batched import JobSet

    sizes, kinds, key64, thr = make_jobs(seed=4, extra=40, hi=600_000)
    print(f"persistent batch, tuning {tuning}: {len(sizes)} jobs", flush=True)
    js = JobSet(sizes, kinds, key64, thr, dev, seed=6).create(**tuning)
    try:
        assert js.workers() >= 256
        _run_and_compare(js, replays=2)
        check_jobs_against_oracle(js)
    finally:
        js.close()


def test_persistent_replays_with_eager_work_between(dev):
    """Graph replays of the persistent launch, with other kernels and
    synchronisations between them, repeat the per-job loop's bits."""
    import torch

    from yallasql_amd.batched import JobSet

    sizes, kinds, key64, thr = make_jobs(seed=9, extra=20, hi=300_000)
    print(f"persistent batch replays: {len(sizes)} jobs", flush=True)
    js = JobSet(sizes, kinds, key64, thr, dev, seed=2).create(split_min=8192, chunk=16384)
    try:
        plain = _run_and_compare(js, replays=1)
        x = torch.rand(1 << 22, device=dev)
        for _ in range(3):
            (x * 2).sum().item()
            js.clear_outputs()
            js.run_batched()
            js.run_plain()  # same outputs written again by the loop, after the launch
            js.run_batched()
            assert js.check()[0] == js.check()[1]
            for a, b in zip(plain, js.snapshot()):
                assert np.array_equal(a, b)
    finally:
        js.close()


def test_persistent_empty_and_tiny_jobs(dev):
    from yallasql_amd.batched import JobSet

    sizes = [0, 0, 0, 1, 1, 1, 2, 3, 4096, 4097]
    kinds = [0, 1, 2, 0, 1, 2, 0, 1, 0, 0]
    key64 = [False, True, False, True, False, True, False, True, True, False]
    thr = np.linspace(0.1, 0.9, len(sizes)).astype(np.float32)
    print("persistent batch: empty and tiny jobs", flush=True)
    js = JobSet(sizes, kinds, key64, thr, dev, seed=5).create(split_min=1, split_filter_min=1, split_agg_min=1,
                                                              chunk=4096)
    try:
        _run_and_compare(js, replays=2)
        check_jobs_against_oracle(js)
    finally:
        js.close()
