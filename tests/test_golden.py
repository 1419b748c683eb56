"""Golden fixtures (tests/golden/*.npz, written by make_golden.py from the
oracle): the oracle must keep reproducing them (CPU), and the HIP kernels
must match them bit for bit (GPU)."""
import glob
import os

import numpy as np
import pytest
import torch

from oracle import pyoracle as O

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIXTURES = sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(HERE, "*.npz")))


def load(name):
    with np.load(os.path.join(HERE, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def test_fixtures_present():
    assert len(FIXTURES) >= 12


@pytest.mark.parametrize("name", FIXTURES)
def test_oracle_reproduces_fixture(name):
    import importlib.util

    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(HERE, "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    fresh = mg.cases()[name]
    stored = load(name)
    assert sorted(fresh) == sorted(stored)
    for k in stored:
        assert np.array_equal(np.asarray(fresh[k]), stored[k], equal_nan=stored[k].dtype.kind == "f"), k


def _T(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


@pytest.mark.gpu
@pytest.mark.parametrize("name", FIXTURES)
def test_kernels_match_fixture(name, dev):
    from yallasql_amd import ops

    f = load(name)
    if name.startswith("sort_"):
        desc = name.endswith("_desc")
        keys = ops.sortkey_encode(_T(f["x"], dev), descending=desc)
        perm = ops.radix_argsort(keys)
        assert np.array_equal(perm.cpu().numpy().view(np.uint32), f["perm"])
    elif name == "merge_u32":
        c, idx, isa, _ = ops.merge_path(O.UKEY32, False, _T(f["a"].view(np.int32), dev), _T(f["b"].view(np.int32), dev))
        assert np.array_equal(c.cpu().numpy().view(np.uint32), f["out0"])
        assert np.array_equal(idx.cpu().numpy().view(np.uint32), f["out1"])
        assert np.array_equal(isa.cpu().numpy(), f["out2"])
    elif name == "filter_gt10":
        m = ops.compare(2, _T(f["p"], dev), _T(np.array([10], np.int32), dev), rhs_scalar=True)
        assert np.array_equal(m.cpu().numpy(), f["mask"])
        [(out, _)], c = ops.compact_cols(m, [_T(f["vals"], dev)])
        assert np.array_equal(out.cpu().numpy(), f["out"])
    elif name == "reduce_int32":
        x = _T(f["x"], dev)
        for agg in range(6):
            acc = ops.acc_new(O.INT, agg, device=dev)
            ops.reduce(O.INT, agg, x, acc)
            got = ops.acc_read(acc)[0]
            exp_i, exp_c = f["i_count"][agg]
            assert got["count"] == exp_c and (agg in (4, 5) or got["i"] == exp_i), agg
    elif name.startswith("join_"):
        cmp = 4 if name == "join_eq" else 0
        pairs, c = ops.join_pairs(cmp, _T(f["l"], dev), _T(f["r"], dev))
        assert np.array_equal(pairs.cpu().numpy().view(np.uint32), f["pairs"])
    elif name == "group_sum_i64":
        accs = ops.acc_new(O.DATETIME, 2, nacc=37, device=dev)
        ops.group_reduce(O.DATETIME, 2, _T(f["gid"].view(np.int32), dev), 37, _T(f["v"], dev), accs)
        got = ops.acc_read(accs)
        assert np.array_equal(np.array([a["i"] for a in got]), f["sums"])
        assert np.array_equal(np.array([a["count"] for a in got]), f["counts"])
    else:
        pytest.fail("fixture without a GPU check: " + name)
