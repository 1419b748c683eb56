"""The serving dispatcher on the host (no GPU): a unit that raises returns
its slot and budget, the other units in flight finish, and serve() raises
the failure instead of waiting forever."""
import threading

import pytest
import torch

from yallasql_amd.serving import FILTER_AGG, JOIN_GROUP, TOPK, Query, Server


class HostTable:
    def __init__(self):
        self.device = torch.device("cpu")
        self.n = 1000


def _runners(fail_qid, seen):
    lock = threading.Lock()

    def run(table, qs):
        with lock:
            seen.extend(q.qid for q in qs)
        for q in qs:
            if q.qid == fail_qid:
                raise ValueError(f"query {q.qid} failed")
            q.result = q.qid
        return table.n

    return {FILTER_AGG: run, TOPK: run, JOIN_GROUP: run}


def test_failing_unit_releases_slot_and_raises():
    seen = []
    srv = Server(HostTable(), slots=2, hbm_budget=1 << 40, runners=_runners(3, seen))
    qs = [Query(i, TOPK, 0) for i in range(8)]
    try:
        with pytest.raises(RuntimeError) as ei:
            srv.serve(qs, realtime=False)
        assert isinstance(ei.value.__cause__, ValueError)
        # every slot and every reserved byte came back
        assert sorted(srv.idle) == [0, 1] and srv.outstanding == 0 and srv.free_bytes == 1 << 40
        # the server keeps working after the failure
        ok = [Query(100 + i, JOIN_GROUP, 0) for i in range(4)]
        srv.serve(ok, realtime=False)
        assert [q.result for q in ok] == [100, 101, 102, 103]
    finally:
        srv.close()
    assert 3 in seen


def test_all_units_succeed_on_host():
    seen = []
    srv = Server(HostTable(), slots=3, runners=_runners(-1, seen))
    qs = [Query(i, i % 3, 0) for i in range(12)]
    try:
        srv.serve(qs, realtime=False)
    finally:
        srv.close()
    assert sorted(seen) == list(range(12))
    assert all(q.result == q.qid for q in qs)
