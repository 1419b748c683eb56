"""ysql_filter_compact_noalloc (static scratch, no workspace argument)
against ysql_filter_compact with a caller workspace, through the C ABI:
two calls in a row, the second over fewer rows, each bit-identical to the
workspace form; and the row limit is refused before any device work."""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _cols(L, srcs, dsts):
    cols = (L.Col * len(srcs))()
    for i, (s, d) in enumerate(zip(srcs, dsts)):
        cols[i] = L.Col(s.data_ptr(), None, d.data_ptr(), None, s.element_size(), 0)
    return cols


def _compact(L, ops, noalloc, pred, thr, srcs, n):
    dev = pred.device
    dsts = [torch.full_like(s, -3) for s in srcs]
    cnt = torch.full((1,), -1, dtype=torch.int64, device=dev)
    cols = _cols(L, srcs, dsts)
    if noalloc:
        L.call("ysql_filter_compact_noalloc", L.FLOAT, L.LT, ops._p(pred), ops._p(thr), cols, len(srcs), n,
               ops._p(cnt), ops._s())
    else:
        ws = ops.workspace(L.ws_bytes("ysql_compact_ws_bytes", n), dev)
        L.call("ysql_filter_compact", L.FLOAT, L.LT, ops._p(pred), ops._p(thr), cols, len(srcs), n, ops._p(cnt),
               ops._p(ws), ws.numel(), ops._s())
    torch.cuda.synchronize()
    m = int(cnt.item())
    return m, [d[:m].cpu().numpy() for d in dsts], [d[m:].cpu().numpy() for d in dsts]


def test_noalloc_equals_workspace_form_twice(dev):
    from yallasql_amd import _lib as L
    from yallasql_amd import ops

    g = torch.Generator(device=dev)
    g.manual_seed(21)
    for n in (3_000_017, 1_234_567):  # the second call has fewer rows
        print(f"noalloc compaction: n={n}", flush=True)
        vals = torch.rand(n, generator=g, device=dev)
        keys = torch.randint(-2**62, 2**62, (n,), generator=g, device=dev, dtype=torch.int64)
        thr = torch.tensor([0.37], device=dev)
        m0, a0, t0 = _compact(L, ops, False, vals, thr, [keys, vals], n)
        m1, a1, t1 = _compact(L, ops, True, vals, thr, [keys, vals], n)
        assert m0 == m1 == int((vals < 0.37).sum().item())
        for x, y in zip(a0, a1):
            assert np.array_equal(x, y)
        for y in t1:  # nothing written past the selected rows
            assert (y == y.dtype.type(-3)).all()
        hv = vals.cpu().numpy()
        assert np.array_equal(a1[1], hv[hv < np.float32(0.37)])


def test_noalloc_refuses_oversized_input(dev):
    from yallasql_amd import _lib as L

    cols = (L.Col * 1)()
    cols[0] = L.Col(256, None, 512, None, 4, 0)
    rc = L.lib.ysql_filter_compact_noalloc(L.FLOAT, L.LT, C.c_void_p(256), C.c_void_p(256), cols, 1,
                                           (1 << 32) + 1, C.c_void_p(768), None)
    assert rc == L.EWORKSPACE
    assert b"2^32" in L.lib.ysql_last_error()
