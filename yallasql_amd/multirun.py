"""Out-of-core multi-run sort (A11): the table is cut into runs, each run is
radix-sorted with 8-bit digits, then the sorted runs are merged: either
pairwise with merge path, level by level, as the reference's OrderOperator
does (order_operator.hpp:160-193 sorts each child batch into a run, then
merge_all_sorted merges runs (i, i+1) per level, carrying an odd run;
order_operator.cpp:116-220), or all at once by the k-way merge
(ysql_merge_runs), which moves every row once instead of once per level.

Keys are int64 (non-negative here, so their raw bits order like the
values); the payload is the uint32 row index.  Runs are written into
slices of three shared buffers, so a merge level never allocates.  The
result is one global stable sort: equal keys keep their row order, because
each run is stable, runs are cut in row order and merge path takes the
earlier run first on ties.
"""
import torch

from . import _lib as L
from . import ops


class MultiRunSort:
    """keys_in: int64 device tensor; run_rows: rows per run; key_bits: the
    significant low bits of the keys (8-bit digits cover them)."""

    def __init__(self, keys_in, run_rows, key_bits=63, merge="kway"):
        assert merge in ("kway", "pairwise")
        self.merge_mode = merge
        dev = keys_in.device
        self.n = n = keys_in.numel()
        self.keys_in = keys_in
        self.key_bits = key_bits
        self.runs = [(o, min(run_rows, n - o)) for o in range(0, n, run_rows)]
        self.bk = [torch.empty(n, dtype=torch.int64, device=dev) for _ in range(3)]
        self.bv = [torch.empty(n, dtype=torch.int32, device=dev) for _ in range(3)]
        self.alt_k = torch.empty(run_rows, dtype=torch.int64, device=dev)
        self.alt_v = torch.empty(run_rows, dtype=torch.int32, device=dev)
        self.sort_ws = ops.workspace(L.ws_bytes("ysql_radix_sort_ws_bytes", L.UKEY64, run_rows), dev)
        self.merge_ws = ops.workspace(L.ws_bytes("ysql_merge_ws_bytes", n, n, n), dev)
        offs = [o for o, _ in self.runs] + [n]
        if merge == "kway" and len(self.runs) > 1:
            import ctypes as C

            b = C.c_size_t()
            L.call("ysql_merge_runs_ws_bytes", L.UKEY64, (C.c_uint64 * len(offs))(*offs), len(self.runs), C.byref(b))
            self.merge_ws = ops.workspace(b.value, dev)
        self.run_offsets = offs
        self.passes = (key_bits + 7) // 8
        self.pass_events = []   # (start, end, rows) per digit-pass launch
        self.merge_events = []  # (start, end, rows) per merge
        self.record = False
        self.final = None

    def _ev(self):
        e = torch.cuda.Event(enable_timing=True)
        e.record()
        return e

    def sort_run(self, off, rows):
        """Run [off, off + rows) sorted into slice [off, off + rows) of buffer 0;
        the payload is the global row index (val_base = off)."""
        p = ops._p
        s = ops._s()
        kin = self.keys_in[off:off + rows]
        ko, vo = self.bk[0][off:off + rows], self.bv[0][off:off + rows]
        ka, va = self.alt_k[:rows], self.alt_v[:rows]
        ws = self.sort_ws
        L.call("ysql_radix_begin", L.UKEY64, 0, p(kin), rows, 0, self.key_bits, p(ws), ws.numel(), s)
        src_k, src_v = kin, None
        P = self.passes
        for q in range(P):
            to_out = ((P - 1 - q) & 1) == 0  # the last pass lands in buffer 0
            dk, dv = (ko, vo) if to_out else (ka, va)
            e0 = self._ev() if self.record else None
            L.call("ysql_radix_pass", L.UKEY64, 0, q, p(src_k), p(src_v), off, p(dk), p(dv), rows, 0, self.key_bits,
                   p(ws), ws.numel(), s)
            if self.record:
                self.pass_events.append((e0, self._ev(), rows))
            src_k, src_v = dk, dv
        return (0, off, rows)

    def merge(self, a, b):
        """Two adjacent runs (buffer, offset, rows) merged into the slice they
        cover of a buffer neither occupies."""
        (ba, oa, na), (bb, ob, nb) = a, b
        out = min({0, 1, 2} - {ba, bb})
        off = min(oa, ob)
        ok, ov = self.bk[out][off:off + na + nb], self.bv[out][off:off + na + nb]
        e0 = self._ev() if self.record else None
        ops.merge_pairs(self.bk[ba][oa:oa + na], self.bv[ba][oa:oa + na], self.bk[bb][ob:ob + nb],
                        self.bv[bb][ob:ob + nb], out_keys=ok, out_vals=ov, ws=self.merge_ws)
        if self.record:
            self.merge_events.append((e0, self._ev(), na + nb))
        return (out, off, na + nb)

    def step(self):
        runs = [self.sort_run(o, r) for o, r in self.runs]
        if self.merge_mode == "kway" and len(runs) > 1:
            # every run sits at its own offset of buffer 0: one merge into buffer 1
            e0 = self._ev() if self.record else None
            ops.merge_runs(self.bk[0], self.bv[0], self.run_offsets, out_keys=self.bk[1], out_vals=self.bv[1],
                           ws=self.merge_ws)
            if self.record:
                self.merge_events.append((e0, self._ev(), self.n))
            self.final = (1, 0, self.n)
            return
        while len(runs) > 1:
            nxt = [self.merge(runs[i], runs[i + 1]) for i in range(0, len(runs) - 1, 2)]
            if len(runs) % 2:
                nxt.append(runs[-1])
            runs = nxt
        self.final = runs[0]

    def step_streamed(self, host_keys, copy_stream):
        """step() with the keys arriving from pinned host memory, out of core:
        every run is copied into its slice of keys_in on copy_stream, and the
        current stream sorts a run as soon as its copy has landed, so run r
        sorts while run r+1 crosses PCIe.  The copies first wait for the
        current stream's earlier work (a previous step still reading
        keys_in)."""
        cur = torch.cuda.current_stream()
        copy_stream.wait_stream(cur)
        landed = []
        with torch.cuda.stream(copy_stream):
            for o, r in self.runs:
                self.keys_in[o:o + r].copy_(host_keys[o:o + r], non_blocking=True)
                e = torch.cuda.Event()
                e.record(copy_stream)
                landed.append(e)
        runs = []
        for (o, r), e in zip(self.runs, landed):
            cur.wait_event(e)
            runs.append(self.sort_run(o, r))
        if len(runs) > 1:
            ops.merge_runs(self.bk[0], self.bv[0], self.run_offsets, out_keys=self.bk[1], out_vals=self.bv[1],
                           ws=self.merge_ws)
            self.final = (1, 0, self.n)
        else:
            self.final = runs[0]

    def result(self):
        b, o, n = self.final
        return self.bk[b][o:o + n], self.bv[b][o:o + n]

    def algorithmic_bytes(self):
        """Sort: P * 2 * n * (8 + 4) per run; pairwise merge: 2 * (m + n) * 12
        per merge; k-way merge: 2 * n * 12 once."""
        b = sum(self.passes * 2 * r * 12 for _, r in self.runs)
        if self.merge_mode == "kway":
            return b + (2 * self.n * 12 if len(self.runs) > 1 else 0)
        sizes = [r for _, r in self.runs]
        while len(sizes) > 1:
            nxt = []
            for i in range(0, len(sizes) - 1, 2):
                b += 2 * (sizes[i] + sizes[i + 1]) * 12
                nxt.append(sizes[i] + sizes[i + 1])
            if len(sizes) % 2:
                nxt.append(sizes[-1])
            sizes = nxt
        return b


def verify_global_sort(keys_in, keys_out, vals_out):
    """Sorted, stable, a permutation of the rows, and every key matches its
    row: together these equal one global stable sort exactly (the oracle's
    closed form for the whole pipeline).  Raises AssertionError."""
    n = keys_in.numel()
    assert keys_out.numel() == n and vals_out.numel() == n
    d = keys_out[1:] - keys_out[:-1]
    assert bool((d >= 0).all()), "keys not sorted"
    ties = d == 0
    assert bool((vals_out[1:][ties] > vals_out[:-1][ties]).all()), "ties out of input order"
    del d, ties
    cnt = torch.zeros(n, dtype=torch.int32, device=keys_in.device)
    cnt.index_add_(0, vals_out.long(), torch.ones_like(vals_out))
    assert bool((cnt == 1).all()), "payload is not a permutation"
    del cnt
    assert bool((keys_in[vals_out.long()] == keys_out).all()), "keys do not match their rows"
