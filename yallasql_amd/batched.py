"""Ragged job sets for the batched executor (ysql_batch_* in the C ABI).

A JobSet holds the inputs and outputs of many independent sort / filter /
aggregate jobs in a few large device buffers (one slice per job) and the
ysql_job table that points into them.  It runs the same jobs as one persistent
launch with a device work queue (run_batched), as the captured graph of
parallel launch chains (run_graph), or as the per-job launch loop on one
stream (run_plain).
"""
import ctypes as C

import numpy as np
import torch

from . import _lib as L

SORT, FILTER, AGG = L.JOB_SORT, L.JOB_FILTER, L.JOB_AGG


def config4_jobs(njobs=4096, seed=3, lo=1024, hi=4_194_304):
    """Config-4 job list: sizes log-uniform in [lo, hi], kind = j mod 3,
    key width by coin, per-job threshold uniform in [0, 1)."""
    r = np.random.default_rng(seed)
    u = r.random(njobs)
    sizes = np.floor(lo * (hi / lo) ** u).astype(np.int64)
    kinds = np.arange(njobs) % 3
    key64 = r.random(njobs) < 0.5
    thr = r.random(njobs).astype(np.float32)
    return sizes, kinds, key64, thr


class JobSet:
    def __init__(self, sizes, kinds, key64, thr, device, seed=0):
        self.dev = device
        self.sizes = [int(s) for s in sizes]
        self.kinds = [int(k) for k in kinds]
        self.key64 = [bool(k) for k in key64]
        self.thr = [float(t) for t in thr]
        n = len(self.sizes)
        off32, off64, offv = [], [], []
        t32 = t64 = tv = 0
        for j in range(n):
            if self.key64[j]:
                off64.append(t64)
                off32.append(-1)
                t64 += self.sizes[j]
            else:
                off32.append(t32)
                off64.append(-1)
                t32 += self.sizes[j]
            offv.append(tv)
            tv += self.sizes[j]
        g = torch.Generator(device=device)
        g.manual_seed(seed)
        # keys uniform and non-negative (same order signed or unsigned)
        self.k32 = torch.randint(0, 2**31, (max(t32, 1),), generator=g, device=device, dtype=torch.int32)
        hi = torch.randint(0, 2**31, (max(t64, 1),), generator=g, device=device, dtype=torch.int64)
        lo = torch.randint(0, 2**32, (max(t64, 1),), generator=g, device=device, dtype=torch.int64)
        self.k64 = (hi << 32) | lo
        del hi, lo
        self.v = torch.rand(max(tv, 1), generator=g, device=device, dtype=torch.float32)
        self.o32 = torch.empty_like(self.k32)
        self.o64 = torch.empty_like(self.k64)
        self.ov = torch.empty_like(self.v)
        self.perm = torch.empty(max(tv, 1), dtype=torch.int32, device=device)
        self.counts = torch.zeros(max(n, 1), dtype=torch.int64, device=device)
        self.accs = torch.zeros(max(n, 1) * 3 * 32, dtype=torch.uint8, device=device)
        self.off32, self.off64, self.offv = off32, off64, offv
        self.jobs = (L.Job * max(n, 1))()
        for j in range(n):
            kb = 8 if self.key64[j] else 4
            kt, ot = (self.k64, self.o64) if self.key64[j] else (self.k32, self.o32)
            ko = self.off64[j] if self.key64[j] else self.off32[j]
            self.jobs[j] = L.Job(
                self.kinds[j], L.UKEY64 if self.key64[j] else L.UKEY32, self.sizes[j],
                kt.data_ptr() + ko * kb, self.v.data_ptr() + offv[j] * 4, self.thr[j], 0,
                ot.data_ptr() + ko * kb, self.perm.data_ptr() + offv[j] * 4, self.ov.data_ptr() + offv[j] * 4,
                self.counts.data_ptr() + j * 8, self.accs.data_ptr() + j * 96)
        self.handle = None
        self.total_rows = sum(self.sizes)

    def create(self, chains=None, sort_variants=None, seg_max=None, **persist):
        """sort_variants: {32: [variant per log2 bucket], 64: [...]} (an
        autotuner's table); seg_max: rows up to which a job runs in one
        workgroup of a segmented kernel of the launch-chain graph (0 =
        never), one value for every kind or a (sort, filter, aggregate)
        triple; persist: fields of the persistent launch's tuning
        (split_min, chunk, jobs_first_every, idle_limit, workers_per_cu).
        None keeps the library defaults."""
        tuning = None
        if chains is not None or sort_variants is not None or seg_max is not None or persist:
            tuning = L.BatchTuning()
            L.lib.ysql_batch_tuning_default(C.byref(tuning))
            for k, v in persist.items():
                setattr(tuning, "persist_" + k, int(v))
            if chains is not None:
                tuning.chains = int(chains)
            if seg_max is not None:
                smax = tuple(seg_max) if isinstance(seg_max, (tuple, list)) else (seg_max,) * 3
                tuning.seg_sort_max, tuning.seg_filter_max, tuning.seg_agg_max = (int(x) for x in smax)
            for w, table in (sort_variants or {}).items():
                row = tuning.sort_variant[0 if int(w) == 32 else 1]
                for b, v in enumerate(table[:32]):
                    row[b] = int(v)
            tuning = C.byref(tuning)
        h = C.c_void_p()
        L.call("ysql_batch_create", self.jobs, len(self.sizes), tuning, C.byref(h))
        self.handle = h
        return self

    def info(self):
        t, gsz, ws = C.c_uint64(), C.c_uint32(), C.c_uint64()
        L.call("ysql_batch_info", self.handle, C.byref(t), C.byref(gsz), C.byref(ws))
        return {"jobs": t.value, "chains": gsz.value, "ws_bytes": ws.value}

    def run_batched(self):
        """The persistent launch (one kernel, device work queue)."""
        L.call("ysql_batch_run", self.handle, C.c_void_p(torch.cuda.current_stream().cuda_stream))

    def run_graph(self):
        """The launch-chain graph beside the segmented kernels."""
        L.call("ysql_batch_run_graph", self.handle, C.c_void_p(torch.cuda.current_stream().cuda_stream))

    def check(self):
        """(jobs completed, jobs with rows) of the last persistent launch
        (synchronises the current stream)."""
        d, t = C.c_uint32(), C.c_uint32()
        L.call("ysql_batch_check", self.handle, C.c_void_p(torch.cuda.current_stream().cuda_stream), C.byref(d),
               C.byref(t))
        return d.value, t.value

    def workers(self):
        w = C.c_uint32()
        L.call("ysql_batch_workers", self.handle, C.byref(w))
        return w.value

    def run_plain(self):
        L.call("ysql_batch_run_plain", self.handle, C.c_void_p(torch.cuda.current_stream().cuda_stream))

    def clear_outputs(self):
        for t in (self.o32, self.o64, self.ov, self.perm, self.counts, self.accs):
            t.fill_(0x5A if t.dtype == torch.uint8 else -7)

    def snapshot(self):
        """Every output buffer, copied to host (for bitwise comparisons)."""
        torch.cuda.synchronize()
        return [t.cpu().numpy().copy() for t in (self.o32, self.o64, self.ov, self.perm, self.counts, self.accs)]

    def job_inputs(self, j):
        n = self.sizes[j]
        if self.key64[j]:
            k = self.k64[self.off64[j]:self.off64[j] + n].cpu().numpy()
        else:
            k = self.k32[self.off32[j]:self.off32[j] + n].cpu().numpy()
        return k, self.v[self.offv[j]:self.offv[j] + n].cpu().numpy()

    def job_outputs(self, j):
        n = self.sizes[j]
        kind = self.kinds[j]
        ot = self.o64 if self.key64[j] else self.o32
        ko = self.off64[j] if self.key64[j] else self.off32[j]
        if kind == SORT:
            return ot[ko:ko + n].cpu().numpy(), self.perm[self.offv[j]:self.offv[j] + n].cpu().numpy()
        if kind == FILTER:
            c = int(self.counts[j].item())
            return ot[ko:ko + c].cpu().numpy(), self.ov[self.offv[j]:self.offv[j] + c].cpu().numpy(), c
        raw = self.accs[j * 96:(j + 1) * 96].cpu().numpy().tobytes()
        return [L.Acc.from_buffer_copy(raw[32 * s:32 * s + 32]) for s in range(3)]

    def algorithmic_bytes(self):
        """Sort: P*2*n*(kb+4); filter: n*(kb+4) in + sel*n*(kb+4) out; agg: n*(kb+4)."""
        total = 0
        for j in range(len(self.sizes)):
            n, kb = self.sizes[j], (8 if self.key64[j] else 4)
            if self.kinds[j] == SORT:
                total += (8 if self.key64[j] else 4) * 2 * n * (kb + 4)
            elif self.kinds[j] == FILTER:
                total += n * (kb + 4) * (1 + self.thr[j])
            else:
                total += n * (kb + 4)
        return total

    def close(self):
        if self.handle and L.lib is not None:
            L.lib.ysql_batch_destroy(self.handle)
        self.handle = None

    def __del__(self):
        self.close()
