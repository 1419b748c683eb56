"""Python handle on the host engine (include/ysql_engine.h).

    ctx = Context()
    t = ctx.table("t", id=ids, price=prices)          # numpy columns -> HBM
    p = ctx.scan(t, batch_rows=1 << 18)
    p = p.filter(cmp(GT, ref(1), f32(500.0))).order([1, 0], [True, False])
    cols, nulls = p.execute()                          # numpy columns

Expression helpers build ysql_expr trees; plan methods consume their inputs.
"""
import ctypes as C
import weakref

import numpy as np

from . import _lib as L

lib = L.lib
P = C.c_void_p

_sig = {
    "ysql_ctx_create": ([C.c_uint64, P], P),
    "ysql_ctx_destroy": ([P], None),
    "ysql_ctx_peak_bytes": ([P], C.c_uint64),
    "ysql_ctx_set_spill": ([P, C.c_uint64, C.c_char_p], C.c_int),
    "ysql_ctx_set_fusion": ([P, C.c_int], C.c_int),
    "ysql_ctx_cache_stats": ([P, P], C.c_int),
    "ysql_table_create": ([P, C.c_char_p], P),
    "ysql_table_add_column": ([P, C.c_char_p, C.c_int, P, P, C.c_uint64], C.c_int),
    "ysql_table_add_device_column": ([P, C.c_char_p, C.c_int, P, P, C.c_uint64], C.c_int),
    "ysql_table_rows": ([P], C.c_uint64),
    "ysql_table_destroy": ([P], None),
    "ysql_expr_ref": ([C.c_int], P),
    "ysql_expr_i32": ([C.c_int32], P),
    "ysql_expr_f32": ([C.c_float], P),
    "ysql_expr_i64": ([C.c_int64], P),
    "ysql_expr_cmp": ([C.c_int, P, P], P),
    "ysql_expr_logic": ([C.c_int, P, P], P),
    "ysql_expr_arith": ([C.c_int, P, P], P),
    "ysql_expr_cast": ([C.c_int, P], P),
    "ysql_expr_destroy": ([P], None),
    "ysql_plan_scan": ([P, P, C.c_uint64], P),
    "ysql_plan_filter": ([P, P], P),
    "ysql_plan_project": ([P, C.POINTER(P), C.c_int], P),
    "ysql_plan_order": ([P, C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_int], P),
    "ysql_plan_execute_csv": ([P, C.c_char_p, C.POINTER(C.c_uint64)], C.c_int),
    "ysql_plan_aggregate": ([P, C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_int),
                             C.POINTER(C.c_int32), C.POINTER(C.c_uint32), C.c_int], P),
    "ysql_plan_join": ([P, P, C.c_int, C.c_int, C.c_int], P),
    "ysql_plan_cross": ([P, P], P),
    "ysql_plan_destroy": ([P], None),
    "ysql_plan_execute": ([P, C.POINTER(P)], C.c_int),
    "ysql_result_ncols": ([P], C.c_int),
    "ysql_result_rows": ([P], C.c_uint64),
    "ysql_result_batches": ([P], C.c_uint64),
    "ysql_result_dtype": ([P, C.c_int], C.c_int),
    "ysql_result_data": ([P, C.c_int], P),
    "ysql_result_nulls": ([P, C.c_int], P),
    "ysql_result_destroy": ([P], None),
    "ysql_engine_last_error": ([], C.c_char_p),
}
for _n, (_a, _r) in _sig.items():
    _f = getattr(lib, _n)
    _f.argtypes = _a
    _f.restype = _r

NP_DT = {L.INT: np.int32, L.FLOAT: np.float32, L.DATETIME: np.int64}
DT_OF = {np.dtype(np.int32): L.INT, np.dtype(np.float32): L.FLOAT, np.dtype(np.int64): L.DATETIME}


class EngineError(RuntimeError):
    pass


def _chk_ptr(p, what):
    if not p:
        raise EngineError(f"{what}: {lib.ysql_engine_last_error().decode()}")
    return p


class Expr:
    def __init__(self, h):
        self.h = _chk_ptr(h, "expression")

    def _take(self):
        h, self.h = self.h, None
        if h is None:
            raise EngineError("expression used twice")
        return h


def ref(i):
    return Expr(lib.ysql_expr_ref(i))


def i32(v):
    return Expr(lib.ysql_expr_i32(int(v)))


def f32(v):
    return Expr(lib.ysql_expr_f32(float(v)))


def i64(v):
    return Expr(lib.ysql_expr_i64(int(v)))


def cmp(op, a, b):
    return Expr(lib.ysql_expr_cmp(op, a._take(), b._take()))


def logic(op, a, b):
    return Expr(lib.ysql_expr_logic(op, a._take(), b._take()))


def arith(op, a, b):
    return Expr(lib.ysql_expr_arith(op, a._take(), b._take()))


def cast(dtype, e):
    return Expr(lib.ysql_expr_cast(dtype, e._take()))


class Plan:
    def __init__(self, ctx, h):
        self.ctx = ctx
        self.h = _chk_ptr(h, "plan")
        ctx._plans.add(self)  # destroyed before the context they run in

    def _take(self):
        h, self.h = self.h, None
        if h is None:
            raise EngineError("plan node used twice")
        return h

    def filter(self, pred):
        return Plan(self.ctx, lib.ysql_plan_filter(self._take(), pred._take()))

    def project(self, exprs):
        arr = (P * len(exprs))(*[e._take() for e in exprs])
        return Plan(self.ctx, lib.ysql_plan_project(self._take(), arr, len(exprs)))

    def order(self, cols, desc=None):
        desc = desc or [0] * len(cols)
        return Plan(self.ctx, lib.ysql_plan_order(self._take(), (C.c_int * len(cols))(*cols),
                                                  (C.c_int * len(cols))(*[int(d) for d in desc]), len(cols)))

    def aggregate(self, aggs, groups=()):
        """aggs: [(agg, col)]; groups: [(col, base, card)]."""
        a = (C.c_int * len(aggs))(*[x[0] for x in aggs])
        c = (C.c_int * len(aggs))(*[x[1] for x in aggs])
        ng = len(groups)
        gc = (C.c_int * max(ng, 1))(*[g[0] for g in groups])
        gb = (C.c_int32 * max(ng, 1))(*[g[1] for g in groups])
        gk = (C.c_uint32 * max(ng, 1))(*[g[2] for g in groups])
        return Plan(self.ctx, lib.ysql_plan_aggregate(self._take(), a, c, len(aggs), gc, gb, gk, ng))

    def join(self, right, lcol, rcol, op=L.EQ):
        return Plan(self.ctx, lib.ysql_plan_join(self._take(), right._take(), lcol, rcol, op))

    def cross(self, right):
        return Plan(self.ctx, lib.ysql_plan_cross(self._take(), right._take()))

    def execute_csv(self, path):
        """Runs the plan to exhaustion, writing each result batch to the CSV
        file `path` as it comes (the reference's CsvWriter format); returns
        the number of rows written."""
        rows = C.c_uint64()
        rc = lib.ysql_plan_execute_csv(self.h, str(path).encode(), C.byref(rows))
        if rc != 0:
            raise EngineError(f"execute_csv returned {rc}: {lib.ysql_engine_last_error().decode()}")
        return rows.value

    def execute(self):
        """Runs the plan to exhaustion; returns (columns, nulls, batches) as
        numpy arrays.  The arrays are views of the result's pinned host
        memory (no copy); it is released when the last of them is."""
        out = P()
        rc = lib.ysql_plan_execute(self.h, C.byref(out))
        if rc != 0:
            raise EngineError(f"execute returned {rc}: {lib.ysql_engine_last_error().decode()}")
        holder = _Result(out)
        n = lib.ysql_result_rows(out)
        cols, nulls = [], []
        for c in range(lib.ysql_result_ncols(out)):
            dt = lib.ysql_result_dtype(out, c)
            if dt == L.STRING:
                cols.append(_host_view(holder, lib.ysql_result_data(out, c), n * 256, np.uint8).reshape(n, 256))
            else:
                cols.append(_host_view(holder, lib.ysql_result_data(out, c), n * np.dtype(NP_DT[dt]).itemsize,
                                       NP_DT[dt]))
            nulls.append(_host_view(holder, lib.ysql_result_nulls(out, c), n, np.uint8))
        return cols, nulls, lib.ysql_result_batches(out)

    def __del__(self):
        if getattr(self, "h", None):
            lib.ysql_plan_destroy(self.h)
            self.h = None


class _Result:
    """Owns a ysql_result; the numpy views of its memory keep it alive."""

    def __init__(self, h):
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            lib.ysql_result_destroy(self.h)
            self.h = None


def _host_view(holder, ptr, nbytes, dtype):
    """A numpy array over nbytes of result memory at ptr, holding `holder`."""
    if not nbytes:
        return np.zeros(0, dtype)
    buf = (C.c_uint8 * nbytes).from_address(ptr)
    buf._owner = holder
    return np.frombuffer(buf, dtype=dtype)


class Table:
    def __init__(self, ctx, h, keep):
        self.ctx, self.h, self._keep = ctx, h, keep

    def __del__(self):
        if getattr(self, "h", None):
            lib.ysql_table_destroy(self.h)
            self.h = None


class CacheStats(C.Structure):
    _fields_ = [(f, C.c_uint64) for f in ("hbm_bytes", "host_bytes", "file_bytes", "peak_hbm_bytes", "evictions",
                                           "spills", "restores_from_host", "restores_from_file")]


class Context:
    def __init__(self, hbm_budget=0, stream=None, host_budget=0, spill_dir=None, fusion=True):
        """hbm_budget: bytes of cached batches kept in HBM (0 = unlimited);
        past it the least recently used go to pinned host memory, and past
        host_budget to files in spill_dir (None: no file tier).  fusion:
        predicate and projection trees run as one fused program per batch
        (False: node by node, one kernel per expression node)."""
        self.h = _chk_ptr(lib.ysql_ctx_create(hbm_budget, stream), "context")
        self._tables = []
        self._plans = weakref.WeakSet()
        if host_budget or spill_dir:
            lib.ysql_ctx_set_spill(self.h, host_budget, None if spill_dir is None else str(spill_dir).encode())
        if not fusion:
            lib.ysql_ctx_set_fusion(self.h, 0)

    def cache_stats(self):
        st = CacheStats()
        lib.ysql_ctx_cache_stats(self.h, C.byref(st))
        return {f: getattr(st, f) for f, _ in CacheStats._fields_}

    def table(self, name, columns, nulls=None):
        """columns: list of (name, numpy array | (n, 256) uint8 strings)."""
        h = _chk_ptr(lib.ysql_table_create(self.h, name.encode()), "table")
        keep = []
        for i, (cn, arr) in enumerate(columns):
            arr = np.ascontiguousarray(arr)
            dt = L.STRING if (arr.dtype == np.uint8 and arr.ndim == 2) else DT_OF[arr.dtype]
            nb = None if nulls is None or nulls[i] is None else np.ascontiguousarray(nulls[i], dtype=np.uint8)
            keep += [arr, nb]
            rc = lib.ysql_table_add_column(h, cn.encode(), dt, arr.ctypes.data_as(P),
                                           None if nb is None else nb.ctypes.data_as(P), arr.shape[0])
            if rc != 0:
                raise EngineError(lib.ysql_engine_last_error().decode())
        t = Table(self, h, keep)
        self._tables.append(t)
        return t

    def device_table(self, name, columns, nulls=None):
        """columns: list of (name, torch device tensor) borrowed by the table."""
        h = _chk_ptr(lib.ysql_table_create(self.h, name.encode()), "table")
        keep = []
        for i, (cn, t) in enumerate(columns):
            dt = L.STRING if t.dim() == 2 else {4: (L.FLOAT if t.is_floating_point() else L.INT), 8: L.DATETIME}[t.element_size()]
            nb = None if nulls is None else nulls[i]
            keep += [t, nb]
            rc = lib.ysql_table_add_device_column(h, cn.encode(), dt, P(t.data_ptr()),
                                                  None if nb is None else P(nb.data_ptr()), t.shape[0])
            if rc != 0:
                raise EngineError(lib.ysql_engine_last_error().decode())
        t = Table(self, h, keep)
        self._tables.append(t)
        return t

    def scan(self, table, batch_rows=0):
        return Plan(self, lib.ysql_plan_scan(self.h, table.h, batch_rows))

    def peak_bytes(self):
        return lib.ysql_ctx_peak_bytes(self.h)

    def close(self):
        for p in list(self._plans):  # a plan's operators hold the context's stream and cache
            p.__del__()
        for t in self._tables:
            t.__del__()
        self._tables = []
        if self.h:
            lib.ysql_ctx_destroy(self.h)
            self.h = None
