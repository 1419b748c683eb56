"""CSV file -> column batches: the scan path's feeder over the C ABI
(include/ysql_csv.h), mirroring the reference's GetOperator
(src/engine/operators/get_operator.cpp:16-148): open with the header names
and types of the requested columns, then next() until it returns no rows."""
import ctypes as C

import numpy as np

from . import _lib as L

_NP = {L.INT: np.int32, L.FLOAT: np.float32, L.DATETIME: np.int64}


def batch_rows(types):
    """Rows per batch for these column types (config.hpp:28-38)."""
    t = (C.c_int * len(types))(*types)
    out = C.c_uint32()
    L.call("ysql_csv_batch_rows", len(types), t, C.byref(out))
    return out.value


_TYPESTR = {L.INT: "<i4", L.FLOAT: "<f4", L.DATETIME: "<i8", L.STRING: "|u1"}


class _DeviceView:
    """A device buffer of the reader, exposed to torch without a copy."""

    def __init__(self, ptr, shape, typestr):
        self.__cuda_array_interface__ = {"shape": shape, "typestr": typestr, "data": (ptr, False), "version": 2,
                                         "strides": None}


class CsvOptions(C.Structure):
    _fields_ = [("batch_rows", C.c_uint64), ("slots", C.c_uint32), ("async_copies", C.c_uint32)]


class CsvReader:
    def __init__(self, path, names, types, batch_rows_=0, slots=0, async_copies=False):
        """batch_rows_ / slots / async_copies: the scale mode of
        ysql_csv_open_ex (0 / 0 / False: the reference's sizing, 2 staging
        areas, copies on the caller's stream)."""
        self.types = list(types)
        self.rows = int(batch_rows_) if batch_rows_ else batch_rows(self.types)
        self.async_copies = bool(async_copies)
        n = len(names)
        cn = (C.c_char_p * n)(*[x.encode() for x in names])
        ct = (C.c_int * n)(*self.types)
        h = C.c_void_p()
        opt = CsvOptions(int(batch_rows_), int(slots), int(bool(async_copies)))
        L.call("ysql_csv_open_ex", str(path).encode(), n, cn, ct, C.byref(opt), C.byref(h))
        self.handle = h

    def _host_bufs(self):
        cols = []
        for t in self.types:
            if t == L.STRING:
                cols.append(np.zeros((self.rows, L.STRING_BYTES), np.uint8))
            else:
                cols.append(np.zeros(self.rows, _NP[t]))
        nulls = [np.zeros(self.rows, np.uint8) for _ in self.types]
        return cols, nulls

    def next_host(self):
        """The next batch as numpy arrays: ([column], [null bytes]); None at the end."""
        cols, nulls = self._host_bufs()
        n = len(cols)
        pc = (C.c_void_p * n)(*[a.ctypes.data for a in cols])
        pn = (C.c_void_p * n)(*[a.ctypes.data for a in nulls])
        rows = C.c_uint32()
        L.call("ysql_csv_next_host", self.handle, pc, pn, C.byref(rows))
        if rows.value == 0:
            return None
        return [a[: rows.value] for a in cols], [a[: rows.value] for a in nulls]

    def next_device(self, device, stream=None):
        """The next batch as device tensors ([column], [null bytes]), copied
        asynchronously on `stream` (default: the current stream); None at the end."""
        import torch

        tdt = {L.INT: torch.int32, L.FLOAT: torch.float32, L.DATETIME: torch.int64}
        cols = [torch.empty((self.rows, L.STRING_BYTES), dtype=torch.uint8, device=device) if t == L.STRING
                else torch.empty(self.rows, dtype=tdt[t], device=device) for t in self.types]
        nulls = [torch.empty(self.rows, dtype=torch.uint8, device=device) for _ in self.types]
        n = len(cols)
        pc = (C.c_void_p * n)(*[a.data_ptr() for a in cols])
        pn = (C.c_void_p * n)(*[a.data_ptr() for a in nulls])
        rows = C.c_uint32()
        s = stream if stream is not None else torch.cuda.current_stream(device)
        L.call("ysql_csv_next", self.handle, pc, pn, C.byref(rows), C.c_void_p(s.cuda_stream))
        if rows.value == 0:
            return None
        return [a[: rows.value] for a in cols], [a[: rows.value] for a in nulls]

    def next_async(self, stream=None):
        """The next batch in the reader's own device buffers (scale mode), as
        torch tensors viewing them: ([column], [null bytes]).  They hold the
        batch until the call `slots` calls later; queue all work reading them
        on `stream` (default: the current stream) before the next call.
        None at the end."""
        import torch

        n = len(self.types)
        pc = (C.c_void_p * n)()
        pn = (C.c_void_p * n)()
        rows = C.c_uint32()
        s = stream if stream is not None else torch.cuda.current_stream()
        L.call("ysql_csv_next_async", self.handle, pc, pn, C.byref(rows), C.c_void_p(s.cuda_stream))
        m = rows.value
        if m == 0:
            return None
        cols = [torch.as_tensor(_DeviceView(pc[i], (m, L.STRING_BYTES) if t == L.STRING else (m,), _TYPESTR[t]),
                                device=s.device) for i, t in enumerate(self.types)]
        nulls = [torch.as_tensor(_DeviceView(pn[i], (m,), "|u1"), device=s.device) for i in range(n)]
        return cols, nulls

    def close(self):
        if getattr(self, "handle", None) is not None and L.lib is not None:
            L.lib.ysql_csv_close(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class CsvWriter:
    """Result batches -> CSV file in the reference's CsvWriter format
    (include/engine/operators/csv_writer.hpp:24-77) over ysql_csv_writer_*."""

    def __init__(self, path, names, types):
        self.types = list(types)
        n = len(names)
        h = C.c_void_p()
        L.call("ysql_csv_writer_open", str(path).encode(), n, (C.c_char_p * n)(*[x.encode() for x in names]),
               (C.c_int * n)(*self.types), C.byref(h))
        self.handle = h

    def write_host(self, cols, nulls=None, rows=None):
        """cols: numpy arrays ((rows, 256) uint8 for strings); nulls: uint8 arrays or None."""
        n = len(cols)
        cols = [np.ascontiguousarray(c) for c in cols]
        nulls = None if nulls is None else [None if x is None else np.ascontiguousarray(x, np.uint8) for x in nulls]
        rows = cols[0].shape[0] if rows is None else rows
        pc = (C.c_void_p * n)(*[c.ctypes.data for c in cols])
        pn = None if nulls is None else (C.c_void_p * n)(*[None if x is None else x.ctypes.data for x in nulls])
        L.call("ysql_csv_writer_write_host", self.handle, pc, pn, rows)

    def write_device(self, cols, nulls=None, stream=None):
        """cols: device tensors; nulls: device uint8 tensors or None."""
        import torch

        n = len(cols)
        rows = cols[0].shape[0]
        pc = (C.c_void_p * n)(*[c.data_ptr() for c in cols])
        pn = None if nulls is None else (C.c_void_p * n)(*[None if x is None else x.data_ptr() for x in nulls])
        s = stream if stream is not None else torch.cuda.current_stream(cols[0].device)
        L.call("ysql_csv_writer_write", self.handle, pc, pn, rows, C.c_void_p(s.cuda_stream))

    def close(self):
        if getattr(self, "handle", None) is not None:
            h, self.handle = self.handle, None
            L.call("ysql_csv_writer_close", h)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
