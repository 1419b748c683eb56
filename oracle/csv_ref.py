"""The reference's CSV scan, built from its own sources (oracle/_ref/libref_csv.so
from oracle/ref_csv.cpp + the reference's vendored csv-parser v2.3.0).

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product package.
available() is False where the reference tree was not there to build it.
"""
import ctypes as C
import os

import numpy as np

INT, FLOAT, DATETIME, STRING = 0, 1, 2, 3
_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "_ref", "libref_csv.so")
_lib = None


class RefCsvError(RuntimeError):
    pass


def available():
    return os.path.exists(LIB)


def _load():
    global _lib
    if _lib is None:
        _lib = C.CDLL(LIB)
        _lib.ref_csv_read.restype = C.c_int64
        _lib.ref_csv_read.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_char_p), C.POINTER(C.c_int), C.c_uint64,
                                      C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.c_char_p, C.c_int]
        _lib.ref_csv_dump.restype = C.c_int64
        _lib.ref_csv_dump.argtypes = [C.c_char_p, C.c_void_p, C.c_uint64, C.c_char_p, C.c_int]
    return _lib


def read(path, names, types, max_rows=None):
    """All rows of the named columns, as the reference's GetOperator converts
    them: ([column arrays], [null byte arrays])."""
    lib = _load()
    if max_rows is None:
        with open(path, "rb") as f:
            data = f.read()
        max_rows = data.count(b"\n") + data.count(b"\r") + 2  # every record ends at one of these, or the end
    cols = []
    for t in types:
        if t == STRING:
            cols.append(np.zeros((max_rows, 256), np.uint8))
        else:
            cols.append(np.zeros(max_rows, {INT: np.int32, FLOAT: np.float32, DATETIME: np.int64}[t]))
    nulls = [np.zeros(max_rows, np.uint8) for _ in types]
    n = len(names)
    err = C.create_string_buffer(256)
    rows = lib.ref_csv_read(str(path).encode(), n, (C.c_char_p * n)(*[x.encode() for x in names]),
                            (C.c_int * n)(*types), max_rows, (C.c_void_p * n)(*[a.ctypes.data for a in cols]),
                            (C.c_void_p * n)(*[a.ctypes.data for a in nulls]), err, 256)
    if rows < 0:
        raise RefCsvError(err.value.decode(errors="replace"))
    return [c[:rows] for c in cols], [x[:rows] for x in nulls]


def dump(path):
    """(delimiter byte, header names, rows of raw fields) as the reference's
    reader hands them out (rows of another length than the header already
    skipped); names and fields are bytes."""
    import struct

    lib = _load()
    err = C.create_string_buffer(256)
    need = lib.ref_csv_dump(str(path).encode(), None, 0, err, 256)
    if need < 0:
        raise RefCsvError(err.value.decode(errors="replace"))
    buf = C.create_string_buffer(max(int(need), 1))
    got = lib.ref_csv_dump(str(path).encode(), buf, need, err, 256)
    if got != need:
        raise RefCsvError("the reader changed its output between two calls")
    b = buf.raw[:need]
    pos = 0

    def u32():
        nonlocal pos
        v = struct.unpack_from("<I", b, pos)[0]
        pos += 4
        return v

    def field():
        nonlocal pos
        n = u32()
        pos += n
        return b[pos - n:pos]

    delim = u32()
    names = [field() for _ in range(u32())]
    rows = []
    while pos < len(b):
        rows.append([field() for _ in range(u32())])
    return delim, names, rows
