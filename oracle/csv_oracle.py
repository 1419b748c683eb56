"""CPU restatement of the reference's CSV -> column batch feeder.

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product package.

Follows GetOperator (src/engine/operators/get_operator.cpp:16-148), the batch
sizing of include/config.hpp:28-38, getDateTime (include/enums/data_type_utils.hpp:24-44)
and String::set (include/kernels/string_kernel.hpp:16-30):
  * header row names the columns; fields by name (get_operator.cpp:68-72);
  * empty field -> NULL (null byte 1) with INT_MIN / FLT_MIN / 0 / zero string;
  * INT: std::stoi (leading blanks, optional sign, digits; trailing text ignored);
  * FLOAT: std::stof (the longest decimal / inf / nan prefix, rounded to float32);
  * DATETIME: strptime "%Y-%m-%d %H:%M:%S" as UTC -> epoch seconds, 0 if it does not parse;
  * STRING: first 255 bytes (up to a NUL), NUL-padded to 256.
Tokenising uses Python's csv module (RFC 4180 quoting).  The numeric prefix
rules are restated with regular expressions; the tests check them against
values built by construction.
"""
import calendar
import csv
import re

import numpy as np

INT, FLOAT, DATETIME, STRING = 0, 1, 2, 3
_BYTES = {INT: 4, FLOAT: 4, DATETIME: 8, STRING: 256}
MAX_BYTES_PER_BATCH = 4 << 20
ALIGN = 256

_WS = "[ \t\n\v\f\r]*"
_INT_RE = re.compile(_WS + r"([+-]?[0-9]+)")
_FLT_RE = re.compile(_WS + r"([+-]?(?:(?:[0-9]+\.?[0-9]*|\.[0-9]+)(?:[eE][+-]?[0-9]+)?|inf(?:inity)?|nan))",
                     re.IGNORECASE)
_DT_RE = re.compile(r"\s*([0-9]{1,4})-([0-9]{1,2})-([0-9]{1,2})\s*([0-9]{1,2}):([0-9]{1,2}):([0-9]{1,2})")


class ConversionError(ValueError):
    pass


def batch_rows(types):
    """config.hpp:28-38: floor(4 MiB / row bytes) rounded down to 256, at least 256."""
    row = sum(_BYTES[t] for t in types)
    aligned = (MAX_BYTES_PER_BATCH // row) // ALIGN * ALIGN
    return max(ALIGN, aligned)


def stoi(s):
    m = _INT_RE.match(s)
    if not m:
        raise ConversionError(s)
    v = int(m.group(1))
    if v < -2**31 or v > 2**31 - 1:
        raise ConversionError(s)
    return v


def stof(s):
    m = _FLT_RE.match(s)
    if not m:
        raise ConversionError(s)
    v = np.float32(float(m.group(1)))
    if np.isinf(v) and "inf" not in m.group(1).lower():
        raise ConversionError(s)  # overflow: std::stof throws out_of_range
    return v


def datetime_seconds(s):
    m = _DT_RE.match(s)
    if not m:
        return 0
    y, mo, d, h, mi, se = (int(x) for x in m.groups())
    if not (1 <= mo <= 12 and 1 <= d <= 31 and h <= 23 and mi <= 59 and se <= 61):
        return 0
    return calendar.timegm((y, mo, d, h, mi, se, 0, 0, 0))


def string_bytes(s):
    b = s.encode()
    cut = b.find(b"\0")
    if cut >= 0:
        b = b[:cut]
    b = b[:255]
    return b + b"\0" * (256 - len(b))


def read(path, names, types):
    """All rows of the requested columns: ([column arrays], [null byte arrays])."""
    with open(path, newline="") as f:
        rd = csv.reader(f)
        header = next(rd)
        idx = [header.index(n) for n in names]
        cols = [[] for _ in names]
        nulls = [[] for _ in names]
        for rec in rd:
            if not rec or (len(rec) == 1 and rec[0] == "" and len(header) > 1):
                continue
            if len(rec) < len(header):
                raise ConversionError("short record")
            for c, (i, t) in enumerate(zip(idx, types)):
                s = rec[i]
                nulls[c].append(1 if s == "" else 0)
                if t == INT:
                    cols[c].append(-2**31 if s == "" else stoi(s))
                elif t == FLOAT:
                    cols[c].append(np.finfo(np.float32).tiny if s == "" else stof(s))
                elif t == DATETIME:
                    cols[c].append(0 if s == "" else datetime_seconds(s))
                else:
                    cols[c].append(np.frombuffer(string_bytes(s), np.uint8))
    out = []
    for c, t in enumerate(types):
        if t == STRING:
            out.append(np.stack(cols[c]) if cols[c] else np.zeros((0, 256), np.uint8))
        else:
            out.append(np.array(cols[c], dtype={INT: np.int32, FLOAT: np.float32, DATETIME: np.int64}[t]))
    return out, [np.array(x, np.uint8) for x in nulls]
