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
Tokenising restates the reference's vendored csv-parser v2.3.0
(vendor/csv-parser/csv.hpp) as csv::CSVReader(path) runs it with its default
format: the delimiter and header row guessed from the first 500,000 bytes
(:7368-7443), the file parsed in 10,000,000-byte chunks (MmapParser::next,
:7174-7201) by the parser's state machine (IBasicCSVParser::parse and
end_feed, :6990-7141), "" pairs undone as CSVRow::get_field does
(:7764-7791), rows of another length than the header skipped (IGNORE_ROW,
CSVReader::read_row :7599-7630).  The numeric prefix rules are restated with
regular expressions; the tests check them against values built by
construction, and the whole reader against the reference's own library
built from its sources (oracle/_ref) and the fixtures it wrote.
"""
import calendar
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
# strtof's hexadecimal form: 0x, hex digits with an optional point, optional p exponent
_HEX_RE = re.compile(_WS + r"([+-]?0[xX](?:[0-9a-fA-F]+\.?[0-9a-fA-F]*|\.[0-9a-fA-F]+)(?:[pP][+-]?[0-9]+)?)")
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
    h = _HEX_RE.match(s)
    if h:
        x = float.fromhex(h.group(1))
        m = h
    else:
        m = _FLT_RE.match(s)
        if not m:
            raise ConversionError(s)
        x = float(m.group(1))
    with np.errstate(over="ignore"):
        v = np.float32(x)
    if np.isinf(v) and "inf" not in m.group(1).lower():
        raise ConversionError(s)  # overflow: std::stof throws out_of_range
    if x != 0.0 and abs(v) < np.finfo(np.float32).tiny:  # (nan compares false)
        raise ConversionError(s)  # underflow to a subnormal or zero: strtof sets ERANGE, std::stof throws
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
    b = s if isinstance(s, bytes) else s.encode()
    cut = b.find(b"\0")
    if cut >= 0:
        b = b[:cut]
    b = b[:255]
    return b + b"\0" * (256 - len(b))


HEAD_BYTES = 500_000
CHUNK_BYTES = 10_000_000
DELIMS = b",|\t;^"
_BOM = b"\xef\xbb\xbf"


class _Chunk:
    """One run of the parser's state machine over `data` (one chunk) with
    delimiter `d`: rows of raw fields, and the offset of the first row left
    incomplete (the remainder the next chunk starts from)."""

    def __init__(self, data, d, skip_bom):
        self.data, self.d = data, d
        self.rows = []
        self.cur = []             # fields of the row being read
        self.row_start = 0
        self.fstart = None        # field start relative to row_start (None: not set)
        self.flen = 0
        self.pairs = False        # the field held a "" pair
        self.quoted = False
        self.pos = 3 if skip_bom and len(data) >= 3 and data[:3] == _BOM else 0
        self.unq = re.compile(b"[" + re.escape(bytes([d])) + b'\r\n"]')
        self.q = re.compile(b'"')

    def kind(self, i):
        c = self.data[i]
        if c == self.d:
            return "D"
        if c in (10, 13):
            return "N"
        if c == 34:
            return "Q"
        return "C"

    def push_field(self):
        s = (self.fstart or 0) + self.row_start
        raw = self.data[s:s + self.flen]
        if self.pairs:  # every second quote character of the field dropped
            out, prev = bytearray(), False
            for ch in raw:
                if ch == 34:
                    if prev:
                        prev = False
                        continue
                    prev = True
                out.append(ch)
            raw = bytes(out)
        self.cur.append(bytes(raw))
        self.fstart, self.flen, self.pairs = None, 0, False

    def push_row(self):
        self.rows.append(self.cur)
        self.cur = []

    def parse(self):
        data, n = self.data, len(self.data)
        while self.pos < n:
            k = self.kind(self.pos)
            if self.quoted and k in "DN":
                k = "C"
            if k == "D":
                self.push_field()
                self.pos += 1
            elif k == "N":
                self.pos += 1
                while self.pos < n and data[self.pos] in (10, 13):
                    self.pos += 1
                self.push_field()
                self.push_row()
                self.row_start = self.pos
            elif k == "C":
                if self.fstart is None:
                    self.fstart = self.pos - self.row_start
                m = (self.q if self.quoted else self.unq).search(data, self.pos)
                self.pos = m.start() if m else n
                self.flen = self.pos - (self.fstart + self.row_start)
            elif self.quoted:  # a quote inside a quoted field
                if self.pos + 1 == n:
                    return self.row_start
                nk = self.kind(self.pos + 1)
                if nk in "DN":
                    self.quoted = False
                    self.pos += 1
                elif nk == "Q":
                    self.pos += 2
                    self.flen += 2
                    self.pairs = True
                else:
                    self.flen += 1
                    self.pos += 1
            elif self.flen == 0:  # opens a quoted field
                self.quoted = True
                self.pos += 1
                if self.fstart is None and self.pos < n:
                    self.fstart = self.pos - self.row_start
            else:
                self.flen += 1
                self.pos += 1
        return self.row_start

    def end_feed(self):
        data = self.data
        last_empty = len(data) > 0 and data[-1] in (self.d, 34)
        if self.flen > 0 or last_empty:
            self.push_field()
        if self.cur:
            self.push_row()


def parse_rows(buf, d, chunk=CHUNK_BYTES):
    """Every row csv-parser reads from the bytes `buf` with delimiter `d`,
    chunk by chunk as MmapParser does (a row crossing a chunk end is read
    again from its start in the next chunk)."""
    rows, pos, first, size = [], 0, True, len(buf)
    while True:
        length = min(size - pos, chunk)
        ch = _Chunk(buf[pos:pos + length], d, first)
        rem = ch.parse()
        last = pos + length == size or size < chunk
        if last:
            ch.end_feed()
        rows.extend(ch.rows)
        if last:
            return rows
        if rem == 0:
            raise ConversionError("a record longer than one parse chunk")
        pos += rem
        first = False


class _TallyOrder:
    """The order in which csv-parser's row-length tally, a libstdc++
    std::unordered_map<size_t, size_t> built from {{0, 0}}, iterates its
    lengths (csv.hpp:7370-7414): the nodes form one list with each bucket's
    nodes together; a new node goes to the front of its bucket's nodes, or
    to the front of the list when its bucket is empty (hashtable.h
    _M_insert_bucket_begin); growing re-inserts the nodes in list order by
    the same rule (_M_rehash_aux).  The map starts with 13 buckets and grows
    through the counts below when an insert would exceed the count (counts
    read off such a map)."""

    GROW = (29, 59, 127, 257, 541, 1109, 2357, 5087, 10273, 20753)

    def __init__(self):
        self.nb, self.order, self.front = 13, [], {}
        self._put(0)

    def _put(self, k):
        b = k % self.nb
        at = self.order.index(self.front[b]) if b in self.front else 0
        self.order.insert(at, k)
        self.front[b] = k

    def add(self, k):
        if len(self.order) + 1 > self.nb:
            self.nb = next(c for c in self.GROW if c > self.nb)
            old, self.order, self.front = self.order, [], {}
            for x in old:
                self._put(x)
        self._put(k)


def guess_format(buf):
    """(delimiter, header row) as csv::CSVReader guesses them from the first
    500,000 bytes up to a NUL: per candidate delimiter, the score is the
    largest (row length x rows of that length), equal scores going to the
    length the tally iterates first; the delimiter with the strictly highest
    score wins (first in ',|\\t;^' on ties) and the header is the first row
    of its length."""
    head = buf[:HEAD_BYTES]
    cut = head.find(b"\0")
    if cut >= 0:
        head = head[:cut]
    best, delim, header = 0, DELIMS[0], 0
    for d in DELIMS:
        ch = _Chunk(head, d, True)
        ch.parse()
        ch.end_feed()
        count, first, tally = {0: 0}, {}, _TallyOrder()
        for i, r in enumerate(ch.rows):
            if len(r) not in first:
                first[len(r)] = i
                count[len(r)] = 0
                tally.add(len(r))
            count[len(r)] += 1
        score, hdr = 0, 0
        for length in tally.order:
            if length * count[length] > score:
                score, hdr = length * count[length], first[length]
        if score > best:
            best, delim, header = score, d, hdr
    return delim, header


def read_fields(path):
    """(header names, data rows of raw fields) as the reference's reader
    yields them."""
    with open(path, "rb") as f:
        buf = f.read()
    if not buf:
        raise ConversionError("no header row")
    d, header = guess_format(buf)
    rows = parse_rows(buf, d)
    if header >= len(rows):
        raise ConversionError("no header row")
    names = rows[header]
    return names, [r for r in rows[header + 1:] if len(r) == len(names)]


def read(path, names, types):
    """All rows of the requested columns: ([column arrays], [null byte arrays])."""
    header, recs = read_fields(path)
    idx = []
    for n in names:
        hits = [i for i, h in enumerate(header) if h == n.encode()]
        if not hits:
            raise ConversionError("no column named %r" % n)
        idx.append(hits[-1])  # a repeated name means its last occurrence
    cols = [[] for _ in names]
    nulls = [[] for _ in names]
    for rec in recs:
        for c, (i, t) in enumerate(zip(idx, types)):
            raw = rec[i]
            s = raw.decode("latin-1")
            nulls[c].append(1 if raw == b"" else 0)
            if t == INT:
                cols[c].append(-2**31 if raw == b"" else stoi(s.split("\0")[0]))
            elif t == FLOAT:
                cols[c].append(np.finfo(np.float32).tiny if raw == b"" else stof(s.split("\0")[0]))
            elif t == DATETIME:
                cols[c].append(0 if raw == b"" else datetime_seconds(s.split("\0")[0]))
            else:
                cols[c].append(np.frombuffer(string_bytes(raw), np.uint8))
    out = []
    for c, t in enumerate(types):
        if t == STRING:
            out.append(np.stack(cols[c]) if cols[c] else np.zeros((0, 256), np.uint8))
        else:
            out.append(np.array(cols[c], dtype={INT: np.int32, FLOAT: np.float32, DATETIME: np.int64}[t]))
    return out, [np.array(x, np.uint8) for x in nulls]


def write_rows(names, types, cols, nulls=None, header=True):
    """The bytes the reference's CsvWriter writes for one batch
    (include/engine/operators/csv_writer.hpp:24-77): names joined by ',' on
    the first line, then per row the fields joined by ',' and '\\n': NULL ->
    "", STRING -> "text up to the NUL" (not escaped), INT -> decimal, FLOAT ->
    fixed with 4 decimals, DATETIME -> "%Y-%m-%d %H:%M:%S" of the UTC time
    (getDateTimeStr, include/enums/data_type_utils.hpp:46-58).  Formatted
    with Python's own routines: '%.4f' of the float's exact value (correctly
    rounded, as std::to_chars is), time.strftime of time.gmtime."""
    import time

    out = []
    if header:
        out.append(",".join(names).encode() + b"\n")
    n = cols[0].shape[0] if cols else 0
    for r in range(n):
        fields = []
        for c, t in enumerate(types):
            if nulls is not None and nulls[c] is not None and nulls[c][r]:
                fields.append(b'""')
            elif t == STRING:
                raw = bytes(cols[c][r])
                cut = raw.find(b"\0")
                fields.append(b'"' + (raw if cut < 0 else raw[:cut]) + b'"')
            elif t == INT:
                fields.append(str(int(cols[c][r])).encode())
            elif t == FLOAT:
                v = np.float32(cols[c][r])
                if np.isnan(v):
                    fields.append(b"-nan" if np.signbit(v) else b"nan")
                else:
                    fields.append(("%.4f" % float(v)).encode())
            else:
                fields.append(b'"' + time.strftime("%Y-%m-%d %H:%M:%S", time.gmtime(int(cols[c][r]))).encode() + b'"')
        out.append(b",".join(fields) + b"\n")
    return b"".join(out)
