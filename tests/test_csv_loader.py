"""CSV -> column batches (include/ysql_csv.h): batch sizing in closed form,
parsed columns against values built by construction, against the fixtures
the reference's own csv-parser wrote (tests/golden/csv), against the
oracle's restatement of that parser and, where oracle/_ref is built, against
the reference's parser itself on seeded random files; the 10,000,000-byte
parse-chunk records; the error paths; and (GPU) the pinned, overlapped
device path against the reference's parser."""
import calendar
import importlib.util
import os
import time

import numpy as np
import pytest

from oracle import csv_oracle as O
from oracle import csv_ref as R

_GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "csv")
_spec = importlib.util.spec_from_file_location("make_csv_golden", os.path.join(_GOLDEN, "make_csv_golden.py"))
G = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(G)
CASES = G.load()

INT, FLOAT, DATETIME, STRING = 0, 1, 2, 3


@pytest.fixture(scope="module")
def loader():
    from yallasql_amd import loader as _loader

    return _loader


def _quote(s):
    return '"' + s.replace('"', '""') + '"' if any(ch in s for ch in ',"\n\r') else s


def _gen(path, n, seed, crlf=False):
    """A CSV with INT, FLOAT, DATETIME, STRING columns (plus an ignored one)
    and the values it holds, 5% of fields empty."""
    r = np.random.default_rng(seed)
    ints = r.integers(-2**31, 2**31, n, dtype=np.int64).astype(np.int32)
    ints[:2] = [-2**31 + 1, 2**31 - 1]
    flts = (r.standard_normal(n) * 10.0 ** r.integers(-8, 9, n)).astype(np.float32)
    secs = r.integers(0, 4_000_000_000, n, dtype=np.int64)
    alphabet = np.array(list('abcXYZ 019,";-'))
    strs = ["".join(r.choice(alphabet, int(r.integers(0, 300)))) for _ in range(n)]
    # a quoted newline in one string of 50: the reference's parser guesses
    # the delimiter from line counts, and a file whose quoted fields hold
    # more newlines than records would read as one '|'-separated column
    strs = [s[:len(s) // 2] + "\n" + s[len(s) // 2:] if i % 50 == 7 else s for i, s in enumerate(strs)]
    empty = r.random((n, 4)) < 0.05
    eol = "\r\n" if crlf else "\n"
    with open(path, "w", newline="") as f:
        f.write("extra,id,price,ts,name" + eol)
        for i in range(n):
            fi = "" if empty[i, 0] else str(int(ints[i]))
            ff = "" if empty[i, 1] else "%.9g" % float(flts[i])
            ft = "" if empty[i, 2] else time.strftime("%Y-%m-%d %H:%M:%S", time.gmtime(int(secs[i])))
            fs = "" if empty[i, 3] or strs[i] == "" else _quote(strs[i])
            f.write(",".join(["x%d" % i, fi, ff, ft, fs]) + eol)
    nul = empty.copy()
    nul[:, 3] |= np.array([s == "" for s in strs])
    return ints, flts, secs, strs, nul


def _read_all(rd):
    cols, nulls = None, None
    while True:
        b = rd.next_host()
        if b is None:
            break
        c, nl = b
        cols = c if cols is None else [np.concatenate([a, x]) for a, x in zip(cols, c)]
        nulls = nl if nulls is None else [np.concatenate([a, x]) for a, x in zip(nulls, nl)]
    return cols, nulls


def test_batch_rows_closed_form(loader):
    for types, rows in (([INT, FLOAT], 524288), ([STRING], 16384), ([DATETIME, STRING, INT], 15616),
                        ([STRING] * 8, 2048), ([STRING] * 20, 768), ([STRING] * 70, 256)):
        assert loader.batch_rows(types) == rows == O.batch_rows(types)


@pytest.mark.parametrize("crlf", [False, True])
def test_values_by_construction(loader, tmp_path, crlf):
    n = 40_000  # 3 batches of 15,360 rows for this schema (272 B per row)
    p = tmp_path / "t.csv"
    ints, flts, secs, strs, nul = _gen(p, n, 5 + crlf, crlf)
    names, types = ["id", "price", "ts", "name"], [INT, FLOAT, DATETIME, STRING]
    rd = loader.CsvReader(p, names, types)
    assert rd.rows == 15360
    cols, nulls = _read_all(rd)
    rd.close()
    for c in range(4):
        assert np.array_equal(nulls[c], nul[:, c].astype(np.uint8)), c
    e0 = nul[:, 0]
    assert np.array_equal(cols[0][~e0], ints[~e0]) and np.all(cols[0][e0] == -2**31)
    e1 = nul[:, 1]
    assert np.array_equal(cols[1][~e1].view(np.uint32), flts[~e1].view(np.uint32))
    assert np.all(cols[1][e1] == np.finfo(np.float32).tiny)
    e2 = nul[:, 2]
    assert np.array_equal(cols[2][~e2], secs[~e2]) and np.all(cols[2][e2] == 0)
    for i in range(n):
        b = b"" if nul[i, 3] else strs[i].encode()[:255]
        assert bytes(cols[3][i][: len(b)]) == b and not cols[3][i][len(b):].any(), i
    # the oracle's restatement agrees on the same file
    oc, on = O.read(p, names, types)
    for c in range(4):
        assert np.array_equal(cols[c], oc[c]) and np.array_equal(nulls[c], on[c]), c


def test_reference_rules_edge_cases(loader, tmp_path):
    p = tmp_path / "e.csv"
    rows = [
        ["a", "b", "c", "d"],
        ["12abc", "3.5e2xyz", "2001-02-03 04:05:06", '"quoted, with ""quotes"""'],
        ["  +7", "  -0.25", "not a date", ""],
        ["-0", ".5", "1970-01-01 00:00:00", "x" * 400],
        ["", "", "", '""'],
        ["2147483647", "inf", " 2020-12-31 23:59:59", '"line\nbreak"'],
    ]
    with open(p, "w", newline="") as f:
        for r in rows:
            f.write(",".join(r) + "\r\n")
    names, types = ["a", "b", "c", "d"], [INT, FLOAT, DATETIME, STRING]
    rd = loader.CsvReader(p, names, types)
    cols, nulls = _read_all(rd)
    rd.close()
    assert list(cols[0]) == [12, 7, 0, -2**31, 2147483647]
    assert list(cols[1][:3]) == [np.float32(350.0), np.float32(-0.25), np.float32(0.5)]
    assert cols[1][3] == np.finfo(np.float32).tiny and np.isinf(cols[1][4])
    assert list(cols[2]) == [calendar.timegm((2001, 2, 3, 4, 5, 6)), 0, 0, 0, calendar.timegm((2020, 12, 31, 23, 59, 59))]
    assert list(nulls[2]) == [0, 0, 0, 1, 0]  # "not a date" is 0 but not NULL
    assert bytes(cols[3][0]).rstrip(b"\0") == b'quoted, with "quotes"'
    assert bytes(cols[3][2]) == b"x" * 255 + b"\0"
    assert bytes(cols[3][4]).rstrip(b"\0") == b"line\nbreak"
    assert list(nulls[3]) == [0, 1, 0, 1, 0]
    oc, on = O.read(p, names, types)
    for c in range(4):
        assert np.array_equal(cols[c], oc[c]) and np.array_equal(nulls[c], on[c]), c


def test_errors(loader, tmp_path):
    from yallasql_amd import _lib as L

    p = tmp_path / "bad.csv"
    p.write_text("a,b\nabc,1\n")
    rd = loader.CsvReader(p, ["a"], [INT])
    with pytest.raises(L.YsqlError):
        rd.next_host()  # no digits: std::stoi throws
    rd.close()
    p.write_text("a,b\n99999999999,1\n")
    rd = loader.CsvReader(p, ["a"], [INT])
    with pytest.raises(L.YsqlError):
        rd.next_host()  # out of int32 range
    rd.close()
    p.write_text("a,b\n1\n")
    rd = loader.CsvReader(p, ["b"], [INT])
    assert rd.next_host() is None  # a record shorter than the header is skipped
    rd.close()
    with pytest.raises(L.YsqlError):
        loader.CsvReader(p, ["zz"], [INT])  # no such column
    with pytest.raises(L.YsqlError):
        loader.CsvReader(os.path.join(str(tmp_path), "missing.csv"), ["a"], [INT])


@pytest.mark.gpu
def test_device_path_matches_reference(loader, tmp_path, dev):
    """Device columns (pinned staging, copies overlapping the next parse)
    equal what the reference's own csv-parser reads (oracle/_ref), or the
    oracle's restatement of it where that library was not built."""
    import torch

    n = 100_000
    p = tmp_path / "g.csv"
    _gen(p, n, 11)
    with open(p, "a", newline="") as f:  # records the parser skips or splits
        f.write('short\r\r\n\nx1,5,2.5,2001-02-03 04:05:06,"p"q,r"\rx2,6,,,z\n')
    names, types = ["id", "price", "ts", "name"], [INT, FLOAT, DATETIME, STRING]
    want_c, want_n = R.read(p, names, types) if R.available() else O.read(p, names, types)
    assert len(want_c[0]) == n + 2
    rd = loader.CsvReader(p, names, types)
    s = torch.cuda.Stream(device=dev)
    got = []
    while True:  # several batches: each parse overlaps the previous batch's copies
        b = rd.next_device(dev, stream=s)
        if b is None:
            break
        got.append(b)
    s.synchronize()
    rd.close()
    assert len(got) > 1
    for c in range(4):
        dc = torch.cat([b[0][c] for b in got]).cpu().numpy()
        dn = torch.cat([b[1][c] for b in got]).cpu().numpy()
        assert np.array_equal(dc, want_c[c]) and np.array_equal(dn, want_n[c]), c


@pytest.mark.gpu
@pytest.mark.parametrize("batch,slots", [(7_001, 2), (40_000, 3), (1 << 20, 4)])
def test_async_ring_matches_reference(loader, tmp_path, dev, batch, slots):
    """Scale mode (ysql_csv_open_ex: batch size of our choosing, a ring of
    pinned staging areas with device buffers of the reader, copies on its
    own stream): each batch is consumed by work queued on the caller's
    stream before the next call, while the next copies run; every column
    equals what the reference's parser reads."""
    import torch

    n = 100_000
    p = tmp_path / "g.csv"
    _gen(p, n, 12)
    names, types = ["id", "price", "ts", "name"], [INT, FLOAT, DATETIME, STRING]
    want_c, want_n = R.read(p, names, types) if R.available() else O.read(p, names, types)
    rd = loader.CsvReader(p, names, types, batch_rows_=batch, slots=slots, async_copies=True)
    s = torch.cuda.Stream(device=dev)
    outs = [torch.empty((n,) + ((256,) if t == STRING else ()), dtype=torch.uint8 if t == STRING else
                        {INT: torch.int32, FLOAT: torch.float32, DATETIME: torch.int64}[t], device=dev) for t in types]
    onul = [torch.empty(n, dtype=torch.uint8, device=dev) for _ in types]
    pos, nb = 0, 0
    with torch.cuda.stream(s):
        while True:
            b = rd.next_async(stream=s)
            if b is None:
                break
            m = b[0][0].shape[0]
            for c in range(4):  # the consumer: kernels on the caller's stream
                outs[c][pos:pos + m].copy_(b[0][c])
                onul[c][pos:pos + m].copy_(b[1][c])
            pos, nb = pos + m, nb + 1
    s.synchronize()
    rd.close()
    assert pos == n and nb == -(-n // batch)
    for c in range(4):
        assert np.array_equal(outs[c].cpu().numpy(), want_c[c]), c
        assert np.array_equal(onul[c].cpu().numpy(), want_n[c]), c


def test_parallel_parse_quotes_and_blank_lines(loader, tmp_path):
    """A batch of 200,000 rows (two 4-byte columns: 524,288 rows per batch)
    is split over several parsing threads; quoted fields holding commas,
    quotes and newlines, blank and "" lines and mixed line ends must not
    move a record boundary."""
    n = 200_000
    r = np.random.default_rng(21)
    ids = r.integers(-10**9, 10**9, n).astype(np.int32)
    flts = r.standard_normal(n).astype(np.float32)
    notes = ['"a,b"', '"x\ny"', '"say ""hi"""', "plain", '""', '"\r\n"', 'q"mid']
    pick = r.integers(0, len(notes), n)
    eols = r.integers(0, 2, n)
    blanks = set(r.choice(n, 300, replace=False).tolist())
    p = tmp_path / "par.csv"
    with open(p, "w", newline="") as f:
        f.write("note,id,price\n")
        for i in range(n):
            if i in blanks:
                f.write("\n" if i % 2 else '""\r\n')
            f.write("%s,%d,%.9g%s" % (notes[pick[i]], ids[i], flts[i], "\r\n" if eols[i] else "\n"))
    rd = loader.CsvReader(p, ["id", "price"], [INT, FLOAT])
    assert rd.rows == 524288
    cols, nulls = _read_all(rd)
    rd.close()
    assert np.array_equal(cols[0], ids)
    assert np.array_equal(cols[1].view(np.uint32), flts.view(np.uint32))
    assert not nulls[0].any() and not nulls[1].any()
    oc, on = O.read(p, ["id", "price"], [INT, FLOAT])
    for c in range(2):
        assert np.array_equal(cols[c], oc[c]) and np.array_equal(nulls[c], on[c]), c
    # a bad field in a late row range is still reported
    with open(p, "a", newline="") as f:
        f.write("z,notanumber,1\n")
    rd = loader.CsvReader(p, ["id", "price"], [INT, FLOAT])
    from yallasql_amd import _lib as L

    with pytest.raises(L.YsqlError):
        _read_all(rd)
    rd.close()


def test_quote_free_fast_path(loader, tmp_path):
    """A file with no quotes (newlines alone end records): blank "\\n" and
    "\\r\\n" lines, mixed line ends, one quoted record with an embedded
    newline late in the file, and no final newline."""
    n = 70_000
    r = np.random.default_rng(33)
    ids = r.integers(-10**6, 10**6, n).astype(np.int32)
    p = tmp_path / "qf.csv"
    with open(p, "w", newline="") as f:
        f.write("id,tag\n")
        for i in range(n):
            if i % 997 == 5:
                f.write("\n" if i % 2 else "\r\n")
            tag = "a b" if i % 13 == 0 else ('"q,\n"' if i == n - 100 else "t")
            f.write("%d,%s%s" % (ids[i], tag, "" if i == n - 1 else ("\r\n" if i % 3 == 0 else "\n")))
    rd = loader.CsvReader(p, ["id"], [INT])
    cols, nulls = _read_all(rd)
    rd.close()
    assert np.array_equal(cols[0], ids) and not nulls[0].any()
    oc, on = O.read(p, ["id"], [INT])
    assert np.array_equal(cols[0], oc[0]) and np.array_equal(nulls[0], on[0])


def _read_file(loader, path, names, types):
    rd = loader.CsvReader(path, names, types)
    try:
        cols, nulls = _read_all(rd)
    finally:
        rd.close()
    if cols is None:  # no rows
        cols = [np.zeros((0, 256), np.uint8) if t == STRING else
                np.zeros(0, {INT: np.int32, FLOAT: np.float32, DATETIME: np.int64}[t]) for t in types]
        nulls = [np.zeros(0, np.uint8) for _ in types]
    return cols, nulls


def _same(a, b):
    ca, na = a
    cb, nb = b
    return all(np.array_equal(x.view(np.uint8), y.view(np.uint8)) for x, y in zip(ca, cb)) and \
        all(np.array_equal(x, y) for x, y in zip(na, nb))


@pytest.mark.parametrize("case", sorted(CASES))
def test_reference_parser_fixture(loader, tmp_path, case):
    """Record ends, quotes, skipped rows, delimiter and header guessing, BOM,
    repeated names and conversions, as the reference's csv-parser read them
    (fixture); the oracle's restatement agrees too."""
    from yallasql_amd import _lib as L

    data, names, types, cols, nulls = CASES[case]
    p = tmp_path / "c.csv"
    p.write_bytes(data)
    if cols is None:  # the reference raised (std::stoi / std::stof)
        with pytest.raises(L.YsqlError):
            _read_file(loader, p, names, types)
        with pytest.raises(O.ConversionError):
            O.read(p, names, types)
        return
    assert _same(_read_file(loader, p, names, types), (cols, nulls))
    assert _same(O.read(p, names, types), (cols, nulls))


@pytest.mark.skipif(not R.available(), reason="oracle/_ref not built (needs the reference tree)")
def test_fixture_matches_reference_build(tmp_path):
    """The fixtures are what the reference's parser, built from its
    sources, reads today."""
    for case, (data, names, types, cols, nulls) in CASES.items():
        p = tmp_path / (case + ".csv")
        p.write_bytes(data)
        if cols is None:
            with pytest.raises(R.RefCsvError):
                R.read(p, names, types)
        else:
            assert _same(R.read(p, names, types), (cols, nulls)), case


def test_random_files_three_way(loader, tmp_path):
    """600 seeded random files over delimiters, quotes, "" pairs and every
    newline form: the loader equals the oracle's restatement and, where
    oracle/_ref is built, the reference's parser, column by column (every
    header name read as a string)."""
    r = np.random.default_rng(1234)
    alpha = [b"a", b"b", b",", b";", b'"', b"\r", b"\n", b"|", b"\t", b"^", b"xy", b'""', b"\r\n"]
    p = tmp_path / "f.csv"
    checked = 0
    for t in range(600):
        n = int(r.integers(1, 80 if t < 500 else 4000))
        data = b"".join(alpha[i] for i in r.integers(0, len(alpha), n))
        p.write_bytes(data)
        try:
            header, _ = O.read_fields(p)
        except O.ConversionError:
            continue
        names = []
        for h in header:
            if h.isascii() and b"\0" not in h and h.decode() not in names:
                names.append(h.decode())
        if not names:
            continue
        types = [STRING] * len(names)
        ours = _read_file(loader, p, names, types)
        assert _same(ours, O.read(p, names, types)), data
        if R.available():
            assert _same(ours, R.read(p, names, types)), data
        checked += 1
    assert checked > 400


def _chunk_file(suffix=b""):
    """A CRLF file whose "\r\n" crosses the first parse-chunk end, one
    record crosses the second, and a blank line's run crosses the third
    chunk's end, which lies 10,000,000 bytes past that record's start.
    Returns (bytes, data lines, header bytes)."""
    C = O.CHUNK_BYTES
    head = b"v,w\r\n" if suffix else b"v\r\n"
    out = bytearray(head)
    n = 0

    def line(k):
        nonlocal n
        body = ((b"%d-" % n) + b"x" * 80)[:k - len(suffix)] + suffix
        assert len(body) == k
        out.extend(body + b"\r\n")
        n += 1

    while len(out) + 62 < C - 1:
        line(58)
    line(C - 1 - len(out))  # '\r' at C - 1, '\n' at C
    while len(out) + 62 < 2 * C - 30:
        line(58)
    rs = len(out)
    line(58)  # crosses 2C
    assert rs < 2 * C < len(out) - 2
    while len(out) + 62 < rs + C - 1:
        line(58)
    line(rs + C - 1 - len(out))
    out.extend(b"\r\n")  # a blank line: the run crosses rs + C
    for _ in range(1000):
        line(40)
    return bytes(out), n, head


def test_chunk_boundary_records(loader, tmp_path):
    """The reference's parser reads the file in 10,000,000-byte chunks; a
    newline run crossing a chunk end yields an empty record there, which a
    one-column table keeps as a NULL row (two here) and a two-column table
    skips."""
    data, n, head = _chunk_file()
    p = tmp_path / "big.csv"
    p.write_bytes(data)
    cols, nulls = _read_file(loader, p, ["v"], [STRING])
    assert len(cols[0]) == n + 2
    # data rows before the first empty record: the lines ended before C - 1
    # (the header's not counted) plus the line whose "\r\n" crosses C
    before = data.count(b"\r\n", len(head), O.CHUNK_BYTES - 1) + 1
    empty = list(np.nonzero(nulls[0])[0])
    assert len(empty) == 2 and empty[0] == before
    assert _same((cols, nulls), O.read(p, ["v"], [STRING]))
    if R.available():
        assert _same((cols, nulls), R.read(p, ["v"], [STRING]))
    data2, n2, _ = _chunk_file(b",7")
    q = tmp_path / "big2.csv"
    q.write_bytes(data2)
    c2, nl2 = _read_file(loader, q, ["w"], [INT])
    assert len(c2[0]) == n2 and not nl2[0].any() and (c2[0] == 7).all()
    if R.available():
        assert _same((c2, nl2), R.read(q, ["w"], [INT]))


def _writer_table(r, n):
    ints = r.integers(-2**31, 2**31, n, dtype=np.int64).astype(np.int32)
    flts = (r.standard_normal(n) * 10.0 ** r.integers(-6, 32, n)).astype(np.float32)
    special = np.array([0.1, 2.5, 1.00005, 0.00015, -0.0, np.inf, -np.inf, np.nan], np.float32)
    flts[:min(n, 8)] = special[:min(n, 8)]
    secs = r.integers(-2_000_000_000, 253_402_300_799, n, dtype=np.int64)
    secs[:min(n, 3)] = np.array([0, -1, 951782400])[:min(n, 3)]
    strs = np.zeros((n, 256), np.uint8)
    alphabet = np.frombuffer(b'ab, "\nxyz\x80', np.uint8)
    for i in range(n):
        k = int(r.integers(0, 40))
        strs[i, :k] = alphabet[r.integers(0, len(alphabet), k)]
    if n > 3:
        strs[3, :] = 0x41  # no NUL in 256 bytes: cut at 256
    nulls = [(r.random(n) < 0.05).astype(np.uint8) for _ in range(4)]
    return [ints, flts, secs, strs], nulls


def test_writer_known_answers(loader, tmp_path):
    """Hand-computed lines: the exact values of 0.1f (0.100000001...),
    1.00005f (1.0000499...) and 0.00015f (0.000150000007...) rounded to 4
    decimals, -0.0, inf, NULL as "", quoted strings and UTC datetimes."""
    p = tmp_path / "k.csv"
    w = loader.CsvWriter(p, ["f", "t", "s", "i"], [FLOAT, DATETIME, STRING, INT])
    f = np.array([0.1, 2.5, 1.00005, 0.00015, -0.0, np.inf, 1e30], np.float32)
    t = np.array([0, -1, 951782400, 0, 86399, 0, 0], np.int64)
    s = np.zeros((7, 256), np.uint8)
    s[0, :5] = list(b'a,"b')[:4] + [0]
    s[1, :3] = list(b"xyz")
    i = np.array([1, -2, 2147483647, -2147483648, 0, 5, 6], np.int32)
    nl = [np.zeros(7, np.uint8), np.zeros(7, np.uint8), np.zeros(7, np.uint8), np.array([0, 0, 0, 0, 0, 1, 0], np.uint8)]
    w.write_host([f, t, s, i], nl)
    w.close()
    assert p.read_bytes().split(b"\n") == [
        b"f,t,s,i",
        b'0.1000,"1970-01-01 00:00:00","a,"b",1',
        b'2.5000,"1969-12-31 23:59:59","xyz",-2',
        b'1.0000,"2000-02-29 00:00:00","",2147483647',
        b'0.0002,"1970-01-01 00:00:00","",-2147483648',
        b'-0.0000,"1970-01-01 23:59:59","",0',
        b'inf,"1970-01-01 00:00:00","",""',
        b'1000000015047466219876688855040.0000,"1970-01-01 00:00:00","",6',
        b"",
    ]


def test_writer_matches_oracle(loader, tmp_path):
    """Several batches (one large enough for the threaded formatter) against
    the oracle's restatement; the header is written once."""
    r = np.random.default_rng(8)
    names, types = ["i", "f", "t", "s"], [INT, FLOAT, DATETIME, STRING]
    p = tmp_path / "w.csv"
    w = loader.CsvWriter(p, names, types)
    want = []
    for k, n in enumerate((0, 5, 40_000, 3)):
        cols, nulls = _writer_table(r, n) if n else ([np.zeros(0, np.int32), np.zeros(0, np.float32),
                                                      np.zeros(0, np.int64), np.zeros((0, 256), np.uint8)], None)
        w.write_host(cols, nulls)
        want.append(O.write_rows(names, types, cols, nulls, header=(k == 0)))
    w.close()
    assert p.read_bytes() == b"".join(want)


@pytest.mark.gpu
def test_writer_device_path_matches_oracle(loader, tmp_path, dev):
    """Batches from device memory (pinned staging, grown between batches)
    give the oracle's bytes."""
    import torch

    r = np.random.default_rng(12)
    names, types = ["i", "f", "t", "s"], [INT, FLOAT, DATETIME, STRING]
    p = tmp_path / "d.csv"
    w = loader.CsvWriter(p, names, types)
    want = []
    for k, n in enumerate((1000, 70_000, 10)):
        cols, nulls = _writer_table(r, n)
        w.write_device([torch.from_numpy(c).to(dev) for c in cols], [torch.from_numpy(x).to(dev) for x in nulls])
        want.append(O.write_rows(names, types, cols, nulls, header=(k == 0)))
    w.close()
    assert p.read_bytes() == b"".join(want)


@pytest.mark.gpu
def test_engine_execute_csv(tmp_path, dev):
    """A plan (filter, ORDER BY a string column, projection) written batch by
    batch to CSV equals the oracle's rendering of the same plan's rows."""
    from yallasql_amd import engine as E

    r = np.random.default_rng(13)
    n = 30_000
    names = np.zeros((n, 256), np.uint8)
    for i in range(n):
        k = int(r.integers(0, 8))
        names[i, :k] = r.integers(97, 101, k)
    ids = r.integers(0, 1000, n).astype(np.int32)
    price = (r.random(n) * 1000).astype(np.float32)
    when = r.integers(0, 2_000_000_000, n).astype(np.int64)
    ctx = E.Context()
    t = ctx.table("t", [("id", ids), ("name", names), ("price", price), ("when", when)])

    def plan():
        p = ctx.scan(t, 8192).filter(E.cmp(2, E.ref(2), E.f32(300.0)))
        return p.order([1, 0], [False, True])

    cols, nulls, _ = plan().execute()
    out = tmp_path / "res.csv"
    rows = plan().execute_csv(out)
    assert rows == cols[0].shape[0] > 0
    want = O.write_rows(["id", "name", "price", "when"], [INT, STRING, FLOAT, DATETIME], cols, nulls)
    assert out.read_bytes() == want
