"""Writes tests/golden/csv/cases.npz: small CSV files and what the
reference's CSV scan reads from them, taken from the reference's own
csv-parser (oracle/_ref/libref_csv.so, built from the reference's sources by
oracle/Makefile; see oracle/ref_csv.cpp).

Each case is the file's bytes, the requested column names and types, and
either the columns and null bytes the reference produced or the mark that it
raised (a field std::stoi / std::stof rejects).  tests/test_csv_loader.py
checks the loader and the oracle's restatement against them.
Run from the repository root:  python tests/golden/csv/make_csv_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, ROOT)

from oracle import csv_ref as R  # noqa: E402

INT, FLOAT, DATETIME, STRING = 0, 1, 2, 3
S2 = [STRING, STRING]

# name: (file bytes, column names, column types)
CASES = {
    # record ends: any run of \r and \n ends a record; blank lines vanish
    "plain": (b"a,b\n1,2\n3,4\n", ["a", "b"], S2),
    "crlf": (b"a,b\r\n1,2\r\n", ["a", "b"], S2),
    "lone_cr": (b"a,b\r1,2\r3,4", ["a", "b"], S2),
    "cr_in_field": (b"a,b\n1,x\ry\n5,6\n", ["a", "b"], S2),
    "blank_lines": (b"a,b\n\n1,2\n\r\n\n3,4\n\n", ["a", "b"], S2),
    "cr_runs": (b"a,b\r\r\r1,2\r\r", ["a", "b"], S2),
    "leading_newline": (b"\na,b\n1,2\n", ["a", "b"], S2),
    "one_column": (b'a\n1\n\n2\r\n""\n3', ["a"], [STRING]),
    "one_column_quotes": (b'a\n""\n"x"\n\n"\n"\n', ["a"], [STRING]),
    # quotes
    "quote_mid_field": (b'a,b\n1,x"y"z\n', ["a", "b"], S2),
    "quoted_newline": (b'a,b\n1,"x\ny"\n', ["a", "b"], S2),
    "quoted_crlf": (b'a,b\n1,"x\r\ny"\n', ["a", "b"], S2),
    "stray_quotes": (b'a,b\n"x""y",2\n"p"q,3\n"r"s"t",4\n5,6\n', ["a", "b"], S2),
    "stray_then_text": (b'a,b\n1,"x"y,2\n3,4\n', ["a", "b"], S2),
    "strays_and_pair": (b'a,b\n1,"a"b"c""d"\n', ["a", "b"], S2),
    "unterminated_quote": (b'a,b\n1,"xyz', ["a", "b"], S2),
    "closing_quote_at_eof": (b'a,b\n1,"xyz"', ["a", "b"], S2),
    "pair_at_eof": (b'a,b\n1,"ab""', ["a", "b"], S2),
    "open_quote_at_eof": (b'a,b\n1,"', ["a", "b"], S2),
    "empty_quoted": (b'a,b\n"",""\n', ["a", "b"], S2),
    "rfc4180": (b'a,b,c\n"x,y","say ""hi""",plain\n', ["a", "b", "c"], [STRING] * 3),
    # rows of another length than the header are skipped
    "trailing_comma": (b"a,b\n1,2,\n3,4\n", ["a", "b"], S2),
    "short_row": (b"a,b\n1\n2,3\n", ["a", "b"], S2),
    "only_header": (b"a,b\n", ["a", "b"], S2),
    # header and delimiter guessing
    "spaces_kept": (b"a, b\n1, 2\n", ["a", " b"], S2),
    "semicolon": (b"a;b\n1;2\n3;4\n", ["a", "b"], S2),
    "tab": (b"a\tb\n1\t2\n", ["a", "b"], S2),
    "pipe_over_comma": (b"a|b\n1,5|2\n", ["a", "b"], S2),
    "caret": (b"a^b^c\n1^2^3\n", ["a", "c"], S2),
    "title_row": (b"title line\na,b\n1,2\n3,4\n", ["a", "b"], S2),
    "header_length_mode": (b"a,b,c\n1,2\n3,4\n5,6,7\n", ["a"], [STRING]),
    "length_tie": (b"x\ny\nz\na,b,c\n", ["x"], [STRING]),
    "delimiter_tie": (b"a;b,c\n", ["a;b", "c"], S2),
    "bom": (b"\xef\xbb\xbfa,b\n1,2\n", ["a", "b"], S2),
    "repeated_name": (b"a,a,b\n1,2,3\n", ["a", "b"], S2),
    "nul_byte": (b"a,b\n1,x\x00y\n3,4\n", ["a", "b"], S2),
    # conversions (get_operator.cpp:74-95)
    "ints": (b"i\n12abc\n  +7\n-0\n2147483647\n-2147483648\n0x1A\n\n 42 \n", ["i"], [INT]),
    "floats": (b"f\n3.5e2xyz\n  -0.25\n.5\ninf\n-inf\nnan\n0x1p3\n1.17549435e-38\n3.4028234e38\n1e-7\n-0\n",
               ["f"], [FLOAT]),
    "datetimes": (b"t\n2001-02-03 04:05:06\n2001-2-3 4:5:6\nnot a date\n2020-12-31 23:59:59 extra\n"
                  b" 2020-12-31 23:59:59\n1969-12-31 23:59:59\n2000-02-30 00:00:00\n\"\"\n", ["t"], [DATETIME]),
    "long_string": (b"s\n" + b"x" * 400 + b"\n" + b"y" * 255 + b"\n" + b"z" * 254 + b"\n", ["s"], [STRING]),
    "mixed_types": (b"a,b,c,d\n7,1.5,2001-02-03 04:05:06,w\n,,,\n-3,2e3,,\"q\"\"\"\n",
                    ["d", "a", "c", "b"], [STRING, INT, DATETIME, FLOAT]),
    "int_no_digits": (b"i\n1\nabc\n", ["i"], [INT]),
    "int_overflow": (b"i\n99999999999\n", ["i"], [INT]),
    "float_overflow": (b"f\n1e39\n", ["f"], [FLOAT]),
    "float_subnormal": (b"f\n1e-40\n", ["f"], [FLOAT]),
}


def main():
    import tempfile

    assert R.available(), "build oracle/_ref first (make -C oracle)"
    out = {}
    tmp = tempfile.mkdtemp()
    for name, (data, names, types) in CASES.items():
        p = os.path.join(tmp, name + ".csv")
        with open(p, "wb") as f:
            f.write(data)
        out[name + "__data"] = np.frombuffer(data, np.uint8)
        out[name + "__names"] = np.frombuffer(b"\0".join(n.encode() for n in names), np.uint8)
        out[name + "__types"] = np.array(types, np.int32)
        try:
            cols, nulls = R.read(p, names, types)
        except R.RefCsvError:
            out[name + "__error"] = np.ones(1, np.uint8)
            continue
        for c in range(len(names)):
            out[f"{name}__c{c}"] = cols[c]
            out[f"{name}__n{c}"] = nulls[c]
    np.savez_compressed(os.path.join(HERE, "cases.npz"), **out)
    print("wrote", len(CASES), "cases")


def load():
    """{name: (data, names, types, cols or None, nulls or None)} from cases.npz."""
    with np.load(os.path.join(HERE, "cases.npz"), allow_pickle=False) as z:
        keys = set(z.files)
        res = {}
        for name in sorted({k.split("__")[0] for k in keys}):
            names = [x.decode() for x in z[name + "__names"].tobytes().split(b"\0")]
            types = [int(t) for t in z[name + "__types"]]
            if name + "__error" in keys:
                cols = nulls = None
            else:
                cols = [z[f"{name}__c{c}"] for c in range(len(names))]
                nulls = [z[f"{name}__n{c}"] for c in range(len(names))]
            res[name] = (z[name + "__data"].tobytes(), names, types, cols, nulls)
    return res


if __name__ == "__main__":
    main()
