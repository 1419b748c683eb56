"""ctypes binding of libyallasql_amd.so (the C ABI in include/ysql_kernels.h).

The library is the product: there is no Python or CPU fallback.  Importing
this module raises if the built library is missing.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libyallasql_amd.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"{LIB_PATH} is not built; run `python -m yallasql_amd.build` "
        "(or __graft_entry__.build()) first"
    )

# torch (ROCm build) brings its own HIP runtime; load it first so that the
# library binds to that same runtime instance instead of a second one (the
# Python layers hand torch's device memory and streams to the library).
import torch  # noqa: E402,F401

lib = C.CDLL(LIB_PATH)

# ---- constants (mirror include/ysql_kernels.h) ---------------------------
OK, EINVAL, EWORKSPACE, EOVERFLOW, EHIP, EUNSUPPORTED = 0, -1, -2, -3, -4, -5
INT, FLOAT, DATETIME, STRING, BOOL, UKEY32, UKEY64 = 0, 1, 2, 3, 4, 16, 17
STRING_BYTES = 256
LT, LE, GT, GE, EQ, NE = 0, 1, 2, 3, 4, 5
AND, OR = 0, 1
ADD, SUB, MUL, DIV, REM = 0, 1, 2, 3, 4
AGG_MAX, AGG_MIN, AGG_SUM, AGG_AVG, AGG_COUNT, AGG_COUNT_STAR = 0, 1, 2, 3, 4, 5
MAX_COLS = 8
MAX_DIMS = 4
MAX_GROUPS = 16384


class Acc(C.Structure):
    _fields_ = [("i", C.c_int64), ("d", C.c_double), ("f", C.c_float), ("pad", C.c_uint32), ("count", C.c_int64)]


class Col(C.Structure):
    _fields_ = [
        ("src", C.c_void_p), ("src_null", C.c_void_p), ("dst", C.c_void_p), ("dst_null", C.c_void_p),
        ("elem_bytes", C.c_uint32), ("pad", C.c_uint32),
    ]


class SortCol(C.Structure):
    _fields_ = [("dtype", C.c_int32), ("descending", C.c_int32), ("a", C.c_void_p), ("b", C.c_void_p)]


class StarDim(C.Structure):
    _fields_ = [
        ("fk", C.c_void_p), ("rows", C.c_uint32), ("group_card", C.c_uint32), ("pass_", C.c_void_p),
        ("group", C.c_void_p), ("group_base", C.c_int32), ("pad", C.c_uint32),
    ]


class XInsn(C.Structure):
    _fields_ = [("op", C.c_int16), ("type", C.c_int16), ("sub", C.c_int16), ("dst", C.c_int8), ("a", C.c_int8),
                ("b", C.c_int8), ("nsrc", C.c_int8), ("imm", C.c_int64)]


class XInput(C.Structure):
    _fields_ = [("data", C.c_void_p), ("nulls", C.c_void_p), ("dtype", C.c_int32), ("pad", C.c_int32)]


class XOutput(C.Structure):
    _fields_ = [("data", C.c_void_p), ("nulls", C.c_void_p), ("reg", C.c_int32), ("dtype", C.c_int32)]


XOP_CONST, XOP_CMP, XOP_LOGIC, XOP_ARITH, XOP_CAST, XOP_NOT = 1, 2, 3, 4, 5, 6
XMAX_REGS, XMAX_INSN, XMAX_INPUTS = 16, 48, 8

assert C.sizeof(Acc) == 32 and C.sizeof(Col) == 40 and C.sizeof(StarDim) == 40
assert C.sizeof(XInsn) == 24 and C.sizeof(XInput) == 24 and C.sizeof(XOutput) == 24

P = C.c_void_p
U64 = C.c_uint64
I32 = C.c_int
U32 = C.c_uint32
SZP = C.POINTER(C.c_size_t)

_SIGS = {
    "ysql_sortkey_encode": [I32, I32, P, P, U64, P],
    "ysql_sortkey_decode": [I32, I32, P, P, U64, P],
    "ysql_radix_sort_ws_bytes": [I32, U64, SZP],
    "ysql_radix_argsort": [I32, I32, P, P, U64, I32, I32, P, C.c_size_t, P],
    "ysql_radix_sort_pairs": [I32, I32, P, P, U64, I32, I32, P, C.c_size_t, P],
    "ysql_radix_sort_ex": [I32, I32, P, P, U32, P, P, U64, I32, I32, P, C.c_size_t, P],
    "ysql_radix_sort_variant": [I32, I32, I32, P, P, U32, P, P, U64, I32, I32, P, C.c_size_t, P],
    "ysql_radix_begin": [I32, I32, P, U64, I32, I32, P, C.c_size_t, P],
    "ysql_radix_pass": [I32, I32, I32, P, P, U32, P, P, U64, I32, I32, P, C.c_size_t, P],
    "ysql_scan_ws_bytes": [U64, SZP],
    "ysql_scan_u32_inclusive": [P, P, U64, P, C.c_size_t, P],
    "ysql_scan_mask_inclusive": [P, P, U64, P, C.c_size_t, P],
    "ysql_merge_ws_bytes": [U64, U64, U64, SZP],
    "ysql_merge_path": [I32, I32, P, U64, P, U64, U64, P, P, P, P, P, C.c_size_t, P],
    "ysql_merge_runs_ws_bytes": [I32, C.POINTER(U64), I32, SZP],
    "ysql_merge_runs": [I32, P, P, C.POINTER(U64), I32, P, P, P, C.c_size_t, P],
    "ysql_csv_writer_open": [C.c_char_p, I32, C.POINTER(C.c_char_p), C.POINTER(C.c_int), C.POINTER(C.c_void_p)],
    "ysql_csv_writer_write_host": [P, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), U64],
    "ysql_csv_writer_write": [P, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), U64, P],
    "ysql_csv_writer_close": [P],
    "ysql_string_sort_ws_bytes": [U64, SZP],
    "ysql_string_sort": [I32, P, P, U64, P, C.c_size_t, P],
    "ysql_string_argsort": [I32, P, P, U64, P, C.c_size_t, P],
    "ysql_merge_rows": [C.POINTER(SortCol), I32, U64, U64, U64, P, P, P, P, P],
    "ysql_merge_pairs": [I32, I32, P, P, U64, P, P, U64, P, P, P, C.c_size_t, P],
    "ysql_gather": [U32, P, P, P, U64, P, P, P],
    "ysql_gather2": [U32, P, P, P, P, P, P, U64, I32, P, P, P],
    "ysql_gather_pairs": [U32, P, P, P, I32, U64, P, P, P],
    "ysql_compact": [U32, P, P, P, P, U64, P, P, P],
    "ysql_compact_ws_bytes": [U64, SZP],
    "ysql_compact_cols": [P, C.POINTER(Col), I32, U64, P, P, C.c_size_t, P],
    "ysql_csv_batch_rows": [I32, C.POINTER(I32), C.POINTER(U32)],
    "ysql_csv_open": [C.c_char_p, I32, C.POINTER(C.c_char_p), C.POINTER(I32), C.POINTER(P)],
    "ysql_csv_next_host": [P, C.POINTER(P), C.POINTER(P), C.POINTER(U32)],
    "ysql_csv_next": [P, C.POINTER(P), C.POINTER(P), C.POINTER(U32), P],
    "ysql_csv_open_ex": [C.c_char_p, I32, C.POINTER(C.c_char_p), C.POINTER(I32), P, C.POINTER(P)],
    "ysql_csv_next_async": [P, C.POINTER(P), C.POINTER(P), C.POINTER(U32), P],
    "ysql_filter_compact": [I32, I32, P, P, C.POINTER(Col), I32, U64, P, P, C.c_size_t, P],
    "ysql_filter_compact_noalloc": [I32, I32, P, P, C.POINTER(Col), I32, U64, P, P],
    "ysql_filter_compact_range": [I32, P, P, C.POINTER(Col), I32, U64, U64, P, P, C.c_size_t, P],
    "ysql_expr_eval": [C.POINTER(XInsn), I32, C.POINTER(XInput), I32, C.POINTER(XOutput), I32, U64, P],
    "ysql_expr_filter_compact": [C.POINTER(XInsn), I32, I32, C.POINTER(XInput), I32, C.POINTER(Col), I32, U64, P, P,
                                 C.c_size_t, P],
    "ysql_compare": [I32, I32, P, I32, P, I32, I32, P, U64, P],
    "ysql_mask_logic": [I32, P, I32, P, I32, P, U64, P],
    "ysql_arith": [I32, I32, P, I32, P, I32, P, U64, P],
    "ysql_cast": [I32, I32, P, P, U64, P],
    "ysql_acc_init": [I32, I32, P, I32, P],
    "ysql_reduce_ws_bytes": [U64, SZP],
    "ysql_reduce": [I32, I32, P, P, U64, P, P, C.c_size_t, P],
    "ysql_agg_finalize": [I32, I32, P, P, P],
    "ysql_acc_from_value": [I32, I32, P, P, P],
    "ysql_acc_to_value": [I32, I32, P, P, P],
    "ysql_div_avg": [I32, P, P, P, P],
    "ysql_double_to_float": [P, P, P],
    "ysql_group_reduce": [I32, I32, P, U32, P, P, U64, P, P],
    "ysql_join_ws_bytes": [I32, I32, U64, U64, SZP],
    "ysql_join_pairs": [I32, I32, P, U64, P, U64, P, U64, P, P, C.c_size_t, P],
    "ysql_pairs_intersect_ws_bytes": [U64, U64, SZP],
    "ysql_pairs_intersect": [P, U64, P, U64, P, U64, P, P, C.c_size_t, P],
    "ysql_cross_expand": [U32, P, P, U64, U64, I32, P, P, P],
    "ysql_star_join_group_sum": [C.POINTER(StarDim), I32, P, P, U64, P, P, U32, P],
    "ysql_star_join_group_sum_multi": [C.POINTER(StarDim), I32, I32, P, P, U64, P, P, U32, P],
}

for _name, _args in _SIGS.items():
    _f = getattr(lib, _name)
    _f.argtypes = _args
    _f.restype = C.c_int

lib.ysql_version.restype = C.c_char_p
lib.ysql_version.argtypes = []
lib.ysql_last_error.restype = C.c_char_p
lib.ysql_last_error.argtypes = []
lib.ysql_csv_close.restype = None
lib.ysql_csv_close.argtypes = [P]


class YsqlError(RuntimeError):
    def __init__(self, fn, rc):
        msg = lib.ysql_last_error().decode(errors="replace")
        super().__init__(f"{fn} returned {rc}" + (f" ({msg})" if msg else ""))
        self.rc = rc


def check(fn, rc):
    if rc != OK:
        raise YsqlError(fn, rc)
    return rc


def call(name, *args):
    """Call a C-ABI entry point and raise YsqlError on a nonzero status."""
    return check(name, getattr(lib, name)(*args))


def ws_bytes(name, *args):
    out = C.c_size_t(0)
    call(name, *args, C.byref(out))
    return out.value


class Job(C.Structure):
    _fields_ = [
        ("kind", C.c_int), ("key_dtype", C.c_int), ("n", C.c_uint64), ("keys", C.c_void_p), ("vals", C.c_void_p),
        ("threshold", C.c_float), ("pad", C.c_uint32), ("keys_out", C.c_void_p), ("perm_out", C.c_void_p),
        ("vals_out", C.c_void_p), ("count_out", C.c_void_p), ("accs_out", C.c_void_p),
    ]


class AggSpec(C.Structure):
    _fields_ = [("dtype", C.c_int), ("agg", C.c_int), ("col", C.c_void_p), ("acc", C.c_void_p)]


FA_OVERWRITE = 1
FA_GENERIC = 2


class BatchTuning(C.Structure):
    _fields_ = [("chains", C.c_int), ("seg_sort_max", C.c_int), ("seg_filter_max", C.c_int), ("seg_agg_max", C.c_int),
                ("sort_variant", (C.c_int8 * 32) * 2), ("persist_split_min", C.c_int), ("persist_chunk", C.c_int),
                ("persist_jobs_first_every", C.c_int), ("persist_idle_limit", C.c_int),
                ("persist_workers_per_cu", C.c_int), ("persist_split_filter_min", C.c_int),
                ("persist_split_agg_min", C.c_int)]


JOB_SORT, JOB_FILTER, JOB_AGG = 0, 1, 2
assert C.sizeof(Job) == 80
for _name, _args, _res in [
    ("ysql_filter_aggregate_ws_bytes", [U64, I32, SZP], C.c_int),
    ("ysql_filter_aggregate", [I32, I32, P, P, P, I32, U64, P, C.c_size_t, P], C.c_int),
    ("ysql_filter_aggregate_ex", [I32, I32, P, P, P, I32, U64, I32, P, C.c_size_t, P], C.c_int),
    ("ysql_filter_aggregate_multi_ws_bytes", [U64, I32, I32, SZP], C.c_int),
    ("ysql_filter_aggregate_multi", [I32, I32, P, P, I32, P, I32, U64, I32, P, C.c_size_t, P], C.c_int),
    ("ysql_group_ids", [P, P, P, I32, U64, P, P], C.c_int),
    ("ysql_sortkey_pack", [I32, I32, P, U64, P, P], C.c_int),
    ("ysql_batch_tuning_default", [C.POINTER(BatchTuning)], None),
    ("ysql_batch_create", [C.POINTER(Job), I32, C.POINTER(BatchTuning), C.POINTER(P)], C.c_int),
    ("ysql_batch_run", [P, P], C.c_int),
    ("ysql_batch_run_plain", [P, P], C.c_int),
    ("ysql_batch_run_graph", [P, P], C.c_int),
    ("ysql_batch_check", [P, P, C.POINTER(U32), C.POINTER(U32)], C.c_int),
    ("ysql_batch_workers", [P, C.POINTER(U32)], C.c_int),
    ("ysql_batch_info", [P, C.POINTER(U64), C.POINTER(U32), C.POINTER(U64)], C.c_int),
    ("ysql_batch_destroy", [P], None),
]:
    _f = getattr(lib, _name)
    _f.argtypes = _args
    _f.restype = _res
