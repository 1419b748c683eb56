"""The C-ABI library loads here (no GPU needed) and exports every symbol
that include/*.h declares; argument checks answer without touching a GPU."""
import ctypes as C
import glob
import os
import re

from tests.conftest import ROOT


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(ysql_\w+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_header_declares_entry_points():
    names = declared_symbols()
    for must in ("ysql_radix_argsort", "ysql_merge_path", "ysql_filter_compact", "ysql_reduce", "ysql_join_pairs",
                 "ysql_cross_expand", "ysql_star_join_group_sum"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from yallasql_amd import _lib

    missing = [n for n in sorted(declared_symbols()) if not hasattr(_lib.lib, n)]
    assert not missing, missing


def test_version_and_ws_queries_without_gpu():
    from yallasql_amd import _lib as L

    assert b"gfx950" in L.lib.ysql_version()
    sz = C.c_size_t(0)
    assert L.lib.ysql_radix_sort_ws_bytes(L.UKEY64, 1 << 25, C.byref(sz)) == 0
    # alternate keys + alternate payloads at least
    assert sz.value >= (1 << 25) * 12
    assert L.lib.ysql_radix_sort_ws_bytes(99, 10, C.byref(sz)) == L.EINVAL
    assert L.lib.ysql_scan_ws_bytes(1000, C.byref(sz)) == 0 and sz.value > 0
    assert L.lib.ysql_merge_ws_bytes(10, 10, 20, C.byref(sz)) == 0 and sz.value > 0
    assert L.lib.ysql_join_ws_bytes(L.INT, L.EQ, 100, 100, C.byref(sz)) == 0 and sz.value > 0


def test_invalid_arguments_rejected_before_launch():
    from yallasql_amd import _lib as L

    # bad dtype / op codes and null pointers return EINVAL without a device call
    assert L.lib.ysql_sortkey_encode(L.STRING, 0, C.c_void_p(8), C.c_void_p(8), 10, None) == L.EINVAL
    assert L.lib.ysql_compare(L.INT, 17, C.c_void_p(8), 0, C.c_void_p(8), 0, 0, C.c_void_p(8), 10, None) == L.EINVAL
    assert L.lib.ysql_arith(L.INT, L.ADD, None, 0, None, 0, None, 10, None) == L.EINVAL
    assert L.lib.ysql_radix_argsort(L.UKEY32, 0, None, None, 10, 0, 32, None, 0, None) == L.EINVAL
    assert L.lib.ysql_radix_argsort(L.UKEY32, 0, C.c_void_p(8), C.c_void_p(8), 1 << 31, 0, 32, None, 0,
                                    None) == L.EINVAL
    assert L.lib.ysql_radix_argsort(L.UKEY32, 0, C.c_void_p(8), C.c_void_p(8), 10, 0, 33, C.c_void_p(8), 1 << 20,
                                    None) == L.EINVAL
    assert L.lib.ysql_radix_argsort(L.UKEY32, 0, C.c_void_p(8), C.c_void_p(8), 10, 0, 32, C.c_void_p(8), 16,
                                    None) == L.EWORKSPACE
    cols = (L.Col * 1)()
    cols[0] = L.Col(8, None, 8, None, 3, 0)  # 3-byte elements are not supported
    assert L.lib.ysql_compact_cols(C.c_void_p(8), cols, 1, 10, C.c_void_p(8), C.c_void_p(8), 1 << 20,
                                   None) == L.EINVAL
    assert L.lib.ysql_star_join_group_sum(None, 0, None, None, 10, None, None, 10, None) == L.EINVAL
    # a digit pass scatters into buffers other tiles may still read: keys or
    # payloads in place are refused
    k, v, k2 = C.c_void_p(256), C.c_void_p(512), C.c_void_p(768)
    assert L.lib.ysql_radix_pass(L.UKEY64, 0, 0, k, None, 0, k, v, 10, 0, 64, C.c_void_p(8), 1 << 20,
                                 None) == L.EINVAL
    assert L.lib.ysql_radix_pass(L.UKEY64, 0, 1, k, v, 0, k2, v, 10, 0, 64, C.c_void_p(8), 1 << 20,
                                 None) == L.EINVAL


def test_reference_shaped_launchers_compile(tmp_path):
    """include/ysql_kernels.hpp instantiates for the reference's types."""
    import shutil
    import subprocess

    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    src = tmp_path / "compat.cpp"
    src.write_text(r'''
#include "ysql_kernels.hpp"
using namespace YallaSQL::Kernel;
void use(hipStream_t s) {
  int *i = nullptr; float *f = nullptr; int64_t *l = nullptr; bool *m = nullptr; uint32_t *u = nullptr;
  uint64_t *u64 = nullptr; char *nb = nullptr; int *cnt = nullptr;
  launch_int_to_uint32(i, u, 10); launch_float_to_uint32(f, u, 10);
  launch_radix_sort<DescOp<uint32_t>, uint32_t, 256>(u, u, 10, s);
  launch_radix_sort<AscOp<uint64_t>, uint64_t, 256>(u64, u, 10, s);
  launch_prefix_sum_mask(m, u, 10, s); launch_prefix_sum<uint32_t>(u, u, 10, s);
  launch_merge_sorted_array_kernel<float, DescSortMergeOp<float>>(f, f, f, u, m, cnt, 10, 5, 5, s);
  launch_move_one_col<int>(i, i, nb, nb, u, 10, s);
  launch_move_one_col_merge<int64_t>(l, l, l, nb, nb, nb, u, m, true, 10, s);
  launch_move_rows_filter_kernel<float>(f, f, u, m, nb, 10, s);
  lanch_move_rows_join_kernel<int>(i, i, u, nb, nb, 10, true, s);
  launch_conditional_operators<float, GEOperator<float>>(f, f, OperandType::SCALAR, OperandType::VECTOR, m, 10, s, false);
  launch_conditional_operators<bool, ANDOperator>(m, m, OperandType::VECTOR, OperandType::VECTOR, m, 10, s);
  launch_binary_operators<int, MulOperator<int>>(i, i, OperandType::SCALAR, OperandType::VECTOR, i, 10, s);
  launch_numerical_cast<int, float>(i, f, 10, s);
  launch_outer_join_operators<int64_t, EQOperator<int64_t>>(l, l, u, cnt, 10, 10, s);
  launch_and_join_operators(u, u, u, u, cnt, 10, 10, s);
  launch_cross_product_col<int>(i, i, nb, nb, 3, 4, s, true);
}
''')
    r = subprocess.run([hipcc, "-fsyntax-only", "-std=c++17", "-I", f"{ROOT}/include", str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
