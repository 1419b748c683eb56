"""Python entry points over the C ABI, on torch device tensors.

Device memory and streams come from torch (plumbing only); every op runs the
HIP kernels of libyallasql_amd.so on torch's current stream.  Sort keys and
unsigned values are carried in int32/int64 tensors holding the unsigned bits.
"""
import ctypes as C

import torch

from . import _lib as L
from ._lib import call, ws_bytes

_DT = {torch.int32: L.INT, torch.float32: L.FLOAT, torch.int64: L.DATETIME, torch.uint8: L.BOOL}


def dtype_code(t):
    if t.dtype == torch.uint8 and t.dim() == 2 and t.shape[1] == L.STRING_BYTES:
        return L.STRING
    return _DT[t.dtype]


def elem_bytes(t):
    if t.dim() == 2:
        return t.shape[1] * t.element_size()
    return t.element_size()


def _p(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def _s():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _dev(t):
    return t.device


def workspace(nbytes, device):
    return torch.empty(max(int(nbytes), 256), dtype=torch.uint8, device=device)


def _rows(t):
    return t.shape[0]


# ---- A1 -----------------------------------------------------------------
def sortkey_encode(x, descending=False):
    out = torch.empty_like(x) if x.dtype != torch.float32 else torch.empty(x.shape, dtype=torch.int32, device=x.device)
    call("ysql_sortkey_encode", dtype_code(x), int(descending), _p(x), _p(out), x.numel(), _s())
    return out


def sortkey_decode(k, dtype, descending=False):
    out = torch.empty(k.shape, dtype=dtype, device=k.device)
    call("ysql_sortkey_decode", _DT[dtype], int(descending), _p(k), _p(out), k.numel(), _s())
    return out


# ---- A2-A6 --------------------------------------------------------------
def _kdt(keys):
    return L.UKEY64 if keys.element_size() == 8 else L.UKEY32


def radix_argsort(keys, descending=False, begin_bit=0, end_bit=None):
    """Sorts `keys` (unsigned bits) in place; returns perm (int32 bits of u32)."""
    n = keys.numel()
    kdt = _kdt(keys)
    end_bit = 8 * keys.element_size() if end_bit is None else end_bit
    perm = torch.empty(n, dtype=torch.int32, device=keys.device)
    ws = workspace(ws_bytes("ysql_radix_sort_ws_bytes", kdt, n), keys.device)
    call("ysql_radix_argsort", kdt, int(descending), _p(keys), _p(perm), n, begin_bit, end_bit, _p(ws), ws.numel(), _s())
    return perm


def radix_sort_pairs(keys, vals, descending=False, begin_bit=0, end_bit=None, ws=None):
    n = keys.numel()
    kdt = _kdt(keys)
    end_bit = 8 * keys.element_size() if end_bit is None else end_bit
    if ws is None:
        ws = workspace(ws_bytes("ysql_radix_sort_ws_bytes", kdt, n), keys.device)
    call("ysql_radix_sort_pairs", kdt, int(descending), _p(keys), _p(vals), n, begin_bit, end_bit, _p(ws), ws.numel(),
         _s())
    return keys, vals


def radix_sort_ex(keys_in, keys_out, vals_out, vals_in=None, val_base=0, descending=False, begin_bit=0,
                  end_bit=None, ws=None):
    n = keys_in.numel()
    kdt = _kdt(keys_in)
    end_bit = 8 * keys_in.element_size() if end_bit is None else end_bit
    if ws is None:
        ws = workspace(ws_bytes("ysql_radix_sort_ws_bytes", kdt, n), keys_in.device)
    call("ysql_radix_sort_ex", kdt, int(descending), _p(keys_in), _p(vals_in), val_base, _p(keys_out), _p(vals_out), n,
         begin_bit, end_bit, _p(ws), ws.numel(), _s())
    return keys_out, vals_out


# ---- A4 / A13 -----------------------------------------------------------
def scan_u32_inclusive(x):
    n = x.numel()
    out = torch.empty(n, dtype=torch.int32, device=x.device)
    ws = workspace(ws_bytes("ysql_scan_ws_bytes", n), x.device)
    call("ysql_scan_u32_inclusive", _p(x), _p(out), n, _p(ws), ws.numel(), _s())
    return out


def scan_mask_inclusive(mask):
    n = mask.numel()
    out = torch.empty(n, dtype=torch.int32, device=mask.device)
    ws = workspace(ws_bytes("ysql_scan_ws_bytes", n), mask.device)
    call("ysql_scan_mask_inclusive", _p(mask), _p(out), n, _p(ws), ws.numel(), _s())
    return out


# ---- A8-A10 -------------------------------------------------------------
def merge_path(dtype, descending, A, B, k_max=None, want_keys=True):
    """Merge path of sorted A and B; STRING runs are (rows, 256) uint8."""
    m, n = _rows(A), _rows(B)
    kmax = m + n if k_max is None else min(k_max, m + n)
    dev = A.device
    C_ = None
    if want_keys:
        C_ = torch.empty((kmax,) + tuple(A.shape[1:]), dtype=A.dtype, device=dev)
    idx = torch.empty(kmax, dtype=torch.int32, device=dev)
    isa = torch.empty(kmax, dtype=torch.uint8, device=dev)
    cons = torch.zeros(1, dtype=torch.int64, device=dev)
    ws = workspace(ws_bytes("ysql_merge_ws_bytes", m, n, kmax), dev)
    call("ysql_merge_path", dtype, int(descending), _p(A), m, _p(B), n, kmax, _p(C_), _p(idx), _p(isa), _p(cons),
         _p(ws), ws.numel(), _s())
    return C_, idx, isa, cons


def string_argsort(strs, descending=False):
    """Stable strcmp-order permutation of a (rows, 256) uint8 string column."""
    n = strs.shape[0]
    perm = torch.empty(n, dtype=torch.int32, device=strs.device)
    ws = workspace(ws_bytes("ysql_string_sort_ws_bytes", n) if n > 1 else 0, strs.device)
    call("ysql_string_argsort", int(descending), _p(strs), _p(perm), n, _p(ws), ws.numel(), _s())
    return perm


def string_sort(strs, perm, descending=False):
    """perm (int32 row indices of strs) reordered stably by strs[perm], in place."""
    n = strs.shape[0]
    assert perm.numel() == n
    ws = workspace(ws_bytes("ysql_string_sort_ws_bytes", n) if n > 1 else 0, strs.device)
    call("ysql_string_sort", int(descending), _p(strs), _p(perm), n, _p(ws), ws.numel(), _s())
    return perm


def merge_rows(keys, m, n, k_max=None, want_keys=False):
    """Stable merge of two runs sorted by key columns: keys = [(dtype code,
    descending, a column, b column)].  Returns (src_idx, src_is_a, consumed_a,
    merged strings or None)."""
    kmax = m + n if k_max is None else min(k_max, m + n)
    dev = keys[0][2].device
    arr = (L.SortCol * len(keys))(*[L.SortCol(t, int(d), a.data_ptr(), b.data_ptr()) for t, d, a, b in keys])
    idx = torch.empty(kmax, dtype=torch.int32, device=dev)
    isa = torch.empty(kmax, dtype=torch.uint8, device=dev)
    cons = torch.zeros(1, dtype=torch.int64, device=dev)
    out = torch.empty((kmax, L.STRING_BYTES), dtype=torch.uint8, device=dev) if want_keys else None
    call("ysql_merge_rows", arr, len(keys), m, n, kmax, _p(idx), _p(isa), _p(out), _p(cons), _s())
    return idx, isa, cons, out


def merge_pairs(a_keys, a_vals, b_keys, b_vals, descending=False, out_keys=None, out_vals=None, ws=None):
    m, n = a_keys.numel(), b_keys.numel()
    dev = a_keys.device
    if out_keys is None:
        out_keys = torch.empty(m + n, dtype=a_keys.dtype, device=dev)
    if out_vals is None:
        out_vals = torch.empty(m + n, dtype=torch.int32, device=dev)
    if ws is None:
        ws = workspace(ws_bytes("ysql_merge_ws_bytes", m, n, m + n), dev)
    call("ysql_merge_pairs", _kdt(a_keys), int(descending), _p(a_keys), _p(a_vals), m, _p(b_keys), _p(b_vals), n,
         _p(out_keys), _p(out_vals), _p(ws), ws.numel(), _s())
    return out_keys, out_vals


def merge_runs(keys, vals, run_offsets, out_keys=None, out_vals=None, ws=None):
    """k-way merge of the sorted runs keys[run_offsets[r]:run_offsets[r+1]]
    (ascending uint32/uint64 keys as int32/int64 tensors, uint32 payloads):
    ties by run, then position."""
    import ctypes as C

    nr = len(run_offsets) - 1
    offs = (C.c_uint64 * (nr + 1))(*[int(o) for o in run_offsets])
    n = int(run_offsets[-1]) - int(run_offsets[0])
    dev = keys.device
    if out_keys is None:
        out_keys = torch.empty(n, dtype=keys.dtype, device=dev)
    if out_vals is None:
        out_vals = torch.empty(n, dtype=torch.int32, device=dev)
    kdt = _kdt(keys)
    if ws is None:
        b = C.c_size_t()
        call("ysql_merge_runs_ws_bytes", kdt, offs, nr, C.byref(b))
        ws = workspace(b.value, dev)
    call("ysql_merge_runs", kdt, _p(keys), _p(vals), offs, nr, _p(out_keys), _p(out_vals), _p(ws), ws.numel(), _s())
    return out_keys, out_vals


# ---- row movement --------------------------------------------------------
def _null_out(src_null, n, dev):
    return torch.empty(n, dtype=torch.uint8, device=dev)


def gather(src, idx, src_null=None):
    n = idx.numel()
    dst = torch.empty((n,) + tuple(src.shape[1:]), dtype=src.dtype, device=src.device)
    dn = _null_out(src_null, n, src.device)
    call("ysql_gather", elem_bytes(src), _p(src), _p(src_null), _p(idx), n, _p(dst), _p(dn), _s())
    return dst, dn


def gather2(a, b, idx, is_a, reverse=False, a_null=None, b_null=None):
    n = idx.numel()
    dst = torch.empty((n,) + tuple(a.shape[1:]), dtype=a.dtype, device=a.device)
    dn = _null_out(None, n, a.device)
    call("ysql_gather2", elem_bytes(a), _p(a), _p(a_null), _p(b), _p(b_null), _p(idx), _p(is_a), n, int(reverse),
         _p(dst), _p(dn), _s())
    return dst, dn


def gather_pairs(src, pairs, side, n, src_null=None):
    dst = torch.empty((n,) + tuple(src.shape[1:]), dtype=src.dtype, device=src.device)
    dn = _null_out(src_null, n, src.device)
    call("ysql_gather_pairs", elem_bytes(src), _p(src), _p(src_null), _p(pairs), side, n, _p(dst), _p(dn), _s())
    return dst, dn


def compact(src, mask, mapping, count, src_null=None):
    n = mask.numel()
    dst = torch.empty((count,) + tuple(src.shape[1:]), dtype=src.dtype, device=src.device)
    dn = _null_out(src_null, count, src.device)
    call("ysql_compact", elem_bytes(src), _p(src), _p(src_null), _p(mask), _p(mapping), n, _p(dst), _p(dn), _s())
    return dst, dn


def _cols(srcs, nulls, n):
    outs, cols = [], (L.Col * max(1, len(srcs)))()
    for i, s in enumerate(srcs):
        d = torch.empty_like(s)
        dn = torch.empty(n, dtype=torch.uint8, device=s.device)
        sn = nulls[i] if nulls else None
        cols[i] = L.Col(s.data_ptr(), sn.data_ptr() if sn is not None else None, d.data_ptr(), dn.data_ptr(),
                        elem_bytes(s), 0)
        outs.append((d, dn))
    return cols, outs


def compact_cols(mask, srcs, nulls=None):
    """Fused mask -> scan -> compaction of every column.  Returns ([(dst, dst_null)], count)."""
    n = mask.numel()
    cols, outs = _cols(srcs, nulls, n)
    cnt = torch.zeros(1, dtype=torch.int64, device=mask.device)
    ws = workspace(ws_bytes("ysql_compact_ws_bytes", n), mask.device)
    call("ysql_compact_cols", _p(mask), cols, len(srcs), n, _p(cnt), _p(ws), ws.numel(), _s())
    c = int(cnt.item())
    return [(d[:c], dn[:c]) for d, dn in outs], c


def filter_compact(cmp, pred, scalar, srcs, nulls=None):
    n = pred.numel()
    cols, outs = _cols(srcs, nulls, n)
    cnt = torch.zeros(1, dtype=torch.int64, device=pred.device)
    ws = workspace(ws_bytes("ysql_compact_ws_bytes", n), pred.device)
    call("ysql_filter_compact", dtype_code(pred), cmp, _p(pred), _p(scalar), cols, len(srcs), n, _p(cnt), _p(ws),
         ws.numel(), _s())
    c = int(cnt.item())
    return [(d[:c], dn[:c]) for d, dn in outs], c


def filter_compact_range(pred, lo, hi, srcs, cap=None, retry=True):
    """Rows with lo <= pred < hi, every column of srcs compacted (no nulls).
    Output buffers hold `cap` rows (default n); on overflow the call is
    repeated with the exact count, or with retry=False the first `cap` rows
    in row order are returned.  Returns ([dst], count)."""
    n = pred.numel()
    dev = pred.device
    cap = n if cap is None else max(int(cap), 1)
    bounds = torch.tensor([lo, hi], dtype=pred.dtype, device=dev)
    outs = [torch.empty((cap,) + tuple(s.shape[1:]), dtype=s.dtype, device=dev) for s in srcs]
    cols = (L.Col * max(1, len(srcs)))()
    for i, (s, d) in enumerate(zip(srcs, outs)):
        cols[i] = L.Col(s.data_ptr(), None, d.data_ptr(), None, elem_bytes(s), 0)
    cnt = torch.zeros(1, dtype=torch.int64, device=dev)
    ws = workspace(ws_bytes("ysql_compact_ws_bytes", n), dev)
    call("ysql_filter_compact_range", dtype_code(pred), _p(pred), _p(bounds), cols, len(srcs), n, cap, _p(cnt),
         _p(ws), ws.numel(), _s())
    c = int(cnt.item())
    if c > cap and retry:
        return filter_compact_range(pred, lo, hi, srcs, cap=c)
    return [d[:min(c, cap)] for d in outs], c


# ---- fused expression programs -------------------------------------------
def xinsn(op, dst, a=0, b=0, type=0, sub=0, nsrc=-1, imm=0):
    """One program instruction (ysql_xinsn); FLOAT constants pass their bits in imm."""
    return L.XInsn(op, type, sub, dst, a, b, nsrc, imm)


def _xprog(insns, inputs, nulls):
    prog = (L.XInsn * max(1, len(insns)))(*insns)
    xin = (L.XInput * max(1, len(inputs)))()
    for i, t in enumerate(inputs):
        nl = nulls[i] if nulls else None
        xin[i] = L.XInput(t.data_ptr(), nl.data_ptr() if nl is not None else None, dtype_code(t), 0)
    return prog, xin


def expr_eval(insns, inputs, outputs, nulls=None, n=None):
    """Runs a program over n rows.  inputs: tensors (registers 0..); outputs:
    list of (register, torch dtype, with_nulls).  Returns [(values, nulls or None)]."""
    n = inputs[0].shape[0] if n is None else n
    dev = inputs[0].device
    prog, xin = _xprog(insns, inputs, nulls)
    res, xout = [], (L.XOutput * max(1, len(outputs)))()
    for i, (reg, dtype, with_nulls) in enumerate(outputs):
        d = torch.empty(n, dtype=dtype, device=dev)
        dn = torch.empty(n, dtype=torch.uint8, device=dev) if with_nulls else None
        xout[i] = L.XOutput(d.data_ptr(), dn.data_ptr() if dn is not None else None, reg, _DT[dtype])
        res.append((d, dn))
    call("ysql_expr_eval", prog, len(insns), xin, len(inputs), xout, len(outputs), n, _s())
    return res


def expr_filter_compact(insns, pred_reg, inputs, srcs, nulls=None, src_nulls=None):
    """Rows where BOOL register pred_reg holds, every column of srcs compacted.
    Returns ([(dst, dst_null)], count)."""
    n = inputs[0].shape[0]
    dev = inputs[0].device
    prog, xin = _xprog(insns, inputs, nulls)
    cols, outs = _cols(srcs, src_nulls, n)
    cnt = torch.zeros(1, dtype=torch.int64, device=dev)
    ws = workspace(ws_bytes("ysql_compact_ws_bytes", n), dev)
    call("ysql_expr_filter_compact", prog, len(insns), pred_reg, xin, len(inputs), cols, len(srcs), n, _p(cnt),
         _p(ws), ws.numel(), _s())
    c = int(cnt.item())
    return [(d[:c], dn[:c]) for d, dn in outs], c


# ---- A12 / A15 ---------------------------------------------------------
def compare(cmp, lhs, rhs, lhs_scalar=False, rhs_scalar=False, negate=False, n=None):
    n = n if n is not None else (rhs.shape[0] if lhs_scalar else lhs.shape[0])
    dev = lhs.device
    out = torch.empty(n, dtype=torch.uint8, device=dev)
    dt = dtype_code(lhs if not lhs_scalar else rhs) if lhs.dim() == 2 or rhs.dim() == 2 else _DT[lhs.dtype]
    call("ysql_compare", dt, cmp, _p(lhs), int(lhs_scalar), _p(rhs), int(rhs_scalar), int(negate), _p(out), n, _s())
    return out


def mask_logic(op, a, b, a_scalar=False, b_scalar=False, n=None):
    n = n if n is not None else (b.numel() if a_scalar else a.numel())
    out = torch.empty(n, dtype=torch.uint8, device=a.device)
    call("ysql_mask_logic", op, _p(a), int(a_scalar), _p(b), int(b_scalar), _p(out), n, _s())
    return out


def arith(op, lhs, rhs, lhs_scalar=False, rhs_scalar=False, n=None):
    n = n if n is not None else (rhs.numel() if lhs_scalar else lhs.numel())
    out = torch.empty(n, dtype=lhs.dtype, device=lhs.device)
    call("ysql_arith", _DT[lhs.dtype], op, _p(lhs), int(lhs_scalar), _p(rhs), int(rhs_scalar), _p(out), n, _s())
    return out


def cast(x, dst_dtype):
    out = torch.empty(x.shape, dtype=dst_dtype, device=x.device)
    call("ysql_cast", _DT[x.dtype], _DT[dst_dtype], _p(x), _p(out), x.numel(), _s())
    return out


# ---- A16 / A17 ---------------------------------------------------------
ACC_DTYPE = torch.uint8  # accumulators are raw 32-byte structs


def acc_new(dtype_c, agg, nacc=1, device="cuda"):
    acc = torch.empty(32 * nacc, dtype=torch.uint8, device=device)
    call("ysql_acc_init", dtype_c, agg, _p(acc), nacc, _s())
    return acc


def acc_read(acc):
    """Host view of accumulators: list of dicts."""
    raw = acc.cpu().numpy().tobytes()
    out = []
    for k in range(len(raw) // 32):
        a = L.Acc.from_buffer_copy(raw[32 * k:32 * k + 32])
        out.append({"i": a.i, "d": a.d, "f": a.f, "count": a.count})
    return out


def reduce(dtype_c, agg, x, acc, null=None, n=None):
    n = x.numel() if n is None else n
    ws = workspace(ws_bytes("ysql_reduce_ws_bytes", n), acc.device)
    call("ysql_reduce", dtype_c, agg, _p(x), _p(null), n, _p(acc), _p(ws), ws.numel(), _s())
    return acc


def filter_aggregate(pred, cmp, scalar, specs, overwrite=False, n=None, generic=False):
    """Fused WHERE pred <cmp> scalar + aggregates.  specs: list of
    (dtype_c, agg, column tensor or None, acc tensor).  With overwrite the
    accumulators receive this call's result; otherwise it is folded in.
    generic: run the general tile kernel (same bits as the two-column one)."""
    n = pred.numel() if n is None else n
    dev = pred.device
    sc = torch.tensor([scalar], dtype=pred.dtype, device=dev)
    arr = (L.AggSpec * len(specs))()
    for k, (dc, agg, col, acc) in enumerate(specs):
        arr[k] = L.AggSpec(dc, agg, _p(col) if col is not None else None, _p(acc))
    ws = workspace(ws_bytes("ysql_filter_aggregate_ws_bytes", n, len(specs)), dev)
    flags = (L.FA_OVERWRITE if overwrite else 0) | (L.FA_GENERIC if generic else 0)
    call("ysql_filter_aggregate_ex", _DT[pred.dtype], cmp, _p(pred), _p(sc), arr, len(specs), n, flags, _p(ws),
         ws.numel(), _s())
    return [acc for (_, _, _, acc) in specs]


def filter_aggregate_multi(pred, cmp, scalars, columns, accs, overwrite=True, n=None, generic=False):
    """Shared scan: query q is WHERE pred <cmp> scalars[q] with aggregates
    columns[s] = (dtype_c, agg, column or None) into accs[q][s]."""
    n = pred.numel() if n is None else n
    dev = pred.device
    nq, ns = len(scalars), len(columns)
    sc = torch.tensor(list(scalars), dtype=pred.dtype, device=dev)
    arr = (L.AggSpec * (nq * ns))()
    for q in range(nq):
        for k, (dc, agg, col) in enumerate(columns):
            arr[q * ns + k] = L.AggSpec(dc, agg, _p(col) if col is not None else None, _p(accs[q][k]))
    ws = workspace(ws_bytes("ysql_filter_aggregate_multi_ws_bytes", n, nq, ns), dev)
    flags = (L.FA_OVERWRITE if overwrite else 0) | (L.FA_GENERIC if generic else 0)
    call("ysql_filter_aggregate_multi", _DT[pred.dtype], cmp, _p(pred), _p(sc), nq, arr, ns, n, flags, _p(ws),
         ws.numel(), _s())
    return accs


def agg_finalize(dtype_c, agg, acc):
    out = torch.zeros(8, dtype=torch.uint8, device=acc.device)
    call("ysql_agg_finalize", dtype_c, agg, _p(acc), _p(out), _s())
    return out


def group_reduce(dtype_c, agg, gid, n_groups, x, acc, null=None):
    call("ysql_group_reduce", dtype_c, agg, _p(gid), n_groups, _p(x), _p(null), gid.numel(), _p(acc), _s())
    return acc


# ---- A18-A21 -------------------------------------------------------------
def join_pairs(cmp, left, right, cap=None):
    """Returns (pairs[count, 2] int32, count).  Retries once with the exact capacity on overflow."""
    Ln, Rn = left.shape[0], right.shape[0]
    dt = dtype_code(left)
    dev = left.device
    ws = workspace(ws_bytes("ysql_join_ws_bytes", dt, cmp, Ln, Rn), dev)
    cnt = torch.zeros(1, dtype=torch.int64, device=dev)
    cap = max(Ln, 1024) if cap is None else cap
    pairs = torch.empty((max(cap, 1), 2), dtype=torch.int32, device=dev)
    call("ysql_join_pairs", dt, cmp, _p(left), Ln, _p(right), Rn, _p(pairs), cap, _p(cnt), _p(ws), ws.numel(), _s())
    c = int(cnt.item())
    if c > cap:
        return join_pairs(cmp, left, right, cap=c)
    return pairs[:c], c


def pairs_intersect(p1, p2, cap=None):
    n1, n2 = p1.shape[0], p2.shape[0]
    dev = p1.device
    cap = max(n1, 1) if cap is None else cap
    ws = workspace(ws_bytes("ysql_pairs_intersect_ws_bytes", n1, n2), dev)
    cnt = torch.zeros(1, dtype=torch.int64, device=dev)
    out = torch.empty((max(cap, 1), 2), dtype=torch.int32, device=dev)
    call("ysql_pairs_intersect", _p(p1), n1, _p(p2), n2, _p(out), cap, _p(cnt), _p(ws), ws.numel(), _s())
    c = int(cnt.item())
    return out[:c], c


def cross_expand(src, Ln, Rn, is_left, src_null=None):
    total = Ln * Rn
    dst = torch.empty((total,) + tuple(src.shape[1:]), dtype=src.dtype, device=src.device)
    dn = torch.empty(total, dtype=torch.uint8, device=src.device)
    call("ysql_cross_expand", elem_bytes(src), _p(src), _p(src_null), Ln, Rn, int(is_left), _p(dst), _p(dn), _s())
    return dst, dn


def star_join_group_sum(dims, m1, m2, n_groups, sums=None, counts=None):
    """dims: list of dicts {fk, rows, pass (uint8 or None), group (int32 or None), group_base, group_card}."""
    dev = m1.device
    arr = (L.StarDim * len(dims))()
    for i, d in enumerate(dims):
        arr[i] = L.StarDim(d["fk"].data_ptr(), d["rows"], d.get("group_card", 0),
                           d["pass"].data_ptr() if d.get("pass") is not None else None,
                           d["group"].data_ptr() if d.get("group") is not None else None,
                           d.get("group_base", 0), 0)
    if sums is None:
        sums = torch.zeros(n_groups, dtype=torch.int64, device=dev)
    if counts is None:
        counts = torch.zeros(n_groups, dtype=torch.int64, device=dev)
    call("ysql_star_join_group_sum", arr, len(dims), _p(m1), _p(m2), m1.numel(), _p(sums), _p(counts), n_groups, _s())
    return sums, counts


def star_join_group_sum_multi(dims, m1, m2, n_groups, nq, sums=None, counts=None):
    """nq star-join queries in one read of the facts: bit q of each dims[d]
    "pass" byte is query q's filter.  sums / counts: (nq, n_groups)."""
    dev = m1.device
    arr = (L.StarDim * len(dims))()
    for i, d in enumerate(dims):
        arr[i] = L.StarDim(d["fk"].data_ptr(), d["rows"], d.get("group_card", 0),
                           d["pass"].data_ptr() if d.get("pass") is not None else None,
                           d["group"].data_ptr() if d.get("group") is not None else None,
                           d.get("group_base", 0), 0)
    if sums is None:
        sums = torch.zeros((nq, n_groups), dtype=torch.int64, device=dev)
    if counts is None:
        counts = torch.zeros((nq, n_groups), dtype=torch.int64, device=dev)
    call("ysql_star_join_group_sum_multi", arr, len(dims), nq, _p(m1), _p(m2), m1.numel(), _p(sums), _p(counts),
         n_groups, _s())
    return sums, counts
