// ysql_kernels.hpp — the reference's kernel-launcher shapes, over the C ABI.
//
// The reference's operators call templated host launchers in namespace
// YallaSQL::Kernel (reference include/kernels/*.hpp).  This header gives the
// same names and argument meaning on top of include/ysql_kernels.h, with
// hipStream_t in place of cudaStream_t, so those call sites compile against
// the gfx950 library unchanged apart from the stream type.  Every wrapper
// throws YallaSQL::Kernel::KernelError on a nonzero status (the reference
// exited the process, utils/macros.hpp:6-22).
//
// Memory: the C ABI under these wrappers never allocates.  The wrappers
// take the workspaces the C ABI asks for as stream-ordered temporaries
// (hipMallocAsync / hipFreeAsync on the caller's stream), because the
// reference's signatures carry no workspace; the batch-level helpers
// allocate their output columns the same way, as the reference's did.
// launch_radix_sort still returns a device index buffer by reference, now
// the caller's to free with hipFreeAsync; the reference leaked it and every
// temporary (radix_sort_kernel.cu:139-158).
//
// Where the reference silently wrote past a fixed buffer, the wrapper
// throws instead: launch_outer_join_operators throws KernelError with
// YSQL_EOVERFLOW when the pair count exceeds the pair buffer's capacity
// (the reference's JoinOperator then gathered with that count,
// join_operator.hpp:86-89).
#pragma once

#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "ysql_kernels.h"

namespace YallaSQL::Kernel {

struct KernelError : std::runtime_error {
  int code;
  KernelError(const char *fn, int rc)
      : std::runtime_error(std::string(fn) + " failed with status " + std::to_string(rc) + ": " + ysql_last_error()),
        code(rc) {}
};

inline void check(const char *fn, int rc) {
  if (rc != YSQL_OK) throw KernelError(fn, rc);
}

// Stream-ordered temporary device buffer.
struct TempBuf {
  void *p = nullptr;
  hipStream_t s;
  TempBuf(size_t bytes, hipStream_t st) : s(st) {
    if (hipMallocAsync(&p, bytes ? bytes : 1, s) != hipSuccess) throw KernelError("hipMallocAsync", YSQL_EHIP);
  }
  ~TempBuf() {
    if (p) (void)hipFreeAsync(p, s);
  }
  TempBuf(const TempBuf &) = delete;
  TempBuf &operator=(const TempBuf &) = delete;
};

// ---- operand tags (comparison_operators_kernel.hpp:25-104,
//      binary_operators_kernel.hpp:27-51, reduction_kernel.hpp:18-37,
//      radix_sort_kernel.hpp:31-45, merge_batchs.hpp:20-34) -------------
enum class OperandType : uint8_t { SCALAR = 0, VECTOR };
template <typename T> struct LEOperator { static constexpr int code = YSQL_LT; };
template <typename T> struct LEQOperator { static constexpr int code = YSQL_LE; };
template <typename T> struct GEOperator { static constexpr int code = YSQL_GT; };
template <typename T> struct GEQOperator { static constexpr int code = YSQL_GE; };
template <typename T> struct EQOperator { static constexpr int code = YSQL_EQ; };
template <typename T> struct NEQOperator { static constexpr int code = YSQL_NE; };
struct ANDOperator { static constexpr int code = YSQL_AND; };
struct OROperator { static constexpr int code = YSQL_OR; };
template <typename T> struct AddOperator { static constexpr int code = YSQL_ADD; };
template <typename T> struct MinusOperator { static constexpr int code = YSQL_SUB; };
template <typename T> struct MulOperator { static constexpr int code = YSQL_MUL; };
template <typename T> struct DivOperator { static constexpr int code = YSQL_DIV; };
template <typename T> struct RemOperator { static constexpr int code = YSQL_REM; };
template <typename T> struct MaxOperator { static constexpr int code = YSQL_AGG_MAX; };
template <typename T> struct MinOperator { static constexpr int code = YSQL_AGG_MIN; };
template <typename T> struct SumOperator { static constexpr int code = YSQL_AGG_SUM; };
template <typename T> struct AscOp { static constexpr bool desc = false, is_string = false; };
template <typename T> struct DescOp { static constexpr bool desc = true, is_string = false; };
template <typename T> struct AscSortMergeOp { static constexpr bool desc = false, is_string = false; };
template <typename T> struct DescSortMergeOp { static constexpr bool desc = true, is_string = false; };
// STRING orders (radix_sort_kernel.hpp:45-73, merge_batchs.hpp:36-48): the
// element type is the reference's 256-byte String.
struct AscOpString { static constexpr bool desc = false, is_string = true; };
struct DescOpString { static constexpr bool desc = true, is_string = true; };
struct AscSortStrOp { static constexpr bool desc = false, is_string = true; };
struct DescSortStrOp { static constexpr bool desc = true, is_string = true; };

template <typename T> struct TypeCode;
template <> struct TypeCode<int> { static constexpr int v = YSQL_INT; };
template <> struct TypeCode<float> { static constexpr int v = YSQL_FLOAT; };
template <> struct TypeCode<int64_t> { static constexpr int v = YSQL_DATETIME; };
template <> struct TypeCode<bool> { static constexpr int v = YSQL_BOOL; };
template <> struct TypeCode<uint32_t> { static constexpr int v = YSQL_UKEY32; };
template <> struct TypeCode<uint64_t> { static constexpr int v = YSQL_UKEY64; };
template <> struct TypeCode<double> { static constexpr int v = YSQL_FLOAT; };  // a FLOAT column's double sum

// Element bytes of a DataType code (data_type_utils.hpp:7-13); 0 = unknown.
inline uint32_t type_bytes(int type) {
  switch (type) {
    case YSQL_INT:
    case YSQL_FLOAT: return 4;
    case YSQL_DATETIME: return 8;
    case YSQL_STRING: return YSQL_STRING_BYTES;
    default: return 0;
  }
}


// ---- A1: radix_sort_utils.cu:39-53 ---------------------------------------
inline void launch_int_to_uint32(int *arr, uint32_t *res, int N, hipStream_t s = nullptr) {
  check("ysql_sortkey_encode", ysql_sortkey_encode(YSQL_INT, 0, arr, res, (uint64_t)N, s));
}
inline void launch_float_to_uint32(float *arr, uint32_t *res, int N, hipStream_t s = nullptr) {
  check("ysql_sortkey_encode", ysql_sortkey_encode(YSQL_FLOAT, 0, arr, res, (uint64_t)N, s));
}
inline void launch_uint32_to_int(uint32_t *arr, int *res, int N, hipStream_t s = nullptr) {
  check("ysql_sortkey_decode", ysql_sortkey_decode(YSQL_INT, 0, arr, res, (uint64_t)N, s));
}
inline void launch_uint32_to_float(uint32_t *arr, float *res, int N, hipStream_t s = nullptr) {
  check("ysql_sortkey_decode", ysql_sortkey_decode(YSQL_FLOAT, 0, arr, res, (uint64_t)N, s));
}

// ---- A2-A6: radix_sort_kernel.cu:139-178 ---------------------------------
// T = uint32_t or uint64_t keys (already normalised); OP = AscOp / DescOp.
// T = the 256-byte String with OP = AscOpString / DescOpString sorts the
// strings in place (strcmp order, stable) through ysql_string_argsort and a
// gather.
template <typename OP, typename T, size_t BlockDim>
void launch_radix_sort(T *d_input, uint32_t *&d_idxs_in, const uint32_t N, hipStream_t stream) {
  if constexpr (OP::is_string) {
    static_assert(sizeof(T) == YSQL_STRING_BYTES, "String is 256 bytes");
    size_t ws_bytes = 0;
    if (N > 1) check("ysql_string_sort_ws_bytes", ysql_string_sort_ws_bytes(N, &ws_bytes));
    TempBuf ws(ws_bytes, stream), sorted((size_t)N * YSQL_STRING_BYTES, stream);
    if (hipMallocAsync(&d_idxs_in, (size_t)N * sizeof(uint32_t) + 4, stream) != hipSuccess)
      throw KernelError("hipMallocAsync", YSQL_EHIP);
    check("ysql_string_argsort", ysql_string_argsort(OP::desc, d_input, d_idxs_in, N, ws.p, ws_bytes, stream));
    check("ysql_gather", ysql_gather(YSQL_STRING_BYTES, d_input, nullptr, d_idxs_in, N, sorted.p, nullptr, stream));
    if (N && hipMemcpyAsync(d_input, sorted.p, (size_t)N * YSQL_STRING_BYTES, hipMemcpyDeviceToDevice, stream) !=
                 hipSuccess)
      throw KernelError("hipMemcpyAsync", YSQL_EHIP);
  } else {
    size_t ws_bytes = 0;
    check("ysql_radix_sort_ws_bytes", ysql_radix_sort_ws_bytes(TypeCode<T>::v, N, &ws_bytes));
    TempBuf ws(ws_bytes, stream);
    if (hipMallocAsync(&d_idxs_in, (size_t)N * sizeof(uint32_t) + 4, stream) != hipSuccess)
      throw KernelError("hipMallocAsync", YSQL_EHIP);
    check("ysql_radix_argsort", ysql_radix_argsort(TypeCode<T>::v, OP::desc, d_input, d_idxs_in, N, 0, 8 * sizeof(T),
                                                   ws.p, ws_bytes, stream));
  }
}

// ---- A4 / A13: prefix_sum_kernel.cu:140-178 --------------------------------
inline void launch_prefix_sum_mask(bool *mask, uint32_t *out, uint32_t sz, hipStream_t stream) {
  size_t b = 0;
  check("ysql_scan_ws_bytes", ysql_scan_ws_bytes(sz, &b));
  TempBuf ws(b, stream);
  check("ysql_scan_mask_inclusive",
        ysql_scan_mask_inclusive(reinterpret_cast<const uint8_t *>(mask), out, sz, ws.p, b, stream));
}
template <typename T>
void launch_prefix_sum(T *in, T *out, uint32_t sz, hipStream_t stream) {
  static_assert(sizeof(T) == 4, "uint32 scan");
  size_t b = 0;
  check("ysql_scan_ws_bytes", ysql_scan_ws_bytes(sz, &b));
  TempBuf ws(b, stream);
  check("ysql_scan_u32_inclusive", ysql_scan_u32_inclusive(reinterpret_cast<const uint32_t *>(in),
                                                           reinterpret_cast<uint32_t *>(out), sz, ws.p, b, stream));
}

// ---- A8-A9: merge_batchs.cu:143-158 (lasti receives corank(k_mx)) --------
template <typename T, typename OP>
void launch_merge_sorted_array_kernel(T *A, T *B, T *C, uint32_t *new_idx, bool *table_idx, int *lasti,
                                      const uint32_t k_mx, uint32_t m, uint32_t n, hipStream_t stream) {
  size_t b = 0;
  check("ysql_merge_ws_bytes", ysql_merge_ws_bytes(m, n, k_mx, &b));
  TempBuf ws(b, stream), cons(8, stream);
  int dtype;
  if constexpr (OP::is_string) {
    static_assert(sizeof(T) == YSQL_STRING_BYTES, "String is 256 bytes");
    dtype = YSQL_STRING;
  } else {
    dtype = TypeCode<T>::v;
  }
  check("ysql_merge_path", ysql_merge_path(dtype, OP::desc, A, m, B, n, k_mx, C, new_idx,
                                           reinterpret_cast<uint8_t *>(table_idx), static_cast<int64_t *>(cons.p),
                                           ws.p, b, stream));
  if (lasti) {  // consumed_a <= m fits the int the reference reads back (order_operator.cpp:175)
    int64_t h = 0;
    if (hipMemcpyAsync(&h, cons.p, 8, hipMemcpyDeviceToHost, stream) != hipSuccess ||
        hipStreamSynchronize(stream) != hipSuccess)
      throw KernelError("hipMemcpyAsync", YSQL_EHIP);
    const int v = (int)h;
    if (hipMemcpyAsync(lasti, &v, sizeof(int), hipMemcpyHostToDevice, stream) != hipSuccess ||
        hipStreamSynchronize(stream) != hipSuccess)
      throw KernelError("hipMemcpyAsync", YSQL_EHIP);
  }
}

// ---- A7 / A10 / A14 / A20: move_rows_kernel.cu:136-180 ---------------------
template <typename T>
void launch_move_one_col(const T *src, T *res, const char *src_isnull, char *res_isnull, uint32_t *map,
                         const uint32_t srcSz, hipStream_t stream) {
  check("ysql_gather", ysql_gather(sizeof(T), src, reinterpret_cast<const uint8_t *>(src_isnull), map, srcSz, res,
                                   reinterpret_cast<uint8_t *>(res_isnull), stream));
}
template <typename T>
void launch_move_one_col_merge(const T *Asrc, const T *Bsrc, T *res, const char *Aisnull, const char *Bisnull,
                               char *res_isnull, const uint32_t *map, const bool *tableIdxs, bool isAsc,
                               const uint32_t srcSz, hipStream_t stream) {
  check("ysql_gather2", ysql_gather2(sizeof(T), Asrc, reinterpret_cast<const uint8_t *>(Aisnull), Bsrc,
                                     reinterpret_cast<const uint8_t *>(Bisnull), map,
                                     reinterpret_cast<const uint8_t *>(tableIdxs), srcSz, isAsc ? 0 : 1, res,
                                     reinterpret_cast<uint8_t *>(res_isnull), stream));
}
template <typename T>
void launch_move_rows_filter_kernel(T *src, T *res, uint32_t *map, bool *mask, char *isnull, const uint32_t srcSz,
                                    hipStream_t stream, char *res_isnull = nullptr) {
  check("ysql_compact", ysql_compact(sizeof(T), src, reinterpret_cast<const uint8_t *>(isnull),
                                     reinterpret_cast<const uint8_t *>(mask), map, srcSz, res,
                                     reinterpret_cast<uint8_t *>(res_isnull), stream));
}

// ---- A18 / A19: comparison_operators_kernel.cu:135-167 ---------------------
// The pair buffer holds `cap` pairs (the reference's was fixed at
// MAX_ROWS_OUT_JOIN_OP = 1,048,576 and unchecked, comparison_join.hpp:87);
// *actualSz receives the full count, so actualSz > cap means overflow.
template <typename T, typename Op>
void launch_outer_join_operators(T *d_rhs, T *d_lhs, uint32_t *pairs, int *actualSz, unsigned int left_rows,
                                 unsigned int right_rows, hipStream_t stream, uint64_t cap = 1048576) {
  size_t b = 0;
  check("ysql_join_ws_bytes", ysql_join_ws_bytes(TypeCode<T>::v, Op::code, left_rows, right_rows, &b));
  TempBuf ws(b, stream), cnt(8, stream);
  if (cap > (uint64_t)INT_MAX) cap = (uint64_t)INT_MAX;  // actualSz is an int
  check("ysql_join_pairs", ysql_join_pairs(TypeCode<T>::v, Op::code, d_lhs, left_rows, d_rhs, right_rows, pairs, cap,
                                           static_cast<uint64_t *>(cnt.p), ws.p, b, stream));
  uint64_t total = 0;
  if (hipMemcpyAsync(&total, cnt.p, 8, hipMemcpyDeviceToHost, stream) != hipSuccess ||
      hipStreamSynchronize(stream) != hipSuccess)
    throw KernelError("hipMemcpyAsync", YSQL_EHIP);
  // only the first `cap` pairs were written: a caller that gathered with the
  // full count would read past them
  if (total > cap) throw KernelError("launch_outer_join_operators (pair buffer too small)", YSQL_EOVERFLOW);
  const int v = (int)total;
  if (hipMemcpyAsync(actualSz, &v, sizeof(int), hipMemcpyHostToDevice, stream) != hipSuccess ||
      hipStreamSynchronize(stream) != hipSuccess)
    throw KernelError("hipMemcpyAsync", YSQL_EHIP);
}
inline void launch_and_join_operators(const uint32_t *l_pairs, const uint32_t *r_pairs, uint32_t *pairs,
                                      uint32_t * /*mask: not needed*/, int *actualSz, unsigned int left_rows,
                                      unsigned int right_rows, hipStream_t stream) {
  size_t b = 0;
  check("ysql_pairs_intersect_ws_bytes", ysql_pairs_intersect_ws_bytes(left_rows, right_rows, &b));
  TempBuf ws(b, stream), cnt(8, stream);
  check("ysql_pairs_intersect", ysql_pairs_intersect(l_pairs, left_rows, r_pairs, right_rows, pairs, left_rows,
                                                     static_cast<uint64_t *>(cnt.p), ws.p, b, stream));
  // the intersection holds at most left_rows pairs: its count fits an int;
  // low 32 bits of the little-endian count
  if (hipMemcpyAsync(actualSz, cnt.p, sizeof(int), hipMemcpyDeviceToDevice, stream) != hipSuccess)
    throw KernelError("hipMemcpyAsync", YSQL_EHIP);
}
template <typename T>
void lanch_move_rows_join_kernel(const T *src, T *res, const uint32_t *pairs, const char *isnull, char *out_isnull,
                                 const uint32_t batchSize, const bool isright, hipStream_t stream) {
  check("ysql_gather_pairs", ysql_gather_pairs(sizeof(T), src, reinterpret_cast<const uint8_t *>(isnull), pairs,
                                               isright ? 1 : 0, batchSize, res,
                                               reinterpret_cast<uint8_t *>(out_isnull), stream));
}

// ---- A12: comparison_operators_kernel.cu:114-132 (op(lhs, rhs)) ----------
template <typename T, typename Op>
void launch_conditional_operators(T *d_rhs, T *d_lhs, OperandType t_rhs, OperandType t_lhs, bool *d_res,
                                  unsigned int sz, hipStream_t &stream, bool isneg = false) {
  if constexpr (std::is_same<T, bool>::value) {
    check("ysql_mask_logic", ysql_mask_logic(Op::code, reinterpret_cast<const uint8_t *>(d_lhs),
                                             t_lhs == OperandType::SCALAR, reinterpret_cast<const uint8_t *>(d_rhs),
                                             t_rhs == OperandType::SCALAR, reinterpret_cast<uint8_t *>(d_res), sz,
                                             stream));
  } else {
    check("ysql_compare", ysql_compare(TypeCode<T>::v, Op::code, d_lhs, t_lhs == OperandType::SCALAR, d_rhs,
                                       t_rhs == OperandType::SCALAR, isneg, reinterpret_cast<uint8_t *>(d_res), sz,
                                       stream));
  }
}

// ---- A15: binary_operators_kernel.cu:58-75, cast_operators_kernel.cu:24-43 -
template <typename T, typename Op>
void launch_binary_operators(T *d_rhs, T *d_lhs, OperandType t_rhs, OperandType t_lhs, T *d_res, unsigned int sz,
                             hipStream_t &stream) {
  check("ysql_arith", ysql_arith(TypeCode<T>::v, Op::code, d_lhs, t_lhs == OperandType::SCALAR, d_rhs,
                                 t_rhs == OperandType::SCALAR, d_res, sz, stream));
}
template <typename S, typename D>
void launch_numerical_cast(S *src, D *dist, const unsigned int sz, hipStream_t &stream) {
  constexpr int sd = std::is_same<S, bool>::value ? YSQL_BOOL : TypeCode<S>::v;
  constexpr int dd = std::is_same<D, unsigned int>::value ? YSQL_INT : TypeCode<D>::v;
  check("ysql_cast", ysql_cast(sd, dd, src, dist, sz, stream));
}

// ---- A21: cross_product_kernel.cu:29-40 ------------------------------------
template <typename T>
void launch_cross_product_col(const T *col, T *out_col, const char *src_nullset, char *out_nullset,
                              const uint32_t leftBs, const uint32_t rightBs, hipStream_t stream, const bool isleft) {
  check("ysql_cross_expand", ysql_cross_expand(sizeof(T), col, reinterpret_cast<const uint8_t *>(src_nullset), leftBs,
                                               rightBs, isleft, out_col, reinterpret_cast<uint8_t *>(out_nullset),
                                               stream));
}

// ---- A21: type dispatch, cross_product_kernel.hpp:13-127 ------------------
// `type` is the engine's DataType (INT, FLOAT, DATETIME, STRING); like the
// reference, an unknown type launches nothing.
template <typename DataTypeT>
void launch_cross_product_type(void *src, void *dist, char *src_nullset, char *dist_nullset, DataTypeT type,
                               const int lbs, const int rbs, hipStream_t stream, bool isleft) {
  const uint32_t eb = type_bytes((int)type);
  if (!eb) return;
  check("ysql_cross_expand", ysql_cross_expand(eb, src, reinterpret_cast<const uint8_t *>(src_nullset), (uint64_t)lbs,
                                               (uint64_t)rbs, isleft, dist,
                                               reinterpret_cast<uint8_t *>(dist_nullset), stream));
}
// Same, starting `offset` rows into the source column (the right side's
// chunk of MAX_RIGHT_SZ rows, cross_product_operator.hpp:87-125).
template <typename DataTypeT>
void launch_cross_product_one_type(void *src, void *dist, uint32_t offset, char *src_nullset, char *dist_nullset,
                                   DataTypeT type, const int lbs, const int rbs, hipStream_t stream, bool isleft) {
  const uint32_t eb = type_bytes((int)type);
  if (!eb) return;
  check("ysql_cross_expand",
        ysql_cross_expand(eb, static_cast<char *>(src) + (size_t)offset * eb,
                          src_nullset ? reinterpret_cast<const uint8_t *>(src_nullset) + offset : nullptr,
                          (uint64_t)lbs, (uint64_t)rbs, isleft, dist, reinterpret_cast<uint8_t *>(dist_nullset),
                          stream));
}

// ---- A16 / A17: reduction_kernel.hpp:7-15 -> reduction_kernel.cu:124-205 ---
// Folds a batch into the caller's typed accumulator *res (device), as the
// reference does per batch (aggregate.hpp:79-137): T_res = int (INT),
// float (FLOAT MIN/MAX), double (FLOAT SUM), int64_t (DATETIME).  NULL rows
// are skipped (the reference folded `inital` for them, the same value when
// it is the operator's identity).  The fold order is fixed, so a float sum
// has the same bits on every run (the reference's block atomics did not).
template <typename T, typename T_res, typename Op>
void launch_reduction_operators(T *d_arr, T_res *res, char *isnull, const uint32_t sz, hipStream_t &stream,
                                const T_res /*inital*/) {
  constexpr int dt = TypeCode<T>::v;
  static_assert(dt == YSQL_INT || dt == YSQL_FLOAT || dt == YSQL_DATETIME, "INT, FLOAT or DATETIME column");
  static_assert(sizeof(T_res) == ((dt == YSQL_FLOAT && Op::code == YSQL_AGG_SUM) ? 8 : sizeof(T)),
                "the reference's accumulator type for this column and operator");
  size_t b = 0;
  check("ysql_reduce_ws_bytes", ysql_reduce_ws_bytes(sz, &b));
  TempBuf ws(b, stream), acc(sizeof(ysql_acc), stream);
  ysql_acc *a = static_cast<ysql_acc *>(acc.p);
  check("ysql_acc_from_value", ysql_acc_from_value(dt, Op::code, res, a, stream));
  check("ysql_reduce", ysql_reduce(dt, Op::code, d_arr, reinterpret_cast<const uint8_t *>(isnull), sz, a, ws.p, b,
                                   stream));
  check("ysql_acc_to_value", ysql_acc_to_value(dt, Op::code, a, res, stream));
}
// COUNT: adds the batch's non-NULL rows to *res (an int32 count; d_arr is
// not read, the reference passed the null set itself, aggregate.hpp:116-119).
inline void launch_reduction_count_notnull(char * /*d_arr*/, int *res, char *isnull, const uint32_t sz,
                                           hipStream_t &stream, const int /*inital*/) {
  size_t b = 0;
  check("ysql_reduce_ws_bytes", ysql_reduce_ws_bytes(sz, &b));
  TempBuf ws(b, stream), acc(sizeof(ysql_acc), stream);
  ysql_acc *a = static_cast<ysql_acc *>(acc.p);
  check("ysql_acc_from_value", ysql_acc_from_value(YSQL_INT, YSQL_AGG_COUNT, res, a, stream));
  check("ysql_reduce", ysql_reduce(YSQL_INT, YSQL_AGG_COUNT, nullptr, reinterpret_cast<const uint8_t *>(isnull), sz,
                                   a, ws.p, b, stream));
  check("ysql_acc_to_value", ysql_acc_to_value(YSQL_INT, YSQL_AGG_COUNT, a, res, stream));
}
// AVG = (double)sum / (double)count -> float; T = int (INT), double (FLOAT),
// int64_t (DATETIME).  The reference launched these two on the default stream.
template <typename T>
void launch_div_avg(const int *counter, const T *sum, float *avg) {
  check("ysql_div_avg", ysql_div_avg(TypeCode<T>::v, counter, sum, avg, nullptr));
}
inline void launch_convert_double_to_float_kernel(const double *input, float *output) {
  check("ysql_double_to_float", ysql_double_to_float(input, output, nullptr));
}

// ---- A7 / A10: batch-level gathers, move_rows_kernel.hpp:51-143 ------------
// BatchT is the engine's Batch (reference include/batch.hpp): columnData,
// nullset[c]->bitset (one byte per row), columns[c]->bytes, batchSize,
// stream, `location`, and the constructor
// Batch(columnData, Device::GPU, rows, columns, nullset, stream); its null
// set type has the constructor NullBitSet(rows, stream).  New columns and
// null sets are allocated on the batch's stream, as the reference did.
template <typename BatchT>
std::unique_ptr<BatchT> move_rows_batch(BatchT &src_batch, uint32_t *map) {
  using NullT = typename std::decay<decltype(*src_batch.nullset[0])>::type;
  using DeviceT = decltype(src_batch.location);
  const size_t ncols = src_batch.columns.size();
  const uint32_t rows = (uint32_t)src_batch.batchSize;
  hipStream_t stream = src_batch.stream;
  std::vector<void *> data(ncols);
  std::vector<std::shared_ptr<NullT>> nullset(ncols);
  for (size_t c = 0; c < ncols; c++) {
    const uint32_t eb = src_batch.columns[c]->bytes;
    if (hipMallocAsync(&data[c], (size_t)rows * eb, stream) != hipSuccess)
      throw KernelError("hipMallocAsync", YSQL_EHIP);
    nullset[c] = std::make_shared<NullT>((int)rows, stream);
    check("ysql_gather", ysql_gather(eb, src_batch.columnData[c],
                                     reinterpret_cast<const uint8_t *>(src_batch.nullset[c]->bitset), map, rows,
                                     data[c], reinterpret_cast<uint8_t *>(nullset[c]->bitset), stream));
  }
  return std::unique_ptr<BatchT>(new BatchT(data, DeviceT::GPU, rows, src_batch.columns, nullset, stream));
}
// Output row i = (tableIdxs[i] ? A : B)[newIdxs[i]] for i < srcSz (the
// merge's gather; isAsc = false writes the rows in reverse).
template <typename BatchT>
std::unique_ptr<BatchT> move_rows_batch_merge(BatchT &Abatch, BatchT &Bbatch, uint32_t *newIdxs, bool *tableIdxs,
                                              const uint32_t srcSz, hipStream_t stream, bool isAsc = true) {
  using NullT = typename std::decay<decltype(*Abatch.nullset[0])>::type;
  using DeviceT = decltype(Abatch.location);
  const size_t ncols = Abatch.columns.size();
  std::vector<void *> data(ncols);
  std::vector<std::shared_ptr<NullT>> nullset(ncols);
  for (size_t c = 0; c < ncols; c++) {
    const uint32_t eb = Abatch.columns[c]->bytes;
    if (hipMallocAsync(&data[c], (size_t)srcSz * eb, stream) != hipSuccess)
      throw KernelError("hipMallocAsync", YSQL_EHIP);
    nullset[c] = std::make_shared<NullT>((int)srcSz, stream);
    check("ysql_gather2",
          ysql_gather2(eb, Abatch.columnData[c], reinterpret_cast<const uint8_t *>(Abatch.nullset[c]->bitset),
                       Bbatch.columnData[c], reinterpret_cast<const uint8_t *>(Bbatch.nullset[c]->bitset), newIdxs,
                       reinterpret_cast<const uint8_t *>(tableIdxs), srcSz, isAsc ? 0 : 1, data[c],
                       reinterpret_cast<uint8_t *>(nullset[c]->bitset), stream));
  }
  return std::unique_ptr<BatchT>(new BatchT(data, DeviceT::GPU, srcSz, Abatch.columns, nullset, stream));
}

}  // namespace YallaSQL::Kernel
