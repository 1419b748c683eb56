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
// Ownership differs in one place on purpose: launch_radix_sort still returns
// a device index buffer by reference, but it is allocated with
// hipMallocAsync on the given stream and belongs to the caller (free it with
// hipFreeAsync); the reference leaked it and all temporaries
// (radix_sort_kernel.cu:139-158).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

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
template <typename T> struct AscOp { static constexpr bool desc = false; };
template <typename T> struct DescOp { static constexpr bool desc = true; };
template <typename T> struct AscSortMergeOp { static constexpr bool desc = false; };
template <typename T> struct DescSortMergeOp { static constexpr bool desc = true; };

template <typename T> struct TypeCode;
template <> struct TypeCode<int> { static constexpr int v = YSQL_INT; };
template <> struct TypeCode<float> { static constexpr int v = YSQL_FLOAT; };
template <> struct TypeCode<int64_t> { static constexpr int v = YSQL_DATETIME; };
template <> struct TypeCode<bool> { static constexpr int v = YSQL_BOOL; };
template <> struct TypeCode<uint32_t> { static constexpr int v = YSQL_UKEY32; };
template <> struct TypeCode<uint64_t> { static constexpr int v = YSQL_UKEY64; };

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
template <typename OP, typename T, size_t BlockDim>
void launch_radix_sort(T *d_input, uint32_t *&d_idxs_in, const uint32_t N, hipStream_t stream) {
  size_t ws_bytes = 0;
  check("ysql_radix_sort_ws_bytes", ysql_radix_sort_ws_bytes(TypeCode<T>::v, N, &ws_bytes));
  void *ws = nullptr;
  if (hipMallocAsync(&d_idxs_in, (size_t)N * sizeof(uint32_t) + 4, stream) != hipSuccess ||
      hipMallocAsync(&ws, ws_bytes + 256, stream) != hipSuccess)
    throw KernelError("hipMallocAsync", YSQL_EHIP);
  int rc = ysql_radix_argsort(TypeCode<T>::v, OP::desc, d_input, d_idxs_in, N, 0, 8 * sizeof(T), ws, ws_bytes + 256,
                              stream);
  hipFreeAsync(ws, stream);
  check("ysql_radix_argsort", rc);
}

// ---- A4 / A13: prefix_sum_kernel.cu:140-178 --------------------------------
inline void launch_prefix_sum_mask(bool *mask, uint32_t *out, uint32_t sz, hipStream_t stream) {
  size_t b = 0;
  check("ysql_scan_ws_bytes", ysql_scan_ws_bytes(sz, &b));
  void *ws = nullptr;
  if (hipMallocAsync(&ws, b, stream) != hipSuccess) throw KernelError("hipMallocAsync", YSQL_EHIP);
  int rc = ysql_scan_mask_inclusive(reinterpret_cast<const uint8_t *>(mask), out, sz, ws, b, stream);
  hipFreeAsync(ws, stream);
  check("ysql_scan_mask_inclusive", rc);
}
template <typename T>
void launch_prefix_sum(T *in, T *out, uint32_t sz, hipStream_t stream) {
  static_assert(sizeof(T) == 4, "uint32 scan");
  size_t b = 0;
  check("ysql_scan_ws_bytes", ysql_scan_ws_bytes(sz, &b));
  void *ws = nullptr;
  if (hipMallocAsync(&ws, b, stream) != hipSuccess) throw KernelError("hipMallocAsync", YSQL_EHIP);
  int rc = ysql_scan_u32_inclusive(reinterpret_cast<const uint32_t *>(in), reinterpret_cast<uint32_t *>(out), sz, ws,
                                   b, stream);
  hipFreeAsync(ws, stream);
  check("ysql_scan_u32_inclusive", rc);
}

// ---- A8-A9: merge_batchs.cu:143-158 (lasti receives corank(k_mx)) --------
template <typename T, typename OP>
void launch_merge_sorted_array_kernel(T *A, T *B, T *C, uint32_t *new_idx, bool *table_idx, int *lasti,
                                      const uint32_t k_mx, uint32_t m, uint32_t n, hipStream_t stream) {
  size_t b = 0;
  check("ysql_merge_ws_bytes", ysql_merge_ws_bytes(m, n, k_mx, &b));
  void *ws = nullptr;
  int64_t *cons = nullptr;
  if (hipMallocAsync(&ws, b + 256, stream) != hipSuccess || hipMallocAsync((void **)&cons, 8, stream) != hipSuccess)
    throw KernelError("hipMallocAsync", YSQL_EHIP);
  int rc = ysql_merge_path(TypeCode<T>::v, OP::desc, A, m, B, n, k_mx, C, new_idx,
                           reinterpret_cast<uint8_t *>(table_idx), cons, ws, b + 256, stream);
  if (rc == YSQL_OK && lasti) {
    int64_t h = 0;
    hipMemcpyAsync(&h, cons, 8, hipMemcpyDeviceToHost, stream);
    hipStreamSynchronize(stream);
    int v = (int)h;
    hipMemcpyAsync(lasti, &v, sizeof(int), hipMemcpyHostToDevice, stream);
    hipStreamSynchronize(stream);
  }
  hipFreeAsync(ws, stream);
  hipFreeAsync(cons, stream);
  check("ysql_merge_path", rc);
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
  void *ws = nullptr;
  uint64_t *cnt = nullptr;
  if (hipMallocAsync(&ws, b + 256, stream) != hipSuccess || hipMallocAsync((void **)&cnt, 8, stream) != hipSuccess)
    throw KernelError("hipMallocAsync", YSQL_EHIP);
  int rc = ysql_join_pairs(TypeCode<T>::v, Op::code, d_lhs, left_rows, d_rhs, right_rows, pairs, cap, cnt, ws, b + 256,
                           stream);
  if (rc == YSQL_OK)  // low 32 bits of the little-endian count
    hipMemcpyAsync(actualSz, cnt, sizeof(int), hipMemcpyDeviceToDevice, stream);
  hipFreeAsync(ws, stream);
  hipFreeAsync(cnt, stream);
  check("ysql_join_pairs", rc);
}
inline void launch_and_join_operators(const uint32_t *l_pairs, const uint32_t *r_pairs, uint32_t *pairs,
                                      uint32_t * /*mask: not needed*/, int *actualSz, unsigned int left_rows,
                                      unsigned int right_rows, hipStream_t stream) {
  size_t b = 0;
  check("ysql_pairs_intersect_ws_bytes", ysql_pairs_intersect_ws_bytes(left_rows, right_rows, &b));
  void *ws = nullptr;
  uint64_t *cnt = nullptr;
  if (hipMallocAsync(&ws, b + 256, stream) != hipSuccess || hipMallocAsync((void **)&cnt, 8, stream) != hipSuccess)
    throw KernelError("hipMallocAsync", YSQL_EHIP);
  int rc = ysql_pairs_intersect(l_pairs, left_rows, r_pairs, right_rows, pairs, left_rows, cnt, ws, b + 256, stream);
  if (rc == YSQL_OK) hipMemcpyAsync(actualSz, cnt, sizeof(int), hipMemcpyDeviceToDevice, stream);
  hipFreeAsync(ws, stream);
  hipFreeAsync(cnt, stream);
  check("ysql_pairs_intersect", rc);
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

}  // namespace YallaSQL::Kernel
