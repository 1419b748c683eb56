// compat_launchers.cpp — calls every launcher of include/ysql_kernels.hpp
// the way the reference's operators call theirs, behind one extern "C"
// function per launcher family, so tests/test_gpu_compat.py can drive them
// with device buffers and compare the results with the oracle.  Each
// function returns 0, or the KernelError code a launcher threw.
//
// Batch, NullBitSet, Column, Device and DataType below are test doubles with
// the public shape of the reference's engine types (include/batch.hpp,
// include/null_bit_set.hpp, include/db/column.hpp, include/enums/*.hpp),
// which is all the batch-level helpers rely on.
#include <cstring>
#include <memory>
#include <vector>

#include "ysql_kernels.hpp"

using namespace YallaSQL::Kernel;

enum class Device : uint8_t { CPU, GPU, FS };
enum class DataType : uint8_t { INT = 0, FLOAT, DATETIME, STRING };

struct NullBitSet {
  char *bitset = nullptr;
  NullBitSet(int sz, hipStream_t stream) {
    if (hipMallocAsync((void **)&bitset, sz > 0 ? sz : 1, stream) != hipSuccess) bitset = nullptr;
  }
  explicit NullBitSet(char *borrowed) : bitset(borrowed), owned(false) {}
  ~NullBitSet() {
    if (owned && bitset) (void)hipFree(bitset);
  }
  bool owned = true;
};

// The reference's fixed-width string (include/kernels/string_kernel.hpp:6-8).
struct String {
  char data[256];
};

struct Column {
  DataType type;
  unsigned int bytes;
  Column(DataType t, unsigned int b) : type(t), bytes(b) {}
};

struct Batch {
  std::vector<void *> columnData;
  std::vector<std::shared_ptr<NullBitSet>> nullset;
  Device location;
  size_t batchSize;
  hipStream_t stream;
  std::vector<std::shared_ptr<Column>> columns;
  Batch(const std::vector<void *> &data, Device loc, uint32_t rows, std::vector<std::shared_ptr<Column>> &cols,
        std::vector<std::shared_ptr<NullBitSet>> &nulls, hipStream_t st)
      : columnData(data), nullset(nulls), location(loc), batchSize(rows), stream(st), columns(cols) {}
};

template <typename F>
static int guarded(F &&f) {
  try {
    f();
    return 0;
  } catch (const KernelError &e) {
    return e.code;
  } catch (...) {
    return -100;
  }
}

static hipStream_t S(void *s) { return reinterpret_cast<hipStream_t>(s); }

extern "C" {

// A1: which = 0 int->key, 1 float->key, 2 key->int, 3 key->float
int compat_sortkey(int which, void *in, void *out, int n, void *stream) {
  return guarded([&] {
    hipStream_t s = S(stream);
    if (which == 0) launch_int_to_uint32((int *)in, (uint32_t *)out, n, s);
    if (which == 1) launch_float_to_uint32((float *)in, (uint32_t *)out, n, s);
    if (which == 2) launch_uint32_to_int((uint32_t *)in, (int *)out, n, s);
    if (which == 3) launch_uint32_to_float((uint32_t *)in, (float *)out, n, s);
  });
}

// A2-A6: keys sorted in place; the launcher's index buffer copied to idx_out
int compat_radix_sort(int key64, int desc, void *keys, int n, uint32_t *idx_out, void *stream) {
  return guarded([&] {
    hipStream_t s = S(stream);
    uint32_t *idx = nullptr;
    if (key64) {
      if (desc) launch_radix_sort<DescOp<uint64_t>, uint64_t, 256>((uint64_t *)keys, idx, n, s);
      else launch_radix_sort<AscOp<uint64_t>, uint64_t, 256>((uint64_t *)keys, idx, n, s);
    } else {
      if (desc) launch_radix_sort<DescOp<uint32_t>, uint32_t, 256>((uint32_t *)keys, idx, n, s);
      else launch_radix_sort<AscOp<uint32_t>, uint32_t, 256>((uint32_t *)keys, idx, n, s);
    }
    if (n) (void)hipMemcpyAsync(idx_out, idx, (size_t)n * 4, hipMemcpyDeviceToDevice, s);
    (void)hipFreeAsync(idx, s);
  });
}

// STRING keys: the reference's 256-byte String sorted in place, and merged
int compat_sort_strings(int desc, void *strs, int n, uint32_t *idx_out, void *stream) {
  return guarded([&] {
    hipStream_t s = S(stream);
    uint32_t *idx = nullptr;
    if (desc) launch_radix_sort<DescOpString, String, 256>((String *)strs, idx, n, s);
    else launch_radix_sort<AscOpString, String, 256>((String *)strs, idx, n, s);
    if (n) (void)hipMemcpyAsync(idx_out, idx, (size_t)n * 4, hipMemcpyDeviceToDevice, s);
    (void)hipFreeAsync(idx, s);
  });
}
int compat_merge_strings(int desc, void *A, void *B, void *C, uint32_t *new_idx, bool *table_idx, int *lasti,
                         uint32_t k, uint32_t m, uint32_t n, void *stream) {
  return guarded([&] {
    hipStream_t s = S(stream);
    if (desc)
      launch_merge_sorted_array_kernel<String, DescSortStrOp>((String *)A, (String *)B, (String *)C, new_idx, table_idx,
                                                              lasti, k, m, n, s);
    else
      launch_merge_sorted_array_kernel<String, AscSortStrOp>((String *)A, (String *)B, (String *)C, new_idx, table_idx,
                                                             lasti, k, m, n, s);
  });
}

// A4 / A13
int compat_prefix_sum(int mask, void *in, uint32_t *out, int n, void *stream) {
  return guarded([&] {
    if (mask) launch_prefix_sum_mask((bool *)in, out, n, S(stream));
    else launch_prefix_sum<uint32_t>((uint32_t *)in, out, n, S(stream));
  });
}

// A8-A9: dtype 0 int, 1 float, 2 int64
int compat_merge(int dtype, int desc, void *A, void *B, void *C, uint32_t *new_idx, bool *table_idx, int *lasti,
                 uint32_t k, uint32_t m, uint32_t n, void *stream) {
  return guarded([&] {
    hipStream_t s = S(stream);
#define M(T)                                                                                                \
  if (desc) launch_merge_sorted_array_kernel<T, DescSortMergeOp<T>>((T *)A, (T *)B, (T *)C, new_idx, table_idx, lasti, k, m, n, s); \
  else launch_merge_sorted_array_kernel<T, AscSortMergeOp<T>>((T *)A, (T *)B, (T *)C, new_idx, table_idx, lasti, k, m, n, s);
    if (dtype == 0) { M(int) }
    else if (dtype == 1) { M(float) }
    else { M(int64_t) }
#undef M
  });
}

// A7 gather, A10 merge gather, A14 filter scatter, A20 join gather (4-byte columns)
int compat_move(int which, void *src, void *src2, void *res, char *isnull, char *isnull2, char *res_isnull,
                uint32_t *map, bool *flags, int is_asc_or_right, uint32_t n, void *stream) {
  return guarded([&] {
    hipStream_t s = S(stream);
    if (which == 0) launch_move_one_col<int>((int *)src, (int *)res, isnull, res_isnull, map, n, s);
    if (which == 1)
      launch_move_one_col_merge<int>((int *)src, (int *)src2, (int *)res, isnull, isnull2, res_isnull, map, flags,
                                     is_asc_or_right != 0, n, s);
    if (which == 2) launch_move_rows_filter_kernel<int>((int *)src, (int *)res, map, flags, isnull, n, s, res_isnull);
    if (which == 3) lanch_move_rows_join_kernel<int>((int *)src, (int *)res, map, isnull, res_isnull, n,
                                                     is_asc_or_right != 0, s);
  });
}

// A12: dtype 0 int, 1 float, 2 int64, 3 bool masks (op 0 AND, 1 OR)
int compat_conditional(int dtype, int op, void *rhs, void *lhs, int rhs_scalar, int lhs_scalar, bool *res, uint32_t n,
                       int isneg, void *stream) {
  return guarded([&] {
    hipStream_t s = S(stream);
    const OperandType tr = rhs_scalar ? OperandType::SCALAR : OperandType::VECTOR;
    const OperandType tl = lhs_scalar ? OperandType::SCALAR : OperandType::VECTOR;
    if (dtype == 3) {
      if (op == 0) launch_conditional_operators<bool, ANDOperator>((bool *)rhs, (bool *)lhs, tr, tl, res, n, s);
      else launch_conditional_operators<bool, OROperator>((bool *)rhs, (bool *)lhs, tr, tl, res, n, s);
      return;
    }
#define C6(T)                                                                                              \
  switch (op) {                                                                                            \
    case YSQL_LT: launch_conditional_operators<T, LEOperator<T>>((T *)rhs, (T *)lhs, tr, tl, res, n, s, isneg); break; \
    case YSQL_LE: launch_conditional_operators<T, LEQOperator<T>>((T *)rhs, (T *)lhs, tr, tl, res, n, s, isneg); break; \
    case YSQL_GT: launch_conditional_operators<T, GEOperator<T>>((T *)rhs, (T *)lhs, tr, tl, res, n, s, isneg); break; \
    case YSQL_GE: launch_conditional_operators<T, GEQOperator<T>>((T *)rhs, (T *)lhs, tr, tl, res, n, s, isneg); break; \
    case YSQL_EQ: launch_conditional_operators<T, EQOperator<T>>((T *)rhs, (T *)lhs, tr, tl, res, n, s, isneg); break; \
    default: launch_conditional_operators<T, NEQOperator<T>>((T *)rhs, (T *)lhs, tr, tl, res, n, s, isneg); break; \
  }
    if (dtype == 0) { C6(int) }
    else if (dtype == 1) { C6(float) }
    else { C6(int64_t) }
#undef C6
  });
}

// A15: arithmetic on int columns (op ADD..REM), then a cast int -> float
int compat_binary(int op, int *rhs, int *lhs, int rhs_scalar, int lhs_scalar, int *res, float *res_f, uint32_t n,
                  void *stream) {
  return guarded([&] {
    hipStream_t s = S(stream);
    const OperandType tr = rhs_scalar ? OperandType::SCALAR : OperandType::VECTOR;
    const OperandType tl = lhs_scalar ? OperandType::SCALAR : OperandType::VECTOR;
    switch (op) {
      case YSQL_ADD: launch_binary_operators<int, AddOperator<int>>(rhs, lhs, tr, tl, res, n, s); break;
      case YSQL_SUB: launch_binary_operators<int, MinusOperator<int>>(rhs, lhs, tr, tl, res, n, s); break;
      case YSQL_MUL: launch_binary_operators<int, MulOperator<int>>(rhs, lhs, tr, tl, res, n, s); break;
      case YSQL_DIV: launch_binary_operators<int, DivOperator<int>>(rhs, lhs, tr, tl, res, n, s); break;
      default: launch_binary_operators<int, RemOperator<int>>(rhs, lhs, tr, tl, res, n, s); break;
    }
    launch_numerical_cast<int, float>(res, res_f, n, s);
  });
}

// A18: int64 join, op EQ or LT; cap = pair buffer capacity
int compat_outer_join(int op, int64_t *rhs, int64_t *lhs, uint32_t *pairs, int *actual, uint32_t L, uint32_t R,
                      uint64_t cap, void *stream) {
  return guarded([&] {
    hipStream_t s = S(stream);
    if (op == YSQL_EQ) launch_outer_join_operators<int64_t, EQOperator<int64_t>>(rhs, lhs, pairs, actual, L, R, s, cap);
    else launch_outer_join_operators<int64_t, LEOperator<int64_t>>(rhs, lhs, pairs, actual, L, R, s, cap);
  });
}

// A19
int compat_and_join(const uint32_t *l_pairs, const uint32_t *r_pairs, uint32_t *pairs, int *actual, uint32_t np1,
                    uint32_t np2, void *stream) {
  return guarded([&] { launch_and_join_operators(l_pairs, r_pairs, pairs, nullptr, actual, np1, np2, S(stream)); });
}

// A21: which 0 = launch_cross_product_col<int>, 1 = _type(type), 2 = _one_type(type, offset)
int compat_cross(int which, int type, void *src, void *dist, char *src_null, char *dist_null, uint32_t offset, int lbs,
                 int rbs, int isleft, void *stream) {
  return guarded([&] {
    hipStream_t s = S(stream);
    if (which == 0) launch_cross_product_col<int>((const int *)src, (int *)dist, src_null, dist_null, lbs, rbs, s, isleft);
    if (which == 1) launch_cross_product_type(src, dist, src_null, dist_null, (DataType)type, lbs, rbs, s, isleft != 0);
    if (which == 2)
      launch_cross_product_one_type(src, dist, offset, src_null, dist_null, (DataType)type, lbs, rbs, s, isleft != 0);
  });
}

// A16-A17: dtype 0 int, 1 float, 2 int64; agg MAX/MIN/SUM into *res (the
// reference's accumulator type), COUNT into an int counter
int compat_reduce(int dtype, int agg, void *arr, void *res, char *isnull, uint32_t n, void *stream) {
  return guarded([&] {
    hipStream_t s = S(stream);
    if (agg == YSQL_AGG_COUNT) {
      launch_reduction_count_notnull(isnull, (int *)res, isnull, n, s, 0);
      return;
    }
#define R3(T, TS)                                                                                              \
  if (agg == YSQL_AGG_MAX) launch_reduction_operators<T, T, MaxOperator<T>>((T *)arr, (T *)res, isnull, n, s, T()); \
  else if (agg == YSQL_AGG_MIN) launch_reduction_operators<T, T, MinOperator<T>>((T *)arr, (T *)res, isnull, n, s, T()); \
  else launch_reduction_operators<T, TS, SumOperator<TS>>((T *)arr, (TS *)res, isnull, n, s, TS());
    if (dtype == 0) { R3(int, int) }
    else if (dtype == 1) { R3(float, double) }
    else { R3(int64_t, int64_t) }
#undef R3
  });
}

// getAggregate's finalisers: AVG of a (counter, sum) pair and double -> float
int compat_finalize(int dtype, const int *counter, const void *sum, float *avg, const double *dsum, float *fsum) {
  return guarded([&] {
    if (dtype == 0) launch_div_avg<int>(counter, (const int *)sum, avg);
    else if (dtype == 1) launch_div_avg<double>(counter, (const double *)sum, avg);
    else launch_div_avg<int64_t>(counter, (const int64_t *)sum, avg);
    launch_convert_double_to_float_kernel(dsum, fsum);
    (void)hipDeviceSynchronize();
  });
}

// A7 / A10 at batch level over a 3-column batch (INT, DATETIME, STRING);
// outputs copied into the caller's buffers, then the new batch is freed.
int compat_move_batch(int merge, void **a_cols, char **a_nulls, void **b_cols, char **b_nulls, uint32_t rows_a,
                      uint32_t *map, bool *table_idx, uint32_t out_rows, int is_asc, void **out_cols, char **out_nulls,
                      void *stream) {
  return guarded([&] {
    hipStream_t s = S(stream);
    std::vector<std::shared_ptr<Column>> cols = {std::make_shared<Column>(DataType::INT, 4),
                                                 std::make_shared<Column>(DataType::DATETIME, 8),
                                                 std::make_shared<Column>(DataType::STRING, 256)};
    auto make = [&](void **c, char **nl, uint32_t rows) {
      std::vector<void *> d(c, c + 3);
      std::vector<std::shared_ptr<NullBitSet>> ns;
      for (int i = 0; i < 3; i++) ns.push_back(std::make_shared<NullBitSet>(nl[i]));
      return std::unique_ptr<Batch>(new Batch(d, Device::GPU, rows, cols, ns, s));
    };
    auto A = make(a_cols, a_nulls, merge ? rows_a : out_rows);
    std::unique_ptr<Batch> out;
    if (merge) {
      auto B = make(b_cols, b_nulls, out_rows);
      out = move_rows_batch_merge(*A, *B, map, table_idx, out_rows, s, is_asc != 0);
    } else {
      out = move_rows_batch(*A, map);
    }
    if (out->batchSize != out_rows || out->location != Device::GPU || out->columns.size() != 3)
      throw KernelError("move_rows_batch shape", YSQL_EINVAL);
    for (int i = 0; i < 3; i++) {
      (void)hipMemcpyAsync(out_cols[i], out->columnData[i], (size_t)out_rows * cols[i]->bytes,
                           hipMemcpyDeviceToDevice, s);
      (void)hipMemcpyAsync(out_nulls[i], out->nullset[i]->bitset, out_rows, hipMemcpyDeviceToDevice, s);
    }
    (void)hipStreamSynchronize(s);
    for (int i = 0; i < 3; i++) (void)hipFree(out->columnData[i]);
  });
}

}  // extern "C"
