// ops_dev.hpp — the scalar rules of projection arithmetic and casts, shared
// by the per-node kernels (elementwise.hip) and the fused expression
// programs (expr.hip), so both give the same bits.
#pragma once

#include <math.h>

#include <limits>
#include <type_traits>

#include "common.hpp"

namespace ysql {

// Arithmetic (binary_operators_kernel.hpp:27-51).  Integer add/sub/mul
// wrap; DIV/REM by 0 give 0; INT_MIN / -1 wraps and INT_MIN % -1 is 0.
template <typename T>
struct Unsigned { typedef T type; };
template <> struct Unsigned<int32_t> { typedef uint32_t type; };
template <> struct Unsigned<int64_t> { typedef uint64_t type; };

template <typename T, int OP>
__device__ __forceinline__ T arith_op(T a, T b) {
  typedef typename Unsigned<T>::type U;
  if constexpr (OP == YSQL_ADD) return (T)((U)a + (U)b);
  else if constexpr (OP == YSQL_SUB) return (T)((U)a - (U)b);
  else if constexpr (OP == YSQL_MUL) return (T)((U)a * (U)b);
  else if constexpr (OP == YSQL_DIV) {
    if (b == 0) return 0;
    if constexpr (std::numeric_limits<T>::is_integer) {
      if (b == (T)-1) return (T)((U)0 - (U)a);
    }
    return a / b;
  } else {
    if (b == 0) return 0;
    if (b == (T)-1) return 0;
    return a % b;
  }
}

// Casts (cast_operators_kernel.cu:5-17).  float -> integer saturates and
// maps NaN to 0 (what v_cvt_i32_f32 does; made explicit for int64).
template <typename D>
__device__ __forceinline__ D sat_from_float(float v) {
  if (isnan(v)) return 0;
  const float lo = (float)std::numeric_limits<D>::min();
  const float hi = (float)std::numeric_limits<D>::max();
  if (v <= lo) return std::numeric_limits<D>::min();
  if (v >= hi) return std::numeric_limits<D>::max();
  return (D)v;
}

}  // namespace ysql
