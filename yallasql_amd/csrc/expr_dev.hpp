// expr_dev.hpp — device interpreter of fused expression programs.
//
// A predicate or projection tree is compiled on the host into a short
// register program (ysql_xinsn, include/ysql_kernels.h).  Each thread runs
// the program over R rows at a time.  The program lives in the kernel
// arguments, so every instruction field is wave-uniform: the dispatch on
// opcode, dtype and register number compiles to scalar branches, and the
// per-row work is plain vector arithmetic on registers.  A register is an
// array S[R] (S = uint32_t when no 64-bit dtype occurs in the program, else
// uint64_t) plus a word of NULL flags (bit e = row e).  Register numbers are
// resolved by switch statements, so the register file stays in VGPRs.
//
// Value encoding in a slot: INT = int32 bits, FLOAT = float bits, DATETIME =
// int64, BOOL = 0 or 1.  The scalar rules (comparison, wrap/zero-divide
// arithmetic, saturating casts) are those of the per-node kernels, so a
// fused tree gives the same bits as the nodes run one by one.
#pragma once

#include "ops_dev.hpp"

namespace ysql {

constexpr int kXRegs = YSQL_XMAX_REGS;

struct XProg {
  ysql_xinsn insn[YSQL_XMAX_INSN];
  ysql_xinput in[YSQL_XMAX_INPUTS];
  ysql_xoutput out[YSQL_MAX_COLS];
  int ninsn, nin, nout, pred;
};

// Host: checks a program over n rows (register bounds, operand dtypes,
// defined before use, input pointers) and copies it with its inputs into
// *out.  *wide: a 64-bit dtype occurs (the kernels then use 64-bit slots).
// reg_types (may be NULL) receives each register's dtype after the program
// (-1: undefined).
int xprog_prepare(const ysql_xinsn *prog, int ninsn, const ysql_xinput *inputs, int nin, uint64_t n, XProg *out,
                  bool *wide, int8_t *reg_types);

// ---- slot encoding -----------------------------------------------------
template <typename S>
__device__ __forceinline__ int32_t xs_i32(S v) { return (int32_t)(uint32_t)v; }
template <typename S>
__device__ __forceinline__ float xs_f32(S v) { return __uint_as_float((uint32_t)v); }
template <typename S>
__device__ __forceinline__ int64_t xs_i64(S v) { return (int64_t)(uint64_t)v; }
template <typename S>
__device__ __forceinline__ S xs_from_i32(int32_t x) { return (S)(uint32_t)x; }
template <typename S>
__device__ __forceinline__ S xs_from_f32(float x) { return (S)__float_as_uint(x); }
template <typename S>
__device__ __forceinline__ S xs_from_i64(int64_t x) { return (S)(uint64_t)x; }

template <typename T, typename S>
__device__ __forceinline__ T xs_get(S v) {
  if constexpr (std::is_same<T, int32_t>::value) return xs_i32(v);
  else if constexpr (std::is_same<T, float>::value) return xs_f32(v);
  else return xs_i64(v);
}
template <typename T, typename S>
__device__ __forceinline__ S xs_put(T x) {
  if constexpr (std::is_same<T, int32_t>::value) return xs_from_i32<S>(x);
  else if constexpr (std::is_same<T, float>::value) return xs_from_f32<S>(x);
  else return xs_from_i64<S>(x);
}

// ---- the register file ----------------------------------------------------
// Per row e, the 32-bit halves of all registers form one 16-lane vector, so a
// register number that is wave-uniform indexes VGPRs directly (M0-relative
// moves) instead of spilling the file to private memory as an array would.
typedef uint32_t XVec __attribute__((ext_vector_type(16)));
static_assert(YSQL_XMAX_REGS == 16, "register file is one 16-lane vector per row");

template <typename S, int R>
struct XRegs {
  XVec lo[R];
  XVec hi[sizeof(S) == 8 ? R : 1];
  XVec nv;  // lane i = NULL flags of register i (bit e = row e)
  __device__ __forceinline__ S get(int i, int e) const {
    if constexpr (sizeof(S) == 8) return ((uint64_t)hi[e][i] << 32) | lo[e][i];
    else return lo[e][i];
  }
  __device__ __forceinline__ void set(int i, int e, S v) {
    lo[e][i] = (uint32_t)v;
    if constexpr (sizeof(S) == 8) hi[e][i] = (uint32_t)((uint64_t)v >> 32);
  }
  __device__ __forceinline__ void clear() {
#pragma unroll
    for (int e = 0; e < R; e++) {
      lo[e] = (XVec)0u;
      if constexpr (sizeof(S) == 8) hi[e] = (XVec)0u;
    }
    nv = (XVec)0u;
  }
};

template <typename S, int R>
__device__ __forceinline__ void xget(const XRegs<S, R> &rf, int i, S (&o)[R]) {
#pragma unroll
  for (int e = 0; e < R; e++) o[e] = rf.get(i, e);
}
template <typename S, int R>
__device__ __forceinline__ void xset(XRegs<S, R> &rf, int i, const S (&v)[R]) {
#pragma unroll
  for (int e = 0; e < R; e++) rf.set(i, e, v[e]);
}

// ---- operations over R rows ---------------------------------------------
template <typename T, typename S, int R>
__device__ __forceinline__ void xcmp_t(int sub, const S (&a)[R], const S (&b)[R], S (&r)[R]) {
  switch (sub) {
#define YSQL_XCMP(C)                                                                    \
  case C:                                                                               \
    _Pragma("unroll") for (int e = 0; e < R; e++) r[e] =                                \
        (S)cmp_static<C>(xs_get<T>(a[e]), xs_get<T>(b[e]));                             \
    break;
    YSQL_XCMP(YSQL_LT)
    YSQL_XCMP(YSQL_LE)
    YSQL_XCMP(YSQL_GT)
    YSQL_XCMP(YSQL_GE)
    YSQL_XCMP(YSQL_EQ)
    YSQL_XCMP(YSQL_NE)
#undef YSQL_XCMP
  }
}

template <typename T, typename S, int R>
__device__ __forceinline__ void xarith_t(int sub, const S (&a)[R], const S (&b)[R], S (&r)[R]) {
  switch (sub) {
#define YSQL_XAR(OP)                                                                     \
  case OP:                                                                               \
    _Pragma("unroll") for (int e = 0; e < R; e++) r[e] =                                 \
        xs_put<T, S>(arith_op<T, OP>(xs_get<T>(a[e]), xs_get<T>(b[e])));                 \
    break;
    YSQL_XAR(YSQL_ADD)
    YSQL_XAR(YSQL_SUB)
    YSQL_XAR(YSQL_MUL)
    YSQL_XAR(YSQL_DIV)
#undef YSQL_XAR
    case YSQL_REM:
      if constexpr (std::numeric_limits<T>::is_integer) {
#pragma unroll
        for (int e = 0; e < R; e++) r[e] = xs_put<T, S>(arith_op<T, YSQL_REM>(xs_get<T>(a[e]), xs_get<T>(b[e])));
      }
      break;
  }
}

// ysql_cast's table: INT/FLOAT/DATETIME to each other, BOOL to INT.
template <typename S, int R>
__device__ __forceinline__ void xcast(int from, int to, const S (&a)[R], S (&r)[R]) {
#pragma unroll
  for (int e = 0; e < R; e++) r[e] = a[e];
  if (from == to) return;
  if (from == YSQL_BOOL) {  // to INT
#pragma unroll
    for (int e = 0; e < R; e++) r[e] = xs_from_i32<S>(a[e] != 0);
    return;
  }
  if (from == YSQL_INT && to == YSQL_FLOAT) {
#pragma unroll
    for (int e = 0; e < R; e++) r[e] = xs_from_f32<S>((float)xs_i32(a[e]));
  } else if (from == YSQL_FLOAT && to == YSQL_INT) {
#pragma unroll
    for (int e = 0; e < R; e++) r[e] = xs_from_i32<S>(sat_from_float<int32_t>(xs_f32(a[e])));
  }
  if constexpr (sizeof(S) == 8) {
    if (from == YSQL_INT && to == YSQL_DATETIME) {
#pragma unroll
      for (int e = 0; e < R; e++) r[e] = xs_from_i64<S>((int64_t)xs_i32(a[e]));
    } else if (from == YSQL_FLOAT && to == YSQL_DATETIME) {
#pragma unroll
      for (int e = 0; e < R; e++) r[e] = xs_from_i64<S>(sat_from_float<int64_t>(xs_f32(a[e])));
    } else if (from == YSQL_DATETIME && to == YSQL_INT) {
#pragma unroll
      for (int e = 0; e < R; e++) r[e] = xs_from_i32<S>((int32_t)xs_i64(a[e]));
    } else if (from == YSQL_DATETIME && to == YSQL_FLOAT) {
#pragma unroll
      for (int e = 0; e < R; e++) r[e] = xs_from_f32<S>((float)xs_i64(a[e]));
    }
  }
}

template <typename S>
__device__ __forceinline__ S xconst(int type, int64_t imm) {
  switch (type) {
    case YSQL_FLOAT: return xs_from_f32<S>(__uint_as_float((uint32_t)(uint64_t)imm));
    case YSQL_DATETIME: return xs_from_i64<S>(imm);
    case YSQL_BOOL: return (S)(imm != 0);
    default: return xs_from_i32<S>((int32_t)imm);
  }
}

// Runs every instruction of `p` over the R rows held in `reg`.  NUL: track
// the NULL flags (projections); a filter's selection ignores them.
template <typename S, int R, bool NUL>
__device__ __forceinline__ void xrun(const XProg &p, XRegs<S, R> &reg) {
  for (int pc = 0; pc < p.ninsn; pc++) {
    const ysql_xinsn I = p.insn[pc];
    S a[R], b[R], r[R];
    switch (I.op) {
      case YSQL_XOP_CONST: {
        const S c = xconst<S>(I.type, I.imm);
#pragma unroll
        for (int e = 0; e < R; e++) r[e] = c;
        break;
      }
      case YSQL_XOP_CMP:
        xget(reg, I.a, a);
        xget(reg, I.b, b);
        if (I.type == YSQL_FLOAT) xcmp_t<float>(I.sub, a, b, r);
        else if (sizeof(S) == 8 && I.type == YSQL_DATETIME) xcmp_t<int64_t>(I.sub, a, b, r);
        else xcmp_t<int32_t>(I.sub, a, b, r);
        break;
      case YSQL_XOP_LOGIC:
        xget(reg, I.a, a);
        xget(reg, I.b, b);
        if (I.sub == YSQL_AND) {
#pragma unroll
          for (int e = 0; e < R; e++) r[e] = a[e] & b[e];
        } else {
#pragma unroll
          for (int e = 0; e < R; e++) r[e] = a[e] | b[e];
        }
        break;
      case YSQL_XOP_ARITH:
        xget(reg, I.a, a);
        xget(reg, I.b, b);
        if (I.type == YSQL_FLOAT) xarith_t<float>(I.sub, a, b, r);
        else if (sizeof(S) == 8 && I.type == YSQL_DATETIME) xarith_t<int64_t>(I.sub, a, b, r);
        else xarith_t<int32_t>(I.sub, a, b, r);
        break;
      case YSQL_XOP_CAST:
        xget(reg, I.a, a);
        xcast(I.type, I.sub, a, r);
        break;
      default:  // YSQL_XOP_NOT
        xget(reg, I.a, a);
#pragma unroll
        for (int e = 0; e < R; e++) r[e] = a[e] ^ (S)1;
        break;
    }
    xset(reg, I.dst, r);
    if constexpr (NUL) reg.nv[I.dst] = I.nsrc >= 0 ? (uint32_t)reg.nv[I.nsrc] : 0u;
  }
}

// Loads R rows starting at `row0` of input column c into register c (and
// its NULL flags when NUL).  `full`: all R rows exist.
template <typename S, int R, bool NUL>
__device__ __forceinline__ void xload_inputs(const XProg &p, uint64_t row0, uint64_t n, XRegs<S, R> &reg) {
  const bool full = row0 + R <= n;
#pragma unroll
  for (int c = 0; c < YSQL_XMAX_INPUTS; c++) {
    if (c < p.nin) {
      const ysql_xinput in = p.in[c];
      if (sizeof(S) == 8 && in.dtype == YSQL_DATETIME) {
        const uint64_t *src = (const uint64_t *)in.data + row0;
        if (full && (((uintptr_t)src & 15) == 0) && (R % 2) == 0) {
#pragma unroll
          for (int q = 0; q < R / 2; q++) {
            const uint4 v = ((const uint4 *)src)[q];
            reg.set(c, 2 * q, (S)(((uint64_t)v.y << 32) | v.x));
            reg.set(c, 2 * q + 1, (S)(((uint64_t)v.w << 32) | v.z));
          }
        } else {
#pragma unroll
          for (int e = 0; e < R; e++) reg.set(c, e, (row0 + e < n) ? (S)src[e] : (S)0);
        }
      } else {
        const uint32_t *src = (const uint32_t *)in.data + row0;
        if (full && (((uintptr_t)src & 15) == 0) && (R % 4) == 0) {
#pragma unroll
          for (int q = 0; q < R / 4; q++) {
            const uint4 v = ((const uint4 *)src)[q];
            reg.set(c, 4 * q, (S)v.x);
            reg.set(c, 4 * q + 1, (S)v.y);
            reg.set(c, 4 * q + 2, (S)v.z);
            reg.set(c, 4 * q + 3, (S)v.w);
          }
        } else {
#pragma unroll
          for (int e = 0; e < R; e++) reg.set(c, e, (row0 + e < n) ? (S)src[e] : (S)0);
        }
      }
      if constexpr (NUL) {
        uint32_t f = 0;
        if (in.nulls) {
#pragma unroll
          for (int e = 0; e < R; e++) f |= (uint32_t)((row0 + e < n) && in.nulls[row0 + e] != 0) << e;
        }
        reg.nv[c] = f;
      }
    }
  }
}

}  // namespace ysql
