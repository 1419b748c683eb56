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

// Operand codes of the device program besides register numbers (the host
// check rewrites ABI programs into these forms, see xprog_prepare):
//   kXAcc  as an operand: the previous instruction's result, held in static
//          registers (no indexed move); as a destination: keep the result
//          only there.  As a NULL source (nsrc): the previous result's flags.
//   kXImm  as operand b: the instruction's immediate, of the operand dtype.
constexpr int kXAcc = 0xFE;
constexpr int kXImm = 0xFD;
constexpr int kXNone = 0xFF;  // nsrc: no NULL flags

// An instruction as the kernels read it: four aligned 32-bit words, so a
// wave fetches it with one scalar load (the C ABI's packed 8- and 16-bit
// fields would need byte loads, which go through the vector memory path
// and wait for every column load in flight).
struct alignas(16) XInsnDev {
  uint32_t code;  // op | type << 8 | sub << 16
  uint32_t regs;  // dst | a << 8 | b << 16 | nsrc << 24 (register numbers or kXAcc / kXImm / kXNone)
  uint32_t imm_lo, imm_hi;
  __host__ __device__ int op() const { return (int)(code & 0xff); }
  __host__ __device__ int type() const { return (int)((code >> 8) & 0xff); }
  __host__ __device__ int sub() const { return (int)(code >> 16); }
  __host__ __device__ int dst() const { return (int)(regs & 0xff); }
  __host__ __device__ int a() const { return (int)((regs >> 8) & 0xff); }
  __host__ __device__ int b() const { return (int)((regs >> 16) & 0xff); }
  __host__ __device__ int nsrc() const { return (int)(regs >> 24); }
  __host__ __device__ int64_t imm() const { return (int64_t)(((uint64_t)imm_hi << 32) | imm_lo); }
};

struct XProg {
  XInsnDev insn[YSQL_XMAX_INSN];
  ysql_xinput in[YSQL_XMAX_INPUTS];
  ysql_xoutput out[YSQL_MAX_COLS];
  int ninsn, nin, nout, pred;  // pred: a register, or kXAcc (the last result)
  int aligned;  // every input column 16-byte aligned and every NULL column 4-byte aligned
};

// Host: checks a program over n rows (register bounds, operand dtypes,
// defined before use, input pointers) and copies it with its inputs into
// *out.  *wide: a 64-bit dtype occurs (the kernels then use 64-bit slots).
// reg_types (may be NULL) receives each register's dtype after the program
// (-1: undefined).
// live_mask: registers read after the program (outputs); pred: the
// predicate register or -1.  The device program it writes folds constants
// into immediate operands and passes single-use results through the
// accumulator (out->pred becomes kXAcc when the predicate is the last result).
int xprog_prepare(const ysql_xinsn *prog, int ninsn, const ysql_xinput *inputs, int nin, uint64_t n,
                  uint32_t live_mask, int pred, XProg *out, bool *wide, int8_t *reg_types);

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
__device__ __forceinline__ void xrun(const XProg &p, XRegs<S, R> &reg, S (&acc)[R], uint32_t &nacc) {
  for (int pc = 0; pc < p.ninsn; pc++) {
    const XInsnDev I = p.insn[pc];
    const int ia = I.a(), ib = I.b(), id = I.dst();
    S a[R], b[R], r[R];
    // operands: the accumulator, the immediate or a register of the file
    auto fetch = [&](int i, S(&o)[R]) {
      if (i == kXAcc) {
#pragma unroll
        for (int e = 0; e < R; e++) o[e] = acc[e];
      } else if (i == kXImm) {
        const S c = xconst<S>(I.type(), I.imm());
#pragma unroll
        for (int e = 0; e < R; e++) o[e] = c;
      } else {
        xget(reg, i, o);
      }
    };
    switch (I.op()) {
      case YSQL_XOP_CONST: {
        const S c = xconst<S>(I.type(), I.imm());
#pragma unroll
        for (int e = 0; e < R; e++) r[e] = c;
        break;
      }
      case YSQL_XOP_CMP:
        fetch(ia, a);
        fetch(ib, b);
        if (I.type() == YSQL_FLOAT) xcmp_t<float>(I.sub(), a, b, r);
        else if (sizeof(S) == 8 && I.type() == YSQL_DATETIME) xcmp_t<int64_t>(I.sub(), a, b, r);
        else xcmp_t<int32_t>(I.sub(), a, b, r);
        break;
      case YSQL_XOP_LOGIC:
        fetch(ia, a);
        fetch(ib, b);
        if (I.sub() == YSQL_AND) {
#pragma unroll
          for (int e = 0; e < R; e++) r[e] = a[e] & b[e];
        } else {
#pragma unroll
          for (int e = 0; e < R; e++) r[e] = a[e] | b[e];
        }
        break;
      case YSQL_XOP_ARITH:
        fetch(ia, a);
        fetch(ib, b);
        if (I.type() == YSQL_FLOAT) xarith_t<float>(I.sub(), a, b, r);
        else if (sizeof(S) == 8 && I.type() == YSQL_DATETIME) xarith_t<int64_t>(I.sub(), a, b, r);
        else xarith_t<int32_t>(I.sub(), a, b, r);
        break;
      case YSQL_XOP_CAST:
        fetch(ia, a);
        xcast(I.type(), I.sub(), a, r);
        break;
      default:  // YSQL_XOP_NOT
        fetch(ia, a);
#pragma unroll
        for (int e = 0; e < R; e++) r[e] = a[e] ^ (S)1;
        break;
    }
    if constexpr (NUL) {
      const int ns = I.nsrc();
      const uint32_t f = ns == kXNone ? 0u : ns == kXAcc ? nacc : (uint32_t)reg.nv[ns];
      nacc = f;
      if (id != kXAcc) reg.nv[id] = f;
    }
#pragma unroll
    for (int e = 0; e < R; e++) acc[e] = r[e];
    if (id != kXAcc) xset(reg, id, r);
  }
}

// Loads NB bytes (4, 8 or a multiple of 16, aligned to min(NB, 16)) as words.
template <int NB>
__device__ __forceinline__ void xload_words(const void *p, uint32_t (&w)[NB / 4]) {
  if constexpr (NB % 16 == 0) {
#pragma unroll
    for (int q = 0; q < NB / 16; q++) {
      const uint4 v = ((const uint4 *)p)[q];
      w[4 * q] = v.x;
      w[4 * q + 1] = v.y;
      w[4 * q + 2] = v.z;
      w[4 * q + 3] = v.w;
    }
  } else if constexpr (NB == 8) {
    const uint2 v = *(const uint2 *)p;
    w[0] = v.x;
    w[1] = v.y;
  } else {
    w[0] = *(const uint32_t *)p;
  }
}

// Loads R rows starting at `row0` of every input column c into register c
// (and its NULL flags when NUL).  `fast` (wave-uniform: every row of the
// group exists and every column is aligned): unguarded vector loads.  All
// loads are issued before any is used, so the columns' loads are in flight
// together; only then are they moved into the register file.
template <typename S, int R, bool NUL>
__device__ __forceinline__ void xload_inputs(const XProg &p, uint64_t row0, uint64_t n, bool fast,
                                             XRegs<S, R> &reg) {
  constexpr int WW = R * ((int)sizeof(S) / 4);  // words of a column's R values at the widest dtype
  if (fast) {
    uint32_t raw[YSQL_XMAX_INPUTS][WW];
    uint32_t nraw[YSQL_XMAX_INPUTS];
#pragma unroll
    for (int c = 0; c < YSQL_XMAX_INPUTS; c++) {
      if (c < p.nin) {
        const ysql_xinput in = p.in[c];
        if (sizeof(S) == 8 && in.dtype == YSQL_DATETIME) {
          xload_words<R * 8>((const uint64_t *)in.data + row0, reinterpret_cast<uint32_t(&)[R * 2]>(raw[c]));
        } else {
          xload_words<R * 4>((const uint32_t *)in.data + row0, reinterpret_cast<uint32_t(&)[R]>(raw[c]));
        }
        if constexpr (NUL) {
          if (in.nulls) {
            if constexpr (R == 4) nraw[c] = *(const uint32_t *)(in.nulls + row0);
            else if constexpr (R == 2) nraw[c] = *(const uint16_t *)(in.nulls + row0);
            else nraw[c] = in.nulls[row0];
          } else {
            nraw[c] = 0;
          }
        }
      }
    }
#pragma unroll
    for (int c = 0; c < YSQL_XMAX_INPUTS; c++) {
      if (c < p.nin) {
        const bool wide = sizeof(S) == 8 && p.in[c].dtype == YSQL_DATETIME;
#pragma unroll
        for (int e = 0; e < R; e++) {
          if (wide) reg.set(c, e, (S)(((uint64_t)raw[c][(2 * e + 1) % WW] << 32) | raw[c][2 * e % WW]));
          else reg.set(c, e, (S)raw[c][e]);
        }
        if constexpr (NUL) {
          uint32_t f = 0;
#pragma unroll
          for (int e = 0; e < R; e++) f |= (uint32_t)(((nraw[c] >> (8 * e)) & 0xff) != 0) << e;
          reg.nv[c] = f;
        }
      }
    }
    return;
  }
#pragma unroll
  for (int c = 0; c < YSQL_XMAX_INPUTS; c++) {
    if (c < p.nin) {
      const ysql_xinput in = p.in[c];
      if (sizeof(S) == 8 && in.dtype == YSQL_DATETIME) {
        const uint64_t *src = (const uint64_t *)in.data + row0;
#pragma unroll
        for (int e = 0; e < R; e++) reg.set(c, e, (row0 + e < n) ? (S)src[e] : (S)0);
      } else {
        const uint32_t *src = (const uint32_t *)in.data + row0;
#pragma unroll
        for (int e = 0; e < R; e++) reg.set(c, e, (row0 + e < n) ? (S)src[e] : (S)0);
      }
      if constexpr (NUL) {
        uint32_t f = 0;
        if (in.nulls) {
#pragma unroll
          for (int e = 0; e < R; e++) {
            const uint8_t v = (row0 + e < n) ? in.nulls[row0 + e] : (uint8_t)0;
            f |= (uint32_t)(v != 0) << e;
          }
        }
        reg.nv[c] = f;
      }
    }
  }
}

// True when every input column is 16-byte aligned and every NULL column
// 4-byte aligned, so a group of rows that starts at a multiple of R rows may
// use xload_inputs' fast path.
__host__ __device__ inline bool xinputs_aligned(const XProg &p) {
  for (int c = 0; c < p.nin; c++) {
    if (((uintptr_t)p.in[c].data & 15) != 0) return false;
    if (p.in[c].nulls && ((uintptr_t)p.in[c].nulls & 3) != 0) return false;
  }
  return true;
}

}  // namespace ysql
