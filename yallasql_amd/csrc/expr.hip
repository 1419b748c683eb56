// expr.hip — fused projection programs (ysql_expr_eval) and the host-side
// check shared with the fused filter (ysql_expr_filter_compact, scan.hip).
//
// The reference evaluates a projection tree node by node: every
// BoundFuncExpression / CastExpression / ComparisonExpression allocates a
// column and launches one kernel (bound_func.hpp:45-167, cast.hpp:23-90,
// filters/comparison.hpp:51-215).  Here the whole tree is one pass: each
// thread loads R rows of every input column once, runs the program in
// registers (expr_dev.hpp) and writes each output column once.
#include <algorithm>
#include <vector>

#include "expr_dev.hpp"

namespace ysql {

static bool x_value_type(int t) { return t == YSQL_INT || t == YSQL_FLOAT || t == YSQL_DATETIME; }

namespace {

// A decoded instruction of the host passes; register fields hold register
// numbers or kXAcc / kXImm / kXNone.
struct HInsn {
  int op, type, sub, dst, a, b, nsrc;
  int64_t imm;
};

bool commutes(int op, int sub) {
  return (op == YSQL_XOP_ARITH && (sub == YSQL_ADD || sub == YSQL_MUL)) ||
         (op == YSQL_XOP_CMP) || op == YSQL_XOP_LOGIC;
}
int swapped_cmp(int c) {
  switch (c) {
    case YSQL_LT: return YSQL_GT;
    case YSQL_LE: return YSQL_GE;
    case YSQL_GT: return YSQL_LT;
    case YSQL_GE: return YSQL_LE;
  }
  return c;
}

// Uses of the value that instruction i defines (in its register), up to
// the next definition of that register: (instruction, field) pairs, field
// 0 = a, 1 = b, 2 = nsrc.  *live_out: the value is still there at the end.
void uses_of(const std::vector<HInsn> &v, size_t i, uint32_t live_mask, int pred,
             std::vector<std::pair<size_t, int>> *uses, bool *live_out) {
  const int r = v[i].dst;
  uses->clear();
  *live_out = false;
  for (size_t k = i + 1; k < v.size(); k++) {
    const HInsn &J = v[k];
    if (J.a == r && J.op != YSQL_XOP_CONST) uses->push_back({k, 0});
    if (J.b == r && (J.op == YSQL_XOP_CMP || J.op == YSQL_XOP_LOGIC || J.op == YSQL_XOP_ARITH)) uses->push_back({k, 1});
    if (J.nsrc == r) uses->push_back({k, 2});
    if (J.dst == r) return;  // redefined: later reads see the new value
  }
  *live_out = ((live_mask >> r) & 1u) || pred == r;
}

// Constants folded into immediate operands: a CONST whose every use is an
// operand of a comparison or arithmetic instruction (the left one only when
// the operation may swap its operands) becomes that instruction's immediate
// operand b; a NULL source naming it becomes "none" (constants are never
// NULL).  Single-use results pass through the accumulator: when the value
// instruction i defines is read only by instruction i + 1 (and is not needed
// after the program), i keeps it in the accumulator and i + 1 reads it there.
void peephole(std::vector<HInsn> &v, uint32_t live_mask, int *pred) {
  std::vector<std::pair<size_t, int>> uses;
  bool live = false;
  for (size_t i = 0; i < v.size(); i++) {
    if (v[i].op != YSQL_XOP_CONST) continue;
    uses_of(v, i, live_mask, *pred, &uses, &live);
    if (live) continue;
    bool ok = true;
    for (auto &u : uses) {
      const HInsn &J = v[u.first];
      if (u.second == 2) continue;
      if (J.op != YSQL_XOP_CMP && J.op != YSQL_XOP_ARITH) ok = false;
      else if (J.a == J.b) ok = false;                        // const op const
      else if (J.b == kXImm) ok = false;                       // b already an immediate
      else if (u.second == 0 && !commutes(J.op, J.sub)) ok = false;
    }
    if (!ok) continue;
    const int r = v[i].dst;
    for (auto &u : uses) {
      HInsn &J = v[u.first];
      if (u.second == 2) {
        J.nsrc = kXNone;
        continue;
      }
      if (u.second == 0) {  // the constant on the left: swap the operands
        std::swap(J.a, J.b);
        if (J.op == YSQL_XOP_CMP) J.sub = swapped_cmp(J.sub);
      }
      J.b = kXImm;
      J.imm = v[i].imm;
      (void)r;
    }
    v.erase(v.begin() + (long)i);
    i--;
  }
  for (size_t i = 0; i + 1 < v.size(); i++) {
    if (v[i].dst == kXAcc) continue;
    uses_of(v, i, live_mask, *pred, &uses, &live);
    if (live || uses.empty()) continue;
    bool next_only = true;
    for (auto &u : uses) next_only = next_only && u.first == i + 1;
    if (!next_only) continue;
    HInsn &J = v[i + 1];
    const int r = v[i].dst;
    if (J.a == r) J.a = kXAcc;
    if (J.b == r) J.b = kXAcc;
    if (J.nsrc == r) J.nsrc = kXAcc;
    v[i].dst = kXAcc;
  }
  // the predicate as the last result
  if (*pred >= 0 && !v.empty() && v.back().dst == *pred && !((live_mask >> *pred) & 1u)) {
    v.back().dst = kXAcc;
    *pred = kXAcc;
  }
}

}  // namespace

// Checks a program and copies it into the kernel argument block.  Registers
// 0..nin-1 are the inputs; every operand must be defined earlier with the
// dtype the instruction names.  *wide: a 64-bit dtype occurs.
int xprog_prepare(const ysql_xinsn *prog, int ninsn, const ysql_xinput *inputs, int nin, uint64_t n,
                  uint32_t live_mask, int pred, XProg *out, bool *wide, int8_t *reg_types) {
  if (ninsn < 0 || ninsn > YSQL_XMAX_INSN || (ninsn > 0 && !prog)) return YSQL_EINVAL;
  if (nin < 0 || nin > YSQL_XMAX_INPUTS || (nin > 0 && !inputs)) return YSQL_EINVAL;
  memset(out, 0, sizeof(*out));
  int8_t ty[kXRegs];
  for (int r = 0; r < kXRegs; r++) ty[r] = -1;
  bool w = false;
  for (int c = 0; c < nin; c++) {
    if (!x_value_type(inputs[c].dtype) || (n > 0 && !inputs[c].data)) return YSQL_EINVAL;
    ty[c] = (int8_t)inputs[c].dtype;
    w |= inputs[c].dtype == YSQL_DATETIME;
    out->in[c] = inputs[c];
  }
  auto defined = [&](int r) { return r >= 0 && r < kXRegs && ty[r] >= 0; };
  std::vector<HInsn> v;
  for (int i = 0; i < ninsn; i++) {
    const ysql_xinsn I = prog[i];
    if (I.dst < 0 || I.dst >= kXRegs) return YSQL_EINVAL;
    if (I.nsrc != -1 && !defined(I.nsrc)) return YSQL_EINVAL;
    int res;
    switch (I.op) {
      case YSQL_XOP_CONST:
        if (!x_value_type(I.type) && I.type != YSQL_BOOL) return YSQL_EINVAL;
        res = I.type;
        break;
      case YSQL_XOP_CMP:
        if (!x_value_type(I.type) || !defined(I.a) || !defined(I.b) || ty[I.a] != I.type || ty[I.b] != I.type)
          return YSQL_EINVAL;
        if (I.sub < YSQL_LT || I.sub > YSQL_NE) return YSQL_EINVAL;
        res = YSQL_BOOL;
        break;
      case YSQL_XOP_LOGIC:
        if (!defined(I.a) || !defined(I.b) || ty[I.a] != YSQL_BOOL || ty[I.b] != YSQL_BOOL) return YSQL_EINVAL;
        if (I.sub != YSQL_AND && I.sub != YSQL_OR) return YSQL_EINVAL;
        res = YSQL_BOOL;
        break;
      case YSQL_XOP_ARITH:
        if (!x_value_type(I.type) || !defined(I.a) || !defined(I.b) || ty[I.a] != I.type || ty[I.b] != I.type)
          return YSQL_EINVAL;
        if (I.sub < YSQL_ADD || I.sub > YSQL_REM) return YSQL_EINVAL;
        if (I.sub == YSQL_REM && I.type == YSQL_FLOAT) return YSQL_EUNSUPPORTED;
        res = I.type;
        break;
      case YSQL_XOP_CAST:
        if (!defined(I.a) || ty[I.a] != I.type) return YSQL_EINVAL;
        if (I.type == YSQL_BOOL) {
          if (I.sub != YSQL_INT) return YSQL_EINVAL;
        } else if (!x_value_type(I.type) || !x_value_type(I.sub)) {
          return YSQL_EINVAL;
        }
        res = I.sub;
        break;
      case YSQL_XOP_NOT:
        if (!defined(I.a) || ty[I.a] != YSQL_BOOL) return YSQL_EINVAL;
        res = YSQL_BOOL;
        break;
      default:
        return YSQL_EINVAL;
    }
    w |= I.type == YSQL_DATETIME || res == YSQL_DATETIME;
    ty[I.dst] = (int8_t)res;
    // fields an operation does not read are cleared, so the passes see only real uses
    const bool binary = I.op == YSQL_XOP_CMP || I.op == YSQL_XOP_LOGIC || I.op == YSQL_XOP_ARITH;
    v.push_back(HInsn{I.op, I.type, I.sub, I.dst, I.op == YSQL_XOP_CONST ? kXNone : I.a, binary ? I.b : kXNone,
                      I.nsrc < 0 ? kXNone : I.nsrc, I.imm});
  }
  if (pred >= 0 && (pred >= kXRegs || ty[pred] != YSQL_BOOL)) return YSQL_EINVAL;
  peephole(v, live_mask, &pred);
  for (size_t i = 0; i < v.size(); i++) {
    const HInsn &I = v[i];
    XInsnDev &d = out->insn[i];
    d.code = (uint32_t)(I.op & 0xff) | (uint32_t)(I.type & 0xff) << 8 | (uint32_t)(uint16_t)I.sub << 16;
    d.regs = (uint32_t)(I.dst & 0xff) | (uint32_t)(I.a & 0xff) << 8 | (uint32_t)(I.b & 0xff) << 16 |
             (uint32_t)(I.nsrc & 0xff) << 24;
    d.imm_lo = (uint32_t)(uint64_t)I.imm;
    d.imm_hi = (uint32_t)((uint64_t)I.imm >> 32);
  }
  out->ninsn = (int)v.size();
  out->nin = nin;
  out->pred = pred;
  out->aligned = xinputs_aligned(*out);
  *wide = w;
  if (reg_types) memcpy(reg_types, ty, sizeof(ty));
  return YSQL_OK;
}

constexpr int kXBlock = 256;
// Rows per thread per pass of the program: 4 with 32-bit slots, 1 with
// 64-bit ones (the register file is then twice as wide; 2 rows measured 193
// VGPRs and spills, 1 row 76).
template <typename S>
constexpr int xrows() { return sizeof(S) == 8 ? 1 : 4; }

template <typename S, int R>
__device__ __forceinline__ void xstore_outputs(const XProg &p, uint64_t row0, uint64_t n, const XRegs<S, R> &reg) {
  const bool full = row0 + R <= n;
#pragma unroll
  for (int o = 0; o < YSQL_MAX_COLS; o++) {
    if (o < p.nout) {
      const ysql_xoutput out = p.out[o];
      S v[R];
      xget(reg, out.reg, v);
      if (sizeof(S) == 8 && out.dtype == YSQL_DATETIME) {
        uint64_t *dst = (uint64_t *)out.data + row0;
        if (full && ((uintptr_t)dst & 15) == 0 && (R % 2) == 0) {
#pragma unroll
          for (int q = 0; q < R / 2; q++) {
            uint4 w;
            w.x = (uint32_t)(uint64_t)v[2 * q];
            w.y = (uint32_t)((uint64_t)v[2 * q] >> 32);
            w.z = (uint32_t)(uint64_t)v[2 * q + 1];
            w.w = (uint32_t)((uint64_t)v[2 * q + 1] >> 32);
            ((uint4 *)dst)[q] = w;
          }
        } else {
#pragma unroll
          for (int e = 0; e < R; e++)
            if (row0 + e < n) dst[e] = (uint64_t)v[e];
        }
      } else {
        uint32_t *dst = (uint32_t *)out.data + row0;
        if (full && ((uintptr_t)dst & 15) == 0 && (R % 4) == 0) {
#pragma unroll
          for (int q = 0; q < R / 4; q++)
            ((uint4 *)dst)[q] = make_uint4((uint32_t)v[4 * q], (uint32_t)v[4 * q + 1], (uint32_t)v[4 * q + 2],
                                           (uint32_t)v[4 * q + 3]);
        } else {
#pragma unroll
          for (int e = 0; e < R; e++)
            if (row0 + e < n) dst[e] = (uint32_t)v[e];
        }
      }
      if (out.nulls) {
        const uint32_t f = reg.nv[out.reg];
        uint8_t *dn = out.nulls + row0;
        if (full && R == 4 && ((uintptr_t)dn & 3) == 0) {
          *(uint32_t *)dn = (f & 1) | ((f >> 1) & 1) << 8 | ((f >> 2) & 1) << 16 | ((f >> 3) & 1) << 24;
        } else if (full && R == 2 && ((uintptr_t)dn & 1) == 0) {
          *(uint16_t *)dn = (uint16_t)((f & 1) | ((f >> 1) & 1) << 8);
        } else {
#pragma unroll
          for (int e = 0; e < R; e++)
            if (row0 + e < n) dn[e] = (uint8_t)((f >> e) & 1);
        }
      }
    }
  }
}

// Grid-stride over groups of kXBlock * R rows; thread t of a group owns R
// consecutive rows (16- or 32-byte loads and stores per column).
template <typename S, int R>
__global__ __launch_bounds__(kXBlock) void k_xeval(XProg p, uint64_t n) {
  XRegs<S, R> reg;
  reg.clear();
  S acc[R] = {};
  uint32_t nacc = 0;
  const uint64_t step = (uint64_t)gridDim.x * kXBlock * R;
  for (uint64_t g0 = (uint64_t)blockIdx.x * kXBlock * R; g0 < n; g0 += step) {
    const uint64_t row0 = g0 + (uint64_t)threadIdx.x * R;
    const bool fast = p.aligned && g0 + (uint64_t)kXBlock * R <= n;  // wave-uniform
    if (row0 < n) {
      xload_inputs<S, R, true>(p, row0, n, fast, reg);
      xrun<S, R, true>(p, reg, acc, nacc);
      xstore_outputs<S, R>(p, row0, n, reg);
    }
  }
}

}  // namespace ysql

using namespace ysql;

extern "C" {

int ysql_expr_eval(const ysql_xinsn *prog, int ninsn, const ysql_xinput *inputs, int ninputs,
                   const ysql_xoutput *outputs, int noutputs, uint64_t n, ysql_stream_t stream) {
  if (noutputs < 0 || noutputs > YSQL_MAX_COLS || (noutputs > 0 && !outputs)) return YSQL_EINVAL;
  XProg p;
  bool wide = false;
  int8_t ty[kXRegs];
  uint32_t live = 0;
  for (int o = 0; o < noutputs; o++) {
    if (outputs[o].reg < 0 || outputs[o].reg >= kXRegs) return YSQL_EINVAL;
    live |= 1u << outputs[o].reg;
  }
  int rc = xprog_prepare(prog, ninsn, inputs, ninputs, n, live, -1, &p, &wide, ty);
  if (rc) return rc;
  for (int o = 0; o < noutputs; o++) {
    const ysql_xoutput out = outputs[o];
    if (ty[out.reg] < 0) return YSQL_EINVAL;
    const int want = ty[out.reg] == YSQL_BOOL ? YSQL_INT : ty[out.reg];
    if (out.dtype != want || (n > 0 && !out.data)) return YSQL_EINVAL;
    p.out[o] = out;
  }
  p.nout = noutputs;
  if (n == 0 || noutputs == 0) return YSQL_OK;
  if (n >= (1ull << 40)) return YSQL_EINVAL;
  hipStream_t s = as_stream(stream);
  const uint64_t cap = (uint64_t)grid_cap() * 8;
  if (wide) {
    const uint64_t g = std::min<uint64_t>(ceil_div(n, (uint64_t)kXBlock * xrows<uint64_t>()), cap);
    k_xeval<uint64_t, xrows<uint64_t>()><<<(unsigned)g, kXBlock, 0, s>>>(p, n);
  } else {
    const uint64_t g = std::min<uint64_t>(ceil_div(n, (uint64_t)kXBlock * xrows<uint32_t>()), cap);
    k_xeval<uint32_t, xrows<uint32_t>()><<<(unsigned)g, kXBlock, 0, s>>>(p, n);
  }
  YSQL_LAUNCH_CHECK();
  return YSQL_OK;
}

}  // extern "C"
