// scan_dev.hpp — device-side stream compaction of one tile (used by the
// compaction kernels of scan.hip).
//
// Each thread owns 16 consecutive rows of a tile of BLOCK * 16 rows: it
// computes their selection flags (from a byte mask or a fused predicate),
// the block scans the per-thread counts, wave 0 finds the tile's exclusive
// prefix by decoupled look-back over 64-bit status words, the selected rows'
// tile offsets are listed in LDS in order, and every column is copied with
// one output row per thread per step (coalesced stores, loads inside the
// tile's cache lines).  An output column that is the predicate column is
// copied from the predicate values kept in LDS (no second read of it).
// Stable: selected rows keep their order.
#pragma once

#include <type_traits>

#include "common.hpp"

namespace ysql {

// VB: bytes of a predicate value kept in LDS (0: none: mask selection, and
// 8-byte predicates, whose 32 KB would halve the workgroups per CU).
template <int BLOCK, int VB = 0>
struct CompactSmem {
  using V = typename std::conditional<VB == 8, uint64_t, uint32_t>::type;
  uint16_t src[BLOCK * 16];
  uint64_t red[BLOCK / kWave + 1];
  uint64_t prefix;
  alignas(16) V pv[VB ? BLOCK * 16 : 1];
};

template <typename PT>
__device__ __forceinline__ void load16(const PT *p, uint64_t base, uint64_t rows, int t, PT (&v)[16]) {
  const uint64_t r0 = (uint64_t)t * 16;
  if (r0 + 16 <= rows && ((uintptr_t)(p + base + r0) & 15) == 0) {
    const uint4 *q = (const uint4 *)(p + base + r0);
    constexpr int NQ = 16 * sizeof(PT) / 16;
    uint4 tmp[NQ];
#pragma unroll
    for (int i = 0; i < NQ; i++) tmp[i] = q[i];
    memcpy(&v[0], &tmp[0], sizeof(tmp));
  } else {
#pragma unroll
    for (int e = 0; e < 16; e++) v[e] = (r0 + e < rows) ? p[base + r0 + e] : PT();
  }
}

// Copies the listed rows of one column: dst[prefix + r] = src[base + idx[r]]
// for r < tot.  Each thread loads its next 8 rows (r = t, t + BLOCK, ...)
// before it stores any, so 8 loads per thread are in flight at once instead
// of one load's latency per row.
template <int BLOCK, typename E>
__device__ __forceinline__ void gather_listed(const E *src, E *dst, const uint16_t *idx, uint64_t base,
                                              uint64_t prefix, uint32_t tot) {
  constexpr int U = 8;
  for (uint32_t r0 = threadIdx.x; r0 < tot; r0 += BLOCK * U) {
    E v[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const uint32_t r = r0 + u * BLOCK;
      v[u] = r < tot ? src[base + idx[r]] : E();
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const uint32_t r = r0 + u * BLOCK;
      if (r < tot) dst[prefix + r] = v[u];
    }
  }
}

// Second half of a compaction tile: given each thread's selection flags for
// its 16 consecutive rows (bit e = row t*16 + e of the tile), scans the
// counts, finds the tile's prefix (look-back over `status`, or `running`
// when status is nullptr), lists the selected rows in LDS and copies every
// column.  VB != 0: the predicate values `pred` of the thread's rows were
// stored in sm.pv, and an output column reading `pred` is copied from there.
// Every thread of the block calls it; contains barriers.
template <int BLOCK, int VB, typename PT>
__device__ __forceinline__ uint64_t compact_scatter(CompactSmem<BLOCK, VB> &sm, uint32_t tile, uint32_t flags,
                                                    const PT *pred, const ysql_col *cols, int ncols, uint64_t n,
                                                    uint64_t cap, uint64_t running, uint64_t *count_out,
                                                    uint64_t *status, uint32_t ntiles) {
  constexpr int TILE = BLOCK * 16;
  const uint64_t base = (uint64_t)tile * TILE;
  const int t = threadIdx.x;
  const uint64_t rows = (n - base) < (uint64_t)TILE ? (n - base) : (uint64_t)TILE;
  {
    const uint64_t r0 = (uint64_t)t * 16;
    if (r0 + 16 > rows) flags &= (r0 >= rows) ? 0u : ((1u << (rows - r0)) - 1u);
  }
  const uint32_t cnt = __builtin_popcount(flags);
  uint64_t total;
  const uint32_t excl = (uint32_t)block_exclusive_sum<BLOCK, uint64_t>(cnt, sm.red, &total);
  if (t < kWave) {
    // with no status array the caller processes tiles in order and passes
    // the running prefix
    const uint64_t p = status ? lookback_exclusive(status, tile, total) : running;
    if (t == 0) sm.prefix = p;
  }
  {
    uint32_t k = excl;
    uint32_t f = flags;
    while (f) {
      const int e = __builtin_ctz(f);
      f &= f - 1;
      sm.src[k++] = (uint16_t)(t * 16 + e);
    }
  }
  __syncthreads();
  const uint64_t prefix = sm.prefix;
  // rows whose output index reaches `cap` are not written (the count still is)
  const uint32_t tot = prefix >= cap ? 0u : (uint32_t)(cap - prefix < total ? cap - prefix : total);
  for (int c = 0; c < ncols; c++) {
    const ysql_col col = cols[c];
    if constexpr (VB != 0 && VB == (int)sizeof(PT)) {
      if (col.src == (const void *)pred && col.elem_bytes == VB) {
        using V = typename CompactSmem<BLOCK, VB>::V;
        V *dst = (V *)col.dst;
        for (uint32_t r = t; r < tot; r += BLOCK) dst[prefix + r] = sm.pv[sm.src[r]];
        if (col.dst_null) {
          for (uint32_t r = t; r < tot; r += BLOCK)
            col.dst_null[prefix + r] = col.src_null ? col.src_null[base + sm.src[r]] : 0;
        }
        continue;
      }
    }
    switch (col.elem_bytes) {
      case 4: {
        gather_listed<BLOCK, uint32_t>((const uint32_t *)col.src, (uint32_t *)col.dst, sm.src, base, prefix, tot);
        break;
      }
      case 8: {
        gather_listed<BLOCK, uint64_t>((const uint64_t *)col.src, (uint64_t *)col.dst, sm.src, base, prefix, tot);
        break;
      }
      case 1: {
        gather_listed<BLOCK, uint8_t>((const uint8_t *)col.src, (uint8_t *)col.dst, sm.src, base, prefix, tot);
        break;
      }
      case 2: {
        gather_listed<BLOCK, uint16_t>((const uint16_t *)col.src, (uint16_t *)col.dst, sm.src, base, prefix, tot);
        break;
      }
      default: {  // 256-byte strings: 16 chunks of 16 bytes per row
        const uint4 *src = (const uint4 *)col.src;
        uint4 *dst = (uint4 *)col.dst;
        for (uint32_t x = t; x < tot * 16; x += BLOCK) {
          const uint32_t r = x >> 4, ch = x & 15;
          dst[(prefix + r) * 16 + ch] = src[(base + sm.src[r]) * 16 + ch];
        }
        break;
      }
    }
    if (col.dst_null) {
      if (col.src_null) {
        gather_listed<BLOCK, uint8_t>(col.src_null, col.dst_null, sm.src, base, prefix, tot);
      } else {
        for (uint32_t r = t; r < tot; r += BLOCK) col.dst_null[prefix + r] = 0;
      }
    }
  }
  if (tile == ntiles - 1 && t == 0 && count_out) *count_out = prefix + total;
  __syncthreads();  // LDS is reused by the caller's next tile
  return total;
}

// MODE 0: selection from mask bytes; MODE 1: pred[i] <CMP> scalar;
// MODE 2: scalar <= pred[i] < hi.  Output rows at index >= cap are not
// written; *count_out receives the full count.  status == nullptr: tiles are
// processed in order by one caller that passes the running prefix (no
// look-back).  Returns the tile's selected-row count.
// Every thread of the block calls it; contains barriers.
template <int BLOCK, int MODE, typename PT, int CMP, int VB>
__device__ __forceinline__ uint64_t compact_tile(CompactSmem<BLOCK, VB> &sm, uint32_t tile, const uint8_t *mask,
                                             const PT *pred, PT scalar, PT hi, const ysql_col *cols, int ncols,
                                             uint64_t n, uint64_t cap, uint64_t running,
                                             uint64_t *count_out, uint64_t *status, uint32_t ntiles) {
  constexpr int TILE = BLOCK * 16;
  const uint64_t base = (uint64_t)tile * TILE;
  const int t = threadIdx.x;
  const uint64_t rows = (n - base) < (uint64_t)TILE ? (n - base) : (uint64_t)TILE;

  uint32_t flags = 0;
  if constexpr (MODE == 0) {
    uint8_t m[16];
    load16<uint8_t>(mask, base, rows, t, m);
#pragma unroll
    for (int e = 0; e < 16; e++) flags |= (uint32_t)(m[e] != 0) << e;
  } else {
    PT v[16];
    load16<PT>(pred, base, rows, t, v);
#pragma unroll
    for (int e = 0; e < 16; e++) {
      const bool sel = MODE == 1 ? cmp_static<CMP>(v[e], scalar) : (!(v[e] < scalar) && v[e] < hi);
      flags |= (uint32_t)sel << e;
    }
    if constexpr (VB == (int)sizeof(PT)) {
      // 16-byte LDS stores of the thread's 16 values, when it selected any
      // (only selected rows are read back)
      if (flags) {
        constexpr int NQ = 16 * VB / 16;
        uint4 q[NQ];
        memcpy(&q[0], &v[0], sizeof(q));
        uint4 *d = reinterpret_cast<uint4 *>(&sm.pv[t * 16]);
#pragma unroll
        for (int i = 0; i < NQ; i++) d[i] = q[i];
      }
    }
  }
  // with mask selection no column is copied from the LDS values
  return compact_scatter<BLOCK, VB, PT>(sm, tile, flags, MODE != 0 ? pred : (const PT *)nullptr, cols, ncols, n, cap,
                                        running, count_out, status, ntiles);
}

}  // namespace ysql
