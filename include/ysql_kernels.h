/*
 * ysql_kernels.h — C ABI of the YallaSQL operator hot path for AMD MI355X (gfx950).
 *
 * Every entry point replaces one host launcher of the reference's
 * `YallaSQL::Kernel` namespace (the reference's headers under include/kernels/),
 * which the reference's operators (under include/engine/operators/) call with
 * device pointers and a stream.  The reference launchers are C++ templates over
 * element type and functor; here they are type-tagged C functions.
 *
 * Conventions (all entry points):
 *   - Data pointers are device pointers unless the parameter says "host".
 *   - Scalars of the vector ops are 1-element device buffers, as in the
 *     reference (`bound_value.hpp:21-48`).  Null masks are 1 byte per row
 *     (1 = NULL), as the reference's NullBitSet (`null_bit_set.hpp:5-31`).
 *   - Outputs and workspaces are owned by the caller.  Nothing allocates
 *     device memory inside a call and nothing leaks (the reference's
 *     `launch_radix_sort` allocated and leaked, `radix_sort_kernel.cu:139-158`).
 *   - Everything is asynchronous on `stream` (a hipStream_t; NULL = the null
 *     stream).  Device-side counts (filter count, pair count, consumed_a) are
 *     written to device memory; the caller decides when to synchronise.
 *   - Every entry point returns an int status (YSQL_OK = 0, negative = error)
 *     and never exits the process (the reference's CUDA_CHECK called exit(),
 *     `utils/macros.hpp:6-22`).
 *   - Entry points keep no global mutable state: they are reentrant across
 *     streams.  Calls that share one workspace on two streams at once race.
 *     The one exception is ysql_filter_compact_noalloc, whose scratch is a
 *     static device array shared by every call on a GPU (see there).
 */
#ifndef YSQL_KERNELS_H
#define YSQL_KERNELS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void *ysql_stream_t; /* hipStream_t */

/* ---- status codes ---------------------------------------------------- */
enum {
  YSQL_OK = 0,
  YSQL_EINVAL = -1,      /* bad argument (null pointer, unsupported dtype/op, n too large) */
  YSQL_EWORKSPACE = -2,  /* workspace smaller than the matching *_ws_bytes() */
  YSQL_EOVERFLOW = -3,   /* capacity exceeded (reported by host-side wrappers) */
  YSQL_EHIP = -4,        /* a HIP runtime call failed */
  YSQL_EUNSUPPORTED = -5 /* combination not implemented */
};

/* ---- data types: DataType of the reference (enums/data_type.hpp:2-7) -- */
enum {
  YSQL_INT = 0,      /* int32                                   */
  YSQL_FLOAT = 1,    /* float32                                 */
  YSQL_DATETIME = 2, /* int64 (epoch seconds)                   */
  YSQL_STRING = 3,   /* 256-byte NUL-padded (string_kernel.hpp) */
  YSQL_BOOL = 4,     /* uint8 mask                              */
  YSQL_UKEY32 = 16,  /* uint32 order-preserving sort key        */
  YSQL_UKEY64 = 17   /* uint64 order-preserving sort key        */
};
#define YSQL_STRING_BYTES 256

/* ---- comparison ops: comparison_operators_kernel.hpp:25-104 ----------
 * Note the reference names: LEOperator is '<', GEOperator is '>'.         */
enum { YSQL_LT = 0, YSQL_LE = 1, YSQL_GT = 2, YSQL_GE = 3, YSQL_EQ = 4, YSQL_NE = 5 };
/* ---- mask logic: ANDOperator / OROperator ---------------------------- */
enum { YSQL_AND = 0, YSQL_OR = 1 };
/* ---- arithmetic: binary_operators_kernel.hpp:27-51 ------------------- */
enum { YSQL_ADD = 0, YSQL_SUB = 1, YSQL_MUL = 2, YSQL_DIV = 3, YSQL_REM = 4 };
/* ---- aggregates: AggType, aggregate.hpp:7-9 --------------------------- */
enum {
  YSQL_AGG_MAX = 0,
  YSQL_AGG_MIN = 1,
  YSQL_AGG_SUM = 2,
  YSQL_AGG_AVG = 3,
  YSQL_AGG_COUNT = 4,
  YSQL_AGG_COUNT_STAR = 5
};

/* Accumulator of one aggregate, resident in device memory and folded into
 * across batches (the reference keeps one device accumulator per
 * AggregateExpression, aggregate.hpp:21-30,201-205).
 *   INT   MIN/MAX -> i (int32 range), SUM/AVG -> i (int64; the reference's
 *                    int32 sum wraps, aggregate.hpp:122-124 — divergent on purpose)
 *   FLOAT MIN/MAX -> f, SUM/AVG -> d (double, starting at exactly 0)
 *   DATETIME      -> i
 *   count         -> rows folded that were not NULL (AVG's divisor, COUNT). */
typedef struct ysql_acc {
  int64_t i;
  double d;
  float f;
  uint32_t pad;
  int64_t count;
} ysql_acc;

/* Column descriptor for the fused multi-column ops (host struct).  dst must
 * not overlap src (nor another column's dst); src == dst is rejected. */
typedef struct ysql_col {
  const void *src;
  const uint8_t *src_null; /* may be NULL: no NULLs (output nulls zeroed) */
  void *dst;
  uint8_t *dst_null; /* may be NULL: output null bytes not written */
  uint32_t elem_bytes; /* 1, 4, 8 or 256 */
  uint32_t pad;
} ysql_col;
#define YSQL_MAX_COLS 8

/* ======================================================================
 * Sort-key normalisation (A1)
 * Replaces launch_int_to_uint32 / launch_float_to_uint32
 * (radix_sort_kernel.hpp:83-89 -> radix_sort_utils.cu:39-53), called from
 * BoundRefSortExpression::evaluate (bound_ref_sort.hpp:46,50).
 *   INT      -> uint32  x ^ 0x80000000
 *   FLOAT    -> uint32  sign ? ~bits : bits ^ 0x80000000   (IEEE total order)
 *   DATETIME -> uint64  x ^ 2^63  (the reference used raw bits,
 *                       bound_ref_sort.hpp:53-55: divergent for negatives)
 * `descending` != 0 additionally complements the key, so an ascending stable
 * sort of the output yields the descending order with ties in input order. */
int ysql_sortkey_encode(int dtype, int descending, const void *in, void *out, uint64_t n,
                        ysql_stream_t stream);
/* Inverse of ysql_sortkey_encode (uint32_to_int / uint32_to_float,
 * radix_sort_utils.cu:12-37). */
int ysql_sortkey_decode(int dtype, int descending, const void *in, void *out, uint64_t n,
                        ysql_stream_t stream);
/* Appends one more ORDER BY column to a composite 64-bit key:
 *     key[i] = (key[i] << w) | encode(in[i])   (w = 32 for INT/FLOAT, 64 for
 * DATETIME), so an ascending stable sort of the composite orders rows by
 * the first column, then the second, ... (the reference sorted on orders[0]
 * only, order_operator.hpp:277).                                           */
int ysql_sortkey_pack(int dtype, int descending, const void *in, uint64_t n, uint64_t *key_inout,
                      ysql_stream_t stream);

/* ======================================================================
 * LSD radix sort, 8-bit digits (A2-A6)
 * Replaces launch_radix_sort<OP,T,BlockDim>(T* d_in, uint32_t*& d_idx, N, stream)
 * (radix_sort_kernel.hpp:22-23 -> radix_sort_kernel.cu:139-178), called from
 * OrderOperator::sortLocallyBatch (order_operator.hpp:196-265).
 * Keys are order-preserving unsigned keys (YSQL_UKEY32 / YSQL_UKEY64, the
 * output of ysql_sortkey_encode).  Stable.  On return keys_inout holds the
 * keys in sorted order (descending if `descending`, like DescOp) and
 * perm_out[sorted_pos] = original row.  Only bits [begin_bit, end_bit) are
 * ranked (pass 0, key_bits for the full key).  n < 2^30.               */
int ysql_radix_sort_ws_bytes(int key_dtype, uint64_t n, size_t *bytes);
int ysql_radix_argsort(int key_dtype, int descending, void *keys_inout, uint32_t *perm_out,
                       uint64_t n, int begin_bit, int end_bit, void *ws, size_t ws_bytes,
                       ysql_stream_t stream);
/* Same sort with a caller-supplied uint32 payload (vals_inout) carried along
 * in place of the row index (used for runs whose payload is a global row id). */
int ysql_radix_sort_pairs(int key_dtype, int descending, void *keys_inout, uint32_t *vals_inout,
                          uint64_t n, int begin_bit, int end_bit, void *ws, size_t ws_bytes,
                          ysql_stream_t stream);
/* Out-of-place form: (keys_in, vals_in) -> (keys_out, vals_out).  keys_in
 * may equal keys_out (in place) or be disjoint from it.  vals_in == NULL
 * gives payload val_base + row (a run of a larger table carries its global
 * row ids this way).  Saves the D2D copy that BoundRefSortExpression made
 * before sorting (bound_ref_sort.hpp:35-44).                               */
int ysql_radix_sort_ex(int key_dtype, int descending, const void *keys_in, const uint32_t *vals_in,
                       uint32_t val_base, void *keys_out, uint32_t *vals_out, uint64_t n,
                       int begin_bit, int end_bit, void *ws, size_t ws_bytes, ysql_stream_t stream);
/* Same sort with an explicit tile shape of the digit passes (variant 0:
 * 512 threads x 16 keys, 1: 256 x 16, 2: 256 x 8; -1 = the default for n,
 * which ysql_radix_sort_ex uses).  Every variant returns the same stable
 * permutation; they differ in speed by size, which an autotuner measures
 * per size bucket.  Sorts of n <= 8192 keys run in one workgroup whatever
 * the variant.                                                            */
int ysql_radix_variant_count(void);
int ysql_radix_sort_variant(int key_dtype, int variant, int descending, const void *keys_in,
                            const uint32_t *vals_in, uint32_t val_base, void *keys_out,
                            uint32_t *vals_out, uint64_t n, int begin_bit, int end_bit, void *ws,
                            size_t ws_bytes, ysql_stream_t stream);
/* Stepwise form of the same sort, for callers that schedule or time the
 * passes themselves (the reference's launcher loops over digit passes,
 * radix_sort_kernel.cu:161-167): ysql_radix_begin builds the digit
 * histograms of keys_in; then ysql_radix_pass for pass = 0 .. P-1,
 * P = ceil((end_bit - begin_bit) / 8), each reading the previous pass's
 * output (pass 0: keys_in, vals_in or val_base + row) into a distinct
 * buffer: keys_out == keys_in, or a non-null vals_in == vals_out, returns
 * YSQL_EINVAL (tiles scatter while other tiles still read their input).
 * All calls of one sort share `ws` and the stream.                        */
int ysql_radix_begin(int key_dtype, int descending, const void *keys_in, uint64_t n, int begin_bit,
                     int end_bit, void *ws, size_t ws_bytes, ysql_stream_t stream);
int ysql_radix_pass(int key_dtype, int descending, int pass, const void *keys_in,
                    const uint32_t *vals_in, uint32_t val_base, void *keys_out, uint32_t *vals_out,
                    uint64_t n, int begin_bit, int end_bit, void *ws, size_t ws_bytes,
                    ysql_stream_t stream);

/* k-way merge of nruns <= 16 sorted runs in one pass (the reference merges
 * runs pairwise, level by level, OrderOperator::merge_all_sorted,
 * order_operator.hpp:160-193, order_operator.cpp:116-220, so each row moves
 * ceil(log2 nruns) times; here once).  Run r is keys[run_offsets[r],
 * run_offsets[r+1]) with its payloads at the same positions of vals
 * (run_offsets: host array of nruns + 1 ascending offsets).  Keys ascending
 * (YSQL_UKEY32 / YSQL_UKEY64, encoded order-preserving keys: a descending
 * order is an ascending one of complemented keys).  Output: keys_out /
 * vals_out[0, total) in merged order, ties by run, then by position: for
 * runs cut in row order and sorted stably, one global stable sort.  Output
 * buffers must differ from the inputs.  total < 2^32.                   */
int ysql_merge_runs_ws_bytes(int key_dtype, const uint64_t *run_offsets, int nruns, size_t *bytes);
int ysql_merge_runs(int key_dtype, const void *keys, const uint32_t *vals, const uint64_t *run_offsets,
                    int nruns, void *keys_out, uint32_t *vals_out, void *ws, size_t ws_bytes,
                    ysql_stream_t stream);

/* ======================================================================
 * STRING sort keys and merges
 * Replace launch_radix_sort<AscOpString / DescOpString, String, ...>
 * (radix_sort_kernel.hpp:45-73, radix_sort_kernel.cu:174,178), called from
 * OrderOperator::sortLocallyBatch (order_operator.hpp:220,250), and
 * launch_merge_sorted_array_kernel<String, AscSortStrOp / DescSortStrOp>
 * (merge_batchs.hpp:36-51).  Order: strcmp_device's (string_kernel.hpp:46-56):
 * unsigned bytes up to the first NUL, at most 256; stable (ties keep their
 * order; DESC reverses the order, not the ties).  String columns are
 * YSQL_STRING_BYTES-byte rows, 8-byte aligned.
 *
 * ysql_string_sort: perm_inout holds row indices of strs (each < n, e.g. a
 * permutation); on return it is stably reordered by strs[perm_inout[i]].
 * Sorting an earlier sort's permutation by another column this way is one
 * LSD step of a multi-column ORDER BY.  The call reads the longest string's
 * length back to the host once (it sets the number of 8-byte chunk passes),
 * so it synchronises `stream`.  n < 2^30.
 * ysql_string_argsort: perm_out = the stable sorting permutation of strs.  */
int ysql_string_sort_ws_bytes(uint64_t n, size_t *bytes);
int ysql_string_sort(int descending, const void *strs, uint32_t *perm_inout, uint64_t n, void *ws,
                     size_t ws_bytes, ysql_stream_t stream);
int ysql_string_argsort(int descending, const void *strs, uint32_t *perm_out, uint64_t n, void *ws,
                        size_t ws_bytes, ysql_stream_t stream);
/* One ORDER BY key column of two sorted runs: a (m rows) and b (n rows) of
 * type YSQL_INT, YSQL_FLOAT, YSQL_DATETIME or YSQL_STRING.  Numeric columns
 * compare by ysql_sortkey_encode's order (the order the sorts use). */
typedef struct ysql_sortcol {
  int dtype;
  int descending;
  const void *a;
  const void *b;
} ysql_sortcol;
/* Stable merge of two runs sorted by nkeys (<= YSQL_MAX_COLS) key columns,
 * compared lexicographically, ties from A first: the first k_max outputs as
 * src_is_a[k] / src_idx[k] (ysql_merge_path's outputs, for ysql_gather2);
 * *consumed_a_out (device, may be NULL) = A rows among them.  keys_out
 * (may be NULL; one STRING key only, 16-byte aligned) receives the merged
 * strings.  No workspace.  ysql_merge_path with dtype YSQL_STRING is the
 * one-column form.                                                         */
int ysql_merge_rows(const ysql_sortcol *keys, int nkeys, uint64_t m, uint64_t n, uint64_t k_max,
                    uint32_t *src_idx, uint8_t *src_is_a, void *keys_out, int64_t *consumed_a_out,
                    ysql_stream_t stream);

/* ======================================================================
 * Inclusive scans (A4, A13)
 * Replace launch_prefix_sum<uint32_t> (prefix_sum.hpp:28-29 -> .cu:158-178)
 * and launch_prefix_sum_mask (prefix_sum.hpp:26 -> .cu:140-156): out[i] =
 * in[0] + ... + in[i]; for the mask form out[i] = #nonzero bytes in [0..i]. */
int ysql_scan_ws_bytes(uint64_t n, size_t *bytes);
int ysql_scan_u32_inclusive(const uint32_t *in, uint32_t *out, uint64_t n, void *ws,
                            size_t ws_bytes, ysql_stream_t stream);
int ysql_scan_mask_inclusive(const uint8_t *mask, uint32_t *out, uint64_t n, void *ws,
                             size_t ws_bytes, ysql_stream_t stream);

/* ======================================================================
 * Merge path (A8, A9)
 * Replaces launch_merge_sorted_array_kernel<T,OP>(A,B,C,new_idx,table_idx,
 * lasti,k_mx,m,n,stream) (merge_batchs.hpp:50-51 -> merge_batchs.cu:143-158),
 * called from OrderOperator::merge_two_runs (order_operator.cpp:116-220).
 * Emits the first k_max outputs of the stable merge of sorted A (m) and B (n):
 * src_is_a[k] = 1 if output k comes from A, src_idx[k] = its index in its
 * source, C_opt[k] (optional) the key.  Ties take A first (the rule of
 * merge_batchs.cu:61; the reference's co-rank used a strict compare,
 * merge_batchs.cu:22, which duplicates and drops rows on ties — fixed here).
 * *consumed_a_out (device int64) = number of A elements among the outputs.
 * dtype: INT/FLOAT/DATETIME compare typed values (FLOAT with IEEE '<', as the
 * reference), UKEY32/UKEY64 compare encoded keys, STRING compares with
 * strcmp_device's order (C_opt then receives the 256-byte strings; A, B and
 * C_opt 16-byte aligned).  `descending` merges descending sequences
 * (DescSortMergeOp / DescSortStrOp).                                      */
int ysql_merge_ws_bytes(uint64_t m, uint64_t n, uint64_t k_max, size_t *bytes);
int ysql_merge_path(int dtype, int descending, const void *A, uint64_t m, const void *B,
                    uint64_t n, uint64_t k_max, void *C_opt, uint32_t *src_idx,
                    uint8_t *src_is_a, int64_t *consumed_a_out, void *ws, size_t ws_bytes,
                    ysql_stream_t stream);
/* Fused merge of (key, uint32 payload) runs: out = stable merge, ties from A
 * first.  This is merge_path + the two gathers of move_rows_batch_merge
 * (move_rows_kernel.hpp:85-143) in one pass, for encoded keys only.        */
int ysql_merge_pairs(int key_dtype, int descending, const void *a_keys, const uint32_t *a_vals,
                     uint64_t m, const void *b_keys, const uint32_t *b_vals, uint64_t n,
                     void *out_keys, uint32_t *out_vals, void *ws, size_t ws_bytes,
                     ysql_stream_t stream);

/* ======================================================================
 * Row movement (A7, A10, A14, A20)
 * launch_move_one_col (move_rows_kernel.cu:51-70,136-143):
 *     dst[i] = src[idx[i]], dst_null[i] = src_null[idx[i]]                 */
int ysql_gather(uint32_t elem_bytes, const void *src, const uint8_t *src_null,
                const uint32_t *idx, uint64_t n, void *dst, uint8_t *dst_null,
                ysql_stream_t stream);
/* launch_move_one_col_merge (move_rows_kernel.cu:72-105,146-156):
 *     dst[d] = (is_a[i] ? a : b)[idx[i]], d = reverse ? n-1-i : i           */
int ysql_gather2(uint32_t elem_bytes, const void *a, const uint8_t *a_null, const void *b,
                 const uint8_t *b_null, const uint32_t *idx, const uint8_t *is_a, uint64_t n,
                 int reverse, void *dst, uint8_t *dst_null, ysql_stream_t stream);
/* lanch_move_rows_join_kernel (move_rows_kernel.cu:7-30,170-180):
 *     dst[i] = src[pairs[2i + side]]  (side 0 = left, 1 = right)             */
int ysql_gather_pairs(uint32_t elem_bytes, const void *src, const uint8_t *src_null,
                      const uint32_t *pairs, int side, uint64_t n, void *dst, uint8_t *dst_null,
                      ysql_stream_t stream);
/* launch_move_rows_filter_kernel (move_rows_kernel.cu:32-49,158-167):
 *     if mask[i]: dst[map[i]-1] = src[i]   (map = inclusive mask scan).
 * NULL rows are written with their null byte (the reference skipped them and
 * left holes, move_rows_kernel.cu:45 — divergent on purpose).              */
int ysql_compact(uint32_t elem_bytes, const void *src, const uint8_t *src_null,
                 const uint8_t *mask, const uint32_t *map, uint64_t n, void *dst,
                 uint8_t *dst_null, ysql_stream_t stream);
/* Fused FilterOperator::next (filter_operator.hpp:31-98): mask -> scan ->
 * compaction of every column in one pass with no host-visible intermediate.
 * Stable.  *count_out (device uint64) = number of selected rows.           */
int ysql_compact_ws_bytes(uint64_t n, size_t *bytes);
int ysql_compact_cols(const uint8_t *mask, const ysql_col *cols, int ncols, uint64_t n,
                      uint64_t *count_out, void *ws, size_t ws_bytes, ysql_stream_t stream);
/* Same, with the predicate `pred_col <cmp> scalar` evaluated in the same
 * pass (no mask in HBM).  `scalar` is a 1-element device buffer of dtype. */
int ysql_filter_compact(int dtype, int cmp, const void *pred_col, const void *scalar,
                        const ysql_col *cols, int ncols, uint64_t n, uint64_t *count_out,
                        void *ws, size_t ws_bytes, ysql_stream_t stream);
/* ysql_filter_compact without a workspace argument, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid allocation).  No counterpart in the reference.  Its
 * look-back scratch is a static device array of the library (about 8 MiB,
 * held by each GPU it runs on until the process ends), so it allocates
 * nothing; n above 2^32 rows returns YSQL_EWORKSPACE.  Every call on one GPU
 * shares that array: calls that overlap in time on one GPU (other streams,
 * other threads) race and must be serialised by the caller.               */
int ysql_filter_compact_noalloc(int dtype, int cmp, const void *pred_col, const void *scalar,
                                const ysql_col *cols, int ncols, uint64_t n,
                                uint64_t *count_out, ysql_stream_t stream);
/* Same with a range predicate bounds[0] <= pred_col[i] < bounds[1] (bounds:
 * 2-element device array; the reference expresses it as two comparisons
 * ANDed, conjunction.hpp:76-113, and compacts from the mask) and an output
 * capacity: rows past `cap` are not written, *count_out is the full count
 * (the caller retries with a larger capacity, like ysql_join_pairs).      */
int ysql_filter_compact_range(int dtype, const void *pred_col, const void *bounds,
                              const ysql_col *cols, int ncols, uint64_t n, uint64_t cap,
                              uint64_t *count_out, void *ws, size_t ws_bytes, ysql_stream_t stream);

/* ======================================================================
 * Predicates (A12)
 * Replaces launch_conditional_operators<T,Op>(d_rhs, d_lhs, t_rhs, t_lhs,
 * d_res, sz, stream, isneg) (comparison_operators_kernel.hpp:7-8 ->
 * comparison_operators_kernel.cu:114-199):
 *     mask[i] = op(lhs[i or 0], rhs[i or 0]) XOR negate   (operand order op(lhs, rhs)) */
int ysql_compare(int dtype, int cmp, const void *lhs, int lhs_is_scalar, const void *rhs,
                 int rhs_is_scalar, int negate, uint8_t *mask_out, uint64_t n,
                 ysql_stream_t stream);
/* AND / OR of masks (launch_conditional_operators<bool, ANDOperator/OROperator>). */
int ysql_mask_logic(int op, const uint8_t *a, int a_is_scalar, const uint8_t *b,
                    int b_is_scalar, uint8_t *out, uint64_t n, ysql_stream_t stream);

/* ======================================================================
 * Projection arithmetic and casts (A15)
 * launch_binary_operators<T,Op> (binary_operators_kernel.hpp:15-16 -> .cu:58-75):
 *     out[i] = op(lhs[i or 0], rhs[i or 0]); DIV/REM by 0 -> 0; REM int only.  */
int ysql_arith(int dtype, int op, const void *lhs, int lhs_is_scalar, const void *rhs,
               int rhs_is_scalar, void *out, uint64_t n, ysql_stream_t stream);
/* launch_numerical_cast<S,D> (cast_operators_kernel.hpp:8-9 -> .cu:24-43):
 *     out[i] = (D) in[i]  for INT/FLOAT/DATETIME/BOOL sources.              */
int ysql_cast(int src_dtype, int dst_dtype, const void *in, void *out, uint64_t n,
              ysql_stream_t stream);

/* ======================================================================
 * Fused expression programs (expression fusion)
 * Replace the node-at-a-time evaluation of a predicate or projection tree:
 * BoundFuncExpression::evaluate (bound_func.hpp:45-167), CastExpression
 * (cast.hpp:22-90), ComparisonExpression (filters/comparison.hpp:51-215) and
 * ConjuntionExpression (filters/conjunction.hpp:76-113), each of which
 * launches one kernel into a freshly allocated column per tree node.  Here a
 * tree is a short register program evaluated per row in one pass: input
 * columns are read once, no intermediate column is written to HBM.
 *
 * Registers 0 .. ninputs-1 hold the input columns; the other registers up to
 * YSQL_XMAX_REGS - 1 are temporaries.  A register holds one typed value per
 * row (dtype INT, FLOAT, DATETIME or BOOL) and a NULL flag per row.
 *   YSQL_XOP_CONST  r[dst] = imm (dtype `type`: INT/DATETIME take imm as an
 *                   integer, FLOAT the float whose bits are imm's low 32 bits);
 *                   NULL flag 0
 *   YSQL_XOP_CMP    r[dst] = r[a] <sub> r[b], both of dtype `type` -> BOOL
 *                   (ysql_compare's rule, sub = YSQL_LT..YSQL_NE)
 *   YSQL_XOP_LOGIC  r[dst] = r[a] AND/OR r[b] (sub = YSQL_AND/YSQL_OR), BOOL
 *   YSQL_XOP_ARITH  r[dst] = r[a] <sub> r[b] over dtype `type` (ysql_arith's
 *                   rules: wrap, /0 -> 0, REM for integers only)
 *   YSQL_XOP_CAST   r[dst] = r[a] converted from dtype `type` to dtype `sub`
 *                   (ysql_cast's rules; BOOL sources to INT only)
 *   YSQL_XOP_NOT    r[dst] = !r[a] (BOOL; the `isneg` of the reference's
 *                   comparison and conjunction nodes)
 * Every instruction's result takes the NULL flags of register `nsrc`
 * (-1: none), the rule of the node it replaces (the non-constant operand's
 * nullset, bound_func.hpp:165, filters/comparison.hpp:213; the child's, cast.hpp:87;
 * the left operand's, filters/conjunction.hpp:111).  Operations run on the stored
 * values of NULL rows, as the per-node kernels do.                          */
enum {
  YSQL_XOP_CONST = 1,
  YSQL_XOP_CMP = 2,
  YSQL_XOP_LOGIC = 3,
  YSQL_XOP_ARITH = 4,
  YSQL_XOP_CAST = 5,
  YSQL_XOP_NOT = 6
};
#define YSQL_XMAX_REGS 16
#define YSQL_XMAX_INSN 48
#define YSQL_XMAX_INPUTS 8
typedef struct ysql_xinsn {
  int16_t op;   /* YSQL_XOP_* */
  int16_t type; /* operand dtype (CONST: the constant's dtype) */
  int16_t sub;  /* comparison / logic / arithmetic op, or CAST's target dtype */
  int8_t dst, a, b;
  int8_t nsrc;  /* register whose NULL flags the result takes, -1 = none */
  int64_t imm;  /* CONST value */
} ysql_xinsn;
typedef struct ysql_xinput {
  const void *data;     /* device column of `dtype` (INT, FLOAT, DATETIME) */
  const uint8_t *nulls; /* device null bytes, or NULL: no NULLs */
  int32_t dtype;
  int32_t pad;
} ysql_xinput;
typedef struct ysql_xoutput {
  void *data;     /* device column of `dtype`, n rows */
  uint8_t *nulls; /* device null bytes written from the register's flags, or NULL */
  int32_t reg;    /* register written; a BOOL register is written as INT 0/1 */
  int32_t dtype;  /* the register's dtype (INT for a BOOL register) */
} ysql_xoutput;
/* Projection: evaluates the program over n rows and writes every output.
 * The program is checked on the host (register bounds, operand dtypes,
 * defined-before-use); a bad program returns YSQL_EINVAL.               */
int ysql_expr_eval(const ysql_xinsn *prog, int ninsn, const ysql_xinput *inputs, int ninputs,
                   const ysql_xoutput *outputs, int noutputs, uint64_t n, ysql_stream_t stream);
/* Filter: the rows where BOOL register `pred_reg` is true are compacted
 * (ysql_compact_cols' contract: stable, up to YSQL_MAX_COLS columns, output
 * NULL bytes copied, *count_out = selected rows) in the same pass that
 * evaluates the program; the workspace is ysql_compact_ws_bytes(n).  As in
 * the per-node filter, the predicate's NULL flags do not affect selection.  */
int ysql_expr_filter_compact(const ysql_xinsn *prog, int ninsn, int pred_reg,
                             const ysql_xinput *inputs, int ninputs, const ysql_col *cols,
                             int ncols, uint64_t n, uint64_t *count_out, void *ws,
                             size_t ws_bytes, ysql_stream_t stream);

/* ======================================================================
 * Reductions (A16, A17)
 * Replace launch_reduction_operators<T,Tres,Op> / launch_reduction_count_notnull
 * (reduction_kernel.hpp:7-15 -> reduction_kernel.cu:78-205): fold n values
 * (NULL rows skipped) into the caller-owned accumulator *acc_inout, which
 * persists across batches as in the reference.  Deterministic: block partials
 * are combined in block order, so a float SUM depends only on the input.
 * COUNT_STAR counts all rows.                                              */
int ysql_acc_init(int dtype, int agg, ysql_acc *acc, int nacc, ysql_stream_t stream);
int ysql_reduce_ws_bytes(uint64_t n, size_t *bytes);
int ysql_reduce(int dtype, int agg, const void *in, const uint8_t *null_opt, uint64_t n,
                ysql_acc *acc_inout, void *ws, size_t ws_bytes, ysql_stream_t stream);
/* AggregateExpression::getAggregate (aggregate.hpp:149-192): writes the
 * final value: MIN/MAX -> dtype; SUM -> int64 (INT, DATETIME) or float
 * (FLOAT, from the double sum); AVG -> float = (double)sum / (double)count;
 * COUNT / COUNT_STAR -> int64.  Replaces div_avg_kernel and
 * convert_double_to_float_kernel (reduction_kernel.cu:5-25).             */
int ysql_agg_finalize(int dtype, int agg, const ysql_acc *acc, void *out, ysql_stream_t stream);

/* The reference's own accumulator form, for callers that keep it: one plain
 * typed value per aggregate (aggregate.hpp:201-205: INT -> int32, FLOAT
 * MIN/MAX -> float, FLOAT SUM/AVG -> double, DATETIME -> int64, COUNT and
 * AVG's counter -> int32).  ysql_acc_from_value loads such a value into a
 * ysql_acc; ysql_acc_to_value stores it back (an INT sum keeps its low 32
 * bits, wrapping as the reference's int32 accumulator does).  With
 * ysql_reduce between them they replace launch_reduction_operators and
 * launch_reduction_count_notnull (reduction_kernel.hpp:7-15 ->
 * reduction_kernel.cu:124-205), called from AggregateExpression::evaluate
 * (aggregate.hpp:79-137).                                                 */
int ysql_acc_from_value(int dtype, int agg, const void *value, ysql_acc *acc, ysql_stream_t stream);
int ysql_acc_to_value(int dtype, int agg, const ysql_acc *acc, void *value, ysql_stream_t stream);
/* launch_div_avg<T> (reduction_kernel.cu:5-15): *avg = (float)((double)*sum /
 * (double)*counter), sum int32 (INT), double (FLOAT) or int64 (DATETIME);
 * launch_convert_double_to_float_kernel (:17-25).  Called from
 * AggregateExpression::getAggregate (aggregate.hpp:149-192).             */
int ysql_div_avg(int dtype, const int32_t *counter, const void *sum, float *avg, ysql_stream_t stream);
int ysql_double_to_float(const double *in, float *out, ysql_stream_t stream);

/* Filtered multi-aggregate in one pass (FilterOperator feeding
 * AggregateOperator, filter_operator.hpp:31-98 + aggregate.hpp:55-146,
 * without materialising the filtered batch):
 *     for rows with pred_col[i] <cmp> scalar: fold specs[s].col[i] into
 *     *specs[s].acc under specs[s].agg   (s < nspecs <= 4).
 * COUNT and COUNT_STAR count the qualifying rows (this form takes no null
 * masks).  Deterministic: canonical tiles (4096..32768 rows, a function of n)
 * folded in a fixed order,
 * so the float SUM is a function of the input alone (the same bits from
 * every launch shape and from the batched executor's aggregate jobs).      */
typedef struct ysql_agg_spec {
  int dtype; /* YSQL_INT / YSQL_FLOAT / YSQL_DATETIME of col */
  int agg;   /* YSQL_AGG_* */
  const void *col;
  ysql_acc *acc; /* device accumulator, initialised with ysql_acc_init */
} ysql_agg_spec;
#define YSQL_MAX_AGG_SPECS 4
int ysql_filter_aggregate_ws_bytes(uint64_t n, int nspecs, size_t *bytes);
int ysql_filter_aggregate(int pred_dtype, int cmp, const void *pred_col, const void *scalar,
                          const ysql_agg_spec *specs, int nspecs, uint64_t n, void *ws,
                          size_t ws_bytes, ysql_stream_t stream);
/* Same, with flags: YSQL_FA_OVERWRITE stores this call's result into the
 * accumulators instead of folding it in (no ysql_acc_init needed; an empty
 * input stores the identities). */
#define YSQL_FA_OVERWRITE 1
/* YSQL_FA_GENERIC runs the general tile kernel even where the two-column
 * kernel applies (every spec reading the predicate column, one other column
 * or nothing); both give the same bits.  For tests and measurements.      */
#define YSQL_FA_GENERIC 2
int ysql_filter_aggregate_ex(int pred_dtype, int cmp, const void *pred_col, const void *scalar,
                             const ysql_agg_spec *specs, int nspecs, uint64_t n, int flags,
                             void *ws, size_t ws_bytes, ysql_stream_t stream);
/* Shared scan for up to 4 queries (no reference counterpart: operator
 * batching across in-flight queries).  Query q is `WHERE pred_col <cmp>
 * scalars[q]` with the aggregates specs[q * nspecs + s]; every query must
 * name the same columns, dtypes and aggregates (only the accumulators
 * differ).  The columns are read once for all queries, and each query's
 * result is bit-identical to running it alone through ysql_filter_aggregate_ex. */
#define YSQL_MAX_SHARED_QUERIES 4
int ysql_filter_aggregate_multi_ws_bytes(uint64_t n, int nqueries, int nspecs, size_t *bytes);
int ysql_filter_aggregate_multi(int pred_dtype, int cmp, const void *pred_col, const void *scalars,
                                int nqueries, const ysql_agg_spec *specs, int nspecs, uint64_t n,
                                int flags, void *ws, size_t ws_bytes, ysql_stream_t stream);

/* GROUP BY (no reference function: the reference's AggregateOperator never
 * reads `groups`, aggregate_operator.hpp:92-103).  acc_inout[g] folds the
 * rows with group_id == g; group_id >= n_groups skips the row.           */
int ysql_group_reduce(int dtype, int agg, const uint32_t *group_id, uint32_t n_groups,
                      const void *in, const uint8_t *null_opt, uint64_t n, ysql_acc *acc_inout,
                      ysql_stream_t stream);
/* Dense group ids over up to 4 int32 columns (host array of device
 * pointers; no reference function):
 *     gid[i] = sum_c (cols[c][i] - base[c]) * stride_c,
 * strides in column order (last column stride 1); a value outside
 * [base[c], base[c] + card[c]) gives UINT32_MAX, which ysql_group_reduce
 * skips.                                                                   */
int ysql_group_ids(const int32_t *const *cols, const int32_t *base, const uint32_t *card, int ncols,
                   uint64_t n, uint32_t *gid_out, ysql_stream_t stream);

/* ======================================================================
 * Joins (A18-A21)
 * Replaces launch_outer_join_operators<T,Op> (comparison_operators_kernel.hpp
 * :10-16 -> .cu:135-150), called from ComparisonJoinExpression::evaluate
 * (comparison_join.hpp:70-236): every (l, r) with op(left[l], right[r])
 * is a pair.  Pairs are written as 2 x uint32 (l, r), ordered by l then r
 * (the reference's order was nondeterministic; the pair SET is the same).
 * At most `cap` pairs are written; *count_out (device uint64) is the full
 * count, so count > cap means overflow (the reference wrote past its
 * 8 MiB buffer, comparison_join.hpp:87).  EQ uses a sort/hash build over
 * `right`; the other ops are a tiled nested loop.                         */
int ysql_join_ws_bytes(int dtype, int cmp, uint64_t L, uint64_t R, size_t *bytes);
int ysql_join_pairs(int dtype, int cmp, const void *left, uint64_t L, const void *right,
                    uint64_t R, uint32_t *pairs_out, uint64_t cap, uint64_t *count_out,
                    void *ws, size_t ws_bytes, ysql_stream_t stream);
/* launch_and_join_operators (comparison_operators_kernel.hpp:17-23 ->
 * .cu:24-54,152-167): pairs of p1 that also occur in p2, each p1 entry at
 * most once, in p1 order.                                                   */
int ysql_pairs_intersect_ws_bytes(uint64_t n1, uint64_t n2, size_t *bytes);
int ysql_pairs_intersect(const uint32_t *p1, uint64_t n1, const uint32_t *p2, uint64_t n2,
                         uint32_t *pairs_out, uint64_t cap, uint64_t *count_out, void *ws,
                         size_t ws_bytes, ysql_stream_t stream);
/* launch_cross_product_col<T> (cross_product_kernel.hpp:10-11 ->
 * cross_product_kernel.cu:29-40): for l < L, r < R, row = l*R + r:
 *     dst[row] = is_left ? src[l] : src[r]                                  */
int ysql_cross_expand(uint32_t elem_bytes, const void *src, const uint8_t *src_null, uint64_t L,
                      uint64_t R, int is_left, void *dst, uint8_t *dst_null,
                      ysql_stream_t stream);

/* Fused star join + grouped SUM (no reference function; composition of
 * join_pairs on dense primary keys + gather + arith + group_reduce).
 * For each fact row i: every dimension d must match (fk_d[i] in [0, rows_d)
 * and, if pass_d != NULL, pass_d[fk_d[i]] != 0); the group is
 *     g = sum_d (group_d[fk_d[i]] - group_base_d) * stride_d
 * over the dimensions with group_d != NULL (strides in dimension order,
 * stride of the last grouped dimension = 1), and the row adds
 * (int64)m1[i] * m2[i] to sums[g] and 1 to counts[g] (sums/counts are
 * folded into, not overwritten).                                          */
typedef struct ysql_star_dim {
  const int32_t *fk;      /* fact foreign-key column (n rows)              */
  uint32_t rows;          /* dimension rows; primary key = row index       */
  uint32_t group_card;    /* number of group values (0 = not grouped)      */
  const uint8_t *pass;    /* dimension filter result (rows bytes) or NULL  */
  const int32_t *group;   /* dimension group attribute (rows) or NULL      */
  int32_t group_base;     /* subtracted from group values                  */
  uint32_t pad;
} ysql_star_dim;
#define YSQL_MAX_DIMS 4
#define YSQL_MAX_GROUPS 16384
int ysql_star_join_group_sum(const ysql_star_dim *dims, int ndims, const int32_t *m1,
                             const int32_t *m2, uint64_t n, int64_t *sums, uint64_t *counts,
                             uint32_t n_groups, ysql_stream_t stream);
/* nq (1..8) star-join queries over the same facts and dimensions that differ
 * only in their dimension filters, in one read of the facts: bit q of
 * dims[d].pass[row] says whether query q keeps that dimension row (a NULL
 * pass keeps every row for every query).  Query q folds into
 * sums[q * n_groups + g] and counts[q * n_groups + g]; each query's result
 * equals its own ysql_star_join_group_sum.  No counterpart in the reference
 * (it runs one query at a time); the serving loop batches concurrent star
 * joins with it.                                                          */
int ysql_star_join_group_sum_multi(const ysql_star_dim *dims, int ndims, int nq, const int32_t *m1,
                                   const int32_t *m2, uint64_t n, int64_t *sums, uint64_t *counts,
                                   uint32_t n_groups, ysql_stream_t stream);

/* ======================================================================
 * Batched ragged jobs (no reference counterpart: the reference launched
 * every operator of every batch separately, one after another).
 * ysql_batch_run is ONE persistent launch (captured with its reset kernel
 * in a hipGraph): a fixed grid of workgroups takes tasks from device-side
 * job lists and a phase queue until every job is done.  Filters,
 * aggregates and sorts up to their split sizes run whole in one workgroup;
 * larger sorts run as digit-count and scatter phases over chunks of
 * persist_chunk rows, larger filters as count and compaction phases, larger
 * aggregates as chunk tasks folded by the last one; each phase is pushed by
 * the workgroup that completes the previous one, so no workgroup ever waits
 * for another.
 * ysql_batch_run_graph replays another form: the jobs split over K launch
 * chains (largest first onto the least-loaded chain) beside segmented
 * one-workgroup-per-job kernels, captured as one hipGraph.
 * ysql_batch_run_plain issues the per-call entry points job after job on one
 * stream.  Sorts and compactions are stable (unique results) and every
 * aggregate folds the same canonical tiles in the same order, so all three
 * produce the same bits.
 *   SORT:   stable argsort of keys (UKEY32/UKEY64): keys_out = sorted keys,
 *           perm_out[pos] = row.
 *   FILTER: rows with vals[i] < threshold; keys_out / vals_out = selected
 *           (key, value) in row order; *count_out = count.
 *   AGG:    rows with vals[i] < threshold; accs_out[0] = SUM(vals) (its
 *           count is COUNT, d / count is AVG), accs_out[1] = MIN(keys),
 *           accs_out[2] = MAX(keys) (keys read as signed INT / DATETIME).
 * The batch object owns the queues, job tables and every workspace (device
 * memory is allocated by ysql_batch_create only); inputs and outputs stay
 * the caller's and must outlive the batch.  Runs of one batch must not
 * overlap in time.                                                        */
enum { YSQL_JOB_SORT = 0, YSQL_JOB_FILTER = 1, YSQL_JOB_AGG = 2 };
typedef struct ysql_job {
  int kind;
  int key_dtype; /* YSQL_UKEY32 or YSQL_UKEY64 */
  uint64_t n;
  const void *keys;
  const float *vals;
  float threshold;
  uint32_t pad;
  void *keys_out;
  uint32_t *perm_out;
  float *vals_out;
  uint64_t *count_out;
  ysql_acc *accs_out;
} ysql_job;
typedef struct ysql_batch ysql_batch;
/* Tuning of the batched executor (ysql_batch_tuning_default fills it). */
typedef struct ysql_batch_tuning {
  int chains;         /* independent launch chains in the graph (1..64) */
  int seg_sort_max;   /* sort jobs up to this many rows run one per workgroup
                         (default 2^19) */
  int seg_filter_max; /* filter jobs likewise (default 2^21; 0 = every job gets a chain) */
  int seg_agg_max;    /* aggregate jobs likewise (default 2^22) */
  /* digit-pass tile variant of a sort job (see ysql_radix_sort_variant) by
   * key width (0: UKEY32, 1: UKEY64) and size bucket floor(log2(n));
   * -1 = the library default.  Filled by an autotuner. */
  signed char sort_variant[2][32];
  /* persistent launch (ysql_batch_run) */
  int persist_split_min;        /* sorts above this many rows run as chunk phases (default 2^18) */
  int persist_chunk;            /* rows per chunk, rounded up to 4096 (default 2^16) */
  int persist_jobs_first_every; /* every k-th worker takes whole jobs before phase tasks
                                   (default 4; 0 = phase tasks first everywhere) */
  int persist_idle_limit;       /* empty rounds (~1 us each) before an idle worker leaves */
  int persist_workers_per_cu;   /* 0 = as many as fit (occupancy) */
  int persist_split_filter_min; /* filters above this many rows run as count + compact chunk
                                   phases (default 2^20) */
  int persist_split_agg_min;    /* aggregates above this many rows run as chunk tasks (default 2^20) */
} ysql_batch_tuning;
void ysql_batch_tuning_default(ysql_batch_tuning *t);
/* Validates the jobs, allocates every workspace and captures the graph.
 * Inputs and outputs must not move while the batch exists. */
int ysql_batch_create(const ysql_job *jobs, int njobs, const ysql_batch_tuning *tuning,
                      ysql_batch **out);
/* The persistent launch on `stream` (one host call, one graph replay). */
int ysql_batch_run(ysql_batch *b, ysql_stream_t stream);
/* The launch-chain graph on `stream`. */
int ysql_batch_run_graph(ysql_batch *b, ysql_stream_t stream);
/* After a ysql_batch_run on `stream`: synchronises it and reports how many
 * jobs with rows the persistent launch completed, of how many
 * (*jobs_done == *jobs_total when it finished them all). */
int ysql_batch_check(ysql_batch *b, ysql_stream_t stream, uint32_t *jobs_done, uint32_t *jobs_total);
/* Workgroups of the persistent launch. */
int ysql_batch_workers(const ysql_batch *b, uint32_t *workers);
/* The same jobs as a per-job launch loop on `stream`. */
int ysql_batch_run_plain(ysql_batch *b, ysql_stream_t stream);
/* Job count, chain count and workspace bytes of a batch. */
int ysql_batch_info(const ysql_batch *b, uint64_t *tasks, uint32_t *grid, uint64_t *ws_bytes);
void ysql_batch_destroy(ysql_batch *b);

/* ======================================================================
 * Utilities                                                              */
/* Library version string (static storage). */
const char *ysql_version(void);
/* Last HIP error text for this thread (static thread-local storage). */
const char *ysql_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* YSQL_KERNELS_H */
