/*
 * ysql_engine.h — C API of the host engine above the kernel ABI.
 *
 * The engine keeps the reference's pull-based design: every operator
 * implements init() / next(CacheManager&) -> BatchID (0 = exhausted), over
 * columnar Batches of device buffers with one null byte per row (reference
 * include/engine/operators/operator.hpp:44-49, include/batch.hpp:21-145).
 * The reference built operator trees from DuckDB logical plans
 * (src/engine/operators/operator.cpp:14-43); DuckDB is not part of this
 * build, so trees are built with the calls below.
 *
 * Objects: a context (one HIP stream, an HBM budget and a CacheManager),
 * tables resident in device memory, expression nodes and plan nodes.  A plan
 * node owns its children and expressions; destroy the root only.  Functions
 * returning pointers return NULL on error, int functions a status (0 = ok);
 * ysql_engine_last_error() gives the reason.  A context is used by one host
 * thread at a time.
 */
#ifndef YSQL_ENGINE_H
#define YSQL_ENGINE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ysql_ctx ysql_ctx;
typedef struct ysql_table ysql_table;
typedef struct ysql_expr ysql_expr;
typedef struct ysql_plan ysql_plan;
typedef struct ysql_result ysql_result;

/* Context: a HIP stream of its own (or `stream` if not NULL) and an HBM
 * budget for cached batches (bytes; 0 = unlimited).                       */
ysql_ctx *ysql_ctx_create(uint64_t hbm_budget_bytes, void *stream);
void ysql_ctx_destroy(ysql_ctx *ctx);
/* Peak bytes of batches held in HBM by the context's CacheManager so far. */
uint64_t ysql_ctx_peak_bytes(const ysql_ctx *ctx);
/* Expression fusion (on by default): a filter predicate or projection
 * tree runs as one fused program per batch (ysql_expr_filter_compact /
 * ysql_expr_eval, include/ysql_kernels.h) instead of one kernel and one
 * intermediate column per node (the reference's evaluate() chain,
 * include/engine/operators/expressions/).  on = 0 runs the nodes one by
 * one; both give the same columns.  A tree the fused program cannot express
 * (strings, mixed operand types) always runs node by node.               */
int ysql_ctx_set_fusion(ysql_ctx *ctx, int on);
/* Residency tiers below HBM (the reference's CPU and FS caches,
 * cache_manager.hpp:67-88): when the HBM budget would be exceeded, least
 * recently used cached batches (ORDER BY's sorted runs among them) move to
 * pinned host memory, copied on a side stream; past `host_budget_bytes`
 * (0 = unlimited) the least recently used host batches go to files in
 * `spill_dir` (NULL: no file tier).  Batches come back into HBM when an
 * operator takes them.                                                   */
int ysql_ctx_set_spill(ysql_ctx *ctx, uint64_t host_budget_bytes, const char *spill_dir);
typedef struct ysql_cache_stats {
  uint64_t hbm_bytes, host_bytes, file_bytes, peak_hbm_bytes;
  uint64_t evictions, spills, restores_from_host, restores_from_file;
} ysql_cache_stats;
int ysql_ctx_cache_stats(const ysql_ctx *ctx, ysql_cache_stats *out);

/* Tables.  add_column copies `rows` host values (and optional host null
 * bytes) to the device; add_device_column borrows device buffers that must
 * outlive the table.  dtype: YSQL_INT / YSQL_FLOAT / YSQL_DATETIME /
 * YSQL_STRING (256-byte rows).                                             */
ysql_table *ysql_table_create(ysql_ctx *ctx, const char *name);
int ysql_table_add_column(ysql_table *t, const char *name, int dtype, const void *host_data,
                          const uint8_t *host_nulls, uint64_t rows);
int ysql_table_add_device_column(ysql_table *t, const char *name, int dtype, void *dev_data,
                                 uint8_t *dev_nulls, uint64_t rows);
uint64_t ysql_table_rows(const ysql_table *t);
void ysql_table_destroy(ysql_table *t);

/* Expressions (BoundRef, BoundValue, Comparison, Conjunction, BoundFunc,
 * Cast of the reference, include/engine/operators/expressions/).         */
ysql_expr *ysql_expr_ref(int column);
ysql_expr *ysql_expr_i32(int32_t v);
ysql_expr *ysql_expr_f32(float v);
ysql_expr *ysql_expr_i64(int64_t v);
ysql_expr *ysql_expr_cmp(int cmp, ysql_expr *lhs, ysql_expr *rhs);
ysql_expr *ysql_expr_logic(int op, ysql_expr *lhs, ysql_expr *rhs);
ysql_expr *ysql_expr_arith(int op, ysql_expr *lhs, ysql_expr *rhs);
ysql_expr *ysql_expr_cast(int dtype, ysql_expr *e);
void ysql_expr_destroy(ysql_expr *e); /* only for expressions not given to a plan */

/* Plan nodes. */
ysql_plan *ysql_plan_scan(ysql_ctx *ctx, ysql_table *t, uint64_t batch_rows);
ysql_plan *ysql_plan_filter(ysql_plan *child, ysql_expr *predicate);
ysql_plan *ysql_plan_project(ysql_plan *child, ysql_expr **exprs, int nexprs);
/* ORDER BY cols[0] (desc[0]), cols[1] ...; rows equal on every key keep
 * their input order (stable).  The concatenated encoded keys must fit 64
 * bits.  One sorted run per child batch, then pairwise merge-path levels. */
ysql_plan *ysql_plan_order(ysql_plan *child, const int *cols, const int *desc, int nkeys);
/* Aggregates aggs[i] (YSQL_AGG_*) over cols[i] (ignored for COUNT_STAR).
 * ngroups > 0: GROUP BY int32 columns gcols with values in
 * [gbase, gbase + gcard); output rows = non-empty groups, group columns
 * first, ordered by group id.                                              */
ysql_plan *ysql_plan_aggregate(ysql_plan *child, const int *aggs, const int *cols, int naggs,
                               const int *gcols, const int32_t *gbase, const uint32_t *gcard,
                               int ngroups);
/* Inner join on left[lcol] <cmp> right[rcol]; output = left cols ++ right cols. */
ysql_plan *ysql_plan_join(ysql_plan *left, ysql_plan *right, int lcol, int rcol, int cmp);
ysql_plan *ysql_plan_cross(ysql_plan *left, ysql_plan *right);
void ysql_plan_destroy(ysql_plan *root);

/* Pulls the root's next() until it returns 0 and copies every batch to host
 * memory.  `out` receives a result to free with ysql_result_destroy.      */
int ysql_plan_execute(ysql_plan *root, ysql_result **out);
/* Pulls the root's next() until it returns 0 and writes every batch to the
 * CSV file `path` as it comes (ExecutorEngine::execute with its CsvWriter,
 * src/engine/executor_engine.cpp:17-58; the format is ysql_csv_writer's,
 * include/ysql_csv.h).  *rows_out (may be NULL) = rows written. */
int ysql_plan_execute_csv(ysql_plan *root, const char *path, uint64_t *rows_out);
int ysql_result_ncols(const ysql_result *r);
uint64_t ysql_result_rows(const ysql_result *r);
uint64_t ysql_result_batches(const ysql_result *r);
int ysql_result_dtype(const ysql_result *r, int col);
const void *ysql_result_data(const ysql_result *r, int col);
const uint8_t *ysql_result_nulls(const ysql_result *r, int col);
void ysql_result_destroy(ysql_result *r);

const char *ysql_engine_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* YSQL_ENGINE_H */
