// engine_capi.cpp — C API of the host engine (include/ysql_engine.h).
#include <cstring>

#include "engine.hpp"
#include "ysql_csv.h"
#include "ysql_engine.h"

using namespace YallaSQL;

struct ysql_ctx {
  ExecContext ctx;
  ysql_ctx(uint64_t budget, hipStream_t s) : ctx(budget, s) {}
};
struct ysql_table {
  ysql_ctx *owner;
  std::shared_ptr<Table> t;
};
struct ysql_expr {
  ExprPtr e;
};
struct ysql_plan {
  ysql_ctx *owner;
  OpPtr op;
};
struct ysql_result {
  HostResult r;
};

static thread_local std::string g_err;

template <typename F>
static auto guard(F f, decltype(f()) fail) -> decltype(f()) {
  try {
    return f();
  } catch (const std::exception &e) {
    g_err = e.what();
    return fail;
  }
}

static bool dtype_ok(int d) { return d >= YSQL_INT && d <= YSQL_STRING; }

extern "C" {

const char *ysql_engine_last_error(void) { return g_err.c_str(); }

ysql_ctx *ysql_ctx_create(uint64_t budget, void *stream) {
  return guard([&]() { return new ysql_ctx(budget, (hipStream_t)stream); }, (ysql_ctx *)nullptr);
}
void ysql_ctx_destroy(ysql_ctx *c) { delete c; }
uint64_t ysql_ctx_peak_bytes(const ysql_ctx *c) { return c ? const_cast<ysql_ctx *>(c)->ctx.cache().peakBytes() : 0; }
int ysql_ctx_set_fusion(ysql_ctx *c, int on) {
  if (!c) return YSQL_EINVAL;
  c->ctx.setFusion(on != 0);
  return YSQL_OK;
}
int ysql_ctx_set_spill(ysql_ctx *c, uint64_t host_budget, const char *dir) {
  if (!c) return YSQL_EINVAL;
  c->ctx.cache().configure(host_budget, dir ? std::string(dir) : std::string());
  return YSQL_OK;
}
int ysql_ctx_cache_stats(const ysql_ctx *c, ysql_cache_stats *out) {
  if (!c || !out) return YSQL_EINVAL;
  const auto st = const_cast<ysql_ctx *>(c)->ctx.cache().stats();
  *out = ysql_cache_stats{st.hbm_bytes, st.host_bytes, st.file_bytes, st.peak_hbm_bytes,
                          st.evictions, st.spills, st.restores_from_host, st.restores_from_file};
  return YSQL_OK;
}

ysql_table *ysql_table_create(ysql_ctx *ctx, const char *name) {
  if (!ctx) {
    g_err = "null context";
    return nullptr;
  }
  auto *t = new ysql_table{ctx, std::make_shared<Table>()};
  t->t->name = name ? name : "";
  return t;
}

static int add_column(ysql_table *t, const char *name, int dtype, uint64_t rows, BufPtr data, BufPtr nulls) {
  if (!t->t->columns.empty() && rows != t->t->rows) throw EngineError("column length differs from the table's");
  t->t->columns.push_back(std::make_shared<Column>(name ? name : "", (DataType)dtype));
  t->t->data.push_back(data);
  t->t->nulls.push_back(nulls);
  t->t->rows = rows;
  return YSQL_OK;
}

int ysql_table_add_column(ysql_table *t, const char *name, int dtype, const void *host_data, const uint8_t *host_nulls,
                          uint64_t rows) {
  if (!t || !dtype_ok(dtype) || (rows && !host_data)) return YSQL_EINVAL;
  return guard(
      [&]() {
        hipStream_t s = t->owner->ctx.stream();
        const size_t eb = getDataTypeNumBytes((DataType)dtype);
        auto d = std::make_shared<DeviceBuffer>(rows * eb, s);
        if (rows) hipCheck(hipMemcpyAsync(d->ptr, host_data, rows * eb, hipMemcpyHostToDevice, s), "hipMemcpyAsync");
        BufPtr nb;
        if (host_nulls) {
          nb = std::make_shared<DeviceBuffer>(rows, s);
          if (rows) hipCheck(hipMemcpyAsync(nb->ptr, host_nulls, rows, hipMemcpyHostToDevice, s), "hipMemcpyAsync");
        }
        t->owner->ctx.sync();
        return add_column(t, name, dtype, rows, d, nb);
      },
      (int)YSQL_EHIP);
}

int ysql_table_add_device_column(ysql_table *t, const char *name, int dtype, void *dev_data, uint8_t *dev_nulls,
                                 uint64_t rows) {
  if (!t || !dtype_ok(dtype) || (rows && !dev_data)) return YSQL_EINVAL;
  return guard(
      [&]() {
        const size_t eb = getDataTypeNumBytes((DataType)dtype);
        auto d = std::make_shared<DeviceBuffer>(dev_data, rows * eb);
        BufPtr nb = dev_nulls ? std::make_shared<DeviceBuffer>(dev_nulls, rows) : nullptr;
        return add_column(t, name, dtype, rows, d, nb);
      },
      (int)YSQL_EINVAL);
}

uint64_t ysql_table_rows(const ysql_table *t) { return t ? t->t->rows : 0; }
void ysql_table_destroy(ysql_table *t) { delete t; }

ysql_expr *ysql_expr_ref(int column) { return new ysql_expr{std::make_unique<BoundRefExpression>(column)}; }
ysql_expr *ysql_expr_i32(int32_t v) { return new ysql_expr{std::make_unique<BoundValueExpression>(DataType::INT, v, 0.f)}; }
ysql_expr *ysql_expr_f32(float v) { return new ysql_expr{std::make_unique<BoundValueExpression>(DataType::FLOAT, 0, v)}; }
ysql_expr *ysql_expr_i64(int64_t v) {
  return new ysql_expr{std::make_unique<BoundValueExpression>(DataType::DATETIME, v, 0.f)};
}

static ExprPtr take(ysql_expr *e) {
  if (!e) throw EngineError("null expression");
  ExprPtr p = std::move(e->e);
  delete e;
  return p;
}

ysql_expr *ysql_expr_cmp(int cmp, ysql_expr *l, ysql_expr *r) {
  return guard(
      [&]() {
        if (cmp < YSQL_LT || cmp > YSQL_NE) throw EngineError("bad comparison");
        return new ysql_expr{std::make_unique<ComparisonExpression>(cmp, take(l), take(r))};
      },
      (ysql_expr *)nullptr);
}
ysql_expr *ysql_expr_logic(int op, ysql_expr *l, ysql_expr *r) {
  return guard(
      [&]() {
        if (op != YSQL_AND && op != YSQL_OR) throw EngineError("bad logic op (OR/AND only)");
        return new ysql_expr{std::make_unique<ConjunctionExpression>(op, take(l), take(r))};
      },
      (ysql_expr *)nullptr);
}
ysql_expr *ysql_expr_arith(int op, ysql_expr *l, ysql_expr *r) {
  return guard(
      [&]() {
        if (op < YSQL_ADD || op > YSQL_REM) throw EngineError("bad arithmetic op");
        return new ysql_expr{std::make_unique<BoundFuncExpression>(op, take(l), take(r))};
      },
      (ysql_expr *)nullptr);
}
ysql_expr *ysql_expr_cast(int dtype, ysql_expr *e) {
  return guard(
      [&]() {
        if (dtype < YSQL_INT || dtype > YSQL_DATETIME) throw EngineError("bad cast type");
        return new ysql_expr{std::make_unique<CastExpression>((DataType)dtype, take(e))};
      },
      (ysql_expr *)nullptr);
}
void ysql_expr_destroy(ysql_expr *e) { delete e; }

static OpPtr takeOp(ysql_plan *p) {
  if (!p) throw EngineError("null plan");
  OpPtr o = std::move(p->op);
  delete p;
  return o;
}

ysql_plan *ysql_plan_scan(ysql_ctx *ctx, ysql_table *t, uint64_t batch_rows) {
  return guard(
      [&]() {
        if (!ctx || !t) throw EngineError("null context or table");
        return new ysql_plan{ctx, std::make_unique<TableScanOperator>(ctx->ctx, t->t, batch_rows)};
      },
      (ysql_plan *)nullptr);
}
ysql_plan *ysql_plan_filter(ysql_plan *child, ysql_expr *pred) {
  return guard(
      [&]() {
        ysql_ctx *o = child->owner;
        return new ysql_plan{o, std::make_unique<FilterOperator>(o->ctx, takeOp(child), take(pred))};
      },
      (ysql_plan *)nullptr);
}
ysql_plan *ysql_plan_project(ysql_plan *child, ysql_expr **exprs, int n) {
  return guard(
      [&]() {
        if (n <= 0 || !exprs) throw EngineError("empty projection");
        ysql_ctx *o = child->owner;
        std::vector<ExprPtr> ex;
        for (int i = 0; i < n; i++) ex.push_back(take(exprs[i]));
        return new ysql_plan{o, std::make_unique<ProjectionOperator>(o->ctx, takeOp(child), std::move(ex))};
      },
      (ysql_plan *)nullptr);
}
ysql_plan *ysql_plan_order(ysql_plan *child, const int *cols, const int *desc, int nkeys) {
  return guard(
      [&]() {
        if (nkeys <= 0 || !cols) throw EngineError("ORDER BY without keys");
        ysql_ctx *o = child->owner;
        std::vector<OrderKey> k;
        for (int i = 0; i < nkeys; i++) k.push_back({cols[i], desc && desc[i] != 0});
        return new ysql_plan{o, std::make_unique<OrderOperator>(o->ctx, takeOp(child), k)};
      },
      (ysql_plan *)nullptr);
}
ysql_plan *ysql_plan_aggregate(ysql_plan *child, const int *aggs, const int *cols, int naggs, const int *gcols,
                               const int32_t *gbase, const uint32_t *gcard, int ngroups) {
  return guard(
      [&]() {
        if (naggs <= 0 || !aggs || !cols) throw EngineError("aggregate without aggregates");
        if (ngroups < 0 || (ngroups > 0 && (!gcols || !gbase || !gcard))) throw EngineError("bad GROUP BY");
        ysql_ctx *o = child->owner;
        std::vector<AggSpec> a;
        for (int i = 0; i < naggs; i++) a.push_back({aggs[i], cols[i]});
        std::vector<GroupSpec> g;
        for (int i = 0; i < ngroups; i++) g.push_back({gcols[i], gbase[i], gcard[i]});
        return new ysql_plan{o, std::make_unique<AggregateOperator>(o->ctx, takeOp(child), a, g)};
      },
      (ysql_plan *)nullptr);
}
ysql_plan *ysql_plan_join(ysql_plan *left, ysql_plan *right, int lcol, int rcol, int cmp) {
  return guard(
      [&]() {
        if (cmp < YSQL_LT || cmp > YSQL_NE) throw EngineError("bad join comparison");
        ysql_ctx *o = left->owner;
        return new ysql_plan{o, std::make_unique<JoinOperator>(o->ctx, takeOp(left), takeOp(right), lcol, rcol, cmp)};
      },
      (ysql_plan *)nullptr);
}
ysql_plan *ysql_plan_cross(ysql_plan *left, ysql_plan *right) {
  return guard(
      [&]() {
        ysql_ctx *o = left->owner;
        return new ysql_plan{o, std::make_unique<CrossProductOperator>(o->ctx, takeOp(left), takeOp(right))};
      },
      (ysql_plan *)nullptr);
}
void ysql_plan_destroy(ysql_plan *p) {
  if (p) {
    ysql_ctx *o = p->owner;
    delete p;
    if (o) (void)hipStreamSynchronize(o->ctx.stream());
  }
}

int ysql_plan_execute(ysql_plan *root, ysql_result **out) {
  if (!root || !out) return YSQL_EINVAL;
  try {
    auto *r = new ysql_result{executeToHost(*root->op, root->owner->ctx)};
    *out = r;
    return YSQL_OK;
  } catch (const EngineError &e) {
    g_err = e.what();
    return e.code ? e.code : YSQL_EINVAL;
  } catch (const std::exception &e) {
    g_err = e.what();
    return YSQL_EINVAL;
  }
}

int ysql_plan_execute_csv(ysql_plan *root, const char *path, uint64_t *rows_out) {
  if (!root || !path) return YSQL_EINVAL;
  ysql_csv_writer *w = nullptr;
  try {
    Operator &op = *root->op;
    ExecContext &ctx = root->owner->ctx;
    const size_t nc = op.columns.size();
    std::vector<std::string> names(nc);
    std::vector<const char *> np(nc);
    std::vector<int> types(nc);
    for (size_t c = 0; c < nc; c++) {
      names[c] = op.columns[c]->name;
      np[c] = names[c].c_str();
      types[c] = (int)op.columns[c]->type;
    }
    int rc = ysql_csv_writer_open(path, (int)nc, np.data(), types.data(), &w);
    if (rc) {
      g_err = ysql_last_error();
      return rc;
    }
    op.init();
    CacheManager &cache = ctx.cache();
    uint64_t rows = 0;
    std::vector<const void *> cols(nc);
    std::vector<const uint8_t *> nulls(nc);
    while (true) {
      const BatchID id = op.next(cache);
      if (!id) break;
      auto b = cache.getBatch(id);
      for (size_t c = 0; c < nc; c++) {
        cols[c] = b->col(c);
        nulls[c] = b->nul(c);
      }
      rc = ysql_csv_writer_write(w, cols.data(), nulls.data(), b->batchSize, ctx.stream());
      if (rc) {
        g_err = ysql_last_error();
        (void)ysql_csv_writer_close(w);
        return rc;
      }
      rows += b->batchSize;
    }
    rc = ysql_csv_writer_close(w);
    w = nullptr;
    if (rows_out) *rows_out = rows;
    return rc;
  } catch (const EngineError &e) {
    g_err = e.what();
    if (w) (void)ysql_csv_writer_close(w);
    return e.code ? e.code : YSQL_EINVAL;
  } catch (const std::exception &e) {
    g_err = e.what();
    if (w) (void)ysql_csv_writer_close(w);
    return YSQL_EINVAL;
  }
}

int ysql_result_ncols(const ysql_result *r) { return r ? (int)r->r.columns.size() : 0; }
uint64_t ysql_result_rows(const ysql_result *r) { return r ? r->r.rows : 0; }
uint64_t ysql_result_batches(const ysql_result *r) { return r ? r->r.batches : 0; }
int ysql_result_dtype(const ysql_result *r, int c) {
  return (r && c >= 0 && (size_t)c < r->r.columns.size()) ? (int)r->r.columns[c]->type : -1;
}
const void *ysql_result_data(const ysql_result *r, int c) {
  return (r && c >= 0 && (size_t)c < r->r.data.size()) ? r->r.data[c] : nullptr;
}
const uint8_t *ysql_result_nulls(const ysql_result *r, int c) {
  return (r && c >= 0 && (size_t)c < r->r.nulls.size()) ? r->r.nulls[c] : nullptr;
}
void ysql_result_destroy(ysql_result *r) { delete r; }

}  // extern "C"
