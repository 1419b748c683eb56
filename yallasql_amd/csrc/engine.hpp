// engine.hpp — host engine above the kernel ABI.
//
// Mirrors the reference's data contract and operator interface:
//   DataType / Column / Batch  (include/enums/data_type.hpp:2-7,
//                               include/db/column.hpp:7-25, include/batch.hpp:21-145)
//   CacheManager               (include/engine/cache_manager.hpp:11-89), here the
//                               HBM residency manager with a byte budget
//   Operator::init / next(CacheManager&) -> BatchID, 0 = exhausted
//                              (include/engine/operators/operator.hpp:44-49)
//   Expression::evaluate       (include/engine/operators/expressions/expression.hpp:31-67)
// All device work of one context runs on its HIP stream; host
// synchronisation happens only where a row count must reach the host (filter
// and join output sizes), as in the reference.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "ysql_kernels.h"

namespace YallaSQL {

enum class DataType : uint8_t { INT = 0, FLOAT, DATETIME, STRING };

inline unsigned getDataTypeNumBytes(DataType t) {
  switch (t) {
    case DataType::INT: return 4;
    case DataType::FLOAT: return 4;
    case DataType::DATETIME: return 8;
    case DataType::STRING: return YSQL_STRING_BYTES;
  }
  return 0;
}
inline int abiType(DataType t) { return (int)t; }  // YSQL_INT.. share the numbering

struct EngineError : std::runtime_error {
  int code;
  EngineError(const std::string &m, int c = YSQL_EINVAL) : std::runtime_error(m), code(c) {}
};

void hipCheck(hipError_t e, const char *what);
void abiCheck(int rc, const char *what);

struct Column {
  std::string name;
  DataType type;
  unsigned bytes;
  Column(std::string n, DataType t) : name(std::move(n)), type(t), bytes(getDataTypeNumBytes(t)) {}
};
using ColumnPtr = std::shared_ptr<Column>;

// Device allocation on a stream (stream-ordered allocator), freed on the
// same stream when the last owner drops it.
// Device memory of the engine comes from one stream-ordered pool per device
// that keeps freed memory for reuse (a plain hipMallocAsync from the default
// pool maps new memory after each release, tens of microseconds per call).
// The pool is trimmed when the last ExecContext of the process is destroyed.
hipMemPool_t enginePool();

// Pinned host memory for query results, reused across queries: taking a
// block reuses a free one of at least the size (and at most twice it) or
// pins a new one; up to 512 MB of free blocks stay pinned.  Thread-safe.
struct PinnedBlock {
  void *ptr = nullptr;
  size_t bytes = 0;
};
PinnedBlock pinnedTake(size_t bytes);
void pinnedGive(PinnedBlock b);

struct DeviceBuffer {
  void *ptr = nullptr;
  size_t bytes = 0;
  hipStream_t stream = nullptr;
  bool owned = true;
  DeviceBuffer(size_t n, hipStream_t s);
  DeviceBuffer(void *borrowed, size_t n) : ptr(borrowed), bytes(n), owned(false) {}
  ~DeviceBuffer();
  DeviceBuffer(const DeviceBuffer &) = delete;
  DeviceBuffer &operator=(const DeviceBuffer &) = delete;
};
using BufPtr = std::shared_ptr<DeviceBuffer>;

// A batch: one device buffer per column (batchSize x bytes) and one null
// byte per row per column.  Buffers are shared, so a batch can be a zero-copy
// view of a table or of another batch (the reference copied instead,
// expressions/bound_ref.hpp:30-37).
class Batch {
 public:
  std::vector<ColumnPtr> columns;
  std::vector<BufPtr> data;     // data[c]->ptr + offset[c]
  std::vector<BufPtr> nulls;    // may hold nullptr (no NULLs in that column)
  std::vector<size_t> offset;   // row offset into the shared buffers
  size_t batchSize = 0;
  hipStream_t stream = nullptr;

  size_t numCols() const { return columns.size(); }
  void *col(size_t c) const {
    return static_cast<char *>(data[c]->ptr) + offset[c] * columns[c]->bytes;
  }
  uint8_t *nul(size_t c) const {
    return nulls[c] ? static_cast<uint8_t *>(nulls[c]->ptr) + offset[c] : nullptr;
  }
  size_t totalBytes() const;
  // Rows [start, start + rows) as a view sharing this batch's buffers.
  std::unique_ptr<Batch> slice(size_t start, size_t rows) const;
};

using BatchID = unsigned;

// BatchID -> Batch, and the batches' residency: HBM, pinned host memory, or
// a spill file (the reference's GPU -> CPU -> FS tiers, cache_manager.hpp:
// 11-89, batch.cu:6-108, with budgets sized for 288 GB of HBM instead of its
// 1 GiB / 2 GiB).  When a batch enters and the HBM budget would be exceeded,
// the least recently used batches are copied to pinned host memory on the
// manager's own copy stream (after the compute stream's work on them, so the
// copies overlap later kernels) and their device buffers are released once
// the copies are done; when the host budget is exceeded the least recently
// used host batches go to files in the spill directory.  Taking a batch
// (getBatch / refBatch) brings it back into HBM on the compute stream first.
// Budgets of 0 are unlimited; with no spill directory the host tier is
// unlimited.  A batch larger than what the budget leaves once every other
// batch has moved out stays in HBM over the budget (it is already there).
class CacheManager {
 public:
  explicit CacheManager(size_t hbm_budget, size_t host_budget = 0, std::string spill_dir = std::string());
  ~CacheManager();
  CacheManager(const CacheManager &) = delete;
  CacheManager &operator=(const CacheManager &) = delete;
  void setComputeStream(hipStream_t s) { compute_ = s; }
  void configure(size_t host_budget, std::string spill_dir);
  BatchID putBatch(std::unique_ptr<Batch> b);
  std::unique_ptr<Batch> getBatch(BatchID id);
  Batch &refBatch(BatchID id);
  size_t bytesInUse() const { return dev_bytes_; }  // in HBM
  size_t peakBytes() const { return peak_; }
  size_t budget() const { return budget_; }
  struct Stats {
    size_t hbm_bytes, host_bytes, file_bytes, peak_hbm_bytes;
    uint64_t evictions, spills, restores_from_host, restores_from_file;
  };
  Stats stats() const;
  void clear();  // drops every batch (device buffers freed on their streams)

 private:
  enum Tier { HBM, HOST, FILE_TIER };
  struct Entry {
    std::unique_ptr<Batch> batch;
    Tier tier = HBM;
    uint64_t last_use = 0;
    size_t bytes = 0;
    std::vector<void *> host_cols, host_nulls;  // HOST tier (nullptr: no null column)
    hipEvent_t copied = nullptr;                // the copies into the host buffers are done
    std::string file;                           // FILE tier
  };
  void makeRoom(size_t bytes, BatchID keep);
  void evictToHost(Entry &e);
  void spillToFile(Entry &e, BatchID id);
  void restore(Entry &e);
  void reapHost();
  Entry &find(BatchID id);
  std::unordered_map<BatchID, Entry> cache_;
  BatchID next_ = 1;
  uint64_t tick_ = 0;
  size_t dev_bytes_ = 0, host_bytes_ = 0, file_bytes_ = 0, peak_ = 0;
  size_t budget_ = 0, host_budget_ = 0;
  std::string spill_dir_;
  hipStream_t compute_ = nullptr, copy_ = nullptr;
  uint64_t evictions_ = 0, spills_ = 0, from_host_ = 0, from_file_ = 0;
  // pinned host buffers released once the copies reading them are done
  std::vector<std::pair<hipEvent_t, std::vector<void *>>> host_retired_;
};

// Per-context execution state: the stream, a growable scratch workspace for
// kernel calls, and small pinned host slots for device-to-host counts.
class ExecContext {
 public:
  ExecContext(size_t budget, hipStream_t stream);
  ~ExecContext();
  hipStream_t stream() const { return stream_; }
  CacheManager &cache() { return cache_; }
  void *workspace(size_t bytes);   // valid until the next call
  uint64_t *deviceCounter();        // one device uint64
  uint64_t readCounter();           // synchronises the stream
  void sync();
  // Expression fusion: predicate and projection trees run as one fused
  // program per batch (on by default); off runs them node by node.
  bool fusion() const { return fusion_; }
  void setFusion(bool on) { fusion_ = on; }

 private:
  bool fusion_ = true;
  hipStream_t stream_ = nullptr;
  bool own_stream_ = false;
  CacheManager cache_;
  void *ws_ = nullptr;
  size_t ws_bytes_ = 0;
  uint64_t *dcount_ = nullptr;
  uint64_t *hcount_ = nullptr;
};

// In-HBM table.
class Table {
 public:
  std::string name;
  std::vector<ColumnPtr> columns;
  std::vector<BufPtr> data, nulls;
  size_t rows = 0;
};

// ---------------------------------------------------------------- expressions
struct ExprValue {
  BufPtr buf;          // device data
  size_t offset = 0;   // rows into buf (views of batch columns)
  BufPtr nullbuf;      // device null bytes or nullptr
  size_t rows = 0;     // 1 for a scalar
  DataType type = DataType::INT;
  bool scalar = false;
  bool mask = false;   // uint8 predicate result
  void *ptr() const { return static_cast<char *>(buf->ptr) + offset * (mask ? 1 : getDataTypeNumBytes(type)); }
  uint8_t *nul() const { return nullbuf ? static_cast<uint8_t *>(nullbuf->ptr) + offset : nullptr; }
};

class Expression {
 public:
  virtual ~Expression() = default;
  // Result type given the input schema.
  virtual DataType returnType(const std::vector<ColumnPtr> &in) const = 0;
  virtual bool isMask() const { return false; }
  virtual ExprValue evaluate(const Batch &b, ExecContext &ctx) = 0;
  virtual std::string name() const = 0;
};
using ExprPtr = std::unique_ptr<Expression>;

class BoundRefExpression : public Expression {
 public:
  explicit BoundRefExpression(int idx) : idx(idx) {}
  int idx;
  DataType returnType(const std::vector<ColumnPtr> &in) const override;
  ExprValue evaluate(const Batch &b, ExecContext &ctx) override;
  std::string name() const override { return "#" + std::to_string(idx); }
};

class BoundValueExpression : public Expression {
 public:
  BoundValueExpression(DataType t, int64_t iv, float fv) : type(t), ival(iv), fval(fv) {}
  DataType type;
  int64_t ival;
  float fval;
  BufPtr dev;  // 1-element device buffer, created on first use
  DataType returnType(const std::vector<ColumnPtr> &) const override { return type; }
  ExprValue evaluate(const Batch &b, ExecContext &ctx) override;
  std::string name() const override { return "const"; }
};

class ComparisonExpression : public Expression {
 public:
  ComparisonExpression(int cmp, ExprPtr l, ExprPtr r) : cmp(cmp), lhs(std::move(l)), rhs(std::move(r)) {}
  int cmp;
  ExprPtr lhs, rhs;
  DataType returnType(const std::vector<ColumnPtr> &) const override { return DataType::INT; }
  bool isMask() const override { return true; }
  ExprValue evaluate(const Batch &b, ExecContext &ctx) override;
  std::string name() const override { return "cmp"; }
};

class ConjunctionExpression : public Expression {
 public:
  ConjunctionExpression(int op, ExprPtr l, ExprPtr r) : op(op), lhs(std::move(l)), rhs(std::move(r)) {}
  int op;
  ExprPtr lhs, rhs;
  DataType returnType(const std::vector<ColumnPtr> &) const override { return DataType::INT; }
  bool isMask() const override { return true; }
  ExprValue evaluate(const Batch &b, ExecContext &ctx) override;
  std::string name() const override { return "logic"; }
};

class BoundFuncExpression : public Expression {
 public:
  BoundFuncExpression(int op, ExprPtr l, ExprPtr r) : op(op), lhs(std::move(l)), rhs(std::move(r)) {}
  int op;
  ExprPtr lhs, rhs;
  DataType returnType(const std::vector<ColumnPtr> &in) const override { return lhs->returnType(in); }
  ExprValue evaluate(const Batch &b, ExecContext &ctx) override;
  std::string name() const override { return "func"; }
};

class CastExpression : public Expression {
 public:
  CastExpression(DataType t, ExprPtr c) : type(t), child(std::move(c)) {}
  DataType type;
  ExprPtr child;
  DataType returnType(const std::vector<ColumnPtr> &) const override { return type; }
  ExprValue evaluate(const Batch &b, ExecContext &ctx) override;
  std::string name() const override { return "cast"; }
};

// ---------------------------------------------------------------- operators
class Operator {
 public:
  std::vector<ColumnPtr> columns;  // output schema
  bool isFinished = false;
  explicit Operator(ExecContext &ctx) : ctx(ctx) {}
  virtual ~Operator() = default;
  virtual void init() {
    for (auto &c : children) c->init();
  }
  virtual BatchID next(CacheManager &cache) = 0;
  std::vector<std::unique_ptr<Operator>> children;
  ExecContext &ctx;
};
using OpPtr = std::unique_ptr<Operator>;

class TableScanOperator : public Operator {
 public:
  TableScanOperator(ExecContext &ctx, std::shared_ptr<Table> t, size_t batch_rows);
  BatchID next(CacheManager &cache) override;

 private:
  std::shared_ptr<Table> table_;
  size_t batch_rows_, pos_ = 0;
};

class FilterOperator : public Operator {
 public:
  FilterOperator(ExecContext &ctx, OpPtr child, ExprPtr pred);
  BatchID next(CacheManager &cache) override;

 private:
  ExprPtr pred_;
  // the predicate as one fused program inside the compaction; false: not fusible
  bool fusedFilter(const Batch &in, const std::vector<ysql_col> &cols, size_t n, size_t wsb);
};

class ProjectionOperator : public Operator {
 public:
  ProjectionOperator(ExecContext &ctx, OpPtr child, std::vector<ExprPtr> exprs);
  BatchID next(CacheManager &cache) override;

 private:
  std::vector<ExprPtr> exprs_;
  void fuseExprs(const Batch &in, std::vector<ExprValue> &vals, std::vector<char> &done);
};

struct OrderKey {
  int col;
  bool desc;
};

class OrderOperator : public Operator {
 public:
  OrderOperator(ExecContext &ctx, OpPtr child, std::vector<OrderKey> keys);
  BatchID next(CacheManager &cache) override;

 private:
  void firstPass(CacheManager &cache);
  std::unique_ptr<Batch> sortBatch(const Batch &b);
  BatchID mergeRuns(BatchID a, BatchID b, CacheManager &cache);
  BufPtr compositeKeys(const Batch &b);
  void sortStep(const Batch &b, const OrderKey &key, BufPtr &perm, bool first, void *ws, size_t wsb);
  std::vector<OrderKey> keys_;
  int key_bits_ = 0;
  bool wide_ = false;     // composite key is 64 bits
  bool by_rows_ = false;  // string keys or keys wider than 64 bits: column by column
  bool done_first_ = false;
  std::vector<BatchID> runs_;     // sorted runs, held by the cache (may leave HBM)
  std::shared_ptr<Batch> final_;  // the merged result being emitted
  size_t max_batch_ = 0, emit_pos_ = 0;
};

struct AggSpec {
  int agg;
  int col;
};
struct GroupSpec {
  int col;
  int32_t base;
  uint32_t card;
};

class AggregateOperator : public Operator {
 public:
  AggregateOperator(ExecContext &ctx, OpPtr child, std::vector<AggSpec> aggs, std::vector<GroupSpec> groups);
  BatchID next(CacheManager &cache) override;

 private:
  std::vector<AggSpec> aggs_;
  std::vector<GroupSpec> groups_;
  uint32_t ngroups_ = 1;
};

class JoinOperator : public Operator {
 public:
  JoinOperator(ExecContext &ctx, OpPtr left, OpPtr right, int lcol, int rcol, int cmp);
  BatchID next(CacheManager &cache) override;

 private:
  void cacheLeftChild(CacheManager &cache);
  int lcol_, rcol_, cmp_;
  bool left_cached_ = false;
  std::vector<std::shared_ptr<Batch>> left_;
  std::shared_ptr<Batch> right_;
  size_t li_ = 0;
  size_t cap_ = 1 << 20;
};

class CrossProductOperator : public Operator {
 public:
  CrossProductOperator(ExecContext &ctx, OpPtr left, OpPtr right);
  BatchID next(CacheManager &cache) override;

 private:
  bool left_cached_ = false;
  std::vector<std::shared_ptr<Batch>> left_;
  std::shared_ptr<Batch> right_;
  size_t li_ = 0, roff_ = 0, rchunk_ = 0;
};

// Copies every batch the root produces to host memory.
// A query's result in host memory: every column's data and null bytes
// (1 = NULL, zeros for a column without nulls) in one pinned block.
struct HostResult {
  std::vector<ColumnPtr> columns;
  std::vector<uint8_t *> data, nulls;  // into block; null when rows == 0
  size_t rows = 0, batches = 0;
  PinnedBlock block;
  HostResult() = default;
  HostResult(const HostResult &) = delete;
  HostResult &operator=(const HostResult &) = delete;
  HostResult(HostResult &&o) noexcept
      : columns(std::move(o.columns)), data(std::move(o.data)), nulls(std::move(o.nulls)), rows(o.rows),
        batches(o.batches), block(o.block) {
    o.block = PinnedBlock{};
  }
  ~HostResult() { pinnedGive(block); }
};
HostResult executeToHost(Operator &root, ExecContext &ctx);

}  // namespace YallaSQL
