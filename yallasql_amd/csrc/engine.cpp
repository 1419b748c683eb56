// engine.cpp — operators and expressions of the host engine.
#include "engine.hpp"

#include <algorithm>
#include <climits>
#include <map>
#include <mutex>
#include <cstdio>
#include <cstring>

namespace YallaSQL {

void hipCheck(hipError_t e, const char *what) {
  if (e != hipSuccess) throw EngineError(std::string(what) + ": " + hipGetErrorString(e), YSQL_EHIP);
}
void abiCheck(int rc, const char *what) {
  if (rc != YSQL_OK) throw EngineError(std::string(what) + " returned " + std::to_string(rc) + " (" + ysql_last_error() + ")", rc);
}

// ---------------------------------------------------------------- memory
namespace {
constexpr int kMaxDevices = 64;
std::mutex g_pool_mu;
hipMemPool_t g_pool[kMaxDevices] = {};
int g_live_contexts = 0;

std::mutex g_pin_mu;
constexpr size_t kPinnedCacheBytes = size_t(512) << 20;
constexpr size_t kPinnedGrain = size_t(1) << 20;
// Free pinned blocks by size.  Never destroyed: blocks still cached when the
// process ends are released with it (a static destructor could run after
// the HIP runtime has shut down).
std::multimap<size_t, void *> &pinFree() {
  static auto *m = new std::multimap<size_t, void *>();
  return *m;
}
size_t g_pin_free_bytes = 0;
}  // namespace

hipMemPool_t enginePool() {
  int dev = 0;
  hipCheck(hipGetDevice(&dev), "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) throw EngineError("device index out of range");
  std::lock_guard<std::mutex> g(g_pool_mu);
  if (!g_pool[dev]) {
    hipMemPoolProps props = {};
    props.allocType = hipMemAllocationTypePinned;
    props.handleTypes = hipMemHandleTypeNone;
    props.location.type = hipMemLocationTypeDevice;
    props.location.id = dev;
    hipMemPool_t pool = nullptr;
    hipCheck(hipMemPoolCreate(&pool, &props), "hipMemPoolCreate");
    uint64_t keep = UINT64_MAX;
    hipCheck(hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep), "hipMemPoolSetAttribute");
    g_pool[dev] = pool;
  }
  return g_pool[dev];
}

PinnedBlock pinnedTake(size_t bytes) {
  bytes = (bytes + kPinnedGrain - 1) / kPinnedGrain * kPinnedGrain;
  {
    std::lock_guard<std::mutex> g(g_pin_mu);
    auto &m = pinFree();
    auto it = m.lower_bound(bytes);
    if (it != m.end() && it->first <= 2 * bytes) {
      PinnedBlock b{it->second, it->first};
      g_pin_free_bytes -= it->first;
      m.erase(it);
      return b;
    }
  }
  PinnedBlock b;
  hipCheck(hipHostMalloc(&b.ptr, bytes, hipHostMallocDefault), "hipHostMalloc");
  b.bytes = bytes;
  return b;
}

void pinnedGive(PinnedBlock b) {
  if (!b.ptr) return;
  {
    std::lock_guard<std::mutex> g(g_pin_mu);
    if (g_pin_free_bytes + b.bytes <= kPinnedCacheBytes) {
      pinFree().emplace(b.bytes, b.ptr);
      g_pin_free_bytes += b.bytes;
      return;
    }
  }
  (void)hipHostFree(b.ptr);
}

DeviceBuffer::DeviceBuffer(size_t n, hipStream_t s) : bytes(n), stream(s) {
  hipCheck(hipMallocFromPoolAsync(&ptr, n ? n : 1, enginePool(), s), "hipMallocFromPoolAsync");
}
DeviceBuffer::~DeviceBuffer() {
  if (owned && ptr) (void)hipFreeAsync(ptr, stream);
}

static BufPtr newBuf(size_t n, hipStream_t s) { return std::make_shared<DeviceBuffer>(n, s); }

size_t Batch::totalBytes() const {
  size_t b = 0;
  for (size_t c = 0; c < columns.size(); c++) b += batchSize * (columns[c]->bytes + (nulls[c] ? 1 : 0));
  return b;
}

std::unique_ptr<Batch> Batch::slice(size_t start, size_t rows) const {
  auto v = std::make_unique<Batch>(*this);
  for (auto &o : v->offset) o += start;
  v->batchSize = rows;
  return v;
}

// ---------------------------------------------------------------- cache
CacheManager::CacheManager(size_t hbm_budget, size_t host_budget, std::string spill_dir)
    : budget_(hbm_budget), host_budget_(host_budget), spill_dir_(std::move(spill_dir)) {}

CacheManager::~CacheManager() {
  clear();
  if (copy_) (void)hipStreamDestroy(copy_);
}

void CacheManager::clear() {
  if (copy_) (void)hipStreamSynchronize(copy_);
  for (auto &kv : cache_) {
    Entry &e = kv.second;
    if (e.copied) (void)hipEventSynchronize(e.copied);
    if (e.tier == HOST) {  // (in the file tier the null entries are only markers)
      for (void *p : e.host_cols)
        if (p) (void)hipHostFree(p);
      for (void *p : e.host_nulls)
        if (p) (void)hipHostFree(p);
    }
    if (e.copied) (void)hipEventDestroy(e.copied);
    if (!e.file.empty()) std::remove(e.file.c_str());
  }
  for (auto &r : host_retired_) {
    (void)hipEventSynchronize(r.first);
    for (void *p : r.second) (void)hipHostFree(p);
    (void)hipEventDestroy(r.first);
  }
  cache_.clear();
  host_retired_.clear();
  dev_bytes_ = host_bytes_ = file_bytes_ = 0;
}

void CacheManager::configure(size_t host_budget, std::string spill_dir) {
  host_budget_ = host_budget;
  spill_dir_ = std::move(spill_dir);
}

CacheManager::Stats CacheManager::stats() const {
  return Stats{dev_bytes_, host_bytes_, file_bytes_, peak_, evictions_, spills_, from_host_, from_file_};
}

CacheManager::Entry &CacheManager::find(BatchID id) {
  auto it = cache_.find(id);
  if (it == cache_.end()) throw EngineError("unknown batch id " + std::to_string(id));
  return it->second;
}

// Host buffers whose host-to-device copies have completed are freed.
void CacheManager::reapHost() {
  size_t k = 0;
  for (auto &r : host_retired_) {
    if (hipEventQuery(r.first) == hipSuccess) {
      for (void *p : r.second) (void)hipHostFree(p);
      (void)hipEventDestroy(r.first);
    } else {
      host_retired_[k++] = std::move(r);
    }
  }
  host_retired_.resize(k);
}

// HBM -> pinned host on the copy stream, ordered after the compute stream's
// work so far; device buffers held only by this batch are freed on the copy
// stream after the copies (views shared with other batches stay alive).
void CacheManager::evictToHost(Entry &e) {
  Batch &b = *e.batch;
  if (!copy_) hipCheck(hipStreamCreateWithFlags(&copy_, hipStreamNonBlocking), "hipStreamCreate");
  hipEvent_t produced;
  hipCheck(hipEventCreateWithFlags(&produced, hipEventDisableTiming), "hipEventCreate");
  hipCheck(hipEventRecord(produced, compute_), "hipEventRecord");
  hipCheck(hipStreamWaitEvent(copy_, produced, 0), "hipStreamWaitEvent");
  (void)hipEventDestroy(produced);
  const size_t nc = b.numCols();
  e.host_cols.assign(nc, nullptr);
  e.host_nulls.assign(nc, nullptr);
  for (size_t c = 0; c < nc; c++) {
    const size_t bytes = b.batchSize * b.columns[c]->bytes;
    hipCheck(hipHostMalloc(&e.host_cols[c], bytes ? bytes : 1, hipHostMallocDefault), "hipHostMalloc");
    if (bytes) hipCheck(hipMemcpyAsync(e.host_cols[c], b.col(c), bytes, hipMemcpyDeviceToHost, copy_), "hipMemcpyAsync");
    if (b.nulls[c]) {
      hipCheck(hipHostMalloc(&e.host_nulls[c], b.batchSize ? b.batchSize : 1, hipHostMallocDefault), "hipHostMalloc");
      if (b.batchSize)
        hipCheck(hipMemcpyAsync(e.host_nulls[c], b.nul(c), b.batchSize, hipMemcpyDeviceToHost, copy_),
                 "hipMemcpyAsync");
    }
  }
  hipCheck(hipEventCreateWithFlags(&e.copied, hipEventDisableTiming), "hipEventCreate");
  hipCheck(hipEventRecord(e.copied, copy_), "hipEventRecord");
  for (auto *v : {&b.data, &b.nulls})
    for (auto &buf : *v)
      if (buf && buf.use_count() == 1 && buf->owned) buf->stream = copy_;  // freed after the copies
  b.data.assign(nc, nullptr);
  b.nulls.assign(nc, nullptr);
  b.offset.assign(nc, 0);
  e.tier = HOST;
  dev_bytes_ -= e.bytes;
  host_bytes_ += e.bytes;
  evictions_++;
}

// Pinned host -> one file (columns, then null columns, in order).
void CacheManager::spillToFile(Entry &e, BatchID id) {
  const Batch &b = *e.batch;
  hipCheck(hipEventSynchronize(e.copied), "hipEventSynchronize");
  const std::string path = spill_dir_ + "/ysql_batch_" + std::to_string(reinterpret_cast<uintptr_t>(this)) + "_" +
                           std::to_string(id) + ".bin";
  FILE *f = std::fopen(path.c_str(), "wb");
  if (!f) throw EngineError("cannot create spill file " + path);
  bool ok = true;
  for (size_t c = 0; c < b.numCols(); c++) {
    const size_t bytes = b.batchSize * b.columns[c]->bytes;
    ok = ok && std::fwrite(e.host_cols[c], 1, bytes, f) == bytes;
    if (e.host_nulls[c]) ok = ok && std::fwrite(e.host_nulls[c], 1, b.batchSize, f) == b.batchSize;
  }
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) {
    std::remove(path.c_str());
    throw EngineError("cannot write spill file " + path);
  }
  for (auto &p : e.host_cols) {
    (void)hipHostFree(p);
    p = nullptr;
  }
  for (auto &p : e.host_nulls)
    if (p) {
      (void)hipHostFree(p);
      p = reinterpret_cast<void *>(1);  // marks "this column has null bytes" for the reload
    }
  (void)hipEventDestroy(e.copied);
  e.copied = nullptr;
  e.file = path;
  e.tier = FILE_TIER;
  host_bytes_ -= e.bytes;
  file_bytes_ += e.bytes;
  spills_++;
}

// Back into HBM on the compute stream (from a file through pinned memory).
void CacheManager::restore(Entry &e) {
  Batch &b = *e.batch;
  const size_t nc = b.numCols();
  if (e.tier == FILE_TIER) {
    FILE *f = std::fopen(e.file.c_str(), "rb");
    if (!f) throw EngineError("cannot open spill file " + e.file);
    bool ok = true;
    for (size_t c = 0; c < nc; c++) {
      const size_t bytes = b.batchSize * b.columns[c]->bytes;
      hipCheck(hipHostMalloc(&e.host_cols[c], bytes ? bytes : 1, hipHostMallocDefault), "hipHostMalloc");
      ok = ok && std::fread(e.host_cols[c], 1, bytes, f) == bytes;
      if (e.host_nulls[c]) {
        hipCheck(hipHostMalloc(&e.host_nulls[c], b.batchSize ? b.batchSize : 1, hipHostMallocDefault),
                 "hipHostMalloc");
        ok = ok && std::fread(e.host_nulls[c], 1, b.batchSize, f) == b.batchSize;
      }
    }
    std::fclose(f);
    if (!ok) throw EngineError("short spill file " + e.file);
    std::remove(e.file.c_str());
    e.file.clear();
    file_bytes_ -= e.bytes;
    host_bytes_ += e.bytes;
    from_file_++;
  } else {
    from_host_++;
    hipCheck(hipStreamWaitEvent(compute_, e.copied, 0), "hipStreamWaitEvent");
  }
  std::vector<void *> release;
  for (size_t c = 0; c < nc; c++) {
    const size_t bytes = b.batchSize * b.columns[c]->bytes;
    b.data[c] = std::make_shared<DeviceBuffer>(bytes, compute_);
    if (bytes)
      hipCheck(hipMemcpyAsync(b.data[c]->ptr, e.host_cols[c], bytes, hipMemcpyHostToDevice, compute_),
               "hipMemcpyAsync");
    release.push_back(e.host_cols[c]);
    if (e.host_nulls[c]) {
      b.nulls[c] = std::make_shared<DeviceBuffer>(b.batchSize, compute_);
      if (b.batchSize)
        hipCheck(hipMemcpyAsync(b.nulls[c]->ptr, e.host_nulls[c], b.batchSize, hipMemcpyHostToDevice, compute_),
                 "hipMemcpyAsync");
      release.push_back(e.host_nulls[c]);
    }
    b.offset[c] = 0;
  }
  hipEvent_t done;
  hipCheck(hipEventCreateWithFlags(&done, hipEventDisableTiming), "hipEventCreate");
  hipCheck(hipEventRecord(done, compute_), "hipEventRecord");
  host_retired_.push_back({done, std::move(release)});
  if (e.copied) (void)hipEventDestroy(e.copied);
  e.copied = nullptr;
  e.host_cols.clear();
  e.host_nulls.clear();
  e.tier = HBM;
  host_bytes_ -= e.bytes;
  dev_bytes_ += e.bytes;
  peak_ = std::max(peak_, dev_bytes_);
}

// Evicts least recently used HBM batches (not `keep`) until `bytes` more fit.
void CacheManager::makeRoom(size_t bytes, BatchID keep) {
  reapHost();
  while (budget_ && dev_bytes_ + bytes > budget_) {
    Entry *victim = nullptr;
    for (auto &kv : cache_)
      if (kv.first != keep && kv.second.tier == HBM && kv.second.bytes &&
          (!victim || kv.second.last_use < victim->last_use))
        victim = &kv.second;
    if (!victim) break;  // nothing left to move out: this batch stays in HBM over the budget
    evictToHost(*victim);
  }
  while (!spill_dir_.empty() && host_budget_ && host_bytes_ > host_budget_) {
    BatchID vid = 0;
    Entry *victim = nullptr;
    for (auto &kv : cache_)
      if (kv.second.tier == HOST && (!victim || kv.second.last_use < victim->last_use)) {
        victim = &kv.second;
        vid = kv.first;
      }
    if (!victim) break;
    spillToFile(*victim, vid);
  }
}

BatchID CacheManager::putBatch(std::unique_ptr<Batch> b) {
  if (!b) throw EngineError("Invalid batch provided.");
  const size_t sz = b->totalBytes();
  makeRoom(sz, 0);
  const BatchID id = next_++;
  Entry &e = cache_[id];
  e.batch = std::move(b);
  e.bytes = sz;
  e.last_use = ++tick_;
  dev_bytes_ += sz;
  peak_ = std::max(peak_, dev_bytes_);
  return id;
}

std::unique_ptr<Batch> CacheManager::getBatch(BatchID id) {
  Entry &e = find(id);
  if (e.tier != HBM) {
    makeRoom(e.bytes, id);
    restore(e);
  }
  auto b = std::move(e.batch);
  dev_bytes_ -= e.bytes;
  cache_.erase(id);
  return b;
}

Batch &CacheManager::refBatch(BatchID id) {
  Entry &e = find(id);
  if (e.tier != HBM) {
    makeRoom(e.bytes, id);
    restore(e);
  }
  e.last_use = ++tick_;
  return *e.batch;
}

// ---------------------------------------------------------------- context
ExecContext::ExecContext(size_t budget, hipStream_t stream) : cache_(budget) {
  if (stream) {
    stream_ = stream;
  } else {
    hipCheck(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
    own_stream_ = true;
  }
  cache_.setComputeStream(stream_);
  hipCheck(hipMalloc(&dcount_, 64), "hipMalloc");
  hipCheck(hipHostMalloc(&hcount_, 64, hipHostMallocDefault), "hipHostMalloc");
  (void)enginePool();
  std::lock_guard<std::mutex> g(g_pool_mu);
  g_live_contexts++;
}
ExecContext::~ExecContext() {
  (void)hipStreamSynchronize(stream_);
  cache_.clear();  // cached batches free their buffers on this stream
  (void)hipStreamSynchronize(stream_);
  if (ws_) (void)hipFreeAsync(ws_, stream_);
  (void)hipStreamSynchronize(stream_);
  (void)hipFree(dcount_);
  (void)hipHostFree(hcount_);
  if (own_stream_) (void)hipStreamDestroy(stream_);
  std::lock_guard<std::mutex> g(g_pool_mu);
  if (--g_live_contexts == 0)  // the last context: hand the pools' cached memory back
    for (hipMemPool_t p : g_pool)
      if (p) (void)hipMemPoolTrimTo(p, 0);
}
void *ExecContext::workspace(size_t bytes) {
  if (bytes > ws_bytes_) {
    if (ws_) hipCheck(hipFreeAsync(ws_, stream_), "hipFreeAsync");
    ws_bytes_ = std::max(bytes, ws_bytes_ + ws_bytes_ / 2) + 256;
    hipCheck(hipMallocFromPoolAsync(&ws_, ws_bytes_, enginePool(), stream_), "hipMallocFromPoolAsync");
  }
  return ws_;
}
uint64_t *ExecContext::deviceCounter() { return dcount_; }
uint64_t ExecContext::readCounter() {
  hipCheck(hipMemcpyAsync(hcount_, dcount_, 8, hipMemcpyDeviceToHost, stream_), "hipMemcpyAsync");
  hipCheck(hipStreamSynchronize(stream_), "hipStreamSynchronize");
  return *hcount_;
}
void ExecContext::sync() { hipCheck(hipStreamSynchronize(stream_), "hipStreamSynchronize"); }

// ---------------------------------------------------------------- expressions
DataType BoundRefExpression::returnType(const std::vector<ColumnPtr> &in) const {
  if (idx < 0 || (size_t)idx >= in.size()) throw EngineError("column index out of range: " + std::to_string(idx));
  return in[idx]->type;
}
ExprValue BoundRefExpression::evaluate(const Batch &b, ExecContext &) {
  if (idx < 0 || (size_t)idx >= b.numCols()) throw EngineError("column index out of range: " + std::to_string(idx));
  ExprValue v;
  v.buf = b.data[idx];
  v.offset = b.offset[idx];
  v.nullbuf = b.nulls[idx];
  v.rows = b.batchSize;
  v.type = b.columns[idx]->type;
  return v;
}

ExprValue BoundValueExpression::evaluate(const Batch &, ExecContext &ctx) {
  if (!dev) {
    dev = newBuf(8, ctx.stream());
    const void *src = type == DataType::FLOAT ? (const void *)&fval : (const void *)&ival;
    size_t n = type == DataType::FLOAT || type == DataType::INT ? 4 : 8;
    int32_t i32 = (int32_t)ival;
    if (type == DataType::INT) src = &i32;
    hipCheck(hipMemcpyAsync(dev->ptr, src, n, hipMemcpyHostToDevice, ctx.stream()), "hipMemcpyAsync");
    ctx.sync();  // the host source is a local
  }
  ExprValue v;
  v.buf = dev;
  v.rows = 1;
  v.type = type;
  v.scalar = true;
  return v;
}

// A constant compared or combined with a column of another numeric type is
// converted on the host (the reference relied on DuckDB's literal casts).
static void coerce(Expression *e, DataType want) {
  auto *bv = dynamic_cast<BoundValueExpression *>(e);
  if (!bv || bv->type == want) return;
  if (want == DataType::FLOAT) bv->fval = bv->type == DataType::FLOAT ? bv->fval : (float)bv->ival;
  else bv->ival = bv->type == DataType::FLOAT ? (int64_t)bv->fval : bv->ival;
  bv->type = want;
  bv->dev.reset();
}

// Derived values get their own copy of the source's null bytes, so data and
// nulls share one offset (0) like every other batch column.
static BufPtr copyNulls(const ExprValue &v, ExecContext &ctx) {
  if (!v.nullbuf || v.scalar) return nullptr;
  auto nb = newBuf(v.rows, ctx.stream());
  hipCheck(hipMemcpyAsync(nb->ptr, v.nul(), v.rows, hipMemcpyDeviceToDevice, ctx.stream()), "hipMemcpyAsync");
  return nb;
}

static size_t outRows(const ExprValue &a, const ExprValue &b) {
  if (a.scalar && b.scalar) return 1;
  return a.scalar ? b.rows : a.rows;
}

ExprValue ComparisonExpression::evaluate(const Batch &b, ExecContext &ctx) {
  std::vector<ColumnPtr> schema = b.columns;
  if (dynamic_cast<BoundValueExpression *>(rhs.get())) coerce(rhs.get(), lhs->returnType(schema));
  if (dynamic_cast<BoundValueExpression *>(lhs.get())) coerce(lhs.get(), rhs->returnType(schema));
  ExprValue l = lhs->evaluate(b, ctx), r = rhs->evaluate(b, ctx);
  if (l.type != r.type) throw EngineError("comparison of different types");
  ExprValue out;
  out.rows = outRows(l, r);
  out.mask = true;
  out.scalar = l.scalar && r.scalar;
  out.buf = newBuf(out.rows, ctx.stream());
  out.nullbuf = copyNulls(l.scalar ? r : l, ctx);
  abiCheck(ysql_compare(abiType(l.type), cmp, l.ptr(), l.scalar, r.ptr(), r.scalar, 0, (uint8_t *)out.buf->ptr,
                        out.rows, ctx.stream()),
           "ysql_compare");
  return out;
}

ExprValue ConjunctionExpression::evaluate(const Batch &b, ExecContext &ctx) {
  ExprValue l = lhs->evaluate(b, ctx), r = rhs->evaluate(b, ctx);
  if (!l.mask || !r.mask) throw EngineError("AND/OR of non-boolean operands");
  ExprValue out;
  out.rows = outRows(l, r);
  out.mask = true;
  out.scalar = l.scalar && r.scalar;
  out.buf = newBuf(out.rows, ctx.stream());
  out.nullbuf = copyNulls(l, ctx);  // the left operand's nullset (filters/conjunction.hpp:111)
  abiCheck(ysql_mask_logic(op, (const uint8_t *)l.ptr(), l.scalar, (const uint8_t *)r.ptr(), r.scalar,
                           (uint8_t *)out.buf->ptr, out.rows, ctx.stream()),
           "ysql_mask_logic");
  return out;
}

ExprValue BoundFuncExpression::evaluate(const Batch &b, ExecContext &ctx) {
  std::vector<ColumnPtr> schema = b.columns;
  if (dynamic_cast<BoundValueExpression *>(rhs.get())) coerce(rhs.get(), lhs->returnType(schema));
  if (dynamic_cast<BoundValueExpression *>(lhs.get())) coerce(lhs.get(), rhs->returnType(schema));
  ExprValue l = lhs->evaluate(b, ctx), r = rhs->evaluate(b, ctx);
  if (l.type != r.type) throw EngineError("arithmetic on different types");
  ExprValue out;
  out.rows = outRows(l, r);
  out.type = l.type;
  out.scalar = l.scalar && r.scalar;
  out.buf = newBuf(out.rows * getDataTypeNumBytes(l.type), ctx.stream());
  out.nullbuf = copyNulls(l.scalar ? r : l, ctx);  // bound_func.hpp: nulls of the non-constant operand
  abiCheck(ysql_arith(abiType(l.type), op, l.ptr(), l.scalar, r.ptr(), r.scalar, out.buf->ptr, out.rows,
                      ctx.stream()),
           "ysql_arith");
  return out;
}

ExprValue CastExpression::evaluate(const Batch &b, ExecContext &ctx) {
  ExprValue c = child->evaluate(b, ctx);
  ExprValue out;
  out.rows = c.rows;
  out.type = type;
  out.scalar = c.scalar;
  out.nullbuf = copyNulls(c, ctx);
  out.buf = newBuf(c.rows * getDataTypeNumBytes(type), ctx.stream());
  abiCheck(ysql_cast(c.mask ? YSQL_BOOL : abiType(c.type), abiType(type), c.ptr(), out.buf->ptr, c.rows, ctx.stream()),
           "ysql_cast");
  return out;
}

// ---------------------------------------------------------------- expression fusion
// Compiles predicate and projection trees into one register program per
// batch (ysql_expr_filter_compact / ysql_expr_eval), replacing one kernel
// and one intermediate column per node.  The node rules are those of the
// evaluate() methods above, hence of the reference's expression nodes:
// constants are coerced to the other operand's type; a comparison or
// arithmetic node takes the NULL flags of its non-constant operand (the left
// one when neither is constant), a cast its child's, a conjunction its left
// operand's.  A tree the program cannot express (a string operand, a mask
// used as a number, operands of different types, REM on floats, more than
// 16 registers or 48 instructions) is evaluated node by node, so it gives
// the same result or the same error as before.
namespace {

struct XVal {
  int reg = -1;
  int type = YSQL_INT;
  bool scalar = false;
  bool nullable = false;
};

static void collectCols(Expression *e, std::vector<int> &cols) {
  if (auto *r = dynamic_cast<BoundRefExpression *>(e)) {
    if (std::find(cols.begin(), cols.end(), r->idx) == cols.end()) cols.push_back(r->idx);
  } else if (auto *c = dynamic_cast<ComparisonExpression *>(e)) {
    collectCols(c->lhs.get(), cols);
    collectCols(c->rhs.get(), cols);
  } else if (auto *j = dynamic_cast<ConjunctionExpression *>(e)) {
    collectCols(j->lhs.get(), cols);
    collectCols(j->rhs.get(), cols);
  } else if (auto *f = dynamic_cast<BoundFuncExpression *>(e)) {
    collectCols(f->lhs.get(), cols);
    collectCols(f->rhs.get(), cols);
  } else if (auto *k = dynamic_cast<CastExpression *>(e)) {
    collectCols(k->child.get(), cols);
  }
}

class XProgram {
 public:
  // `cols`: the batch columns the trees read; they become registers 0..n-1.
  XProgram(const Batch &b, const std::vector<int> &cols) : b_(b) {
    for (int c : cols) {
      if (c < 0 || (size_t)c >= b.numCols() || inputs.size() >= YSQL_XMAX_INPUTS) {
        bad_ = true;
        return;
      }
      const DataType t = b.columns[c]->type;
      if (t == DataType::STRING) {
        bad_ = true;
        return;
      }
      colreg_.push_back({c, (int)inputs.size()});
      inputs.push_back(ysql_xinput{b.col(c), b.nul(c), abiType(t), 0});
      used_ |= 1u << (inputs.size() - 1);
    }
  }
  std::vector<ysql_xinsn> insn;
  std::vector<ysql_xinput> inputs;

  // Compiles one tree; false (program unchanged) when it cannot be fused.
  bool add(Expression *e, XVal *out) {
    if (bad_) return false;
    const size_t n0 = insn.size();
    const uint32_t u0 = used_;
    if (!compile(e, out) || out->scalar) {
      insn.resize(n0);
      used_ = u0;
      return false;
    }
    return true;
  }

 private:
  const Batch &b_;
  std::vector<std::pair<int, int>> colreg_;
  uint32_t used_ = 0;
  bool bad_ = false;

  int alloc() {
    for (int r = 0; r < YSQL_XMAX_REGS; r++)
      if (!(used_ & (1u << r))) {
        used_ |= 1u << r;
        return r;
      }
    return -1;
  }
  void release(const XVal &v) {
    if (v.reg >= (int)inputs.size()) used_ &= ~(1u << v.reg);
  }
  bool emit(int op, int type, int sub, int dst, int a, int b, int nsrc, int64_t imm) {
    if (insn.size() >= YSQL_XMAX_INSN || dst < 0) return false;
    ysql_xinsn I;
    memset(&I, 0, sizeof(I));
    I.op = (int16_t)op;
    I.type = (int16_t)type;
    I.sub = (int16_t)sub;
    I.dst = (int8_t)dst;
    I.a = (int8_t)a;
    I.b = (int8_t)b;
    I.nsrc = (int8_t)nsrc;
    I.imm = imm;
    insn.push_back(I);
    return true;
  }
  // the operand whose NULL flags a binary node takes
  static const XVal &nullSide(const XVal &l, const XVal &r) { return l.scalar ? r : l; }

  bool compile(Expression *e, XVal *out) {
    const std::vector<ColumnPtr> &schema = b_.columns;
    if (auto *ref = dynamic_cast<BoundRefExpression *>(e)) {
      for (auto &cr : colreg_)
        if (cr.first == ref->idx) {
          out->reg = cr.second;
          out->type = inputs[cr.second].dtype;
          out->scalar = false;
          out->nullable = inputs[cr.second].nulls != nullptr;
          return true;
        }
      return false;
    }
    if (auto *val = dynamic_cast<BoundValueExpression *>(e)) {
      if (val->type == DataType::STRING) return false;
      int64_t imm = val->ival;
      if (val->type == DataType::FLOAT) {
        uint32_t bits;
        memcpy(&bits, &val->fval, 4);
        imm = bits;
      } else if (val->type == DataType::INT) {
        imm = (int32_t)val->ival;
      }
      const int r = alloc();
      if (!emit(YSQL_XOP_CONST, abiType(val->type), 0, r, 0, 0, -1, imm)) return false;
      *out = XVal{r, abiType(val->type), true, false};
      return true;
    }
    if (auto *c = dynamic_cast<ComparisonExpression *>(e)) {
      if (dynamic_cast<BoundValueExpression *>(c->rhs.get())) coerce(c->rhs.get(), c->lhs->returnType(schema));
      if (dynamic_cast<BoundValueExpression *>(c->lhs.get())) coerce(c->lhs.get(), c->rhs->returnType(schema));
      XVal l, r;
      if (!compile(c->lhs.get(), &l) || !compile(c->rhs.get(), &r)) return false;
      if (l.type == YSQL_BOOL || r.type == YSQL_BOOL || l.type != r.type) return false;
      const XVal &ns = nullSide(l, r);
      release(l);
      release(r);
      const int d = alloc();
      if (!emit(YSQL_XOP_CMP, l.type, c->cmp, d, l.reg, r.reg, ns.nullable ? ns.reg : -1, 0)) return false;
      *out = XVal{d, YSQL_BOOL, l.scalar && r.scalar, ns.nullable};
      return true;
    }
    if (auto *j = dynamic_cast<ConjunctionExpression *>(e)) {
      XVal l, r;
      if (!compile(j->lhs.get(), &l) || !compile(j->rhs.get(), &r)) return false;
      if (l.type != YSQL_BOOL || r.type != YSQL_BOOL) return false;
      release(l);
      release(r);
      const int d = alloc();
      if (!emit(YSQL_XOP_LOGIC, YSQL_BOOL, j->op, d, l.reg, r.reg, l.nullable ? l.reg : -1, 0)) return false;
      *out = XVal{d, YSQL_BOOL, l.scalar && r.scalar, l.nullable};
      return true;
    }
    if (auto *f = dynamic_cast<BoundFuncExpression *>(e)) {
      if (dynamic_cast<BoundValueExpression *>(f->rhs.get())) coerce(f->rhs.get(), f->lhs->returnType(schema));
      if (dynamic_cast<BoundValueExpression *>(f->lhs.get())) coerce(f->lhs.get(), f->rhs->returnType(schema));
      XVal l, r;
      if (!compile(f->lhs.get(), &l) || !compile(f->rhs.get(), &r)) return false;
      if (l.type == YSQL_BOOL || r.type == YSQL_BOOL || l.type != r.type) return false;
      if (f->op == YSQL_REM && l.type == YSQL_FLOAT) return false;
      if (f->op < YSQL_ADD || f->op > YSQL_REM) return false;
      const XVal &ns = nullSide(l, r);
      release(l);
      release(r);
      const int d = alloc();
      if (!emit(YSQL_XOP_ARITH, l.type, f->op, d, l.reg, r.reg, ns.nullable ? ns.reg : -1, 0)) return false;
      *out = XVal{d, l.type, l.scalar && r.scalar, ns.nullable};
      return true;
    }
    if (auto *k = dynamic_cast<CastExpression *>(e)) {
      XVal c;
      if (!compile(k->child.get(), &c)) return false;
      const int to = abiType(k->type);
      if (k->type == DataType::STRING) return false;
      if (c.type == YSQL_BOOL && to != YSQL_INT) return false;
      release(c);
      const int d = alloc();
      if (!emit(YSQL_XOP_CAST, c.type, to, d, c.reg, 0, c.nullable ? c.reg : -1, 0)) return false;
      *out = XVal{d, to, c.scalar, c.nullable};
      return true;
    }
    return false;
  }
};

}  // namespace

// ---------------------------------------------------------------- scan
TableScanOperator::TableScanOperator(ExecContext &ctx, std::shared_ptr<Table> t, size_t batch_rows)
    : Operator(ctx), table_(std::move(t)), batch_rows_(batch_rows ? batch_rows : (size_t)1 << 22) {
  columns = table_->columns;
}
BatchID TableScanOperator::next(CacheManager &cache) {
  if (isFinished || pos_ >= table_->rows) {
    isFinished = true;
    return 0;
  }
  const size_t n = std::min(batch_rows_, table_->rows - pos_);
  auto b = std::make_unique<Batch>();
  b->columns = table_->columns;
  b->data = table_->data;
  b->nulls = table_->nulls;
  b->offset.assign(b->columns.size(), pos_);
  b->batchSize = n;
  b->stream = ctx.stream();
  pos_ += n;
  return cache.putBatch(std::move(b));
}

// ---------------------------------------------------------------- filter
FilterOperator::FilterOperator(ExecContext &ctx, OpPtr child, ExprPtr pred) : Operator(ctx), pred_(std::move(pred)) {
  columns = child->columns;
  children.push_back(std::move(child));
}

static int flipCmp(int c) {
  switch (c) {
    case YSQL_LT: return YSQL_GT;
    case YSQL_LE: return YSQL_GE;
    case YSQL_GT: return YSQL_LT;
    case YSQL_GE: return YSQL_LE;
  }
  return c;
}

bool FilterOperator::fusedFilter(const Batch &in, const std::vector<ysql_col> &cols, size_t n, size_t wsb) {
  std::vector<int> pcols;
  collectCols(pred_.get(), pcols);
  XProgram prog(in, pcols);
  XVal pv;
  if (!prog.add(pred_.get(), &pv) || pv.type != YSQL_BOOL) return false;
  for (size_t c0 = 0; c0 < cols.size(); c0 += YSQL_MAX_COLS) {
    const int nc = (int)std::min<size_t>(YSQL_MAX_COLS, cols.size() - c0);
    abiCheck(ysql_expr_filter_compact(prog.insn.data(), (int)prog.insn.size(), pv.reg, prog.inputs.data(),
                                      (int)prog.inputs.size(), cols.data() + c0, nc, n, ctx.deviceCounter(),
                                      ctx.workspace(wsb), wsb, ctx.stream()),
             "ysql_expr_filter_compact");
  }
  return true;
}

BatchID FilterOperator::next(CacheManager &cache) {
  while (true) {
    if (isFinished) return 0;
    const BatchID cid = children[0]->next(cache);
    if (cid == 0) {
      isFinished = true;
      return 0;
    }
    auto in = cache.getBatch(cid);
    const size_t n = in->batchSize;
    if (n == 0) continue;  // the reference's Aggregate spun here on empty batches
    hipStream_t s = ctx.stream();
    // outputs sized for every row; the batch reports the selected count
    auto out = std::make_unique<Batch>();
    out->columns = in->columns;
    out->stream = s;
    std::vector<ysql_col> cols(in->numCols());
    for (size_t c = 0; c < in->numCols(); c++) {
      out->data.push_back(newBuf(n * in->columns[c]->bytes, s));
      out->nulls.push_back(in->nulls[c] ? newBuf(n, s) : nullptr);
      out->offset.push_back(0);
      cols[c] = ysql_col{in->col(c), in->nul(c), out->data[c]->ptr,
                         out->nulls[c] ? (uint8_t *)out->nulls[c]->ptr : nullptr, in->columns[c]->bytes, 0};
    }
    size_t wsb = 0;
    abiCheck(ysql_compact_ws_bytes(n, &wsb), "ysql_compact_ws_bytes");
    // fused path: <column> <cmp> <constant> (either side)
    auto *cmp = dynamic_cast<ComparisonExpression *>(pred_.get());
    BoundRefExpression *ref = nullptr;
    BoundValueExpression *val = nullptr;
    int op = -1;
    if (cmp) {
      if ((ref = dynamic_cast<BoundRefExpression *>(cmp->lhs.get())) &&
          (val = dynamic_cast<BoundValueExpression *>(cmp->rhs.get())))
        op = cmp->cmp;
      else if ((ref = dynamic_cast<BoundRefExpression *>(cmp->rhs.get())) &&
               (val = dynamic_cast<BoundValueExpression *>(cmp->lhs.get())))
        op = flipCmp(cmp->cmp);
    }
    if (op >= 0 && ref && ref->idx >= 0 && (size_t)ref->idx < in->numCols() &&
        in->columns[ref->idx]->type != DataType::STRING) {
      coerce(val, in->columns[ref->idx]->type);
      ExprValue sv = val->evaluate(*in, ctx);
      for (size_t c0 = 0; c0 < cols.size(); c0 += YSQL_MAX_COLS) {
        const int nc = (int)std::min<size_t>(YSQL_MAX_COLS, cols.size() - c0);
        abiCheck(ysql_filter_compact(abiType(in->columns[ref->idx]->type), op, in->col(ref->idx), sv.ptr(),
                                     cols.data() + c0, nc, n, ctx.deviceCounter(), ctx.workspace(wsb), wsb, s),
                 "ysql_filter_compact");
      }
    } else if (ctx.fusion() && fusedFilter(*in, cols, n, wsb)) {
      // the predicate tree ran as one program inside the compaction pass
    } else {
      ExprValue m = pred_->evaluate(*in, ctx);
      if (!m.mask) throw EngineError("filter predicate is not boolean");
      if (m.scalar) throw EngineError("constant filter predicate");
      for (size_t c0 = 0; c0 < cols.size(); c0 += YSQL_MAX_COLS) {
        const int nc = (int)std::min<size_t>(YSQL_MAX_COLS, cols.size() - c0);
        abiCheck(ysql_compact_cols((const uint8_t *)m.ptr(), cols.data() + c0, nc, n, ctx.deviceCounter(),
                                   ctx.workspace(wsb), wsb, s),
                 "ysql_compact_cols");
      }
    }
    const uint64_t cnt = ctx.readCounter();
    if (cnt == 0) continue;
    out->batchSize = cnt;
    return cache.putBatch(std::move(out));
  }
}

// ---------------------------------------------------------------- projection
ProjectionOperator::ProjectionOperator(ExecContext &ctx, OpPtr child, std::vector<ExprPtr> exprs)
    : Operator(ctx), exprs_(std::move(exprs)) {
  for (size_t i = 0; i < exprs_.size(); i++) {
    DataType t = exprs_[i]->isMask() ? DataType::INT : exprs_[i]->returnType(child->columns);
    columns.push_back(std::make_shared<Column>(exprs_[i]->name(), t));
  }
  children.push_back(std::move(child));
}
// Runs the non-reference expressions as fused programs: consecutive trees
// share one program while their columns fit the inputs (8) and their
// registers the register file; a tree that cannot be fused is left to the
// node-by-node path.
void ProjectionOperator::fuseExprs(const Batch &in, std::vector<ExprValue> &vals, std::vector<char> &done) {
  std::vector<size_t> cand;
  for (size_t i = 0; i < exprs_.size(); i++)
    if (!dynamic_cast<BoundRefExpression *>(exprs_[i].get())) cand.push_back(i);
  const size_t n = in.batchSize;
  size_t k = 0;
  while (k < cand.size()) {
    std::vector<size_t> group;
    std::vector<int> cols;
    for (size_t j = k; j < cand.size() && group.size() < YSQL_MAX_COLS; j++) {
      std::vector<int> c2 = cols;
      collectCols(exprs_[cand[j]].get(), c2);
      if (c2.size() > YSQL_XMAX_INPUTS) break;
      XProgram trial(in, c2);
      XVal v;
      bool ok = true;
      for (size_t g : group) ok = ok && trial.add(exprs_[g].get(), &v);
      ok = ok && trial.add(exprs_[cand[j]].get(), &v);
      if (!ok) break;
      group.push_back(cand[j]);
      cols = c2;
    }
    if (group.empty()) {  // this tree alone cannot be fused
      k++;
      continue;
    }
    XProgram prog(in, cols);
    std::vector<ysql_xoutput> outs;
    for (size_t g : group) {
      XVal v;
      if (!prog.add(exprs_[g].get(), &v)) throw EngineError("expression fusion: a checked tree failed to compile");
      const int dt = v.type == YSQL_BOOL ? YSQL_INT : v.type;
      ExprValue ev;
      ev.type = (DataType)dt;
      ev.rows = n;
      ev.buf = newBuf(n * getDataTypeNumBytes(ev.type), ctx.stream());
      ev.nullbuf = v.nullable ? newBuf(n, ctx.stream()) : nullptr;
      outs.push_back(ysql_xoutput{ev.buf->ptr, ev.nullbuf ? (uint8_t *)ev.nullbuf->ptr : nullptr, v.reg, dt});
      vals[g] = ev;
      done[g] = 1;
    }
    abiCheck(ysql_expr_eval(prog.insn.data(), (int)prog.insn.size(), prog.inputs.data(), (int)prog.inputs.size(),
                            outs.data(), (int)outs.size(), n, ctx.stream()),
             "ysql_expr_eval");
    k += group.size();
  }
}

BatchID ProjectionOperator::next(CacheManager &cache) {
  if (isFinished) return 0;
  const BatchID cid = children[0]->next(cache);
  if (cid == 0) {
    isFinished = true;
    return 0;
  }
  auto in = cache.getBatch(cid);
  auto out = std::make_unique<Batch>();
  out->stream = ctx.stream();
  out->batchSize = in->batchSize;
  std::vector<ExprValue> fused(exprs_.size());
  std::vector<char> is_fused(exprs_.size(), 0);
  if (ctx.fusion() && in->batchSize > 0) fuseExprs(*in, fused, is_fused);
  for (size_t i = 0; i < exprs_.size(); i++) {
    ExprValue v = is_fused[i] ? fused[i] : exprs_[i]->evaluate(*in, ctx);
    if (v.scalar && in->batchSize != 1) throw EngineError("constant projection is not supported");
    if (v.mask) {  // boolean -> INT column
      auto buf = newBuf(v.rows * 4, ctx.stream());
      abiCheck(ysql_cast(YSQL_BOOL, YSQL_INT, v.ptr(), buf->ptr, v.rows, ctx.stream()), "ysql_cast");
      v.buf = buf;
      v.offset = 0;
      v.mask = false;
      v.type = DataType::INT;
    }
    out->columns.push_back(columns[i]);
    out->data.push_back(v.buf);
    out->offset.push_back(v.offset);
    out->nulls.push_back(v.nullbuf);
  }
  return cache.putBatch(std::move(out));
}

// ---------------------------------------------------------------- order
OrderOperator::OrderOperator(ExecContext &ctx, OpPtr child, std::vector<OrderKey> keys)
    : Operator(ctx), keys_(std::move(keys)) {
  columns = child->columns;
  if (keys_.empty()) throw EngineError("ORDER BY without keys");
  for (auto &k : keys_) {
    if (k.col < 0 || (size_t)k.col >= columns.size()) throw EngineError("ORDER BY column out of range");
    const DataType t = columns[k.col]->type;
    if (t == DataType::STRING) by_rows_ = true;
    key_bits_ += t == DataType::DATETIME ? 64 : 32;
  }
  if (keys_.size() > YSQL_MAX_COLS) throw EngineError("at most 8 ORDER BY columns");
  // string keys, or numeric keys wider than one 64-bit composite: sorted
  // column by column (last key first, each sort stable) and merged by
  // comparing the key rows themselves
  if (key_bits_ > 64) by_rows_ = true;
  wide_ = key_bits_ > 32;
  children.push_back(std::move(child));
}

// One stable sort step of the column-by-column form: perm (row indices of
// b) reordered by column `key` of the rows it names.  `first`: perm is not
// yet set (the step writes the sorting permutation of the column itself).
void OrderOperator::sortStep(const Batch &b, const OrderKey &key, BufPtr &perm, bool first, void *ws, size_t wsb) {
  const size_t n = b.batchSize;
  hipStream_t s = ctx.stream();
  const DataType t = columns[key.col]->type;
  if (t == DataType::STRING) {
    if (first)
      abiCheck(ysql_string_argsort(key.desc, b.col(key.col), (uint32_t *)perm->ptr, n, ws, wsb, s),
               "ysql_string_argsort");
    else
      abiCheck(ysql_string_sort(key.desc, b.col(key.col), (uint32_t *)perm->ptr, n, ws, wsb, s), "ysql_string_sort");
    return;
  }
  const uint32_t eb = b.columns[key.col]->bytes;
  const int kdt = eb == 8 ? YSQL_UKEY64 : YSQL_UKEY32;
  const void *src = b.col(key.col);
  BufPtr tmp;
  if (!first) {  // the column in the current order
    tmp = newBuf(n * eb, s);
    abiCheck(ysql_gather(eb, src, nullptr, (const uint32_t *)perm->ptr, n, tmp->ptr, nullptr, s), "ysql_gather");
    src = tmp->ptr;
  }
  auto k = newBuf(n * eb, s), k2 = newBuf(n * eb, s);
  abiCheck(ysql_sortkey_encode(abiType(t), key.desc, src, k->ptr, n, s), "ysql_sortkey_encode");
  auto out = newBuf(n * 4, s);
  abiCheck(ysql_radix_sort_ex(kdt, 0, k->ptr, first ? nullptr : (const uint32_t *)perm->ptr, 0, k2->ptr,
                              (uint32_t *)out->ptr, n, 0, (int)eb * 8, ws, wsb, s),
           "ysql_radix_sort_ex");
  perm = out;
}

BufPtr OrderOperator::compositeKeys(const Batch &b) {
  const size_t n = b.batchSize;
  hipStream_t s = ctx.stream();
  if (!wide_) {
    auto k = newBuf(n * 4, s);
    const auto &key = keys_[0];
    abiCheck(ysql_sortkey_encode(abiType(columns[key.col]->type), key.desc, b.col(key.col), k->ptr, n, s),
             "ysql_sortkey_encode");
    return k;
  }
  auto k = newBuf(n * 8, s);
  for (auto &key : keys_)
    abiCheck(ysql_sortkey_pack(abiType(columns[key.col]->type), key.desc, b.col(key.col), n, (uint64_t *)k->ptr, s),
             "ysql_sortkey_pack");
  return k;
}

// A child batch sorted into a run.  Unless sorting column by column, the
// sorted composite keys ride along as a last, hidden column (UKEY32 or
// UKEY64 bits, typed INT / DATETIME for their width), so the run keeps
// everything its merges need while the cache may move it out of HBM.
std::unique_ptr<Batch> OrderOperator::sortBatch(const Batch &b) {
  const size_t n = b.batchSize;
  hipStream_t s = ctx.stream();
  BufPtr keys;
  auto perm = newBuf(n * 4, s);
  if (by_rows_) {
    size_t wsb = 0, w = 0;
    abiCheck(ysql_radix_sort_ws_bytes(YSQL_UKEY64, n, &w), "ysql_radix_sort_ws_bytes");
    wsb = w;
    if (n > 1) {
      abiCheck(ysql_string_sort_ws_bytes(n, &w), "ysql_string_sort_ws_bytes");
      wsb = std::max(wsb, w);
    }
    void *ws = ctx.workspace(wsb);
    for (size_t q = keys_.size(); q-- > 0;) sortStep(b, keys_[q], perm, q + 1 == keys_.size(), ws, wsb);
  } else {
    keys = compositeKeys(b);
    const int kdt = wide_ ? YSQL_UKEY64 : YSQL_UKEY32;
    size_t wsb = 0;
    abiCheck(ysql_radix_sort_ws_bytes(kdt, n, &wsb), "ysql_radix_sort_ws_bytes");
    // the packed key occupies its low key_bits_ bits
    abiCheck(ysql_radix_argsort(kdt, 0, keys->ptr, (uint32_t *)perm->ptr, n, 0, key_bits_, ctx.workspace(wsb), wsb,
                                s),
             "ysql_radix_argsort");
  }
  auto out = std::make_unique<Batch>();
  out->columns = b.columns;
  out->stream = s;
  out->batchSize = n;
  for (size_t c = 0; c < b.numCols(); c++) {
    out->data.push_back(newBuf(n * b.columns[c]->bytes, s));
    out->nulls.push_back(b.nulls[c] ? newBuf(n, s) : nullptr);
    out->offset.push_back(0);
    abiCheck(ysql_gather(b.columns[c]->bytes, b.col(c), b.nul(c), (const uint32_t *)perm->ptr, n, out->data[c]->ptr,
                         out->nulls[c] ? (uint8_t *)out->nulls[c]->ptr : nullptr, s),
             "ysql_gather");
  }
  if (keys) {
    out->columns.push_back(std::make_shared<Column>("", wide_ ? DataType::DATETIME : DataType::INT));
    out->data.push_back(keys);
    out->nulls.push_back(nullptr);
    out->offset.push_back(0);
  }
  return out;
}

// Two runs taken from the cache, merged (stable, the first run first on
// ties), the result put back.
BatchID OrderOperator::mergeRuns(BatchID ia, BatchID ib, CacheManager &cache) {
  hipStream_t s = ctx.stream();
  auto a = cache.getBatch(ia);
  auto b = cache.getBatch(ib);
  const size_t m = a->batchSize, n = b->batchSize, k = m + n;
  const size_t nc = columns.size();  // user columns; a composite key column follows them
  auto idx = newBuf(k * 4, s), isa = newBuf(k, s);
  BufPtr keys;
  if (by_rows_) {  // the runs' key columns compared row against row
    ysql_sortcol sc[YSQL_MAX_COLS];
    for (size_t q = 0; q < keys_.size(); q++)
      sc[q] = ysql_sortcol{abiType(columns[keys_[q].col]->type), keys_[q].desc ? 1 : 0, a->col(keys_[q].col),
                           b->col(keys_[q].col)};
    abiCheck(ysql_merge_rows(sc, (int)keys_.size(), m, n, k, (uint32_t *)idx->ptr, (uint8_t *)isa->ptr, nullptr,
                             nullptr, s),
             "ysql_merge_rows");
  } else {
    const int kdt = wide_ ? YSQL_UKEY64 : YSQL_UKEY32;
    keys = newBuf(k * (wide_ ? 8 : 4), s);
    size_t wsb = 0;
    abiCheck(ysql_merge_ws_bytes(m, n, k, &wsb), "ysql_merge_ws_bytes");
    abiCheck(ysql_merge_path(kdt, 0, a->col(nc), m, b->col(nc), n, k, keys->ptr, (uint32_t *)idx->ptr,
                             (uint8_t *)isa->ptr, (int64_t *)ctx.deviceCounter(), ctx.workspace(wsb), wsb, s),
             "ysql_merge_path");
  }
  auto out = std::make_unique<Batch>();
  out->columns = a->columns;
  out->stream = s;
  out->batchSize = k;
  for (size_t c = 0; c < nc; c++) {
    const bool hasn = a->nulls[c] || b->nulls[c];
    out->data.push_back(newBuf(k * out->columns[c]->bytes, s));
    out->nulls.push_back(hasn ? newBuf(k, s) : nullptr);
    out->offset.push_back(0);
    abiCheck(ysql_gather2(out->columns[c]->bytes, a->col(c), a->nul(c), b->col(c), b->nul(c),
                          (const uint32_t *)idx->ptr, (const uint8_t *)isa->ptr, k, 0, out->data[c]->ptr,
                          hasn ? (uint8_t *)out->nulls[c]->ptr : nullptr, s),
             "ysql_gather2");
  }
  if (keys) {
    out->data.push_back(keys);
    out->nulls.push_back(nullptr);
    out->offset.push_back(0);
  }
  a.reset();
  b.reset();
  return cache.putBatch(std::move(out));
}

void OrderOperator::firstPass(CacheManager &cache) {
  // pass 1: every child batch becomes one sorted run, kept in the cache
  // (order_operator.hpp:89-121)
  while (true) {
    const BatchID id = children[0]->next(cache);
    if (!id) break;
    auto b = cache.getBatch(id);
    if (b->batchSize == 0) continue;
    max_batch_ = std::max(max_batch_, b->batchSize);
    runs_.push_back(cache.putBatch(sortBatch(*b)));
  }
  // pass 2: pairwise merges, level by level, carrying an odd run
  // (order_operator.hpp:160-193)
  while (runs_.size() > 1) {
    std::vector<BatchID> nxt;
    for (size_t i = 0; i + 1 < runs_.size(); i += 2) nxt.push_back(mergeRuns(runs_[i], runs_[i + 1], cache));
    if (runs_.size() % 2) nxt.push_back(runs_.back());
    runs_ = std::move(nxt);
  }
  if (!runs_.empty()) final_ = std::shared_ptr<Batch>(cache.getBatch(runs_[0]));
  runs_.clear();
  done_first_ = true;
}

BatchID OrderOperator::next(CacheManager &cache) {
  if (isFinished) return 0;
  if (!done_first_) firstPass(cache);
  if (!final_ || emit_pos_ >= final_->batchSize) {
    isFinished = true;
    final_.reset();
    return 0;
  }
  const size_t n = std::min(max_batch_, final_->batchSize - emit_pos_);
  auto v = final_->slice(emit_pos_, n);
  while (v->columns.size() > columns.size()) {  // the composite key column stays behind
    v->columns.pop_back();
    v->data.pop_back();
    v->nulls.pop_back();
    v->offset.pop_back();
  }
  emit_pos_ += n;
  return cache.putBatch(std::move(v));
}

// ---------------------------------------------------------------- aggregate
AggregateOperator::AggregateOperator(ExecContext &ctx, OpPtr child, std::vector<AggSpec> aggs,
                                     std::vector<GroupSpec> groups)
    : Operator(ctx), aggs_(std::move(aggs)), groups_(std::move(groups)) {
  if (groups_.size() > 4) throw EngineError("at most 4 GROUP BY columns");
  uint64_t ng = 1;
  for (auto &g : groups_) {
    if (g.col < 0 || (size_t)g.col >= child->columns.size() || child->columns[g.col]->type != DataType::INT)
      throw EngineError("GROUP BY needs INT columns");
    if (g.card == 0) throw EngineError("GROUP BY cardinality 0");
    ng *= g.card;
  }
  if (ng > YSQL_MAX_GROUPS) throw EngineError("too many groups");
  ngroups_ = (uint32_t)ng;
  for (auto &g : groups_) columns.push_back(std::make_shared<Column>(child->columns[g.col]->name, DataType::INT));
  for (auto &a : aggs_) {
    if (a.agg < YSQL_AGG_MAX || a.agg > YSQL_AGG_COUNT_STAR) throw EngineError("bad aggregate");
    DataType src = DataType::INT;
    if (a.agg != YSQL_AGG_COUNT_STAR) {
      if (a.col < 0 || (size_t)a.col >= child->columns.size()) throw EngineError("aggregate column out of range");
      src = child->columns[a.col]->type;
      if (src == DataType::STRING) throw EngineError("aggregate over strings");
    }
    DataType out;
    switch (a.agg) {
      case YSQL_AGG_MIN: case YSQL_AGG_MAX: out = src; break;
      case YSQL_AGG_SUM: out = src == DataType::FLOAT ? DataType::FLOAT : DataType::DATETIME; break;
      case YSQL_AGG_AVG: out = DataType::FLOAT; break;
      default: out = DataType::DATETIME; break;
    }
    columns.push_back(std::make_shared<Column>("agg", out));
  }
  children.push_back(std::move(child));
}

BatchID AggregateOperator::next(CacheManager &cache) {
  if (isFinished) return 0;
  hipStream_t s = ctx.stream();
  const size_t na = aggs_.size();
  const uint32_t G = ngroups_;
  // accumulators: one per aggregate per group, plus a row count per group
  auto accs = newBuf((na + 1) * G * sizeof(ysql_acc), s);
  ysql_acc *A = (ysql_acc *)accs->ptr;
  const auto &in_cols = children[0]->columns;
  for (size_t i = 0; i < na; i++) {
    const int t = aggs_[i].agg == YSQL_AGG_COUNT_STAR ? YSQL_INT : abiType(in_cols[aggs_[i].col]->type);
    abiCheck(ysql_acc_init(t, aggs_[i].agg, A + i * G, (int)G, s), "ysql_acc_init");
  }
  abiCheck(ysql_acc_init(YSQL_INT, YSQL_AGG_COUNT_STAR, A + na * G, (int)G, s), "ysql_acc_init");
  while (true) {
    const BatchID id = children[0]->next(cache);
    if (!id) break;
    auto b = cache.getBatch(id);
    const size_t n = b->batchSize;
    if (n == 0) continue;  // reference spun forever here (aggregate_operator.hpp:46-48)
    if (groups_.empty()) {
      size_t wsb = 0;
      abiCheck(ysql_reduce_ws_bytes(n, &wsb), "ysql_reduce_ws_bytes");
      for (size_t i = 0; i < na; i++) {
        const int agg = aggs_[i].agg;
        const bool star = agg == YSQL_AGG_COUNT_STAR;
        abiCheck(ysql_reduce(star ? YSQL_INT : abiType(b->columns[aggs_[i].col]->type), agg,
                             star ? nullptr : b->col(aggs_[i].col), star ? nullptr : b->nul(aggs_[i].col), n, A + i,
                             ctx.workspace(wsb), wsb, s),
                 "ysql_reduce");
      }
      abiCheck(ysql_reduce(YSQL_INT, YSQL_AGG_COUNT_STAR, nullptr, nullptr, n, A + na, ctx.workspace(wsb), wsb, s),
               "ysql_reduce");
    } else {
      auto gid = newBuf(n * 4, s);
      const int32_t *gcols[4];
      int32_t base[4];
      uint32_t card[4];
      for (size_t g = 0; g < groups_.size(); g++) {
        gcols[g] = (const int32_t *)b->col(groups_[g].col);
        base[g] = groups_[g].base;
        card[g] = groups_[g].card;
      }
      abiCheck(ysql_group_ids(gcols, base, card, (int)groups_.size(), n, (uint32_t *)gid->ptr, s), "ysql_group_ids");
      for (size_t i = 0; i < na; i++) {
        const int agg = aggs_[i].agg;
        const bool star = agg == YSQL_AGG_COUNT_STAR;
        abiCheck(ysql_group_reduce(star ? YSQL_INT : abiType(b->columns[aggs_[i].col]->type), agg,
                                   (const uint32_t *)gid->ptr, G, star ? gid->ptr : b->col(aggs_[i].col),
                                   star ? nullptr : b->nul(aggs_[i].col), n, A + i * G, s),
                 "ysql_group_reduce");
      }
      abiCheck(ysql_group_reduce(YSQL_INT, YSQL_AGG_COUNT_STAR, (const uint32_t *)gid->ptr, G, gid->ptr, nullptr, n,
                                 A + na * G, s),
               "ysql_group_reduce");
    }
  }
  // finalise on the host (getAggregate, aggregate.hpp:149-192)
  std::vector<ysql_acc> h((na + 1) * G);
  hipCheck(hipMemcpyAsync(h.data(), A, h.size() * sizeof(ysql_acc), hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
  ctx.sync();
  std::vector<uint32_t> live;
  for (uint32_t g = 0; g < G; g++)
    if (groups_.empty() || h[na * G + g].count > 0) live.push_back(g);
  const size_t rows = live.size();
  auto out = std::make_unique<Batch>();
  out->stream = s;
  out->batchSize = rows;
  out->columns = columns;
  std::vector<std::vector<uint8_t>> host(columns.size());
  std::vector<std::vector<uint8_t>> hnul(columns.size(), std::vector<uint8_t>(rows, 0));
  for (size_t c = 0; c < columns.size(); c++) host[c].resize(rows * columns[c]->bytes);
  // group key columns from the dense id
  std::vector<uint64_t> stride(groups_.size());
  uint64_t st = 1;
  for (int g = (int)groups_.size() - 1; g >= 0; g--) {
    stride[g] = st;
    st *= groups_[g].card;
  }
  for (size_t r = 0; r < rows; r++) {
    const uint32_t g = live[r];
    for (size_t k = 0; k < groups_.size(); k++) {
      int32_t v = (int32_t)((g / stride[k]) % groups_[k].card) + groups_[k].base;
      memcpy(&host[k][r * 4], &v, 4);
    }
    for (size_t i = 0; i < na; i++) {
      const ysql_acc &a = h[i * G + g];
      const size_t c = groups_.size() + i;
      const DataType src = aggs_[i].agg == YSQL_AGG_COUNT_STAR ? DataType::INT : in_cols[aggs_[i].col]->type;
      uint8_t *dst = &host[c][r * columns[c]->bytes];
      const int agg = aggs_[i].agg;
      if (agg == YSQL_AGG_COUNT || agg == YSQL_AGG_COUNT_STAR) {
        int64_t v = a.count;
        memcpy(dst, &v, 8);
        continue;
      }
      if (a.count == 0) hnul[c][r] = 1;  // MIN/MAX/SUM/AVG of no rows is NULL
      if (agg == YSQL_AGG_MIN || agg == YSQL_AGG_MAX) {
        if (src == DataType::INT) { int32_t v = (int32_t)a.i; memcpy(dst, &v, 4); }
        else if (src == DataType::DATETIME) memcpy(dst, &a.i, 8);
        else memcpy(dst, &a.f, 4);
      } else if (agg == YSQL_AGG_SUM) {
        if (src == DataType::FLOAT) { float v = (float)a.d; memcpy(dst, &v, 4); }
        else memcpy(dst, &a.i, 8);
      } else {  // AVG: (double)sum / (double)count as float (div_avg_kernel)
        const double sum = src == DataType::FLOAT ? a.d : (double)a.i;
        float v = a.count ? (float)(sum / (double)a.count) : 0.f;
        memcpy(dst, &v, 4);
      }
    }
  }
  for (size_t c = 0; c < columns.size(); c++) {
    auto d = newBuf(host[c].size(), s);
    auto nb = newBuf(rows, s);
    if (rows) {
      hipCheck(hipMemcpyAsync(d->ptr, host[c].data(), host[c].size(), hipMemcpyHostToDevice, s), "hipMemcpyAsync");
      hipCheck(hipMemcpyAsync(nb->ptr, hnul[c].data(), rows, hipMemcpyHostToDevice, s), "hipMemcpyAsync");
    }
    out->data.push_back(d);
    out->nulls.push_back(nb);
    out->offset.push_back(0);
  }
  ctx.sync();  // host staging goes out of scope
  isFinished = true;
  if (rows == 0) return 0;
  return cache.putBatch(std::move(out));
}

// ---------------------------------------------------------------- join
JoinOperator::JoinOperator(ExecContext &ctx, OpPtr left, OpPtr right, int lcol, int rcol, int cmp)
    : Operator(ctx), lcol_(lcol), rcol_(rcol), cmp_(cmp) {
  if (lcol < 0 || (size_t)lcol >= left->columns.size() || rcol < 0 || (size_t)rcol >= right->columns.size())
    throw EngineError("join column out of range");
  if (left->columns[lcol]->type != right->columns[rcol]->type) throw EngineError("join on different types");
  columns = left->columns;
  columns.insert(columns.end(), right->columns.begin(), right->columns.end());
  children.push_back(std::move(left));
  children.push_back(std::move(right));
}

void JoinOperator::cacheLeftChild(CacheManager &cache) {
  while (true) {
    const BatchID id = children[0]->next(cache);
    if (!id) break;
    auto b = cache.getBatch(id);
    if (b->batchSize) left_.push_back(std::shared_ptr<Batch>(std::move(b)));
  }
  left_cached_ = true;
}

BatchID JoinOperator::next(CacheManager &cache) {
  if (isFinished) return 0;
  if (!left_cached_) cacheLeftChild(cache);
  hipStream_t s = ctx.stream();
  while (true) {
    if (!right_ || li_ >= left_.size()) {
      const BatchID id = children[1]->next(cache);
      if (!id || left_.empty()) {
        isFinished = true;
        return 0;
      }
      right_ = std::shared_ptr<Batch>(cache.getBatch(id));
      li_ = 0;
      if (right_->batchSize == 0) continue;
    }
    const Batch &L = *left_[li_++];
    const Batch &R = *right_;
    const int dt = abiType(L.columns[lcol_]->type);
    size_t wsb = 0;
    abiCheck(ysql_join_ws_bytes(dt, cmp_, L.batchSize, R.batchSize, &wsb), "ysql_join_ws_bytes");
    BufPtr pairs;
    uint64_t cnt = 0;
    for (int attempt = 0; attempt < 2; attempt++) {
      pairs = newBuf(cap_ * 8, s);
      abiCheck(ysql_join_pairs(dt, cmp_, L.col(lcol_), L.batchSize, R.col(rcol_), R.batchSize, (uint32_t *)pairs->ptr,
                               cap_, ctx.deviceCounter(), ctx.workspace(wsb), wsb, s),
               "ysql_join_pairs");
      cnt = ctx.readCounter();
      if (cnt <= cap_) break;
      cap_ = cnt;  // capacity grows to the exact count and the pair set is rebuilt
    }
    if (cnt == 0) continue;
    auto out = std::make_unique<Batch>();
    out->columns = columns;
    out->stream = s;
    out->batchSize = cnt;
    auto emit = [&](const Batch &src, int side) {
      for (size_t c = 0; c < src.numCols(); c++) {
        auto d = newBuf(cnt * src.columns[c]->bytes, s);
        auto nb = src.nulls[c] ? newBuf(cnt, s) : nullptr;
        abiCheck(ysql_gather_pairs(src.columns[c]->bytes, src.col(c), src.nul(c), (const uint32_t *)pairs->ptr, side,
                                   cnt, d->ptr, nb ? (uint8_t *)nb->ptr : nullptr, s),
                 "ysql_gather_pairs");
        out->data.push_back(d);
        out->nulls.push_back(nb);
        out->offset.push_back(0);
      }
    };
    emit(L, 0);
    emit(R, 1);
    return cache.putBatch(std::move(out));
  }
}

// ---------------------------------------------------------------- cross product
CrossProductOperator::CrossProductOperator(ExecContext &ctx, OpPtr left, OpPtr right) : Operator(ctx) {
  columns = left->columns;
  columns.insert(columns.end(), right->columns.begin(), right->columns.end());
  children.push_back(std::move(left));
  children.push_back(std::move(right));
}

BatchID CrossProductOperator::next(CacheManager &cache) {
  if (isFinished) return 0;
  hipStream_t s = ctx.stream();
  if (!left_cached_) {
    while (true) {
      const BatchID id = children[0]->next(cache);
      if (!id) break;
      auto b = cache.getBatch(id);
      if (b->batchSize) left_.push_back(std::shared_ptr<Batch>(std::move(b)));
    }
    left_cached_ = true;
  }
  constexpr size_t kMaxOut = (size_t)1 << 24;  // rows per output batch
  while (true) {
    if (!right_ || roff_ >= right_->batchSize) {
      const BatchID id = children[1]->next(cache);
      if (!id || left_.empty()) {
        isFinished = true;
        return 0;
      }
      right_ = std::shared_ptr<Batch>(cache.getBatch(id));
      roff_ = 0;
      li_ = 0;
      size_t lmax = 1;
      for (auto &l : left_) lmax = std::max(lmax, l->batchSize);
      rchunk_ = std::max<size_t>(1, kMaxOut / lmax);
      if (right_->batchSize == 0) continue;
    }
    const Batch &L = *left_[li_];
    const size_t rlen = std::min(rchunk_, right_->batchSize - roff_);
    auto Rv = right_->slice(roff_, rlen);
    const size_t rows = L.batchSize * rlen;
    auto out = std::make_unique<Batch>();
    out->columns = columns;
    out->stream = s;
    out->batchSize = rows;
    auto emit = [&](const Batch &src, int is_left) {
      for (size_t c = 0; c < src.numCols(); c++) {
        auto d = newBuf(rows * src.columns[c]->bytes, s);
        auto nb = src.nulls[c] ? newBuf(rows, s) : nullptr;
        abiCheck(ysql_cross_expand(src.columns[c]->bytes, src.col(c), src.nul(c), L.batchSize, rlen, is_left, d->ptr,
                                   nb ? (uint8_t *)nb->ptr : nullptr, s),
                 "ysql_cross_expand");
        out->data.push_back(d);
        out->nulls.push_back(nb);
        out->offset.push_back(0);
      }
    };
    emit(L, 1);
    emit(*Rv, 0);
    if (++li_ >= left_.size()) {
      li_ = 0;
      roff_ += rlen;
    }
    return cache.putBatch(std::move(out));
  }
}

// ---------------------------------------------------------------- execution
// Layout of a result block of n rows: every column's data, then every
// column's null bytes, each at a 256-byte boundary.  Returns the block size.
static size_t resultLayout(const std::vector<ColumnPtr> &cols, size_t n, std::vector<size_t> &doff,
                           std::vector<size_t> &noff) {
  const size_t nc = cols.size();
  doff.resize(nc);
  noff.resize(nc);
  size_t off = 0;
  for (size_t c = 0; c < nc; c++) {
    doff[c] = off;
    off = (off + n * cols[c]->bytes + 255) / 256 * 256;
  }
  for (size_t c = 0; c < nc; c++) {
    noff[c] = off;
    off = (off + n + 255) / 256 * 256;
  }
  return off;
}

HostResult executeToHost(Operator &root, ExecContext &ctx) {
  HostResult r;
  r.columns = root.columns;
  const size_t nc = r.columns.size();
  root.init();
  CacheManager &cache = ctx.cache();
  struct Part {
    PinnedBlock blk;
    size_t rows;
  };
  std::vector<Part> parts;
  std::vector<size_t> doff, noff;
  try {
    while (true) {
      const BatchID id = root.next(cache);
      if (!id) break;
      auto b = cache.getBatch(id);
      const size_t n = b->batchSize;
      r.batches++;
      if (!n) continue;
      // each batch's columns go straight into pinned memory, all copies queued
      // before one wait at the end (a batch's device buffers are freed in
      // stream order, after its copies)
      Part p{pinnedTake(resultLayout(r.columns, n, doff, noff)), n};
      parts.push_back(p);
      char *base = static_cast<char *>(p.blk.ptr);
      for (size_t c = 0; c < nc; c++) {
        hipCheck(hipMemcpyAsync(base + doff[c], b->col(c), n * r.columns[c]->bytes, hipMemcpyDeviceToHost,
                                ctx.stream()),
                 "hipMemcpyAsync");
        if (b->nul(c))
          hipCheck(hipMemcpyAsync(base + noff[c], b->nul(c), n, hipMemcpyDeviceToHost, ctx.stream()),
                   "hipMemcpyAsync");
        else
          std::memset(base + noff[c], 0, n);
      }
      r.rows += n;
    }
    ctx.sync();
  } catch (...) {
    (void)hipStreamSynchronize(ctx.stream());
    for (auto &p : parts) pinnedGive(p.blk);
    throw;
  }
  r.data.assign(nc, nullptr);
  r.nulls.assign(nc, nullptr);
  if (parts.empty()) return r;
  if (parts.size() == 1) {
    r.block = parts[0].blk;
  } else {  // several batches: one block in row order
    r.block = pinnedTake(resultLayout(r.columns, r.rows, doff, noff));
    std::vector<size_t> pd, pn;
    size_t at = 0;
    for (auto &p : parts) {
      resultLayout(r.columns, p.rows, pd, pn);
      const char *src = static_cast<const char *>(p.blk.ptr);
      char *dst = static_cast<char *>(r.block.ptr);
      for (size_t c = 0; c < nc; c++) {
        const size_t eb = r.columns[c]->bytes;
        std::memcpy(dst + doff[c] + at * eb, src + pd[c], p.rows * eb);
        std::memcpy(dst + noff[c] + at, src + pn[c], p.rows);
      }
      at += p.rows;
      pinnedGive(p.blk);
    }
  }
  resultLayout(r.columns, r.rows, doff, noff);
  for (size_t c = 0; c < nc; c++) {
    r.data[c] = static_cast<uint8_t *>(r.block.ptr) + doff[c];
    r.nulls[c] = static_cast<uint8_t *>(r.block.ptr) + noff[c];
  }
  return r;
}

}  // namespace YallaSQL
