// csv_loader.cpp — CSV file -> column batches (ysql_csv.h): the feeder of
// the scan path, replacing the reference's GetOperator
// (src/engine/operators/get_operator.cpp:16-148).
//
// The file is mapped; one pass over a batch's bytes finds its record starts
// (RFC 4180 quotes), then up to 16 threads tokenise and convert disjoint row
// ranges with the reference's rules into one of two pinned staging buffers,
// whose host-to-device copies are queued on the caller's stream while the
// next batch is parsed.
#include <hip/hip_runtime.h>

#include <emmintrin.h>
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <charconv>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <limits>
#include <string>
#include <thread>
#include <vector>

#include "common.hpp"
#include "ysql_csv.h"

namespace {

constexpr size_t kMaxBytesPerBatch = 4u << 20;  // config.hpp: MAX_BYTES_PER_BATCH
constexpr uint32_t kAlign = 256;                // config.hpp: GPU_ALIGNMENT = BLOCK_DIM

uint32_t type_bytes(int t) {
  switch (t) {
    case YSQL_INT: return 4;
    case YSQL_FLOAT: return 4;
    case YSQL_DATETIME: return 8;
    case YSQL_STRING: return YSQL_STRING_BYTES;
    default: return 0;
  }
}

// One staging area: a pinned buffer per column and per null column, and the
// event recorded after the copies queued from it.
struct Stage {
  std::vector<void *> cols;
  std::vector<uint8_t *> nulls;
  hipEvent_t done = nullptr;
  bool pending = false;
};

}  // namespace

// Reads one record starting at p (RFC 4180: fields separated by ',', records
// by "\n" or "\r\n", a field in quotes may hold separators and newlines, ""
// inside quotes is one quote; a quote that does not open a field is an
// ordinary character) into rec[0 .. nrec); returns the start of the next
// record.  Runs of ordinary characters are copied in bulk.
static const char *parse_record(const char *p, const char *e, std::vector<std::string> &rec, size_t &nrec) {
  nrec = 0;
  auto next_field = [&]() {
    if (nrec == rec.size()) rec.emplace_back();
    std::string *f = &rec[nrec++];
    f->clear();
    return f;
  };
  std::string *fld = next_field();
  bool quoted = false;
  while (p < e) {  // the end of the file ends the record
    const char *q = p;
    if (!quoted) {
      while (q < e && *q != ',' && *q != '\n' && *q != '\r' && *q != '"') q++;
      fld->append(p, (size_t)(q - p));
      p = q;
      if (p == e) break;
      const char c = *p++;
      if (c == ',') {
        fld = next_field();
      } else if (c == '\n') {
        break;
      } else if (c == '\r') {
        if (p < e && *p == '\n') {
          p++;
          break;
        }
        fld->push_back('\r');
      } else if (fld->empty()) {  // '"' opening a field
        quoted = true;
      } else {
        fld->push_back('"');
      }
    } else {
      while (q < e && *q != '"') q++;
      fld->append(p, (size_t)(q - p));
      p = q;
      if (p == e) break;
      p++;  // the quote
      if (p < e && *p == '"') {
        p++;
        fld->push_back('"');
      } else {
        quoted = false;
      }
    }
  }
  return p;
}

struct ysql_csv_reader {
  int fd = -1;
  const char *data = nullptr;  // the whole file, mapped read-only
  size_t size = 0;
  const char *pos = nullptr, *end = nullptr;
  const char *next_quote = nullptr;  // first '"' at or after the scan position (end: none)
  std::vector<int> types;
  std::vector<uint32_t> bytes;
  std::vector<int> field;  // CSV field index of each requested column
  int nfields = 0;
  uint32_t batch_rows = 0;
  int threads = 1;
  std::vector<const char *> starts;  // record starts of the batch being parsed
  Stage st[2];
  int cur = 0;

  const char *quote_at_or_after(const char *p) {
    if (next_quote < p) {
      const void *q = memchr(p, '"', (size_t)(end - p));
      next_quote = q ? (const char *)q : end;
    }
    return next_quote;
  }

  // Finds the end of the record starting at p (the start of the next one),
  // following parse_record's quote rules, and whether the record is blank
  // (no bytes, or a lone empty quoted field, before its terminator).
  const char *scan_record(const char *p, bool *blank) {
    const char *rs = p;
    const char *content_end = nullptr;
    while (p < end) {
      const void *nlv = memchr(p, '\n', (size_t)(end - p));
      const char *nl = nlv ? (const char *)nlv : end;
      const char *q = quote_at_or_after(p);
      while (q < nl && !(q == rs || q[-1] == ',')) q = quote_at_or_after(q + 1);  // ordinary quotes
      if (q >= nl) {  // no field opens with a quote before the newline
        content_end = (nl > rs && nl < end && nl[-1] == '\r') ? nl - 1 : nl;
        p = nl < end ? nl + 1 : end;
        break;
      }
      p = q + 1;  // inside a quoted field: find its closing quote
      while (true) {
        const char *c = quote_at_or_after(p);
        if (c == end) {
          p = end;
          break;
        }
        if (c + 1 < end && c[1] == '"') {
          p = c + 2;
          continue;
        }
        p = c + 1;
        break;
      }
    }
    if (!content_end) content_end = p;
    const size_t n = (size_t)(content_end - rs);
    *blank = n == 0 || (n == 2 && rs[0] == '"' && rs[1] == '"');
    return p;
  }
};

namespace {

// Converts one field into element `row` of a column; false if the reference
// would throw (std::stoi / std::stof).  get_operator.cpp:74-95.
bool convert(int type, const std::string &s, void *col, uint8_t *nulls, uint32_t row) {
  const bool empty = s.empty();
  if (nulls) nulls[row] = empty ? 1 : 0;
  switch (type) {
    case YSQL_INT: {
      int32_t v = std::numeric_limits<int32_t>::min();
      if (!empty) {
        // plain [-]digits prefix: from_chars gives what strtol gives; any
        // other form (blanks, '+') takes strtol itself
        long x = 0;
        const auto fc = std::from_chars(s.data(), s.data() + s.size(), x, 10);
        if (fc.ec == std::errc::result_out_of_range) return false;
        if (fc.ec != std::errc()) {
          errno = 0;
          char *end = nullptr;
          x = strtol(s.c_str(), &end, 10);
          if (end == s.c_str() || errno == ERANGE) return false;
        }
        if (x < INT_MIN || x > INT_MAX) return false;
        v = (int32_t)x;
      }
      memcpy((char *)col + (size_t)row * 4, &v, 4);
      return true;
    }
    case YSQL_FLOAT: {
      float v = std::numeric_limits<float>::min();
      if (!empty) {
        // plain decimal prefix: from_chars rounds correctly, as strtof does;
        // hex forms, blanks, '+', range errors and zero or subnormal results
        // take strtof itself (its errno rules decide those)
        const char *b = s.data() + (s[0] == '-');
        const bool hex = s.size() > (size_t)(b - s.data()) + 1 && b[0] == '0' && (b[1] == 'x' || b[1] == 'X');
        const auto fc = std::from_chars(s.data(), s.data() + s.size(), v);
        if (hex || fc.ec != std::errc() || !(std::fabs(v) >= std::numeric_limits<float>::min())) {
          errno = 0;
          char *end = nullptr;
          v = strtof(s.c_str(), &end);
          if (end == s.c_str() || errno == ERANGE) return false;
        }
      }
      memcpy((char *)col + (size_t)row * 4, &v, 4);
      return true;
    }
    case YSQL_DATETIME: {
      int64_t v = 0;
      if (!empty) {  // data_type_utils.hpp:24-44
        struct tm tmv;
        memset(&tmv, 0, sizeof(tmv));
        if (strptime(s.c_str(), "%Y-%m-%d %H:%M:%S", &tmv) != nullptr) v = (int64_t)timegm(&tmv);
      }
      memcpy((char *)col + (size_t)row * 8, &v, 8);
      return true;
    }
    default: {  // STRING: string_kernel.hpp String::set
      char *d = (char *)col + (size_t)row * YSQL_STRING_BYTES;
      const size_t n = s.size() < YSQL_STRING_BYTES - 1 ? s.size() : YSQL_STRING_BYTES - 1;
      size_t i = 0;
      for (; i < n && s[i] != '\0'; i++) d[i] = s[i];
      memset(d + i, 0, YSQL_STRING_BYTES - i);
      return true;
    }
  }
}

// Record starts in quote-free text [p, lim), where every '\n' ends a record:
// appends the starts of the complete records found (blank ones, "\n" or
// "\r\n", skipped when `skip_blank`) until `want` starts are held, and
// returns the start of the first record not consumed.  Newlines are found
// 16 bytes at a time.
const char *quote_free_starts(const char *p, const char *lim, size_t want, bool skip_blank,
                              std::vector<const char *> &starts) {
  const char *rs = p, *c = p;
  auto record_end = [&](const char *nl) {
    const bool blank = nl == rs || (nl == rs + 1 && *rs == '\r');
    if (!(blank && skip_blank)) starts.push_back(rs);
    rs = nl + 1;
  };
  const __m128i nlv = _mm_set1_epi8('\n');
  for (; c + 16 <= lim; c += 16) {
    unsigned m = (unsigned)_mm_movemask_epi8(_mm_cmpeq_epi8(_mm_loadu_si128((const __m128i *)c), nlv));
    while (m) {
      record_end(c + __builtin_ctz(m));
      m &= m - 1;
      if (starts.size() >= want) return rs;
    }
  }
  for (; c < lim; c++) {
    if (*c == '\n') {
      record_end(c);
      if (starts.size() >= want) return rs;
    }
  }
  return rs;
}

// Converts records [r0, r1) of the batch (starts[]) into rows r0.. of the
// column buffers; returns the first failing row (r1: none) and its error.
uint32_t parse_rows(const ysql_csv_reader *r, uint32_t r0, uint32_t r1, void *const *cols, uint8_t *const *nulls,
                    const char **err) {
  std::vector<std::string> rec;
  size_t nrec = 0;
  const int nc = (int)r->types.size();
  for (uint32_t i = r0; i < r1; i++) {
    parse_record(r->starts[i], r->end, rec, nrec);
    if ((int)nrec < r->nfields) {
      *err = "csv: record with fewer fields than the header";
      return i;
    }
    for (int c = 0; c < nc; c++) {
      if (!convert(r->types[c], rec[r->field[c]], cols[c], nulls ? nulls[c] : nullptr, i)) {
        *err = "csv: field does not convert to the column type";
        return i;
      }
    }
  }
  return r1;
}

// Parses up to batch_rows records into the given column buffers.  One pass
// over the bytes finds the record starts (skipping blank lines); then up to
// r->threads threads convert disjoint row ranges.
int parse_batch(ysql_csv_reader *r, void *const *cols, uint8_t *const *nulls, uint32_t *rows) {
  r->starts.clear();
  while (r->starts.size() < r->batch_rows && r->pos < r->end) {
    const char *q = r->quote_at_or_after(r->pos);
    if (q > r->pos) {  // no quote before q: newlines alone end records
      r->pos = quote_free_starts(r->pos, q, r->batch_rows, r->nfields > 1, r->starts);
      if (r->starts.size() >= r->batch_rows || r->pos >= r->end) break;
    }
    // the record at r->pos holds the quote at q, or ends the file unterminated
    bool blank = false;
    const char *rs = r->pos;
    r->pos = r->scan_record(rs, &blank);
    if (!(blank && r->nfields > 1)) r->starts.push_back(rs);
  }
  const uint32_t n = (uint32_t)r->starts.size();
  constexpr uint32_t kMinRowsPerThread = 16384;
  const uint32_t nt = std::max<uint32_t>(1, std::min<uint32_t>((uint32_t)r->threads, n / kMinRowsPerThread));
  std::vector<uint32_t> bad(nt, 0);
  std::vector<const char *> err(nt, nullptr);
  const uint32_t per = (n + nt - 1) / nt;
  auto work = [&](uint32_t t) {
    const uint32_t r0 = std::min(n, t * per), r1 = std::min(n, r0 + per);
    bad[t] = parse_rows(r, r0, r1, cols, nulls, &err[t]) < r1 ? 1 : 0;
  };
  if (nt == 1) {
    work(0);
  } else {
    std::vector<std::thread> pool;
    pool.reserve(nt - 1);
    for (uint32_t t = 1; t < nt; t++) pool.emplace_back(work, t);
    work(0);
    for (auto &th : pool) th.join();
  }
  for (uint32_t t = 0; t < nt; t++)
    if (bad[t]) {  // the error of the earliest failing row range
      ysql::set_last_error(err[t]);
      return YSQL_EINVAL;
    }
  *rows = n;
  return YSQL_OK;
}

void free_stage(Stage &s) {
  if (s.pending && s.done) (void)hipEventSynchronize(s.done);
  for (void *p : s.cols)
    if (p) (void)hipHostFree(p);
  for (uint8_t *p : s.nulls)
    if (p) (void)hipHostFree(p);
  if (s.done) (void)hipEventDestroy(s.done);
  s = Stage();
}

}  // namespace

extern "C" {

int ysql_csv_batch_rows(int ncols, const int *types, uint32_t *rows) {
  if (ncols < 1 || !types || !rows) return YSQL_EINVAL;
  size_t row_bytes = 0;
  for (int c = 0; c < ncols; c++) {
    const uint32_t b = type_bytes(types[c]);
    if (!b) return YSQL_EINVAL;
    row_bytes += b;
  }
  const size_t max_rows = kMaxBytesPerBatch / row_bytes;
  const uint32_t aligned = (uint32_t)(max_rows / kAlign) * kAlign;
  *rows = aligned > kAlign ? aligned : kAlign;
  return YSQL_OK;
}

int ysql_csv_open(const char *path, int ncols, const char *const *names, const int *types,
                  ysql_csv_reader **out) {
  if (!path || !names || !types || !out || ncols < 1) return YSQL_EINVAL;
  *out = nullptr;
  auto *r = new ysql_csv_reader();
  int rc = ysql_csv_batch_rows(ncols, types, &r->batch_rows);
  if (rc) {
    delete r;
    return rc;
  }
  r->fd = open(path, O_RDONLY);
  struct stat sb;
  if (r->fd < 0 || fstat(r->fd, &sb) != 0) {
    ysql::set_last_error("csv: cannot open the file");
    ysql_csv_close(r);
    return YSQL_EINVAL;
  }
  r->size = (size_t)sb.st_size;
  if (r->size == 0) {
    ysql::set_last_error("csv: no header row");
    ysql_csv_close(r);
    return YSQL_EINVAL;
  }
  void *m = mmap(nullptr, r->size, PROT_READ, MAP_PRIVATE, r->fd, 0);
  if (m == MAP_FAILED) {
    ysql::set_last_error("csv: cannot map the file");
    ysql_csv_close(r);
    return YSQL_EINVAL;
  }
  (void)madvise(m, r->size, MADV_SEQUENTIAL);
  r->data = (const char *)m;
  r->end = r->data + r->size;
  r->next_quote = nullptr;
  const unsigned hw = std::thread::hardware_concurrency();
  r->threads = (int)std::max(1u, std::min(hw, 16u));
  std::vector<std::string> header;
  size_t nh = 0;
  bool blank = false;
  r->pos = r->scan_record(r->data, &blank);
  parse_record(r->data, r->end, header, nh);
  header.resize(nh);
  r->nfields = (int)header.size();
  for (int c = 0; c < ncols; c++) {
    if (!names[c]) rc = YSQL_EINVAL;
    int idx = -1;
    for (int i = 0; names[c] && i < r->nfields; i++)
      if (header[i] == names[c]) {
        idx = i;
        break;
      }
    if (idx < 0) rc = YSQL_EINVAL;
    r->types.push_back(types[c]);
    r->bytes.push_back(type_bytes(types[c]));
    r->field.push_back(idx);
  }
  if (rc) {
    ysql::set_last_error("csv: a requested column is not in the header");
    ysql_csv_close(r);
    return rc;
  }
  *out = r;
  return YSQL_OK;
}

int ysql_csv_next_host(ysql_csv_reader *r, void *const *host_cols, uint8_t *const *host_nulls, uint32_t *rows) {
  if (!r || !host_cols || !rows) return YSQL_EINVAL;
  for (size_t c = 0; c < r->types.size(); c++)
    if (!host_cols[c]) return YSQL_EINVAL;
  return parse_batch(r, host_cols, host_nulls, rows);
}

int ysql_csv_next(ysql_csv_reader *r, void *const *dev_cols, uint8_t *const *dev_nulls, uint32_t *rows,
                  ysql_stream_t stream) {
  if (!r || !dev_cols || !rows) return YSQL_EINVAL;
  const size_t nc = r->types.size();
  for (size_t c = 0; c < nc; c++)
    if (!dev_cols[c]) return YSQL_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  Stage &st = r->st[r->cur];
  if (st.cols.empty()) {  // first use: pinned staging for one batch
    st.cols.assign(nc, nullptr);
    st.nulls.assign(nc, nullptr);
    for (size_t c = 0; c < nc; c++) {
      hipError_t e = hipHostMalloc(&st.cols[c], (size_t)r->batch_rows * r->bytes[c], hipHostMallocDefault);
      if (e == hipSuccess) e = hipHostMalloc((void **)&st.nulls[c], r->batch_rows, hipHostMallocDefault);
      if (e != hipSuccess) {
        std::string msg = std::string("csv: pinned staging allocation failed: ") + hipGetErrorString(e);
        ysql::set_last_error(msg.c_str());
        return YSQL_EHIP;
      }
    }
    if (hipEventCreateWithFlags(&st.done, hipEventDisableTiming) != hipSuccess) return YSQL_EHIP;
  }
  if (st.pending) {  // the copies queued from this staging area must have landed
    if (hipEventSynchronize(st.done) != hipSuccess) return YSQL_EHIP;
    st.pending = false;
  }
  int rc = parse_batch(r, st.cols.data(), st.nulls.data(), rows);
  if (rc) return rc;
  if (*rows == 0) return YSQL_OK;
  for (size_t c = 0; c < nc; c++) {
    if (hipMemcpyAsync(dev_cols[c], st.cols[c], (size_t)*rows * r->bytes[c], hipMemcpyHostToDevice, s) !=
        hipSuccess)
      return YSQL_EHIP;
    if (dev_nulls && dev_nulls[c] &&
        hipMemcpyAsync(dev_nulls[c], st.nulls[c], *rows, hipMemcpyHostToDevice, s) != hipSuccess)
      return YSQL_EHIP;
  }
  if (hipEventRecord(st.done, s) != hipSuccess) return YSQL_EHIP;
  st.pending = true;
  r->cur ^= 1;
  return YSQL_OK;
}

void ysql_csv_close(ysql_csv_reader *r) {
  if (!r) return;
  free_stage(r->st[0]);
  free_stage(r->st[1]);
  if (r->data) munmap((void *)r->data, r->size);
  if (r->fd >= 0) close(r->fd);
  delete r;
}

}  // extern "C"
