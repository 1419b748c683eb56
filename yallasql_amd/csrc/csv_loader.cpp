// csv_loader.cpp — CSV file -> column batches (ysql_csv.h): the feeder of
// the scan path, replacing the reference's GetOperator
// (src/engine/operators/get_operator.cpp:16-148).
//
// The file is read in 4 MiB chunks and tokenised by hand (RFC 4180 quotes);
// each requested column is converted with the reference's rules into one of
// two pinned staging buffers, whose host-to-device copies are queued on the
// caller's stream while the next batch is parsed.
#include <hip/hip_runtime.h>

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
#include <vector>

#include "common.hpp"
#include "ysql_csv.h"

namespace {

constexpr size_t kMaxBytesPerBatch = 4u << 20;  // config.hpp: MAX_BYTES_PER_BATCH
constexpr uint32_t kAlign = 256;                // config.hpp: GPU_ALIGNMENT = BLOCK_DIM
constexpr size_t kChunk = 4u << 20;             // file read size

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

struct ysql_csv_reader {
  FILE *f = nullptr;
  std::vector<char> buf;
  size_t pos = 0, len = 0;
  bool eof = false;
  std::vector<int> types;
  std::vector<uint32_t> bytes;
  std::vector<int> field;  // CSV field index of each requested column
  int nfields = 0;
  uint32_t batch_rows = 0;
  Stage st[2];
  int cur = 0;
  std::vector<std::string> rec;  // the current record's fields: rec[0 .. nrec)
  size_t nrec = 0;

  // refills the buffer; false at the end of the file
  bool refill() {
    if (eof) return false;
    len = fread(buf.data(), 1, buf.size(), f);
    pos = 0;
    if (len == 0) eof = true;
    return len > 0;
  }
  // next byte of the file, -1 at the end
  int get() {
    if (pos == len && !refill()) return -1;
    return (unsigned char)buf[pos++];
  }
  int peek() {
    if (pos == len && !refill()) return -1;
    return (unsigned char)buf[pos];
  }

  // Reads one record into rec; false at the end of the file.  RFC 4180:
  // fields separated by ',', records by "\n" or "\r\n", a field in quotes may
  // hold separators and newlines, "" inside quotes is one quote; a quote
  // that does not open a field is an ordinary character.  Runs of ordinary
  // characters are copied in bulk.
  bool read_record() {
    nrec = 0;
    if (pos == len && !refill()) return false;
    std::string *fld = next_field();
    bool quoted = false;
    while (true) {
      if (pos == len && !refill()) break;  // end of file ends the record
      const char *p = buf.data() + pos, *e = buf.data() + len, *q = p;
      if (!quoted) {
        while (q < e && *q != ',' && *q != '\n' && *q != '\r' && *q != '"') q++;
        fld->append(p, (size_t)(q - p));
        pos += (size_t)(q - p);
        if (pos == len) continue;
        const char c = buf[pos++];
        if (c == ',') {
          fld = next_field();
        } else if (c == '\n') {
          break;
        } else if (c == '\r') {
          if (peek() == '\n') {
            pos++;
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
        pos += (size_t)(q - p);
        if (pos == len) continue;
        pos++;  // the quote
        if (peek() == '"') {
          pos++;
          fld->push_back('"');
        } else {
          quoted = false;
        }
      }
    }
    return true;
  }
  // the next field of the current record (strings are reused across records)
  std::string *next_field() {
    if (nrec == rec.size()) rec.emplace_back();
    std::string *f = &rec[nrec++];
    f->clear();
    return f;
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

// Parses up to batch_rows records into the given column buffers.
int parse_batch(ysql_csv_reader *r, void *const *cols, uint8_t *const *nulls, uint32_t *rows) {
  uint32_t n = 0;
  const int nc = (int)r->types.size();
  while (n < r->batch_rows && r->read_record()) {
    if (r->nrec == 1 && r->rec[0].empty() && r->nfields > 1) continue;  // blank line
    if ((int)r->nrec < r->nfields) {
      ysql::set_last_error("csv: record with fewer fields than the header");
      return YSQL_EINVAL;
    }
    for (int c = 0; c < nc; c++) {
      if (!convert(r->types[c], r->rec[r->field[c]], cols[c], nulls ? nulls[c] : nullptr, n)) {
        ysql::set_last_error("csv: field does not convert to the column type");
        return YSQL_EINVAL;
      }
    }
    n++;
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
  r->f = fopen(path, "rb");
  if (!r->f) {
    ysql::set_last_error("csv: cannot open the file");
    delete r;
    return YSQL_EINVAL;
  }
  r->buf.resize(kChunk);
  if (!r->read_record()) {
    ysql::set_last_error("csv: no header row");
    fclose(r->f);
    delete r;
    return YSQL_EINVAL;
  }
  const std::vector<std::string> header(r->rec.begin(), r->rec.begin() + r->nrec);
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
    fclose(r->f);
    delete r;
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
  if (r->f) fclose(r->f);
  delete r;
}

}  // extern "C"
