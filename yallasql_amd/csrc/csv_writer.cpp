// csv_writer.cpp — result batches -> CSV file (ysql_csv.h), replacing the
// reference's CsvWriter (include/engine/operators/csv_writer.hpp:12-160),
// which ExecutorEngine feeds with every result batch
// (src/engine/executor_engine.cpp:19-58).
//
// Format (csv_writer.hpp:24-77): the column names joined by ',' on the first
// line (written with the first batch); then per row, fields joined by ','
// and a '\n':
//   NULL (null byte set)  ""
//   STRING                "<bytes up to the first NUL>"   (no escaping)
//   INT                   std::to_chars decimal
//   FLOAT                 std::to_chars fixed, 4 decimals (correctly rounded
//                         from the float's exact value)
//   DATETIME              "<strftime %Y-%m-%d %H:%M:%S of gmtime_r>"
// Two defects of the reference are not reproduced: its 32-byte buffer makes
// to_chars fail for floats whose fixed form needs more than 32 characters
// (magnitudes from about 1e27 up), after which it appends the buffer's
// unspecified contents; a buffer of 64 here always holds the value.  And a
// string with no NUL in its 256 bytes is read past its end there
// (operator+= on a char*); here it is cut at 256 bytes.
//
// Rows are formatted on up to 16 host threads (row ranges) and written in
// order.  The device form copies a batch into pinned staging on the caller's
// stream first.
#include <hip/hip_runtime.h>

#include <charconv>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <string>
#include <thread>
#include <vector>

#include "common.hpp"
#include "ysql_csv.h"

struct ysql_csv_writer {
  FILE *f = nullptr;
  std::vector<std::string> names;
  std::vector<int> types;
  bool header_done = false;
  bool failed = false;
  int threads = 1;
  std::vector<std::string> parts;  // one formatted row range per thread
  // pinned staging of the device form
  std::vector<void *> hcols;
  std::vector<uint8_t *> hnulls;
  uint64_t cap = 0;
};

namespace {

uint32_t elem_bytes(int t) {
  switch (t) {
    case YSQL_INT: case YSQL_FLOAT: return 4;
    case YSQL_DATETIME: return 8;
    case YSQL_STRING: return YSQL_STRING_BYTES;
    default: return 0;
  }
}

void append_field(std::string &out, int type, const void *col, uint64_t row) {
  char buf[64];
  switch (type) {
    case YSQL_STRING: {
      const char *s = (const char *)col + row * YSQL_STRING_BYTES;
      const void *z = memchr(s, 0, YSQL_STRING_BYTES);
      out += '"';
      out.append(s, z ? (size_t)((const char *)z - s) : (size_t)YSQL_STRING_BYTES);
      out += '"';
      return;
    }
    case YSQL_INT: {
      int32_t v;
      memcpy(&v, (const char *)col + row * 4, 4);
      const auto r = std::to_chars(buf, buf + sizeof(buf), v);
      out.append(buf, (size_t)(r.ptr - buf));
      return;
    }
    case YSQL_FLOAT: {
      float v;
      memcpy(&v, (const char *)col + row * 4, 4);
      const auto r = std::to_chars(buf, buf + sizeof(buf), v, std::chars_format::fixed, 4);
      out.append(buf, (size_t)(r.ptr - buf));
      return;
    }
    default: {  // DATETIME: getDateTimeStr (data_type_utils.hpp:46-58)
      int64_t v;
      memcpy(&v, (const char *)col + row * 8, 8);
      const time_t t = (time_t)v;
      struct tm tmv;
      memset(&tmv, 0, sizeof(tmv));
      size_t k = 0;
      if (gmtime_r(&t, &tmv)) k = strftime(buf, 26, "%Y-%m-%d %H:%M:%S", &tmv);
      out += '"';
      out.append(buf, k);
      out += '"';
      return;
    }
  }
}

void format_rows(const ysql_csv_writer *w, const void *const *cols, const uint8_t *const *nulls, uint64_t r0,
                 uint64_t r1, std::string &out) {
  out.clear();
  const size_t nc = w->types.size();
  for (uint64_t r = r0; r < r1; r++) {
    for (size_t c = 0; c < nc; c++) {
      if (nulls && nulls[c] && nulls[c][r])
        out += "\"\"";
      else
        append_field(out, w->types[c], cols[c], r);
      if (c + 1 < nc) out += ',';
    }
    out += '\n';
  }
}

int write_out(ysql_csv_writer *w, const std::string &s) {
  if (!s.empty() && fwrite(s.data(), 1, s.size(), w->f) != s.size()) {
    w->failed = true;
    ysql::set_last_error("csv writer: write failed");
    return YSQL_EINVAL;
  }
  return YSQL_OK;
}

int write_rows(ysql_csv_writer *w, const void *const *cols, const uint8_t *const *nulls, uint64_t rows) {
  if (!w->header_done) {
    std::string h;
    for (size_t c = 0; c < w->names.size(); c++) {
      h += w->names[c];
      if (c + 1 < w->names.size()) h += ',';
    }
    h += '\n';
    w->header_done = true;
    if (int rc = write_out(w, h)) return rc;
  }
  constexpr uint64_t kMinRowsPerThread = 16384;
  const uint64_t nt = std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)w->threads, rows / kMinRowsPerThread));
  w->parts.resize(nt);
  const uint64_t per = (rows + nt - 1) / nt;
  auto work = [&](uint64_t t) {
    const uint64_t r0 = std::min(rows, t * per), r1 = std::min(rows, r0 + per);
    format_rows(w, cols, nulls, r0, r1, w->parts[t]);
  };
  if (nt == 1) {
    work(0);
  } else {
    std::vector<std::thread> pool;
    for (uint64_t t = 1; t < nt; t++) pool.emplace_back(work, t);
    work(0);
    for (auto &th : pool) th.join();
  }
  for (uint64_t t = 0; t < nt; t++)
    if (int rc = write_out(w, w->parts[t])) return rc;
  return YSQL_OK;
}

void free_staging(ysql_csv_writer *w) {
  for (void *p : w->hcols)
    if (p) (void)hipHostFree(p);
  for (uint8_t *p : w->hnulls)
    if (p) (void)hipHostFree(p);
  w->hcols.clear();
  w->hnulls.clear();
  w->cap = 0;
}

}  // namespace

extern "C" {

int ysql_csv_writer_open(const char *path, int ncols, const char *const *names, const int *types,
                         ysql_csv_writer **out) {
  if (!path || ncols < 1 || !names || !types || !out) return YSQL_EINVAL;
  *out = nullptr;
  auto *w = new ysql_csv_writer();
  for (int c = 0; c < ncols; c++) {
    if (!names[c] || !elem_bytes(types[c])) {
      delete w;
      return YSQL_EINVAL;
    }
    w->names.emplace_back(names[c]);
    w->types.push_back(types[c]);
  }
  w->f = fopen(path, "wb");
  if (!w->f) {
    ysql::set_last_error("csv writer: cannot open the file");
    delete w;
    return YSQL_EINVAL;
  }
  w->threads = (int)std::max(1u, std::min(std::thread::hardware_concurrency(), 16u));
  *out = w;
  return YSQL_OK;
}

int ysql_csv_writer_write_host(ysql_csv_writer *w, const void *const *host_cols, const uint8_t *const *host_nulls,
                               uint64_t rows) {
  if (!w || !host_cols) return YSQL_EINVAL;
  for (size_t c = 0; c < w->types.size(); c++)
    if (rows && !host_cols[c]) return YSQL_EINVAL;
  return write_rows(w, host_cols, host_nulls, rows);
}

int ysql_csv_writer_write(ysql_csv_writer *w, const void *const *dev_cols, const uint8_t *const *dev_nulls,
                          uint64_t rows, ysql_stream_t stream) {
  if (!w || !dev_cols) return YSQL_EINVAL;
  const size_t nc = w->types.size();
  for (size_t c = 0; c < nc; c++)
    if (rows && !dev_cols[c]) return YSQL_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  if (rows > w->cap) {  // grow the pinned staging
    YSQL_HIP_TRY(hipStreamSynchronize(s));
    free_staging(w);
    w->hcols.assign(nc, nullptr);
    w->hnulls.assign(nc, nullptr);
    for (size_t c = 0; c < nc; c++) {
      hipError_t e = hipHostMalloc(&w->hcols[c], rows * elem_bytes(w->types[c]), hipHostMallocDefault);
      if (e == hipSuccess) e = hipHostMalloc((void **)&w->hnulls[c], rows, hipHostMallocDefault);
      if (e != hipSuccess) {
        free_staging(w);
        ysql::set_last_error("csv writer: pinned staging allocation failed");
        return YSQL_EHIP;
      }
    }
    w->cap = rows;
  }
  std::vector<const uint8_t *> nl(nc, nullptr);
  for (size_t c = 0; c < nc && rows; c++) {
    YSQL_HIP_TRY(hipMemcpyAsync(w->hcols[c], dev_cols[c], rows * elem_bytes(w->types[c]), hipMemcpyDeviceToHost, s));
    if (dev_nulls && dev_nulls[c]) {
      YSQL_HIP_TRY(hipMemcpyAsync(w->hnulls[c], dev_nulls[c], rows, hipMemcpyDeviceToHost, s));
      nl[c] = w->hnulls[c];
    }
  }
  YSQL_HIP_TRY(hipStreamSynchronize(s));
  return write_rows(w, (const void *const *)w->hcols.data(), nl.data(), rows);
}

int ysql_csv_writer_close(ysql_csv_writer *w) {
  if (!w) return YSQL_EINVAL;
  int rc = w->failed ? YSQL_EINVAL : YSQL_OK;
  if (w->f && fclose(w->f) != 0) rc = YSQL_EINVAL;
  free_staging(w);
  delete w;
  return rc;
}

}  // extern "C"
