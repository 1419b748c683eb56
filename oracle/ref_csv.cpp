// ref_csv.cpp — TEST INFRASTRUCTURE ONLY (the checker, never the product).
//
// The reference's CSV scan (GetOperator::next, src/engine/operators/
// get_operator.cpp:56-109) on the CPU, built against the reference's own
// vendored csv-parser v2.3.0 (vendor/csv-parser/csv.hpp, included from where
// it lies in the reference tree by oracle/Makefile; nothing of it is copied
// here):
//   * the file is opened as `csv::CSVReader reader(path)` with the default
//     format, as get_operator.cpp:25 does (delimiter and header row guessed);
//   * each requested field is fetched by column name,
//     row[name].get<std::string>() (get_operator.cpp:68-72);
//   * conversions restate get_operator.cpp:77-87 with the same libstdc++
//     calls: empty -> NULL with INT_MIN / FLT_MIN / 0 / zero string;
//     INT std::stoi, FLOAT std::stof, DATETIME strptime("%Y-%m-%d %H:%M:%S")
//     + timegm with 0 on failure (data_type_utils.hpp:24-44), STRING the
//     first 255 bytes up to a NUL, NUL-padded to 256 (string_kernel.hpp:16-30).
// A conversion that throws in the reference (std::stoi / std::stof on text
// with no number, or out of range) is reported as an error here.
#include <climits>
#include <cfloat>
#include <cstdint>
#include <cstring>
#include <ctime>
#include <string>
#include <vector>

#include "csv.hpp"

namespace {

constexpr int kInt = 0, kFloat = 1, kDatetime = 2, kString = 3;
constexpr int kStrBytes = 256;

int64_t datetime_of(const std::string &s) {
  std::tm tm = {};
  if (strptime(s.c_str(), "%Y-%m-%d %H:%M:%S", &tm) == nullptr) return 0;
  return (int64_t)timegm(&tm);
}

void set_err(char *err, int errlen, const std::string &msg) {
  if (err && errlen > 0) {
    strncpy(err, msg.c_str(), (size_t)errlen - 1);
    err[errlen - 1] = 0;
  }
}

}  // namespace

// Reads up to max_rows rows of the named columns into outs[c] (int32,
// float32, int64 or 256-byte strings by types[c]) and nulls[c] (1 byte per
// row).  Returns the row count, or -1 with `err` set.
extern "C" int64_t ref_csv_read(const char *path, int ncols, const char *const *names, const int *types,
                                uint64_t max_rows, void *const *outs, uint8_t *const *nulls, char *err, int errlen) {
  try {
    csv::CSVReader reader(path);
    csv::CSVRow row;
    uint64_t r = 0;
    std::vector<std::string> vals((size_t)ncols);
    while (r < max_rows && reader.read_row(row)) {
      for (int c = 0; c < ncols; c++) vals[(size_t)c] = row[names[c]].get<std::string>();
      for (int c = 0; c < ncols; c++) {
        const std::string &s = vals[(size_t)c];
        nulls[c][r] = s.empty() ? 1 : 0;
        switch (types[c]) {
          case kInt: ((int32_t *)outs[c])[r] = s.empty() ? INT_MIN : std::stoi(s); break;
          case kFloat: ((float *)outs[c])[r] = s.empty() ? FLT_MIN : std::stof(s); break;
          case kDatetime: ((int64_t *)outs[c])[r] = s.empty() ? 0 : datetime_of(s); break;
          case kString: {
            char *d = (char *)outs[c] + r * kStrBytes;
            memset(d, 0, kStrBytes);
            size_t i = 0;
            while (i < kStrBytes - 1 && i < s.size() && s[i] != '\0') {
              d[i] = s[i];
              i++;
            }
            break;
          }
          default: set_err(err, errlen, "unknown column type"); return -1;
        }
      }
      r++;
    }
    return (int64_t)r;
  } catch (const std::exception &e) {
    set_err(err, errlen, e.what());
    return -1;
  }
}

// Every row the reader hands out, as raw fields: the guessed delimiter, the
// header names, then each row.  Serialised into `out` as little-endian
// uint32 counts and lengths: delimiter byte, nnames, (len, bytes)..., then
// per row nfields, (len, bytes)...  Returns the bytes written (or needed,
// when above `cap`), -1 with `err` set on an exception.
extern "C" int64_t ref_csv_dump(const char *path, uint8_t *out, uint64_t cap, char *err, int errlen) {
  try {
    csv::CSVReader reader(path);
    std::string buf;
    auto put32 = [&](uint32_t v) { buf.append((const char *)&v, 4); };
    auto put_str = [&](const std::string &s) {
      put32((uint32_t)s.size());
      buf += s;
    };
    put32((uint32_t)(unsigned char)reader.get_format().get_delim());
    const std::vector<std::string> names = reader.get_col_names();
    put32((uint32_t)names.size());
    for (const auto &n : names) put_str(n);
    csv::CSVRow row;
    while (reader.read_row(row)) {
      put32((uint32_t)row.size());
      for (size_t i = 0; i < row.size(); i++) put_str(row[i].get<std::string>());
    }
    if (buf.size() <= cap) memcpy(out, buf.data(), buf.size());
    return (int64_t)buf.size();
  } catch (const std::exception &e) {
    set_err(err, errlen, e.what());
    return -1;
  }
}
