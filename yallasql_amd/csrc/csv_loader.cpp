// csv_loader.cpp — CSV file -> column batches (ysql_csv.h): the feeder of
// the scan path, replacing the reference's GetOperator
// (src/engine/operators/get_operator.cpp:16-148), which reads the file
// through csv::CSVReader (vendored csv-parser v2.3.0,
// vendor/csv-parser/csv.hpp) with its default, guessing format.
//
// The tokenising rules are that library's, restated here:
//   * the delimiter is guessed among ',', '|', '\t', ';', '^' from the first
//     500,000 bytes (up to a NUL byte), and so is the header row
//     (csv.hpp:5007-5014 guess_csv, :7368-7443 calculate_score and
//     _guess_format): for each candidate the head is split into rows and
//     scored as the largest (row length x rows of that length), ties going
//     to the length the library's hash map iterates first; the first
//     candidate with the strictly highest score wins, and the header is the
//     first row of the winning length;
//   * any run of '\r' and '\n' ends a record, so blank lines vanish
//     (csv.hpp:7080-7092); a newline at the very start of the data is an
//     empty one-field record;
//   * a '"' opens a quoted field only as the first byte of a field; inside
//     one, '"' closes it when a delimiter or newline follows (or at the end
//     of the file), "" is kept as a pair, and any other '"' is a literal
//     (csv.hpp:7099-7135); a field that held a "" pair drops every second
//     quote character of its text (csv.hpp:7764-7791);
//   * records whose field count differs from the header's are skipped
//     (VariableColumnPolicy::IGNORE_ROW, csv.hpp:7616-7625); a UTF-8 BOM at
//     the start is skipped (csv.hpp:7150-7162); a column name that repeats
//     in the header means its last occurrence (csv.hpp:7217-7222);
//   * the library parses the file in 10,000,000-byte chunks, each starting
//     at the first record the previous one left incomplete, or at the chunk
//     end when a newline run reaches it (csv.hpp:7174-7201); a run that
//     crosses a chunk end therefore yields one more empty one-field record
//     at that offset (kept as a NULL row by a one-column table, skipped
//     otherwise), and this reader yields it at the same offset.
//
// The file is mapped; one pass over a batch's bytes finds its record starts
// (16 bytes at a time through quote-free text), then up to 16 threads
// tokenise and convert disjoint row ranges with the reference's rules into
// one of two pinned staging buffers, whose host-to-device copies are queued
// on the caller's stream while the next batch is parsed.
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
constexpr size_t kHeadBytes = 500000;           // csv.hpp:6960 get_csv_head
constexpr size_t kChunkBytes = 10000000;        // csv.hpp:4792 ITERATION_CHUNK_SIZE
constexpr char kDelims[5] = {',', '|', '\t', ';', '^'};

uint32_t type_bytes(int t) {
  switch (t) {
    case YSQL_INT: return 4;
    case YSQL_FLOAT: return 4;
    case YSQL_DATETIME: return 8;
    case YSQL_STRING: return YSQL_STRING_BYTES;
    default: return 0;
  }
}

inline bool is_nl(char c) { return c == '\n' || c == '\r'; }

// First '\r' or '\n' in [p, e) (e: none), 16 bytes at a time.
const char *find_nl(const char *p, const char *e) {
  const __m128i n = _mm_set1_epi8('\n'), r = _mm_set1_epi8('\r');
  for (; p + 16 <= e; p += 16) {
    const __m128i v = _mm_loadu_si128((const __m128i *)p);
    const unsigned m = (unsigned)_mm_movemask_epi8(_mm_or_si128(_mm_cmpeq_epi8(v, n), _mm_cmpeq_epi8(v, r)));
    if (m) return p + __builtin_ctz(m);
  }
  for (; p < e; p++)
    if (is_nl(*p)) return p;
  return e;
}

// A field's text once it has held a "" pair: every second quote character
// is dropped, counting all of the field's quotes (csv.hpp:7774-7789).
void drop_pair_quotes(std::string &f) {
  size_t o = 0;
  bool prev = false;
  for (size_t i = 0; i < f.size(); i++) {
    if (f[i] == '"') {
      if (prev) {
        prev = false;
        continue;
      }
      prev = true;
    }
    f[o++] = f[i];
  }
  f.resize(o);
}

// One staging area: a pinned buffer per column and per null column, and the
// event recorded after the copies queued from it.  In the asynchronous mode
// it also owns the batch's device buffers, and `used` is recorded on the
// caller's stream at the call after the one that handed the batch out (its
// work on the batch is queued by then); a copy into the area waits on it.
struct Stage {
  std::vector<void *> cols;
  std::vector<uint8_t *> nulls;
  std::vector<void *> dcols;
  std::vector<uint8_t *> dnulls;
  hipEvent_t done = nullptr;
  hipEvent_t used = nullptr;
  bool pending = false;
  bool used_pending = false;
};

}  // namespace

// Reads the record starting at p into rec[0 .. nrec) with delimiter d, by
// the rules above; the end of the data ends the record.  Returns the start
// of the next record (past the newline run).  Runs of ordinary characters
// are copied in bulk.
static const char *parse_record(const char *p, const char *e, char d, std::vector<std::string> &rec,
                                size_t &nrec) {
  nrec = 0;
  bool pairs = false;  // the current field held a "" pair
  auto next_field = [&]() {
    if (nrec == rec.size()) rec.emplace_back();
    std::string *f = &rec[nrec++];
    f->clear();
    pairs = false;
    return f;
  };
  auto finish = [&](std::string *f) {
    if (pairs) drop_pair_quotes(*f);
  };
  std::string *fld = next_field();
  bool quoted = false;
  while (p < e) {
    const char *q = p;
    if (!quoted) {
      while (q < e && *q != d && *q != '\n' && *q != '\r' && *q != '"') q++;
      fld->append(p, (size_t)(q - p));
      p = q;
      if (p == e) break;
      const char c = *p++;
      if (c == d) {
        finish(fld);
        fld = next_field();
      } else if (is_nl(c)) {
        while (p < e && is_nl(*p)) p++;
        finish(fld);
        return p;
      } else if (fld->empty()) {  // '"' as a field's first byte opens it
        quoted = true;
      } else {
        fld->push_back('"');
      }
    } else {
      while (q < e && *q != '"') q++;
      fld->append(p, (size_t)(q - p));
      p = q;
      if (p == e) break;
      if (p + 1 == e) {  // a closing quote ends the data
        p++;
        break;
      }
      const char n = p[1];
      if (n == d || is_nl(n)) {  // closes the field
        quoted = false;
        p++;
      } else if (n == '"') {
        fld->append(2, '"');
        pairs = true;
        p += 2;
      } else {
        fld->push_back('"');
        p++;
      }
    }
  }
  finish(fld);
  return p;
}

// Iteration order of the row-length tally csv-parser keeps in a
// std::unordered_map<size_t, size_t> built from {{0, 0}} (csv.hpp:7370-7414):
// among row lengths of equal score the first one iterated wins.  libstdc++
// keeps one singly linked list of nodes, each bucket's nodes in one run; a
// node whose bucket is empty goes to the list's front, any other to the
// front of its bucket's run (hashtable.h _M_insert_bucket_begin); growing
// re-threads the list front to back by the same rule (_M_rehash_aux).  Such
// a map holds 13 buckets and grows to the next count below when an insert
// would exceed the count (counts read off such a map).
class TallyOrder {
 public:
  TallyOrder() : bucket_(13, kEmpty) { add(0); }
  // The length is new to the tally.
  void add(size_t len) {
    if (key_.size() + 1 > bucket_.size()) grow();
    key_.push_back(len);
    nxt_.push_back(-1);
    link((int)key_.size() - 1);
  }
  // Lengths in iteration order.
  std::vector<size_t> order() const {
    std::vector<size_t> out;
    for (int n = head_; n >= 0; n = nxt_[n]) out.push_back(key_[n]);
    return out;
  }

 private:
  static constexpr int kEmpty = -2, kHead = -1;  // bucket: no node / the node before is the list head
  std::vector<size_t> key_;
  std::vector<int> nxt_;
  std::vector<int> bucket_;  // node before the bucket's first node
  int head_ = -1;

  int &next_of(int n) { return n == kHead ? head_ : nxt_[n]; }
  void link(int n) {
    const size_t b = key_[n] % bucket_.size();
    if (bucket_[b] != kEmpty) {
      nxt_[n] = next_of(bucket_[b]);
      next_of(bucket_[b]) = n;
    } else {
      nxt_[n] = head_;
      head_ = n;
      if (nxt_[n] >= 0) bucket_[key_[nxt_[n]] % bucket_.size()] = n;
      bucket_[b] = kHead;
    }
  }
  void grow() {
    static const size_t kCounts[] = {29, 59, 127, 257, 541, 1109, 2357, 5087, 10273, 20753};
    size_t nb = bucket_.size() * 2 + 1;
    for (size_t c : kCounts)
      if (c > bucket_.size()) {
        nb = c;
        break;
      }
    std::vector<int> nbk(nb, kEmpty);
    int p = head_;
    head_ = -1;
    size_t first_bucket = 0;
    while (p >= 0) {
      const int next = nxt_[p];
      const size_t b = key_[p] % nb;
      if (nbk[b] == kEmpty) {
        nxt_[p] = head_;
        head_ = p;
        nbk[b] = kHead;
        if (nxt_[p] >= 0) nbk[first_bucket] = p;
        first_bucket = b;
      } else {
        int &before = nbk[b] == kHead ? head_ : nxt_[nbk[b]];
        nxt_[p] = before;
        before = p;
      }
      p = next;
    }
    bucket_.swap(nbk);
  }
};

// The delimiter and header row csv::CSVReader guesses for this data
// (csv.hpp:7368-7443): `head` is the first 500,000 bytes up to a NUL, less a
// UTF-8 BOM.
static void guess_format(const char *head, const char *e, char *delim, size_t *header_row) {
  size_t best = 0;
  *delim = kDelims[0];
  *header_row = 0;
  std::vector<std::string> rec;
  std::vector<size_t> count, first;  // per row length
  for (char d : kDelims) {
    count.assign(1, 0);
    first.assign(1, 0);
    TallyOrder tally;
    size_t row = 0, nrec = 0;
    for (const char *p = head; p < e; row++) {
      p = parse_record(p, e, d, rec, nrec);
      if (nrec >= count.size()) {
        count.resize(nrec + 1, 0);
        first.resize(nrec + 1, 0);
      }
      if (count[nrec]++ == 0) {
        first[nrec] = row;
        tally.add(nrec);
      }
    }
    size_t score = 0, hdr = 0;
    for (size_t len : tally.order())
      if (len * count[len] > score) {
        score = len * count[len];
        hdr = first[len];
      }
    if (score > best) {
      best = score;
      *delim = d;
      *header_row = hdr;
    }
  }
}

struct ysql_csv_reader {
  int fd = -1;
  const char *data = nullptr;  // the whole file, mapped read-only
  size_t size = 0;
  const char *pos = nullptr, *end = nullptr;
  const char *next_quote = nullptr;  // first '"' at or after the scan position (end: none)
  char delim = ',';
  size_t chunk_end = kChunkBytes;  // file offset where the library's current parse chunk ends
  const char *pending = nullptr;   // a chunk-end record not yet handed out
  std::vector<int> types;
  std::vector<uint32_t> bytes;
  std::vector<int> field;  // CSV field index of each requested column
  int nfields = 0;
  uint32_t batch_rows = 0;
  int threads = 1;
  std::vector<const char *> starts;  // record starts of the batch being parsed
  std::vector<uint8_t> keep;         // per record of the batch: its field count matches the header
  std::vector<Stage> st = std::vector<Stage>(2);
  int cur = 0;
  hipStream_t copy_stream = nullptr;  // asynchronous mode: the reader's copy stream
  int handed = -1;                    // asynchronous mode: the area handed out by the last call

  const char *quote_at_or_after(const char *p) {
    if (next_quote < p) {
      const void *q = memchr(p, '"', (size_t)(end - p));
      next_quote = q ? (const char *)q : end;
    }
    return next_quote;
  }

  // The record [rs, a) ended by the newline run [a, b) (a == b == end: the
  // file ends it) is handed out, with the chunk bookkeeping of the library:
  // a record whose newline lies at or past the chunk end starts the next
  // chunk; a run reaching the chunk end starts the next chunk there, and one
  // crossing it makes an empty record at that offset.
  void end_record(const char *rs, const char *a, const char *b, size_t want) {
    const size_t oa = (size_t)(a - data), ob = (size_t)(b - data);
    if (oa >= chunk_end) chunk_end = (size_t)(rs - data) + kChunkBytes;
    starts.push_back(rs);
    if (oa < chunk_end && ob >= chunk_end) {
      if (ob > chunk_end) {
        const char *extra = data + chunk_end;
        if (starts.size() < want)
          starts.push_back(extra);
        else
          pending = extra;
      }
      chunk_end += kChunkBytes;
    }
  }

  // Record starts in quote-free text [p, lim), p a record start: every
  // record ended there is handed out until `want` are held; returns the
  // start of the first record not handed out.  Newline runs are found as
  // the edges of the 16-byte newline masks.
  const char *quote_free_records(const char *p, const char *lim, size_t want) {
    const char *rs = p, *a = nullptr;
    bool in_run = false;
    const __m128i n = _mm_set1_epi8('\n'), r = _mm_set1_epi8('\r');
    unsigned carry = 0;
    const char *c = p;
    for (; c + 16 <= lim; c += 16) {
      const __m128i v = _mm_loadu_si128((const __m128i *)c);
      const unsigned m = (unsigned)_mm_movemask_epi8(_mm_or_si128(_mm_cmpeq_epi8(v, n), _mm_cmpeq_epi8(v, r)));
      unsigned t = (m ^ ((m << 1) | carry)) & 0xffffu;  // bytes whose newline-ness differs from the previous one
      carry = m >> 15;
      while (t) {
        const int i = __builtin_ctz(t);
        t &= t - 1;
        if ((m >> i) & 1) {  // a run starts
          a = c + i;
          in_run = true;
        } else {  // a run ends: the next record starts here
          end_record(rs, a, c + i, want);
          rs = c + i;
          in_run = false;
          if (starts.size() >= want) return rs;
        }
      }
    }
    for (; c < lim; c++) {
      const bool nl = is_nl(*c);
      if (nl && !in_run) {
        a = c;
        in_run = true;
      } else if (!nl && in_run) {
        end_record(rs, a, c, want);
        rs = c;
        in_run = false;
        if (starts.size() >= want) return rs;
      }
    }
    if (in_run) {  // lim (a quote or the end) is not a newline: the run ends there
      end_record(rs, a, lim, want);
      rs = lim;
    }
    return rs;
  }

  // The end of the record starting at rs by parse_record's rules: *a = its
  // newline run's first byte (end: the file ends it); returns the run's end.
  const char *scan_record(const char *rs, const char **a) {
    const char *p = rs;
    while (true) {
      const char *nl = find_nl(p, end);
      const char *q = quote_at_or_after(p);
      while (q < nl && !(q == rs || q[-1] == delim)) q = quote_at_or_after(q + 1);  // literal quotes
      if (q >= nl) {
        *a = nl;
        const char *b = nl;
        while (b < end && is_nl(*b)) b++;
        return b;
      }
      p = q + 1;  // inside a quoted field: find the quote that closes it
      while (true) {
        const char *c = quote_at_or_after(p);
        if (c + 1 >= end) {  // no closing quote, or it ends the file
          *a = end;
          return end;
        }
        const char n = c[1];
        if (n == delim || is_nl(n)) {
          p = c + 1;
          break;
        }
        p = c + (n == '"' ? 2 : 1);
      }
    }
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

// Converts records [r0, r1) of the batch (starts[]) into rows row0 + r0..
// of the column buffers; a record whose field count is not the header's is
// marked skipped (keep[i] = 0) and its row left as it was.  Returns the
// first failing record (r1: none) and its error.
uint32_t parse_rows(ysql_csv_reader *r, uint32_t r0, uint32_t r1, uint32_t row0, void *const *cols,
                    uint8_t *const *nulls, const char **err) {
  std::vector<std::string> rec;
  size_t nrec = 0;
  const int nc = (int)r->types.size();
  for (uint32_t i = r0; i < r1; i++) {
    parse_record(r->starts[i], r->end, r->delim, rec, nrec);
    r->keep[i] = (int)nrec == r->nfields;
    if (!r->keep[i]) continue;
    for (int c = 0; c < nc; c++) {
      if (!convert(r->types[c], rec[r->field[c]], cols[c], nulls ? nulls[c] : nullptr, row0 + i)) {
        *err = "csv: field does not convert to the column type";
        return i;
      }
    }
  }
  return r1;
}

// Moves the kept rows of records [0, n) (rows row0..) down over the skipped
// ones; returns how many were kept.
uint32_t drop_skipped(ysql_csv_reader *r, uint32_t n, uint32_t row0, void *const *cols, uint8_t *const *nulls) {
  uint32_t o = 0;
  for (uint32_t i = 0; i < n;) {
    if (!r->keep[i]) {
      i++;
      continue;
    }
    uint32_t j = i;
    while (j < n && r->keep[j]) j++;
    if (o != i) {
      for (size_t c = 0; c < r->types.size(); c++) {
        const size_t b = r->bytes[c];
        memmove((char *)cols[c] + (size_t)(row0 + o) * b, (char *)cols[c] + (size_t)(row0 + i) * b, (j - i) * b);
        if (nulls && nulls[c]) memmove(nulls[c] + row0 + o, nulls[c] + row0 + i, j - i);
      }
    }
    o += j - i;
    i = j;
  }
  return o;
}

// Parses up to batch_rows rows into the given column buffers.  One pass
// over the bytes finds record starts; up to r->threads threads convert
// disjoint record ranges; skipped records are squeezed out and more records
// read until the batch is full or the file ends.
int parse_batch(ysql_csv_reader *r, void *const *cols, uint8_t *const *nulls, uint32_t *rows) {
  uint32_t have = 0;
  while (have < r->batch_rows && (r->pos < r->end || r->pending)) {
    const size_t want = r->batch_rows - have;
    r->starts.clear();
    if (r->pending) {
      r->starts.push_back(r->pending);
      r->pending = nullptr;
    }
    while (r->starts.size() < want && r->pos < r->end) {
      const char *q = r->quote_at_or_after(r->pos);
      if (q > r->pos) {  // no quote before q: newline runs alone end records
        r->pos = r->quote_free_records(r->pos, q, want);
        if (r->starts.size() >= want || r->pos >= r->end) break;
      }
      // the record at r->pos holds the quote at q
      const char *rs = r->pos, *a = nullptr;
      r->pos = r->scan_record(rs, &a);
      r->end_record(rs, a, r->pos, want);
    }
    const uint32_t n = (uint32_t)r->starts.size();
    if (n == 0) break;
    r->keep.assign(n, 1);
    constexpr uint32_t kMinRowsPerThread = 16384;
    const uint32_t nt = std::max<uint32_t>(1, std::min<uint32_t>((uint32_t)r->threads, n / kMinRowsPerThread));
    std::vector<uint32_t> bad(nt, 0);
    std::vector<const char *> err(nt, nullptr);
    const uint32_t per = (n + nt - 1) / nt;
    auto work = [&](uint32_t t) {
      const uint32_t r0 = std::min(n, t * per), r1 = std::min(n, r0 + per);
      bad[t] = parse_rows(r, r0, r1, have, cols, nulls, &err[t]) < r1 ? 1 : 0;
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
      if (bad[t]) {  // the error of the earliest failing record range
        ysql::set_last_error(err[t]);
        return YSQL_EINVAL;
      }
    have += drop_skipped(r, n, have, cols, nulls);
  }
  *rows = have;
  return YSQL_OK;
}

void free_stage(Stage &s) {
  if (s.pending && s.done) (void)hipEventSynchronize(s.done);
  if (s.used_pending && s.used) (void)hipEventSynchronize(s.used);
  for (void *p : s.cols)
    if (p) (void)hipHostFree(p);
  for (uint8_t *p : s.nulls)
    if (p) (void)hipHostFree(p);
  for (void *p : s.dcols)
    if (p) (void)hipFree(p);
  for (uint8_t *p : s.dnulls)
    if (p) (void)hipFree(p);
  if (s.done) (void)hipEventDestroy(s.done);
  if (s.used) (void)hipEventDestroy(s.used);
  s = Stage();
}

// Pinned staging (and, with `device`, device buffers) for one batch.
int alloc_stage(ysql_csv_reader *r, Stage &st, bool device) {
  const size_t nc = r->types.size();
  st.cols.assign(nc, nullptr);
  st.nulls.assign(nc, nullptr);
  if (device) {
    st.dcols.assign(nc, nullptr);
    st.dnulls.assign(nc, nullptr);
  }
  for (size_t c = 0; c < nc; c++) {
    const size_t b = (size_t)r->batch_rows * r->bytes[c];
    hipError_t e = hipHostMalloc(&st.cols[c], b, hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc((void **)&st.nulls[c], r->batch_rows, hipHostMallocDefault);
    if (e == hipSuccess && device) e = hipMalloc(&st.dcols[c], b);
    if (e == hipSuccess && device) e = hipMalloc((void **)&st.dnulls[c], r->batch_rows);
    if (e != hipSuccess) {
      std::string msg = std::string("csv: staging allocation failed: ") + hipGetErrorString(e);
      ysql::set_last_error(msg.c_str());
      return YSQL_EHIP;
    }
  }
  if (hipEventCreateWithFlags(&st.done, hipEventDisableTiming) != hipSuccess) return YSQL_EHIP;
  if (device && hipEventCreateWithFlags(&st.used, hipEventDisableTiming) != hipSuccess) return YSQL_EHIP;
  return YSQL_OK;
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

  // delimiter and header row from the head, as csv::CSVReader guesses them
  const bool bom = r->size >= 3 && memcmp(r->data, "\xEF\xBB\xBF", 3) == 0;
  size_t head_len = std::min(r->size, kHeadBytes);
  if (const void *nul = memchr(r->data, 0, head_len)) head_len = (size_t)((const char *)nul - r->data);
  size_t header_row = 0;
  const char *hb = r->data + (head_len >= 3 && bom ? 3 : 0);
  guess_format(hb, r->data + std::max<size_t>(head_len, (size_t)(hb - r->data)), &r->delim, &header_row);

  // records before the header row are skipped; the header names the fields
  r->pos = r->data + (bom ? 3 : 0);
  std::vector<std::string> header;
  size_t nh = 0;
  for (size_t i = 0; i <= header_row && r->pos < r->end; i++) {
    const char *rs = r->pos, *a = nullptr;
    r->pos = r->scan_record(rs, &a);
    if (i == header_row) parse_record(rs, r->end, r->delim, header, nh);
  }
  if (nh == 0) {
    ysql::set_last_error("csv: no header row");
    ysql_csv_close(r);
    return YSQL_EINVAL;
  }
  header.resize(nh);
  r->nfields = (int)header.size();
  for (int c = 0; c < ncols; c++) {
    if (!names[c]) rc = YSQL_EINVAL;
    int idx = -1;
    for (int i = 0; names[c] && i < r->nfields; i++)
      if (header[i] == names[c]) idx = i;  // a repeated name means its last occurrence
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
  if (r->copy_stream) {
    ysql::set_last_error("csv: a reader opened for ysql_csv_next_async");
    return YSQL_EINVAL;
  }
  if (st.cols.empty()) {  // first use: pinned staging for one batch
    int rc = alloc_stage(r, st, false);
    if (rc) return rc;
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
  r->cur = (r->cur + 1) % (int)r->st.size();
  return YSQL_OK;
}

int ysql_csv_open_ex(const char *path, int ncols, const char *const *names, const int *types,
                     const ysql_csv_options *opt, ysql_csv_reader **out) {
  if (!out) return YSQL_EINVAL;
  *out = nullptr;
  if (opt && opt->slots != 0 && (opt->slots < 2 || opt->slots > 64)) return YSQL_EINVAL;
  if (opt && opt->batch_rows >= (1ull << 31)) return YSQL_EINVAL;
  ysql_csv_reader *r = nullptr;
  int rc = ysql_csv_open(path, ncols, names, types, &r);
  if (rc) return rc;
  if (opt && opt->batch_rows) r->batch_rows = (uint32_t)opt->batch_rows;
  if (opt && opt->slots) r->st.resize(opt->slots);
  if (opt && opt->async_copies) {
    if (hipStreamCreateWithFlags(&r->copy_stream, hipStreamNonBlocking) != hipSuccess) {
      ysql::set_last_error("csv: cannot create the copy stream");
      ysql_csv_close(r);
      return YSQL_EHIP;
    }
  }
  *out = r;
  return YSQL_OK;
}

int ysql_csv_next_async(ysql_csv_reader *r, void **dev_cols, uint8_t **dev_nulls, uint32_t *rows,
                        ysql_stream_t stream) {
  if (!r || !dev_cols || !rows) return YSQL_EINVAL;
  if (!r->copy_stream) {
    ysql::set_last_error("csv: ysql_csv_next_async needs a reader opened with async_copies");
    return YSQL_EINVAL;
  }
  hipStream_t s = (hipStream_t)stream;
  const size_t nc = r->types.size();
  // the caller's work on the batch handed out last is queued by now
  if (r->handed >= 0) {
    Stage &h = r->st[r->handed];
    if (hipEventRecord(h.used, s) != hipSuccess) return YSQL_EHIP;
    h.used_pending = true;
    r->handed = -1;
  }
  Stage &st = r->st[r->cur];
  if (st.cols.empty()) {
    int rc = alloc_stage(r, st, true);
    if (rc) return rc;
  }
  if (st.pending) {  // the staging area's previous copies have landed
    if (hipEventSynchronize(st.done) != hipSuccess) return YSQL_EHIP;
    st.pending = false;
  }
  int rc = parse_batch(r, st.cols.data(), st.nulls.data(), rows);
  if (rc) return rc;
  if (*rows == 0) return YSQL_OK;
  // the copies wait, on the copy stream, for the caller's work on the
  // batch that last used these device buffers; the caller's stream then
  // waits for the copies, so its earlier work overlaps them
  if (st.used_pending) {
    if (hipStreamWaitEvent(r->copy_stream, st.used, 0) != hipSuccess) return YSQL_EHIP;
    st.used_pending = false;
  }
  for (size_t c = 0; c < nc; c++) {
    if (hipMemcpyAsync(st.dcols[c], st.cols[c], (size_t)*rows * r->bytes[c], hipMemcpyHostToDevice,
                       r->copy_stream) != hipSuccess ||
        hipMemcpyAsync(st.dnulls[c], st.nulls[c], *rows, hipMemcpyHostToDevice, r->copy_stream) != hipSuccess)
      return YSQL_EHIP;
    dev_cols[c] = st.dcols[c];
    if (dev_nulls) dev_nulls[c] = st.dnulls[c];
  }
  if (hipEventRecord(st.done, r->copy_stream) != hipSuccess) return YSQL_EHIP;
  if (hipStreamWaitEvent(s, st.done, 0) != hipSuccess) return YSQL_EHIP;
  st.pending = true;
  r->handed = r->cur;
  r->cur = (r->cur + 1) % (int)r->st.size();
  return YSQL_OK;
}

void ysql_csv_close(ysql_csv_reader *r) {
  if (!r) return;
  // the caller's work on the last batch handed out may still read the
  // reader's device buffers: let the device finish before freeing them
  if (r->copy_stream) (void)hipDeviceSynchronize();
  for (Stage &st : r->st) free_stage(st);
  if (r->copy_stream) (void)hipStreamDestroy(r->copy_stream);
  if (r->data) munmap((void *)r->data, r->size);
  if (r->fd >= 0) close(r->fd);
  delete r;
}

}  // extern "C"
