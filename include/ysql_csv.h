/* ysql_csv.h — CSV file -> column batches (the feeder of the scan path).
 *
 * Replaces the reference's GetOperator (src/engine/operators/get_operator.cpp:16-148) and its
 * batch sizing (include/config.hpp:28-38): a header row names the columns; each requested
 * column is parsed with the reference's rules (get_operator.cpp:68-95,
 * include/enums/data_type_utils.hpp:24-44):
 *   INT      std::stoi of the field (leading blanks, optional sign, digits; trailing text
 *            ignored; no digits or out of int32 range is an error)
 *   FLOAT    std::stof of the field (likewise, hexadecimal forms included; a result that
 *            overflows or underflows to a subnormal or zero is an error)
 *   DATETIME strptime "%Y-%m-%d %H:%M:%S" as UTC (timegm) -> int64 epoch seconds; a field
 *            that does not parse gives 0
 *   STRING   the first 255 bytes, NUL-padded to 256
 * An empty field is NULL: null byte 1 and the value INT_MIN / FLT_MIN / 0 / all-zero string.
 *
 * Records are split as the reference's csv::CSVReader (vendored csv-parser v2.3.0,
 * vendor/csv-parser/csv.hpp) splits them with its default format: the delimiter (one of
 * ',', '|', '\t', ';', '^') and the header row are guessed from the first 500,000 bytes; any
 * run of '\r' and '\n' ends a record (blank lines vanish); a '"' that starts a field quotes
 * it until a '"' followed by the delimiter, a newline or the end of the file; "" inside quotes
 * is a quote; records whose field count differs from the header's are skipped; a UTF-8 BOM
 * is skipped; a repeated header name means its last column.  The library's parse chunks of
 * 10,000,000 bytes are followed too: a newline run crossing a chunk end adds an empty
 * one-field record there (a NULL row in a one-column table).
 *
 * Batches hold ysql_csv_batch_rows() rows (the last one fewer): floor(4 MiB / row bytes),
 * rounded down to a multiple of 256, at least 256.
 *
 * The file is mapped read-only while the reader is open: it must not be truncated or
 * rewritten meanwhile.  A batch of 32,768 rows or more is converted on up to 16 host
 * threads (row ranges; results identical to one thread).  One reader serves one caller
 * thread at a time.
 *
 * Device path: ysql_csv_next parses the next batch into one of two pinned staging buffers
 * and queues its host-to-device copies on `stream`; it returns once the copies are queued,
 * so parsing the next batch overlaps the previous batch's copies.  A staging buffer is
 * reused only after the copies queued from it have completed.  The device buffers of a
 * batch are valid once `stream` reaches that call's copies.
 */
#ifndef YSQL_CSV_H
#define YSQL_CSV_H

#include <stdint.h>

#include "ysql_kernels.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ysql_csv_reader ysql_csv_reader;

/* Rows per batch for columns of these YSQL_* types (config.hpp:28-38). */
int ysql_csv_batch_rows(int ncols, const int *types, uint32_t *rows);

/* Opens `path`, guesses its delimiter and header row and reads the header; names[c] is the
 * CSV header name of requested column c, types[c] its YSQL_* type (get_operator.cpp:16-53).
 * YSQL_EINVAL for a missing or empty file, an unknown name or a bad type. */
int ysql_csv_open(const char *path, int ncols, const char *const *names, const int *types,
                  ysql_csv_reader **out);

/* Next batch into host memory: host_cols[c] holds batch_rows elements of column c,
 * host_nulls[c] batch_rows null bytes (may be NULL: not written).  *rows = rows read
 * (0: end of file).  YSQL_EINVAL for a field that does not convert. */
int ysql_csv_next_host(ysql_csv_reader *r, void *const *host_cols, uint8_t *const *host_nulls,
                       uint32_t *rows);

/* Next batch into device memory (get_operator.cpp:56-108, storeBuffer :125-148), through
 * pinned staging and asynchronous copies on `stream` (see above). */
int ysql_csv_next(ysql_csv_reader *r, void *const *dev_cols, uint8_t *const *dev_nulls, uint32_t *rows,
                  ysql_stream_t stream);

/* ----------------------------------------------------------------------
 * Scale mode: batches sized for HBM rather than the reference's 4 MiB, a ring
 * of pinned staging areas, and copies on a copy stream of the reader that
 * overlap the caller's kernels (the reference copied each column
 * synchronously, get_operator.cpp:125-148).                              */
typedef struct ysql_csv_options {
  uint64_t batch_rows;   /* rows per batch (< 2^31); 0 = ysql_csv_batch_rows' sizing */
  uint32_t slots;        /* staging areas in the ring, 2..64; 0 = 2 */
  uint32_t async_copies; /* nonzero: the reader owns device buffers per area and copies on its own
                            stream (ysql_csv_next_async); zero: ysql_csv_next as above */
} ysql_csv_options;
/* ysql_csv_open with options (opt may be NULL: the defaults).  Pinned host
 * memory = slots x batch_rows x (row bytes + 1 null byte per column), and as
 * much device memory with async_copies.                                    */
int ysql_csv_open_ex(const char *path, int ncols, const char *const *names, const int *types,
                     const ysql_csv_options *opt, ysql_csv_reader **out);
/* Next batch into the reader's device buffers (reader opened with
 * async_copies): parses into the next staging area, queues its host-to-device
 * copies on the reader's copy stream, makes `stream` wait for them, and
 * returns the batch's device column and null pointers in dev_cols[c] /
 * dev_nulls[c] (dev_nulls may be NULL).  Work queued on `stream` before the
 * call runs while the copies do; work queued after it sees the batch.
 * Contract: queue every piece of work that reads a batch on `stream` before
 * the next call.  A batch's buffers are overwritten by the call `slots`
 * calls later, after that work has completed (the copy stream waits for it
 * on the device, not the host).                                            */
int ysql_csv_next_async(ysql_csv_reader *r, void **dev_cols, uint8_t **dev_nulls, uint32_t *rows,
                        ysql_stream_t stream);

/* Waits for the reader's queued copies (and, for a reader opened with
 * async_copies, for the device), frees its staging and closes the file. */
void ysql_csv_close(ysql_csv_reader *r);

/* ----------------------------------------------------------------------
 * Result writer: replaces the reference's CsvWriter
 * (include/engine/operators/csv_writer.hpp:12-160), fed by ExecutorEngine
 * with every result batch (src/engine/executor_engine.cpp:19-58).  The
 * first write puts the column names joined by ',' on the first line; each
 * row is its fields joined by ',' and ended by '\n':
 *   NULL      ""                          STRING    "text up to the NUL" (not escaped)
 *   INT       decimal                     FLOAT     fixed, 4 decimals (std::to_chars)
 *   DATETIME  "YYYY-MM-DD HH:MM:SS" (UTC, strftime of gmtime_r)
 * Unlike the reference, floats whose fixed form is longer than 32 characters
 * (|x| >= about 1e27) are written correctly, not as undefined buffer
 * contents, and a string with no NUL in its 256 bytes is cut at 256.  Rows
 * are formatted on up to 16 host threads and written in order. */
typedef struct ysql_csv_writer ysql_csv_writer;
/* Creates (truncates) `path`; names[c] / types[c] (YSQL_INT, YSQL_FLOAT,
 * YSQL_DATETIME, YSQL_STRING) describe the result columns. */
int ysql_csv_writer_open(const char *path, int ncols, const char *const *names, const int *types,
                         ysql_csv_writer **out);
/* Appends `rows` rows from host buffers; host_nulls (or any entry) may be
 * NULL: no NULLs in that column. */
int ysql_csv_writer_write_host(ysql_csv_writer *w, const void *const *host_cols,
                               const uint8_t *const *host_nulls, uint64_t rows);
/* Appends `rows` rows from device buffers: copied on `stream` into the
 * writer's pinned staging (grown as needed), then formatted.  Returns after
 * the rows are written (synchronises `stream`). */
int ysql_csv_writer_write(ysql_csv_writer *w, const void *const *dev_cols, const uint8_t *const *dev_nulls,
                          uint64_t rows, ysql_stream_t stream);
/* Flushes and closes the file, frees the writer; YSQL_EINVAL if a write failed. */
int ysql_csv_writer_close(ysql_csv_writer *w);

#ifdef __cplusplus
}
#endif

#endif /* YSQL_CSV_H */
