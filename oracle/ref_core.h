// CPU ORACLE — test infrastructure only. This is a C++ restatement of
// MiniLake's CPU executor for the hot path (hash, SwissTable, expression
// evaluation, filter, hash aggregate, hash join, TopN, morsel scheduler).
// It is the parity checker for the HIP library and the timed CPU baseline;
// nothing in the product links, imports or calls it.
//
// Every function cites the reference file:line it restates (paths relative
// to the MiniLake repository root).
#pragma once

#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../include/minilake_gpu.h"

namespace ref {

// MiniLakeError (crates/minilake-core/src/error.rs:8-48) as a C++ exception.
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
[[noreturn]] inline void fail(int code, const std::string& m) { throw Error(code, m); }

// DataType (crates/minilake-core/src/types.rs:15-28).
enum DT : int { Boolean = 0, Int32 = 1, Int64 = 2, Float64 = 3, Date = 4, Utf8 = 5, Null = -1 };

inline const char* dt_name(int t) {
  switch (t) {
    case Boolean: return "BOOLEAN";
    case Int32: return "INT32";
    case Int64: return "INT64";
    case Float64: return "FLOAT64";
    case Date: return "DATE";
    case Utf8: return "VARCHAR";
    default: return "NULL";
  }
}
inline bool is_numeric(int t) { return t == Int32 || t == Int64 || t == Float64; }

// Dictionary of a string column (DictVec, crates/minilake-core/src/column.rs:126-131).
struct Dict {
  std::vector<std::string> entries;
};

// Column (crates/minilake-core/src/column.rs:135-210). Values are a typed
// array (bool as u8, i32/Date, i64, f64, u32 dictionary codes for Utf8);
// `valid` is null when every row is valid (Column::new drops all-valid bitmaps).
struct Column {
  int type = Int64;
  size_t len = 0;
  const void* data = nullptr;
  std::shared_ptr<void> owner;
  std::shared_ptr<std::vector<uint8_t>> valid;
  std::shared_ptr<const Dict> dict;

  bool is_valid(size_t r) const { return !valid || (*valid)[r]; }
  template <class T> const T* as() const { return static_cast<const T*>(data); }
  int width() const {
    switch (type) {
      case Boolean: return 1;
      case Int32: case Date: case Utf8: return 4;
      default: return 8;
    }
  }
};
using ColumnRef = std::shared_ptr<const Column>;

template <class T>
inline ColumnRef make_column(int type, std::vector<T>&& v, std::shared_ptr<std::vector<uint8_t>> valid = nullptr,
                             std::shared_ptr<const Dict> dict = nullptr) {
  auto c = std::make_shared<Column>();
  auto store = std::make_shared<std::vector<T>>(std::move(v));
  c->type = type;
  c->len = store->size();
  c->data = store->data();
  c->owner = store;
  // Column::new drops a bitmap without nulls (column.rs:204-210).
  if (valid) {
    bool any_null = false;
    for (uint8_t x : *valid) any_null |= (x == 0);
    if (any_null) c->valid = std::move(valid);
  }
  c->dict = std::move(dict);
  return c;
}

// Batch (crates/minilake-core/src/batch.rs:25-29) with optional ascending selection.
struct Batch {
  std::vector<ColumnRef> columns;
  size_t num_rows = 0;
  std::shared_ptr<const std::vector<uint32_t>> sel;
  size_t active_rows() const { return sel ? sel->size() : num_rows; }
};

// Rows (crates/minilake-exec/src/operators/aggregate/accumulator.rs:55-84).
struct Rows {
  const uint32_t* sel = nullptr;
  size_t n = 0;
  size_t len() const { return n; }
  size_t row(size_t k) const { return sel ? sel[k] : k; }
};
inline Rows batch_rows(const Batch& b) {  // aggregate/mod.rs:60-65
  Rows r;
  if (b.sel) {
    r.sel = b.sel->data();
    r.n = b.sel->size();
  } else {
    r.n = b.num_rows;
  }
  return r;
}

// Column::gather (crates/minilake-core/src/column.rs:380-398).
ColumnRef gather(const Column& c, const uint32_t* idx, size_t n);
// Batch::compact (batch.rs:121-134).
Batch compact(const Batch& b);
// Batch::concat (batch.rs:168-184).
Batch concat(const std::vector<Batch>& batches);
// Column::from_scalar broadcast helper.
struct Scalar;
ColumnRef column_from_scalar(const Scalar& s, size_t n, int dt);

// ScalarValue (crates/minilake-core/src/scalar.rs:11-30), runtime subset.
struct Scalar {
  int type = Null;  // Null means SQL NULL
  int64_t i = 0;    // Int32/Int64/Date/Boolean
  double f = 0;     // Float64
  std::string s;    // Utf8
  bool is_null() const { return type == Null; }
  bool as_f64(double* out) const {
    if (type == Int32 || type == Int64 || type == Date) { *out = (double)i; return true; }
    if (type == Float64) { *out = f; return true; }
    return false;
  }
  bool as_i64(int64_t* out) const {
    if (type == Int32 || type == Int64 || type == Date) { *out = i; return true; }
    return false;
  }
};

}  // namespace ref
