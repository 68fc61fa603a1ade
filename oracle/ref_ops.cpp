// CPU ORACLE — test infrastructure only (see ref_core.h).
// Operators, sinks and the morsel scheduler, restating
//   crates/minilake-exec/src/operators/aggregate/{accumulator,group_keys,hash}.rs
//   crates/minilake-exec/src/operators/{join,sort,filter,projection,collect}.rs
//   crates/minilake-exec/src/{pipeline,scheduler,executor}.rs
// and the extern "C" entry points the tests and the CPU baseline call.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <functional>
#include <map>
#include <mutex>
#include <thread>

#include "ref_expr.h"
#include "ref_hash.h"
#include "ref_oracle.h"

namespace ref {

// ===================================================================== GroupKeys
// group_keys.rs:27-150: distinct keys stored column-wise, group id = insertion index.
struct GroupKeys {
  struct KeyCol {
    int dt;
    std::vector<int32_t> i32;
    std::vector<int64_t> i64;
    std::vector<double> f64;
    std::vector<uint8_t> b;
    std::vector<std::string> s;
    std::vector<uint8_t> valid;
  };
  std::vector<KeyCol> cols;
  size_t len = 0;

  explicit GroupKeys(const std::vector<int>& types) {
    for (int t : types) {
      KeyCol k;
      k.dt = t;
      cols.push_back(std::move(k));
    }
  }
  uint32_t push(const std::vector<ColumnRef>& in, size_t row) {  // group_keys.rs:67-101
    for (size_t i = 0; i < cols.size(); i++) {
      KeyCol& kc = cols[i];
      const Column& c = *in[i];
      bool v = c.is_valid(row);
      kc.valid.push_back(v);
      switch (kc.dt) {
        case Int32: case Date: kc.i32.push_back(c.as<int32_t>()[row]); break;
        case Int64: kc.i64.push_back(c.as<int64_t>()[row]); break;
        case Float64: kc.f64.push_back(c.as<double>()[row]); break;
        case Boolean: kc.b.push_back(c.as<uint8_t>()[row]); break;
        case Utf8: kc.s.push_back(c.dict->entries[c.as<uint32_t>()[row]]); break;
      }
    }
    len += 1;
    return (uint32_t)(len - 1);
  }
  bool equals(uint32_t g, const std::vector<ColumnRef>& in, size_t row) const {  // group_keys.rs:104-125
    for (size_t i = 0; i < cols.size(); i++) {
      const KeyCol& kc = cols[i];
      const Column& c = *in[i];
      bool v = c.is_valid(row);
      if (v != (bool)kc.valid[g]) return false;
      if (!v) continue;
      bool eq = false;
      switch (kc.dt) {
        case Int32: case Date: eq = kc.i32[g] == c.as<int32_t>()[row]; break;
        case Int64: eq = kc.i64[g] == c.as<int64_t>()[row]; break;
        case Float64: eq = kc.f64[g] == c.as<double>()[row]; break;
        case Boolean: eq = kc.b[g] == c.as<uint8_t>()[row]; break;
        case Utf8: eq = kc.s[g] == c.dict->entries[c.as<uint32_t>()[row]]; break;
      }
      if (!eq) return false;
    }
    return true;
  }
  std::vector<ColumnRef> to_columns() const {  // group_keys.rs:128-144
    std::vector<ColumnRef> out;
    for (const KeyCol& kc : cols) {
      auto valid = std::make_shared<std::vector<uint8_t>>(kc.valid);
      switch (kc.dt) {
        case Int32: case Date: out.push_back(make_column(kc.dt, std::vector<int32_t>(kc.i32), valid)); break;
        case Int64: out.push_back(make_column(Int64, std::vector<int64_t>(kc.i64), valid)); break;
        case Float64: out.push_back(make_column(Float64, std::vector<double>(kc.f64), valid)); break;
        case Boolean: out.push_back(make_column(Boolean, std::vector<uint8_t>(kc.b), valid)); break;
        case Utf8: {
          auto d = std::make_shared<Dict>();
          d->entries = kc.s;
          std::vector<uint32_t> codes(kc.s.size());
          for (size_t i = 0; i < codes.size(); i++) codes[i] = (uint32_t)i;
          out.push_back(make_column(Utf8, std::move(codes), valid, d));
          break;
        }
      }
    }
    return out;
  }
};

// =================================================================== Accumulator
// accumulator.rs:99-544
enum StateKind { S_COUNT, S_SUM_I64, S_SUM_F64, S_AVG, S_MINMAX_I64, S_MINMAX_F64, S_MINMAX_STR };

struct Accumulator {
  int func;
  int input_type = Null;  // -1 = none
  int output_type;
  StateKind kind;
  std::vector<int64_t> i;    // count / sum i64 / min-max i64 / avg count
  std::vector<double> f;     // sum f64 / avg sum / min-max f64
  std::vector<std::string> s; // min-max of strings (State::MinMaxStr, accumulator.rs:94)
  std::vector<uint8_t> seen;

  Accumulator(int fn, int arg_type, bool has_arg) : func(fn) {  // accumulator.rs:104-165
    auto bad = [&]() { fail(ML_ERR_PLAN, "aggregate not supported for this argument type"); };
    switch (fn) {
      case ML_AGG_COUNT_STAR: output_type = Int64; kind = S_COUNT; break;
      case ML_AGG_COUNT: input_type = has_arg ? arg_type : Null; output_type = Int64; kind = S_COUNT; break;
      case ML_AGG_SUM:
        if (arg_type == Int32 || arg_type == Int64 || arg_type == Boolean) { input_type = Int64; output_type = Int64; kind = S_SUM_I64; }
        else if (arg_type == Float64) { input_type = Float64; output_type = Float64; kind = S_SUM_F64; }
        else bad();
        break;
      case ML_AGG_AVG:
        if (!is_numeric(arg_type)) bad();
        input_type = Float64; output_type = Float64; kind = S_AVG;
        break;
      case ML_AGG_MIN: case ML_AGG_MAX:
        if (arg_type == Int32 || arg_type == Int64 || arg_type == Date) { input_type = Int64; output_type = arg_type; kind = S_MINMAX_I64; }
        else if (arg_type == Float64) { input_type = Float64; output_type = Float64; kind = S_MINMAX_F64; }
        else if (arg_type == Utf8) { input_type = Utf8; output_type = Utf8; kind = S_MINMAX_STR; }  // :156-160
        else bad();
        break;
      default: bad();
    }
  }
  size_t num_groups() const {
    if (kind == S_MINMAX_STR) return s.size();
    return kind == S_SUM_F64 || kind == S_MINMAX_F64 ? f.size() : i.size();
  }
  void resize(size_t n) {  // accumulator.rs:205-230
    switch (kind) {
      case S_COUNT: i.resize(n, 0); break;
      case S_SUM_I64: case S_MINMAX_I64: i.resize(n, 0); seen.resize(n, 0); break;
      case S_SUM_F64: case S_MINMAX_F64: f.resize(n, 0.0); seen.resize(n, 0); break;
      case S_AVG: f.resize(n, 0.0); i.resize(n, 0); break;
      case S_MINMAX_STR: s.resize(n); seen.resize(n, 0); break;
    }
  }
  ColumnRef prepare_input(ColumnRef c) const {  // accumulator.rs:229-236
    if (input_type != Null && func != ML_AGG_COUNT && c->type != input_type) return cast_column(c, input_type);
    return c;
  }
  // accumulator.rs:240-379 (grouped path; gids always present here)
  void update(const Column* in, Rows rows, const std::vector<uint32_t>& gids) {
    bool is_min = func == ML_AGG_MIN;
    if (kind == S_COUNT) {
      if (func == ML_AGG_COUNT_STAR || !in || !in->valid) {
        for (uint32_t g : gids) i[g] += 1;
      } else {
        for (size_t k = 0; k < rows.len(); k++)
          if (in->is_valid(rows.row(k))) i[gids[k]] += 1;
      }
      return;
    }
    if (!in) fail(ML_ERR_INTERNAL, "aggregate without input");
    switch (kind) {
      case S_SUM_I64: {
        const int64_t* v = in->as<int64_t>();
        for (size_t k = 0; k < rows.len(); k++) {
          size_t r = rows.row(k);
          if (!in->is_valid(r)) continue;
          uint32_t g = gids[k];
          i[g] = (int64_t)((uint64_t)i[g] + (uint64_t)v[r]);
          seen[g] = 1;
        }
        break;
      }
      case S_SUM_F64: {
        const double* v = in->as<double>();
        for (size_t k = 0; k < rows.len(); k++) {
          size_t r = rows.row(k);
          if (!in->is_valid(r)) continue;
          uint32_t g = gids[k];
          f[g] += v[r];
          seen[g] = 1;
        }
        break;
      }
      case S_AVG: {
        const double* v = in->as<double>();
        for (size_t k = 0; k < rows.len(); k++) {
          size_t r = rows.row(k);
          if (!in->is_valid(r)) continue;
          uint32_t g = gids[k];
          f[g] += v[r];
          i[g] += 1;
        }
        break;
      }
      case S_MINMAX_I64: {
        const int64_t* v = in->as<int64_t>();
        for (size_t k = 0; k < rows.len(); k++) {
          size_t r = rows.row(k);
          if (!in->is_valid(r)) continue;
          uint32_t g = gids[k];
          int64_t x = v[r];
          if (!seen[g] || (is_min && x < i[g]) || (!is_min && x > i[g])) { i[g] = x; seen[g] = 1; }
        }
        break;
      }
      case S_MINMAX_F64: {
        const double* v = in->as<double>();
        for (size_t k = 0; k < rows.len(); k++) {
          size_t r = rows.row(k);
          if (!in->is_valid(r)) continue;
          uint32_t g = gids[k];
          double x = v[r];
          if (!seen[g] || (is_min && x < f[g]) || (!is_min && x > f[g])) { f[g] = x; seen[g] = 1; }
        }
        break;
      }
      case S_MINMAX_STR: {  // accumulator.rs:352-368: bytewise (std::string compares as unsigned bytes)
        const uint32_t* v = in->as<uint32_t>();
        for (size_t k = 0; k < rows.len(); k++) {
          size_t r = rows.row(k);
          if (!in->is_valid(r)) continue;
          uint32_t g = gids[k];
          const std::string& x = in->dict->entries[v[r]];
          if (!seen[g] || (is_min && x < s[g]) || (!is_min && x > s[g])) { s[g] = x; seen[g] = 1; }
        }
        break;
      }
      default: break;
    }
  }
  // str_column (accumulator.rs:546-553): the strings, "" for a group never seen, NULL there
  ColumnRef str_column() const {
    auto d = std::make_shared<Dict>();
    d->entries = s;
    std::vector<uint32_t> codes(s.size());
    for (size_t g = 0; g < s.size(); g++) codes[g] = (uint32_t)g;
    return make_column(Utf8, std::move(codes), std::make_shared<std::vector<uint8_t>>(seen), d);
  }
  // state (accumulator.rs:382-404)
  std::vector<ColumnRef> state() const {
    auto bm = [&]() { return std::make_shared<std::vector<uint8_t>>(seen); };
    switch (kind) {
      case S_COUNT: return {make_column(Int64, std::vector<int64_t>(i))};
      case S_SUM_I64: case S_MINMAX_I64: return {make_column(Int64, std::vector<int64_t>(i), bm())};
      case S_SUM_F64: case S_MINMAX_F64: return {make_column(Float64, std::vector<double>(f), bm())};
      case S_AVG: return {make_column(Float64, std::vector<double>(f)), make_column(Int64, std::vector<int64_t>(i))};
      case S_MINMAX_STR: return {str_column()};
    }
    return {};
  }
  size_t num_state_cols() const { return kind == S_AVG ? 2 : 1; }
  // merge (accumulator.rs:407-497)
  void merge(const ColumnRef* st, const std::vector<uint32_t>& gids) {
    bool is_min = func == ML_AGG_MIN;
    const Column& first = *st[0];
    for (size_t k = 0; k < gids.size(); k++) {
      uint32_t g = gids[k];
      switch (kind) {
        case S_COUNT: i[g] += first.as<int64_t>()[k]; break;
        case S_SUM_I64:
          if (first.is_valid(k)) { i[g] = (int64_t)((uint64_t)i[g] + (uint64_t)first.as<int64_t>()[k]); seen[g] = 1; }
          break;
        case S_SUM_F64:
          if (first.is_valid(k)) { f[g] += first.as<double>()[k]; seen[g] = 1; }
          break;
        case S_AVG: f[g] += first.as<double>()[k]; i[g] += st[1]->as<int64_t>()[k]; break;
        case S_MINMAX_I64: {
          int64_t x = first.as<int64_t>()[k];
          if (first.is_valid(k) && (!seen[g] || (is_min && x < i[g]) || (!is_min && x > i[g]))) { i[g] = x; seen[g] = 1; }
          break;
        }
        case S_MINMAX_F64: {
          double x = first.as<double>()[k];
          if (first.is_valid(k) && (!seen[g] || (is_min && x < f[g]) || (!is_min && x > f[g]))) { f[g] = x; seen[g] = 1; }
          break;
        }
        case S_MINMAX_STR: {  // accumulator.rs:470-484
          if (!first.is_valid(k)) break;
          const std::string& x = first.dict->entries[first.as<uint32_t>()[k]];
          if (!seen[g] || (is_min && x < s[g]) || (!is_min && x > s[g])) { s[g] = x; seen[g] = 1; }
          break;
        }
      }
    }
  }
  // finish (accumulator.rs:500-526)
  ColumnRef finish() const {
    auto bm = [&]() { return std::make_shared<std::vector<uint8_t>>(seen); };
    switch (kind) {
      case S_COUNT: return make_column(Int64, std::vector<int64_t>(i));
      case S_SUM_I64: return make_column(Int64, std::vector<int64_t>(i), bm());
      case S_SUM_F64: return make_column(Float64, std::vector<double>(f), bm());
      case S_AVG: {
        std::vector<double> v(f.size());
        auto valid = std::make_shared<std::vector<uint8_t>>(f.size());
        for (size_t g = 0; g < f.size(); g++) {
          v[g] = i[g] == 0 ? 0.0 : f[g] / (double)i[g];
          (*valid)[g] = i[g] > 0;
        }
        return make_column(Float64, std::move(v), valid);
      }
      case S_MINMAX_I64: {
        if (output_type == Int32 || output_type == Date) {
          std::vector<int32_t> v(i.size());
          for (size_t g = 0; g < i.size(); g++) v[g] = (int32_t)i[g];
          return make_column(output_type, std::move(v), bm());
        }
        return make_column(Int64, std::vector<int64_t>(i), bm());
      }
      case S_MINMAX_F64: return make_column(Float64, std::vector<double>(f), bm());
      case S_MINMAX_STR: return str_column();
    }
    return nullptr;
  }
};

// ================================================================= HashAggState
struct AggDef {
  int func;
  int arg_type;
  ExprPtr arg;  // null for COUNT(*)
};

struct HashAggState {  // aggregate/hash.rs:55-172
  std::vector<int> key_types;
  GroupKeys keys;
  SwissTable table;
  std::vector<Accumulator> accs;
  std::vector<uint64_t> hashes;
  std::vector<uint32_t> gids;

  HashAggState(const std::vector<int>& kt, const std::vector<AggDef>& aggs)
      : key_types(kt), keys(kt), table(SwissTable::with_capacity(0)) {
    for (auto& a : aggs) accs.emplace_back(a.func, a.arg_type, (bool)a.arg);
  }
  size_t num_groups() const { return keys.len; }

  struct Store {  // hash.rs:36-52
    GroupKeys* keys;
    const std::vector<ColumnRef>* cols;
    Rows rows;
    bool equals(uint32_t p, size_t k) const { return keys->equals(p, *cols, rows.row(k)); }
    uint32_t insert(size_t k) { return keys->push(*cols, rows.row(k)); }
  };

  void assign_groups(const std::vector<ColumnRef>& key_cols, Rows rows) {  // hash.rs:86-101
    hash_columns(key_cols, rows, hashes);
    Store store{&keys, &key_cols, rows};
    table.find_or_insert_batch(hashes.data(), hashes.size(), store, gids);
    size_t n = keys.len;
    for (auto& acc : accs)
      if (acc.num_groups() < n) acc.resize(n);
  }
  void update(const Batch& b, const std::vector<ExprPtr>& group_exprs, const std::vector<AggDef>& aggs) {  // hash.rs:104-122
    std::vector<ColumnRef> key_cols;
    for (size_t i = 0; i < group_exprs.size(); i++) key_cols.push_back(evaluate_to_column(*group_exprs[i], b, key_types[i]));
    Rows rows = batch_rows(b);
    assign_groups(key_cols, rows);
    for (size_t a = 0; a < aggs.size(); a++) {
      ColumnRef arg;
      if (aggs[a].arg) arg = accs[a].prepare_input(evaluate_to_column(*aggs[a].arg, b, aggs[a].arg_type));
      accs[a].update(arg.get(), rows, gids);
    }
  }
  Batch partial_batch() const {  // hash.rs:149-155
    Batch b;
    b.columns = keys.to_columns();
    for (auto& acc : accs)
      for (auto& c : acc.state()) b.columns.push_back(c);
    b.num_rows = num_groups();
    return b;
  }
  void merge_partial(const Batch& batch) {  // hash.rs:125-137
    Batch b = compact(batch);
    size_t nk = key_types.size();
    std::vector<ColumnRef> kc(b.columns.begin(), b.columns.begin() + nk);
    Rows all;
    all.n = b.num_rows;
    assign_groups(kc, all);
    size_t c = nk;
    for (auto& acc : accs) {
      acc.merge(&b.columns[c], gids);
      c += acc.num_state_cols();
    }
  }
  void merge_state(const HashAggState& other) {  // hash.rs:140-146
    if (other.num_groups() == 0) return;
    merge_partial(other.partial_batch());
  }
  Batch final_batch() const {  // hash.rs:158-164 (one batch; split only matters for batch streaming)
    Batch b;
    b.columns = keys.to_columns();
    for (auto& acc : accs) b.columns.push_back(acc.finish());
    b.num_rows = num_groups();
    return b;
  }
};

// ======================================================================= sinks
struct LocalSink {
  virtual ~LocalSink() = default;
  virtual void sink(const Batch& b) = 0;
  virtual void combine() = 0;
};
struct Sink {
  virtual ~Sink() = default;
  virtual std::unique_ptr<LocalSink> create_local() = 0;
  virtual void finalize() = 0;
};

// HashAggregateSink (aggregate/hash.rs:285-484), no spill. Partial mode emits
// each group's states (partial_batch); Final mode reads rows of states,
// keys first, and merges them (HashAggLocal::sink, hash.rs:429-437).
struct HashAggSink : Sink {
  std::vector<ExprPtr> group_exprs;
  std::vector<int> key_types;
  std::vector<AggDef> aggs;
  int mode = ML_AGG_MODE_SINGLE;
  std::mutex mu;
  std::unique_ptr<HashAggState> global;
  Batch output;
  int resizes = 0;
  size_t capacity = 0;

  struct Local : LocalSink {
    HashAggSink* s;
    HashAggState state;
    Local(HashAggSink* sink) : s(sink), state(sink->key_types, sink->aggs) {}
    void sink(const Batch& b) override {
      if (s->mode == ML_AGG_MODE_FINAL) state.merge_partial(b);
      else state.update(b, s->group_exprs, s->aggs);
    }
    void combine() override {  // hash.rs:439-483
      std::lock_guard<std::mutex> g(s->mu);
      s->resizes += state.table.resizes;
      if (!s->global) s->global = std::make_unique<HashAggState>(std::move(state));
      else s->global->merge_state(state);
    }
  };
  std::unique_ptr<LocalSink> create_local() override { return std::make_unique<Local>(this); }
  void finalize() override {  // hash.rs:369-425
    HashAggState empty(key_types, aggs);
    const HashAggState& st = global ? *global : empty;
    output = mode == ML_AGG_MODE_PARTIAL ? st.partial_batch() : st.final_batch();
    if (global) capacity = global->table.capacity();
  }
};

// UngroupedAggregateSink (aggregate/ungrouped.rs:36-139), Single mode: one
// group; per-thread accumulators merged into the global ones under a mutex,
// once per thread; always one output row (no input at all: COUNT = 0 and
// SUM/AVG/MIN/MAX NULL, ungrouped.rs:82-91).
struct UngroupedAggSink : Sink {
  std::vector<AggDef> aggs;
  int mode = ML_AGG_MODE_SINGLE;
  std::mutex mu;
  std::unique_ptr<std::vector<Accumulator>> global;
  Batch output;

  std::vector<Accumulator> fresh() const {  // ungrouped.rs:23-32
    std::vector<Accumulator> a;
    for (auto& d : aggs) {
      a.emplace_back(d.func, d.arg_type, (bool)d.arg);
      a.back().resize(1);
    }
    return a;
  }
  struct Local : LocalSink {
    UngroupedAggSink* s;
    std::vector<Accumulator> accs;
    std::vector<uint32_t> gids;
    Local(UngroupedAggSink* sink) : s(sink), accs(sink->fresh()) {}
    void sink(const Batch& b) override {  // ungrouped.rs:100-121
      if (s->mode == ML_AGG_MODE_FINAL) {  // rows of states, all into the one group
        Batch c = compact(b);
        gids.assign(c.num_rows, 0u);
        size_t col = 0;
        for (auto& acc : accs) {
          acc.merge(&c.columns[col], gids);
          col += acc.num_state_cols();
        }
        return;
      }
      Rows rows = batch_rows(b);
      gids.assign(rows.len(), 0u);
      for (size_t a = 0; a < s->aggs.size(); a++) {
        ColumnRef arg;
        if (s->aggs[a].arg) arg = accs[a].prepare_input(evaluate_to_column(*s->aggs[a].arg, b, s->aggs[a].arg_type));
        accs[a].update(arg.get(), rows, gids);
      }
    }
    void combine() override {  // ungrouped.rs:123-139
      std::lock_guard<std::mutex> g(s->mu);
      if (!s->global) {
        s->global = std::make_unique<std::vector<Accumulator>>(std::move(accs));
        return;
      }
      const std::vector<uint32_t> zero(1, 0u);
      for (size_t a = 0; a < accs.size(); a++) {
        auto st = accs[a].state();
        (*s->global)[a].merge(st.data(), zero);
      }
    }
  };
  std::unique_ptr<LocalSink> create_local() override { return std::make_unique<Local>(this); }
  void finalize() override {  // ungrouped.rs:75-96
    std::vector<Accumulator> accs = global ? std::move(*global) : fresh();
    output.columns.clear();
    for (auto& acc : accs) {
      if (mode == ML_AGG_MODE_PARTIAL)
        for (auto& c : acc.state()) output.columns.push_back(c);
      else
        output.columns.push_back(acc.finish());
    }
    output.num_rows = 1;
  }
};

// ================================================================== hash join
struct JoinTable {  // join.rs:36-45
  Batch batch;
  std::vector<ColumnRef> keys;
  std::vector<int> key_types;
  SwissTable table;
  std::vector<uint32_t> next;
  std::vector<int> schema;
};
constexpr uint32_t NONE = 0xFFFFFFFFu;

// build_table (join.rs:160-203)
static std::unique_ptr<JoinTable> build_table(const std::vector<Batch>& batches, const std::vector<ExprPtr>& key_exprs,
                                              const std::vector<int>& key_types, const std::vector<int>& schema) {
  auto jt = std::make_unique<JoinTable>();
  if (batches.empty()) {
    jt->batch.num_rows = 0;
    for (int t : schema) {
      (void)t;
      jt->batch.columns.push_back(make_column(Int64, std::vector<int64_t>()));
    }
  } else {
    jt->batch = concat(batches);
  }
  size_t n = jt->batch.num_rows;
  for (size_t i = 0; i < key_exprs.size(); i++) jt->keys.push_back(evaluate_to_column(*key_exprs[i], jt->batch, key_types[i]));
  std::vector<uint64_t> hashes;
  Rows all;
  all.n = n;
  hash_columns(jt->keys, all, hashes);
  jt->table = SwissTable::with_capacity(n);
  jt->next.assign(n, NONE);
  for (size_t r = 0; r < n; r++) {
    bool null_key = false;
    for (auto& k : jt->keys) null_key |= !k->is_valid(r);
    if (null_key) continue;
    uint32_t head;
    size_t slot;
    bool found = jt->table.find_or_vacant(
        hashes[r],
        [&](uint32_t h) {
          for (auto& k : jt->keys)
            if (!values_equal(*k, h, *k, r)) return false;
          return true;
        },
        &head, &slot);
    if (found) {
      jt->next[r] = jt->next[head];
      jt->next[head] = (uint32_t)r;
    } else {
      jt->table.insert_at(slot, hashes[r], (uint32_t)r);
    }
  }
  jt->key_types = key_types;
  jt->schema = schema;
  return jt;
}

struct JoinBuildSink : Sink {  // join.rs:72-157,216-242
  std::vector<ExprPtr> keys;
  std::vector<int> key_types;
  std::vector<int> schema;
  std::mutex mu;
  std::vector<Batch> batches;
  std::unique_ptr<JoinTable> table;
  struct Local : LocalSink {
    JoinBuildSink* s;
    std::vector<Batch> batches;
    explicit Local(JoinBuildSink* sink) : s(sink) {}
    void sink(const Batch& b) override { batches.push_back(compact(b)); }
    void combine() override {
      std::lock_guard<std::mutex> g(s->mu);
      for (auto& b : batches) s->batches.push_back(std::move(b));
    }
  };
  std::unique_ptr<LocalSink> create_local() override { return std::make_unique<Local>(this); }
  void finalize() override { table = build_table(batches, keys, key_types, schema); batches.clear(); }
};

// HashJoinProbe::execute / emit (join.rs:263-337)
static void probe_execute(const JoinTable& jt, const std::vector<ExprPtr>& key_exprs, bool build_is_left,
                          size_t batch_size, const Batch& batch, std::vector<Batch>& out,
                          std::vector<std::pair<uint32_t, uint32_t>>* pairs_only = nullptr) {
  if (jt.table.len == 0) return;
  std::vector<ColumnRef> keys;
  for (size_t i = 0; i < key_exprs.size(); i++) keys.push_back(evaluate_to_column(*key_exprs[i], batch, jt.key_types[i]));
  Rows rows = batch_rows(batch);
  std::vector<uint64_t> hashes;
  hash_columns(keys, rows, hashes);
  std::vector<uint32_t> pi, bi;
  auto emit = [&]() {
    if (pairs_only) {
      for (size_t i = 0; i < pi.size(); i++) pairs_only->push_back({pi[i], bi[i]});
    } else {
      Batch o;
      o.num_rows = pi.size();
      std::vector<ColumnRef> p, b;
      for (auto& c : batch.columns) p.push_back(gather(*c, pi.data(), pi.size()));
      for (auto& c : jt.batch.columns) b.push_back(gather(*c, bi.data(), bi.size()));
      if (build_is_left) { o.columns = b; o.columns.insert(o.columns.end(), p.begin(), p.end()); }
      else { o.columns = p; o.columns.insert(o.columns.end(), b.begin(), b.end()); }
      out.push_back(std::move(o));
    }
    pi.clear();
    bi.clear();
  };
  for (size_t k = 0; k < hashes.size(); k++) {
    size_t r = rows.row(k);
    bool null_key = false;
    for (auto& c : keys) null_key |= !c->is_valid(r);
    if (null_key) continue;
    uint32_t b;
    bool found = jt.table.find(
        hashes[k],
        [&](uint32_t cand) {
          for (size_t i = 0; i < keys.size(); i++)
            if (!values_equal(*jt.keys[i], cand, *keys[i], r)) return false;
          return true;
        },
        &b);
    if (!found) continue;
    for (;;) {
      pi.push_back((uint32_t)r);
      bi.push_back(b);
      if (pi.size() >= batch_size) emit();
      b = jt.next[b];
      if (b == NONE) break;
    }
  }
  if (!pi.empty()) emit();
}

// ================================================================= Sort / TopN
struct SortKeyDef {
  ExprPtr expr;
  bool asc;
  bool nulls_first;
};

// total_cmp ordering key for f64 (Rust f64::total_cmp)
static int64_t total_key(double x) {
  int64_t b;
  std::memcpy(&b, &x, 8);
  b ^= (int64_t)(((uint64_t)(b >> 63)) >> 1);
  return b;
}

// compare_rows (sort.rs:43-82)
static int compare_rows(const std::vector<ColumnRef>& cols, const std::vector<SortKeyDef>& keys, size_t a, size_t b) {
  for (size_t i = 0; i < keys.size(); i++) {
    const Column& c = *cols[i];
    bool va = c.is_valid(a), vb = c.is_valid(b);
    if (!va && !vb) continue;
    if (!va) return keys[i].nulls_first ? -1 : 1;
    if (!vb) return keys[i].nulls_first ? 1 : -1;
    int o = 0;
    switch (c.type) {
      case Int32: case Date: { auto x = c.as<int32_t>()[a], y = c.as<int32_t>()[b]; o = (x > y) - (x < y); break; }
      case Int64: { auto x = c.as<int64_t>()[a], y = c.as<int64_t>()[b]; o = (x > y) - (x < y); break; }
      case Float64: { auto x = total_key(c.as<double>()[a]), y = total_key(c.as<double>()[b]); o = (x > y) - (x < y); break; }
      case Boolean: { auto x = c.as<uint8_t>()[a], y = c.as<uint8_t>()[b]; o = (x > y) - (x < y); break; }
      case Utf8: { int cc = c.dict->entries[c.as<uint32_t>()[a]].compare(c.dict->entries[c.as<uint32_t>()[b]]); o = (cc > 0) - (cc < 0); break; }
    }
    if (!keys[i].asc) o = -o;
    if (o != 0) return o;
  }
  return 0;
}

static std::vector<int> batch_schema(const Batch& b) {
  std::vector<int> s;
  for (auto& c : b.columns) s.push_back(c->type);
  return s;
}

// sort_batches (sort.rs:85-132)
static bool sort_batches(const std::vector<Batch>& batches, const std::vector<SortKeyDef>& keys, long long limit,
                         Batch* out) {
  if (batches.empty()) return false;
  Batch all = concat(batches);
  size_t n = all.num_rows;
  std::vector<int> schema = batch_schema(all);
  std::vector<ColumnRef> kc;
  for (auto& k : keys) kc.push_back(evaluate_to_column(*k.expr, all, expr_type(*k.expr, schema)));
  std::vector<uint32_t> idx(n);
  for (size_t i = 0; i < n; i++) idx[i] = (uint32_t)i;
  auto less = [&](uint32_t a, uint32_t b) { return compare_rows(kc, keys, a, b) < 0; };
  if (limit >= 0 && (size_t)limit < n) {
    std::nth_element(idx.begin(), idx.begin() + limit, idx.end(), less);  // select_nth_unstable_by
    idx.resize((size_t)limit);
    std::stable_sort(idx.begin(), idx.end(), less);
  } else {
    std::stable_sort(idx.begin(), idx.end(), less);
  }
  out->num_rows = idx.size();
  out->sel = nullptr;
  out->columns.clear();
  for (auto& c : all.columns) out->columns.push_back(gather(*c, idx.data(), idx.size()));
  return true;
}

struct SortSink : Sink {  // sort.rs:145-268
  std::vector<SortKeyDef> keys;
  long long limit = -1;
  std::mutex mu;
  std::vector<Batch> global;
  Batch output;
  bool has_output = false;
  struct Local : LocalSink {
    SortSink* s;
    std::vector<Batch> batches;
    size_t rows = 0;
    explicit Local(SortSink* sink) : s(sink) {}
    void truncate() {  // sort.rs:187-196
      if (s->limit < 0) return;
      Batch b;
      if (sort_batches(batches, s->keys, s->limit, &b)) {
        rows = b.num_rows;
        batches.clear();
        batches.push_back(std::move(b));
      }
    }
    void sink(const Batch& b) override {  // sort.rs:239-250
      rows += b.active_rows();
      batches.push_back(compact(b));
      if (s->limit >= 0 && rows > std::max<size_t>(4 * (size_t)s->limit, 16384)) truncate();
    }
    void combine() override {
      truncate();
      std::lock_guard<std::mutex> g(s->mu);
      for (auto& b : batches) s->global.push_back(std::move(b));
    }
  };
  std::unique_ptr<LocalSink> create_local() override { return std::make_unique<Local>(this); }
  void finalize() override { has_output = sort_batches(global, keys, limit, &output); }
};

// CollectSink (operators/collect.rs:11-53), compacting.
struct CollectSink : Sink {
  std::mutex mu;
  std::vector<Batch> batches;
  struct Local : LocalSink {
    CollectSink* s;
    std::vector<Batch> batches;
    explicit Local(CollectSink* sink) : s(sink) {}
    void sink(const Batch& b) override { batches.push_back(compact(b)); }
    void combine() override {
      std::lock_guard<std::mutex> g(s->mu);
      for (auto& b : batches) s->batches.push_back(std::move(b));
    }
  };
  std::unique_ptr<LocalSink> create_local() override { return std::make_unique<Local>(this); }
  void finalize() override {}
};

// LimitSink (operators/limit.rs:26-117): workers append compacted batches
// under a mutex until limit + offset rows are collected (a flag then stops
// the others early); finalize concatenates and keeps rows [offset, offset +
// limit). With one thread the morsels arrive in order, so these are the
// first rows in source order.
struct LimitSink : Sink {
  size_t limit = 0, offset = 0;
  std::mutex mu;
  std::vector<Batch> batches;
  size_t have = 0;
  std::atomic<bool> done{false};
  struct Local : LocalSink {
    LimitSink* s;
    explicit Local(LimitSink* sink) : s(sink) {}
    void sink(const Batch& b) override {  // limit.rs:88-109
      if (s->done.load(std::memory_order_relaxed)) return;
      const size_t needed = s->limit > SIZE_MAX - s->offset ? SIZE_MAX : s->limit + s->offset;
      std::lock_guard<std::mutex> g(s->mu);
      if (s->have >= needed) return;
      Batch c = compact(b);
      s->have += c.num_rows;
      s->batches.push_back(std::move(c));
      if (s->have >= needed) s->done.store(true, std::memory_order_relaxed);
    }
    void combine() override {}
  };
  std::unique_ptr<LocalSink> create_local() override { return std::make_unique<Local>(this); }
  void finalize() override {}
};

// ============================================================ pipeline + scheduler
struct OpDef {
  int kind;
  ExprPtr pred;
  const JoinTable* table = nullptr;
  std::vector<ExprPtr> keys;
  bool build_is_left = false;
  std::vector<ExprPtr> exprs;
  std::vector<int> types;
};

struct SourceDef {
  Batch table;  // dense (a selection is compacted first)
  size_t morsel_rows;
  size_t batch_size;
  size_t num_morsels() const { return table.num_rows == 0 ? 0 : (table.num_rows + morsel_rows - 1) / morsel_rows; }
};

static Batch slice(const Batch& b, size_t off, size_t len) {  // Batch::slice (batch.rs:137-149)
  Batch o;
  o.num_rows = len;
  for (auto& c : b.columns) {
    auto s = std::make_shared<Column>(*c);
    s->len = len;
    s->data = static_cast<const uint8_t*>(c->data) + off * (size_t)c->width();
    if (c->valid) s->valid = std::make_shared<std::vector<uint8_t>>(c->valid->begin() + off, c->valid->begin() + off + len);
    o.columns.push_back(s);
  }
  return o;
}

// Pipeline::push (pipeline.rs:139-161)
static void push(const std::vector<OpDef>& ops, size_t idx, const Batch& batch, LocalSink& local, size_t batch_size) {
  if (batch.active_rows() == 0) return;
  if (idx == ops.size()) {
    local.sink(batch);
    return;
  }
  const OpDef& op = ops[idx];
  switch (op.kind) {
    case ML_PIPE_FILTER: {  // filter.rs:19-25
      auto sel = std::make_shared<std::vector<uint32_t>>(select(*op.pred, batch, batch.sel.get()));
      Batch b = batch;
      b.sel = sel;
      push(ops, idx + 1, b, local, batch_size);
      break;
    }
    case ML_PIPE_PROJECT: {  // projection.rs:21-32
      Batch b;
      b.num_rows = batch.num_rows;
      b.sel = batch.sel;
      for (size_t i = 0; i < op.exprs.size(); i++) b.columns.push_back(evaluate_to_column(*op.exprs[i], batch, op.types[i]));
      push(ops, idx + 1, b, local, batch_size);
      break;
    }
    case ML_PIPE_PROBE: {
      std::vector<Batch> out;
      probe_execute(*op.table, op.keys, op.build_is_left, batch_size, batch, out);
      for (auto& b : out) push(ops, idx + 1, b, local, batch_size);
      break;
    }
    default: fail(ML_ERR_PLAN, "unknown operator");
  }
}

// run_pipeline + run_parallel + AtomicMorselQueue (executor.rs:270-288,
// scheduler.rs:44-68,102-144, pipeline.rs:109-136).
static void run_pipeline(const SourceDef& src, const std::vector<OpDef>& ops, Sink& sink, int threads) {
  size_t n = src.num_morsels();
  std::atomic<size_t> next{0};
  std::atomic<bool> cancelled{false};
  int workers = (int)std::max<size_t>(1, std::min<size_t>((size_t)std::max(threads, 1), std::max<size_t>(n, 1)));
  std::vector<std::string> errors(workers);
  std::vector<int> codes(workers, 0);
  auto work = [&](int w) {
    try {
      auto local = sink.create_local();
      for (;;) {
        size_t m = next.fetch_add(1, std::memory_order_relaxed);
        if (m >= n || cancelled.load(std::memory_order_relaxed)) break;
        size_t begin = m * src.morsel_rows;
        size_t end = std::min(begin + src.morsel_rows, src.table.num_rows);
        for (size_t off = begin; off < end; off += src.batch_size) {
          size_t len = std::min(src.batch_size, end - off);
          push(ops, 0, slice(src.table, off, len), *local, src.batch_size);
        }
      }
      local->combine();
    } catch (const Error& e) {
      codes[w] = e.code;
      errors[w] = e.what();
      cancelled = true;
    } catch (const std::exception& e) {
      codes[w] = ML_ERR_INTERNAL;
      errors[w] = e.what();
      cancelled = true;
    }
  };
  if (workers <= 1) {
    work(0);
  } else {
    std::vector<std::thread> ts;
    for (int w = 0; w < workers; w++) ts.emplace_back(work, w);
    for (auto& t : ts) t.join();
  }
  for (int w = 0; w < workers; w++)  // first error by worker id wins
    if (codes[w]) fail(codes[w], errors[w]);
  sink.finalize();
}

}  // namespace ref

// ======================================================================= C API
using namespace ref;

struct mlo_table {
  Batch batch;
  std::vector<std::shared_ptr<std::vector<uint64_t>>> bitmaps;
  std::vector<std::shared_ptr<ml_dict>> dicts;
  std::vector<std::shared_ptr<std::vector<uint32_t>>> dict_offsets;
  std::vector<std::shared_ptr<std::string>> dict_bytes;
  std::vector<ml_column> cols;
  int32_t resizes = 0;
  int64_t capacity = 0;
};
struct mlo_join {
  std::unique_ptr<JoinTable> table;
};

static thread_local std::string g_err;

template <class F>
static int guarded(F f) {
  try {
    f();
    return ML_OK;
  } catch (const Error& e) {
    g_err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    g_err = e.what();
    return ML_ERR_INTERNAL;
  }
}

// ml_column (host pointers) -> oracle column; dictionaries are converted once per ml_dict.
static std::shared_ptr<const Dict> convert_dict(const ml_dict* d) {
  static std::mutex mu;
  static std::map<const ml_dict*, std::shared_ptr<const Dict>> cache;
  std::lock_guard<std::mutex> g(mu);
  auto it = cache.find(d);
  auto nd = std::make_shared<Dict>();
  for (int i = 0; i < d->len; i++) nd->entries.emplace_back((const char*)d->bytes + d->offsets[i], d->offsets[i + 1] - d->offsets[i]);
  if (it != cache.end() && it->second->entries == nd->entries) return it->second;
  cache[d] = nd;
  return nd;
}

static Batch import_batch(const ml_batch* b) {
  Batch out;
  out.num_rows = (size_t)b->num_rows;
  for (int i = 0; i < b->num_columns; i++) {
    const ml_column& mc = b->columns[i];
    auto c = std::make_shared<Column>();
    c->type = mc.type;
    c->len = (size_t)mc.len;
    c->data = mc.data;
    if ((int64_t)c->len != b->num_rows) fail(ML_ERR_INTERNAL, "column length differs from batch rows");
    if (mc.type == Utf8) {
      if (!mc.dict) fail(ML_ERR_UNSUPPORTED, "UTF8 columns must be dictionary-encoded");
      c->dict = convert_dict(mc.dict);
    }
    if (mc.validity) {
      auto v = std::make_shared<std::vector<uint8_t>>(c->len);
      bool any = false;
      for (size_t r = 0; r < c->len; r++) {
        (*v)[r] = (mc.validity[r / 64] >> (r % 64)) & 1;
        any |= !(*v)[r];
      }
      if (any) c->valid = v;
    }
    out.columns.push_back(c);
  }
  if (b->selection) out.sel = std::make_shared<std::vector<uint32_t>>(b->selection, b->selection + b->selection_len);
  return out;
}

static std::vector<int> schema_of(const Batch& b) { return batch_schema(b); }

static std::vector<OpDef> import_ops(const ml_operator* ops, int n, std::vector<int>& schema) {
  std::vector<OpDef> out;
  for (int i = 0; i < n; i++) {
    OpDef d;
    d.kind = ops[i].kind;
    switch (d.kind) {
      case ML_PIPE_FILTER:
        d.pred = build_expr(ops[i].predicate);
        if (expr_type(*d.pred, schema) != Boolean) fail(ML_ERR_EXECUTION, "filter predicate is not boolean");
        break;
      case ML_PIPE_PROBE: {
        auto* j = reinterpret_cast<const mlo_join*>(ops[i].table);
        d.table = j->table.get();
        d.build_is_left = ops[i].build_is_left != 0;
        for (int k = 0; k < ops[i].num_keys; k++) d.keys.push_back(build_expr(ops[i].keys[k]));
        if (d.keys.size() != d.table->key_types.size()) fail(ML_ERR_PLAN, "probe and build key counts differ");
        std::vector<int> ns;
        if (d.build_is_left) { ns = d.table->schema; ns.insert(ns.end(), schema.begin(), schema.end()); }
        else { ns = schema; ns.insert(ns.end(), d.table->schema.begin(), d.table->schema.end()); }
        schema = ns;
        break;
      }
      case ML_PIPE_PROJECT: {
        std::vector<int> ns;
        for (int k = 0; k < ops[i].num_exprs; k++) {
          d.exprs.push_back(build_expr(ops[i].exprs[k]));
          d.types.push_back(expr_type(*d.exprs.back(), schema));
        }
        schema = d.types;
        d.types = ns.empty() ? d.types : ns;
        break;
      }
      default: fail(ML_ERR_PLAN, "unknown operator kind");
    }
    out.push_back(std::move(d));
  }
  return out;
}

static SourceDef make_source(const ml_batch* b, const mlo_options* opt) {
  SourceDef s;
  s.table = compact(import_batch(b));
  s.morsel_rows = opt && opt->morsel_rows > 0 ? (size_t)opt->morsel_rows : 122880;
  s.batch_size = opt && opt->batch_size > 0 ? (size_t)opt->batch_size : 2048;
  return s;
}
static int threads_of(const mlo_options* opt) { return opt && opt->threads > 0 ? opt->threads : 1; }

static mlo_table* export_batch(Batch b) {
  auto* t = new mlo_table();
  t->batch = compact(b);
  for (auto& c : t->batch.columns) {
    ml_column mc{};
    mc.type = c->type;
    mc.data = c->data;
    mc.len = (int64_t)c->len;
    if (c->valid) {
      auto bm = std::make_shared<std::vector<uint64_t>>((c->len + 63) / 64, 0);
      for (size_t r = 0; r < c->len; r++)
        if ((*c->valid)[r]) (*bm)[r / 64] |= 1ull << (r % 64);
      t->bitmaps.push_back(bm);
      mc.validity = bm->data();
    }
    if (c->type == Utf8) {
      auto offs = std::make_shared<std::vector<uint32_t>>();
      auto bytes = std::make_shared<std::string>();
      offs->push_back(0);
      for (auto& e : c->dict->entries) {
        *bytes += e;
        offs->push_back((uint32_t)bytes->size());
      }
      auto d = std::make_shared<ml_dict>();
      d->len = (int32_t)c->dict->entries.size();
      d->offsets = offs->data();
      d->bytes = (const uint8_t*)bytes->data();
      t->dict_offsets.push_back(offs);
      t->dict_bytes.push_back(bytes);
      t->dicts.push_back(d);
      mc.dict = d.get();
    }
    t->cols.push_back(mc);
  }
  return t;
}

extern "C" {

const char* mlo_last_error(void) { return g_err.c_str(); }

uint64_t mlo_fold_mul(uint64_t a, uint64_t b) { return fold_mul(a, b); }
uint64_t mlo_hash_u64(uint64_t x) { return hash_u64(x); }
uint64_t mlo_combine(uint64_t h, uint64_t v) { return combine(h, v); }
uint64_t mlo_hash_bytes(const uint8_t* b, uint64_t n) { return hash_bytes(b, (size_t)n); }
uint64_t mlo_match_tag(uint64_t group, uint8_t tag) { return match_tag(group, tag); }
uint64_t mlo_match_empty(uint64_t group) { return match_empty(group); }

// SwissTable group-by over u64 keys in batches (the shape of
// crates/minilake-hashtable/tests/proptest_tables.rs:14-40). weak != 0 uses the
// collision hash x % 4. out_ids[i] = payload of key i; returns distinct count.
int64_t mlo_swiss_group_ids(const uint64_t* keys, int64_t n, int64_t batch, int32_t weak, uint32_t* out_ids,
                            int64_t* out_capacity, int32_t* out_resizes) {
  SwissTable t = SwissTable::with_capacity(0);
  std::vector<uint64_t> stored;
  std::vector<uint32_t> out;
  struct U64Keys {  // lib.rs:39-57
    const uint64_t* input;
    std::vector<uint64_t>* stored;
    bool equals(uint32_t p, size_t k) const { return (*stored)[p] == input[k]; }
    uint32_t insert(size_t k) { stored->push_back(input[k]); return (uint32_t)(stored->size() - 1); }
  };
  if (batch <= 0) batch = n > 0 ? n : 1;
  for (int64_t off = 0; off < n; off += batch) {
    int64_t len = std::min(batch, n - off);
    std::vector<uint64_t> hs((size_t)len);
    for (int64_t k = 0; k < len; k++) hs[k] = weak ? keys[off + k] % 4 : hash_u64(keys[off + k]);
    U64Keys ks{keys + off, &stored};
    t.find_or_insert_batch(hs.data(), (size_t)len, ks, out);
    for (int64_t k = 0; k < len; k++) out_ids[off + k] = out[k];
  }
  // every stored key must be found again with its id (proptest_tables.rs:33-39)
  for (size_t p = 0; p < stored.size(); p++) {
    uint64_t h = weak ? stored[p] % 4 : hash_u64(stored[p]);
    uint32_t got;
    if (!t.find(h, [&](uint32_t q) { return stored[q] == stored[p]; }, &got) || got != p) return -1;
  }
  if (out_capacity) *out_capacity = (int64_t)t.capacity();
  if (out_resizes) *out_resizes = t.resizes;
  return (int64_t)stored.size();
}

int mlo_expr_type(const ml_expr* e, const ml_batch* b, int32_t* out) {
  return guarded([&] { *out = expr_type(*build_expr(*e), schema_of(import_batch(b))); });
}

int mlo_filter(const ml_expr* pred, const ml_batch* in, uint32_t* out_sel, int64_t* out_len) {
  return guarded([&] {
    Batch b = import_batch(in);
    auto sel = select(*build_expr(*pred), b, b.sel.get());
    std::memcpy(out_sel, sel.data(), sel.size() * 4);
    *out_len = (int64_t)sel.size();
  });
}

int mlo_project(const ml_expr* exprs, int32_t n, const ml_batch* in, mlo_table** out) {
  return guarded([&] {
    Batch b = import_batch(in);
    std::vector<int> schema = schema_of(b);
    Batch o;
    o.num_rows = b.num_rows;
    for (int i = 0; i < n; i++) {
      auto e = build_expr(exprs[i]);
      o.columns.push_back(evaluate_to_column(*e, b, expr_type(*e, schema)));
    }
    *out = export_batch(o);
  });
}

int mlo_hash_columns(const ml_expr* keys, int32_t nkeys, const ml_batch* in, uint64_t* out) {
  return guarded([&] {
    Batch b = import_batch(in);
    std::vector<int> schema = schema_of(b);
    std::vector<ColumnRef> cols;
    for (int i = 0; i < nkeys; i++) {
      auto e = build_expr(keys[i]);
      cols.push_back(evaluate_to_column(*e, b, expr_type(*e, schema)));
    }
    std::vector<uint64_t> h;
    hash_columns(cols, batch_rows(b), h);
    std::memcpy(out, h.data(), h.size() * 8);
  });
}

int mlo_pipeline_hashagg(const ml_batch* src, const ml_operator* ops, int32_t nops, const ml_agg_spec* spec,
                         const mlo_options* opt, mlo_table** out) {
  return guarded([&] {
    SourceDef s = make_source(src, opt);
    std::vector<int> schema = schema_of(s.table);
    auto opdefs = import_ops(ops, nops, schema);
    std::vector<AggDef> aggs;
    for (int a = 0; a < spec->num_aggs; a++) {
      AggDef d;
      d.func = spec->aggs[a].func;
      d.arg_type = spec->aggs[a].arg_type;
      if (d.func != ML_AGG_COUNT_STAR) d.arg = build_expr(spec->aggs[a].arg);
      aggs.push_back(d);
    }
    if (spec->num_keys == 0) {  // no GROUP BY: the planner's UngroupedAggregate
      UngroupedAggSink sink;
      sink.aggs = aggs;
      sink.mode = spec->mode;
      run_pipeline(s, opdefs, sink, threads_of(opt));
      *out = export_batch(sink.output);
      return;
    }
    HashAggSink sink;
    for (int k = 0; k < spec->num_keys; k++) {
      sink.group_exprs.push_back(build_expr(spec->group_exprs[k]));
      sink.key_types.push_back(spec->key_types[k]);
    }
    sink.aggs = aggs;
    sink.mode = spec->mode;
    run_pipeline(s, opdefs, sink, threads_of(opt));
    *out = export_batch(sink.output);
    (*out)->resizes = sink.resizes;
    (*out)->capacity = (int64_t)sink.capacity;
  });
}

int mlo_pipeline_join_build(const ml_batch* src, const ml_operator* ops, int32_t nops, const ml_join_build_spec* spec,
                            const mlo_options* opt, mlo_join** out) {
  return guarded([&] {
    SourceDef s = make_source(src, opt);
    std::vector<int> schema = schema_of(s.table);
    auto opdefs = import_ops(ops, nops, schema);
    JoinBuildSink sink;
    for (int k = 0; k < spec->num_keys; k++) {
      sink.keys.push_back(build_expr(spec->keys[k]));
      sink.key_types.push_back(spec->key_types[k]);
    }
    sink.schema = schema;
    run_pipeline(s, opdefs, sink, threads_of(opt));
    auto* j = new mlo_join();
    j->table = std::move(sink.table);
    *out = j;
  });
}

int64_t mlo_join_num_rows(const mlo_join* j) { return (int64_t)j->table->batch.num_rows; }
int64_t mlo_join_num_keys(const mlo_join* j) { return (int64_t)j->table->table.len; }
void mlo_join_free(mlo_join* j) { delete j; }

int mlo_join_probe(const mlo_join* j, const ml_expr* keys, int32_t nkeys, const ml_batch* probe, mlo_table** out) {
  return guarded([&] {
    Batch b = import_batch(probe);
    std::vector<ExprPtr> ks;
    for (int k = 0; k < nkeys; k++) ks.push_back(build_expr(keys[k]));
    std::vector<std::pair<uint32_t, uint32_t>> pairs;
    std::vector<Batch> unused;
    probe_execute(*j->table, ks, false, 2048, b, unused, &pairs);
    std::vector<int64_t> p(pairs.size()), q(pairs.size());
    for (size_t i = 0; i < pairs.size(); i++) { p[i] = pairs[i].first; q[i] = pairs[i].second; }
    Batch o;
    o.num_rows = pairs.size();
    o.columns.push_back(make_column(Int64, std::move(p)));
    o.columns.push_back(make_column(Int64, std::move(q)));
    *out = export_batch(o);
  });
}

int mlo_pipeline_topn(const ml_batch* src, const ml_operator* ops, int32_t nops, const ml_sort_spec* spec,
                      const mlo_options* opt, mlo_table** out) {
  return guarded([&] {
    SourceDef s = make_source(src, opt);
    std::vector<int> schema = schema_of(s.table);
    auto opdefs = import_ops(ops, nops, schema);
    SortSink sink;
    for (int k = 0; k < spec->num_keys; k++)
      sink.keys.push_back({build_expr(spec->keys[k].expr), spec->keys[k].asc != 0, spec->keys[k].nulls_first != 0});
    sink.limit = spec->limit;
    run_pipeline(s, opdefs, sink, threads_of(opt));
    if (sink.has_output) {
      *out = export_batch(sink.output);
    } else {
      Batch empty;
      for (int t : schema) {
        switch (t) {
          case Int32: case Date: empty.columns.push_back(make_column(t, std::vector<int32_t>())); break;
          case Boolean: empty.columns.push_back(make_column(t, std::vector<uint8_t>())); break;
          case Float64: empty.columns.push_back(make_column(t, std::vector<double>())); break;
          case Utf8: empty.columns.push_back(make_column(t, std::vector<uint32_t>(), nullptr, std::make_shared<Dict>())); break;
          default: empty.columns.push_back(make_column(Int64, std::vector<int64_t>())); break;
        }
      }
      *out = export_batch(empty);
    }
  });
}

static Batch empty_batch(const std::vector<int>& schema) {
  Batch empty;
  for (int t : schema) {
    if (t == Int32 || t == Date) empty.columns.push_back(make_column(t, std::vector<int32_t>()));
    else if (t == Boolean) empty.columns.push_back(make_column(t, std::vector<uint8_t>()));
    else if (t == Float64) empty.columns.push_back(make_column(t, std::vector<double>()));
    else if (t == Utf8) empty.columns.push_back(make_column(t, std::vector<uint32_t>(), nullptr, std::make_shared<Dict>()));
    else empty.columns.push_back(make_column(Int64, std::vector<int64_t>()));
  }
  return empty;
}

int mlo_pipeline_collect(const ml_batch* src, const ml_operator* ops, int32_t nops, const mlo_options* opt,
                         mlo_table** out) {
  return guarded([&] {
    SourceDef s = make_source(src, opt);
    std::vector<int> schema = schema_of(s.table);
    auto opdefs = import_ops(ops, nops, schema);
    CollectSink sink;
    run_pipeline(s, opdefs, sink, threads_of(opt));
    *out = export_batch(sink.batches.empty() ? empty_batch(schema) : concat(sink.batches));
  });
}

int mlo_pipeline_limit(const ml_batch* src, const ml_operator* ops, int32_t nops, int64_t limit, int64_t offset,
                       const mlo_options* opt, mlo_table** out) {
  return guarded([&] {
    if (limit < 0 || offset < 0) throw std::runtime_error("negative LIMIT or OFFSET");
    SourceDef s = make_source(src, opt);
    std::vector<int> schema = schema_of(s.table);
    auto opdefs = import_ops(ops, nops, schema);
    LimitSink sink;
    sink.limit = (size_t)limit;
    sink.offset = (size_t)offset;
    run_pipeline(s, opdefs, sink, threads_of(opt));
    if (sink.batches.empty()) {
      *out = export_batch(empty_batch(schema));
      return;
    }
    Batch all = concat(sink.batches);  // limit.rs:64-83
    const size_t start = std::min(sink.offset, all.num_rows);
    const size_t len = std::min(sink.limit, all.num_rows - start);
    *out = export_batch(slice(all, start, len));
  });
}

int64_t mlo_table_num_rows(const mlo_table* t) { return (int64_t)t->batch.num_rows; }
int32_t mlo_table_num_columns(const mlo_table* t) { return (int32_t)t->cols.size(); }
const ml_column* mlo_table_column(const mlo_table* t, int32_t i) { return &t->cols[(size_t)i]; }
int32_t mlo_table_resizes(const mlo_table* t) { return t->resizes; }
int64_t mlo_table_capacity(const mlo_table* t) { return t->capacity; }
void mlo_table_free(mlo_table* t) { delete t; }

}  // extern "C"

// ---------------------------------------------------------------- MLB1 format
// batch := "MLB1" num_cols:u32 num_rows:u64 column*; column := type:u8
// has_validity:u8 [validity words] payload, little-endian (ipc.rs:1-13).
static int mlb1_tag(int t) {  // type_tag (ipc.rs:21-30)
  switch (t) {
    case Boolean: return 0;
    case Int32: return 1;
    case Int64: return 2;
    case Float64: return 3;
    case Date: return 4;
    case Utf8: return 5;
  }
  fail(ML_ERR_INTERNAL, "bad column type");
}
static int mlb1_type(int tag) {  // tag_type (ipc.rs:32-42)
  if (tag < 0 || tag > 5) fail(ML_ERR_INTERNAL, "bad type tag " + std::to_string(tag));
  static const int types[] = {Boolean, Int32, Int64, Float64, Date, Utf8};
  return types[tag];
}
template <class T>
static void put_le(std::string& w, T x) {
  w.append(reinterpret_cast<const char*>(&x), sizeof(T));  // x86: already little-endian
}

// write_batch (ipc.rs:44-87): the selection is applied first; a column has
// a validity section when any kept row is NULL (Column::new drops all-valid
// bitmaps); dictionary columns are decoded to plain strings.
static void mlb1_write(const Batch& batch, std::string& w) {
  Batch b = compact(batch);
  const size_t n = b.num_rows;
  w += "MLB1";
  put_le<uint32_t>(w, (uint32_t)b.columns.size());
  put_le<uint64_t>(w, (uint64_t)n);
  for (auto& c : b.columns) {
    w.push_back((char)mlb1_tag(c->type));
    w.push_back(c->valid ? 1 : 0);
    if (c->valid) {
      std::vector<uint64_t> words((n + 63) / 64, 0);
      for (size_t r = 0; r < n; r++)
        if ((*c->valid)[r]) words[r / 64] |= 1ull << (r % 64);
      for (uint64_t x : words) put_le(w, x);
    }
    switch (c->type) {
      case Boolean:
        for (size_t r = 0; r < n; r++) w.push_back(c->as<uint8_t>()[r] != 0 ? 1 : 0);
        break;
      case Int32: case Date:
        for (size_t r = 0; r < n; r++) put_le(w, c->as<int32_t>()[r]);
        break;
      case Int64:
        for (size_t r = 0; r < n; r++) put_le(w, c->as<int64_t>()[r]);
        break;
      case Float64:
        for (size_t r = 0; r < n; r++) put_le(w, c->as<double>()[r]);
        break;
      case Utf8: {
        std::string bytes;
        put_le<uint32_t>(w, 0);
        for (size_t r = 0; r < n; r++) {
          bytes += c->dict->entries[c->as<uint32_t>()[r]];
          put_le<uint32_t>(w, (uint32_t)bytes.size());
        }
        put_le<uint64_t>(w, (uint64_t)bytes.size());
        w += bytes;
        break;
      }
    }
  }
}

// read_batch (ipc.rs:111-170). A stream shorter than the magic is a clean end
// (nullopt); a batch cut short is an I/O error; a bad magic, type tag or
// string offsets are internal errors. Strings come back as a dictionary
// column with one entry per row.
static bool mlb1_read(const uint8_t* p, size_t len, size_t* used, Batch* out) {
  size_t pos = 0;
  auto take = [&](size_t k) {
    if (len - pos < k) fail(ML_ERR_IO, "failed to fill whole buffer");
    const uint8_t* q = p + pos;
    pos += k;
    return q;
  };
  auto u32 = [&]() { uint32_t x; std::memcpy(&x, take(4), 4); return x; };
  auto u64 = [&]() { uint64_t x; std::memcpy(&x, take(8), 8); return x; };
  if (len < 4) return false;
  if (std::memcmp(take(4), "MLB1", 4) != 0) fail(ML_ERR_INTERNAL, "bad batch magic");
  const size_t ncols = u32();
  const size_t n = (size_t)u64();
  out->columns.clear();
  out->sel = nullptr;
  out->num_rows = n;
  for (size_t ci = 0; ci < ncols; ci++) {
    const uint8_t* hdr = take(2);
    const int t = mlb1_type(hdr[0]);
    std::shared_ptr<std::vector<uint8_t>> valid;
    if (hdr[1] == 1) {
      valid = std::make_shared<std::vector<uint8_t>>(n);
      for (size_t wi = 0; wi < (n + 63) / 64; wi++) {
        const uint64_t x = u64();
        for (size_t r = wi * 64; r < std::min(n, wi * 64 + 64); r++) (*valid)[r] = (x >> (r % 64)) & 1;
      }
    }
    ColumnRef c;
    switch (t) {
      case Boolean: {
        const uint8_t* q = take(n);
        std::vector<uint8_t> v(n);
        for (size_t r = 0; r < n; r++) v[r] = q[r] != 0;
        c = make_column(t, std::move(v), valid);
        break;
      }
      case Int32: case Date: {
        std::vector<int32_t> v(n);
        if (n) std::memcpy(v.data(), take(4 * n), 4 * n);
        c = make_column(t, std::move(v), valid);
        break;
      }
      case Int64: {
        std::vector<int64_t> v(n);
        if (n) std::memcpy(v.data(), take(8 * n), 8 * n);
        c = make_column(t, std::move(v), valid);
        break;
      }
      case Float64: {
        std::vector<double> v(n);
        if (n) std::memcpy(v.data(), take(8 * n), 8 * n);
        c = make_column(t, std::move(v), valid);
        break;
      }
      default: {  // Utf8: StringVec::from_parts (column.rs:101-110) checks the offsets
        std::vector<uint32_t> offs(n + 1);
        for (auto& o : offs) o = u32();
        const uint64_t blen = u64();
        const uint8_t* bytes = take((size_t)blen);
        bool ok = offs[0] == 0 && offs[n] == blen;
        for (size_t r = 0; r < n && ok; r++) ok = offs[r] <= offs[r + 1];
        if (!ok) fail(ML_ERR_INTERNAL, "corrupt string offsets");
        auto d = std::make_shared<Dict>();
        std::vector<uint32_t> codes(n);
        for (size_t r = 0; r < n; r++) {
          d->entries.emplace_back(reinterpret_cast<const char*>(bytes) + offs[r], offs[r + 1] - offs[r]);
          codes[r] = (uint32_t)r;
        }
        c = make_column(Utf8, std::move(codes), valid, d);
        break;
      }
    }
    out->columns.push_back(c);
  }
  *used = pos;
  return true;
}

extern "C" {

int mlo_mlb1_write(const ml_batch* src, uint8_t** out, int64_t* len) {
  return guarded([&] {
    std::string w;
    mlb1_write(import_batch(src), w);
    *out = static_cast<uint8_t*>(std::malloc(std::max<size_t>(w.size(), 1)));
    if (!*out) fail(ML_ERR_RESOURCES_EXHAUSTED, "out of memory");
    std::memcpy(*out, w.data(), w.size());
    *len = (int64_t)w.size();
  });
}

void mlo_bytes_free(uint8_t* p) { std::free(p); }

int mlo_mlb1_read(const uint8_t* src, int64_t len, int64_t* consumed, mlo_table** out) {
  *out = nullptr;
  *consumed = 0;
  return guarded([&] {
    Batch b;
    size_t used = 0;
    if (!mlb1_read(src, (size_t)std::max<int64_t>(len, 0), &used, &b)) return;
    *out = export_batch(std::move(b));
    *consumed = (int64_t)used;
  });
}

}  // extern "C"
