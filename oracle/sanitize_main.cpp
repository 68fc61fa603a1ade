// CPU ORACLE — test infrastructure only: a driver that runs the oracle's
// operators (hash, filter, projection, hash aggregate with partial-state
// merges across threads, join build with duplicate and NULL keys + probe,
// TopN, full sort, LIMIT, MLB1 round trip, SwissTable with colliding hashes)
// on seeded inputs, so the oracle can be built with AddressSanitizer +
// UndefinedBehaviorSanitizer and with ThreadSanitizer (oracle/Makefile,
// target `sanitize`). It checks return codes and a few invariants; the
// sanitizers check the memory and threading behaviour. Exit 0 = clean.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "ref_oracle.h"

namespace {

int failures = 0;
void check(bool ok, const char* what) {
  if (!ok) {
    std::fprintf(stderr, "FAILED: %s (%s)\n", what, mlo_last_error());
    failures++;
  }
}

ml_expr_node col_node(int c) {
  ml_expr_node n{};
  n.kind = ML_EXPR_COLUMN;
  n.column = c;
  return n;
}
ml_expr_node lit_i64(int type, int64_t v) {
  ml_expr_node n{};
  n.kind = ML_EXPR_LITERAL;
  n.type = type;
  n.i64 = v;
  return n;
}
ml_expr_node lit_f64(double v) {
  ml_expr_node n{};
  n.kind = ML_EXPR_LITERAL;
  n.type = ML_TYPE_FLOAT64;
  n.f64 = v;
  return n;
}
ml_expr_node binary(int op) {
  ml_expr_node n{};
  n.kind = ML_EXPR_BINARY;
  n.op = op;
  return n;
}
ml_expr expr_of(const std::vector<ml_expr_node>& v) {
  ml_expr e{};
  e.nodes = v.data();
  e.num_nodes = (int32_t)v.size();
  return e;
}

struct Table {
  std::vector<int32_t> i32;
  std::vector<int64_t> i64;
  std::vector<double> f64;
  std::vector<uint32_t> codes;
  std::vector<uint64_t> valid;  // i64 column validity
  std::vector<uint32_t> dict_off{0};
  std::vector<uint8_t> dict_bytes;
  ml_dict dict{};
  std::vector<ml_column> cols;
  ml_batch batch{};
};

// columns: 0 i32 in [-50, 50), 1 i64 key in [0, keys) with ~3% NULLs,
// 2 f64 (some -0.0 and NaN), 3 dictionary string with entries not in byte order
void make_table(Table& t, int64_t n, int64_t keys, uint64_t seed) {
  std::mt19937_64 rng(seed);
  t.i32.resize(n);
  t.i64.resize(n);
  t.f64.resize(n);
  t.codes.resize(n);
  t.valid.assign((n + 63) / 64, ~0ull);
  for (int64_t r = 0; r < n; r++) {
    t.i32[r] = (int32_t)(rng() % 100) - 50;
    t.i64[r] = (int64_t)(rng() % (uint64_t)keys);
    t.f64[r] = (double)(rng() % 100000) / 7.0 - 5000.0;
    if (r % 97 == 0) t.f64[r] = -0.0;
    if (r % 211 == 0) t.f64[r] = std::nan("");
    t.codes[r] = (uint32_t)(rng() % 4);
    if (rng() % 33 == 0) t.valid[r / 64] &= ~(1ull << (r % 64));
  }
  const char* entries[4] = {"pear", "apple", "", "Fig"};
  for (const char* e : entries) {
    t.dict_bytes.insert(t.dict_bytes.end(), e, e + std::strlen(e));
    t.dict_off.push_back((uint32_t)t.dict_bytes.size());
  }
  t.dict.len = 4;
  t.dict.offsets = t.dict_off.data();
  t.dict.bytes = t.dict_bytes.data();
  t.cols.resize(4);
  t.cols[0] = ml_column{ML_TYPE_INT32, 0, t.i32.data(), n, nullptr, nullptr};
  t.cols[1] = ml_column{ML_TYPE_INT64, 0, t.i64.data(), n, t.valid.data(), nullptr};
  t.cols[2] = ml_column{ML_TYPE_FLOAT64, 0, t.f64.data(), n, nullptr, nullptr};
  t.cols[3] = ml_column{ML_TYPE_UTF8, 0, t.codes.data(), n, nullptr, &t.dict};
  t.batch.columns = t.cols.data();
  t.batch.num_columns = 4;
  t.batch.num_rows = n;
}

}  // namespace

int main() {
  const int64_t n = 200000;
  Table t;
  make_table(t, n, 5000, 7);
  mlo_options opt{4, 2048, 8192};  // four threads, small morsels: many partial-state merges

  // hashes of the four columns
  {
    std::vector<ml_expr_node> k0{col_node(0)}, k1{col_node(1)}, k2{col_node(2)}, k3{col_node(3)};
    ml_expr keys[4] = {expr_of(k0), expr_of(k1), expr_of(k2), expr_of(k3)};
    std::vector<uint64_t> h(n);
    check(mlo_hash_columns(keys, 4, &t.batch, h.data()) == 0, "hash_columns");
  }
  // filter i32 < 0 AND f64 > 0.0, then a projection with a division by i32
  std::vector<ml_expr_node> pred{col_node(0), lit_i64(ML_TYPE_INT32, 0), binary(ML_OP_LT), col_node(2),
                                 lit_f64(0.0), binary(ML_OP_GT), binary(ML_OP_AND)};
  {
    std::vector<uint32_t> sel(n);
    int64_t len = 0;
    ml_expr p = expr_of(pred);
    check(mlo_filter(&p, &t.batch, sel.data(), &len) == 0 && len > 0 && len < n, "filter");
    std::vector<ml_expr_node> e0{col_node(2), col_node(0), binary(ML_OP_DIV)}, e1{col_node(1), col_node(0),
                                                                                 binary(ML_OP_MOD)};
    ml_expr ex[2] = {expr_of(e0), expr_of(e1)};
    mlo_table* out = nullptr;
    check(mlo_project(ex, 2, &t.batch, &out) == 0 && mlo_table_num_rows(out) == n, "project");
    mlo_table_free(out);
  }
  // GROUP BY (i64, string) with COUNT(*), SUM(i32), AVG(f64), MIN(f64), MAX(i64), after the filter
  {
    std::vector<ml_expr_node> g0{col_node(1)}, g1{col_node(3)}, a0{col_node(0)}, a1{col_node(2)}, a2{col_node(1)};
    ml_expr groups[2] = {expr_of(g0), expr_of(g1)};
    int32_t key_types[2] = {ML_TYPE_INT64, ML_TYPE_UTF8};
    ml_aggregate aggs[5] = {{ML_AGG_COUNT_STAR, ML_TYPE_INT64, ml_expr{}}, {ML_AGG_SUM, ML_TYPE_INT32, expr_of(a0)},
                            {ML_AGG_AVG, ML_TYPE_FLOAT64, expr_of(a1)}, {ML_AGG_MIN, ML_TYPE_FLOAT64, expr_of(a1)},
                            {ML_AGG_MAX, ML_TYPE_INT64, expr_of(a2)}};
    for (int mode : {ML_AGG_MODE_SINGLE, ML_AGG_MODE_PARTIAL}) {
      ml_agg_spec spec{groups, key_types, 2, 5, aggs, mode, ML_SIZING_PLAIN, 0};
      ml_operator op{};
      op.kind = ML_PIPE_FILTER;
      op.predicate = expr_of(pred);
      mlo_table* out = nullptr;
      check(mlo_pipeline_hashagg(&t.batch, &op, 1, &spec, &opt, &out) == 0 && mlo_table_num_rows(out) > 0,
            "hash aggregate");
      mlo_table_free(out);
    }
    ml_agg_spec ungrouped{nullptr, nullptr, 0, 5, aggs, ML_AGG_MODE_SINGLE, ML_SIZING_PLAIN, 0};
    mlo_table* out = nullptr;
    check(mlo_pipeline_hashagg(&t.batch, nullptr, 0, &ungrouped, &opt, &out) == 0 && mlo_table_num_rows(out) == 1,
          "ungrouped aggregate");
    mlo_table_free(out);
  }
  // join: build on the i64 key (duplicates and NULL keys), probe a second table
  {
    Table b;
    make_table(b, 20000, 5000, 11);
    std::vector<ml_expr_node> k{col_node(1)};
    ml_expr keys[1] = {expr_of(k)};
    int32_t kt[1] = {ML_TYPE_INT64};
    ml_join_build_spec spec{keys, kt, 1, 0};
    mlo_join* j = nullptr;
    check(mlo_pipeline_join_build(&b.batch, nullptr, 0, &spec, &opt, &j) == 0 && mlo_join_num_rows(j) == 20000,
          "join build");
    mlo_table* pairs = nullptr;
    check(mlo_join_probe(j, keys, 1, &t.batch, &pairs) == 0 && mlo_table_num_rows(pairs) > n, "join probe");
    mlo_table_free(pairs);
    ml_operator op{};
    op.kind = ML_PIPE_PROBE;
    op.table = (const ml_join_table*)j;
    op.keys = keys;
    op.num_keys = 1;
    mlo_table* out = nullptr;
    check(mlo_pipeline_collect(&t.batch, &op, 1, &opt, &out) == 0 && mlo_table_num_columns(out) == 8,
          "probe pipeline collect");
    mlo_table_free(out);
    mlo_join_free(j);
  }
  // TopN and a full sort over (f64 DESC NULLS FIRST, string ASC), LIMIT/OFFSET
  {
    std::vector<ml_expr_node> s0{col_node(2)}, s1{col_node(3)};
    ml_sort_key keys[2] = {{expr_of(s0), 0, 1}, {expr_of(s1), 1, 0}};
    for (int64_t limit : {int64_t(10), int64_t(-1)}) {
      ml_sort_spec spec{keys, 2, 0, limit};
      mlo_table* out = nullptr;
      check(mlo_pipeline_topn(&t.batch, nullptr, 0, &spec, &opt, &out) == 0 &&
                mlo_table_num_rows(out) == (limit < 0 ? n : limit),
            "topn / sort");
      mlo_table_free(out);
    }
    mlo_table* out = nullptr;
    check(mlo_pipeline_limit(&t.batch, nullptr, 0, 100, n - 30, &opt, &out) == 0 && mlo_table_num_rows(out) == 30,
          "limit");
    mlo_table_free(out);
  }
  // MLB1 write and read back, with a selection
  {
    std::vector<uint32_t> sel;
    for (uint32_t r = 0; r < (uint32_t)n; r += 3) sel.push_back(r);
    ml_batch b = t.batch;
    b.selection = sel.data();
    b.selection_len = (int64_t)sel.size();
    uint8_t* bytes = nullptr;
    int64_t len = 0;
    check(mlo_mlb1_write(&b, &bytes, &len) == 0 && len > 0, "mlb1 write");
    mlo_table* back = nullptr;
    int64_t used = 0;
    check(mlo_mlb1_read(bytes, len, &used, &back) == 0 && used == len &&
              mlo_table_num_rows(back) == (int64_t)sel.size(),
          "mlb1 read");
    mlo_table_free(back);
    mlo_bytes_free(bytes);
  }
  // SwissTable with a hash that collides constantly (x % 4), across growth
  {
    std::vector<uint64_t> keys(6000);
    for (size_t i = 0; i < keys.size(); i++) keys[i] = (uint64_t)(i * 2654435761u % 2003);
    std::vector<uint32_t> ids(keys.size());
    int64_t cap = 0;
    int32_t resizes = 0;
    check(mlo_swiss_group_ids(keys.data(), (int64_t)keys.size(), 1024, 1, ids.data(), &cap, &resizes) == 2003,
          "swiss table (colliding hash)");
  }
  if (failures) {
    std::fprintf(stderr, "%d check(s) failed\n", failures);
    return 1;
  }
  std::printf("sanitize driver: all checks passed\n");
  return 0;
}
