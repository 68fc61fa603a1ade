// CPU ORACLE — test infrastructure only (see ref_core.h).
// Batch-at-a-time expression evaluation, restating
// crates/minilake-exec/src/expr/eval.rs:61-346, kernels/{cmp,arith,boolean,cast}.rs
// and kernels/hash.rs.
#include "ref_expr.h"

#include <cmath>
#include <limits>

#include "ref_hash.h"

namespace ref {

// ------------------------------------------------------------------ columns
ColumnRef gather(const Column& c, const uint32_t* idx, size_t n) {  // column.rs:380-398
  auto out = std::make_shared<Column>();
  out->type = c.type;
  out->len = n;
  out->dict = c.dict;
  int w = c.width();
  auto store = std::make_shared<std::vector<uint8_t>>(n * (size_t)w);
  const uint8_t* src = static_cast<const uint8_t*>(c.data);
  for (size_t i = 0; i < n; i++) std::memcpy(store->data() + i * w, src + (size_t)idx[i] * w, w);
  out->data = store->data();
  out->owner = store;
  if (c.valid) {
    auto v = std::make_shared<std::vector<uint8_t>>(n);
    bool any_null = false;
    for (size_t i = 0; i < n; i++) {
      (*v)[i] = (*c.valid)[idx[i]];
      any_null |= (*v)[i] == 0;
    }
    if (any_null) out->valid = v;
  }
  return out;
}

Batch compact(const Batch& b) {  // batch.rs:121-134
  if (!b.sel) return b;
  Batch out;
  out.num_rows = b.sel->size();
  for (auto& c : b.columns) out.columns.push_back(gather(*c, b.sel->data(), b.sel->size()));
  return out;
}

Batch concat(const std::vector<Batch>& batches) {  // batch.rs:168-184
  if (batches.empty()) fail(ML_ERR_INTERNAL, "concat of zero batches");
  std::vector<Batch> dense;
  dense.reserve(batches.size());
  for (auto& b : batches) dense.push_back(compact(b));
  size_t rows = 0;
  for (auto& b : dense) rows += b.num_rows;
  Batch out;
  out.num_rows = rows;
  for (size_t ci = 0; ci < dense[0].columns.size(); ci++) {
    const Column& first = *dense[0].columns[ci];
    int w = first.width();
    auto c = std::make_shared<Column>();
    c->type = first.type;
    c->len = rows;
    c->dict = first.dict;
    auto store = std::make_shared<std::vector<uint8_t>>(rows * (size_t)w);
    std::shared_ptr<std::vector<uint8_t>> valid;
    bool any_valid_map = false;
    for (auto& b : dense) any_valid_map |= (bool)b.columns[ci]->valid;
    if (any_valid_map) valid = std::make_shared<std::vector<uint8_t>>(rows, 1);
    size_t off = 0;
    for (auto& b : dense) {
      const Column& part = *b.columns[ci];
      if (part.type != first.type) fail(ML_ERR_INTERNAL, "concat of different column types");
      if (part.len) std::memcpy(store->data() + off * w, part.data, part.len * (size_t)w);
      if (valid && part.valid)
        for (size_t i = 0; i < part.len; i++) (*valid)[off + i] = (*part.valid)[i];
      off += part.len;
    }
    c->data = store->data();
    c->owner = store;
    c->valid = valid;
    out.columns.push_back(c);
  }
  return out;
}

ColumnRef column_from_scalar(const Scalar& s, size_t n, int dt) {  // Column::from_scalar
  switch (dt) {
    case Boolean: {
      std::vector<uint8_t> v(n, (uint8_t)(s.i != 0));
      auto valid = s.is_null() ? std::make_shared<std::vector<uint8_t>>(n, 0) : nullptr;
      return make_column(Boolean, std::move(v), valid);
    }
    case Int32: case Date: {
      std::vector<int32_t> v(n, (int32_t)s.i);
      auto valid = s.is_null() ? std::make_shared<std::vector<uint8_t>>(n, 0) : nullptr;
      return make_column(dt, std::move(v), valid);
    }
    case Int64: {
      std::vector<int64_t> v(n, s.i);
      auto valid = s.is_null() ? std::make_shared<std::vector<uint8_t>>(n, 0) : nullptr;
      return make_column(Int64, std::move(v), valid);
    }
    case Float64: {
      double f = 0;
      if (!s.is_null() && !s.as_f64(&f)) fail(ML_ERR_EXECUTION, "not a float constant");
      std::vector<double> v(n, f);
      auto valid = s.is_null() ? std::make_shared<std::vector<uint8_t>>(n, 0) : nullptr;
      return make_column(Float64, std::move(v), valid);
    }
    case Utf8: {
      auto d = std::make_shared<Dict>();
      d->entries.push_back(s.s);
      std::vector<uint32_t> v(n, 0);
      auto valid = s.is_null() ? std::make_shared<std::vector<uint8_t>>(n, 0) : nullptr;
      return make_column(Utf8, std::move(v), valid, d);
    }
  }
  fail(ML_ERR_INTERNAL, "bad scalar type");
}

// -------------------------------------------------------------- expression tree
ExprPtr build_expr(const ml_expr& e) {
  std::vector<ExprPtr> stack;
  for (int i = 0; i < e.num_nodes; i++) {
    const ml_expr_node& n = e.nodes[i];
    auto x = std::make_shared<Expr>();
    x->kind = n.kind;
    x->op = n.op;
    x->type = n.type;
    x->column = n.column;
    switch (n.kind) {
      case ML_EXPR_COLUMN: break;
      case ML_EXPR_LITERAL:
        x->lit.type = n.type;
        x->lit.i = n.i64;
        x->lit.f = n.f64;
        if (n.type == Utf8) x->lit.s.assign(n.str ? n.str : "", (size_t)n.str_len);
        if (n.type == Boolean) x->lit.i = n.i64 != 0;
        if (n.type == Int32 || n.type == Date) x->lit.i = (int32_t)n.i64;
        break;
      case ML_EXPR_BINARY:
        if (stack.size() < 2) fail(ML_ERR_PLAN, "postfix program underflow");
        x->right = stack.back(); stack.pop_back();
        x->left = stack.back(); stack.pop_back();
        break;
      case ML_EXPR_NOT: case ML_EXPR_NEGATE: case ML_EXPR_CAST:
        if (stack.empty()) fail(ML_ERR_PLAN, "postfix program underflow");
        x->left = stack.back(); stack.pop_back();
        break;
      default: fail(ML_ERR_PLAN, "unknown expression node");
    }
    stack.push_back(x);
  }
  if (stack.size() != 1) fail(ML_ERR_PLAN, "postfix program must leave exactly one value");
  return stack[0];
}

static bool is_cmp(int op) { return op >= ML_OP_EQ && op <= ML_OP_GTEQ; }
static bool is_arith(int op) { return op <= ML_OP_MOD; }
static bool is_logical(int op) { return op == ML_OP_AND || op == ML_OP_OR; }
static int flip(int op) {  // expr/mod.rs:76-84
  switch (op) {
    case ML_OP_LT: return ML_OP_GT;
    case ML_OP_LTEQ: return ML_OP_GTEQ;
    case ML_OP_GT: return ML_OP_LT;
    case ML_OP_GTEQ: return ML_OP_LTEQ;
    default: return op;
  }
}

// numeric_supertype (crates/minilake-core/src/types.rs:54-62)
static int numeric_supertype(int a, int b) {
  if (a == b && is_numeric(a)) return a;
  if ((a == Float64 && is_numeric(b)) || (b == Float64 && is_numeric(a))) return Float64;
  if ((a == Int64 && is_numeric(b)) || (b == Int64 && is_numeric(a))) return Int64;
  return -2;
}
// common_type (eval.rs:218-228)
int common_type(int l, int r) {
  if (l == r) return l;
  if ((l == Date && r == Utf8) || (l == Utf8 && r == Date)) return Date;
  if ((l == Date && (r == Int32 || r == Int64)) || ((l == Int32 || l == Int64) && r == Date)) return Int32;
  return numeric_supertype(l, r);
}

// PhysicalExpr::data_type restated for the node kinds supported here.
int expr_type(const Expr& e, const std::vector<int>& schema) {
  switch (e.kind) {
    case ML_EXPR_COLUMN:
      if (e.column < 0 || (size_t)e.column >= schema.size()) fail(ML_ERR_INTERNAL, "column index out of range");
      return schema[e.column];
    case ML_EXPR_LITERAL: return e.lit.type;
    case ML_EXPR_NOT: return Boolean;
    case ML_EXPR_NEGATE: return expr_type(*e.left, schema);
    case ML_EXPR_CAST: return e.type;
    case ML_EXPR_BINARY: {
      if (is_cmp(e.op) || is_logical(e.op)) return Boolean;
      if (e.op == ML_OP_DIV) return Float64;
      int lt = expr_type(*e.left, schema), rt = expr_type(*e.right, schema);
      if (lt == Null) lt = Int64;
      if (rt == Null) rt = Int64;
      if (e.op == ML_OP_SUB && lt == Date && rt == Date) return Int32;
      if (lt == Date || rt == Date) return Date;
      int ct = common_type(lt, rt);
      if (ct < 0) fail(ML_ERR_EXECUTION, std::string("incompatible types ") + dt_name(lt) + " and " + dt_name(rt));
      return ct;
    }
  }
  fail(ML_ERR_INTERNAL, "bad expression");
}

// ------------------------------------------------------------------ casts
// Rust `f as i64` / `f as i32`: saturating, NaN -> 0.
static int64_t sat_i64(double f) {
  if (std::isnan(f)) return 0;
  if (f >= 9223372036854775807.0) return INT64_MAX;
  if (f <= -9223372036854775808.0) return INT64_MIN;
  return (int64_t)f;
}
static int32_t sat_i32(double f) {
  if (std::isnan(f)) return 0;
  if (f >= 2147483647.0) return INT32_MAX;
  if (f <= -2147483648.0) return INT32_MIN;
  return (int32_t)f;
}

// cast_column (kernels/cast.rs:6-73), numeric/date/bool subset.
ColumnRef cast_column(const ColumnRef& c, int to) {
  if (c->type == to) return c;
  size_t n = c->len;
  auto valid = c->valid;
  int from = c->type;
  auto unsupported = [&]() -> ColumnRef {
    fail(ML_ERR_UNSUPPORTED, std::string("cast from ") + dt_name(from) + " to " + dt_name(to));
  };
  if (from == Int32 || from == Date) {
    const int32_t* v = c->as<int32_t>();
    if (to == Int64) { std::vector<int64_t> o(v, v + n); return make_column(Int64, std::move(o), valid); }
    if (to == Float64) { std::vector<double> o(n); for (size_t i = 0; i < n; i++) o[i] = (double)v[i]; return make_column(Float64, std::move(o), valid); }
    if (to == Date || to == Int32) { std::vector<int32_t> o(v, v + n); return make_column(to, std::move(o), valid); }
    return unsupported();
  }
  if (from == Int64) {
    const int64_t* v = c->as<int64_t>();
    if (to == Int32) { std::vector<int32_t> o(n); for (size_t i = 0; i < n; i++) o[i] = (int32_t)(uint32_t)(uint64_t)v[i]; return make_column(Int32, std::move(o), valid); }
    if (to == Float64) { std::vector<double> o(n); for (size_t i = 0; i < n; i++) o[i] = (double)v[i]; return make_column(Float64, std::move(o), valid); }
    return unsupported();
  }
  if (from == Float64) {
    const double* v = c->as<double>();
    if (to == Int64) { std::vector<int64_t> o(n); for (size_t i = 0; i < n; i++) o[i] = sat_i64(v[i]); return make_column(Int64, std::move(o), valid); }
    if (to == Int32) { std::vector<int32_t> o(n); for (size_t i = 0; i < n; i++) o[i] = sat_i32(v[i]); return make_column(Int32, std::move(o), valid); }
    return unsupported();
  }
  if (from == Boolean) {
    const uint8_t* v = c->as<uint8_t>();
    if (to == Int32) { std::vector<int32_t> o(n); for (size_t i = 0; i < n; i++) o[i] = v[i]; return make_column(Int32, std::move(o), valid); }
    if (to == Int64) { std::vector<int64_t> o(n); for (size_t i = 0; i < n; i++) o[i] = v[i]; return make_column(Int64, std::move(o), valid); }
    if (to == Float64) { std::vector<double> o(n); for (size_t i = 0; i < n; i++) o[i] = v[i]; return make_column(Float64, std::move(o), valid); }
    return unsupported();
  }
  return unsupported();
}

// ScalarValue::cast_to (crates/minilake-core/src/scalar.rs:83-108), numeric subset.
static Scalar scalar_cast(const Scalar& s, int to) {
  if (s.is_null()) return s;
  Scalar o;
  o.type = to;
  auto err = [&]() -> Scalar { fail(ML_ERR_PLAN, std::string("cannot cast constant to ") + dt_name(to)); };
  switch (to) {
    case Boolean: if (s.type != Boolean) return err(); o.i = s.i; return o;
    case Int32: {
      int64_t v;
      if (s.type == Float64) v = sat_i64(s.f);
      else if (!s.as_i64(&v)) return err();
      if (v < INT32_MIN || v > INT32_MAX) return err();
      o.i = v; return o;
    }
    case Int64: {
      if (s.type == Float64) { o.i = sat_i64(s.f); return o; }
      if (!s.as_i64(&o.i)) return err();
      return o;
    }
    case Float64: if (!s.as_f64(&o.f)) return err(); return o;
    case Date: if (s.type != Date) return err(); o.i = s.i; return o;
    case Utf8: if (s.type != Utf8) return err(); o.s = s.s; return o;
  }
  return err();
}

static Datum cast_datum(const Datum& d, int to) {  // eval.rs:199-205
  if (d.is_scalar) return Datum::scalar(scalar_cast(d.s, to));
  if (d.a->type == to) return d;
  return Datum::array(cast_column(d.a, to));
}

// --------------------------------------------------------------- scalar ops
// scalar_binary (kernels/scalar_ops.rs:12-105), runtime subset (no decimals).
static Scalar scalar_binary(int op, const Scalar& l, const Scalar& r) {
  Scalar o;
  if (op == ML_OP_AND || op == ML_OP_OR) {
    bool lb = l.type == Boolean, rb = r.type == Boolean;
    if (op == ML_OP_AND) {
      if ((lb && !l.i) || (rb && !r.i)) { o.type = Boolean; o.i = 0; }
      else if (lb && rb) { o.type = Boolean; o.i = 1; }
      return o;
    }
    if ((lb && l.i) || (rb && r.i)) { o.type = Boolean; o.i = 1; }
    else if (lb && rb) { o.type = Boolean; o.i = 0; }
    return o;
  }
  if (l.is_null() || r.is_null()) return o;
  if (is_cmp(op)) {
    int ord;
    if (l.type == Utf8 && r.type == Utf8) {
      int c = l.s.compare(r.s);
      ord = c < 0 ? -1 : (c > 0 ? 1 : 0);
    } else if (l.type == Boolean && r.type == Boolean) {
      ord = (l.i > r.i) - (l.i < r.i);
    } else if (l.type == Float64 || r.type == Float64) {
      double a, b;
      if (!l.as_f64(&a) || !r.as_f64(&b)) fail(ML_ERR_PLAN, "cannot compare constants");
      if (std::isnan(a) || std::isnan(b)) fail(ML_ERR_PLAN, "cannot compare NaN constants");
      ord = (a > b) - (a < b);
    } else {
      int64_t a, b;
      if (!l.as_i64(&a) || !r.as_i64(&b)) fail(ML_ERR_PLAN, "cannot compare constants");
      ord = (a > b) - (a < b);
    }
    o.type = Boolean;
    switch (op) {
      case ML_OP_EQ: o.i = ord == 0; break;
      case ML_OP_NOTEQ: o.i = ord != 0; break;
      case ML_OP_LT: o.i = ord < 0; break;
      case ML_OP_LTEQ: o.i = ord <= 0; break;
      case ML_OP_GT: o.i = ord > 0; break;
      default: o.i = ord >= 0; break;
    }
    return o;
  }
  if (l.type == Date && (r.type == Int32 || r.type == Int64) && (op == ML_OP_ADD || op == ML_OP_SUB)) {
    int64_t v = op == ML_OP_ADD ? l.i + r.i : l.i - r.i;
    if (v < INT32_MIN || v > INT32_MAX) fail(ML_ERR_EXECUTION, "overflow in date arithmetic");
    o.type = Date; o.i = v; return o;
  }
  if (l.type == Date && r.type == Date && op == ML_OP_SUB) { o.type = Int64; o.i = l.i - r.i; return o; }
  if (op == ML_OP_DIV) {
    double a, b;
    if (!l.as_f64(&a) || !r.as_f64(&b)) fail(ML_ERR_PLAN, "not numeric");
    if (b == 0.0) return o;
    o.type = Float64; o.f = a / b; return o;
  }
  bool li = l.type == Int32 || l.type == Int64, ri = r.type == Int32 || r.type == Int64;
  if (li && ri) {
    int64_t a = l.i, b = r.i, v;
    bool ovf = false;
    switch (op) {
      case ML_OP_ADD: ovf = __builtin_add_overflow(a, b, &v); break;
      case ML_OP_SUB: ovf = __builtin_sub_overflow(a, b, &v); break;
      case ML_OP_MUL: ovf = __builtin_mul_overflow(a, b, &v); break;
      default:
        if (b == 0) return o;
        if (a == INT64_MIN && b == -1) { ovf = true; v = 0; }
        else v = a % b;
    }
    if (ovf) fail(ML_ERR_EXECUTION, "overflow in constant arithmetic");
    if (l.type == Int32 && r.type == Int32 && v >= INT32_MIN && v <= INT32_MAX) o.type = Int32;
    else o.type = Int64;
    o.i = v;
    return o;
  }
  double a, b;
  if (!l.as_f64(&a) || !r.as_f64(&b)) fail(ML_ERR_PLAN, "not numeric");
  o.type = Float64;
  switch (op) {
    case ML_OP_ADD: o.f = a + b; break;
    case ML_OP_SUB: o.f = a - b; break;
    case ML_OP_MUL: o.f = a * b; break;
    default: o.f = std::fmod(a, b); break;
  }
  return o;
}

// ------------------------------------------------------------ binary kernels
// combine_validity (crates/minilake-core/src/bitmap.rs:153-159)
static std::shared_ptr<std::vector<uint8_t>> combine_validity(const Datum& l, const Datum& r, size_t n) {
  auto lv = l.is_scalar ? nullptr : l.a->valid;
  auto rv = r.is_scalar ? nullptr : r.a->valid;
  if (!lv && !rv) return nullptr;
  auto v = std::make_shared<std::vector<uint8_t>>(n, 1);
  for (size_t i = 0; i < n; i++) (*v)[i] = (!lv || (*lv)[i]) && (!rv || (*rv)[i]);
  return v;
}

// Operand (kernels/mod.rs:32-37): slice or broadcast constant.
template <class T>
struct Operand {
  const T* p = nullptr;
  T s{};
  T at(size_t i) const { return p ? p[i] : s; }
};
template <class T> static Operand<T> operand(const Datum& d);
template <> Operand<int32_t> operand<int32_t>(const Datum& d) {  // eval.rs:386-388
  Operand<int32_t> o;
  if (d.is_scalar) { int64_t v; if (!d.s.as_i64(&v)) fail(ML_ERR_EXECUTION, "expected i32 constant"); o.s = (int32_t)v; }
  else { if (d.a->type != Int32 && d.a->type != Date) fail(ML_ERR_EXECUTION, "expected i32 column"); o.p = d.a->as<int32_t>(); }
  return o;
}
template <> Operand<int64_t> operand<int64_t>(const Datum& d) {
  Operand<int64_t> o;
  if (d.is_scalar) { if (!d.s.as_i64(&o.s)) fail(ML_ERR_EXECUTION, "expected i64 constant"); }
  else { if (d.a->type != Int64) fail(ML_ERR_EXECUTION, "expected i64 column"); o.p = d.a->as<int64_t>(); }
  return o;
}
template <> Operand<double> operand<double>(const Datum& d) {
  Operand<double> o;
  if (d.is_scalar) { if (!d.s.as_f64(&o.s)) fail(ML_ERR_EXECUTION, "expected f64 constant"); }
  else { if (d.a->type != Float64) fail(ML_ERR_EXECUTION, "expected f64 column"); o.p = d.a->as<double>(); }
  return o;
}
template <> Operand<uint8_t> operand<uint8_t>(const Datum& d) {
  Operand<uint8_t> o;
  if (d.is_scalar) { if (d.s.type != Boolean) fail(ML_ERR_EXECUTION, "expected bool constant"); o.s = (uint8_t)d.s.i; }
  else { if (d.a->type != Boolean) fail(ML_ERR_EXECUTION, "expected bool column"); o.p = d.a->as<uint8_t>(); }
  return o;
}

template <class T>
static std::vector<uint8_t> compare_vals(int op, Operand<T> l, Operand<T> r, size_t n) {  // cmp.rs:12-31
  std::vector<uint8_t> m(n);
  for (size_t i = 0; i < n; i++) {
    T a = l.at(i), b = r.at(i);
    bool v;
    switch (op) {
      case ML_OP_EQ: v = a == b; break;
      case ML_OP_NOTEQ: v = a != b; break;
      case ML_OP_LT: v = a < b; break;
      case ML_OP_LTEQ: v = a <= b; break;
      case ML_OP_GT: v = a > b; break;
      case ML_OP_GTEQ: v = a >= b; break;
      default: fail(ML_ERR_INTERNAL, "not a comparison");
    }
    m[i] = v;
  }
  return m;
}

static bool compare_bytes(int op, const std::string& a, const std::string& b) {  // cmp.rs:34-45
  int c = a.compare(b);
  switch (op) {
    case ML_OP_EQ: return c == 0;
    case ML_OP_NOTEQ: return c != 0;
    case ML_OP_LT: return c < 0;
    case ML_OP_LTEQ: return c <= 0;
    case ML_OP_GT: return c > 0;
    case ML_OP_GTEQ: return c >= 0;
  }
  return false;
}

// Wrapping integer arithmetic (kernels/arith.rs:28-93).
template <class T, class U>
static T warith(int op, T a, T b) {
  switch (op) {
    case ML_OP_ADD: return (T)((U)a + (U)b);
    case ML_OP_SUB: return (T)((U)a - (U)b);
    case ML_OP_MUL: return (T)((U)a * (U)b);
    default:  // Mod: 0 when b == 0, wrapping_rem otherwise
      if (b == 0) return 0;
      if (b == -1) return 0;
      return a % b;
  }
}

template <class T, class U>
static std::vector<T> arith_int(int op, Operand<T> l, Operand<T> r, size_t n) {  // arith.rs:95-112
  std::vector<T> o(n);
  for (size_t i = 0; i < n; i++) o[i] = warith<T, U>(op, l.at(i), r.at(i));
  return o;
}
static std::vector<double> arith_f64(int op, Operand<double> l, Operand<double> r, size_t n) {
  std::vector<double> o(n);
  for (size_t i = 0; i < n; i++) {
    double a = l.at(i), b = r.at(i);
    switch (op) {
      case ML_OP_ADD: o[i] = a + b; break;
      case ML_OP_SUB: o[i] = a - b; break;
      case ML_OP_MUL: o[i] = a * b; break;
      default: o[i] = std::fmod(a, b); break;
    }
  }
  return o;
}

// mask_zero_divisors (eval.rs:303-324)
template <class T>
static std::shared_ptr<std::vector<uint8_t>> mask_zero_divisors(std::shared_ptr<std::vector<uint8_t>> validity,
                                                                Operand<T> d, size_t n) {
  if (!d.p) {
    if (d.s == 0) return std::make_shared<std::vector<uint8_t>>(n, 0);
    return validity;
  }
  bool any = false;
  for (size_t i = 0; i < n; i++) any |= d.p[i] == 0;
  if (!any) return validity;
  auto bm = validity ? std::make_shared<std::vector<uint8_t>>(*validity) : std::make_shared<std::vector<uint8_t>>(n, 1);
  for (size_t i = 0; i < n; i++)
    if (d.p[i] == 0) (*bm)[i] = 0;
  return bm;
}

static ColumnRef to_bool_col(const Datum& d, size_t n) {
  if (d.is_scalar) return column_from_scalar(d.s, n, Boolean);
  if (d.a->type != Boolean) fail(ML_ERR_EXECUTION, "AND/OR on non-boolean");
  return d.a;
}

// logical (eval.rs:348-359) + Kleene and/or (kernels/boolean.rs:18-58)
static Datum logical(int op, const Datum& l, const Datum& r, size_t n) {
  ColumnRef a = to_bool_col(l, n), b = to_bool_col(r, n);
  const uint8_t* av = a->as<uint8_t>();
  const uint8_t* bv = b->as<uint8_t>();
  std::vector<uint8_t> vals(n);
  for (size_t i = 0; i < n; i++) vals[i] = op == ML_OP_AND ? (av[i] & bv[i]) : (av[i] | bv[i]);
  std::shared_ptr<std::vector<uint8_t>> valid;
  if (a->valid || b->valid) {
    valid = std::make_shared<std::vector<uint8_t>>(n);
    for (size_t i = 0; i < n; i++) {
      bool va = a->is_valid(i), vb = b->is_valid(i);
      (*valid)[i] = op == ML_OP_AND ? ((va && vb) || (va && !av[i]) || (vb && !bv[i]))
                                    : ((va && vb) || (va && av[i]) || (vb && bv[i]));
    }
  }
  return Datum::array(make_column(Boolean, std::move(vals), valid));
}

// binary (eval.rs:230-301)
static Datum binary(int op, Datum l, Datum r, size_t n) {
  if (l.is_scalar && r.is_scalar) return Datum::scalar(scalar_binary(op, l.s, r.s));
  if (is_logical(op)) return logical(op, l, r, n);
  if ((l.is_scalar && l.s.is_null()) || (r.is_scalar && r.s.is_null())) return Datum::scalar(Scalar{});
  int lt = l.data_type(), rt = r.data_type();
  if (lt == Null) lt = Int64;
  if (rt == Null) rt = Int64;
  int ct = common_type(lt, rt);
  if (ct < 0) fail(ML_ERR_EXECUTION, std::string("incompatible types ") + dt_name(lt) + " and " + dt_name(rt));
  bool date_result = is_arith(op) && (lt == Date || rt == Date);
  bool date_diff = op == ML_OP_SUB && lt == Date && rt == Date;
  l = cast_datum(l, ct);
  r = cast_datum(r, ct);
  auto validity = combine_validity(l, r, n);
  if (is_cmp(op)) {
    std::vector<uint8_t> mask;
    switch (ct) {  // comparison (eval.rs:326-346)
      case Int32: case Date: mask = compare_vals(op, operand<int32_t>(l), operand<int32_t>(r), n); break;
      case Int64: mask = compare_vals(op, operand<int64_t>(l), operand<int64_t>(r), n); break;
      case Float64: mask = compare_vals(op, operand<double>(l), operand<double>(r), n); break;
      case Boolean: mask = compare_vals(op, operand<uint8_t>(l), operand<uint8_t>(r), n); break;
      case Utf8: {
        // compare_string_scalar (cmp.rs:52-61): once per dictionary entry.
        const Datum* col = &l;
        const Datum* lit = &r;
        int o = op;
        if (l.is_scalar) { col = &r; lit = &l; o = flip(op); }
        if (col->is_scalar || !lit->is_scalar) fail(ML_ERR_UNSUPPORTED, "string column vs column comparison");
        const Column& c = *col->a;
        std::vector<uint8_t> per_entry(c.dict->entries.size());
        for (size_t k = 0; k < per_entry.size(); k++) per_entry[k] = compare_bytes(o, c.dict->entries[k], lit->s.s);
        mask.resize(n);
        const uint32_t* codes = c.as<uint32_t>();
        for (size_t i = 0; i < n; i++) mask[i] = per_entry[codes[i]];
        break;
      }
      default: fail(ML_ERR_INTERNAL, "bad comparison type");
    }
    return Datum::array(make_column(Boolean, std::move(mask), validity));
  }
  if (op == ML_OP_DIV) {
    l = cast_datum(l, Float64);
    r = cast_datum(r, Float64);
    auto a = operand<double>(l), b = operand<double>(r);
    std::vector<double> v(n);
    for (size_t i = 0; i < n; i++) v[i] = a.at(i) / b.at(i);  // div_f64 (arith.rs:115-117)
    validity = mask_zero_divisors(validity, b, n);
    return Datum::array(make_column(Float64, std::move(v), validity));
  }
  switch (ct) {
    case Int32: case Date: {
      auto a = operand<int32_t>(l), b = operand<int32_t>(r);
      auto v = arith_int<int32_t, uint32_t>(op, a, b, n);
      if (op == ML_OP_MOD) validity = mask_zero_divisors(validity, b, n);
      int t = date_diff ? Int32 : (date_result ? Date : Int32);
      return Datum::array(make_column(t, std::move(v), validity));
    }
    case Int64: {
      auto a = operand<int64_t>(l), b = operand<int64_t>(r);
      auto v = arith_int<int64_t, uint64_t>(op, a, b, n);
      if (op == ML_OP_MOD) validity = mask_zero_divisors(validity, b, n);
      return Datum::array(make_column(Int64, std::move(v), validity));
    }
    case Float64:
      return Datum::array(make_column(Float64, arith_f64(op, operand<double>(l), operand<double>(r), n), validity));
    default:
      fail(ML_ERR_EXECUTION, std::string("arithmetic on ") + dt_name(ct) + " is not supported");
  }
}

// evaluate (eval.rs:61-151)
Datum evaluate(const Expr& e, const Batch& b) {
  size_t n = b.num_rows;
  switch (e.kind) {
    case ML_EXPR_COLUMN:
      if (e.column < 0 || (size_t)e.column >= b.columns.size()) fail(ML_ERR_INTERNAL, "column missing");
      return Datum::array(b.columns[e.column]);
    case ML_EXPR_LITERAL: return Datum::scalar(e.lit);
    case ML_EXPR_BINARY: {
      Datum l = evaluate(*e.left, b);
      Datum r = evaluate(*e.right, b);
      return binary(e.op, l, r, n);
    }
    case ML_EXPR_NOT: {
      Datum d = evaluate(*e.left, b);
      if (d.is_scalar) {
        Scalar s;
        if (d.s.type == Boolean) { s.type = Boolean; s.i = !d.s.i; }
        return Datum::scalar(s);
      }
      if (d.a->type != Boolean) fail(ML_ERR_EXECUTION, "NOT of non-boolean");
      std::vector<uint8_t> v(n);
      const uint8_t* x = d.a->as<uint8_t>();
      for (size_t i = 0; i < n; i++) v[i] = !x[i];
      return Datum::array(make_column(Boolean, std::move(v), d.a->valid));
    }
    case ML_EXPR_NEGATE: {
      Datum d = evaluate(*e.left, b);
      if (d.is_scalar) {
        Scalar s = d.s;
        if (s.type == Int32 || s.type == Int64) s.i = (int64_t)(0 - (uint64_t)s.i);
        else if (s.type == Float64) s.f = -s.f;
        else if (!s.is_null()) fail(ML_ERR_EXECUTION, "cannot negate");
        return Datum::scalar(s);
      }
      const Column& c = *d.a;
      if (c.type == Int32) { std::vector<int32_t> v(n); for (size_t i = 0; i < n; i++) v[i] = (int32_t)(0u - (uint32_t)c.as<int32_t>()[i]); return Datum::array(make_column(Int32, std::move(v), c.valid)); }
      if (c.type == Int64) { std::vector<int64_t> v(n); for (size_t i = 0; i < n; i++) v[i] = (int64_t)(0ull - (uint64_t)c.as<int64_t>()[i]); return Datum::array(make_column(Int64, std::move(v), c.valid)); }
      if (c.type == Float64) { std::vector<double> v(n); for (size_t i = 0; i < n; i++) v[i] = -c.as<double>()[i]; return Datum::array(make_column(Float64, std::move(v), c.valid)); }
      fail(ML_ERR_EXECUTION, "cannot negate");
    }
    case ML_EXPR_CAST: return cast_datum(evaluate(*e.left, b), e.type);
  }
  fail(ML_ERR_INTERNAL, "bad expression");
}

// evaluate_to_column (eval.rs:154-160)
ColumnRef evaluate_to_column(const Expr& e, const Batch& b, int dt) {
  Datum d = evaluate(e, b);
  ColumnRef c = d.is_scalar ? column_from_scalar(d.s.is_null() ? d.s : scalar_cast(d.s, dt), b.num_rows, dt) : d.a;
  if (c->type != dt) return cast_column(c, dt);
  return c;
}

// mask_to_selection (kernels/cmp.rs:85-118)
std::vector<uint32_t> mask_to_selection(const uint8_t* mask, size_t n, const std::vector<uint8_t>* validity,
                                        const std::vector<uint32_t>* input) {
  std::vector<uint32_t> out;
  if (!input) {
    out.assign(n, 0);
    size_t k = 0;
    for (size_t i = 0; i < n; i++) {  // branch-free dense path
      out[k] = (uint32_t)i;
      k += (size_t)(mask[i] & (validity ? (*validity)[i] : 1));
    }
    out.resize(k);
    return out;
  }
  for (uint32_t i : *input)
    if (mask[i] && (!validity || (*validity)[i])) out.push_back(i);
  return out;
}

// select (eval.rs:167-196)
std::vector<uint32_t> select(const Expr& e, const Batch& b, const std::vector<uint32_t>* input) {
  if (e.kind == ML_EXPR_BINARY && e.op == ML_OP_AND) {
    std::vector<uint32_t> sel = select(*e.left, b, input);
    if (sel.empty()) return sel;
    return select(*e.right, b, &sel);
  }
  Datum d = evaluate(e, b);
  if (d.is_scalar) {
    if (d.s.type == Boolean && d.s.i) {
      if (input) return *input;
      std::vector<uint32_t> all(b.num_rows);
      for (size_t i = 0; i < b.num_rows; i++) all[i] = (uint32_t)i;
      return all;
    }
    return {};
  }
  if (d.a->type != Boolean) fail(ML_ERR_EXECUTION, "filter predicate is not boolean");
  return mask_to_selection(d.a->as<uint8_t>(), b.num_rows, d.a->valid.get(), input);
}

// hash_columns / hash_one (kernels/hash.rs:16-79)
void hash_columns(const std::vector<ColumnRef>& cols, Rows rows, std::vector<uint64_t>& out) {
  size_t n = rows.len();
  out.assign(n, SEED);
  for (auto& cp : cols) {
    const Column& c = *cp;
    auto mix = [&](auto to_u64) {
      for (size_t k = 0; k < n; k++) {
        size_t r = rows.row(k);
        uint64_t x = c.is_valid(r) ? hash_u64(to_u64(r)) : NULL_HASH;
        out[k] = combine(out[k], x);
      }
    };
    switch (c.type) {
      case Int32: case Date: { auto v = c.as<int32_t>(); mix([&](size_t r) { return (uint64_t)(int64_t)v[r]; }); break; }
      case Int64: { auto v = c.as<int64_t>(); mix([&](size_t r) { return (uint64_t)v[r]; }); break; }
      case Float64: {
        auto v = c.as<double>();
        mix([&](size_t r) { double x = v[r]; uint64_t bits; std::memcpy(&bits, &x, 8); return x == 0.0 ? 0ull : bits; });
        break;
      }
      case Boolean: { auto v = c.as<uint8_t>(); mix([&](size_t r) { return (uint64_t)(v[r] != 0); }); break; }
      case Utf8: {
        // dictionary: hash once per entry, look up by code (hash.rs:66-77)
        std::vector<uint64_t> dh(c.dict->entries.size());
        for (size_t e = 0; e < dh.size(); e++)
          dh[e] = hash_bytes((const uint8_t*)c.dict->entries[e].data(), c.dict->entries[e].size());
        auto codes = c.as<uint32_t>();
        for (size_t k = 0; k < n; k++) {
          size_t r = rows.row(k);
          uint64_t x = c.is_valid(r) ? dh[codes[r]] : NULL_HASH;
          out[k] = combine(out[k], x);
        }
        break;
      }
      default: fail(ML_ERR_INTERNAL, "bad key column type");
    }
  }
}

// values_equal (kernels/hash.rs:84-101)
bool values_equal(const Column& x, size_t a, const Column& y, size_t b) {
  bool va = x.is_valid(a), vb = y.is_valid(b);
  if (!va && !vb) return true;
  if (va != vb) return false;
  switch (x.type) {
    case Int32: case Date: return x.as<int32_t>()[a] == y.as<int32_t>()[b];
    case Int64: return x.as<int64_t>()[a] == y.as<int64_t>()[b];
    case Float64: return x.as<double>()[a] == y.as<double>()[b];
    case Boolean: return x.as<uint8_t>()[a] == y.as<uint8_t>()[b];
    case Utf8: return x.dict->entries[x.as<uint32_t>()[a]] == y.dict->entries[y.as<uint32_t>()[b]];
  }
  return false;
}

}  // namespace ref
