// Expression compiler and pipeline planner (host code).
//
// Typing and coercions follow MiniLake's batch evaluator so that every row
// value is the one the reference computes:
//   common_type / binary / comparison   crates/minilake-exec/src/expr/eval.rs:218-346
//   wrapping integer arithmetic         crates/minilake-exec/src/kernels/arith.rs:28-93
//   x / 0 and x % 0 are NULL            eval.rs:260-270, 303-324
//   casts                               crates/minilake-exec/src/kernels/cast.rs:6-73
//   constant folding                    crates/minilake-exec/src/kernels/scalar_ops.rs:12-105
//   dictionary predicates (per entry)   crates/minilake-exec/src/kernels/cmp.rs:52-61
#include <cmath>

#include "engine.h"

namespace mlg {

// ----------------------------------------------------------------- parsing
NodePtr parse_expr(const ml_expr& e) {
  std::vector<NodePtr> stack;
  if (!e.nodes || e.num_nodes <= 0) fail(ML_ERR_PLAN, "empty expression");
  for (int i = 0; i < e.num_nodes; i++) {
    const ml_expr_node& m = e.nodes[i];
    auto n = std::make_shared<Node>();
    n->kind = m.kind;
    n->op = m.op;
    n->type = m.type;
    n->column = m.column;
    switch (m.kind) {
      case ML_EXPR_COLUMN: break;
      case ML_EXPR_LITERAL:
        n->lit.type = m.type;
        if (m.type == ML_TYPE_FLOAT64) n->lit.f = m.f64;
        else if (m.type == ML_TYPE_UTF8) n->lit.s.assign(m.str ? m.str : "", (size_t)m.str_len);
        else if (m.type == ML_TYPE_BOOLEAN) n->lit.i = m.i64 != 0;
        else if (m.type == ML_TYPE_INT32 || m.type == ML_TYPE_DATE) n->lit.i = (int32_t)m.i64;
        else if (m.type == ML_TYPE_INT64) n->lit.i = m.i64;
        else if (m.type != ML_TYPE_NULL) fail(ML_ERR_PLAN, "bad literal type");
        break;
      case ML_EXPR_BINARY:
        if (stack.size() < 2) fail(ML_ERR_PLAN, "postfix program underflow");
        if (m.op < ML_OP_ADD || m.op > ML_OP_OR) fail(ML_ERR_PLAN, "bad binary operator");
        n->right = stack.back(); stack.pop_back();
        n->left = stack.back(); stack.pop_back();
        break;
      case ML_EXPR_NOT: case ML_EXPR_NEGATE: case ML_EXPR_CAST:
        if (stack.empty()) fail(ML_ERR_PLAN, "postfix program underflow");
        n->left = stack.back(); stack.pop_back();
        break;
      default: fail(ML_ERR_PLAN, "unknown expression node kind");
    }
    stack.push_back(n);
  }
  if (stack.size() != 1) fail(ML_ERR_PLAN, "postfix program must leave exactly one value");
  return stack[0];
}

// --------------------------------------------------------------- type rules
static bool is_numeric(int t) { return t == ML_TYPE_INT32 || t == ML_TYPE_INT64 || t == ML_TYPE_FLOAT64; }
static bool is_cmp(int op) { return op >= ML_OP_EQ && op <= ML_OP_GTEQ; }
static bool is_logical(int op) { return op == ML_OP_AND || op == ML_OP_OR; }
static bool is_arith(int op) { return op <= ML_OP_MOD; }
static int flip(int op) {
  switch (op) {
    case ML_OP_LT: return ML_OP_GT;
    case ML_OP_LTEQ: return ML_OP_GTEQ;
    case ML_OP_GT: return ML_OP_LT;
    case ML_OP_GTEQ: return ML_OP_LTEQ;
    default: return op;
  }
}
static int numeric_supertype(int a, int b) {  // crates/minilake-core/src/types.rs:54-62
  if (a == b && is_numeric(a)) return a;
  if ((a == ML_TYPE_FLOAT64 && is_numeric(b)) || (b == ML_TYPE_FLOAT64 && is_numeric(a))) return ML_TYPE_FLOAT64;
  if ((a == ML_TYPE_INT64 && is_numeric(b)) || (b == ML_TYPE_INT64 && is_numeric(a))) return ML_TYPE_INT64;
  return -2;
}
static int common_type(int l, int r) {  // eval.rs:218-228
  if (l == r) return l;
  if ((l == ML_TYPE_DATE && r == ML_TYPE_UTF8) || (l == ML_TYPE_UTF8 && r == ML_TYPE_DATE)) return ML_TYPE_DATE;
  if ((l == ML_TYPE_DATE && (r == ML_TYPE_INT32 || r == ML_TYPE_INT64)) ||
      ((l == ML_TYPE_INT32 || l == ML_TYPE_INT64) && r == ML_TYPE_DATE))
    return ML_TYPE_INT32;
  return numeric_supertype(l, r);
}

static bool is_const(const NodePtr& n) {
  if (n->kind == ML_EXPR_LITERAL) return true;
  if (n->kind == ML_EXPR_COLUMN) return false;
  if (n->left && !is_const(n->left)) return false;
  if (n->right && !is_const(n->right)) return false;
  return true;
}

static int64_t sat_i64(double f) {
  if (std::isnan(f)) return 0;
  if (f >= 9223372036854775808.0) return INT64_MAX;
  if (f <= -9223372036854775808.0) return INT64_MIN;
  return (int64_t)f;
}

static bool as_f64(const Scalar& s, double* out) {
  if (s.type == ML_TYPE_INT32 || s.type == ML_TYPE_INT64 || s.type == ML_TYPE_DATE) { *out = (double)s.i; return true; }
  if (s.type == ML_TYPE_FLOAT64) { *out = s.f; return true; }
  return false;
}
static bool as_i64(const Scalar& s, int64_t* out) {
  if (s.type == ML_TYPE_INT32 || s.type == ML_TYPE_INT64 || s.type == ML_TYPE_DATE) { *out = s.i; return true; }
  return false;
}

// ScalarValue::cast_to (crates/minilake-core/src/scalar.rs:83-108)
static Scalar scalar_cast(const Scalar& s, int to) {
  if (s.is_null() || s.type == to) return s;
  Scalar o;
  o.type = to;
  auto err = [&]() -> Scalar { fail(ML_ERR_PLAN, std::string("cannot cast constant to ") + type_name(to)); };
  switch (to) {
    case ML_TYPE_BOOLEAN: return err();
    case ML_TYPE_INT32: {
      int64_t v;
      if (s.type == ML_TYPE_FLOAT64) v = sat_i64(s.f);
      else if (!as_i64(s, &v)) return err();
      if (v < INT32_MIN || v > INT32_MAX) return err();
      o.i = v;
      return o;
    }
    case ML_TYPE_INT64:
      if (s.type == ML_TYPE_FLOAT64) { o.i = sat_i64(s.f); return o; }
      if (!as_i64(s, &o.i)) return err();
      return o;
    case ML_TYPE_FLOAT64:
      if (!as_f64(s, &o.f)) return err();
      return o;
    case ML_TYPE_DATE: return err();  // only Date -> Date (or strings, unsupported here)
    default: return err();
  }
}

// scalar_binary (scalar_ops.rs:12-105), runtime subset (no decimals).
static Scalar scalar_binary(int op, const Scalar& l, const Scalar& r) {
  Scalar o;
  if (is_logical(op)) {
    const bool lb = l.type == ML_TYPE_BOOLEAN, rb = r.type == ML_TYPE_BOOLEAN;
    if (op == ML_OP_AND) {
      if ((lb && !l.i) || (rb && !r.i)) { o.type = ML_TYPE_BOOLEAN; o.i = 0; }
      else if (lb && rb) { o.type = ML_TYPE_BOOLEAN; o.i = 1; }
    } else {
      if ((lb && l.i) || (rb && r.i)) { o.type = ML_TYPE_BOOLEAN; o.i = 1; }
      else if (lb && rb) { o.type = ML_TYPE_BOOLEAN; o.i = 0; }
    }
    return o;
  }
  if (l.is_null() || r.is_null()) return o;
  if (is_cmp(op)) {
    int ord;
    if (l.type == ML_TYPE_UTF8 && r.type == ML_TYPE_UTF8) {
      const int c = l.s.compare(r.s);
      ord = (c > 0) - (c < 0);
    } else if (l.type == ML_TYPE_BOOLEAN && r.type == ML_TYPE_BOOLEAN) {
      ord = (l.i > r.i) - (l.i < r.i);
    } else if (l.type == ML_TYPE_FLOAT64 || r.type == ML_TYPE_FLOAT64) {
      double a, b;
      if (!as_f64(l, &a) || !as_f64(r, &b) || std::isnan(a) || std::isnan(b)) fail(ML_ERR_PLAN, "cannot compare constants");
      ord = (a > b) - (a < b);
    } else {
      int64_t a, b;
      if (!as_i64(l, &a) || !as_i64(r, &b)) fail(ML_ERR_PLAN, "cannot compare constants");
      ord = (a > b) - (a < b);
    }
    o.type = ML_TYPE_BOOLEAN;
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
  if (l.type == ML_TYPE_DATE && (r.type == ML_TYPE_INT32 || r.type == ML_TYPE_INT64) && (op == ML_OP_ADD || op == ML_OP_SUB)) {
    const int64_t v = op == ML_OP_ADD ? l.i + r.i : l.i - r.i;
    if (v < INT32_MIN || v > INT32_MAX) fail(ML_ERR_EXECUTION, "overflow in date arithmetic");
    o.type = ML_TYPE_DATE;
    o.i = v;
    return o;
  }
  if (l.type == ML_TYPE_DATE && r.type == ML_TYPE_DATE && op == ML_OP_SUB) {
    o.type = ML_TYPE_INT64;
    o.i = l.i - r.i;
    return o;
  }
  if (op == ML_OP_DIV) {
    double a, b;
    if (!as_f64(l, &a) || !as_f64(r, &b)) fail(ML_ERR_PLAN, "not numeric");
    if (b == 0.0) return o;
    o.type = ML_TYPE_FLOAT64;
    o.f = a / b;
    return o;
  }
  const bool li = l.type == ML_TYPE_INT32 || l.type == ML_TYPE_INT64;
  const bool ri = r.type == ML_TYPE_INT32 || r.type == ML_TYPE_INT64;
  if (li && ri) {
    int64_t v = 0;
    bool ovf = false;
    switch (op) {
      case ML_OP_ADD: ovf = __builtin_add_overflow(l.i, r.i, &v); break;
      case ML_OP_SUB: ovf = __builtin_sub_overflow(l.i, r.i, &v); break;
      case ML_OP_MUL: ovf = __builtin_mul_overflow(l.i, r.i, &v); break;
      default:
        if (r.i == 0) return o;
        if (l.i == INT64_MIN && r.i == -1) ovf = true;
        else v = l.i % r.i;
    }
    if (ovf) fail(ML_ERR_EXECUTION, "overflow in constant arithmetic");
    o.type = (l.type == ML_TYPE_INT32 && r.type == ML_TYPE_INT32 && v >= INT32_MIN && v <= INT32_MAX) ? ML_TYPE_INT32
                                                                                                        : ML_TYPE_INT64;
    o.i = v;
    return o;
  }
  double a, b;
  if (!as_f64(l, &a) || !as_f64(r, &b)) fail(ML_ERR_PLAN, "not numeric");
  o.type = ML_TYPE_FLOAT64;
  switch (op) {
    case ML_OP_ADD: o.f = a + b; break;
    case ML_OP_SUB: o.f = a - b; break;
    case ML_OP_MUL: o.f = a * b; break;
    default: o.f = std::fmod(a, b); break;
  }
  return o;
}

static Scalar fold(const NodePtr& n) {
  switch (n->kind) {
    case ML_EXPR_LITERAL: return n->lit;
    case ML_EXPR_BINARY: return scalar_binary(n->op, fold(n->left), fold(n->right));
    case ML_EXPR_NOT: {
      Scalar s = fold(n->left), o;
      if (s.type == ML_TYPE_BOOLEAN) { o.type = ML_TYPE_BOOLEAN; o.i = !s.i; }
      return o;
    }
    case ML_EXPR_NEGATE: {
      Scalar s = fold(n->left);
      if (s.type == ML_TYPE_INT32 || s.type == ML_TYPE_INT64) s.i = (int64_t)(0ull - (uint64_t)s.i);
      else if (s.type == ML_TYPE_FLOAT64) s.f = -s.f;
      else if (!s.is_null()) fail(ML_ERR_EXECUTION, "cannot negate");
      return s;
    }
    case ML_EXPR_CAST: return scalar_cast(fold(n->left), n->type);
    default: fail(ML_ERR_INTERNAL, "fold of a column");
  }
}

// ------------------------------------------------------------------ compiler
int Compiler::type_of(const NodePtr& n) {
  switch (n->kind) {
    case ML_EXPR_COLUMN: return n->col->type;
    case ML_EXPR_LITERAL: return n->lit.type;
    case ML_EXPR_CAST: return n->type;
    case ML_EXPR_NOT: return ML_TYPE_BOOLEAN;
    case ML_EXPR_NEGATE: return type_of(n->left);
    case ML_EXPR_BINARY: {
      if (is_const(n)) return fold(n).type;
      if (is_cmp(n->op) || is_logical(n->op)) return ML_TYPE_BOOLEAN;
      if (n->op == ML_OP_DIV) return ML_TYPE_FLOAT64;
      int lt = type_of(n->left), rt = type_of(n->right);
      if (lt == ML_TYPE_NULL) lt = ML_TYPE_INT64;
      if (rt == ML_TYPE_NULL) rt = ML_TYPE_INT64;
      if (n->op == ML_OP_SUB && lt == ML_TYPE_DATE && rt == ML_TYPE_DATE) return ML_TYPE_INT32;
      const int ct = common_type(lt, rt);
      if (ct < 0) fail(ML_ERR_EXECUTION, std::string("incompatible types ") + type_name(lt) + " and " + type_name(rt));
      if (lt == ML_TYPE_DATE || rt == ML_TYPE_DATE) return (ct == ML_TYPE_INT32 || ct == ML_TYPE_DATE) ? ML_TYPE_DATE : ct;
      return ct;
    }
  }
  fail(ML_ERR_INTERNAL, "bad expression");
}

void Compiler::push(uint16_t op, uint64_t imm, uint64_t imm2, uint8_t stage, uint32_t flags) {
  Instr in{};
  in.op = op;
  in.depth = (uint8_t)depth_;
  in.stage = stage;
  in.flags = flags;
  in.imm = imm;
  in.imm2 = imm2;
  ins_.push_back(in);
  depth_++;
  if (depth_ > max_depth_) max_depth_ = depth_;
}
void Compiler::unary(uint16_t op, uint64_t imm) {
  Instr in{};
  in.op = op;
  in.depth = (uint8_t)depth_;
  in.imm = imm;
  ins_.push_back(in);
}
void Compiler::binary(uint16_t op) {
  Instr in{};
  in.op = op;
  in.depth = (uint8_t)depth_;
  ins_.push_back(in);
  depth_--;
}

Compiler::Val Compiler::materialize(const Val& v) {
  if (!v.is_const) return v;
  const Scalar& s = v.s;
  uint64_t bits = 0;
  uint32_t valid = 1;
  int t = s.type;
  switch (s.type) {
    case ML_TYPE_NULL: valid = 0; nullable_ = true; t = ML_TYPE_INT64; break;
    case ML_TYPE_FLOAT64: std::memcpy(&bits, &s.f, 8); break;
    case ML_TYPE_BOOLEAN: bits = s.i ? 1 : 0; break;
    case ML_TYPE_INT32: case ML_TYPE_DATE: bits = (uint64_t)(int64_t)(int32_t)s.i; break;
    case ML_TYPE_INT64: bits = (uint64_t)s.i; break;
    default: fail(ML_ERR_UNSUPPORTED, "string constants are only supported in dictionary comparisons");
  }
  push(OP_CONST, bits, 0, 0, valid);
  return Val{false, {}, t};
}

// cast_column semantics on a runtime value (kernels/cast.rs:6-73)
void Compiler::emit_cast(int from, int to) {
  if (from == to) return;
  auto unsupported = [&]() { fail(ML_ERR_UNSUPPORTED, std::string("cast from ") + type_name(from) + " to " + type_name(to)); };
  switch (from) {
    case ML_TYPE_INT32: case ML_TYPE_DATE:
      if (to == ML_TYPE_INT64 || to == ML_TYPE_INT32 || to == ML_TYPE_DATE) return;
      if (to == ML_TYPE_FLOAT64) { unary(OP_I2F); return; }
      unsupported();
    case ML_TYPE_INT64:
      if (to == ML_TYPE_INT32) { unary(OP_I64_TO_I32); return; }
      if (to == ML_TYPE_FLOAT64) { unary(OP_I2F); return; }
      unsupported();
    case ML_TYPE_FLOAT64:
      if (to == ML_TYPE_INT64) { unary(OP_F2I64); return; }
      if (to == ML_TYPE_INT32) { unary(OP_F2I32); return; }
      unsupported();
    case ML_TYPE_BOOLEAN:
      if (to == ML_TYPE_INT32 || to == ML_TYPE_INT64) return;
      if (to == ML_TYPE_FLOAT64) { unary(OP_I2F); return; }
      unsupported();
    default: unsupported();
  }
}

Compiler::Val Compiler::cast_val(const Val& v, int to) {
  if (v.is_const) return Val{true, scalar_cast(v.s, to), to};
  emit_cast(v.type, to);
  return Val{false, {}, to};
}

static bool compare_bytes(int op, const std::string& a, const std::string& b) {  // cmp.rs:34-45
  const int c = a.compare(b);
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

const uint8_t* Compiler::dict_predicate(const HostDict& d, int op, const std::string& lit) {
  std::vector<uint8_t> per_entry(d.entries.size() + 1, 0);
  for (size_t k = 0; k < d.entries.size(); k++) per_entry[k] = compare_bytes(op, d.entries[k], lit);
  if (bump) {
    void* at = bump->take(per_entry.size());
    ring_upload(ctx_, at, per_entry.data(), per_entry.size());
    return (const uint8_t*)at;
  }
  consts.push_back(ctx_->const_upload(per_entry.data(), per_entry.size()));
  return consts.back()->as<uint8_t>();
}

Compiler::Val Compiler::gen(const NodePtr& n) {
  if (is_const(n)) return Val{true, fold(n), fold(n).type};
  switch (n->kind) {
    case ML_EXPR_COLUMN: {
      const VCol& c = *n->col;
      uint16_t op;
      switch (c.type) {
        case ML_TYPE_BOOLEAN: op = OP_LOAD_U8; break;
        case ML_TYPE_INT32: case ML_TYPE_DATE: op = OP_LOAD_I32; break;
        case ML_TYPE_UTF8: op = OP_LOAD_U32; break;
        default: op = OP_LOAD_I64; break;
      }
      if (c.valid) nullable_ = true;
      push(op, (uint64_t)(uintptr_t)c.data, (uint64_t)(uintptr_t)c.valid, (uint8_t)c.stage);
      return Val{false, {}, c.type};
    }
    case ML_EXPR_BINARY: return gen_binary(*n);
    case ML_EXPR_NOT: {
      Val v = gen(n->left);
      if (v.type != ML_TYPE_BOOLEAN) fail(ML_ERR_EXECUTION, "NOT of non-boolean");
      unary(OP_NOT);
      return Val{false, {}, ML_TYPE_BOOLEAN};
    }
    case ML_EXPR_NEGATE: {
      Val v = gen(n->left);
      if (v.type == ML_TYPE_INT32) unary(OP_NEG_I32);
      else if (v.type == ML_TYPE_INT64) unary(OP_NEG_I64);
      else if (v.type == ML_TYPE_FLOAT64) unary(OP_NEG_F64);
      else fail(ML_ERR_EXECUTION, "cannot negate");
      return Val{false, {}, v.type};
    }
    case ML_EXPR_CAST: {
      Val v = gen(n->left);
      return cast_val(v, n->type);
    }
  }
  fail(ML_ERR_INTERNAL, "bad expression node");
}

static const VCol* plain_column(const NodePtr& n) { return n->kind == ML_EXPR_COLUMN ? n->col : nullptr; }

// binary (eval.rs:230-301)
Compiler::Val Compiler::gen_binary(const Node& n) {
  const NodePtr& L = n.left;
  const NodePtr& R = n.right;
  const bool lc = is_const(L), rc = is_const(R);
  const int op = n.op;
  if (is_logical(op)) {  // logical() with Kleene semantics (eval.rs:348-359)
    for (const NodePtr* side : {&L, &R}) {
      if (is_const(*side)) {
        Scalar s = fold(*side);
        if (!s.is_null() && s.type != ML_TYPE_BOOLEAN) fail(ML_ERR_EXECUTION, "AND/OR on non-boolean");
        if (s.is_null()) {
          push(OP_CONST, 0, 0, 0, 0);
          nullable_ = true;
        } else {
          push(OP_CONST, s.i ? 1 : 0, 0, 0, 1);
        }
      } else {
        Val v = gen(*side);
        if (v.type != ML_TYPE_BOOLEAN) fail(ML_ERR_EXECUTION, "AND/OR on non-boolean");
      }
    }
    binary(op == ML_OP_AND ? OP_AND : OP_OR);
    return Val{false, {}, ML_TYPE_BOOLEAN};
  }
  if ((lc && fold(L).is_null()) || (rc && fold(R).is_null())) {
    Scalar nul;
    return Val{true, nul, ML_TYPE_NULL};
  }
  int lt = lc ? fold(L).type : type_of(L);
  int rt = rc ? fold(R).type : type_of(R);
  const int ct = common_type(lt, rt);
  if (ct < 0) fail(ML_ERR_EXECUTION, std::string("incompatible types ") + type_name(lt) + " and " + type_name(rt));
  const bool date_result = is_arith(op) && (lt == ML_TYPE_DATE || rt == ML_TYPE_DATE);
  const bool date_diff = op == ML_OP_SUB && lt == ML_TYPE_DATE && rt == ML_TYPE_DATE;

  if (is_cmp(op) && ct == ML_TYPE_UTF8) {
    // string column vs string constant: evaluated once per dictionary entry
    const NodePtr& colside = lc ? R : L;
    const NodePtr& litside = lc ? L : R;
    const VCol* c = plain_column(colside);
    if (!c || !c->dict || !is_const(litside)) fail(ML_ERR_UNSUPPORTED, "string comparison needs a dictionary column and a constant");
    const Scalar s = fold(litside);
    if (s.type != ML_TYPE_UTF8) fail(ML_ERR_EXECUTION, "string comparison with a non-string constant");
    const uint8_t* table = dict_predicate(*c->dict, lc ? flip(op) : op, s.s);
    gen(colside);
    unary(OP_LOOKUP_U8, (uint64_t)(uintptr_t)table);
    return Val{false, {}, ML_TYPE_BOOLEAN};
  }

  // operand types after the Div promotion
  const int opnd = op == ML_OP_DIV ? ML_TYPE_FLOAT64 : ct;
  auto side = [&](const NodePtr& x, bool c, int t) {
    if (c) {
      Scalar s = scalar_cast(fold(x), ct);
      if (op == ML_OP_DIV) s = scalar_cast(s, ML_TYPE_FLOAT64);
      materialize(Val{true, s, s.type});
    } else {
      gen(x);
      emit_cast(t, ct);
      if (op == ML_OP_DIV) emit_cast(ct, ML_TYPE_FLOAT64);
    }
  };
  side(L, lc, lt);
  side(R, rc, rt);
  if (is_cmp(op)) {
    const bool f = ct == ML_TYPE_FLOAT64;
    if (ct != ML_TYPE_FLOAT64 && ct != ML_TYPE_INT32 && ct != ML_TYPE_INT64 && ct != ML_TYPE_DATE && ct != ML_TYPE_BOOLEAN)
      fail(ML_ERR_INTERNAL, "bad comparison type");
    static const uint16_t icmp[] = {OP_EQ_I, OP_NE_I, OP_LT_I, OP_LE_I, OP_GT_I, OP_GE_I};
    static const uint16_t fcmp[] = {OP_EQ_F, OP_NE_F, OP_LT_F, OP_LE_F, OP_GT_F, OP_GE_F};
    binary(f ? fcmp[op - ML_OP_EQ] : icmp[op - ML_OP_EQ]);
    return Val{false, {}, ML_TYPE_BOOLEAN};
  }
  if (op == ML_OP_DIV) {
    binary(OP_DIV_F64);
    nullable_ = true;
    return Val{false, {}, ML_TYPE_FLOAT64};
  }
  (void)opnd;
  switch (ct) {
    case ML_TYPE_INT32: case ML_TYPE_DATE: {
      static const uint16_t o32[] = {OP_ADD_I32, OP_SUB_I32, OP_MUL_I32, 0, OP_MOD_I32};
      binary(o32[op]);
      if (op == ML_OP_MOD) nullable_ = true;
      return Val{false, {}, date_diff ? ML_TYPE_INT32 : (date_result ? ML_TYPE_DATE : ML_TYPE_INT32)};
    }
    case ML_TYPE_INT64: {
      static const uint16_t o64[] = {OP_ADD_I64, OP_SUB_I64, OP_MUL_I64, 0, OP_MOD_I64};
      binary(o64[op]);
      if (op == ML_OP_MOD) nullable_ = true;
      return Val{false, {}, ML_TYPE_INT64};
    }
    case ML_TYPE_FLOAT64: {
      static const uint16_t of[] = {OP_ADD_F64, OP_SUB_F64, OP_MUL_F64, 0, OP_MOD_F64};
      binary(of[op]);
      return Val{false, {}, ML_TYPE_FLOAT64};
    }
    default: fail(ML_ERR_EXECUTION, std::string("arithmetic on ") + type_name(ct) + " is not supported");
  }
}

// Peephole over the program starting at `first`: LOAD (no validity) CONST
// (valid) CMP becomes one OP_CMPC, and LOAD_U32 LOOKUP_U8 one OP_LOOKUPC.
// Each interpreted instruction costs a dispatch, so filters over a column and
// a constant run in one.
void Compiler::fuse(int32_t first) {
  std::vector<Instr> out;
  const size_t n = ins_.size();
  for (size_t i = (size_t)first; i < n; i++) {
    const Instr& a = ins_[i];
    if (a.op < OP_CONST && a.imm2 == 0 && i + 2 < n) {
      const Instr& b = ins_[i + 1];
      const Instr& c = ins_[i + 2];
      const bool icmp = c.op >= OP_EQ_I && c.op <= OP_GE_I, fcmp = c.op >= OP_EQ_F && c.op <= OP_GE_F;
      if (b.op == OP_CONST && (b.flags & 1u) && b.depth == a.depth + 1 && c.depth == a.depth + 2 && (icmp || fcmp) &&
          !(fcmp && a.op != OP_LOAD_I64)) {
        Instr f = a;
        f.op = OP_CMPC;
        f.imm2 = b.imm;
        f.flags = ((uint32_t)(icmp ? c.op - OP_EQ_I : c.op - OP_EQ_F) << 8) | (fcmp ? 0x800u : 0u) |
                  ((uint32_t)a.op << 13);
        out.push_back(f);
        i += 2;
        continue;
      }
    }
    if (a.op == OP_LOAD_U32 && a.imm2 == 0 && i + 1 < n) {
      const Instr& b = ins_[i + 1];
      if (b.op == OP_LOOKUP_U8 && b.depth == a.depth + 1) {
        Instr f = a;
        f.op = OP_LOOKUPC;
        f.imm2 = b.imm;
        out.push_back(f);
        i += 1;
        continue;
      }
    }
    out.push_back(a);
  }
  ins_.resize((size_t)first);
  ins_.insert(ins_.end(), out.begin(), out.end());
}

Prog Compiler::compile(const NodePtr& n, int want, int* out_type) {
  depth_ = 0;
  max_depth_ = 0;
  nullable_ = false;
  Prog p;
  p.first = (int32_t)ins_.size();
  Val v = gen(n);
  if (v.is_const) {
    Scalar s = v.s;
    if (want >= 0 && !s.is_null()) s = scalar_cast(s, want);
    v = materialize(Val{true, s, s.type});
    if (want >= 0) v.type = want;
  } else if (want >= 0) {
    emit_cast(v.type, want);
    v.type = want;
  }
  if (max_depth_ > MAX_DEPTH) fail(ML_ERR_UNSUPPORTED, "expression needs more than 6 stack slots");
  if (depth_ != 1) fail(ML_ERR_INTERNAL, "expression compiled to an unbalanced program");
  fuse(p.first);
  p.count = (int16_t)((int32_t)ins_.size() - p.first);
  p.nullable = nullable_ ? 1 : 0;
  if (out_type) *out_type = v.type;
  return p;
}

Prog Compiler::compile_key(const NodePtr& n, int key_type, uint64_t* nan_ids) {
  Prog p = compile(n, key_type);
  if (key_type == ML_TYPE_FLOAT64) {
    depth_ = 1;
    unary(OP_KEY_F64, (uint64_t)(uintptr_t)nan_ids);
    p.count++;
  }
  return p;
}

Prog Compiler::compile_join_key(const NodePtr& n, int key_type) {
  Prog p = compile(n, key_type);
  if (key_type == ML_TYPE_FLOAT64) {  // values_equal compares f64 with == (kernels/hash.rs:94)
    depth_ = 1;
    unary(OP_JKEY_F64);
    p.count++;
    p.nullable = 1;
  }
  return p;
}

Prog Compiler::compile_hash_input(const NodePtr& n, bool* prehashed) {
  const int t = type_of(n);
  *prehashed = false;
  if (t == ML_TYPE_UTF8) {
    const VCol* c = plain_column(n);
    if (!c || !c->dict) fail(ML_ERR_UNSUPPORTED, "string keys must be dictionary columns");
    std::vector<uint64_t> hs(c->dict->entries.size() + 1, 0);
    for (size_t k = 0; k < c->dict->entries.size(); k++) {  // hash_bytes (hashtable/src/hash.rs:39-55)
      const std::string& e = c->dict->entries[k];
      uint64_t h = HSEED ^ ((uint64_t)e.size() * HMUL);
      size_t full = e.size() / 8;
      for (size_t w = 0; w < full; w++) {
        uint64_t x;
        std::memcpy(&x, e.data() + 8 * w, 8);
        h = fold_mul(h ^ x, HMUL);
      }
      size_t tail = e.size() - 8 * full;
      if (tail) {
        uint64_t x = 0;
        std::memcpy(&x, e.data() + 8 * full, tail);
        h = fold_mul(h ^ x, HMUL2);
      }
      hs[k] = h;
    }
    const uint64_t* tp;
    if (bump) {
      void* at = bump->take(hs.size() * 8);
      ring_upload(ctx_, at, hs.data(), hs.size() * 8);
      tp = (const uint64_t*)at;
    } else {
      consts.push_back(ctx_->const_upload(hs.data(), hs.size() * 8));
      tp = consts.back()->as<uint64_t>();
    }
    Prog p = compile(n, -1);
    depth_ = 1;
    unary(OP_LOOKUP_U64, (uint64_t)(uintptr_t)tp);
    p.count++;
    *prehashed = true;
    return p;
  }
  Prog p = compile(n, t);
  if (t == ML_TYPE_FLOAT64) {
    depth_ = 1;
    unary(OP_HASH_F64);
    p.count++;
  }
  return p;
}

Prog Compiler::compile_load(const VCol& c) {
  auto n = std::make_shared<Node>();
  n->kind = ML_EXPR_COLUMN;
  n->col = &c;
  return compile(n, -1);
}

// ------------------------------------------------------------ pipeline build
NodePtr resolve_layout(const NodePtr& n, const std::vector<const VCol*>& layout) {
  if (n->kind == ML_EXPR_COLUMN) {
    if (n->column < 0 || (size_t)n->column >= layout.size())
      fail(ML_ERR_INTERNAL, "column #" + std::to_string(n->column) + " missing");
    const VCol* c = layout[(size_t)n->column];
    if (c->expr) return c->expr;
    auto m = std::make_shared<Node>(*n);
    m->col = c;
    return m;
  }
  if (n->kind == ML_EXPR_LITERAL) return n;
  auto m = std::make_shared<Node>(*n);
  if (n->left) m->left = resolve_layout(n->left, layout);
  if (n->right) m->right = resolve_layout(n->right, layout);
  return m;
}

}  // namespace mlg
