// CPU ORACLE — test infrastructure only (see ref_core.h).
#pragma once

#include "ref_core.h"

namespace ref {

// PhysicalExpr (crates/minilake-exec/src/expr/mod.rs:109-173), tree form.
struct Expr;
using ExprPtr = std::shared_ptr<const Expr>;
struct Expr {
  int kind = ML_EXPR_COLUMN;
  int op = 0;
  int type = 0;
  int column = 0;
  Scalar lit;
  ExprPtr left, right;
};
ExprPtr build_expr(const ml_expr& e);

// Datum (eval.rs:25-31): a column or a constant.
struct Datum {
  bool is_scalar = false;
  ColumnRef a;
  Scalar s;
  static Datum array(ColumnRef c) { Datum d; d.a = std::move(c); return d; }
  static Datum scalar(Scalar v) { Datum d; d.is_scalar = true; d.s = std::move(v); return d; }
  int data_type() const { return is_scalar ? s.type : a->type; }
};

int common_type(int l, int r);
int expr_type(const Expr& e, const std::vector<int>& schema);
ColumnRef cast_column(const ColumnRef& c, int to);
Datum evaluate(const Expr& e, const Batch& b);
ColumnRef evaluate_to_column(const Expr& e, const Batch& b, int dt);
std::vector<uint32_t> mask_to_selection(const uint8_t* mask, size_t n, const std::vector<uint8_t>* validity,
                                        const std::vector<uint32_t>* input);
std::vector<uint32_t> select(const Expr& e, const Batch& b, const std::vector<uint32_t>* input);
void hash_columns(const std::vector<ColumnRef>& cols, Rows rows, std::vector<uint64_t>& out);
bool values_equal(const Column& x, size_t a, const Column& y, size_t b);

}  // namespace ref
