// tprog.h -- instruction set of the boundary-time programs: postfix code that
// the host compiles from a boundary expression of t (expr.cpp) and the device
// evaluates at each stage time (k_bc_time).  An instruction is op | arg << 8.
#pragma once

namespace tsm {

enum TpOp : int {
  TP_END = 0,
  TP_SLOT,  // push the entry's slot value arg (a subexpression without t)
  TP_T,     // push the stage time
  TP_NEG, TP_NOT,
  TP_ADD, TP_SUB, TP_MUL, TP_DIV, TP_MOD, TP_POW,
  TP_LT, TP_GT, TP_LE, TP_GE, TP_EQ, TP_NE, TP_AND, TP_OR,
  TP_IF,    // c a b -> c ? a : b (fparser's if; both branches are evaluated)
  TP_F1,    // one-argument function arg (TpFn)
  TP_F2,    // two-argument function arg (TpFn)
};

enum TpFn : int {
  TF_EXP = 0, TF_EXP2, TF_LOG, TF_LOG10, TF_LOG2, TF_SQRT, TF_CBRT, TF_SIN, TF_COS, TF_TAN,
  TF_ASIN, TF_ACOS, TF_ATAN, TF_SINH, TF_COSH, TF_TANH, TF_ABS, TF_FLOOR, TF_CEIL, TF_TRUNC,
  TF_INT,
  TF_POW, TF_ATAN2, TF_MIN, TF_MAX, TF_HYPOT,
};

// deepest evaluation stack a program may need
constexpr int TP_MAX_DEPTH = 16;

}  // namespace tsm
