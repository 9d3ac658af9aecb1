// expr.h -- expression evaluator for initial/boundary/exact-solution fields.
//
// Evaluates the function-parser syntax the reference's param.in uses
// (src/fparser.h:20-35 wraps the vendored fparser 4.5.1 with the constant
// `pi`, the user constants of the `constants` section, the variables x, y, t
// and the function pert(type, x, y, k)).  Supported: numbers, + - * / % ^,
// unary - and !, comparisons = != < <= > >= (with fparser's 1e-12 tolerance),
// & and |, and the functions abs acos asin atan atan2 cbrt ceil cos cosh exp
// exp2 floor hypot if int log log10 log2 max min pow sin sinh sqrt tan tanh
// trunc pert.
#pragma once

#include <functional>
#include <map>
#include <memory>
#include <string>
#include <vector>

namespace tsm {

class Expr {
 public:
  // Parses `text`; returns false and sets `err` on a syntax error.
  bool parse(const std::string& text, const std::map<std::string, double>& constants,
             std::string& err);
  // pert(type, x, y, k) callback (src/perturb.cc:16-48)
  void set_pert(std::function<double(int, double, double, double)> f) { pert_ = std::move(f); }
  double eval(double x, double y, double t) const;
  bool uses_time() const { return uses_t_; }
  // Compiles the expression into a boundary-time program (tprog.h): every
  // largest subexpression without t becomes a slot whose value time_slot()
  // gives at a point.  Fails on pert() (the program has no random numbers)
  // and on a wrong argument count.
  bool time_program(std::vector<int>& code, int& depth, std::string& err);
  int n_time_slots() const { return (int)tslots_.size(); }
  double time_slot(int k, double x, double y) const;

 private:
  struct Node;
  std::shared_ptr<Node> root_;
  std::function<double(int, double, double, double)> pert_;
  bool uses_t_ = false;
  std::vector<const Node*> tslots_;
  double eval_node(const Node* n, const double* v) const;
  bool emit_tp(const Node* n, std::vector<int>& code, int& sp, int& depth, std::string& err);
  static bool node_uses_t(const Node* n);
  static bool node_has_pert(const Node* n);
  friend struct ExprParser;
};

}  // namespace tsm
