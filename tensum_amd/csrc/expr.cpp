// expr.cpp -- recursive-descent parser and tree evaluator for param.in fields.
//
// Precedence, lowest first, as in fparser 4.5.1: | ; & ; comparisons ;
// + - ; * / % ; unary - and ! ; ^ (right associative, binds tighter than the
// unary minus).  Comparisons return 1 or 0 with fparser's absolute tolerance
// of 1e-12 (x < y holds when x < y - 1e-12); & | ! treat |v| >= 0.5 as true.
#include "expr.h"
#include "tprog.h"

#include <cctype>
#include <cmath>
#include <cstdlib>
#include <cstring>

namespace tsm {

struct Expr::Node {
  enum Kind { NUM, VAR, UN, BIN, FUN } kind;
  double val = 0.0;
  int var = 0;          // 0 x, 1 y, 2 t
  char op = 0;          // unary: '-', '!'; binary: + - * / % ^ < > l g = n & |
  std::string fname;
  std::vector<std::shared_ptr<Node>> kids;
};

struct ExprParser {
  const std::string& s;
  const std::map<std::string, double>& consts;
  size_t i = 0;
  std::string err;
  bool uses_t = false;
  using P = std::shared_ptr<Expr::Node>;

  ExprParser(const std::string& s_, const std::map<std::string, double>& c) : s(s_), consts(c) {}
  void ws() { while (i < s.size() && std::isspace((unsigned char)s[i])) ++i; }
  bool eat(const char* tok) {
    ws();
    size_t n = std::strlen(tok);
    if (s.compare(i, n, tok) == 0) { i += n; return true; }
    return false;
  }
  P mk(Expr::Node::Kind k) { auto n = std::make_shared<Expr::Node>(); n->kind = k; return n; }
  P bin(char op, P a, P b) { P n = mk(Expr::Node::BIN); n->op = op; n->kids = {a, b}; return n; }

  P parse_or() {
    P a = parse_and();
    while (a && eat("|")) { P b = parse_and(); if (!b) return nullptr; a = bin('|', a, b); }
    return a;
  }
  P parse_and() {
    P a = parse_cmp();
    while (a && eat("&")) { P b = parse_cmp(); if (!b) return nullptr; a = bin('&', a, b); }
    return a;
  }
  P parse_cmp() {
    P a = parse_add();
    while (a) {
      char op = 0;
      if (eat("<=")) op = 'l';
      else if (eat(">=")) op = 'g';
      else if (eat("!=")) op = 'n';
      else if (eat("<")) op = '<';
      else if (eat(">")) op = '>';
      else if (eat("=")) op = '=';
      else break;
      P b = parse_add();
      if (!b) return nullptr;
      a = bin(op, a, b);
    }
    return a;
  }
  P parse_add() {
    P a = parse_mul();
    while (a) {
      ws();
      if (i < s.size() && (s[i] == '+' || s[i] == '-')) {
        char op = s[i++];
        P b = parse_mul();
        if (!b) return nullptr;
        a = bin(op, a, b);
      } else break;
    }
    return a;
  }
  P parse_mul() {
    P a = parse_unary();
    while (a) {
      ws();
      if (i < s.size() && (s[i] == '*' || s[i] == '/' || s[i] == '%')) {
        char op = s[i++];
        P b = parse_unary();
        if (!b) return nullptr;
        a = bin(op, a, b);
      } else break;
    }
    return a;
  }
  P parse_unary() {
    ws();
    if (i < s.size() && (s[i] == '-' || s[i] == '!')) {
      char op = s[i++];
      P a = parse_unary();
      if (!a) return nullptr;
      P n = mk(Expr::Node::UN); n->op = op; n->kids = {a};
      return n;
    }
    if (i < s.size() && s[i] == '+') { ++i; return parse_unary(); }
    return parse_pow();
  }
  P parse_pow() {
    P a = parse_primary();
    if (a && eat("^")) {
      P b = parse_unary();  // right associative; a^-b allowed
      if (!b) return nullptr;
      return bin('^', a, b);
    }
    return a;
  }
  P parse_primary() {
    ws();
    if (i >= s.size()) { err = "unexpected end of expression"; return nullptr; }
    char c = s[i];
    if (c == '(') {
      ++i;
      P a = parse_or();
      if (!a) return nullptr;
      if (!eat(")")) { err = "missing )"; return nullptr; }
      return a;
    }
    if (std::isdigit((unsigned char)c) || c == '.') {
      const char* b = s.c_str() + i;
      char* e = nullptr;
      double v = std::strtod(b, &e);
      if (e == b) { err = "bad number"; return nullptr; }
      i += (size_t)(e - b);
      P n = mk(Expr::Node::NUM); n->val = v;
      return n;
    }
    if (std::isalpha((unsigned char)c) || c == '_') {
      size_t j = i;
      while (j < s.size() && (std::isalnum((unsigned char)s[j]) || s[j] == '_')) ++j;
      std::string name = s.substr(i, j - i);
      i = j;
      ws();
      if (i < s.size() && s[i] == '(') {
        ++i;
        P n = mk(Expr::Node::FUN);
        n->fname = name;
        ws();
        if (!(i < s.size() && s[i] == ')')) {
          do {
            P a = parse_or();
            if (!a) return nullptr;
            n->kids.push_back(a);
          } while (eat(","));
        }
        if (!eat(")")) { err = "missing ) after arguments of " + name; return nullptr; }
        static const char* known[] = {"if", "pert", "pow", "exp", "exp2", "log", "log10", "log2",
                                      "sqrt", "cbrt", "sin", "cos", "tan", "asin", "acos", "atan",
                                      "atan2", "sinh", "cosh", "tanh", "abs", "min", "max", "hypot",
                                      "floor", "ceil", "trunc", "int"};
        bool ok = false;
        for (const char* k : known) ok = ok || name == k;
        if (!ok) { err = "unknown function " + name; return nullptr; }
        return n;
      }
      if (name == "x" || name == "y" || name == "t") {
        P n = mk(Expr::Node::VAR);
        n->var = name == "x" ? 0 : (name == "y" ? 1 : 2);
        if (n->var == 2) uses_t = true;
        return n;
      }
      auto it = consts.find(name);
      if (it != consts.end()) { P n = mk(Expr::Node::NUM); n->val = it->second; return n; }
      if (name == "pi") { P n = mk(Expr::Node::NUM); n->val = M_PI; return n; }
      err = "unknown identifier " + name;
      return nullptr;
    }
    err = std::string("unexpected character '") + c + "'";
    return nullptr;
  }
};

bool Expr::parse(const std::string& text, const std::map<std::string, double>& constants,
                 std::string& err) {
  ExprParser p(text, constants);
  root_ = p.parse_or();
  p.ws();
  if (root_ && p.i != text.size()) { root_ = nullptr; p.err = "trailing characters"; }
  if (!root_) { err = p.err + " in \"" + text + "\""; return false; }
  uses_t_ = p.uses_t;
  return true;
}

namespace {
const double kEps = 1e-12;
inline bool truth(double v) { return std::fabs(v) >= 0.5; }
}  // namespace

double Expr::eval_node(const Node* n, const double* v) const {
  switch (n->kind) {
    case Node::NUM: return n->val;
    case Node::VAR: return v[n->var];
    case Node::UN: {
      double a = eval_node(n->kids[0].get(), v);
      return n->op == '-' ? -a : (truth(a) ? 0.0 : 1.0);
    }
    case Node::BIN: {
      double a = eval_node(n->kids[0].get(), v), b = eval_node(n->kids[1].get(), v);
      switch (n->op) {
        case '+': return a + b;
        case '-': return a - b;
        case '*': return a * b;
        case '/': return a / b;
        case '%': return std::fmod(a, b);
        case '^': return std::pow(a, b);
        case '<': return (a < b - kEps) ? 1.0 : 0.0;
        case '>': return (b < a - kEps) ? 1.0 : 0.0;
        case 'l': return (a <= b + kEps) ? 1.0 : 0.0;
        case 'g': return (b <= a + kEps) ? 1.0 : 0.0;
        case '=': return (std::fabs(a - b) <= kEps) ? 1.0 : 0.0;
        case 'n': return (std::fabs(a - b) > kEps) ? 1.0 : 0.0;
        case '&': return (truth(a) && truth(b)) ? 1.0 : 0.0;
        case '|': return (truth(a) || truth(b)) ? 1.0 : 0.0;
      }
      return 0.0;
    }
    case Node::FUN: {
      const std::string& f = n->fname;
      const size_t na = n->kids.size();
      double a[4] = {0, 0, 0, 0};
      if (f == "if") {  // lazy
        if (na != 3) return NAN;
        return truth(eval_node(n->kids[0].get(), v)) ? eval_node(n->kids[1].get(), v)
                                                       : eval_node(n->kids[2].get(), v);
      }
      for (size_t k = 0; k < na && k < 4; ++k) a[k] = eval_node(n->kids[k].get(), v);
      if (f == "pert") return pert_ ? pert_((int)a[0], a[1], a[2], a[3]) : 0.0;
      if (f == "pow") return std::pow(a[0], a[1]);
      if (f == "exp") return std::exp(a[0]);
      if (f == "exp2") return std::exp2(a[0]);
      if (f == "log") return std::log(a[0]);
      if (f == "log10") return std::log10(a[0]);
      if (f == "log2") return std::log2(a[0]);
      if (f == "sqrt") return std::sqrt(a[0]);
      if (f == "cbrt") return std::cbrt(a[0]);
      if (f == "sin") return std::sin(a[0]);
      if (f == "cos") return std::cos(a[0]);
      if (f == "tan") return std::tan(a[0]);
      if (f == "asin") return std::asin(a[0]);
      if (f == "acos") return std::acos(a[0]);
      if (f == "atan") return std::atan(a[0]);
      if (f == "atan2") return std::atan2(a[0], a[1]);
      if (f == "sinh") return std::sinh(a[0]);
      if (f == "cosh") return std::cosh(a[0]);
      if (f == "tanh") return std::tanh(a[0]);
      if (f == "abs") return std::fabs(a[0]);
      if (f == "min") return a[0] < a[1] ? a[0] : a[1];
      if (f == "max") return a[0] > a[1] ? a[0] : a[1];
      if (f == "hypot") return std::hypot(a[0], a[1]);
      if (f == "floor") return std::floor(a[0]);
      if (f == "ceil") return std::ceil(a[0]);
      if (f == "trunc") return std::trunc(a[0]);
      if (f == "int") return std::floor(a[0] + 0.5);
      return NAN;
    }
  }
  return NAN;
}

double Expr::eval(double x, double y, double t) const {
  const double v[3] = {x, y, t};
  return root_ ? eval_node(root_.get(), v) : NAN;
}

bool Expr::node_uses_t(const Node* n) {
  if (n->kind == Node::VAR) return n->var == 2;
  for (const auto& k : n->kids)
    if (node_uses_t(k.get())) return true;
  return false;
}

bool Expr::node_has_pert(const Node* n) {
  if (n->kind == Node::FUN && n->fname == "pert") return true;
  for (const auto& k : n->kids)
    if (node_has_pert(k.get())) return true;
  return false;
}

bool Expr::emit_tp(const Node* n, std::vector<int>& code, int& sp, int& depth, std::string& err) {
  auto push = [&](int op) { code.push_back(op); if (++sp > depth) depth = sp; };
  if (!node_uses_t(n)) {
    push(TP_SLOT | (int)tslots_.size() << 8);
    tslots_.push_back(n);
    return true;
  }
  switch (n->kind) {
    case Node::NUM: return false;  // (uses no t)
    case Node::VAR: push(TP_T); return true;
    case Node::UN:
      if (!emit_tp(n->kids[0].get(), code, sp, depth, err)) return false;
      code.push_back(n->op == '-' ? TP_NEG : TP_NOT);
      return true;
    case Node::BIN: {
      if (!emit_tp(n->kids[0].get(), code, sp, depth, err) ||
          !emit_tp(n->kids[1].get(), code, sp, depth, err))
        return false;
      static const char ops[] = "+-*/%^<>lg=n&|";
      static const int codes[] = {TP_ADD, TP_SUB, TP_MUL, TP_DIV, TP_MOD, TP_POW, TP_LT,
                                  TP_GT,  TP_LE,  TP_GE,  TP_EQ,  TP_NE,  TP_AND, TP_OR};
      const char* at = std::strchr(ops, n->op);
      if (!at || !*at) { err = "unknown operator"; return false; }
      code.push_back(codes[at - ops]);
      --sp;
      return true;
    }
    case Node::FUN: {
      static const char* f1[] = {"exp",  "exp2", "log",  "log10", "log2",  "sqrt", "cbrt",
                                 "sin",  "cos",  "tan",  "asin",  "acos",  "atan", "sinh",
                                 "cosh", "tanh", "abs",  "floor", "ceil",  "trunc", "int"};
      static const char* f2[] = {"pow", "atan2", "min", "max", "hypot"};
      const std::string& f = n->fname;
      const size_t na = n->kids.size();
      int op = -1, arity = 0;
      if (f == "if") { op = TP_IF; arity = 3; }
      for (int k = 0; k < (int)(sizeof(f1) / sizeof(f1[0])); ++k)
        if (f == f1[k]) { op = TP_F1 | (TF_EXP + k) << 8; arity = 1; }
      for (int k = 0; k < (int)(sizeof(f2) / sizeof(f2[0])); ++k)
        if (f == f2[k]) { op = TP_F2 | (TF_POW + k) << 8; arity = 2; }
      if (op < 0) { err = "function " + f + " is not available in a boundary value of t"; return false; }
      if ((int)na != arity) { err = "wrong number of arguments to " + f; return false; }
      for (const auto& k : n->kids)
        if (!emit_tp(k.get(), code, sp, depth, err)) return false;
      code.push_back(op);
      sp -= arity - 1;
      return true;
    }
  }
  return false;
}

bool Expr::time_program(std::vector<int>& code, int& depth, std::string& err) {
  code.clear();
  tslots_.clear();
  depth = 0;
  if (!root_) { err = "no expression"; return false; }
  if (node_has_pert(root_.get())) { err = "pert() is not available in a boundary value of t"; return false; }
  int sp = 0;
  if (!emit_tp(root_.get(), code, sp, depth, err)) return false;
  code.push_back(TP_END);
  if (depth > TP_MAX_DEPTH) { err = "expression nests too deeply for the boundary-time evaluator"; return false; }
  return true;
}

double Expr::time_slot(int k, double x, double y) const {
  const double v[3] = {x, y, 0.0};
  return eval_node(tslots_[k], v);
}

}  // namespace tsm
