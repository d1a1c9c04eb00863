// latex_expr.cpp — recursive-descent parser and tree evaluator (see latex_expr.h).
#include "latex_expr.h"

#include <cctype>
#include <cmath>
#include <stdexcept>
#include <vector>

namespace pdehip_host {

struct LatexExpr::Node {
  enum Kind { Num, Var, Neg, Add, Sub, Mul, Div, Pow, Func, Cmp } kind = Num;
  double value = 0.0;
  int var = 0;          // 0 x, 1 y, 2 z, 3 t
  int op = 0;           // function id / comparison id
  std::unique_ptr<Node> a, b;
};

namespace {

using NodeP = std::unique_ptr<LatexExpr::Node>;

NodeP make(LatexExpr::Node::Kind k) {
  NodeP n(new LatexExpr::Node());
  n->kind = k;
  return n;
}

NodeP bin(LatexExpr::Node::Kind k, NodeP a, NodeP b) {
  NodeP n = make(k);
  n->a = std::move(a);
  n->b = std::move(b);
  return n;
}

enum Fn { F_SIN, F_COS, F_TAN, F_EXP, F_LN, F_LOG, F_SQRT, F_SINH, F_COSH, F_TANH, F_ATAN, F_ABS };

struct Parser {
  std::string s;
  size_t i = 0;

  [[noreturn]] void fail(const std::string& m) { throw std::runtime_error(m); }
  void ws() { while (i < s.size() && (std::isspace(static_cast<unsigned char>(s[i])) || s.compare(i, 2, "\\,") == 0 || s.compare(i, 2, "\\;") == 0 || s.compare(i, 2, "\\ ") == 0)) i += (s[i] == '\\') ? 2 : 1; }
  bool peek(const char* t) {
    ws();
    return s.compare(i, std::char_traits<char>::length(t), t) == 0;
  }
  bool eat(const char* t) {
    if (peek(t)) { i += std::char_traits<char>::length(t); return true; }
    return false;
  }
  // command word after a backslash, e.g. "sin" for "\sin"
  bool eat_cmd(const char* name) {
    ws();
    const size_t n = std::char_traits<char>::length(name);
    if (i < s.size() && s[i] == '\\' && s.compare(i + 1, n, name) == 0 &&
        (i + 1 + n >= s.size() || !std::isalpha(static_cast<unsigned char>(s[i + 1 + n])))) {
      i += 1 + n;
      return true;
    }
    return false;
  }

  NodeP comparison() {
    NodeP l = expr();
    for (;;) {
      int op = -1;
      if (eat("<=") || eat_cmd("le") || eat_cmd("leq")) op = 1;
      else if (eat(">=") || eat_cmd("ge") || eat_cmd("geq")) op = 3;
      else if (eat("<") || eat_cmd("lt")) op = 0;
      else if (eat(">") || eat_cmd("gt")) op = 2;
      if (op < 0) return l;
      NodeP r = expr();
      NodeP c = bin(LatexExpr::Node::Cmp, std::move(l), std::move(r));
      c->op = op;
      l = std::move(c);
    }
  }

  NodeP expr() {
    NodeP l = term();
    for (;;) {
      if (eat("+")) l = bin(LatexExpr::Node::Add, std::move(l), term());
      else if (eat("-")) l = bin(LatexExpr::Node::Sub, std::move(l), term());
      else return l;
    }
  }

  bool starts_factor() {
    ws();
    if (i >= s.size()) return false;
    const char c = s[i];
    if (std::isdigit(static_cast<unsigned char>(c)) || c == '.' || c == '(' || c == '{' || c == '|') return true;
    if (c == 'x' || c == 'y' || c == 'z' || c == 't') return true;
    if (c == '\\') {
      return !(s.compare(i, 6, "\\right") == 0 || s.compare(i, 3, "\\le") == 0 ||
               s.compare(i, 3, "\\ge") == 0 || s.compare(i, 3, "\\lt") == 0 ||
               s.compare(i, 3, "\\gt") == 0 || s.compare(i, 6, "\\cdot") == 0 ||
               s.compare(i, 6, "\\times") == 0);
    }
    return false;
  }

  NodeP term() {
    NodeP l = unary();
    for (;;) {
      if (eat("*") || eat_cmd("cdot") || eat_cmd("times")) l = bin(LatexExpr::Node::Mul, std::move(l), unary());
      else if (eat("/")) l = bin(LatexExpr::Node::Div, std::move(l), unary());
      else if (starts_factor()) l = bin(LatexExpr::Node::Mul, std::move(l), power());
      else return l;
    }
  }

  NodeP unary() {
    if (eat("-")) { NodeP n = make(LatexExpr::Node::Neg); n->a = unary(); return n; }
    if (eat("+")) return unary();
    return power();
  }

  NodeP power() {
    NodeP base = primary();
    if (eat("^")) {
      NodeP e;
      if (peek("{")) { eat("{"); e = comparison(); if (!eat("}")) fail("expected }"); }
      else e = unary_atom();
      return bin(LatexExpr::Node::Pow, std::move(base), std::move(e));
    }
    return base;
  }

  // exponent without braces: a signed single atom (number, variable, group)
  NodeP unary_atom() {
    if (eat("-")) { NodeP n = make(LatexExpr::Node::Neg); n->a = primary(); return n; }
    return primary();
  }

  NodeP group(char open, char close) {
    (void)open;
    NodeP e = comparison();
    const char cl[2] = {close, 0};
    if (!eat(cl)) fail(std::string("expected ") + close);
    return e;
  }

  NodeP func_arg() {
    if (eat("(")) return group('(', ')');
    if (eat("{")) return group('{', '}');
    if (eat_cmd("left")) {
      if (eat("(")) { NodeP e = comparison(); if (!eat_cmd("right") || !eat(")")) fail("expected \\right)"); return e; }
      if (eat("[")) { NodeP e = comparison(); if (!eat_cmd("right") || !eat("]")) fail("expected \\right]"); return e; }
      fail("unsupported \\left delimiter");
    }
    return power();
  }

  NodeP fn(int id) {
    NodeP n = make(LatexExpr::Node::Func);
    n->op = id;
    if (eat("^")) {   // \sin^2(x)
      NodeP e = unary_atom();
      n->a = func_arg();
      return bin(LatexExpr::Node::Pow, std::move(n), std::move(e));
    }
    n->a = func_arg();
    return n;
  }

  NodeP primary() {
    ws();
    if (i >= s.size()) fail("unexpected end of expression");
    const char c = s[i];
    if (std::isdigit(static_cast<unsigned char>(c)) || c == '.') {
      size_t used = 0;
      const double v = std::stod(s.substr(i), &used);
      i += used;
      NodeP n = make(LatexExpr::Node::Num);
      n->value = v;
      return n;
    }
    if (c == 'x' || c == 'y' || c == 'z' || c == 't') {
      ++i;
      NodeP n = make(LatexExpr::Node::Var);
      n->var = c == 'x' ? 0 : c == 'y' ? 1 : c == 'z' ? 2 : 3;
      return n;
    }
    if (eat("(")) return group('(', ')');
    if (eat("{")) return group('{', '}');
    if (eat("|")) { NodeP n = make(LatexExpr::Node::Func); n->op = F_ABS; n->a = expr(); if (!eat("|")) fail("expected |"); return n; }
    if (eat_cmd("left")) {
      if (eat("(")) { NodeP e = comparison(); if (!eat_cmd("right") || !eat(")")) fail("expected \\right)"); return e; }
      if (eat("|")) { NodeP n = make(LatexExpr::Node::Func); n->op = F_ABS; n->a = comparison(); if (!eat_cmd("right") || !eat("|")) fail("expected \\right|"); return n; }
      fail("unsupported \\left delimiter");
    }
    if (eat_cmd("pi")) { NodeP n = make(LatexExpr::Node::Num); n->value = M_PI; return n; }
    if (eat_cmd("frac")) {
      if (!eat("{")) fail("expected { after \\frac");
      NodeP a = group('{', '}');
      if (!eat("{")) fail("expected second { after \\frac");
      NodeP b = group('{', '}');
      return bin(LatexExpr::Node::Div, std::move(a), std::move(b));
    }
    if (eat_cmd("sqrt")) { NodeP n = make(LatexExpr::Node::Func); n->op = F_SQRT; n->a = func_arg(); return n; }
    static const struct { const char* name; int id; } fns[] = {
        {"sinh", F_SINH}, {"cosh", F_COSH}, {"tanh", F_TANH}, {"sin", F_SIN}, {"cos", F_COS},
        {"tan", F_TAN}, {"exp", F_EXP}, {"ln", F_LN}, {"log", F_LOG}, {"arctan", F_ATAN},
        {"abs", F_ABS}};
    for (const auto& f : fns)
      if (eat_cmd(f.name)) return fn(f.id);
    if (c == 'e' && !(i + 1 < s.size() && std::isalpha(static_cast<unsigned char>(s[i + 1])))) {
      ++i;
      NodeP n = make(LatexExpr::Node::Num);
      n->value = M_E;
      return n;
    }
    fail(std::string("unexpected token at '") + s.substr(i, 12) + "'");
  }
};

double eval(const LatexExpr::Node* n, const double* v) {
  using K = LatexExpr::Node;
  switch (n->kind) {
    case K::Num: return n->value;
    case K::Var: return v[n->var];
    case K::Neg: return -eval(n->a.get(), v);
    case K::Add: return eval(n->a.get(), v) + eval(n->b.get(), v);
    case K::Sub: return eval(n->a.get(), v) - eval(n->b.get(), v);
    case K::Mul: return eval(n->a.get(), v) * eval(n->b.get(), v);
    case K::Div: return eval(n->a.get(), v) / eval(n->b.get(), v);
    case K::Pow: return std::pow(eval(n->a.get(), v), eval(n->b.get(), v));
    case K::Cmp: {
      const double a = eval(n->a.get(), v), b = eval(n->b.get(), v);
      switch (n->op) {
        case 0: return a < b ? 1.0 : 0.0;
        case 1: return a <= b ? 1.0 : 0.0;
        case 2: return a > b ? 1.0 : 0.0;
        default: return a >= b ? 1.0 : 0.0;
      }
    }
    case K::Func: {
      const double a = eval(n->a.get(), v);
      switch (n->op) {
        case F_SIN: return std::sin(a);
        case F_COS: return std::cos(a);
        case F_TAN: return std::tan(a);
        case F_EXP: return std::exp(a);
        case F_LN: return std::log(a);
        case F_LOG: return std::log10(a);
        case F_SQRT: return std::sqrt(a);
        case F_SINH: return std::sinh(a);
        case F_COSH: return std::cosh(a);
        case F_TANH: return std::tanh(a);
        case F_ATAN: return std::atan(a);
        default: return std::abs(a);
      }
    }
  }
  return 0.0;
}

}  // namespace

LatexExpr::LatexExpr() = default;
LatexExpr::~LatexExpr() = default;
LatexExpr::LatexExpr(LatexExpr&&) noexcept = default;
LatexExpr& LatexExpr::operator=(LatexExpr&&) noexcept = default;

LatexExpr LatexExpr::Parse(const std::string& text) {
  LatexExpr e;
  Parser p;
  p.s = text;
  try {
    e.root_ = p.comparison();
    p.ws();
    if (p.i != p.s.size()) throw std::runtime_error("trailing input at '" + p.s.substr(p.i, 12) + "'");
  } catch (const std::exception& ex) {
    e.root_.reset();
    e.error_ = ex.what();
  }
  return e;
}

double LatexExpr::Eval(double x, double y, double z, double t) const {
  const double v[4] = {x, y, z, t};
  return root_ ? eval(root_.get(), v) : 0.0;
}

}  // namespace pdehip_host
