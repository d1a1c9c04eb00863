// latex_expr.h — host-side evaluator for the LaTeX expressions a SolveInput carries (RHS,
// variable coefficients, boundary values), used by the HIP plug-in to sample them onto the
// grid before upload, as the CUDA plug-in does with the reference's evaluator
// (src/backends/cuda/solvers/relaxation.cu:186-253).  Standalone builds use this parser;
// in-tree builds define PDEHIP_REFERENCE_EVALUATOR and sample through the reference's
// ExpressionEvaluator (include/expression_eval.h:48-60) instead.
//
// Supported: numbers, x y z t, \pi, + - * / ^, implicit products, parentheses and braces,
// \left( \right), \frac{}{}, \sqrt{}, \sin \cos \tan \exp \ln \log \sinh \cosh \tanh \arctan
// \abs and |..|, comparisons (< > \le \ge <= >=) returning 1 or 0.
#pragma once

#include <memory>
#include <string>

namespace pdehip_host {

class LatexExpr {
 public:
  static LatexExpr Parse(const std::string& text);
  LatexExpr();
  ~LatexExpr();
  LatexExpr(LatexExpr&&) noexcept;
  LatexExpr& operator=(LatexExpr&&) noexcept;
  bool ok() const { return error_.empty() && root_ != nullptr; }
  const std::string& error() const { return error_; }
  double Eval(double x, double y, double z, double t) const;

  struct Node;

 private:
  std::unique_ptr<Node> root_;
  std::string error_;
};

}  // namespace pdehip_host
