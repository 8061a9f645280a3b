// Test shim: the few MatrixType members the plugin adapter touches, column-major like
// Deformetrica's Armadillo-backed MatrixType (ArmadilloMatrixWrapper.h:99-101).
#pragma once
#include <cstddef>
#include <vector>

class MatrixType {
 public:
  MatrixType() : r_(0), c_(0) {}
  MatrixType(size_t r, size_t c, double fill = 0.0) : r_(r), c_(c), v_(r * c, fill) {}
  size_t rows() const { return r_; }
  size_t cols() const { return c_; }
  size_t columns() const { return c_; }
  double &operator()(size_t i, size_t j) { return v_[j * r_ + i]; }
  double operator()(size_t i, size_t j) const { return v_[j * r_ + i]; }

 private:
  size_t r_, c_;
  std::vector<double> v_;
};
