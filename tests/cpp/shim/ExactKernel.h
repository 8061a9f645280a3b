// Test shim: AbstractKernel / ExactKernel state and CPU ops with the reference's
// signatures (AbstractKernel.h:35-187, ExactKernel.h:33-186), the CPU ops delegating to
// the FP64 oracle (oracle/exact_kernel.c).  Only what the adapter and its test use.
#pragma once
#include <stdexcept>
#include <vector>

#include "LinearAlgebra.h"

extern "C" {
void orc_convolve(int d, int dv, double h, const double *X, long nx, const double *Y,
                  const double *W, long ny, double *V);
void orc_varlin(int d, int dv, double h, const double *X, long nx, const double *Y,
                const double *W, long ny, double *G);
void orc_grad1(int d, int dv, double h, const double *X, const double *alpha, long nx,
               const double *Y, const double *W, long ny, double *out);
int orc_grad_dim(int d, int dv, double h, const double *X, long nx, const double *Y,
                 const double *W, long ny, int dim, double *out);
int orc_special_hessian(int d, double h, const double *Y, const double *beta, const double *xi,
                        long n, double *out);
}

template <class ScalarType, unsigned int PointDim>
class AbstractKernel {
 public:
  AbstractKernel() { SetKernelWidth(1.0); m_Modified = false; }
  AbstractKernel(const MatrixType &X, double h) {
    SetSources(X);
    SetWeights(MatrixType(X.rows(), 1, 1.0));
    SetKernelWidth(h);
  }
  AbstractKernel(const MatrixType &X, const MatrixType &W, double h) {
    SetSources(X);
    SetWeights(W);
    SetKernelWidth(h);
  }
  virtual ~AbstractKernel() {}
  virtual AbstractKernel *Clone() const = 0;
  double GetKernelWidth() const { return m_KernelWidth; }
  virtual void SetKernelWidth(double h) {
    if (h < 1e-10) h = 1e-10;
    m_KernelWidth = h;
    m_KernelWidthSquared = h * h;
    m_Modified = true;
  }
  MatrixType &GetSources() { return m_Sources; }
  void SetSources(const MatrixType &X) { m_Sources = X; m_Modified = true; }
  MatrixType &GetWeights() { return m_Weights; }
  void SetWeights(const MatrixType &W) { m_Weights = W; m_Modified = true; }
  bool IsModified() const { return m_Modified; }
  void SetModified() { m_Modified = true; }
  void UnsetModified() { m_Modified = false; }

  virtual MatrixType Convolve(const MatrixType &X) = 0;
  virtual std::vector<MatrixType> ConvolveGradient(const MatrixType &X) = 0;
  virtual MatrixType ConvolveGradient(const MatrixType &X, const MatrixType &alpha) = 0;
  virtual MatrixType ConvolveGradient(const MatrixType &X, unsigned int dim) = 0;
  virtual MatrixType ConvolveSpecialHessian(const MatrixType &xi) = 0;

 protected:
  MatrixType m_Sources, m_Weights;
  ScalarType m_KernelWidth, m_KernelWidthSquared;
  bool m_Modified;
};

namespace shim {
inline std::vector<double> pack(const MatrixType &M) {
  std::vector<double> v(M.rows() * M.cols());
  for (size_t i = 0; i < M.rows(); i++)
    for (size_t j = 0; j < M.cols(); j++) v[i * M.cols() + j] = M(i, j);
  return v;
}
inline MatrixType unpack(const std::vector<double> &v, size_t r, size_t c) {
  MatrixType M(r, c, 0.0);
  for (size_t i = 0; i < r; i++)
    for (size_t j = 0; j < c; j++) M(i, j) = v[i * c + j];
  return M;
}
}  // namespace shim

template <class ScalarType, unsigned int PointDim>
class ExactKernel : public AbstractKernel<ScalarType, PointDim> {
 public:
  typedef AbstractKernel<ScalarType, PointDim> Superclass;
  ExactKernel() : Superclass() {}
  ExactKernel(const MatrixType &X, double h) : Superclass(X, h) {}
  ExactKernel(const MatrixType &X, const MatrixType &W, double h) : Superclass(X, W, h) {}
  virtual ExactKernel *Clone() const { return new ExactKernel(*this); }

  virtual MatrixType Convolve(const MatrixType &X) {
    check();
    auto x = shim::pack(X), y = shim::pack(this->m_Sources), w = shim::pack(this->m_Weights);
    const size_t dv = this->m_Weights.cols();
    std::vector<double> out(X.rows() * dv);
    orc_convolve(PointDim, dv, this->m_KernelWidth, x.data(), X.rows(), y.data(), w.data(),
                 this->m_Sources.rows(), out.data());
    return shim::unpack(out, X.rows(), dv);
  }
  virtual std::vector<MatrixType> ConvolveGradient(const MatrixType &X) {
    check();
    auto x = shim::pack(X), y = shim::pack(this->m_Sources), w = shim::pack(this->m_Weights);
    const size_t dv = this->m_Weights.cols();
    std::vector<double> g(X.rows() * dv * PointDim);
    orc_varlin(PointDim, dv, this->m_KernelWidth, x.data(), X.rows(), y.data(), w.data(),
               this->m_Sources.rows(), g.data());
    std::vector<MatrixType> out;
    for (size_t i = 0; i < X.rows(); i++) {
      MatrixType Gi(dv, PointDim, 0.0);
      for (size_t k = 0; k < dv; k++)
        for (size_t l = 0; l < PointDim; l++) Gi(k, l) = g[(i * dv + k) * PointDim + l];
      out.push_back(Gi);
    }
    return out;
  }
  virtual MatrixType ConvolveGradient(const MatrixType &X, const MatrixType &alpha) {
    check();
    auto x = shim::pack(X), a = shim::pack(alpha), y = shim::pack(this->m_Sources),
         w = shim::pack(this->m_Weights);
    std::vector<double> out(X.rows() * PointDim);
    orc_grad1(PointDim, this->m_Weights.cols(), this->m_KernelWidth, x.data(), a.data(), X.rows(),
              y.data(), w.data(), this->m_Sources.rows(), out.data());
    return shim::unpack(out, X.rows(), PointDim);
  }
  virtual MatrixType ConvolveGradient(const MatrixType &X, unsigned int dim) {
    check();
    auto x = shim::pack(X), y = shim::pack(this->m_Sources), w = shim::pack(this->m_Weights);
    const size_t dv = this->m_Weights.cols();
    std::vector<double> out(X.rows() * dv);
    if (orc_grad_dim(PointDim, dv, this->m_KernelWidth, x.data(), X.rows(), y.data(), w.data(),
                     this->m_Sources.rows(), (int)dim, out.data()) != 0)
      throw std::runtime_error("dimension index out of bounds");
    return shim::unpack(out, X.rows(), dv);
  }
  virtual MatrixType ConvolveSpecialHessian(const MatrixType &xi) {
    check();
    auto y = shim::pack(this->m_Sources), b = shim::pack(this->m_Weights), e = shim::pack(xi);
    std::vector<double> out(this->m_Sources.rows() * PointDim);
    orc_special_hessian(PointDim, this->m_KernelWidth, y.data(), b.data(), e.data(),
                        this->m_Sources.rows(), out.data());
    return shim::unpack(out, this->m_Sources.rows(), PointDim);
  }

 private:
  void check() {
    if (this->m_Sources.rows() != this->m_Weights.rows())
      throw std::runtime_error("Sources and weights count mismatch");
  }
};
