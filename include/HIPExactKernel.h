/*
 * HIPExactKernel.h — the MI355X exact-kernel type for Deformetrica's KernelFactory.
 *
 * Header-only adapter over the C ABI in dfm_hip.h.  It derives from ExactKernel<S, D>, as
 * CUDAExactKernel does (src/support/kernels/CUDAExactKernel.h:29-81), and overrides the
 * same five operations; everything else (EvaluateKernel*, ConvolveHessian,
 * ConvolveInverse, the *ImageFast variants) stays on the inherited CPU path.
 *
 * Inside the Deformetrica tree it compiles against the real ExactKernel.h and
 * LinearAlgebra.h (Armadillo-backed, column-major MatrixType).  It needs from MatrixType
 * only rows(), cols(), element access M(i, j) and the (rows, cols, fill) constructor, so
 * it packs row-major buffers itself instead of relying on vectorise_row_wise.
 *
 * Errors become std::runtime_error with the library's message, as the reference
 * plugin's checks do (CUDAExactKernel.cxx:50-51, 80-82, 186-189, 200-203, 240-243).
 * Thread safety: each call is synchronous on the calling thread's per-thread stream;
 * kernel objects may be used from several threads at once (one object per thread, as
 * Deformetrica's per-subject clones are).
 */
#pragma once

#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "dfm_hip.h"

template <class ScalarType, unsigned int PointDim>
class HIPExactKernel : public ExactKernel<ScalarType, PointDim> {
 public:
  typedef ExactKernel<ScalarType, PointDim> Superclass;

  HIPExactKernel() : Superclass(), m_Precision(DefaultPrecision()) {}
  HIPExactKernel(const HIPExactKernel &o) : Superclass(), m_Precision(o.m_Precision) {
    this->m_Sources = o.m_Sources;
    this->m_Weights = o.m_Weights;
    this->SetKernelWidth(o.GetKernelWidth());
    if (o.IsModified())
      this->SetModified();
    else
      this->UnsetModified();
  }
  HIPExactKernel(const MatrixType &X, double h) : Superclass(X, h), m_Precision(DefaultPrecision()) {}
  HIPExactKernel(const MatrixType &X, const MatrixType &W, double h)
      : Superclass(X, W, h), m_Precision(DefaultPrecision()) {}

  virtual HIPExactKernel *Clone() const { return new HIPExactKernel(*this); }
  virtual ~HIPExactKernel() {}

  /// DFM_FP64 (the reference arithmetic) or DFM_FP32C.  The process-wide default is
  /// DFM_FP64, or the value of the environment variable DFM_PRECISION ("fp64", "fp32c",
  /// "fp32x" or "bf16mfma").
  void SetPrecision(int precision) { m_Precision = precision; }
  int GetPrecision() const { return m_Precision; }

  static int DefaultPrecision() {
    CheckAbi();
    const char *env = std::getenv("DFM_PRECISION");
    if (env && std::strcmp(env, "fp32c") == 0) return DFM_FP32C;
    if (env && std::strcmp(env, "fp32x") == 0) return DFM_FP32X;
    if (env && std::strcmp(env, "bf16mfma") == 0) return DFM_BF16MFMA;
    return DFM_FP64;
  }
  /// Throws when the loaded libdfm_hip.so was built from another ABI version of dfm_hip.h
  /// than this adapter (struct layouts and signatures would not match).
  static void CheckAbi() {
    const int v = dfm_abi_version();
    if (v != DFM_ABI_VERSION)
      throw std::runtime_error("HIPExactKernel: libdfm_hip.so has ABI version " +
                               std::to_string(v) + ", this header expects " +
                               std::to_string(DFM_ABI_VERSION));
  }


  /// CUDAExactKernel::Convolve (CUDAExactKernel.cxx:41-90).
  virtual MatrixType Convolve(const MatrixType &X) {
    MatrixType &Y = this->GetSources();
    MatrixType &W = this->GetWeights();
    if (Y.rows() != W.rows()) throw std::runtime_error("Sources and weights count mismatch");
    const int dv = (int)W.cols();
    std::vector<double> x = Pack(X), y = Pack(Y), w = Pack(W), g((size_t)X.rows() * dv);
    Check(dfm_gauss_conv(PointDim, dv, this->GetKernelWidth(), x.data(), y.data(), w.data(),
                         g.data(), (int)X.rows(), (int)Y.rows(), m_Precision),
          "Convolve");
    return Unpack(g, X.rows(), dv);
  }

  /// CUDAExactKernel::ConvolveGradient(X, alpha) (CUDAExactKernel.cxx:165-194).
  virtual MatrixType ConvolveGradient(const MatrixType &X, const MatrixType &alpha) {
    MatrixType &Y = this->GetSources();
    MatrixType &W = this->GetWeights();
    if (Y.rows() != W.rows()) throw std::runtime_error("Sources and weights count mismatch");
    const int dv = (int)W.cols();
    if (alpha.rows() != X.rows() || (int)alpha.cols() != dv)
      throw std::runtime_error("In HIPExactKernel::ConvolveGradient(X, alpha) - alpha must be N x Dv");
    std::vector<double> x = Pack(X), a = Pack(alpha), y = Pack(Y), w = Pack(W),
                        g((size_t)X.rows() * PointDim);
    Check(dfm_gauss_grad1conv(PointDim, dv, this->GetKernelWidth(), a.data(), x.data(), y.data(),
                              w.data(), g.data(), (int)X.rows(), (int)Y.rows(), m_Precision),
          "ConvolveGradient(X, alpha)");
    return Unpack(g, X.rows(), PointDim);
  }

  /// CUDAExactKernel::ConvolveGradient(X, dim) (CUDAExactKernel.cxx:228-259), with the
  /// sources taken from the kernel as ExactKernel does (ExactKernel.cxx:294-323).
  virtual MatrixType ConvolveGradient(const MatrixType &X, unsigned int dim) {
    MatrixType &Y = this->GetSources();
    MatrixType &W = this->GetWeights();
    if (Y.rows() != W.rows()) throw std::runtime_error("Sources and weights count mismatch");
    if (dim >= PointDim) throw std::runtime_error("dimension index out of bounds");
    const int dv = (int)W.cols();
    std::vector<double> x = Pack(X), y = Pack(Y), w = Pack(W), g((size_t)X.rows() * dv);
    Check(dfm_gauss_gradconv(PointDim, dv, this->GetKernelWidth(), x.data(), y.data(), w.data(),
                             dim, g.data(), (int)X.rows(), (int)Y.rows(), m_Precision),
          "ConvolveGradient(X, dim)");
    return Unpack(g, X.rows(), dv);
  }

  /// CUDAExactKernel::ConvolveGradient(X) (CUDAExactKernel.cxx:261-320): one Dv x D
  /// matrix per target, G_i(k, l) = sum_j W_j[k] dK/dx_l.
  virtual std::vector<MatrixType> ConvolveGradient(const MatrixType &X) {
    MatrixType &Y = this->GetSources();
    MatrixType &W = this->GetWeights();
    if (Y.rows() != W.rows()) throw std::runtime_error("Sources and weights count mismatch");
    const int dv = (int)W.cols();
    std::vector<double> x = Pack(X), y = Pack(Y), w = Pack(W),
                        g((size_t)X.rows() * dv * PointDim);
    Check(dfm_gauss_gradconv_varlin(PointDim, dv, this->GetKernelWidth(), x.data(), y.data(),
                                    w.data(), g.data(), (int)X.rows(), (int)Y.rows(), m_Precision),
          "ConvolveGradient(X)");
    std::vector<MatrixType> out;
    out.reserve(X.rows());
    for (size_t i = 0; i < (size_t)X.rows(); i++) {
      MatrixType Gi(dv, PointDim, 0.0);
      for (int k = 0; k < dv; k++)
        for (unsigned l = 0; l < PointDim; l++) Gi(k, l) = g[(i * dv + k) * PointDim + l];
      out.push_back(Gi);
    }
    return out;
  }

  /// CUDAExactKernel::ConvolveSpecialHessian (CUDAExactKernel.cxx:196-226).
  virtual MatrixType ConvolveSpecialHessian(const MatrixType &xi) {
    MatrixType &Y = this->GetSources();
    MatrixType &W = this->GetWeights();
    if (Y.rows() != W.rows()) throw std::runtime_error("Sources and weights count mismatch");
    if (Y.rows() != xi.rows()) throw std::runtime_error("Y and Xi count mismatch");
    const int dv = (int)W.cols();
    if (dv != (int)PointDim)
      throw std::runtime_error("In HIPExactKernel::ConvolveSpecialHessian(xi) - Invalid number of columns of beta !");
    std::vector<double> y = Pack(Y), w = Pack(W), e = Pack(xi), g((size_t)Y.rows() * PointDim);
    Check(dfm_gauss_graddiffconv(PointDim, dv, this->GetKernelWidth(), y.data(), w.data(), e.data(),
                                 g.data(), (int)Y.rows(), m_Precision),
          "ConvolveSpecialHessian");
    return Unpack(g, Y.rows(), PointDim);
  }

 private:
  int m_Precision;

  static std::vector<double> Pack(const MatrixType &M) {
    const size_t r = M.rows(), c = M.cols();
    std::vector<double> v(r * c);
    for (size_t i = 0; i < r; i++)
      for (size_t j = 0; j < c; j++) v[i * c + j] = M(i, j);
    return v;
  }

  static MatrixType Unpack(const std::vector<double> &v, size_t r, size_t c) {
    MatrixType M(r, c, 0.0);
    for (size_t i = 0; i < r; i++)
      for (size_t j = 0; j < c; j++) M(i, j) = v[i * c + j];
    return M;
  }

  static void Check(int rc, const char *what) {
    if (rc == DFM_OK) return;
    if (rc == DFM_ERR_DIM)
      throw std::runtime_error(std::string("In HIPExactKernel::") + what +
                               " - Invalid number of columns of beta ! (" + dfm_last_error() + ")");
    if (rc == DFM_ERR_INDEX) throw std::runtime_error("dimension index out of bounds");
    throw std::runtime_error(std::string("HIPExactKernel::") + what + ": " + dfm_last_error());
  }
};
