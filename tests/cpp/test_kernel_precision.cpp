// The reference's GPU-vs-exact kernel precision tests
// (tests/unit_tests/kernels/TestKernelPrecisionCUDA.cxx:18-203 with the fixture of
// AbstractTestKernelPrecision.{h,cxx}), run against the plugin adapter HIPExactKernel.
// The "exact" side is the FP64 oracle behind the test shim's ExactKernel.
//
// Usage: test_kernel_precision [fp64|fp32c]   (exit status 0 = all cases pass)
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "ExactKernel.h"
#include "HIPExactKernel.h"

namespace {

int g_failures = 0;
int g_precision = DFM_FP64;
double g_tol = 1e-10;  // AbstractTestKernelPrecision.h:78 (double)

// AbstractTestKernelPrecision.cxx:21-29
void generate_random_matrix(MatrixType &X) {
  std::mt19937 gen(1);
  for (size_t i = 0; i < X.rows(); ++i)
    for (size_t j = 0; j < X.cols(); ++j)
      X(i, j) = std::exp(-(double)gen() / (double)(gen.max()));
}

MatrixType random_matrix(size_t r, size_t c) {
  MatrixType M(r, c);
  generate_random_matrix(M);
  return M;
}

// AbstractTestKernelPrecision.cxx:48-63 (per element min(rel, abs)) for FP64; for the
// FP32-compensated path the norm-wise relative error.
void compare(const MatrixType &a, const MatrixType &b, const std::string &msg) {
  if (a.rows() != b.rows() || a.cols() != b.cols()) {
    std::printf("FAIL %s: shape\n", msg.c_str());
    g_failures++;
    return;
  }
  double worst = 0.0, num = 0.0, den = 0.0;
  for (size_t i = 0; i < a.rows(); ++i)
    for (size_t j = 0; j < a.cols(); ++j) {
      const double x = a(i, j), y = b(i, j), d = std::abs(x - y);
      worst = std::max(worst, std::min(d / std::abs(x), d));
      num += d * d;
      den += x * x;
    }
  const double nw = den > 0 ? std::sqrt(num / den) : std::sqrt(num);
  const double measured = g_precision == DFM_FP64 ? worst : nw;
  const bool ok = measured <= g_tol;
  if (!ok) g_failures++;
  std::printf("%s %-52s min(rel,abs)=%.3e normwise=%.3e\n", ok ? "ok  " : "FAIL", msg.c_str(),
              worst, nw);
}

struct Fixture {
  double kernel_width = 1.88;  // AbstractTestKernelPrecision.h:26
  MatrixType X2D = random_matrix(501, 2), Y2D = random_matrix(201, 2), W2D = random_matrix(201, 2),
             Z2D = random_matrix(501, 2);
  MatrixType X3D = random_matrix(501, 3), Y3D = random_matrix(201, 3), W3D = random_matrix(201, 3),
             Z3D = random_matrix(501, 3);
  MatrixType W4D = random_matrix(201, 4), W6D = random_matrix(201, 6);
};

template <unsigned D>
struct Pair {
  ExactKernel<double, D> exact;
  HIPExactKernel<double, D> hip;
  Pair(const MatrixType &Y, double w) {
    exact.SetSources(Y);
    exact.SetKernelWidth(w);
    hip.SetSources(Y);
    hip.SetKernelWidth(w);
    hip.SetPrecision(g_precision);
  }
  void weights(const MatrixType &W) {
    exact.SetWeights(W);
    hip.SetWeights(W);
  }
};

void run_reference_cases(const Fixture &f) {
  Pair<2> k2(f.Y2D, f.kernel_width);
  Pair<3> k3(f.Y3D, f.kernel_width);

  k2.weights(f.W2D);  // exact_vs_cuda_Convolve_2
  compare(k2.exact.Convolve(f.X2D), k2.hip.Convolve(f.X2D), "Convolve_2");
  k2.weights(f.W4D);  // exact_vs_cuda_Convolve_4
  compare(k2.exact.Convolve(f.X2D), k2.hip.Convolve(f.X2D), "Convolve_4");
  k2.weights(f.W2D);  // exact_vs_cuda_ConvolveGradient_2
  compare(k2.exact.ConvolveGradient(f.X2D, f.Z2D), k2.hip.ConvolveGradient(f.X2D, f.Z2D),
          "ConvolveGradient_2");
  for (const MatrixType *W : {&f.W4D, &f.W2D}) {  // ConvolveGradientVarlin_4 / _2
    k2.weights(*W);
    auto e = k2.exact.ConvolveGradient(f.X2D);
    auto g = k2.hip.ConvolveGradient(f.X2D);
    for (size_t j = 0; j < e.size(); j++)
      if (j == 0 || j + 1 == e.size() || g_precision == DFM_FP64)
        compare(e[j], g[j], "ConvolveGradientVarlin_" + std::to_string(W->cols()) + "[" +
                                std::to_string(j) + "]");
  }
  k2.weights(f.W2D);  // exact_vs_cuda_ConvolveGradient1_2
  for (unsigned dim = 0; dim < 2; dim++)
    compare(k2.exact.ConvolveGradient(f.Y2D, dim), k2.hip.ConvolveGradient(f.Y2D, dim),
            "ConvolveGradient1_2 dim " + std::to_string(dim));
  compare(k2.exact.ConvolveSpecialHessian(f.W2D), k2.hip.ConvolveSpecialHessian(f.W2D),
          "ConvolveSpecialHessian_2");

  k3.weights(f.W3D);
  compare(k3.exact.Convolve(f.X3D), k3.hip.Convolve(f.X3D), "Convolve_3");
  k3.weights(f.W6D);
  compare(k3.exact.Convolve(f.X3D), k3.hip.Convolve(f.X3D), "Convolve_6");
  for (const MatrixType *W : {&f.W6D, &f.W3D}) {  // ConvolveGradientVarlin_6 / _3
    k3.weights(*W);
    auto e = k3.exact.ConvolveGradient(f.X3D);
    auto g = k3.hip.ConvolveGradient(f.X3D);
    for (size_t j = 0; j < e.size(); j++)
      if (j == 0 || j + 1 == e.size() || g_precision == DFM_FP64)
        compare(e[j], g[j], "ConvolveGradientVarlin_" + std::to_string(W->cols()) + "[" +
                                std::to_string(j) + "]");
  }
  k3.weights(f.W3D);
  compare(k3.exact.ConvolveGradient(f.X3D, f.Z3D), k3.hip.ConvolveGradient(f.X3D, f.Z3D),
          "ConvolveGradient_3");
  for (unsigned dim = 0; dim < 3; dim++)
    compare(k3.exact.ConvolveGradient(f.Y3D, dim), k3.hip.ConvolveGradient(f.Y3D, dim),
            "ConvolveGradient1_3 dim " + std::to_string(dim));
  compare(k3.exact.ConvolveSpecialHessian(f.W3D), k3.hip.ConvolveSpecialHessian(f.W3D),
          "ConvolveSpecialHessian_3");
}

// Plugin behaviour the callers rely on: Clone copies state, errors are thrown as the
// reference's runtime_errors, and per-subject threads can call kernel objects at once.
void run_plugin_cases(const Fixture &f) {
  HIPExactKernel<double, 3> k;
  k.SetPrecision(g_precision);
  k.SetKernelWidth(f.kernel_width);
  k.SetSources(f.Y3D);
  k.SetWeights(f.W3D);
  HIPExactKernel<double, 3> *c = k.Clone();
  compare(k.Convolve(f.X3D), c->Convolve(f.X3D), "Clone gives the same Convolve");
  delete c;

  bool threw = false;
  k.SetWeights(MatrixType(200, 3, 1.0));
  try {
    k.Convolve(f.X3D);
  } catch (const std::runtime_error &e) {
    threw = std::strstr(e.what(), "mismatch") != nullptr;
  }
  std::printf("%s mismatch throws runtime_error\n", threw ? "ok  " : "FAIL");
  if (!threw) g_failures++;

  threw = false;
  k.SetWeights(MatrixType(201, 7, 1.0));
  try {
    k.Convolve(f.X3D);
  } catch (const std::runtime_error &e) {
    threw = std::strstr(e.what(), "Invalid number of columns") != nullptr;
  }
  std::printf("%s unsupported Dv throws runtime_error\n", threw ? "ok  " : "FAIL");
  if (!threw) g_failures++;

  threw = false;
  k.SetWeights(f.W3D);
  try {
    k.ConvolveGradient(f.Y3D, 3u);
  } catch (const std::runtime_error &e) {
    threw = std::strstr(e.what(), "out of bounds") != nullptr;
  }
  std::printf("%s bad dim throws runtime_error\n", threw ? "ok  " : "FAIL");
  if (!threw) g_failures++;

  // concurrent subjects: one kernel object per thread, as AbstractAtlas's threads clone
  ExactKernel<double, 3> ex;
  ex.SetKernelWidth(0.3);
  ex.SetSources(f.Y3D);
  ex.SetWeights(f.W3D);
  const MatrixType expect = ex.Convolve(f.X3D);
  std::vector<MatrixType> got(8);
  std::vector<std::thread> th;
  for (int t = 0; t < 8; t++)
    th.emplace_back([&, t] {
      HIPExactKernel<double, 3> kt;
      kt.SetPrecision(g_precision);
      kt.SetKernelWidth(0.3);
      kt.SetSources(f.Y3D);
      kt.SetWeights(f.W3D);
      for (int rep = 0; rep < 5; rep++) got[t] = kt.Convolve(f.X3D);
    });
  for (auto &t : th) t.join();
  for (int t = 0; t < 8; t++) compare(expect, got[t], "thread " + std::to_string(t) + " Convolve");
}

}  // namespace

int main(int argc, char **argv) {
  if (argc > 1 && std::strcmp(argv[1], "fp32c") == 0) {
    g_precision = DFM_FP32C;
    g_tol = 1e-6;
  } else if (argc > 1 && std::strcmp(argv[1], "fp32x") == 0) {
    g_precision = DFM_FP32X;
    g_tol = 1e-6;
  }
  std::printf("precision %s, tolerance %.0e\n", argc > 1 ? argv[1] : "fp64", g_tol);
  Fixture f;
  run_reference_cases(f);
  run_plugin_cases(f);
  std::printf("%d failure(s)\n", g_failures);
  return g_failures == 0 ? 0 : 1;
}
