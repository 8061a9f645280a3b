// Timing of the plugin path as the reference's landmark shooting calls it: Diffeos::Shoot
// (Diffeos.cxx:617-632) makes, per Euler step, one Convolve(cp) and one
// ConvolveGradient(cp, mom) on a kernel whose sources and weights are the control points
// and momenta, through the host-matrix overloads that CUDAExactKernel implements
// (CUDAExactKernel.cxx:41-90).  Each of T threads shoots on its own HIPExactKernel, as the
// reference's per-subject threads do (AbstractAtlas.cxx:489-557).
//
// Usage: bench_plugin_shoot [threads] [fp64|fp32c] [points] [steps] [shoots per thread]
// Prints one JSON line: wall ms per shoot per thread, calls per second over all threads.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <thread>
#include <vector>

#include "ExactKernel.h"
#include "HIPExactKernel.h"

namespace {

double shoot(HIPExactKernel<double, 3> &k, MatrixType cp, MatrixType mom, int steps) {
  const double dt = 1.0 / steps;
  for (int s = 0; s < steps; s++) {
    k.SetSources(cp);
    k.SetWeights(mom);
    const MatrixType v = k.Convolve(cp);
    const MatrixType g = k.ConvolveGradient(cp, mom);
    for (size_t i = 0; i < cp.rows(); i++)
      for (size_t j = 0; j < 3; j++) {
        cp(i, j) += dt * v(i, j);
        mom(i, j) -= dt * g(i, j);
      }
  }
  return cp(0, 0);
}

}  // namespace

int main(int argc, char **argv) {
  const int threads = argc > 1 ? std::atoi(argv[1]) : 1;
  const int prec = (argc > 2 && std::strcmp(argv[2], "fp32c") == 0) ? DFM_FP32C : DFM_FP64;
  const size_t n = argc > 3 ? (size_t)std::atol(argv[3]) : 1000;
  const int steps = argc > 4 ? std::atoi(argv[4]) : 10;
  const int reps = argc > 5 ? std::atoi(argv[5]) : 20;

  std::vector<MatrixType> cps, moms;
  for (int t = 0; t < threads; t++) {
    std::mt19937 gen(100 + t);
    std::uniform_real_distribution<double> u(0.0, 1.0);
    std::normal_distribution<double> g(0.0, 0.01);
    MatrixType cp(n, 3), mom(n, 3);
    for (size_t i = 0; i < n; i++)
      for (size_t j = 0; j < 3; j++) {
        cp(i, j) = u(gen);
        mom(i, j) = g(gen);
      }
    cps.push_back(cp);
    moms.push_back(mom);
  }
  std::vector<HIPExactKernel<double, 3>> ks(threads);
  for (auto &k : ks) {
    k.SetKernelWidth(0.2);
    k.SetPrecision(prec);
  }
  for (int t = 0; t < threads; t++) shoot(ks[t], cps[t], moms[t], steps);  // warm-up

  std::vector<double> sink(threads);
  const auto t0 = std::chrono::steady_clock::now();
  std::vector<std::thread> th;
  for (int t = 0; t < threads; t++)
    th.emplace_back([&, t] {
      for (int r = 0; r < reps; r++) sink[t] += shoot(ks[t], cps[t], moms[t], steps);
    });
  for (auto &x : th) x.join();
  const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  const double calls = 2.0 * steps * reps * threads;
  std::printf("{\"threads\": %d, \"precision\": \"%s\", \"points\": %zu, \"steps\": %d, "
              "\"shoots_per_thread\": %d, \"ms_per_shoot\": %.4f, \"us_per_call\": %.2f, "
              "\"calls_per_s\": %.1f, \"check\": %.6f}\n",
              threads, prec == DFM_FP32C ? "fp32c" : "fp64", n, steps, reps, s * 1e3 / reps,
              s * 1e6 / (2.0 * steps * reps), calls / s, sink[0]);
  return std::isfinite(sink[0]) ? 0 : 1;
}
