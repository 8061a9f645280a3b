// transport.hip — ExactKernel::ComputeKernelMatrix / ConvolveInverse and
// Diffeos::ParallelTransport on the device.
//
// ConvolveInverse (ExactKernel.cxx:131-164): V = solve(K(Y, Y), X).  The reference's
// Armadillo solve() is LAPACK's dgesv for a square system, LU with partial pivoting; here
// the FP64 kernel matrix is built by a kernel and factored by rocSOLVER's dgetrf / dgetrs
// (the same algorithm), all on the caller's stream.
//
// ParallelTransport (Diffeos.cxx:293-582): the Jacobi-field transport of a momentum along
// the geodesic (P, A) of dfm_shoot, step by step: the velocity's momenta K^-1 v, the
// conservation of the norms by the (alpha, beta) rescaling, and the finite-difference
// (+-eps) Runge-Kutta 2 step of the perturbed geodesic.  Every kernel sum is one pair sum
// of the engine (Convolve; Grad1 with alpha = the momentum, which is the reference's
// varlin list contracted with it), every scalar (the dot products, alpha, beta) stays on
// the device, so the loop runs without host round trips.  The reference writes its
// outputs as parallelTransport[t] for absolute time indices t, which is its per-step
// buffer only when the initial time is t0; here the output of target time t is the
// transport of step t - initial index, the same thing when the initial time is t0.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "dfm_hip.h"
#include "dfm_internal.h"

namespace dfm {
namespace {

unsigned blocks(long n) { return (unsigned)std::max(1L, (n + 255) / 256); }

// K[i][j] = exp(-|y_i - y_j|^2 / sigma^2), n x n (symmetric, so row- and column-major agree)
__global__ __launch_bounds__(256) void k_kernel_matrix(const double *Y, long n, int D, double inv_s2,
                                                       double *K) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n * n) return;
  const long i = idx / n, j = idx - i * n;
  double r2 = 0.0;
  for (int k = 0; k < D; k++) {
    const double d = Y[i * D + k] - Y[j * D + k];
    r2 += d * d;
  }
  K[idx] = exp(-r2 * inv_s2);
}

// column-major B (n x dv) <-> row-major X (n x dv)
__global__ __launch_bounds__(256) void k_transpose(const double *src, double *dst, long rows,
                                                   long cols) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= rows * cols) return;
  const long r = idx / cols, c = idx - r * cols;
  dst[c * rows + r] = src[idx];
}

__global__ void k_flag_info(const rocblas_int *info, double *flag) {
  if (threadIdx.x == 0 && blockIdx.x == 0 && *info != 0) *flag = 1.0;
}

// *out = sum a_i b_i (one workgroup)
__global__ __launch_bounds__(256) void k_dot(const double *a, const double *b, long n, double *out) {
  __shared__ double sh[256];
  double s = 0.0;
  for (long i = threadIdx.x; i < n; i += 256) s += a[i] * b[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sh[0];
}

// out = ca a + cb b (b may be null); coefficients from device scalars when given
__global__ __launch_bounds__(256) void k_axpby(double *out, const double *a, double ca,
                                               const double *b, double cb, long n,
                                               const double *sa, const double *sb) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double A_ = sa ? *sa : ca, B_ = sb ? *sb : cb;
  out[i] = A_ * a[i] + (b ? B_ * b[i] : 0.0);
}

// scalar slots
enum { S_WV, S_W2, S_V2, S_IWV, S_IW2, S_ALPHA, S_BETA, S_PN2, S_PWV, S_BAD, S_COUNT };

// the norm-conserving rescaling (Diffeos.cxx:466-482)
__global__ void k_alpha_beta(double *S) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const double v2 = S[S_V2], wv = S[S_WV];
  S[S_ALPHA] = sqrt((S[S_IW2] * v2 - S[S_IWV] * wv) / (S[S_PN2] * v2 - S[S_PWV] * S[S_PWV]));
  S[S_BETA] = (S[S_IWV] - S[S_ALPHA] * S[S_PWV]) / v2;
}

__global__ void k_set(double *S, int i, double v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) S[i] = v;
}

int launched(const char *what) {
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? DFM_OK : set_error(DFM_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

KParams blank() {
  KParams p;
  std::memset(&p, 0, sizeof(p));
  return p;
}

// one rocBLAS handle per host thread, bound to the call's stream
int handle_for(hipStream_t st, rocblas_handle *h) {
  static thread_local rocblas_handle handle = nullptr;
  if (!handle && rocblas_create_handle(&handle) != rocblas_status_success)
    return set_error(DFM_ERR_HIP, "rocblas_create_handle failed");
  if (rocblas_set_stream(handle, st) != rocblas_status_success)
    return set_error(DFM_ERR_HIP, "rocblas_set_stream failed");
  *h = handle;
  return DFM_OK;
}

size_t a256(size_t v) { return (v + 255) & ~(size_t)255; }

struct InvLayout {
  double *K, *B, *flag;
  rocblas_int *ipiv, *info;
  size_t total;
  InvLayout(char *base, long n, int dv) {
    size_t off = 0;
    auto take = [&](size_t bytes) {
      char *p = base ? base + off : nullptr;
      off += a256(bytes);
      return p;
    };
    K = (double *)take(sizeof(double) * n * n);
    B = (double *)take(sizeof(double) * n * dv);
    flag = (double *)take(sizeof(double));
    ipiv = (rocblas_int *)take(sizeof(rocblas_int) * std::max(1L, n));
    info = (rocblas_int *)take(sizeof(rocblas_int));
    total = off;
  }
};

// V = K(Y)^-1 X on the stream; a singular K sets *flag (read by the caller)
int convolve_inverse(hipStream_t st, int D, int dv, double sigma, const double *Y, long n,
                     const double *X, double *V, const InvLayout &L) {
  if (n == 0) return DFM_OK;
  hipLaunchKernelGGL(k_kernel_matrix, dim3(blocks(n * n)), dim3(256), 0, st, Y, n, D,
                     1.0 / (sigma * sigma), L.K);
  hipLaunchKernelGGL(k_transpose, dim3(blocks(n * dv)), dim3(256), 0, st, X, L.B, n, (long)dv);
  int rc = launched("kernel matrix");
  if (rc) return rc;
  rocblas_handle h;
  if ((rc = handle_for(st, &h))) return rc;
  if (rocsolver_dgetrf(h, (rocblas_int)n, (rocblas_int)n, L.K, (rocblas_int)n, L.ipiv, L.info) !=
      rocblas_status_success)
    return set_error(DFM_ERR_HIP, "rocsolver_dgetrf failed");
  hipLaunchKernelGGL(k_flag_info, dim3(1), dim3(64), 0, st, L.info, L.flag);
  if (rocsolver_dgetrs(h, rocblas_operation_none, (rocblas_int)n, (rocblas_int)dv, L.K,
                       (rocblas_int)n, L.ipiv, L.B, (rocblas_int)n) != rocblas_status_success)
    return set_error(DFM_ERR_HIP, "rocsolver_dgetrs failed");
  hipLaunchKernelGGL(k_transpose, dim3(blocks(n * dv)), dim3(256), 0, st, L.B, V, (long)dv, n);
  return launched("convolve inverse");
}

int read_flag(hipStream_t st, const double *flag, double *out) {
  hipError_t e = hipMemcpyAsync(out, flag, sizeof(double), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  return e == hipSuccess ? DFM_OK : set_error(DFM_ERR_HIP, "flag read: %s", hipGetErrorString(e));
}

struct Transport {
  InvLayout inv;
  double *S, *vel, *veln, *kinv, *dpos, *q, *tr, *mp, *mn, *dmp, *dmn, *cp1, *mom1, *dpos2, *plus;
  void *pw;
  size_t pw_bytes, total;
  Transport(char *base, int prec, int D, long n, long m) : inv(nullptr, n, D) {
    size_t off = 0;
    auto take = [&](size_t bytes) {
      char *p = base ? base + off : nullptr;
      off += a256(bytes);
      return p;
    };
    const InvLayout dry(nullptr, n, D);
    char *ib = take(dry.total);
    inv = InvLayout(ib, n, D);
    S = (double *)take(sizeof(double) * S_COUNT);
    double **v[] = {&vel, &veln, &kinv, &dpos, &q, &tr, &mp, &mn, &dmp, &dmn, &cp1, &mom1, &dpos2, &plus};
    for (double **x : v) *x = (double *)take(sizeof(double) * std::max(n, m) * D);
    pw_bytes = std::max({pairsum_workspace(OP_CONV, prec, D, D, n, n),
                         pairsum_workspace(OP_GRAD1, prec, D, D, n, n),
                         pairsum_workspace(OP_CONV, prec, D, D, n, m)});
    pw = take(pw_bytes);
    total = off;
  }
};

int check_args(int D, int prec, double sigma, long n, long m, int T) {
  if (D != 2 && D != 3) return set_error(DFM_ERR_DIM, "unsupported dimpoint %d (2 or 3)", D);
  if (!valid_precision(prec))
    return set_error(DFM_ERR_ARG, "unknown precision %d", prec);
  if (!(sigma > 0)) return set_error(DFM_ERR_ARG, "kernel width must be positive");
  if (n < 0 || m < 0 || n > 46340 || m > 0x7fffffffL)
    return set_error(DFM_ERR_ARG, "control points must be in [0, 46340] (dense n x n solve)");
  if (T < 1) return set_error(DFM_ERR_ARG, "number of time points must be >= 1");
  return DFM_OK;
}

}  // namespace
}  // namespace dfm

using namespace dfm;

extern "C" {

size_t dfm_convolve_inverse_workspace(int dimpoint, int dimvect, long n) {
  if (n < 0 || n > 46340 || dimvect < 1) return 0;
  (void)dimpoint;
  return InvLayout(nullptr, n, dimvect).total;
}

int dfm_kernel_matrix(void *stream, int D, double sigma, const double *y, long n, double *K) {
  clear_error();
  int rc = check_args(D, DFM_FP64, sigma, n, 0, 1);
  if (rc) return rc;
  if (n == 0) return DFM_OK;
  if (!y || !K) return set_error(DFM_ERR_ARG, "NULL argument");
  hipLaunchKernelGGL(k_kernel_matrix, dim3(blocks(n * n)), dim3(256), 0, (hipStream_t)stream, y,
                     n, D, 1.0 / (sigma * sigma), K);
  return launched("k_kernel_matrix");
}

int dfm_convolve_inverse(void *stream, int D, int dv, double sigma, const double *y, long n,
                         const double *x, double *v, void *work, size_t work_bytes) {
  clear_error();
  int rc = check_args(D, DFM_FP64, sigma, n, 0, 1);
  if (rc) return rc;
  if (dv < 1 || dv > 64) return set_error(DFM_ERR_DIM, "unsupported dimvect %d", dv);
  if (n == 0) return DFM_OK;
  if (!y || !x || !v) return set_error(DFM_ERR_ARG, "NULL argument");
  InvLayout need(nullptr, n, dv);
  if (!work || work_bytes < need.total)
    return set_error(DFM_ERR_ARG, "workspace too small: need %zu bytes", need.total);
  InvLayout L((char *)work, n, dv);
  hipStream_t st = (hipStream_t)stream;
  if ((rc = fill(st, L.flag, 1, 0.0))) return rc;
  if ((rc = convolve_inverse(st, D, dv, sigma, y, n, x, v, L))) return rc;
  double bad = 0.0;
  if ((rc = read_flag(st, L.flag, &bad))) return rc;
  return bad != 0.0 ? set_error(DFM_ERR_ARG, "kernel matrix is singular") : DFM_OK;
}

size_t dfm_parallel_transport_workspace(int dimpoint, long n, long m, int precision) {
  if (check_args(dimpoint, precision, 1.0, n, m, 1)) return 0;
  return Transport(nullptr, precision, dimpoint, n, m).total;
}

int dfm_parallel_transport(void *stream, int D, double sigma, int prec, const double *P,
                           const double *A, long n, int T, double t0, double tN,
                           const double *mom0, const double *cp0, long m, double initial_time,
                           const double *target_times, int ntargets, double *out, void *work,
                           size_t work_bytes) {
  clear_error();
  int rc = check_args(D, prec, sigma, n, m, T);
  if (rc) return rc;
  if (ntargets < 0 || (ntargets > 0 && (!target_times || !out)))
    return set_error(DFM_ERR_ARG, "bad target times");
  if (ntargets == 0) return DFM_OK;
  if (m != n) return set_error(DFM_ERR_ARG, "the transported momenta live on the geodesic's %ld control points (got %ld)", n, m);
  if (!P || !A || !mom0 || !cp0) return set_error(DFM_ERR_ARG, "NULL argument");
  Transport need(nullptr, prec, D, n, m);
  if (!work || work_bytes < need.total)
    return set_error(DFM_ERR_ARG, "workspace too small: need %zu bytes", need.total);
  Transport L((char *)work, prec, D, n, m);
  hipStream_t st = (hipStream_t)stream;
  const long nd = n * D;
  auto axpby = [&](double *o, const double *a, double ca, const double *b, double cb,
                   const double *sa = nullptr, const double *sb = nullptr) {
    hipLaunchKernelGGL(k_axpby, dim3(blocks(nd)), dim3(256), 0, st, o, a, ca, b, cb, nd, sa, sb);
  };
  auto dot = [&](const double *a, const double *b, int slot) {
    hipLaunchKernelGGL(k_dot, dim3(1), dim3(256), 0, st, a, b, nd, L.S + slot);
  };
  auto conv = [&](const double *x, long nx, const double *y, const double *w, long ny, double *o) {
    KParams p = blank();
    p.x = x; p.nx = nx; p.y = y; p.w = w; p.ny = ny; p.out0 = o;
    return pairsum(st, OP_CONV, prec, D, D, sigma, p, nx * D, 0, L.pw, L.pw_bytes);
  };
  auto grad1 = [&](const double *x, const double *alpha, const double *w, double *o) {
    KParams p = blank();
    p.x = x; p.ta = alpha; p.nx = n; p.y = x; p.w = w; p.ny = n; p.out0 = o;
    return pairsum(st, OP_GRAD1, prec, D, D, sigma, p, nd, 0, L.pw, L.pw_bytes);
  };
  const double h = T > 1 ? (tN - t0) / (T - 1) : 0.0;
  const double eps = h;
  auto to_index = [&](double t) {
    const double c = (t - t0) / h;
    const int i = (int)c;
    return (c - i) >= 0.5 ? i + 1 : i;
  };
  // early exits (Diffeos.cxx:296-310): no motion, nothing to transport, one time point
  hipLaunchKernelGGL(k_dot, dim3(1), dim3(256), 0, st, A, A, nd, L.S + S_PN2);
  hipLaunchKernelGGL(k_dot, dim3(1), dim3(256), 0, st, mom0, mom0, nd, L.S + S_PWV);
  double norms[2];
  hipError_t e = hipMemcpyAsync(norms, L.S + S_PN2, 2 * sizeof(double), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return set_error(DFM_ERR_HIP, "norms: %s", hipGetErrorString(e));
  const int initial = T > 1 ? to_index(initial_time) : 0;
  std::vector<int> targets(ntargets);
  for (int i = 0; i < ntargets; i++) targets[i] = T > 1 ? to_index(target_times[i]) : 0;
  const int final_ = targets.back();
  auto copy_initial = [&]() {
    for (int i = 0; i < ntargets; i++) axpby(out + (size_t)i * nd, mom0, 1.0, nullptr, 0.0);
    return launched("copy");
  };
  if (norms[0] < 1e-20 || norms[1] < 1e-20 || T <= 1 || final_ == initial) return copy_initial();
  if (initial < 0 || final_ >= T || targets[0] < initial)
    return set_error(DFM_ERR_ARG, "target times outside the geodesic's time grid");
  for (int i = 1; i < ntargets; i++)
    if (targets[i] < targets[i - 1])
      return set_error(DFM_ERR_ARG, "target times must be increasing");

  auto Pt = [&](int t) { return P + (size_t)t * nd; };
  auto At = [&](int t) { return A + (size_t)t * nd; };
  if ((rc = fill(st, L.inv.flag, 1, 0.0))) return rc;
  hipLaunchKernelGGL(k_set, dim3(1), dim3(64), 0, st, L.S, (int)S_ALPHA, 1.0);
  hipLaunchKernelGGL(k_set, dim3(1), dim3(64), 0, st, L.S, (int)S_BETA, 0.0);
  // the initial momenta's velocity at the geodesic's first control points
  if ((rc = conv(Pt(0), n, cp0, mom0, m, L.vel))) return rc;
  int oi = 0;
  auto emit = [&](int t, const double *tr) {
    while (oi < ntargets && targets[oi] == t) {
      axpby(out + (size_t)oi * nd, tr, 1.0, nullptr, 0.0);
      oi++;
    }
  };
  for (int t = initial; t < final_; t++) {
    if ((rc = conv(Pt(t), n, Pt(t), At(t), n, L.dpos))) return rc;
    if ((rc = convolve_inverse(st, D, D, sigma, Pt(t), n, L.vel, L.kinv, L.inv))) return rc;
    if (t == initial) {
      dot(L.kinv, L.dpos, S_IWV);
      if ((rc = conv(Pt(t), n, Pt(t), L.kinv, n, L.q))) return rc;
      dot(L.kinv, L.q, S_IW2);
    } else {
      if ((rc = conv(Pt(t), n, Pt(t), L.kinv, n, L.q))) return rc;
      dot(L.kinv, L.q, S_PN2);
      dot(At(t), L.q, S_PWV);
      hipLaunchKernelGGL(k_alpha_beta, dim3(1), dim3(64), 0, st, L.S);
    }
    axpby(L.tr, L.kinv, 0.0, At(t), 0.0, L.S + S_ALPHA, L.S + S_BETA);
    emit(t, L.tr);
    if ((rc = conv(Pt(t), n, Pt(t), L.tr, n, L.q))) return rc;
    dot(At(t), L.q, S_WV);
    dot(L.kinv, L.q, S_W2);
    dot(At(t), L.dpos, S_V2);
    // the +-eps perturbed geodesic, one Runge-Kutta 2 step
    axpby(L.mp, At(t), 1.0, L.kinv, eps);
    axpby(L.mn, At(t), 1.0, L.kinv, -eps);
    if ((rc = grad1(Pt(t), L.mp, L.mp, L.dmp)) || (rc = grad1(Pt(t), L.mn, L.mn, L.dmn))) return rc;
    // dmom = sum_j (mom_i.mom_j) grad_1 K: the reference's varlin contraction (Diffeos.cxx:508-523)
    for (int sgn = 0; sgn < 2; sgn++) {
      const double se = sgn == 0 ? eps : -eps;
      axpby(L.cp1, L.dpos, h / 2, L.vel, h / 2 * se);
      axpby(L.cp1, Pt(t), 1.0, L.cp1, 1.0);
      axpby(L.mom1, sgn == 0 ? L.mp : L.mn, 1.0, sgn == 0 ? L.dmp : L.dmn, -h / 2);
      if ((rc = conv(L.cp1, n, L.cp1, L.mom1, n, L.dpos2))) return rc;
      // cp2 / (2 eps) = (P_t + h dpos2) / (2 eps)
      axpby(L.cp1, Pt(t), 1.0 / (2 * eps), L.dpos2, h / (2 * eps));
      if (sgn == 0) axpby(L.plus, L.cp1, 1.0, nullptr, 0.0);
    }
    axpby(L.veln, L.plus, 1.0 / eps, L.cp1, -1.0 / eps);
    std::swap(L.vel, L.veln);
    if ((rc = launched("transport step"))) return rc;
  }
  // last step, without conservation (Diffeos.cxx:563-567)
  if ((rc = convolve_inverse(st, D, D, sigma, Pt(final_ - 1), n, L.vel, L.tr, L.inv))) return rc;
  emit(final_, L.tr);
  if ((rc = launched("transport"))) return rc;
  double bad = 0.0;
  if ((rc = read_flag(st, L.inv.flag, &bad))) return rc;
  return bad != 0.0 ? set_error(DFM_ERR_ARG, "kernel matrix is singular") : DFM_OK;
}

}  // extern "C"
