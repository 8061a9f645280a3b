/*
 * ORACLE — test infrastructure only.
 *
 * A plain-C, FP64 restatement of Deformetrica's CPU ExactKernel (the reference
 * implementation of the Gaussian kernel-sum engine) and of the generic
 * AbstractKernel::ConvolveSpecialHessian.  It exists to check the HIP library:
 * only the test-suite, the smoke check and the CPU-baseline leg of the benchmark
 * load it.  The product path never calls it.
 *
 * Kernel: K(x,y) = exp(-|x-y|^2 / h^2)   (no factor 2; ExactKernel.h:71-86)
 * Layout: every matrix is row-major AoS, x[i*d + k], the same layout the
 *         reference's GPU wrappers receive (CUDAExactKernel.cxx:46-48,
 *         ArmadilloMatrixWrapper vectorise_row_wise).
 * Loops : target-outer, source-inner, weight-component-inner, as in
 *         ExactKernel.cxx, so sums are formed in the reference's order.
 * exp   : libm exp (MathFunctions.h:16-20, USE_FAST_MATH off).
 *
 * Pinning (tests/test_oracle.py): closed forms (one source, sigma -> inf),
 * central finite differences of the oracle's own Convolve for the gradient and
 * Hessian families, symmetry, the reference's two SpecialHessian forms against
 * each other (TestKernelPrecisionCUDA.cxx:97-106), the mt19937 known answer, and
 * the reference-held parallel-transport vectors.
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* AbstractKernel::SetKernelWidth clamps h >= 1e-10 (AbstractKernel.h:85-91). */
static double width_sq(double h) {
  if (h < 1e-10) h = 1e-10;
  return h * h;
}

/* ExactKernel::EvaluateKernel(X, Y, i, j) — ExactKernel.h:78-86. */
static double eval_kernel(const double *x, const double *y, int d, double s2) {
  double dist2 = 0.0;
  for (int k = 0; k < d; k++) {
    double diff = x[k] - y[k];
    dist2 += diff * diff;
  }
  return exp(-dist2 / s2);
}

/* ExactKernel::EvaluateKernelGradient(X, Y, i, j) — ExactKernel.h:98-110.
 * g = (x - y) * (-2 exp(-r^2/h^2) / h^2). */
static void eval_kernel_gradient(const double *x, const double *y, int d, double s2,
                                 double *g) {
  double dist2 = 0.0;
  for (int k = 0; k < d; k++) {
    double diff = x[k] - y[k];
    g[k] = diff;
    dist2 += diff * diff;
  }
  double f = -2.0 * exp(-dist2 / s2) / s2;
  for (int k = 0; k < d; k++) g[k] *= f;
}

/* ExactKernel::EvaluateKernelGradient(x, y) — ExactKernel.h:88-95. */
static void eval_kernel_gradient_v(const double *x, const double *y, int d, double s2,
                                   double *g) {
  double dist2 = 0.0;
  for (int k = 0; k < d; k++) {
    double t = x[k] - y[k];
    dist2 += t * t;
  }
  double kk = exp(-dist2 / s2);
  double f = -2.0 * kk / s2;
  for (int k = 0; k < d; k++) g[k] = (x[k] - y[k]) * f;
}

/* ExactKernel::EvaluateKernelHessian(X, Y, i, j) — ExactKernel.h:136-159.
 * H = k * (4 (x-y)(x-y)^T / h^4 - 2/h^2 I); H is d*d row-major. */
static void eval_kernel_hessian(const double *x, const double *y, int d, double s2,
                                double *H) {
  double xmy[8];
  double dist2 = 0.0;
  for (int k = 0; k < d; k++) {
    double diff = x[k] - y[k];
    xmy[k] = diff;
    dist2 += diff * diff;
  }
  for (int b = 0; b < d; b++)
    for (int a = 0; a < d; a++) H[a * d + b] = 4.0 * xmy[a] * xmy[b] / (s2 * s2);
  for (int a = 0; a < d; a++) H[a * d + a] -= 2.0 / s2;
  double kij = exp(-dist2 / s2);
  for (int a = 0; a < d * d; a++) H[a] *= kij;
}

/* ExactKernel::EvaluateKernelHessian(x, y) — ExactKernel.h:113-133. */
static void eval_kernel_hessian_v(const double *x, const double *y, int d, double s2,
                                  double *H) {
  double xy[8];
  double distsq = 0.0;
  for (int k = 0; k < d; k++) {
    double t = x[k] - y[k];
    xy[k] = t;
    distsq += t * t;
  }
  double kk = exp(-distsq / s2);
  double f = 4.0 * kk / (s2 * s2);
  for (int a = 0; a < d; a++)
    for (int b = 0; b < d; b++) H[a * d + b] = xy[a] * xy[b] * f;
  for (int a = 0; a < d; a++) H[a * d + a] -= 2.0 * kk / s2;
}

/* ExactKernel::Convolve — ExactKernel.cxx:51-74.  V (nx x dv). */
void orc_convolve(int d, int dv, double h, const double *X, long nx, const double *Y,
                  const double *W, long ny, double *V) {
  double s2 = width_sq(h);
  memset(V, 0, sizeof(double) * (size_t)nx * dv);
  for (long i = 0; i < nx; i++)
    for (long j = 0; j < ny; j++) {
      double kij = eval_kernel(X + i * d, Y + j * d, d, s2);
      for (int k = 0; k < dv; k++) V[i * dv + k] += W[j * dv + k] * kij;
    }
}

/* ExactKernel::ConvolveGradient(X) ("varlin" list) — ExactKernel.cxx:166-202.
 * G (nx x dv x d): G[i][k][l] = sum_j W[j][k] * dK/dx_l(x_i, y_j). */
void orc_varlin(int d, int dv, double h, const double *X, long nx, const double *Y,
                const double *W, long ny, double *G) {
  double s2 = width_sq(h);
  double g[8];
  memset(G, 0, sizeof(double) * (size_t)nx * dv * d);
  for (long i = 0; i < nx; i++) {
    double *Gi = G + i * dv * d;
    for (long j = 0; j < ny; j++) {
      eval_kernel_gradient(X + i * d, Y + j * d, d, s2, g);
      for (int k = 0; k < dv; k++) {
        double wjk = W[j * dv + k];
        for (int l = 0; l < d; l++) Gi[k * d + l] += g[l] * wjk;
      }
    }
  }
}

/* ExactKernel::ConvolveGradient(X, alpha) ("Grad1") — ExactKernel.cxx:264-275:
 * builds the varlin list, then result(i,:) = G_i^T alpha_i.  out (nx x d). */
void orc_grad1(int d, int dv, double h, const double *X, const double *alpha, long nx,
               const double *Y, const double *W, long ny, double *out) {
  double s2 = width_sq(h);
  double g[8], Gi[64];
  for (long i = 0; i < nx; i++) {
    memset(Gi, 0, sizeof(Gi));
    for (long j = 0; j < ny; j++) {
      eval_kernel_gradient(X + i * d, Y + j * d, d, s2, g);
      for (int k = 0; k < dv; k++) {
        double wjk = W[j * dv + k];
        for (int l = 0; l < d; l++) Gi[k * d + l] += g[l] * wjk;
      }
    }
    for (int l = 0; l < d; l++) {
      double acc = 0.0;
      for (int k = 0; k < dv; k++) acc += Gi[k * d + l] * alpha[i * dv + k];
      out[i * d + l] = acc;
    }
  }
}

/* ExactKernel::ConvolveGradient(X, dim) — ExactKernel.cxx:294-323.
 * out (nx x dv): out[i][k] = sum_j dK/dx_dim(x_i, y_j) W[j][k]. Returns -1 on bad dim. */
int orc_grad_dim(int d, int dv, double h, const double *X, long nx, const double *Y,
                 const double *W, long ny, int dim, double *out) {
  if (dim < 0 || dim >= d) return -1;
  double s2 = width_sq(h);
  double g[8];
  memset(out, 0, sizeof(double) * (size_t)nx * dv);
  for (long i = 0; i < nx; i++)
    for (long j = 0; j < ny; j++) {
      eval_kernel_gradient_v(X + i * d, Y + j * d, d, s2, g);
      for (int k = 0; k < dv; k++) out[i * dv + k] = out[i * dv + k] + g[dim] * W[j * dv + k];
    }
  return 0;
}

/* ExactKernel::ConvolveHessian(X) — ExactKernel.cxx:365-397.
 * out (nx x dv x d x d): out[i][k] = sum_j H(x_i, y_j) W[j][k]. */
void orc_hessian_list(int d, int dv, double h, const double *X, long nx, const double *Y,
                      const double *W, long ny, double *out) {
  double s2 = width_sq(h);
  double H[64];
  memset(out, 0, sizeof(double) * (size_t)nx * dv * d * d);
  for (long i = 0; i < nx; i++) {
    double *Hi = out + i * dv * d * d;
    for (long j = 0; j < ny; j++) {
      eval_kernel_hessian(X + i * d, Y + j * d, d, s2, H);
      for (int k = 0; k < dv; k++) {
        double wjk = W[j * dv + k];
        for (int a = 0; a < d * d; a++) Hi[k * d * d + a] = Hi[k * d * d + a] + H[a] * wjk;
      }
    }
  }
}

/* ExactKernel::ConvolveHessian(X, row, col) — ExactKernel.cxx:399-428. out (nx x dv). */
int orc_hessian_rc(int d, int dv, double h, const double *X, long nx, const double *Y,
                   const double *W, long ny, int row, int col, double *out) {
  if (row < 0 || col < 0 || row >= d || col >= d) return -1;
  double s2 = width_sq(h);
  double H[64];
  memset(out, 0, sizeof(double) * (size_t)nx * dv);
  for (long i = 0; i < nx; i++)
    for (long j = 0; j < ny; j++) {
      eval_kernel_hessian_v(X + i * d, Y + j * d, d, s2, H);
      for (int k = 0; k < dv; k++)
        out[i * dv + k] = out[i * dv + k] + W[j * dv + k] * H[row * d + col];
    }
  return 0;
}

/* AbstractKernel::ConvolveSpecialHessian(xi), generic form — AbstractKernel.cxx:22-88.
 * Sources Y (n x d), weights beta (n x d), xi (n x d) -> out (n x d).
 * gamma_i = sum_j (beta_i . beta_j) H(y_i, y_j) (xi_j - xi_i). */
int orc_special_hessian(int d, double h, const double *Y, const double *beta,
                        const double *xi, long n, double *out) {
  size_t nd = (size_t)n * d;
  double *hess = malloc(sizeof(double) * nd * d * d);
  double *Hacc = calloc(nd * d, sizeof(double)); /* H[i] (d x d) row-major */
  double *Wd = malloc(sizeof(double) * nd);
  double *Wtmp = malloc(sizeof(double) * nd * 2);
  double *Hrr = malloc(sizeof(double) * nd);
  double *Hrq = malloc(sizeof(double) * nd * 2);
  if (!hess || !Hacc || !Wd || !Wtmp || !Hrr || !Hrq) {
    free(hess); free(Hacc); free(Wd); free(Wtmp); free(Hrr); free(Hrq);
    return -2;
  }
  /* dXi4 first term: -(sum_dim HessA[i][dim] beta(i,dim)) xi_i. */
  orc_hessian_list(d, d, h, Y, n, Y, beta, n, hess);
  for (long i = 0; i < n; i++) {
    double aux[64];
    memset(aux, 0, sizeof(aux));
    for (int dim = 0; dim < d; dim++)
      for (int a = 0; a < d * d; a++)
        aux[a] -= hess[((size_t)i * d + dim) * d * d + a] * beta[i * d + dim];
    for (int a = 0; a < d; a++) {
      double acc = 0.0;
      for (int b = 0; b < d; b++) acc += aux[a * d + b] * xi[i * d + b];
      out[i * d + a] = acc;
    }
  }
  /* dXi4 second term. */
  for (int r = 0; r < d; r++) {
    for (long i = 0; i < n; i++)
      for (int c = 0; c < d; c++) Wd[i * d + c] = beta[i * d + c] * xi[i * d + r];
    orc_hessian_rc(d, d, h, Y, n, Y, Wd, n, r, r, Hrr);
    for (long i = 0; i < n; i++)
      for (int q = 0; q < d; q++) Hacc[(size_t)i * d * d + r * d + q] += Hrr[i * d + q];
    for (int q = r + 1; q < d; q++) {
      for (long i = 0; i < n; i++)
        for (int c = 0; c < d; c++) {
          Wtmp[i * 2 * d + c] = Wd[i * d + c];
          Wtmp[i * 2 * d + c + d] = beta[i * d + c] * xi[i * d + q];
        }
      orc_hessian_rc(d, 2 * d, h, Y, n, Y, Wtmp, n, r, q, Hrq);
      for (long i = 0; i < n; i++)
        for (int c = 0; c < d; c++) {
          Hacc[(size_t)i * d * d + q * d + c] += Hrq[i * 2 * d + c];
          Hacc[(size_t)i * d * d + r * d + c] += Hrq[i * 2 * d + c + d];
        }
    }
  }
  for (long i = 0; i < n; i++)
    for (int a = 0; a < d; a++) {
      double acc = 0.0;
      for (int b = 0; b < d; b++) acc += Hacc[(size_t)i * d * d + a * d + b] * beta[i * d + b];
      out[i * d + a] += acc;
    }
  free(hess); free(Hacc); free(Wd); free(Wtmp); free(Hrr); free(Hrq);
  return 0;
}

/* The reference GPU form of the same operator (CUDAExactKernel.cxx:196-226 with the
 * GradDiff functor, ScalarRadialKernel.h:145-164 and GaussFunction.h:59-63), scaled by
 * -0.5 as the plugin does:
 * gamma_i = sum_j (b_i.b_j) K_ij [ (2/h^2)(xi_i - xi_j) - (4/h^4)((y_i-y_j).(xi_i-xi_j))(y_i-y_j) ]. */
void orc_special_hessian_graddiff(int d, double h, const double *Y, const double *beta,
                                  const double *xi, long n, double *out) {
  double s2 = width_sq(h);
  for (long i = 0; i < n; i++) {
    double acc[8] = {0};
    for (long j = 0; j < n; j++) {
      double r2 = 0.0, dot = 0.0, bb = 0.0, xmy[8], emf[8];
      for (int k = 0; k < d; k++) {
        xmy[k] = Y[i * d + k] - Y[j * d + k];
        emf[k] = xi[i * d + k] - xi[j * d + k];
        r2 += xmy[k] * xmy[k];
        dot += xmy[k] * emf[k];
        bb += beta[i * d + k] * beta[j * d + k];
      }
      double kk = exp(-r2 / s2);
      double d1 = 4.0 * bb * (-kk / s2);
      double d2 = 8.0 * bb * dot * (kk / (s2 * s2));
      for (int k = 0; k < d; k++) acc[k] += d1 * emf[k] + d2 * xmy[k];
    }
    for (int k = 0; k < d; k++) out[i * d + k] = -0.5 * acc[k];
  }
}

/* ExactKernel::ComputeKernelMatrix — ExactKernel.cxx:131-143. out (n x n). */
void orc_kernel_matrix(int d, double h, const double *Y, long n, double *out) {
  double s2 = width_sq(h);
  for (long i = 0; i < n; i++)
    for (long j = 0; j < n; j++) out[i * n + j] = eval_kernel(Y + i * d, Y + j * d, d, s2);
}

/* ExactKernel::ConvolveInverse — ExactKernel.cxx:154-164: solve K V = X with K the kernel
 * matrix of the sources.  X, V: (n x dv).  Returns -3 if K is exactly singular. */
int orc_convolve_inverse(int d, int dv, double h, const double *Y, long n, const double *X,
                         double *V) {
  /* ExactKernel::ConvolveInverse (ExactKernel.cxx:156-164): V = solve(K(Y, Y), X), which
   * Armadillo hands to LAPACK's dgesv for a square system: LU factorisation with partial
   * (row) pivoting, the pivot being the first entry of largest magnitude in the column,
   * then forward and back substitution. */
  double *A = malloc(sizeof(double) * (size_t)n * n);
  long *piv = malloc(sizeof(long) * (size_t)(n > 0 ? n : 1));
  if (!A || !piv) { free(A); free(piv); return -2; }
  orc_kernel_matrix(d, h, Y, n, A);
  for (long k = 0; k < n; k++) {
    long p = k;
    double best = fabs(A[k * n + k]);
    for (long i = k + 1; i < n; i++)
      if (fabs(A[i * n + k]) > best) { best = fabs(A[i * n + k]); p = i; }
    piv[k] = p;
    if (best == 0.0) { free(A); free(piv); return -3; }
    if (p != k)
      for (long c = 0; c < n; c++) { double t = A[k * n + c]; A[k * n + c] = A[p * n + c]; A[p * n + c] = t; }
    const double inv = 1.0 / A[k * n + k];
    for (long i = k + 1; i < n; i++) {
      const double l = A[i * n + k] * inv;
      A[i * n + k] = l;
      for (long c = k + 1; c < n; c++) A[i * n + c] -= l * A[k * n + c];
    }
  }
  for (long i = 0; i < n; i++)
    for (int c = 0; c < dv; c++) V[i * dv + c] = X[i * dv + c];
  for (long k = 0; k < n; k++)          /* the row interchanges, in order */
    if (piv[k] != k)
      for (int c = 0; c < dv; c++) { double t = V[k * dv + c]; V[k * dv + c] = V[piv[k] * dv + c]; V[piv[k] * dv + c] = t; }
  for (int c = 0; c < dv; c++) {
    for (long i = 0; i < n; i++) {      /* L z = P x (unit lower) */
      double t = V[i * dv + c];
      for (long k = 0; k < i; k++) t -= A[i * n + k] * V[k * dv + c];
      V[i * dv + c] = t;
    }
    for (long i = n - 1; i >= 0; i--) { /* U v = z */
      double t = V[i * dv + c];
      for (long k = i + 1; k < n; k++) t -= A[i * n + k] * V[k * dv + c];
      V[i * dv + c] = t / A[i * n + i];
    }
  }
  free(A);
  free(piv);
  return 0;
}

/* ---------------------------------------------------------------------------------- */
/* Multi-threaded CPU baseline: the same restated loops (Convolve then
 * ConvolveGradient(X, alpha)) over contiguous target-row blocks on a worker pool, one
 * worker per host core, as the reference's lib/ThreadPool/ThreadPool.h drives host
 * work.  No per-pair heap allocation (the reference allocates a VectorType per pair in
 * EvaluateKernelGradient), so this is a faster-than-reference baseline. */
typedef struct {
  int d, dv;
  double h;
  const double *X, *alpha, *Y, *W;
  long nx, ny;
  double *conv, *grad;
  long block;
  long next; /* guarded by mu */
  pthread_mutex_t mu;
} pool_job;

static void *pool_worker(void *arg) {
  pool_job *J = (pool_job *)arg;
  for (;;) {
    pthread_mutex_lock(&J->mu);
    long start = J->next;
    J->next += J->block;
    pthread_mutex_unlock(&J->mu);
    if (start >= J->nx) break;
    long cnt = J->nx - start < J->block ? J->nx - start : J->block;
    if (J->conv)
      orc_convolve(J->d, J->dv, J->h, J->X + start * J->d, cnt, J->Y, J->W, J->ny,
                   J->conv + start * J->dv);
    if (J->grad)
      orc_grad1(J->d, J->dv, J->h, J->X + start * J->d, J->alpha + start * J->dv, cnt,
                J->Y, J->W, J->ny, J->grad + start * J->d);
  }
  return NULL;
}

/* Convolve (if conv != NULL) and ConvolveGradient(X, alpha) (if grad != NULL) on
 * nthreads workers.  Returns 0, or -2 if a thread could not be started. */
int orc_convolve_grad1_threaded(int nthreads, int d, int dv, double h, const double *X,
                                const double *alpha, long nx, const double *Y,
                                const double *W, long ny, double *conv, double *grad) {
  if (nthreads < 1) nthreads = 1;
  /* blocks of up to 8 rows, at least four blocks per worker so that small row counts
   * still keep every worker busy */
  long block = nx / (4L * nthreads);
  if (block < 1) block = 1;
  if (block > 8) block = 8;
  pool_job J = {d, dv, h, X, alpha, Y, W, nx, ny, conv, grad, block, 0,
                PTHREAD_MUTEX_INITIALIZER};
  pthread_t *th = malloc(sizeof(pthread_t) * nthreads);
  if (!th) return -2;
  int started = 0, rc = 0;
  for (int t = 0; t < nthreads; t++) {
    if (pthread_create(&th[t], NULL, pool_worker, &J) != 0) { rc = -2; break; }
    started++;
  }
  if (started == 0) pool_worker(&J);
  for (int t = 0; t < started; t++) pthread_join(th[t], NULL);
  free(th);
  return rc;
}

/* ---------------------------------------------------------------------------------- */
/* MT19937 (Matsumoto & Nishimura 1998), seeded as std::mt19937(seed) is: x_0 = seed,
 * x_i = 1812433253 (x_{i-1} ^ (x_{i-1} >> 30)) + i.  Used to regenerate the inputs of
 * the reference's kernel precision tests (AbstractTestKernelPrecision.cxx:21-29). */
typedef struct {
  uint32_t mt[624];
  int idx;
} orc_mt19937;

void orc_mt_seed(orc_mt19937 *g, uint32_t seed) {
  g->mt[0] = seed;
  for (int i = 1; i < 624; i++)
    g->mt[i] = 1812433253u * (g->mt[i - 1] ^ (g->mt[i - 1] >> 30)) + (uint32_t)i;
  g->idx = 624;
}

uint32_t orc_mt_next(orc_mt19937 *g) {
  if (g->idx >= 624) {
    for (int i = 0; i < 624; i++) {
      uint32_t y = (g->mt[i] & 0x80000000u) | (g->mt[(i + 1) % 624] & 0x7fffffffu);
      uint32_t v = g->mt[(i + 397) % 624] ^ (y >> 1);
      if (y & 1u) v ^= 0x9908b0dfu;
      g->mt[i] = v;
    }
    g->idx = 0;
  }
  uint32_t y = g->mt[g->idx++];
  y ^= y >> 11;
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= y >> 18;
  return y;
}

/* n successive outputs of std::mt19937(seed). */
void orc_mt_stream(uint32_t seed, long n, uint32_t *out) {
  orc_mt19937 g;
  orc_mt_seed(&g, seed);
  for (long i = 0; i < n; i++) out[i] = orc_mt_next(&g);
}

/* AbstractTestKernelPrecision::generate_random_matrix: the generator is re-created with
 * seed 1 on every call and fills row by row, X(i,j) = exp(-gen() / gen.max()). */
void orc_reference_fixture(long rows, int cols, double *X) {
  orc_mt19937 g;
  orc_mt_seed(&g, 1u);
  for (long i = 0; i < rows; i++)
    for (int j = 0; j < cols; j++)
      X[i * cols + j] = exp(-(double)orc_mt_next(&g) / (double)4294967295u);
}
