/*
 * dfm_hip.h — C ABI of the MI355X (gfx950) Gaussian kernel-sum engine.
 *
 * K(x, y) = exp(-|x - y|^2 / sigma^2)   (Deformetrica ExactKernel.h:71-86; sigma^2 in the
 * denominator, no factor 2).  Every matrix crossing this boundary is FP64, row-major
 * array-of-structures (x[i*dimpoint + k]) — the layout the reference's CUDA plugin hands
 * its library after vectorise_row_wise (CUDAExactKernel.cxx:46-48).
 *
 * Two families of entry points:
 *   dfm_gauss_*      host pointers, synchronous: drop-in replacements for the
 *                    reference's GaussGpu*Conv1D<double, DP, DV> host wrappers
 *                    (lib/cuda_convolutions/GpuConv1D.h), with the point and weight
 *                    dimensions as run-time arguments and a status code instead of the
 *                    reference's unconditional 0.
 *   dfm_*            device pointers, stream-ordered, caller-provided workspace (size
 *                    from the matching dfm_*_workspace query); they never allocate.
 *
 * Precision (int precision):
 *   DFM_FP64   plain FP64 arithmetic and FP64 exp — the reference's arithmetic.
 *   DFM_FP32C  FP32 exp and multiply-add over LDS-staged source tiles, with targets
 *              spatially sorted and re-centred in FP64 per wavefront, FP32 partial sums
 *              of 64 sources folded into FP64 accumulators (compensated, cascaded
 *              summation).  A wavefront whose targets span more than 4 scaled kernel
 *              widths forms its coordinate differences exactly (FP32 hi + lo parts).
 *   DFM_FP32X  as DFM_FP32C with exact coordinate differences in every wavefront.
 *   DFM_BF16MFMA  Convolve, ConvolveGradient(X, alpha), the fused sum and the varifold
 *              sums of single calls from 8,192 targets and 2,048 sources on, and of ragged
 *              batches from 2,048 sources per segment: the distances' cross term -2 x.y
 *              and |y|^2 on the matrix cores (v_mfma_f32_16x16x32_bf16 over three-part
 *              BF16 splits of the FP32 centred coordinates, FP32 accumulation), the target
 *              norm applied exactly per target, exp2 and accumulation as DFM_FP32C.
 *              Groups of 256 targets spanning more than 8 scaled kernel units (5 in
 *              ragged batches) run on one centre per wavefront of 64 targets, or the
 *              DFM_FP32C wavefront code, in the same launch; every other operation and
 *              smaller calls run as DFM_FP32C.
 *
 * Supported shapes: dimpoint in {2, 3}; dimvect in {1, ..., 6} (the reference supports
 * dimvect in {D, 2D, D(D+1)/2}, CUDAExactKernel.cxx:56-82).  Counts are non-negative and
 * below 2^31.  Status codes below; dfm_last_error() returns the calling thread's last
 * message.  Entry points are thread-safe; calls that share a workspace buffer must be
 * ordered on one stream.
 */
#ifndef DFM_HIP_H
#define DFM_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum dfm_status {
  DFM_OK = 0,
  DFM_ERR_ARG = 1,      /* null pointer, negative count, undersized workspace */
  DFM_ERR_MISMATCH = 2, /* "Sources and weights count mismatch" (ExactKernel.cxx:58-59) */
  DFM_ERR_DIM = 3,      /* unsupported dimpoint / dimvect (CUDAExactKernel.cxx:80-82) */
  DFM_ERR_INDEX = 4,    /* "dimension index out of bounds" (CUDAExactKernel.cxx:242-243) */
  DFM_ERR_HIP = 5,      /* HIP runtime error (message in dfm_last_error) */
  DFM_ERR_NOMEM = 6     /* device allocation failed (host-pointer wrappers only) */
};

enum dfm_precision { DFM_FP64 = 0, DFM_FP32C = 1, DFM_FP32X = 2, DFM_BF16MFMA = 3 };

/* ABI version of this header (bumped on any signature or struct-layout change); callers
 * compare dfm_abi_version() of the loaded library with DFM_ABI_VERSION of the header they
 * were built against.
 *   1  first release
 *   2  dfm_atlas_desc gained dimension and cell_corners (appended); new entry points for
 *      polylines, volume meshes, point clouds, landmarks, ConvolveHessian and the
 *      adjoint's jump conditions; DFM_OP_HESSIAN
 *   3  dfm_memcpy_async; DFM_BF16MFMA also serves ragged batches
 *   4  DFM_OP_VARIFOLD and DFM_OP_VARIFOLD_CONV in the workspace and flops queries */
#define DFM_ABI_VERSION 4
int dfm_abi_version(void);
/* Last error message of the calling thread ("" when none). */
const char *dfm_last_error(void);

/* Device memory and streams for callers that keep data resident between calls (no
 * reference counterpart: the reference's wrappers allocate and copy inside every call,
 * GpuConv1D.cu:98-130).  They use the same HIP runtime as the kernels, so a caller that
 * has no HIP runtime of its own (Python through ctypes, the C++ plugin) needs nothing
 * else.  dfm_memcpy is ordered on `stream` and returns after the copy completes.
 * kind: 0 host->device, 1 device->host, 2 device->device (a plain copy kernel, k_copy,
 * for 16-byte-aligned buffers).  dfm_memcpy_async is the same copy queued on `stream`
 * without waiting: a host buffer must stay valid and unchanged until the stream has passed
 * the copy (for example until the caller's next dfm_stream_synchronize). */
int dfm_device_count(int *count);
int dfm_malloc(void **ptr, size_t bytes);
int dfm_free(void *ptr);
int dfm_memcpy(void *dst, const void *src, size_t bytes, int kind, void *stream);
int dfm_memcpy_async(void *dst, const void *src, size_t bytes, int kind, void *stream);
int dfm_stream_create(void **stream);
int dfm_stream_destroy(void *stream);
int dfm_stream_synchronize(void *stream);

/* Kernel timing for measurement (no reference counterpart: the reference times nothing
 * on the device).  While enabled on the calling thread, every pair-sum kernel launch
 * (with its split reduction, when there is one, or the single launch of a small
 * dfm_shoot) is bracketed by HIP events on the stream it runs on.  dfm_profile_read waits for the recorded events, returns their summed
 * duration and launch count, and forgets them. */
int dfm_profile_enable(int on);
int dfm_profile_read(double *total_ms, long *launches);
/* The algorithmic work of the launches timed since the last call (while profiling is
 * enabled on this thread): (target, source) pairs summed over every launch and segment,
 * and those pairs times dfm_flops_per_pair of each launch's op; then resets both. */
int dfm_profile_work(double *pairs, double *flops);
/* Algorithmic flops of one (target, source) pair of an op (DFM_OP_*), counted on the
 * reference's per-pair arithmetic with exp as one: 3 dimpoint + 1 for the kernel value,
 * then 2 dimvect (Convolve), 2 dimvect + 2 dimpoint + 1 (ConvolveGradient(X, alpha)),
 * both for the fused sum (29 at dimpoint = dimvect = 3), one more product for the
 * varifold's fused sum (30) and 2 dimvect for its Convolve part's weight K (u.u) (22). */
double dfm_flops_per_pair(int op, int dimpoint, int dimvect);

/* Autotuner of the pair-sum work split (no reference counterpart: the reference launches
 * one thread per target in blocks of 192, GpuConv1D.cu:98-132).  Per (op, precision,
 * dimpoint, dimvect, batched or not, log2 of the target and source counts, and the grid's
 * class of target workgroups: < 64, < 1024, < 4096, >= 4096) it chooses
 * the targets per lane (1, 2 or 4), the wavefront count the source splits aim for (4096,
 * 16384 or 65536), the fewest source tiles per split (4 or 1) and, for the FP32C fused
 * Convolve + Grad1 in 3D, whether the per-pair accumulation runs on the matrix cores.
 *   mode 0  heuristics only (default; results deterministic across calls)
 *   mode 1  explore: the first calls of each key run every candidate twice, timed with HIP
 *           events on the call's stream (read back at a later call of that key), then the
 *           fastest is kept.  Plans differ in summation order, so results differ across
 *           candidates by rounding.  Not for use while a stream is being captured.
 *           Threads may explore concurrently; dfm_autotune_reset may run at any time (a
 *           timed launch in flight keeps its events alive and its timing is dropped).
 *   mode 2  use the tuned choices (default where a key was not tuned); capture-safe.
 * Workspace queries always return the maximum over every candidate, so a buffer sized
 * in any mode fits every plan.  dfm_autotune_report writes one line per key (truncated to
 * len - 1 characters) and returns the full length. */
int dfm_autotune(int mode);
int dfm_autotune_reset(void);
int dfm_autotune_report(char *buf, size_t len);

/* ------------------------------------------------------------------------------------ */
/* Host-pointer wrappers (synchronous: copy in, launch, copy back on the calling thread's
 * per-thread stream, one synchronisation).  Their device buffers come from a per-device
 * pool that keeps them resident for the life of the process: concurrent calls from
 * several threads lease separate buffer sets, and a buffer is reallocated only when a call
 * needs more than it holds (DFM_ERR_NOMEM if that fails), so steady call sequences neither
 * allocate nor free.                                                                    */

/* gamma (nx x dv) = sum_j K(x_i, y_j) beta_j.
 * Replaces GaussGpuEvalConv1D<TYPE, DIMPOINT, DIMVECT> (GpuConv1D.h:45-47,
 * GpuConv1D.cu:148-155), called from CUDAExactKernel::Convolve (CUDAExactKernel.cxx:41-90). */
int dfm_gauss_conv(int dimpoint, int dimvect, double sigma, const double *x, const double *y,
                   const double *beta, double *gamma, int nx, int ny, int precision);

/* gamma (nx x dp) = sum_j (alpha_i . beta_j) grad_x K(x_i, y_j).
 * Replaces GaussGpuGrad1Conv1D (GpuConv1D.h:78-80), called from
 * CUDAExactKernel::ConvolveGradient(X, alpha) (CUDAExactKernel.cxx:165-194). */
int dfm_gauss_grad1conv(int dimpoint, int dimvect, double sigma, const double *alpha,
                        const double *x, const double *y, const double *beta, double *gamma,
                        int nx, int ny, int precision);

/* gamma (nx x dv) = sum_j beta_j d/dx_dim K(x_i, y_j).
 * Replaces GaussGpuGradConv1D (GpuConv1D.h:113-114), called from
 * CUDAExactKernel::ConvolveGradient(X, dim) (CUDAExactKernel.cxx:228-259).  The reference
 * form is the self-interaction y = x; this one takes the sources explicitly, as
 * ExactKernel::ConvolveGradient(X, dim) does (ExactKernel.cxx:294-323). */
int dfm_gauss_gradconv(int dimpoint, int dimvect, double sigma, const double *x, const double *y,
                       const double *beta, unsigned dim, double *gamma, int nx, int ny,
                       int precision);

/* gamma (nx x dv x dp), gamma[i][k][l] = sum_j beta_j[k] d/dx_l K(x_i, y_j).
 * Replaces GaussGpuGradConv_varlin_1D (GpuConv1D.h:147-148), called from
 * CUDAExactKernel::ConvolveGradient(X) (CUDAExactKernel.cxx:261-320); output index
 * i*dv*dp + k*dp + l as the reference reads it (CUDAExactKernel.cxx:312). */
int dfm_gauss_gradconv_varlin(int dimpoint, int dimvect, double sigma, const double *x,
                              const double *y, const double *beta, double *gamma, int nx, int ny,
                              int precision);

/* gamma (n x dp) = sum_j (beta_i . beta_j) H_x K(x_i, x_j) (eta_j - eta_i)
 *               = sum_j (b_i.b_j) K_ij [(2/s^2)(eta_i-eta_j) - (4/s^4)((x_i-x_j).(eta_i-eta_j))(x_i-x_j)],
 * i.e. the reference's GaussGpuGradDiffConv1D result already multiplied by -0.5, which is
 * what CUDAExactKernel::ConvolveSpecialHessian returns (CUDAExactKernel.cxx:196-226;
 * GpuConv1D.h:180-181) and what AbstractKernel::ConvolveSpecialHessian computes
 * (AbstractKernel.cxx:22-88).  Self-interaction: sources = x, weights = beta (n x dv). */
int dfm_gauss_graddiffconv(int dimpoint, int dimvect, double sigma, const double *x,
                           const double *beta, const double *eta, double *gamma, int nx,
                           int precision);

/* gamma (nx x dv x dp x dp), gamma[i][k] = sum_j beta_j[k] H_x K(x_i, y_j) with
 * H_x K = K (4 (x - y)(x - y)^T / s^4 - 2 I / s^2) (ExactKernel.h:113-133); index
 * ((i*dv + k)*dp + r)*dp + c, both triangles written.  Replaces
 * ExactKernel::ConvolveHessian(X) (ExactKernel.cxx:365-397), which CUDAExactKernel
 * inherits on the CPU; the (X, row, col) and (X, k, dp, dq) overloads (:399-468) are
 * slices of it. */
int dfm_gauss_hessianconv(int dimpoint, int dimvect, double sigma, const double *x,
                          const double *y, const double *beta, double *gamma, int nx, int ny,
                          int precision);

/* Fused Convolve + ConvolveGradient(X, alpha) over one pass of the pairs: conv (nx x dv)
 * and grad (nx x dp); either output may be NULL.  No single reference counterpart: it
 * replaces the back-to-back GaussGpuEvalConv1D + GaussGpuGrad1Conv1D calls that the
 * reference's callers make on the same sources (e.g. DeterministicAtlas.cxx:246-281). */
int dfm_gauss_conv_grad1(int dimpoint, int dimvect, double sigma, const double *x,
                         const double *alpha, const double *y, const double *beta, double *conv,
                         double *grad, int nx, int ny, int precision);

/* ------------------------------------------------------------------------------------ */
/* Device-pointer, stream-ordered entry points.  `stream` is a hipStream_t (NULL = the  */
/* default stream); `work` is device memory of at least the queried size (may be NULL  */
/* when the query returns 0).  Pointers are device pointers; outputs are overwritten.  */

/* op codes for the workspace query */
enum dfm_op {
  DFM_OP_CONV = 0,
  DFM_OP_GRAD1 = 1,
  DFM_OP_CONV_GRAD1 = 2,
  DFM_OP_VARLIN = 3,
  DFM_OP_GRADDIM = 4,
  DFM_OP_GRADDIFF = 5,
  DFM_OP_SHOOT_VJP = 6,
  DFM_OP_HESSIAN = 7,
  /* the varifold data term's sums over unit-root normals u (dimvect = dimpoint):
   * V_i = sum_j K_ij (u_i . u_j) u_j and its gradient sum_j grad_1 K_ij (u_i . u_j)^2
   * (no entry point of its own: the varifold match and gradient entry points run it) */
  DFM_OP_VARIFOLD = 8,
  /* its Convolve part V_i alone (the varifold match) */
  DFM_OP_VARIFOLD_CONV = 9
};

/* Bytes of workspace the device entry point for `op` needs at these sizes. */
size_t dfm_workspace_size(int op, int dimpoint, int dimvect, long nx, long ny, int precision);

int dfm_conv(void *stream, int dimpoint, int dimvect, double sigma, int precision,
             const double *x, long nx, const double *y, const double *beta, long ny,
             double *gamma, void *work, size_t work_bytes);

int dfm_grad1(void *stream, int dimpoint, int dimvect, double sigma, int precision,
              const double *x, const double *alpha, long nx, const double *y,
              const double *beta, long ny, double *gamma, void *work, size_t work_bytes);

int dfm_conv_grad1(void *stream, int dimpoint, int dimvect, double sigma, int precision,
                   const double *x, const double *alpha, long nx, const double *y,
                   const double *beta, long ny, double *conv, double *grad, void *work,
                   size_t work_bytes);

/* dfm_conv_grad1 without a workspace argument, for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc); no counterpart in the reference.  Its workspace is a static device array of
 * DFM_STATIC_SCRATCH_BYTES per GPU, held until the process ends; the call itself
 * allocates nothing.  It runs the plan dfm_conv_grad1 runs at the same sizes, so results
 * are bitwise equal to dfm_conv_grad1's; a call whose plan needs more scratch than the
 * array holds fails with DFM_ERR_ARG ("static scratch ... too small: need N bytes").
 * Every call on a GPU shares that GPU's array: calls that overlap on one GPU (other
 * streams, other threads) race and must be ordered by the caller. */
#define DFM_STATIC_SCRATCH_BYTES ((size_t)64 << 20)
int dfm_conv_grad1_noalloc(void *stream, int dimpoint, int dimvect, double sigma, int precision,
                           const double *x, const double *alpha, long nx, const double *y,
                           const double *beta, long ny, double *conv, double *grad);

int dfm_varlin(void *stream, int dimpoint, int dimvect, double sigma, int precision,
               const double *x, long nx, const double *y, const double *beta, long ny,
               double *gamma, void *work, size_t work_bytes);

int dfm_graddim(void *stream, int dimpoint, int dimvect, double sigma, int precision,
                const double *x, long nx, const double *y, const double *beta, long ny,
                int dim, double *gamma, void *work, size_t work_bytes);

int dfm_hessian(void *stream, int dimpoint, int dimvect, double sigma, int precision,
                const double *x, long nx, const double *y, const double *beta, long ny,
                double *gamma, void *work, size_t work_bytes);

int dfm_graddiff(void *stream, int dimpoint, int dimvect, double sigma, int precision,
                 const double *x, const double *beta, const double *eta, long n, double *gamma,
                 void *work, size_t work_bytes);

/* ------------------------------------------------------------------------------------ */
/* Device-resident integrators (device pointers, stream-ordered, caller workspace).     */

enum dfm_scheme { DFM_EULER = 0, DFM_HEUN = 1 };

/* Geodesic shooting of n control points: P, A (T x n x d) receive the trajectory,
 * P[0] = p0, A[0] = a0, dt = (tN - t0) / (T - 1).  DFM_EULER is Diffeos::Shoot
 * (Diffeos.cxx:589-659): p += dt K(p,p) a, a -= dt G^T a, one fused
 * Convolve + ConvolveGradient per step; DFM_HEUN is the Heun variant the reference keeps
 * commented out (Diffeos.cxx:634-652).  |a0|_F < 1e-20 leaves the points in place
 * (Diffeos.cxx:609-611). */
size_t dfm_shoot_workspace(int dimpoint, long n, int precision);
int dfm_shoot(void *stream, int dimpoint, double sigma, int precision, int scheme,
              const double *p0, const double *a0, long n, int T, double t0, double tN,
              double *P, double *A, void *work, size_t work_bytes);

/* Flow of m landmark points in the deformation (P, A) (T x n x d): X (T x m x d), X[0] = x0.
 * Diffeos::FlowLandmarkPointsTrajectory (Diffeos.cxx:661-710); improved_euler != 0 is the
 * reference's default Heun scheme. */
size_t dfm_flow_workspace(int dimpoint, long n, long m, int precision);
int dfm_flow(void *stream, int dimpoint, double sigma, int precision, int improved_euler,
             const double *P, const double *A, long n, const double *x0, long m, int T,
             double t0, double tN, double *X, void *work, size_t work_bytes);

/* Transpose of the shooting field's Jacobian, F(p, a) = (K(p,p) a, -G(p,a)^T a), at (p, a)
 * applied to (mu_p, mu_a):  jp = GradSym(mu_p) + SpecialHessian(mu_a),
 * ja = Convolve(mu_p) - Diff(mu_a), one fused pair sum.  No single reference counterpart:
 * the reference's adjoint assembles the same kernel sums from Convolve, ConvolveGradient
 * and ConvolveSpecialHessian calls (AdjointEquationsIntegrator.cxx:476-567).  Workspace
 * from dfm_workspace_size(DFM_OP_SHOOT_VJP, ...). */
int dfm_shoot_vjp(void *stream, int dimpoint, double sigma, int precision, const double *p,
                  const double *a, const double *mu_p, const double *mu_a, long n, double *jp,
                  double *ja, void *work, size_t work_bytes);

/* Gradient of a terminal cost E(x_{T-1}) with respect to the initial control points p0,
 * momenta a0 and landmarks x0, given theta_T = dE/dx_{T-1} (m x d): the exact discrete
 * adjoint (backward pass) of dfm_shoot + dfm_flow with the same scheme, precision and
 * time grid.  The forward pass is run once, storing (p, a, x) every checkpoint_every
 * steps and at the end; the backward pass recomputes the interior states of each segment
 * from its checkpoint.  checkpoint_every = 0 chooses: every state (nothing recomputed)
 * when the trajectory, (T + 1) x (2n + m) x d doubles, takes at most 4 GiB, else every 5th.
 * Results are bitwise independent of the interval.  An alternative to
 * Diffeos::IntegrateAdjointEquations / AdjointEquationsIntegrator::Update
 * (Diffeos.cxx:1390-1443, AdjointEquationsIntegrator.cxx:33-136, 414-567), whose continuous
 * adjoint agrees with this gradient to O(dt).  grad_x0 may be NULL. */
enum dfm_adjoint_form { DFM_ADJOINT_DISCRETE = 0, DFM_ADJOINT_CONTINUOUS = 1 };
size_t dfm_adjoint_workspace(int dimpoint, long n, long m, int T, int checkpoint_every,
                             int precision);
int dfm_shoot_flow_adjoint(void *stream, int dimpoint, double sigma, int precision, int scheme,
                           int improved_euler, const double *p0, const double *a0, long n,
                           const double *x0, long m, int T, double t0, double tN,
                           int checkpoint_every, const double *theta_T, double *grad_p0,
                           double *grad_a0, double *grad_x0, void *work, size_t work_bytes);
/* The same gradient as the reference computes it: Diffeos::IntegrateAdjointEquations +
 * AdjointEquationsIntegrator::Update (Diffeos.cxx:1390-1443, AdjointEquationsIntegrator.cxx:
 * 59-136, 414-567), the continuous adjoint equations of the shoot + flow integrated
 * backward along the forward states with the flow's Euler / improved-Euler choice:
 * theta' = J_x v^T theta, then xi' = J_F^T xi + J_z v^T theta (ComputeUpdateAt).  Returns
 * (xi_pos(0), xi_mom(0), theta(0)), the reference's GetAdjointPosAt0 / GetAdjointMomAt0 /
 * landmark adjoint at 0.  Same arguments and workspace (dfm_adjoint_workspace) as
 * dfm_shoot_flow_adjoint. */
int dfm_shoot_flow_adjoint_equations(void *stream, int dimpoint, double sigma, int precision,
                                     int scheme, int improved_euler, const double *p0,
                                     const double *a0, long n, const double *x0, long m, int T,
                                     double t0, double tN, int checkpoint_every,
                                     const double *theta_T, double *grad_p0, double *grad_a0,
                                     double *grad_x0, void *work, size_t work_bytes);

/* The continuous adjoint with landmark conditions at several time indices: the reference's
 * Diffeos::IntegrateAdjointEquations with a list and jump times (Diffeos.cxx:1393-1443,
 * AdjointEquationsIntegrator.cxx:59-136), the backward pass of the geodesic regression
 * (Regression.cxx:600-662).  conditions (device, njumps x m x d) are the observations'
 * match gradients in the order of jump_times (host, njumps indices in [0, T)).  As in the
 * reference: theta starts from the last condition if its time is T-1 (else from zero);
 * the condition at jump time t is added to theta's rate over the step t -> t-1 (both
 * stages under improved Euler); the reference's list index moves to the next jump only
 * under improved Euler, so under standard Euler at most one jump after the terminal one
 * is applied; a single condition is no jump and starts theta at T-1 whatever its time; a
 * jump at time 0 is never reached.  Outputs and workspace (dfm_adjoint_workspace) as
 * dfm_shoot_flow_adjoint_equations. */
int dfm_shoot_flow_adjoint_jumps(void *stream, int dimpoint, double sigma, int precision,
                                 int scheme, int improved_euler, const double *p0,
                                 const double *a0, long n, const double *x0, long m, int T,
                                 double t0, double tN, int checkpoint_every, int njumps,
                                 const int *jump_times, const double *conditions,
                                 double *grad_p0, double *grad_a0, double *grad_x0, void *work,
                                 size_t work_bytes);

/* ---- ragged batches of subjects (device pointers, host offset arrays) -------------------
 * A batch of nseg subjects (1 <= nseg <= 65535) stores each per-point array concatenated
 * over subjects: subject s owns rows [off[s], off[s + 1]) where off is a host array of
 * nseg + 1 entries, off[0] = 0, non-decreasing.  The library writes device copies of the
 * offsets into the workspace with kernel launches, and batched calls make no allocation,
 * memset, host copy or synchronisation: they can be captured in a HIP graph (stream
 * capture) and replayed.  Each pair sum covers every subject in one launch (grid.z =
 * subject).  FP32 paths skip the Morton ordering in batches and rely on the per-wavefront
 * exact-difference switch for accuracy.
 * No single reference counterpart: the reference runs subjects one after another, or on
 * threads, through per-subject kernel objects (AbstractAtlas.cxx:490-560, 826-865). */

/* HIP graphs of stream-ordered calls: begin capture on a created stream, issue batched
 * calls, end capture to get an instantiated graph, replay it with dfm_graph_launch. */
int dfm_graph_begin_capture(void *stream);
int dfm_graph_end_capture(void *stream, void **graph_exec);
int dfm_graph_launch(void *graph_exec, void *stream);
int dfm_graph_destroy(void *graph_exec);

/* dfm_conv_grad1 for every subject at once: targets x (rows by x_offsets), sources y and
 * weights beta (rows by y_offsets); conv and/or grad may be NULL.  op for the workspace
 * query: DFM_OP_CONV, DFM_OP_GRAD1 or DFM_OP_CONV_GRAD1 (what the call computes). */
size_t dfm_batch_workspace(int op, int dimpoint, int dimvect, int nseg, const long *x_offsets,
                           const long *y_offsets, int precision);
int dfm_batch_conv_grad1(void *stream, int dimpoint, int dimvect, double sigma, int precision,
                         int nseg, const long *x_offsets, const long *y_offsets, const double *x,
                         const double *alpha, const double *y, const double *beta, double *conv,
                         double *grad, void *work, size_t work_bytes);

/* dfm_shoot / dfm_flow / dfm_shoot_flow_adjoint for every subject at once: control points
 * by cp_offsets, landmarks by lm_offsets; trajectories time-major, P, A: T x (all control
 * points) x d and X: T x (all landmarks) x d.  The zero-momenta exit is per subject. */
size_t dfm_batch_shoot_workspace(int dimpoint, int nseg, const long *cp_offsets, int precision);
int dfm_batch_shoot(void *stream, int dimpoint, double sigma, int precision, int scheme, int nseg,
                    const long *cp_offsets, const double *p0, const double *a0, int T, double t0,
                    double tN, double *P, double *A, void *work, size_t work_bytes);
size_t dfm_batch_flow_workspace(int dimpoint, int nseg, const long *cp_offsets,
                                const long *lm_offsets, int precision);
int dfm_batch_flow(void *stream, int dimpoint, double sigma, int precision, int improved_euler,
                   int nseg, const long *cp_offsets, const long *lm_offsets, const double *P,
                   const double *A, const double *x0, int T, double t0, double tN, double *X,
                   void *work, size_t work_bytes);
size_t dfm_batch_adjoint_workspace(int dimpoint, int nseg, const long *cp_offsets,
                                   const long *lm_offsets, int T, int checkpoint_every,
                                   int precision);
int dfm_batch_shoot_flow_adjoint(void *stream, int dimpoint, double sigma, int precision,
                                 int scheme, int improved_euler, int nseg,
                                 const long *cp_offsets, const long *lm_offsets,
                                 const double *p0, const double *a0, const double *x0, int T,
                                 double t0, double tN, int checkpoint_every,
                                 const double *theta_T, double *grad_p0, double *grad_a0,
                                 double *grad_x0, void *work, size_t work_bytes);

/* ---- kernel matrix, ConvolveInverse and parallel transport (device pointers) ----------- */

/* K (n x n) = [exp(-|y_i - y_j|^2 / sigma^2)]; replaces ExactKernel::ComputeKernelMatrix
 * (ExactKernel.cxx:131-143). */
int dfm_kernel_matrix(void *stream, int dimpoint, double sigma, const double *y, long n,
                      double *K);

/* v (n x dv) = K(y, y)^-1 x, FP64: ExactKernel::ConvolveInverse (ExactKernel.cxx:154-164),
 * whose Armadillo solve() is LAPACK dgesv; here the same LU with partial pivoting by
 * rocSOLVER (dgetrf / dgetrs) on the stream.  n <= 46340.  DFM_ERR_ARG if K is singular
 * (the call synchronises the stream to report it). */
size_t dfm_convolve_inverse_workspace(int dimpoint, int dimvect, long n);
int dfm_convolve_inverse(void *stream, int dimpoint, int dimvect, double sigma, const double *y,
                         long n, const double *x, double *v, void *work, size_t work_bytes);

/* Parallel transport of momenta mom0 (on the control points cp0, m == n) along the geodesic
 * P, A (T x n x d, as written by dfm_shoot with the same sigma and time grid), reported at
 * ntargets increasing target times into out (ntargets x n x d): Diffeos::ParallelTransport
 * (Diffeos.cxx:293-582), with its early exits, its norm-conserving rescaling and its
 * finite-difference Runge-Kutta 2 step; kernel sums at `precision`, the K^-1 solves in
 * FP64.  The output of target time t is the transport of step t - (index of
 * initial_time); the reference indexes its per-step buffer by t, which is the same when
 * the initial time is t0. */
size_t dfm_parallel_transport_workspace(int dimpoint, long n, long m, int precision);
int dfm_parallel_transport(void *stream, int dimpoint, double sigma, int precision,
                           const double *P, const double *A, long n, int T, double t0,
                           double tN, const double *mom0, const double *cp0, long m,
                           double initial_time, const double *target_times, int ntargets,
                           double *out, void *work, size_t work_bytes);

/* ---- surface data terms (device pointers) ----------------------------------------------
 * Meshes are vertices (nv x 3, FP64, row-major) and triangles (nf x 3, int32 indices).
 * DFM_CURRENT is OrientedSurfaceMesh (area-weighted normals as currents), DFM_VARIFOLD is
 * NonOrientedSurfaceMesh (normals scaled by |N|^-1/2, upper-triangular N N^T weights). */
enum dfm_data_term { DFM_CURRENT = 0, DFM_VARIFOLD = 1 };

/* Vertex -> corner incidence list of a mesh (host pointers, built once per topology):
 * offsets (nv + 1), incidences (3 nf) holding 3 * face + corner, grouped by vertex in
 * increasing corner order.  Used by dfm_surface_match_gradient's deterministic gather, in
 * place of the reference's serial scatter over faces (OrientedSurfaceMesh.cxx:205-234,
 * NonOrientedSurfaceMesh.cxx:213-256).  DFM_ERR_ARG for an index outside [0, nv). */
int dfm_mesh_incidence(const int *faces, long nf, long nv, int *offsets, int *incidences);

/* Workspace bytes for a match (gradient = 0) or a match gradient (gradient = 1) between a
 * template of nfs faces and a target of nft faces. */
size_t dfm_surface_workspace(int kind, long nfs, long nft, int precision, int gradient);

/* *match (device, one double) = |S|^2 + |T|^2 - 2 <S, T> of the template S (vs, fs) and the
 * target T (vt, ft) at kernel width sigma.  Replaces OrientedSurfaceMesh::ComputeMatch
 * (OrientedSurfaceMesh.cxx:125-163) and NonOrientedSurfaceMesh::ComputeMatch
 * (NonOrientedSurfaceMesh.cxx:124-165).  DFM_ERR_ARG for a face index outside [0, nv). */
int dfm_surface_match(void *stream, int kind, double sigma, int precision, const double *vs,
                      long nvs, const int *fs, long nfs, const double *vt, long nvt,
                      const int *ft, long nft, double *match, void *work, size_t work_bytes);

/* grad (device, nvs x 3) = d match / d vs, with the template's incidence list from
 * dfm_mesh_incidence (device copies of offsets and incidences).  Replaces
 * OrientedSurfaceMesh::ComputeMatchGradient (OrientedSurfaceMesh.cxx:166-237) and
 * NonOrientedSurfaceMesh::ComputeMatchGradient (NonOrientedSurfaceMesh.cxx:169-260). */
int dfm_surface_match_gradient(void *stream, int kind, double sigma, int precision,
                               const double *vs, long nvs, const int *fs, long nfs,
                               const int *offsets, const int *incidences, const double *vt,
                               long nvt, const int *ft, long nft, double *grad, void *work,
                               size_t work_bytes);

/* ---- polyline data terms (device pointers) ---------------------------------------------
 * Polylines are points (nv x dim, dim = 2 or 3, FP64, row-major) and segments (ne x 2,
 * int32 indices).  DFM_CURRENT is OrientedPolyLine (centres (p0 + p1) / 2, tangents
 * p1 - p0), DFM_VARIFOLD is NonOrientedPolyLine (tangents scaled by |p1 - p0|^-1/2,
 * upper-triangular T T^T weights). */

/* Vertex -> corner incidence list of any cell list of `corners` (1..4) points per cell
 * (host pointers): offsets (nv + 1), incidences (corners * ncells) holding
 * corners * cell + corner, grouped by vertex in increasing corner order.  corners = 2 gives
 * the segment list dfm_curve_match_gradient gathers with, in place of the reference's
 * serial scatter over segments (OrientedPolyLine.cxx:216-234, NonOrientedPolyLine.cxx:
 * 215-251).  DFM_ERR_ARG for an index outside [0, nv). */
int dfm_cell_incidence(const int *cells, long ncells, int corners, long nv, int *offsets,
                       int *incidences);

/* Workspace bytes for a polyline match (gradient = 0) or match gradient (gradient = 1)
 * between a template of nes segments and a target of net segments in dimension dim. */
size_t dfm_curve_workspace(int kind, int dim, long nes, long net, int precision, int gradient);

/* *match (device, one double) = |S|^2 + |T|^2 - 2 <S, T> of the template S (vs, es) and the
 * target T (vt, et).  Replaces OrientedPolyLine::ComputeMatch (OrientedPolyLine.cxx:
 * 132-170) and NonOrientedPolyLine::ComputeMatch (NonOrientedPolyLine.cxx:129-170).
 * DFM_ERR_DIM for dim outside {2, 3}; DFM_ERR_ARG for a segment index outside [0, nv). */
int dfm_curve_match(void *stream, int kind, int dim, double sigma, int precision,
                    const double *vs, long nvs, const int *es, long nes, const double *vt,
                    long nvt, const int *et, long net, double *match, void *work,
                    size_t work_bytes);

/* grad (device, nvs x dim) = d match / d vs, with the template's segment incidence list
 * from dfm_cell_incidence (corners = 2; device copies).  Replaces
 * OrientedPolyLine::ComputeMatchGradient (OrientedPolyLine.cxx:177-238) and
 * NonOrientedPolyLine::ComputeMatchGradient (NonOrientedPolyLine.cxx:177-255). */
int dfm_curve_match_gradient(void *stream, int kind, int dim, double sigma, int precision,
                             const double *vs, long nvs, const int *es, long nes,
                             const int *offsets, const int *incidences, const double *vt,
                             long nvt, const int *et, long net, double *grad, void *work,
                             size_t work_bytes);

/* ---- volume-mesh data terms (device pointers) --------------------------------------------
 * OrientedVolumeMesh: points (nv x dim, dim = 2 or 3) and cells of dim + 1 points (int32;
 * triangles in 2D, tetrahedra in 3D), each a Dirac at its centre weighted by its volume
 * |det(p1 - p0, ..., pD - p0)| / D! (OrientedVolumeMesh.cxx:297-334). */

/* Workspace bytes for a volume match (gradient = 0) or match gradient (gradient = 1). */
size_t dfm_volume_workspace(int dim, long ncs, long nct, int precision, int gradient);

/* *match (device, one double) = |S|^2 + |T|^2 - 2 <S, T>.  Replaces
 * OrientedVolumeMesh::ComputeMatch (OrientedVolumeMesh.cxx:114-152).  DFM_ERR_DIM for dim
 * outside {2, 3}; DFM_ERR_ARG for a cell index outside [0, nv). */
int dfm_volume_match(void *stream, int dim, double sigma, int precision, const double *vs,
                     long nvs, const int *cs, long ncs, const double *vt, long nvt,
                     const int *ct, long nct, double *match, void *work, size_t work_bytes);

/* grad (device, nvs x dim) with the template's point -> corner incidence list from
 * dfm_cell_incidence (corners = dim + 1; device copies).  Replaces
 * OrientedVolumeMesh::ComputeMatchGradient (OrientedVolumeMesh.cxx:156-286) as written: the
 * determinant's cofactor vectors carry no sign of the determinant (the derivative of the
 * match for positively oriented cells), and the 3D corner factor is the single-precision
 * 1/3 of the reference. */
int dfm_volume_match_gradient(void *stream, int dim, double sigma, int precision,
                              const double *vs, long nvs, const int *cs, long ncs,
                              const int *offsets, const int *incidences, const double *vt,
                              long nvt, const int *ct, long nct, double *grad, void *work,
                              size_t work_bytes);

/* ---- point-cloud and landmark data terms (device pointers) -------------------------------
 * PointCloud: points (n x dim, dim = 2 or 3), each a Dirac weighted by one scalar (the
 * VTK point scalars, else 1; PointCloud.cxx:203-226).  Landmark: corresponding points, no
 * kernel. */

/* Workspace bytes for a point-cloud match (gradient = 0) or match gradient (gradient = 1)
 * between a template of ns points and a target of nt points. */
size_t dfm_pointcloud_workspace(int dim, long ns, long nt, int precision, int gradient);

/* *match (device, one double) = |S|^2 + |T|^2 - 2 <S, T> with |S|^2 = sum_i ws_i (K ws)(xs_i)
 * and <S, T> = sum_j wt_j (K ws)(xt_j).  Replaces PointCloud::ComputeMatch
 * (PointCloud.cxx:108-148; its print of the value is not made) with UpdateSelfNorm
 * (:231-250).  DFM_ERR_DIM for dim outside {2, 3}. */
int dfm_pointcloud_match(void *stream, int dim, double sigma, int precision, const double *xs,
                         const double *ws, long ns, const double *xt, const double *wt, long nt,
                         double *match, void *work, size_t work_bytes);

/* grad (device, ns x dim) = 2 (ConvolveGradient(xs, ws) from (xs, ws) + ConvolveGradient(xs,
 * ws) from (xt, wt)).  Replaces PointCloud::ComputeMatchGradient (PointCloud.cxx:152-194)
 * as written: the cross term enters with +, so the result is d/dxs |S + T|^2, not the
 * derivative of the match. */
int dfm_pointcloud_match_gradient(void *stream, int dim, double sigma, int precision,
                                  const double *xs, const double *ws, long ns, const double *xt,
                                  const double *wt, long nt, double *grad, void *work,
                                  size_t work_bytes);

/* Workspace bytes for dfm_landmark_match (0 for dim outside {2, 3} or a bad count). */
size_t dfm_landmark_workspace(int dim, long n);

/* *match (device, one double) = |X - Y|_F^2 of n corresponding points.  Replaces
 * Landmark::ComputeMatch (Landmark.cxx:287-304). */
int dfm_landmark_match(void *stream, int dim, const double *x, const double *y, long n,
                       double *match, void *work, size_t work_bytes);

/* grad (device, n x dim) = 2 (X - Y).  Replaces Landmark::ComputeMatchGradient
 * (Landmark.cxx:306-322). */
int dfm_landmark_match_gradient(void *stream, int dim, const double *x, const double *y, long n,
                                double *grad);

/* Data term between two sets of oriented samples (centre c, area-weighted normal N; the
 * reference's m_Centers / m_Normals of a mesh, OrientedSurfaceMesh.cxx:298-326 and
 * NonOrientedSurfaceMesh.cxx:269-308, given directly): *match (device double) =
 * |S|^2 + |T|^2 - 2 <S, T> with the current kernel K(c_i, c_j) N_i.N_j or the varifold
 * kernel K(c_i, c_j) (N_i.N_j)^2 / (|N_i| |N_j|); grad_centers and grad_normals (ns x 3,
 * either may be NULL) are d match / d c_S and d match / d N_S.  The mesh entry points
 * above chain these to vertex positions. */
size_t dfm_samples_workspace(int kind, long ns, long nt, int precision, int gradient);
int dfm_samples_match_gradient(void *stream, int kind, double sigma, int precision,
                               const double *cs, const double *Ns, long ns, const double *ct,
                               const double *Nt, long nt, double *match, double *grad_centers,
                               double *grad_normals, void *work, size_t work_bytes);

/* ---- deterministic atlas of surfaces or polylines ----------------------------------------
 * Log-likelihood and gradient of DeterministicAtlas for surface subjects
 * (DeterministicAtlas::UpdateFixedEffectsAndComputeCompleteLogLikelihood and
 * ::ComputeCompleteLogLikelihoodGradient, DeterministicAtlas.cxx:100-151, 159-281, with
 * AbstractAtlas::ComputeDataTermGradient(Subject) and ::ConvolveGradTemplate,
 * AbstractAtlas.cxx:490-620, 727-751; RKHS regularity):
 *   loglik[0] = -1/2 sum_s r_s / data_sigma_squared, r_s = match(phi_s(template), target_s)
 *   loglik[1] = -1/2 sum_s a_s^T K(cp, cp) a_s
 * and the gradients of loglik[0] + loglik[1] with respect to the momenta a_s, the shared
 * control points cp and the template vertices (optionally smoothed).  The adjoint is the
 * exact discrete adjoint of the shoot + flow scheme (the reference's continuous adjoint
 * agrees with it to O(dt)).  Every subject runs in the same launches.
 * dfm_atlas_create allocates and uploads everything (template topology, target meshes and
 * their squared norms, the working memory; dfm_atlas_device_bytes reports the latter);
 * dfm_atlas_gradient then allocates nothing and launches kernels only, so it can be
 * captured in a HIP graph.  Calls on one atlas must be ordered on one stream. */
typedef struct dfm_atlas dfm_atlas;

typedef struct dfm_atlas_desc {
  int kind;                     /* DFM_CURRENT or DFM_VARIFOLD */
  int precision;                /* DFM_FP64, DFM_FP32C or DFM_FP32X */
  int scheme;                   /* DFM_EULER or DFM_HEUN shooting */
  int improved_euler;           /* Heun (1) or Euler (0) landmark flow */
  int number_of_time_points;    /* T >= 1 */
  int checkpoint_every;         /* adjoint checkpoint interval >= 1, or 0: automatic */
  double t0, tN;
  double deformation_width;     /* kernel width of the deformation */
  double attachment_width;      /* kernel width of the current / varifold norm */
  double data_sigma_squared;    /* noise variance of the data term */
  double smoothing_width;       /* template gradient smoothing width; <= 1e-20: none */
  long n_control_points;
  long n_template_vertices;
  long n_template_faces;
  const int *template_faces;    /* host, n_template_faces x cell_corners */
  int nsubjects;                /* 1 .. 65535 */
  const long *target_vertex_offsets;  /* host, nsubjects + 1 */
  const long *target_face_offsets;    /* host, nsubjects + 1 */
  const double *target_vertices;      /* host, concatenated, x dimension */
  const int *target_faces;            /* host, concatenated, indices local to each target */
  int adjoint;                  /* DFM_ADJOINT_DISCRETE or DFM_ADJOINT_CONTINUOUS */
  int dimension;                /* 2 or 3 (0: 3); every point array below is x dimension */
  int cell_corners;             /* 3: triangle surfaces (dimension 3; 0: 3), 2: polylines
                                   (OrientedPolyLine / NonOrientedPolyLine segments) */
} dfm_atlas_desc;

int dfm_atlas_create(const dfm_atlas_desc *desc, dfm_atlas **atlas);
int dfm_atlas_destroy(dfm_atlas *atlas);
size_t dfm_atlas_device_bytes(const dfm_atlas *atlas);
/* Device pointers (d = the descriptor's dimension): cp (n x d), momenta
 * (nsubjects x n x d), template_vertices (nv x d); outputs loglik (2 doubles), residuals
 * (nsubjects, may be NULL), grad_cp (n x d, may be NULL), grad_momenta (nsubjects x n x d,
 * may be NULL), grad_template (nv x d, may be NULL). */
int dfm_atlas_gradient(dfm_atlas *atlas, void *stream, const double *cp, const double *momenta,
                       const double *template_vertices, double *loglik, double *residuals,
                       double *grad_cp, double *grad_momenta, double *grad_template);

#ifdef __cplusplus
}
#endif

#endif /* DFM_HIP_H */
