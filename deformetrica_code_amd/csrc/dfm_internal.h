// dfm_internal.h — types shared by the kernel file and the C-ABI layer.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "dfm_hip.h"

namespace dfm {

enum : int {
  OP_CONV = 0,
  OP_GRAD1 = 1,
  OP_CONVGRAD1 = 2,
  OP_VARLIN = 3,
  OP_GRADDIM = 4,
  OP_GRADDIFF = 5,
  OP_SHOOTVJP = 6,
  OP_HESS = 7,
  OP_VARIFOLD = 8,
  OP_VARIFOLD_CONV = 9  // the Convolve part of OP_VARIFOLD alone
};

// The pair-sum ops of a current or varifold data term (dimvect = dimpoint): the current's
// Convolve and fused Convolve + Grad1 over the normals N; the varifold's sums over the
// unit-root normals u = N / sqrt|N| (OP_VARIFOLD_CONV and OP_VARIFOLD).  The reference's varifold
// convolves the upper triangle W_j of u_j u_j^T (Dv = D (D + 1) / 2) and contracts its
// varlin list with u_i u_i^T (NonOrientedSurfaceMesh.cxx:169-260, NonOrientedPolyLine.cxx:
// 203-251); every use of the result is S((K W)(c_i)) u_i = sum_j K_ij (u_i . u_j) u_j, the
// Convolve part of OP_VARIFOLD, and the gradient is Grad1 with weights (u_i . u_j)^2.
inline int data_conv_op(int kind) { return kind == DFM_CURRENT ? OP_CONV : OP_VARIFOLD_CONV; }
inline int data_grad_op(int kind) { return kind == DFM_CURRENT ? OP_CONVGRAD1 : OP_VARIFOLD; }

constexpr int kBBoxParts = 256;

// Source tile (LDS) per wavefront (a whole number of 64-record stagings) and, for FP32,
// sources summed in FP32 before each fold into the FP64 accumulators.  A small FP32 tile
// leaves the wavefront count per SIMD to the registers: at N = M = 1M the fused FP32 sum
// runs in 385 ms with 64-record tiles (3 KB) against 423 ms with 256 (12 KB, 13 wavefronts
// per CU); FP64 runs 1087 ms with 128-record tiles and 1098 ms with 64.  Folding once per
// 64-source tile instead of every 32 sources saves 0.9% (375 -> 372 ms, same box) for a 2%
// larger error (conv 8.1e-8 -> 8.3e-8, gradient 1.11e-7 -> 1.12e-7 normwise vs FP64).
constexpr int kTileF64 = 128;
constexpr int kTileF32 = 64;
constexpr int kFoldF32 = 64;

// FP32 single calls Hilbert-order their targets from this many targets on.  Below it the
// sort's launches (bounding box, keys, radix passes) cost more than the compact wavefronts
// save, and wide wavefronts take the exact-difference path instead.
constexpr long kSortMinTargets = 8192;

// DFM_BF16MFMA: targets per workgroup of the matrix-core kernel (4 wavefronts x 4 tiles of
// 16), which also decides, per group of this many sorted targets, which path sums them;
// and sources per LDS tile (one staging thread each)
constexpr int kBfGroup = 256;
constexpr int kBfTile = 256;
// DFM_BF16MFMA: a group of kBfGroup targets spanning more than this many scaled kernel units
// takes the FP32C wavefront path in the same workgroup instead.  The matrix cores add each
// group of 8 K-slot products aligned to the largest with 24 bits below it (measured:
// 2^24 + 1 - 2^24 gives 1, 2^25 + 1 - 2^25 gives 0) and the groups in FP32, so
// r' = |y|^2 - 2 x.y rounds at ~2^-24 of |x|^2 for a target at distance |x| from the group
// centre: a target norm added in the K dimension would not cancel exactly (and
// exp2(-r') reaches 2^|x|^2: 2^48 at this bound, far from FP32 overflow).  The gradient's
// x sum(t) - sum(t y) also loses more to cancellation the farther a target sits from the
// centre.  Single calls (Hilbert-sorted): the 1M-sample sphere varifold at sigma 0.02
// (34% of its groups span more than 6 units, 12% more than 8) runs its match + gradient in
// 810 ms at 8 against 905 at 7 and 950 at 6, with gradient errors against FP64 of 3.2e-7,
// 2.7e-7 and 2.5e-7 (centres; normals 2.5e-7, 2.0e-7, 1.8e-7).  1M uniform points at
// sigma 0.05 have no group above 6.
constexpr double kBfExtent = 8.0;
// Ragged batches (the caller's order): the 64-subject atlas's varifold gradient sums
// (OP_VARIFOLD, weights K (u_i . u_j)^2) put its template gradient 1.08e-6 from FP64 at 6
// and 7.7e-7 at 5 (FP32C 7.5e-7; eager gradient 82.8 against 79.2 ms).
constexpr double kBfExtentBatch = 5.0;

// |a0|_F < 1e-20 leaves the points where they are (Diffeos.cxx:609-611, 684-685)
constexpr double kZeroMomenta2 = 1e-40;

// FP32 paths: a wavefront whose targets span more than this many scaled kernel units
// (sqrt(log2 e) / sigma per unit length) forms exact (hi + lo) coordinate differences.
constexpr double kExactExtent = 4.0;
// checkpoint_every = 0: keep every state (nothing recomputed in the backward) when the
// trajectory takes at most this many bytes, else every kAutoCheckpoint-th state
constexpr double kResidentTrajectoryBytes = 4.0 * 1024 * 1024 * 1024;
constexpr int kAutoCheckpoint = 5;

enum : int { MODE_ADAPTIVE = 0, MODE_EXACT = 1 };

// Launch parameters of the pair-sum kernel (passed by value).
struct KParams {
  const double *x;   // targets (nx x D)
  const double *ta;  // per-target extra: alpha (nx x DV) or beta (GRADDIFF)
  const double *tb;  // per-target extra: eta (GRADDIFF), mu_p (SHOOTVJP)
  const double *tc;  // per-target extra: mu_a (SHOOTVJP)
  const double *y;   // sources (ny x D)
  const double *w;   // weights (ny x DV)
  const double *se;  // per-source extra: eta_j (GRADDIFF), mu_p_j (SHOOTVJP)
  const double *se2; // per-source extra: mu_a_j (SHOOTVJP)
  const uint32_t *perm;
  // ragged batch (blockIdx.z = segment): segment s owns targets [xoff[s], xoff[s + 1]) and
  // sources [yoff[s], yoff[s + 1]) of the concatenated arrays; null for a single segment
  const long *xoff;
  const long *yoff;
  double *partial;
  double *out0;
  double *out1;
  long nx, ny;
  long split_len;
  int nsplit;
  int dim;
  double sigma;
  double s;    // coordinate scale
  double cg;   // gradient factor  -2 / (sigma^2 s)
  double ch1;  // GradDiff:  2 / sigma^2
  double ch2;  // GradDiff: -4 / (sigma^4 s^2)
  double exact_extent;  // FP32C: wavefront extent (scaled units) above which differences are exact
  double exact_extent_bf;  // DFM_BF16MFMA: group extent above which k_pairsum runs the group
  int wave_filter;      // 0 every wavefront; 1 only wide ones (the rest run on the matrix cores)
  // Source splits finished in k_pairsum (null: by k_reduce_splits): the wavefront of each
  // target tile (workgroup along the targets, times segment) that is the last of the
  // p.nsplit splits to store its partials adds them up and finalizes.  One counter per tile,
  // zero at launch and reset by that last wavefront; the caller (the integrators) owns the
  // array, clears it once, and orders the launches that share it on one stream.
  int *done;
  // Fused integrator step (OP_CONVGRAD1 with x = positions, ta = w = momenta):
  //   e_p1 = e_pb [+ e_cx e_px] + e_c conv,   e_a1 = e_ab [- e_cx e_ax] - e_c grad,
  // gated per segment by e_gate (squared norm of the segment's initial momenta below 1e-40:
  // e_p1 = e_pb, e_a1 = 0).  k_pairsum applies it to each target once its sums are complete;
  // other kernels leave it to k_step_update.
  int e_on;
  double *e_p1, *e_a1;
  const double *e_pb, *e_ab, *e_px, *e_ax, *e_gate;
  double e_c, e_cx;
  // Linear epilogue of the Convolve / Grad1 / fused / SHOOTVJP sums (o_on != 0): output k
  // (out0, out1) takes  o_cb[k] base_k + o_cs[k] S + o_ce[k] extra_k  (added in that order)
  // instead of the finished sum S, where a null base / extra term is left out.  base_k and
  // extra_k are indexed like out_k and may alias it (each element is read, then written, by
  // the one thread that finishes its target).  It saves the integrators the elementwise
  // launch that would otherwise follow the sum.
  int o_on;
  const double *o_base[2], *o_extra[2];
  double o_cb[2], o_cs[2], o_ce[2];
};

// Work decomposition and workspace layout of one call.
// Wavefronts per workgroup of the small-grid form of k_pairsum (FP32, one target per lane):
// the W wavefronts of a workgroup share its 64 targets and take W consecutive source
// splits, whose sums are added in LDS (wavefront order, FP64) before one store, so a grid
// of few target workgroups fills the chip without split partials and their reduction
// launch.
constexpr int kWaveSplit = 16;
// SHOOTVJP's fifteen accumulators per target: eight wavefronts (101 VGPRs under a budget of
// 256; at sixteen the budget of 128 spills 53)
constexpr int wave_split_of(int op) { return op == OP_SHOOTVJP ? 8 : kWaveSplit; }

struct Plan {
  int R;            // targets per lane
  int W;            // wavefronts per workgroup sharing its targets (FP32, R = 1: wave_split_of)
  int mx;           // FP32C: accumulate on the matrix cores (k_pairsum_mx), R = 1
  int nseg;         // segments of a ragged batch (grid.z)
  bool sort;        // Hilbert-order the targets first (single-segment FP32 calls)
  long chunks;      // workgroups along the targets (of the largest segment)
  int nsplit;       // source splits
  long split_len;   // sources per split (multiple of the tile)
  size_t off_bbox, off_keys_in, off_keys_out, off_vals_in, off_perm, off_sort_tmp, off_partial;
  size_t sort_tmp_bytes;
  size_t work_bytes;
};

inline bool valid_precision(int prec) { return prec >= DFM_FP64 && prec <= DFM_BF16MFMA; }

int set_error(int code, const char *fmt, ...);
const char *last_error();
void clear_error();

Plan make_plan(int op, int prec, int D, int DV, long nx, long ny);
// A ragged batch of nseg segments: nx_total targets in all, at most max_nx targets and
// max_ny sources in one segment.  No Hilbert sort (batched calls are capture-safe).
Plan make_batch_plan(int op, int prec, int D, int DV, int nseg, long nx_total, long max_nx,
                     long max_ny);
int fill(hipStream_t st, double *dst, long n, double v);

// One pair-sum call on device pointers: plans the work, checks the workspace, clamps the
// kernel width and launches (empty source sets give zero outputs).
// `ws_label` names the workspace in the error of a call that does not fit it.
int pairsum(hipStream_t st, int op, int prec, int D, int DV, double sigma, KParams p,
            long out0_n, long out1_n, void *work, size_t work_bytes,
            const char *ws_label = "workspace");
size_t pairsum_workspace(int op, int prec, int D, int DV, long nx, long ny);

// Segments of a ragged batch: host offsets (n + 1 entries, from 0, non-decreasing) and
// their device copy.  n == 1 with null offsets is one plain segment of `total` points.
struct Seg {
  int n = 1;
  const long *h = nullptr;
  const long *d = nullptr;
  long total = 0;
  long max = 0;
  static Seg single(long count) {
    Seg s;
    s.total = s.max = count;
    return s;
  }
  long count(int i) const { return h ? h[i + 1] - h[i] : total; }
};
// Checks host offsets (first 0, non-decreasing) and fills total and max.
int seg_from_host(Seg &s, int n, const long *h, const long *d);
// Writes n longs from host memory into device memory with kernel launches only (no
// memcpy node, so a captured graph replays it).
int put_longs(hipStream_t st, long *dst, const long *src, int n);
// One pair-sum call over a ragged batch (targets segmented by xs, sources by ys, same
// segment count); p.nx / p.ny are the concatenated totals.
int pairsum_batch(hipStream_t st, int op, int prec, int D, int DV, double sigma, KParams p,
                  const Seg &xs, const Seg &ys, long out0_n, long out1_n, void *work,
                  size_t work_bytes);
size_t pairsum_batch_workspace(int op, int prec, int D, int DV, const Seg &xs, const Seg &ys);
int copy_d2d(hipStream_t st, void *dst, const void *src, size_t bytes);

// surface data-term launchers (surfaces.hip); nsub meshes share one triangle list
int surf_geometry(hipStream_t st, int kind, const double *V, long nv, const int *F, long nf,
                  int nsub, double *C, double *W, double *A, double *N, double *bad);
int surf_contrib(hipStream_t st, int kind, const double *V, long nv, const int *F, long nf,
                 int nsub, const double *N, const double *ktS, const double *gS,
                 const double *ktT, const double *gT, double *contrib);
// the same for cells of `corners` points in dimension D: triangles (3, D = 3) through the
// surface launchers, segments (2, D = 2 or 3) through the polyline kernels; contrib holds
// corners x D values per cell
int cells_geometry(hipStream_t st, int kind, int D, int corners, const double *V, long nv,
                   const int *cells, long nc, int nsub, double *C, double *W, double *A,
                   double *N, double *bad);
int cells_contrib(hipStream_t st, int kind, int D, int corners, const double *V, long nv,
                  const int *cells, long nc, int nsub, const double *N, const double *ktS,
                  const double *gS, const double *ktT, const double *gT, double *contrib);
int cells_gather(hipStream_t st, int D, int corners, const int *offsets, const int *inc, long nv,
                 long nc, int nsub, const double *contrib, double scale, double *grad);
int surf_gather(hipStream_t st, const int *offsets, const int *inc, long nv, long nf, int nsub,
                const double *contrib, double scale, double *grad);
int seg_dot(hipStream_t st, const double *X, const double *Y, int dv, int nsub, long seg_len,
            const long *off, double *out);

// batched shoot + flow adjoint in phases (integrators.hip): phase 1 forward (checkpoints
// and *x_final in the workspace), phase 2 backward from them, phase 3 both; form
// DFM_ADJOINT_DISCRETE or DFM_ADJOINT_CONTINUOUS
int batch_adjoint_phase(hipStream_t st, int D, double sigma, int prec, int scheme, int improved,
                        int nseg, const long *cp_off, const long *lm_off, const double *p0,
                        const double *a0, const double *x0, int T, double t0, double tN, int k,
                        const double *theta_T, double *gp, double *ga, double *gx, void *work,
                        size_t work_bytes, int phase, const double **x_final, int form);
size_t batch_adjoint_workspace(int D, int nseg, const long *cp_off, const long *lm_off, int T,
                               int k, int prec);

// the matrix-core form exists for this op and shape (pairsum_f32c.hip)
bool mx_supported(int op, int D, int DV);
// the BF16 cross-term kernel exists for this op and shape (pairsum_bf16.hip)
bool bf_supported(int op, int D, int DV);
int launch_op_bf16(hipStream_t st, int op, int D, int DV, const KParams &p, const Plan &pl);

// one per precision (pairsum_f64.hip, pairsum_f32c.hip, pairsum_f32x.hip)
int launch_op_f64(hipStream_t st, int op, int D, int DV, const KParams &p, const Plan &pl);
int launch_op_f32c(hipStream_t st, int op, int D, int DV, const KParams &p, const Plan &pl);
int launch_op_f32x(hipStream_t st, int op, int D, int DV, const KParams &p, const Plan &pl);
// autotuner of the plan choice (pairsum.hip): mode 0 heuristics, 1 explore + exploit,
// 2 use tuned choices only
void tune_set_mode(int mode);
void tune_reset();
int tune_report(char *buf, size_t len);
void profile_enable(bool on);
int profile_read(double *total_ms, long *launches);
int profile_work(double *pairs, double *flops);
hipEvent_t profile_begin(hipStream_t st);
void profile_end(hipStream_t st, hipEvent_t e0, double pairs, double flops);
// algorithmic flops of one (target, source) pair of an op (pairsum.hip)
double flops_per_pair(int op, int D, int DV);

}  // namespace dfm
