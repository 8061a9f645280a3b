// pairsum_kernels.h — the pair-sum kernel templates, instantiated once per precision in
// pairsum_f64.hip, pairsum_f32c.hip and pairsum_f32x.hip (compiled in parallel).
//
// One templated wavefront kernel evaluates every op of the engine: each 64-lane
// workgroup owns R*64 targets held in registers, streams the sources through a private
// LDS tile, evaluates K(x_i, y_j) = exp(-|x_i - y_j|^2 / sigma^2) once per pair and
// feeds it to the op's accumulators:
//
//   CONV       sum_j K w_j                                  ExactKernel.cxx:51-74
//   GRAD1      sum_j (a_i.w_j) grad_x K                      ExactKernel.cxx:264-275
//   CONVGRAD1  both of the above in one pass
//   VARLIN     sum_j w_j[k] d/dx_l K                         ExactKernel.cxx:166-202
//   GRADDIM    sum_j w_j d/dx_dim K                          ExactKernel.cxx:294-323
//   GRADDIFF   sum_j (b_i.b_j) H_x K (eta_j - eta_i)         AbstractKernel.cxx:22-88
//   SHOOTVJP   transpose Jacobian of the shooting field (integrators.hip), fused
//
// FP64 path: coordinates scaled by 1/sigma, K = exp(-r^2) in FP64 — the arithmetic of the
// reference's ExactKernel.
// FP32C path: targets are ordered along a Hilbert curve (radix sort of 30-bit keys) so
// each wavefront's targets are spatially compact; each wavefront re-centres its targets
// and every staged source on the centre of its targets' bounding box in FP64, scales by
// sqrt(log2 e)/sigma and only then rounds to FP32, so the FP32 differences carry the
// rounding of the small centred coordinates rather than of the absolute ones.
// K = exp2(-r'^2) on the transcendental unit (v_exp_f32).  Each lane accumulates 32
// sources in FP32 and folds that partial sum into FP64 accumulators (cascaded,
// compensated summation: the FP32 rounding is bounded by a 32-term partial sum, never by
// the running total).  A wavefront whose targets span more than kExactExtent scaled
// units (sparse targets, narrow kernels) instead forms each difference from FP32 hi and
// lo parts, (xh - yh) + (xl - yl); FP32X takes that path in every wavefront.
//
// Ragged batches: grid.z = segment; segment s pairs targets [xoff[s], xoff[s + 1]) with
// sources [yoff[s], yoff[s + 1]) of the concatenated arrays.
//
// Work split: grid.x = target chunks of R*64, grid.y = source splits.  With one split the
// kernel writes the scaled results straight to the caller's output (through the sort
// permutation); with several it writes FP64 partials that k_reduce_splits sums in split
// order, so the result is deterministic.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <type_traits>

#include "dfm_hip.h"
#include "dfm_internal.h"

namespace dfm {

// ------------------------------------------------------------------------------------
// precision traits

template <typename T>
struct Prec;

// kFold: sources summed in the working precision before the partial sum is folded into
// the FP64 accumulator.  In FP32 a one-tile (64-source) partial keeps the rounding of each
// add proportional to 64 terms even when a wide kernel makes every pair significant and
// the sum cancels.
template <>
struct Prec<float> {
  static constexpr bool kCenter = true;
  static constexpr int kTile = kTileF32;
  static constexpr int kFold = kFoldF32;
  __device__ __forceinline__ static float kexp(float r2) { return __builtin_amdgcn_exp2f(-r2); }
};

template <>
struct Prec<double> {
  static constexpr bool kCenter = false;
  static constexpr int kTile = kTileF64;
  static constexpr int kFold = kTileF64;
  __device__ __forceinline__ static double kexp(double r2) { return exp(-r2); }
};

// ------------------------------------------------------------------------------------
// op shapes

template <int OP, int D, int DV>
struct Shape {
  // per-target extra inputs: alpha (GRAD1, CONVGRAD1), beta + eta (GRADDIFF) or
  // alpha + mu_p + mu_a (SHOOTVJP)
  static constexpr int NTE = (OP == OP_GRAD1 || OP == OP_CONVGRAD1 || OP == OP_VARIFOLD ||
                              OP == OP_VARIFOLD_CONV)
                                 ? DV
                             : (OP == OP_GRADDIFF)                  ? DV + D
                             : (OP == OP_SHOOTVJP)                  ? DV + 2 * D
                                                                    : 0;
  // per-source record: y, w (+ eta_j for GRADDIFF, + mu_p_j, mu_a_j for SHOOTVJP)
  static constexpr int NSRC =
      D + DV + (OP == OP_GRADDIFF ? D : 0) + (OP == OP_SHOOTVJP ? 2 * D : 0);
  static constexpr int NSP = (NSRC + 3) / 4 * 4;
  // with the FP32 residuals of the source coordinates appended (exact-difference path)
  static constexpr int NSPX = (NSRC + D + 3) / 4 * 4;
  static constexpr int NACC = OP == OP_CONV        ? DV
                              : OP == OP_GRAD1     ? D
                              : OP == OP_CONVGRAD1 ? DV + D
                              : OP == OP_VARIFOLD  ? DV + D
                              : OP == OP_VARIFOLD_CONV ? DV
                              : OP == OP_VARLIN    ? DV * D
                              : OP == OP_GRADDIM   ? DV
                              : OP == OP_GRADDIFF  ? 2 * D
                              : OP == OP_HESS      ? DV * (1 + D * (D + 1) / 2)
                                                   : 5 * D;
  // FP32 compact wavefronts of CONV, GRAD1 and CONVGRAD1 use the expanded distance (see
  // pair_update_expand): record [y (D) | w (DV) | |y|^2], and GRAD1 accumulates sum_j t_ij
  // beside sum_j t_ij y_j
  static constexpr bool EXPANDABLE =
      OP == OP_CONV || OP == OP_GRAD1 || OP == OP_CONVGRAD1 || OP == OP_VARIFOLD ||
      OP == OP_VARIFOLD_CONV;
  static constexpr int NSPE = (D + DV + 1 + 3) / 4 * 4;
  static constexpr int NACCE = OP == OP_GRAD1 ? D + 1 : NACC;
  static constexpr int NACCMAX = NACCE > NACC ? NACCE : NACC;
};

// ------------------------------------------------------------------------------------
// per-pair update.  xi: scaled (centred) target, xl: its FP32 rounding residual (exact
// path); te: per-target extras; rec: staged source record
// [y (D) | w (DV) | eta_j (D, GRADDIFF) | y residual (D, exact path)]; acc: the op's
// working-precision accumulators.
//
// Exact path: with x = xh + xl and y = yh + yl (FP32 hi parts and residuals of the FP64
// centred coordinates) the difference is formed as (xh - yh) + (xl - yl), which carries
// the rounding of the difference itself instead of the rounding of the coordinates.

template <int OP, int D, int DV, int AUX, bool EXACT, typename T>
__device__ __forceinline__ void pair_update(const T *xi, const T *xl, const T *te, const T *rec,
                                            T *acc) {
  constexpr int LO = Shape<OP, D, DV>::NSRC;
  T d[D];
  T r2 = T(0);
#pragma unroll
  for (int k = 0; k < D; k++) {
    if constexpr (EXACT)
      d[k] = (xi[k] - rec[k]) + (xl[k] - rec[LO + k]);
    else
      d[k] = xi[k] - rec[k];
    r2 += d[k] * d[k];
  }
  const T K = Prec<T>::kexp(r2);
  const T *w = rec + D;
  if constexpr (OP == OP_CONV) {
#pragma unroll
    for (int k = 0; k < DV; k++) acc[k] += K * w[k];
  } else if constexpr (OP == OP_GRAD1 || OP == OP_CONVGRAD1 || OP == OP_VARIFOLD ||
                       OP == OP_VARIFOLD_CONV) {
    // VARIFOLD: the fused sum with K (a_i . w_j) as the Convolve weight (a = w = u);
    // VARIFOLD_CONV its Convolve part alone
    constexpr int G = (OP == OP_GRAD1) ? 0 : DV;
    T sa = te[0] * w[0];
#pragma unroll
    for (int k = 1; k < DV; k++) sa += te[k] * w[k];
    const T kc = (OP == OP_VARIFOLD || OP == OP_VARIFOLD_CONV) ? K * sa : K;
    if constexpr (OP != OP_GRAD1) {
#pragma unroll
      for (int k = 0; k < DV; k++) acc[k] += kc * w[k];
    }
    if constexpr (OP != OP_VARIFOLD_CONV) {
      const T t = kc * sa;
#pragma unroll
      for (int l = 0; l < D; l++) acc[G + l] += t * d[l];
    }
  } else if constexpr (OP == OP_VARLIN) {
#pragma unroll
    for (int k = 0; k < DV; k++) {
      const T t = K * w[k];
#pragma unroll
      for (int l = 0; l < D; l++) acc[k * D + l] += t * d[l];
    }
  } else if constexpr (OP == OP_GRADDIM) {
    const T t = K * d[AUX];
#pragma unroll
    for (int k = 0; k < DV; k++) acc[k] += t * w[k];
  } else if constexpr (OP == OP_HESS) {
    // per weight k: K w_k and K w_k d d^T (upper triangle); the Hessian's two terms are
    // combined in FP64 at the end
    constexpr int NQ = D * (D + 1) / 2;
    T q[NQ];
    int m = 0;
#pragma unroll
    for (int a = 0; a < D; a++)
#pragma unroll
      for (int b = a; b < D; b++) q[m++] = d[a] * d[b];
#pragma unroll
    for (int k = 0; k < DV; k++) {
      const T t = K * w[k];
      acc[k * (NQ + 1)] += t;
#pragma unroll
      for (int j = 0; j < NQ; j++) acc[k * (NQ + 1) + 1 + j] += t * q[j];
    }
  } else if constexpr (OP == OP_SHOOTVJP) {
    // transpose of the shooting field's Jacobian at (p, a) applied to (mu_p, mu_a):
    //   conv:  K mu_p_j                                   -> Convolve(mu_p)
    //   gsym:  K (mu_p_i . a_j + a_i . mu_p_j) d          -> GradSym(mu_p)
    //   diff:  K ((mu_a_i - mu_a_j) . d) a_j               -> Diff(mu_a)
    //   sh1/2: K (a_i . a_j) e,  K (a_i . a_j)(d . e) d   -> SpecialHessian(mu_a)
    const T *a_i = te;
    const T *mp_i = te + DV;
    const T *ma_i = te + DV + D;
    const T *mp_j = rec + D + DV;
    const T *ma_j = rec + D + DV + D;
    T s1 = mp_i[0] * w[0] + a_i[0] * mp_j[0];
    T bb = a_i[0] * w[0];
#pragma unroll
    for (int k = 1; k < D; k++) {
      s1 += mp_i[k] * w[k] + a_i[k] * mp_j[k];
      bb += a_i[k] * w[k];
    }
    T e[D];
    T ed = T(0);
#pragma unroll
    for (int k = 0; k < D; k++) {
      e[k] = ma_i[k] - ma_j[k];
      ed += e[k] * d[k];
    }
    const T t1 = K * s1, t2 = K * ed, t3 = K * bb, t4 = t3 * ed;
#pragma unroll
    for (int k = 0; k < D; k++) {
      acc[k] += K * mp_j[k];
      acc[D + k] += t1 * d[k];
      acc[2 * D + k] += t2 * w[k];
      acc[3 * D + k] += t3 * e[k];
      acc[4 * D + k] += t4 * d[k];
    }
  } else {  // OP_GRADDIFF
    T bb = te[0] * w[0];
#pragma unroll
    for (int k = 1; k < DV; k++) bb += te[k] * w[k];
    const T *eta_i = te + DV;
    const T *eta_j = rec + D + DV;
    T e[D];
    T dot = T(0);
#pragma unroll
    for (int k = 0; k < D; k++) {
      e[k] = eta_i[k] - eta_j[k];
      dot += d[k] * e[k];
    }
    const T t = K * bb;
    const T t2 = t * dot;
#pragma unroll
    for (int k = 0; k < D; k++) {
      acc[k] += t * e[k];
      acc[D + k] += t2 * d[k];
    }
  }
}

// Per-pair update of the expanded-distance form (FP32 compact wavefronts; CONV, GRAD1,
// CONVGRAD1).  With xt, yt the FP32 centred scaled points,
//   |xt - yt|^2 = |xt|^2 + (|yt|^2 - 2 xt . yt)
// and the bracket is three FMAs from the staged |yt|^2 and xm2 = -2 xt, so
//   K = exp2(-|xt|^2) K',  K' = exp2(-(|yt|^2 - 2 xt . yt)),
// with the per-target factor exp2(-|xt|^2) applied once, in FP64, after the sum.  The
// gradient sum_j t_ij (xt_i - yt_j), t_ij = K_ij (a_i . w_j), is formed at the end as
// xt_i sum_j t_ij - sum_j t_ij yt_j, with sum_j t_ij = a_i . conv_i for CONVGRAD1 and an
// accumulator of its own for GRAD1.  Three FMAs replace the three differences and three
// multiply-adds of |xt - yt|^2; the error stays that of the FP32 differences (the
// bracket's rounding is of the order of the coordinates' own FP32 rounding).
template <int OP, int D, int DV>
__device__ __forceinline__ void pair_update_expand(const float *xm2, const float *te,
                                                   const float *rec, float *acc) {
  float e = rec[D + DV];
#pragma unroll
  for (int k = 0; k < D; k++) e = __builtin_fmaf(xm2[k], rec[k], e);
  const float K = Prec<float>::kexp(e);
  const float *w = rec + D;
  if constexpr (OP == OP_CONV) {
#pragma unroll
    for (int k = 0; k < DV; k++) acc[k] += K * w[k];
  } else {  // GRAD1, CONVGRAD1, VARIFOLD (Convolve weight K (a_i . w_j)), VARIFOLD_CONV
    constexpr int G = (OP == OP_GRAD1) ? 0 : DV;
    float sa = te[0] * w[0];
#pragma unroll
    for (int k = 1; k < DV; k++) sa += te[k] * w[k];
    const float kc = (OP == OP_VARIFOLD || OP == OP_VARIFOLD_CONV) ? K * sa : K;
    if constexpr (OP != OP_GRAD1) {
#pragma unroll
      for (int k = 0; k < DV; k++) acc[k] += kc * w[k];
    }
    if constexpr (OP != OP_VARIFOLD_CONV) {
      const float t = kc * sa;
      if constexpr (OP == OP_GRAD1) acc[D] += t;
#pragma unroll
      for (int l = 0; l < D; l++) acc[G + l] += t * rec[l];
    }
  }
}

// The expanded form's sums of one target -> the sums pair_update accumulates (the conv
// part, then sum_j t_ij (xt_i - yt_j) for the gradient), in FP64.
template <int OP, int D, int DV, int NA>
__device__ __forceinline__ void expand_finish(double (&a)[NA], const float *xt, const float *te) {
  double x[D];
  double n2 = 0.0;
#pragma unroll
  for (int k = 0; k < D; k++) {
    x[k] = (double)xt[k];
    n2 += x[k] * x[k];
  }
  const double f = exp2(-n2);
  if constexpr (OP == OP_CONV || OP == OP_VARIFOLD_CONV) {
#pragma unroll
    for (int k = 0; k < DV; k++) a[k] *= f;
  } else if constexpr (OP == OP_GRAD1) {
    const double st = a[D];
#pragma unroll
    for (int l = 0; l < D; l++) a[l] = f * (x[l] * st - a[l]);
  } else {  // OP_CONVGRAD1, OP_VARIFOLD: sum_j t_ij = a_i . conv_i in both
    double st = 0.0;
#pragma unroll
    for (int k = 0; k < DV; k++) st += (double)te[k] * a[k];
#pragma unroll
    for (int l = 0; l < D; l++) a[DV + l] = f * (x[l] * st - a[DV + l]);
#pragma unroll
    for (int k = 0; k < DV; k++) a[k] *= f;
  }
}

// The fused integrator step for target `o` of segment `seg` from its finished sums (KParams
// e_*): the same arithmetic, in the same order, as the separate update (k_lin) it replaces.
template <int D>
__device__ __forceinline__ void step_update(const KParams &p, long o, int seg, const double *conv,
                                            const double *grad) {
  const bool frozen = p.e_gate && p.e_gate[seg] < kZeroMomenta2;
#pragma unroll
  for (int k = 0; k < D; k++) {
    const long i = o * D + k;
    double vp, va;
    if (frozen) {
      vp = p.e_pb[i];
      va = 0.0;
    } else {
      vp = 1.0 * p.e_pb[i];
      va = 1.0 * p.e_ab[i];
      if (p.e_px) {
        vp += p.e_cx * p.e_px[i];
        va += -p.e_cx * p.e_ax[i];
      }
      vp += p.e_c * conv[k];
      va += -p.e_c * grad[k];
    }
    p.e_p1[i] = vp;
    p.e_a1[i] = va;
  }
}

// Store element i of output K (0: out0, 1: out1) with value v, through the linear
// epilogue (KParams o_*) when it is on.
template <int K>
__device__ __forceinline__ void put(const KParams &p, double *out, long i, double v) {
  if (p.o_on) {
    const double s = p.o_cs[K] * v;
    v = p.o_base[K] ? p.o_cb[K] * p.o_base[K][i] + s : s;
    if (p.o_extra[K]) v += p.o_ce[K] * p.o_extra[K][i];
  }
  out[i] = v;
}

// Scale the FP64 accumulators of one target and store them at original index `o`.
template <int OP, int D, int DV>
__device__ __forceinline__ void finalize(const double *a, const KParams &p, long o) {
  if constexpr (OP == OP_CONV || OP == OP_VARIFOLD_CONV) {
#pragma unroll
    for (int k = 0; k < DV; k++) put<0>(p, p.out0, o * DV + k, a[k]);
  } else if constexpr (OP == OP_GRAD1) {
#pragma unroll
    for (int l = 0; l < D; l++) put<0>(p, p.out0, o * D + l, p.cg * a[l]);
  } else if constexpr (OP == OP_CONVGRAD1 || OP == OP_VARIFOLD) {
    if (p.out0) {
#pragma unroll
      for (int k = 0; k < DV; k++) put<0>(p, p.out0, o * DV + k, a[k]);
    }
    if (p.out1) {
#pragma unroll
      for (int l = 0; l < D; l++) put<1>(p, p.out1, o * D + l, p.cg * a[DV + l]);
    }
  } else if constexpr (OP == OP_VARLIN) {
#pragma unroll
    for (int k = 0; k < DV * D; k++) p.out0[o * DV * D + k] = p.cg * a[k];
  } else if constexpr (OP == OP_GRADDIM) {
#pragma unroll
    for (int k = 0; k < DV; k++) p.out0[o * DV + k] = p.cg * a[k];
  } else if constexpr (OP == OP_GRADDIFF) {
#pragma unroll
    for (int k = 0; k < D; k++) p.out0[o * D + k] = p.ch1 * a[k] + p.ch2 * a[D + k];
  } else if constexpr (OP == OP_HESS) {
    // H_k = 4 / sigma^4 sum K w_k dx dx^T - 2 / sigma^2 sum K w_k I, with dx = d / s:
    // 4 / (sigma^4 s^2) = -ch2 and 2 / sigma^2 = ch1; out0 (nx x DV x D x D), both triangles
    constexpr int NQ = D * (D + 1) / 2;
#pragma unroll
    for (int k = 0; k < DV; k++) {
      const double *h = a + k * (NQ + 1);
      double *out = p.out0 + (o * DV + k) * D * D;
      int m = 1;
#pragma unroll
      for (int r = 0; r < D; r++)
#pragma unroll
        for (int c = r; c < D; c++) {
          const double v = -p.ch2 * h[m++] - (r == c ? p.ch1 * h[0] : 0.0);
          out[r * D + c] = v;
          out[c * D + r] = v;
        }
    }
  } else {  // OP_SHOOTVJP: out0 = J_p, out1 = J_a
#pragma unroll
    for (int k = 0; k < D; k++) {
      put<0>(p, p.out0, o * D + k, p.cg * a[D + k] + p.ch1 * a[3 * D + k] + p.ch2 * a[4 * D + k]);
      put<1>(p, p.out1, o * D + k, a[k] - p.cg * a[2 * D + k]);
    }
  }
}

__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v = fmin(v, __shfl_xor(v, m, 64));
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v = fmax(v, __shfl_xor(v, m, 64));
  return v;
}

// ------------------------------------------------------------------------------------
// the pair-sum kernel

// Stream the sources [js, je) through the wavefront's LDS tile and accumulate, in `ntiles`
// tiles (at least enough for [js, je); the wavefronts of a workgroup pass the same number of
// barriers, so they take the same count, and tiles past je stage padding and sum nothing).
// EXPAND selects the expanded-distance form (pair_update_expand; FP32 only).
template <typename T, int OP, int D, int DV, int AUX, int R, bool EXACT, bool EXPAND, int NTE,
          int NA, bool PF>
__device__ __forceinline__ void run_tiles(const KParams &p, T *sm, int lane, long js, long je,
                                          long ntiles, const double *c, double s,
                                          const T (&xt)[R][D],
                                          const T (&xl)[R][D], const T (&te)[R][NTE],
                                          double (&accd)[R][NA]) {
  using S = Shape<OP, D, DV>;
  static_assert(!EXPAND || (S::EXPANDABLE && !EXACT), "expanded form: CONV / GRAD1 / CONVGRAD1");
  constexpr int NSP = EXPAND ? S::NSPE : EXACT ? S::NSPX : S::NSP;
  constexpr int NACC = EXPAND ? S::NACCE : S::NACC;  // accumulators in use
  static_assert(NACC <= NA, "accumulator array holds the form's sums");
  constexpr int LO = S::NSRC;
  constexpr int TS = Prec<T>::kTile;
  constexpr int FOLD = Prec<T>::kFold;
  constexpr int REC_PER_LANE = TS / 64;
  static_assert(TS >= 64 && TS % 64 == 0, "each lane stages whole records: tiles of 64k sources");
  // a padded source sits far outside the support of K: exp(-|x - far|^2) is exactly 0
  const T far = T(1.0e4);

  T acc[R][NACC];
#pragma unroll
  for (int r = 0; r < R; r++)
#pragma unroll
    for (int a = 0; a < NACC; a++) acc[r][a] = T(0);
  T xm2[R][D];  // -2 xt, the expanded form's cross-term factor
#pragma unroll
  for (int r = 0; r < R; r++)
#pragma unroll
    for (int k = 0; k < D; k++) xm2[r][k] = T(-2) * xt[r][k];

  // PF (the small-grid form): one-record-per-lane tiles read the next tile's raw FP64
  // sources into registers before summing the current one, so the global loads are in
  // flight during the sums, where few wavefronts per SIMD would not hide them.  The
  // one-wavefront form leaves it off: its registers would cost the large grids occupancy
  // (the fused FP32C sum at N = M = 1M: 332 against 306 ms).
  constexpr int NRAW = S::NSRC;
  constexpr bool PREFETCH = PF && REC_PER_LANE == 1 && NRAW <= 9;
  double raw[PREFETCH ? NRAW : 1];
  auto fetch = [&](long j) {
    if constexpr (PREFETCH) {
      if (j < je) {
#pragma unroll
        for (int k = 0; k < D; k++) raw[k] = p.y[j * D + k];
#pragma unroll
        for (int k = 0; k < DV; k++) raw[D + k] = p.w[j * DV + k];
        if constexpr (OP == OP_GRADDIFF || OP == OP_SHOOTVJP) {
#pragma unroll
          for (int k = 0; k < D; k++) raw[D + DV + k] = p.se[j * D + k];
        }
        if constexpr (OP == OP_SHOOTVJP) {
#pragma unroll
          for (int k = 0; k < D; k++) raw[D + DV + D + k] = p.se2[j * D + k];
        }
      }
    }
  };
  if constexpr (PREFETCH) {
    if (ntiles > 0) fetch(js + lane);
  }

  for (long t = 0; t < ntiles; t++) {
    const long j0 = js + t * TS;
    __syncthreads();
#pragma unroll
    for (int q = 0; q < REC_PER_LANE; q++) {
      const long j = j0 + q * 64 + lane;
      T rec[NSP];
#pragma unroll
      for (int k = 0; k < NSP; k++) rec[k] = T(0);
      if (j < je) {
        // the record's raw values: prefetched, or read now
        double v[NRAW];
        if constexpr (PREFETCH) {
#pragma unroll
          for (int k = 0; k < NRAW; k++) v[k] = raw[k];
        } else {
#pragma unroll
          for (int k = 0; k < D; k++) v[k] = p.y[j * D + k];
#pragma unroll
          for (int k = 0; k < DV; k++) v[D + k] = p.w[j * DV + k];
          if constexpr (OP == OP_GRADDIFF || OP == OP_SHOOTVJP) {
#pragma unroll
            for (int k = 0; k < D; k++) v[D + DV + k] = p.se[j * D + k];
          }
          if constexpr (OP == OP_SHOOTVJP) {
#pragma unroll
            for (int k = 0; k < D; k++) v[D + DV + D + k] = p.se2[j * D + k];
          }
        }
#pragma unroll
        for (int k = 0; k < D; k++) {
          const double u = (v[k] - c[k]) * s;
          rec[k] = (T)u;
          if constexpr (EXACT) rec[LO + k] = (T)(u - (double)rec[k]);
        }
#pragma unroll
        for (int k = D; k < NRAW; k++) rec[k] = (T)v[k];
      } else {
#pragma unroll
        for (int k = 0; k < D; k++) rec[k] = far;
      }
      if constexpr (EXPAND) {  // |yt|^2 of the FP32 coordinates, rounded once
        double n2 = 0.0;
#pragma unroll
        for (int k = 0; k < D; k++) n2 += (double)rec[k] * (double)rec[k];
        rec[D + DV] = (T)n2;
      }
      T *dst = sm + (q * 64 + lane) * NSP;
#pragma unroll
      for (int k = 0; k < NSP; k++) dst[k] = rec[k];
    }
    __syncthreads();
    if constexpr (PREFETCH) {
      if (t + 1 < ntiles) fetch(j0 + TS + lane);
    }

    const long left = je - j0;
    // records past the end of the sources were staged with zero weight far away
    const int cnt = (int)(left <= 0 ? 0 : left < TS ? ((left + FOLD - 1) / FOLD) * FOLD : TS);
    for (int jf = 0; jf < cnt; jf += FOLD) {
      for (int jj = jf; jj < jf + FOLD; jj += 4) {
#pragma unroll
        for (int u = 0; u < 4; u++) {
          T rec[NSP];
          const T *srcp = sm + (jj + u) * NSP;
#pragma unroll
          for (int k = 0; k < NSP; k++) rec[k] = srcp[k];
#pragma unroll
          for (int r = 0; r < R; r++) {
            if constexpr (EXPAND)
              pair_update_expand<OP, D, DV>(xm2[r], te[r], rec, acc[r]);
            else
              pair_update<OP, D, DV, AUX, EXACT, T>(xt[r], xl[r], te[r], rec, acc[r]);
          }
        }
      }
      // fold the FOLD-source partial sums into the FP64 accumulators
#pragma unroll
      for (int r = 0; r < R; r++)
#pragma unroll
        for (int a = 0; a < NACC; a++) {
          accd[r][a] += (double)acc[r][a];
          acc[r][a] = T(0);
        }
    }
  }
}

// Targets of one wavefront (R per lane, positions base + 64 r + lane of [.., xe)) into
// registers, with the centre of their bounding box and its extent (FP32 paths).  Lanes past
// the end are marked invalid and read nothing.
template <typename T, int D, int R>
__device__ __forceinline__ void load_targets(const KParams &p, int lane, long base, long xe,
                                             long (&pos)[R], long (&orig)[R], bool (&valid)[R],
                                             double (&xd)[R][D], double (&c)[D],
                                             double &extent) {
#pragma unroll
  for (int r = 0; r < R; r++) {
    pos[r] = base + r * 64 + lane;
    valid[r] = pos[r] < xe;
    orig[r] = valid[r] ? (p.perm ? (long)p.perm[pos[r]] : pos[r]) : 0;
#pragma unroll
    for (int k = 0; k < D; k++) xd[r][k] = valid[r] ? p.x[orig[r] * D + k] : 0.0;
  }
  extent = 0.0;
#pragma unroll
  for (int k = 0; k < D; k++) c[k] = 0.0;
  if constexpr (Prec<T>::kCenter) {
#pragma unroll
    for (int k = 0; k < D; k++) {
      double lo = INFINITY, hi = -INFINITY;
#pragma unroll
      for (int r = 0; r < R; r++)
        if (valid[r]) {
          lo = fmin(lo, xd[r][k]);
          hi = fmax(hi, xd[r][k]);
        }
      lo = wave_min(lo);
      hi = wave_max(hi);
      c[k] = (lo <= hi) ? 0.5 * (lo + hi) : 0.0;
      extent = (lo <= hi) ? fmax(extent, hi - lo) : extent;
    }
  }
}

// The sums of one wavefront's targets over the sources [js, je) streamed through its LDS
// tile `sm` in `ntiles` tiles, into accd (FP64, the layout finalize takes).  Every lane of
// the wavefront takes part (invalid lanes sit at the centre and their sums are discarded).
template <typename T, int OP, int D, int DV, int AUX, int R, int MODE, bool PF>
__device__ __forceinline__ void wave_sums(const KParams &p, T *sm, int lane, long js, long je,
                                          long ntiles, const long (&orig)[R],
                                          const bool (&valid)[R], const double (&xd)[R][D],
                                          const double (&c)[D], double extent,
                                          double (&accd)[R][Shape<OP, D, DV>::NACCMAX]) {
  using S = Shape<OP, D, DV>;
  constexpr int NA = S::NACCMAX;
  constexpr int NTE = S::NTE > 0 ? S::NTE : 1;
  const double s = p.s;
  T xt[R][D];
  T xl[R][D];
  T te[R][NTE];
#pragma unroll
  for (int r = 0; r < R; r++) {
#pragma unroll
    for (int k = 0; k < D; k++) {
      // lanes past the end sit at the centre (their sums are discarded)
      const double v = valid[r] ? (xd[r][k] - c[k]) * s : 0.0;
      xt[r][k] = (T)v;
      xl[r][k] = (T)(v - (double)xt[r][k]);
    }
    if constexpr (OP == OP_GRAD1 || OP == OP_CONVGRAD1 || OP == OP_VARIFOLD ||
                  OP == OP_VARIFOLD_CONV) {
#pragma unroll
      for (int k = 0; k < DV; k++) te[r][k] = valid[r] ? (T)p.ta[orig[r] * DV + k] : T(0);
    } else if constexpr (OP == OP_GRADDIFF || OP == OP_SHOOTVJP) {
#pragma unroll
      for (int k = 0; k < DV; k++) te[r][k] = valid[r] ? (T)p.ta[orig[r] * DV + k] : T(0);
#pragma unroll
      for (int k = 0; k < D; k++) te[r][DV + k] = valid[r] ? (T)p.tb[orig[r] * D + k] : T(0);
      if constexpr (OP == OP_SHOOTVJP) {
#pragma unroll
        for (int k = 0; k < D; k++) te[r][DV + D + k] = valid[r] ? (T)p.tc[orig[r] * D + k] : T(0);
      }
    } else {
      te[r][0] = T(0);
    }
  }

#pragma unroll
  for (int r = 0; r < R; r++)
#pragma unroll
    for (int a = 0; a < NA; a++) accd[r][a] = 0.0;

  if constexpr (Prec<T>::kCenter) {
    // wavefront-uniform choice: coordinates re-centred on a box wider than a few kernel
    // widths lose too much to FP32 rounding, so such wavefronts form exact differences
    const bool exact = MODE == MODE_EXACT || extent * s > p.exact_extent;
    if (exact) {
      run_tiles<T, OP, D, DV, AUX, R, true, false, NTE, NA, PF>(p, sm, lane, js, je, ntiles, c, s,
                                                                xt, xl, te, accd);
    } else if constexpr (S::EXPANDABLE) {
      run_tiles<T, OP, D, DV, AUX, R, false, true, NTE, NA, PF>(p, sm, lane, js, je, ntiles, c, s,
                                                                xt, xl, te, accd);
#pragma unroll
      for (int r = 0; r < R; r++) expand_finish<OP, D, DV, NA>(accd[r], xt[r], te[r]);
    } else {
      run_tiles<T, OP, D, DV, AUX, R, false, false, NTE, NA, PF>(p, sm, lane, js, je, ntiles, c,
                                                                 s, xt, xl, te, accd);
    }
  } else {
    run_tiles<T, OP, D, DV, AUX, R, false, false, NTE, NA, PF>(p, sm, lane, js, je, ntiles, c, s,
                                                               xt, xl, te, accd);
  }
}

// A wavefront's finished sums (split `split` of p.nsplit), stored through finalize (one
// split) or as split partials.  With a fused integrator step (p.e_on, OP_CONVGRAD1) each
// target's sums are finished and the step applied here: directly with one split, else by
// the wavefront of target tile `tile` (segment `seg`) that is the last of the p.nsplit
// splits to store its partials.
template <int OP, int D, int DV, int R>
__device__ __forceinline__ void store_sums(const KParams &p, int lane, int split,
                                           const long (&pos)[R], const long (&orig)[R],
                                           const bool (&valid)[R],
                                           double (&accd)[R][Shape<OP, D, DV>::NACCMAX],
                                           long tile, int seg) {
  constexpr int NACC = Shape<OP, D, DV>::NACC;
#pragma unroll
  for (int r = 0; r < R; r++) {
    if (!valid[r]) continue;
    if (p.nsplit == 1) {
      finalize<OP, D, DV>(accd[r], p, orig[r]);
    } else {
      double *dst = p.partial + ((long)split * p.nx + pos[r]) * NACC;
#pragma unroll
      for (int a = 0; a < NACC; a++) dst[a] = accd[r][a];
    }
  }
  if (p.nsplit > 1 && p.done) {
    // publish this split's partials, count it, and let the tile's last split add them up
    // (in the split order of k_reduce_splits) and finalize
    __threadfence();
    int last = 0;
    if (lane == 0) last = atomicAdd(&p.done[tile], 1) == p.nsplit - 1;
    last = __shfl(last, 0);
    if (!last) return;
    __threadfence();
    if (lane == 0) p.done[tile] = 0;  // zero again for the next launch
#pragma unroll
    for (int r = 0; r < R; r++) {
      if (!valid[r]) continue;
#pragma unroll
      for (int a = 0; a < NACC; a++) accd[r][a] = 0.0;
      for (int sp = 0; sp < p.nsplit; sp++) {
        const double *src = p.partial + ((long)sp * p.nx + pos[r]) * NACC;
#pragma unroll
        for (int a = 0; a < NACC; a++) accd[r][a] += src[a];
      }
      finalize<OP, D, DV>(accd[r], p, orig[r]);
    }
  } else if (p.nsplit > 1) {
    return;  // k_reduce_splits finishes
  }
  if constexpr (OP == OP_CONVGRAD1 && DV == D) {
    if (p.e_on) {
#pragma unroll
      for (int r = 0; r < R; r++) {
        if (!valid[r]) continue;
        double grad[D];
#pragma unroll
        for (int l = 0; l < D; l++) grad[l] = p.cg * accd[r][DV + l];
        step_update<D>(p, orig[r], seg, accd[r], grad);
      }
    }
  }
}

// wave_sums over [js, je) and store_sums.  The number of __syncthreads() depends on js, je
// only, so the wavefronts of a larger workgroup that call this with the same source range
// pass the same barriers.
template <typename T, int OP, int D, int DV, int AUX, int R, int MODE>
__device__ __forceinline__ void sum_wavefront(const KParams &p, T *sm, int lane, int split,
                                              long js, long je, const long (&pos)[R],
                                              const long (&orig)[R], const bool (&valid)[R],
                                              const double (&xd)[R][D], const double (&c)[D],
                                              double extent, long tile, int seg) {
  constexpr int TS = Prec<T>::kTile;
  double accd[R][Shape<OP, D, DV>::NACCMAX];
  const long ntiles = je > js ? (je - js + TS - 1) / TS : 0;
  wave_sums<T, OP, D, DV, AUX, R, MODE, false>(p, sm, lane, js, je, ntiles, orig, valid, xd, c,
                                               extent, accd);
  store_sums<OP, D, DV, R>(p, lane, split, pos, orig, valid, accd, tile, seg);
}

// One workgroup = W wavefronts on the same 64 R targets (W = 1: one wavefront, the form of
// the large grids).  grid = (target chunks, source splits of W wavefronts each, segments).
template <typename T, int OP, int D, int DV, int AUX, int R, int MODE, int W>
__global__ __launch_bounds__(64 * W) void k_pairsum(KParams p) {
  using S = Shape<OP, D, DV>;
  constexpr int TS = Prec<T>::kTile;
  constexpr int FOLD = Prec<T>::kFold;
  constexpr int NA = S::NACCMAX;
  constexpr int NACC = S::NACC;
  static_assert(TS % FOLD == 0 && FOLD % 4 == 0, "fold interval divides the tile");
  static_assert(S::NSP % 4 == 0 && S::NSPX % 4 == 0, "records are whole 16-byte groups");
  static_assert(W == 1 || R == 1, "several wavefronts per workgroup: one target per lane");
  constexpr int TILE = TS * (Prec<T>::kCenter ? S::NSPX : S::NSP);  // T per wavefront tile
  constexpr int RCH = 4;  // accumulators per round of the workgroup's LDS sum
  constexpr size_t RED = W > 1 ? sizeof(double) * W * 64 * RCH : 0;
  constexpr size_t BYTES = sizeof(T) * W * TILE > RED ? sizeof(T) * W * TILE : RED;

  __shared__ __attribute__((aligned(16))) unsigned char smraw[BYTES];

  const int wave = W > 1 ? (int)(threadIdx.x >> 6) : 0;
  const int lane = W > 1 ? (int)(threadIdx.x & 63) : (int)threadIdx.x;
  T *sm = reinterpret_cast<T *>(smraw) + wave * TILE;
  // segment of a ragged batch (one segment: the whole target and source sets)
  long xb = 0, xe = p.nx, yb = 0, ye = p.ny;
  if (p.xoff) {
    xb = p.xoff[blockIdx.z];
    xe = p.xoff[blockIdx.z + 1];
  }
  if (p.yoff) {
    yb = p.yoff[blockIdx.z];
    ye = p.yoff[blockIdx.z + 1];
  }
  const long base = xb + (long)blockIdx.x * (64 * R);
  if (base >= xe) return;  // the whole workgroup lies past its segment

  long pos[R], orig[R];
  bool valid[R];
  double xd[R][D], c[D], extent;
  load_targets<T, D, R>(p, lane, base, xe, pos, orig, valid, xd, c, extent);
  if constexpr (Prec<T>::kCenter) {
    // the compact wavefronts of this call run on the matrix cores (k_pairsum_mx)
    if (p.wave_filter == 1 && !(extent * p.s > p.exact_extent)) return;
  }
  const long tile = (long)blockIdx.z * gridDim.x + blockIdx.x;
  if constexpr (W == 1) {
    const long js = yb + (long)blockIdx.y * p.split_len;
    const long je = std::min(ye, js + p.split_len);
    sum_wavefront<T, OP, D, DV, AUX, R, MODE>(p, sm, lane, (int)blockIdx.y, js, je, pos, orig,
                                              valid, xd, c, extent, tile, (int)blockIdx.z);
  } else {
    // every wavefront runs the tiles of a whole split (the workgroup's barriers), those past
    // its segment's sources staging padding
    const long js = yb + ((long)blockIdx.y * W + wave) * p.split_len;
    const long je = std::min(ye, js + p.split_len);
    const long ntiles = (p.split_len + TS - 1) / TS;
    double accd[R][NA];
    wave_sums<T, OP, D, DV, AUX, R, MODE, true>(p, sm, lane, js, je, ntiles, orig, valid, xd, c,
                                                extent, accd);
    double *red = reinterpret_cast<double *>(smraw);
    for (int a0 = 0; a0 < NACC; a0 += RCH) {
      __syncthreads();  // the tiles, or the previous round's sums, are consumed
#pragma unroll
      for (int q = 0; q < RCH; q++)
        if (a0 + q < NACC) red[(wave * RCH + q) * 64 + lane] = accd[0][a0 + q];
      __syncthreads();
      if (wave == 0) {
#pragma unroll
        for (int q = 0; q < RCH; q++) {
          if (a0 + q >= NACC) continue;
          double v = 0.0;
          for (int w = 0; w < W; w++) v += red[(w * RCH + q) * 64 + lane];
          accd[0][a0 + q] = v;
        }
      }
    }
    if (wave != 0) return;
    store_sums<OP, D, DV, R>(p, lane, (int)blockIdx.y, pos, orig, valid, accd, tile,
                             (int)blockIdx.z);
  }
}

template <int OP, int D, int DV>
__global__ __launch_bounds__(256) void k_reduce_splits(KParams p) {
  constexpr int NACC = Shape<OP, D, DV>::NACC;
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= p.nx) return;
  double a[NACC];
#pragma unroll
  for (int k = 0; k < NACC; k++) a[k] = 0.0;
  for (int sp = 0; sp < p.nsplit; sp++) {
    const double *src = p.partial + ((long)sp * p.nx + i) * NACC;
#pragma unroll
    for (int k = 0; k < NACC; k++) a[k] += src[k];
  }
  const long o = p.perm ? (long)p.perm[i] : i;
  finalize<OP, D, DV>(a, p, o);
}

// ------------------------------------------------------------------------------------
// dispatch over (op, D, DV, AUX, R) for one (T, MODE)

template <typename T, int MODE, int OP, int D, int DV, int AUX>
static void launch_one(hipStream_t st, const KParams &p, const Plan &pl) {
  dim3 grid((unsigned)pl.chunks, (unsigned)pl.nsplit, (unsigned)pl.nseg);
  if constexpr (std::is_same<T, float>::value &&
                (Shape<OP, D, DV>::EXPANDABLE || OP == OP_SHOOTVJP) && DV <= 3) {
    if (pl.W > 1) {  // the plan holds R = 1 and W = wave_split_of(OP) for these ops
      constexpr int WS = wave_split_of(OP);
      hipLaunchKernelGGL((k_pairsum<T, OP, D, DV, AUX, 1, MODE, WS>), grid, dim3(64 * WS), 0,
                         st, p);
      if (pl.nsplit > 1 && !p.done)
        hipLaunchKernelGGL((k_reduce_splits<OP, D, DV>), dim3((unsigned)((p.nx + 255) / 256)),
                           dim3(256), 0, st, p);
      return;
    }
  }
  switch (pl.R) {
    case 1: hipLaunchKernelGGL((k_pairsum<T, OP, D, DV, AUX, 1, MODE, 1>), grid, dim3(64), 0, st, p); break;
    case 2: hipLaunchKernelGGL((k_pairsum<T, OP, D, DV, AUX, 2, MODE, 1>), grid, dim3(64), 0, st, p); break;
    default: hipLaunchKernelGGL((k_pairsum<T, OP, D, DV, AUX, 4, MODE, 1>), grid, dim3(64), 0, st, p); break;
  }
  // with completion counters k_pairsum finishes its own source splits
  if (pl.nsplit > 1 && !p.done)
    hipLaunchKernelGGL((k_reduce_splits<OP, D, DV>), dim3((unsigned)((p.nx + 255) / 256)),
                       dim3(256), 0, st, p);
}

template <typename T, int MODE, int OP, int D, int DV>
static int launch_aux(hipStream_t st, const KParams &p, const Plan &pl) {
  if constexpr (OP == OP_GRADDIM) {
    switch (p.dim) {
      case 0: launch_one<T, MODE, OP, D, DV, 0>(st, p, pl); break;
      case 1: launch_one<T, MODE, OP, D, DV, 1>(st, p, pl); break;
      default:
        if constexpr (D == 3) launch_one<T, MODE, OP, D, DV, 2>(st, p, pl);
        break;
    }
  } else {
    launch_one<T, MODE, OP, D, DV, 0>(st, p, pl);
  }
  return DFM_OK;
}

template <typename T, int MODE, int OP, int D>
static int launch_dv(hipStream_t st, int DV, const KParams &p, const Plan &pl) {
  if constexpr (OP == OP_GRADDIFF || OP == OP_SHOOTVJP || OP == OP_VARIFOLD ||
                OP == OP_VARIFOLD_CONV) {
    if (DV != D) return set_error(DFM_ERR_DIM, "this op needs dimvect == dimpoint");
    return launch_aux<T, MODE, OP, D, D>(st, p, pl);
  } else {
    switch (DV) {
      case 1: return launch_aux<T, MODE, OP, D, 1>(st, p, pl);
      case 2: return launch_aux<T, MODE, OP, D, 2>(st, p, pl);
      case 3: return launch_aux<T, MODE, OP, D, 3>(st, p, pl);
      case 4: return launch_aux<T, MODE, OP, D, 4>(st, p, pl);
      case 5: return launch_aux<T, MODE, OP, D, 5>(st, p, pl);
      case 6: return launch_aux<T, MODE, OP, D, 6>(st, p, pl);
    }
  }
  return set_error(DFM_ERR_DIM, "unsupported dimvect %d", DV);
}

template <typename T, int MODE, int OP>
static int launch_d(hipStream_t st, int D, int DV, const KParams &p, const Plan &pl) {
  if (D == 2) return launch_dv<T, MODE, OP, 2>(st, DV, p, pl);
  if (D == 3) return launch_dv<T, MODE, OP, 3>(st, DV, p, pl);
  return set_error(DFM_ERR_DIM, "unsupported dimpoint %d", D);
}

template <typename T, int MODE>
static int launch_op(hipStream_t st, int op, int D, int DV, const KParams &p, const Plan &pl) {
  switch (op) {
    case OP_CONV: return launch_d<T, MODE, OP_CONV>(st, D, DV, p, pl);
    case OP_GRAD1: return launch_d<T, MODE, OP_GRAD1>(st, D, DV, p, pl);
    case OP_CONVGRAD1: return launch_d<T, MODE, OP_CONVGRAD1>(st, D, DV, p, pl);
    case OP_VARLIN: return launch_d<T, MODE, OP_VARLIN>(st, D, DV, p, pl);
    case OP_GRADDIM: return launch_d<T, MODE, OP_GRADDIM>(st, D, DV, p, pl);
    case OP_GRADDIFF: return launch_d<T, MODE, OP_GRADDIFF>(st, D, DV, p, pl);
    case OP_SHOOTVJP: return launch_d<T, MODE, OP_SHOOTVJP>(st, D, DV, p, pl);
    case OP_HESS: return launch_d<T, MODE, OP_HESS>(st, D, DV, p, pl);
    case OP_VARIFOLD: return launch_d<T, MODE, OP_VARIFOLD>(st, D, DV, p, pl);
    case OP_VARIFOLD_CONV: return launch_d<T, MODE, OP_VARIFOLD_CONV>(st, D, DV, p, pl);
  }
  return set_error(DFM_ERR_ARG, "unknown op %d", op);
}

}  // namespace dfm
