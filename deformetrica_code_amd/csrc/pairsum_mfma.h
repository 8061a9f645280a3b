// pairsum_mfma.h — the FP32C fused Convolve + ConvolveGradient(X, alpha) (and Grad1 alone)
// with the per-pair accumulation on the matrix cores.
//
// For targets i and sources j (centred, scaled coordinates, K_ij = exp2(-|x_i - y_j|^2)):
//   conv_i = sum_j K_ij w_j                       = (K B)_i, columns w
//   grad_i ~ sum_j K_ij (a_i.w_j)(x_i - y_j)
//          = sum_k a_ik (x_i S_ik - T_ik.),  S = K w,  T_ik. = sum_j K_ij w_jk y_j
// so both sums are one matrix product K (targets x sources) times B (sources x 4 Dv),
// B_j = [w_j | w_j (x) y_j], staged once per source tile.  The VALU evaluates only
// r^2 and exp2 per pair (one pair per lane: the A operand of v_mfma_f32_16x16x4f32, lane l
// holding K[target l%16][source l/16]); the MFMA accumulates 16 targets x 4 sources x 16
// columns per instruction in FP32, folded into FP64 every 32 sources; the per-target
// combination a.(x S - T) runs once per target at the end.  Same algorithmic sums as
// pair_update's CONVGRAD1 / GRAD1 (ExactKernel.cxx:51-74, 264-275), same FP64 partial
// layout, so the split reduction and finalize are shared.
//
// A wavefront is 64 targets (four 16-row tiles).  Only compact wavefronts (extent below
// the exact-difference threshold) run here; wide ones are left to k_pairsum with R = 1,
// which takes exactly the complementary set of wavefronts (KParams::wave_filter).
#pragma once

#include "pairsum_kernels.h"

namespace dfm {

typedef float floatx4 __attribute__((ext_vector_type(4)));

template <int DV>
struct MxShape {
  static constexpr int NC = 4 * DV;               // columns: w (DV) + w (x) y (3 DV)
  static constexpr int NB = (NC + 15) / 16;       // 16-column MFMA blocks
  static constexpr int REC = 4 + 16 * NB;         // LDS record: y (3) + pad, B (NB x 16)
  static constexpr int TS = DV <= 3 ? 128 : 64;   // sources per LDS tile (<= 10 KB)
  static constexpr int NT = 4;                    // 16-target tiles per wavefront
  static constexpr int FOLD = 32;                 // sources per FP32 partial
};

template <int OP, int DV>
__global__ __launch_bounds__(64) void k_pairsum_mx(KParams p) {
  constexpr int D = 3;
  using S = MxShape<DV>;
  constexpr int NC = S::NC, NB = S::NB, REC = S::REC, TS = S::TS, NT = S::NT;
  // the source tile and, after the loop, the per-target reduction share one buffer
  constexpr int TILE_FLOATS = TS * REC;
  constexpr int RED_FLOATS = NT * 16 * 16 * 2;  // one 16-column block of FP64 sums
  constexpr int BUF = TILE_FLOATS > RED_FLOATS ? TILE_FLOATS : RED_FLOATS;
  __shared__ __attribute__((aligned(16))) float sm[BUF];

  const int lane = threadIdx.x;
  const int split = blockIdx.y;
  long xb = 0, xe = p.nx, yb = 0, ye = p.ny;
  if (p.xoff) {
    xb = p.xoff[blockIdx.z];
    xe = p.xoff[blockIdx.z + 1];
  }
  if (p.yoff) {
    yb = p.yoff[blockIdx.z];
    ye = p.yoff[blockIdx.z + 1];
  }
  const long base = xb + (long)blockIdx.x * 64;
  if (base >= xe) return;

  // ---- the wavefront's 64 targets: centre and extent exactly as k_pairsum with R = 1
  const long pos = base + lane;
  const bool valid = pos < xe;
  const long orig = valid ? (p.perm ? (long)p.perm[pos] : pos) : 0;
  double xd[D];
#pragma unroll
  for (int k = 0; k < D; k++) xd[k] = valid ? p.x[orig * D + k] : 0.0;
  double c[D];
  double extent = 0.0;
#pragma unroll
  for (int k = 0; k < D; k++) {
    double lo = valid ? xd[k] : INFINITY, hi = valid ? xd[k] : -INFINITY;
    lo = wave_min(lo);
    hi = wave_max(hi);
    c[k] = (lo <= hi) ? 0.5 * (lo + hi) : 0.0;
    extent = (lo <= hi) ? fmax(extent, hi - lo) : extent;
  }
  const double s = p.s;
  if (extent * s > p.exact_extent) return;  // wide wavefront: k_pairsum's exact path

  // lane l evaluates K for target (t*16 + l%16), source (4g + l/16)
  const int row = lane & 15, grp = lane >> 4;
  float xt[NT][D];
#pragma unroll
  for (int t = 0; t < NT; t++)
#pragma unroll
    for (int k = 0; k < D; k++) {
      const double v = __shfl((xd[k] - c[k]) * s, t * 16 + row, 64);
      xt[t][k] = (float)v;
    }

  floatx4 acc[NT][NB];
  double accd[NT][NB][4];
#pragma unroll
  for (int t = 0; t < NT; t++)
#pragma unroll
    for (int b = 0; b < NB; b++) {
      acc[t][b] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int v = 0; v < 4; v++) accd[t][b][v] = 0.0;
    }

  const long js = yb + (long)split * p.split_len;
  const long je = std::min(ye, js + p.split_len);
  const float far = 1.0e4f;
  for (long j0 = js; j0 < je; j0 += TS) {
    __syncthreads();
    // stage TS sources: each lane stages TS / 64 records
#pragma unroll
    for (int q = 0; q < TS / 64; q++) {
      const long j = j0 + q * 64 + lane;
      float *rec = sm + (q * 64 + lane) * REC;
      float y[D], w[DV];
      if (j < je) {
#pragma unroll
        for (int k = 0; k < D; k++) y[k] = (float)((p.y[j * D + k] - c[k]) * s);
#pragma unroll
        for (int k = 0; k < DV; k++) w[k] = (float)p.w[j * DV + k];
      } else {
#pragma unroll
        for (int k = 0; k < D; k++) y[k] = far;
#pragma unroll
        for (int k = 0; k < DV; k++) w[k] = 0.f;
      }
      float r[REC];
#pragma unroll
      for (int k = 0; k < REC; k++) r[k] = 0.f;
#pragma unroll
      for (int k = 0; k < D; k++) r[k] = y[k];
#pragma unroll
      for (int k = 0; k < DV; k++) {
        r[4 + k] = w[k];
#pragma unroll
        for (int l = 0; l < D; l++) r[4 + DV + k * D + l] = w[k] * y[l];
      }
#pragma unroll
      for (int k = 0; k < REC; k += 4)
        *reinterpret_cast<float4 *>(rec + k) = make_float4(r[k], r[k + 1], r[k + 2], r[k + 3]);
    }
    __syncthreads();

    const long left = je - j0;
    // whole groups of 4 sources; staged records past the end are far away with zero weight
    const int ngrp = (int)(left < TS ? (left + 3) / 4 : TS / 4);
    for (int g0 = 0; g0 < ngrp; g0 += S::FOLD / 4) {
      const int g1 = min(ngrp, g0 + S::FOLD / 4);
      for (int g = g0; g < g1; g++) {
        const float *rec = sm + (g * 4 + grp) * REC;
        const float4 yv = *reinterpret_cast<const float4 *>(rec);
        float bv[NB];
#pragma unroll
        for (int b = 0; b < NB; b++) bv[b] = rec[4 + b * 16 + row];
#pragma unroll
        for (int t = 0; t < NT; t++) {
          const float d0 = xt[t][0] - yv.x, d1 = xt[t][1] - yv.y, d2 = xt[t][2] - yv.z;
          const float K = Prec<float>::kexp(d0 * d0 + d1 * d1 + d2 * d2);
#pragma unroll
          for (int b = 0; b < NB; b++)
            acc[t][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(K, bv[b], acc[t][b], 0, 0, 0);
        }
      }
#pragma unroll
      for (int t = 0; t < NT; t++)
#pragma unroll
        for (int b = 0; b < NB; b++) {
#pragma unroll
          for (int v = 0; v < 4; v++) accd[t][b][v] += (double)acc[t][b][v];
          acc[t][b] = floatx4{0.f, 0.f, 0.f, 0.f};
        }
    }
  }

  // ---- per-target combination: D[row 4*(l/16)+v][col l%16] of each 16-column block ->
  // LDS, then lane (t, i) = (l/16, l%16) reads its target's 16 columns
  double *red = reinterpret_cast<double *>(sm);
  double mine[16 * NB];
#pragma unroll
  for (int b = 0; b < NB; b++) {
    __syncthreads();
#pragma unroll
    for (int t = 0; t < NT; t++)
#pragma unroll
      for (int v = 0; v < 4; v++) red[(t * 16 + grp * 4 + v) * 16 + row] = accd[t][b][v];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; k++) mine[b * 16 + k] = red[lane * 16 + k];  // target t*16 + i
  }
  if (!valid) return;
  double a[DV + D];
  // S_k = mine[k]; T_kl = mine[DV + k*D + l]; x in the same scaled centred FP32 coordinates
  float xs[D];
#pragma unroll
  for (int k = 0; k < D; k++) xs[k] = (float)((xd[k] - c[k]) * s);
#pragma unroll
  for (int k = 0; k < DV; k++) a[k] = mine[k];
#pragma unroll
  for (int l = 0; l < D; l++) a[DV + l] = 0.0;
#pragma unroll
  for (int k = 0; k < DV; k++) {
    const double al = p.ta[orig * DV + k];
#pragma unroll
    for (int l = 0; l < D; l++) a[DV + l] += al * ((double)xs[l] * mine[k] - mine[DV + k * D + l]);
  }
  constexpr int NACC = Shape<OP, D, DV>::NACC;
  double out[NACC];
  if constexpr (OP == OP_CONVGRAD1) {
#pragma unroll
    for (int k = 0; k < DV + D; k++) out[k] = a[k];
  } else {  // OP_GRAD1
#pragma unroll
    for (int l = 0; l < D; l++) out[l] = a[DV + l];
  }
  if (p.nsplit == 1) {
    finalize<OP, D, DV>(out, p, orig);
  } else {
    double *dst = p.partial + ((long)split * p.nx + pos) * NACC;
#pragma unroll
    for (int k = 0; k < NACC; k++) dst[k] = out[k];
  }
}

}  // namespace dfm
