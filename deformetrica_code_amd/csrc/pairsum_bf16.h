// pairsum_bf16.h — Convolve / ConvolveGradient(X, alpha) / the fused sum / the varifold's
// sums with the coordinate cross term of the distances on the matrix cores (precision
// DFM_BF16MFMA).
//
// For a group of 256 spatially sorted targets the workgroup re-centres targets and sources
// on the group's bounding-box centre (FP64), scales by sqrt(log2 e) / sigma and rounds to
// FP32: xf, yf.  Then
//   |xf - yf|^2 = |xf|^2 + r',   r' = |yf|^2 - 2 xf . yf,
// and r' for 16 sources x 16 targets is one v_mfma_f32_16x16x32_bf16: every FP32 value v
// is split into three BF16 parts v = v0 + v1 + v2 (8 + 8 + 8 significand bits, exact for
// an FP32 value), and the K = 32 dimension carries the split |yf|^2 (against ones) and the
// cross products y_a (-2 x_b) of part levels a + b <= 2 (six per coordinate; the dropped
// levels are below 2^-23 of |x||y|).  The slots are ordered by level, so the large
// leading terms (|y|^2_0 and y_0 x_0) meet in the first 8-wide group of K, where the
// hardware sums them before any FP32 rounding.  The target norm enters exactly: the sums
// run over K' = exp2(-r') and each target's results are multiplied by exp2(-|xf|^2) in
// FP64 at the end (expand_finish, shared with the FP32C expanded form).  The exp2 and every
// accumulation stay on the VALU in FP32, folded into FP64 once per 256-source tile.
//
// Layout: sources on the MFMA rows (A, lane l holds source row l%16, K slots 8(l/16)..+7),
// targets on the columns (B, fixed per wavefront), so lane l receives r' of its target
// column l%16 against sources 4(l/16)..4(l/16)+3 of each 16-source group and accumulates
// per target; the four lane groups' partial sums are added by a reduce-scatter (fold4).
// A wavefront holds 4 tiles of 16 targets; four wavefronts share the workgroup's 256
// targets and one LDS tile of 256 staged sources (A rows 64 B, FP32 y and w).
//
// A group whose targets span more than kBfExtent scaled units (kBfExtentBatch in ragged
// batches of several segments; none of the Hilbert-ordered groups of 1M uniform points at
// sigma 0.05; 12% of those of 1M samples of a sphere at sigma 0.02) runs the same sweep on
// one centre per wavefront when each wavefront's 64 targets fit the bound (each wavefront
// staging the sources on its own centre, 64 to a tile), else the DFM_FP32C wavefront code
// (sum_wavefront) in the same workgroup, one wavefront per 64 targets.
// Ragged batches (grid.z = segment) keep the caller's target order: the groups of a
// triangle mesh's faces in mesh order are as compact as Hilbert-ordered ones.  Same
// accumulator layout, split partials and finalize as k_pairsum: k_reduce_splits is shared.
//
// The reference arithmetic this replaces: the per-pair squared distance of
// ScalarRadialKernel.h:61-73, tiled over threads in GpuConv1D.cu:55-78.
#pragma once

#include "pairsum_kernels.h"

namespace dfm {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));


// K-slot contents: kind 0 zero, 1 a part of |y|^2 (times 1), 2 a cross product y_yp (-2 x_xp)
// of coordinate k.  Levels 0, 1, 2 in order; within a level the |y|^2 part first, then the
// cross products by decreasing source part.
struct BfSlot {
  int kind, yp, xp, k;
};

template <int D>
__host__ __device__ constexpr BfSlot bf_slot(int s) {
  int i = 0;
  for (int lvl = 0; lvl < 3; lvl++) {
    if (i++ == s) return BfSlot{1, lvl, 0, 0};
    for (int yp = lvl; yp >= 0; yp--)
      for (int k = 0; k < D; k++)
        if (i++ == s) return BfSlot{2, yp, lvl - yp, k};
  }
  return BfSlot{0, 0, 0, 0};
}
static_assert(3 + 6 * 3 <= 32, "the split terms fit the K dimension");
static_assert(6 * 6 <= 32 + 8, "the dot product's slots fit two MFMAs, the second in slots 32..39");

__device__ __forceinline__ void bf_split3(float v, __bf16 (&o)[3]) {
  // v = o0 + o1 + o2 exactly for an FP32 v: each residual of a round-to-nearest is exact
  o[0] = (__bf16)v;
  float r = v - (float)o[0];
  o[1] = (__bf16)r;
  r = r - (float)o[1];
  o[2] = (__bf16)r;
}

// v + (v of lane l ^ 32) + (v of lanes l ^ 16, l ^ 48): the sum over the four lane groups
// l / 16, in every lane (gfx950 permlane swaps: rows of 16 lanes exchanged across halves)
__device__ __forceinline__ float sum_lane_groups(float v) {
  const auto a = __builtin_amdgcn_permlane32_swap(__float_as_uint(v), __float_as_uint(v), false, false);
  v = __uint_as_float(a[0]) + __uint_as_float(a[1]);
  const auto b = __builtin_amdgcn_permlane16_swap(__float_as_uint(v), __float_as_uint(v), false, false);
  return __uint_as_float(b[0]) + __uint_as_float(b[1]);
}

// Four tiles' partial sums t0..t3, each summed over the four lane groups, with lane group h
// keeping tile h: a reduce-scatter of two permlane32 swaps (lanes 32..63 of the first operand
// trade places with lanes 0..31 of the second) and one permlane16 swap (odd rows of 16 lanes
// of the first operand with even rows of the second).  Every value is the sum that
// sum_lane_groups forms, (v_h + v_h^2) + (v_h^1 + v_h^3), at 3 swaps and 3 adds for four
// tiles instead of 8 and 8.
__device__ __forceinline__ float fold4(float t0, float t1, float t2, float t3) {
  const auto a = __builtin_amdgcn_permlane32_swap(__float_as_uint(t0), __float_as_uint(t2), false, false);
  const float u0 = __uint_as_float(a[0]) + __uint_as_float(a[1]);  // lanes 0-31 tile 0, 32-63 tile 2
  const auto b = __builtin_amdgcn_permlane32_swap(__float_as_uint(t1), __float_as_uint(t3), false, false);
  const float u1 = __uint_as_float(b[0]) + __uint_as_float(b[1]);  // lanes 0-31 tile 1, 32-63 tile 3
  const auto c = __builtin_amdgcn_permlane16_swap(__float_as_uint(u0), __float_as_uint(u1), false, false);
  return __uint_as_float(c[0]) + __uint_as_float(c[1]);  // row h of 16 lanes: tile h
}

// K slots of the weight dot product a_i . w_j (GRAD1, CONVGRAD1): the products w_wp a_ap of
// part levels wp + ap <= 2 for each weight component k, by level: 6 DV slots, one MFMA for
// DV <= 5, two chained MFMAs for DV = 6 (slots 32..35 in the second).
template <int DV>
__host__ __device__ constexpr BfSlot bf_dot_slot(int s) {
  int i = 0;
  for (int lvl = 0; lvl < 3; lvl++)
    for (int wp = lvl; wp >= 0; wp--)
      for (int k = 0; k < DV; k++)
        if (i++ == s) return BfSlot{2, wp, lvl - wp, k};
  return BfSlot{0, 0, 0, 0};
}

// One 16 x 16 x 32 BF16 product of the A rows of this lane's 16-source group (LDS) with a
// fixed B column.
__device__ __forceinline__ floatx4 bf_mfma(const bf16x8 &a, const bf16x8 &b) {
  const floatx4 zero4 = {0.f, 0.f, 0.f, 0.f};
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, zero4, 0, 0, 0);
}

// The same, added to c.
__device__ __forceinline__ floatx4 bf_mfma_acc(const bf16x8 &a, const bf16x8 &b, const floatx4 &c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// Lane (h, n)'s eight B values of a column, slots 8h .. 8h + 7.
__device__ __forceinline__ bf16x8 bf_column(const __bf16 (&col)[32], int h) {
  bf16x8 b;
#pragma unroll
  for (int v8 = 0; v8 < 8; v8++)
    b[v8] = h == 0 ? col[v8] : h == 1 ? col[8 + v8] : h == 2 ? col[16 + v8] : col[24 + v8];
  return b;
}

// Three wavefronts per SIMD: at most 168 VGPRs.
template <int OP, int D, int DV, int NT>
__global__ __launch_bounds__(256, 3) void k_pairsum_bf(KParams p) {
  using S = Shape<OP, D, DV>;
  constexpr int NACC = S::NACC;
  constexpr int NA = S::NACCMAX;
  constexpr int NU = S::NACCE;  // sums of the expanded form in use
  constexpr int NV = D + DV;    // FP32 values per staged source (y, w)
  constexpr int TS = kBfTile;
  constexpr int NTE = (OP == OP_CONV) ? 1 : DV;
  // the weight dot product a_i . w_j on the matrix cores too (six slots per component, one
  // MFMA up to DV = 5, two from DV = 6)
  constexpr bool DOT = OP != OP_CONV;
  constexpr int NDOT = 6 * DV <= 32 ? 1 : 2;
  constexpr int GROUP = 64 * NT;  // four wavefronts of NT tiles of 16 targets
  static_assert(NT >= 1 && NT <= 4, "a lane group of 16 lanes finishes each tile");
  static_assert(S::EXPANDABLE, "Convolve, Grad1 and the fused sum");

  // A rows by K-slot group: smA[q * TS + j] holds slots 8q .. 8q + 7 of staged source j, so
  // the 16 lanes of one slot group read 16 consecutive 16-byte rows (no bank conflicts)
  __shared__ __attribute__((aligned(16))) bf16x8 smA[TS * 4];  // distance rows: 4 x 8 BF16
  __shared__ __attribute__((aligned(16))) bf16x8 smW[DOT ? TS * 4 : 4];  // dot-product rows
  // slots 32..39 of the dot-product rows (DV = 6)
  __shared__ __attribute__((aligned(16))) bf16x8 smW2[NDOT == 2 ? TS : 1];
  __shared__ __attribute__((aligned(16))) float smV[TS * NV];
  __shared__ double sbox[4][2][3];
  __shared__ int sok[4];

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int split = blockIdx.y;
  // segment of a ragged batch (blockIdx.z; one segment: the whole target and source sets)
  long xb = 0, xe = p.nx, yb = 0, ye = p.ny;
  if (p.xoff) {
    xb = p.xoff[blockIdx.z];
    xe = p.xoff[blockIdx.z + 1];
  }
  if (p.yoff) {
    yb = p.yoff[blockIdx.z];
    ye = p.yoff[blockIdx.z + 1];
  }
  const long gbase = xb + (long)blockIdx.x * GROUP;
  const long gend = std::min(xe, gbase + GROUP);
  if (gbase >= xe) return;  // the whole workgroup lies past its segment

  // ---- the group's bounding box (FP64, exact min / max) and centre
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  {
    const long pos = gbase + tid;
    if (pos < gend) {
      const long o = p.perm ? (long)p.perm[pos] : pos;
#pragma unroll
      for (int k = 0; k < D; k++) lo[k] = hi[k] = p.x[o * D + k];
    }
#pragma unroll
    for (int k = 0; k < D; k++) {
      lo[k] = wave_min(lo[k]);
      hi[k] = wave_max(hi[k]);
    }
    if (lane == 0)
#pragma unroll
      for (int k = 0; k < D; k++) {
        sbox[wave][0][k] = lo[k];
        sbox[wave][1][k] = hi[k];
      }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < D; k++) {
      lo[k] = fmin(fmin(sbox[0][0][k], sbox[1][0][k]), fmin(sbox[2][0][k], sbox[3][0][k]));
      hi[k] = fmax(fmax(sbox[0][1][k], sbox[1][1][k]), fmax(sbox[2][1][k], sbox[3][1][k]));
    }
  }
  double c[D];
  double extent = 0.0;
#pragma unroll
  for (int k = 0; k < D; k++) {
    c[k] = 0.5 * (lo[k] + hi[k]);
    extent = fmax(extent, hi[k] - lo[k]);
  }
  const double s = p.s;
  // A group too wide for one centre: if each of its four wavefronts is compact on a centre
  // of its own, every wavefront runs the matrix-core sweep on that centre over tiles of 64
  // sources that it stages itself (into its own rows of the tile arrays); otherwise the
  // FP32C code below.  Either way the choice is the workgroup's (one barrier sequence).
  bool per_wave = false;
  if (!(extent * s <= p.exact_extent_bf)) {
    double wlo[3] = {INFINITY, INFINITY, INFINITY}, whi[3] = {-INFINITY, -INFINITY, -INFINITY};
    const long wp = gbase + 64 * wave + lane;
    if (wp < gend) {
      const long o = p.perm ? (long)p.perm[wp] : wp;
#pragma unroll
      for (int k = 0; k < D; k++) wlo[k] = whi[k] = p.x[o * D + k];
    }
    double wext = 0.0;
#pragma unroll
    for (int k = 0; k < D; k++) {
      wlo[k] = wave_min(wlo[k]);
      whi[k] = wave_max(whi[k]);
      if (wlo[k] <= whi[k]) wext = fmax(wext, whi[k] - wlo[k]);
    }
    if (lane == 0) sok[wave] = wext * s <= p.exact_extent_bf;
    __syncthreads();
    per_wave = sok[0] && sok[1] && sok[2] && sok[3];
    if (per_wave) {
#pragma unroll
      for (int k = 0; k < D; k++) c[k] = (wlo[k] <= whi[k]) ? 0.5 * (wlo[k] + whi[k]) : 0.0;
    }
  }
  if (!per_wave && !(extent * s <= p.exact_extent_bf)) {
    // a wide group (its sorted targets straddle a fold of the Hilbert curve): each wavefront
    // sums its 64 targets as DFM_FP32C does, on its own centre (exact differences where its
    // own targets are wide), with a slice of smA as its LDS tile.  All four wavefronts pass
    // the same barriers: the same source range, every lane taking part.
    constexpr int NREC = Shape<OP, D, DV>::NSPX;
    static_assert(4 * kTileF32 * NREC * sizeof(float) <= sizeof(smA), "tiles fit smA");
    long wpos[1], worig[1];
    bool wvalid[1];
    double wxd[1][D], wc[D], wext;
    load_targets<float, D, 1>(p, lane, gbase + 64 * wave, gend, wpos, worig, wvalid, wxd, wc,
                              wext);
    const long wjs = yb + (long)split * p.split_len;
    const long wje = std::min(ye, wjs + p.split_len);
    float *wsm = reinterpret_cast<float *>(smA) + wave * (kTileF32 * NREC);
    // (a fused integrator step never runs here: it is applied by k_step_update)
    sum_wavefront<float, OP, D, DV, 0, 1, MODE_ADAPTIVE>(p, wsm, lane, split, wjs, wje, wpos, worig,
                                                        wvalid, wxd, wc, wext, -1, 0);
    return;
  }

  // ---- this wavefront's targets: tile t, column n -> position wbase + 16 t + n
  const int h = lane >> 4, n = lane & 15;
  const long wbase = gbase + 16 * NT * wave;
  bf16x8 bop[NT];               // distance columns
  bf16x8 bdot[DOT ? NT : 1][NDOT];  // dot-product columns (slots 0..31, 32..63)
#pragma unroll
  for (int t = 0; t < NT; t++) {
    const long pos = wbase + 16 * t + n;
    const bool v = pos < gend;
    const long o = v ? (p.perm ? (long)p.perm[pos] : pos) : 0;
    __bf16 xp[D][3];
#pragma unroll
    for (int k = 0; k < D; k++)
      bf_split3(v ? (float)((p.x[o * D + k] - c[k]) * s) : 0.f, xp[k]);
    __bf16 col[32];
#pragma unroll
    for (int q = 0; q < 32; q++) {
      const BfSlot sl = bf_slot<D>(q);
      col[q] = sl.kind == 1   ? (__bf16)1.0f
               : sl.kind == 2 ? (__bf16)(-2.0f * (float)xp[sl.k][sl.xp])
                              : (__bf16)0.0f;
    }
    bop[t] = bf_column(col, h);
    if constexpr (DOT) {
      __bf16 ap[DV][3];
#pragma unroll
      for (int k = 0; k < DV; k++) bf_split3(v ? (float)p.ta[o * DV + k] : 0.f, ap[k]);
#pragma unroll
      for (int m = 0; m < NDOT; m++) {
#pragma unroll
        for (int q = 0; q < 32; q++) {
          const BfSlot sl = bf_dot_slot<DV>(32 * m + q);
          col[q] = sl.kind == 2 ? ap[sl.k][sl.xp] : (__bf16)0.0f;
        }
        bdot[t][m] = bf_column(col, h);
      }
    }
  }

  // FP64 sums of this lane's own target: tile h, column n (the tiles' FP32 partials are
  // summed over the four lane groups before each fold)
  double accd[NA];
#pragma unroll
  for (int a = 0; a < NA; a++) accd[a] = 0.0;

  const long js = yb + (long)split * p.split_len;
  const long je = std::min(ye, js + p.split_len);
  // PW: tiles of 64 sources, each wavefront staging the same sources on its own centre into
  // its rows jb .. jb + 63 (row tid in both forms)
  auto sweep = [&](auto pw) {
  constexpr bool PW = decltype(pw)::value;
  constexpr int TSL = PW ? 64 : TS;
  const int jb = PW ? 64 * wave : 0;
  // FP32 partial sums over 512 sources (two 256-source tiles) between folds into the FP64
  // accumulators: the fused sum at N = M = 1M ran 239.6 against 244.1 ms folding every
  // tile, norm-wise errors 1.10e-7 / 1.74e-7 against 1.07e-7 / 1.68e-7 vs FP64.  The
  // per-wavefront sweep folds every 64-source tile: over 512 sources it put the 64-subject
  // atlas's template gradient at 9.7e-7 of FP64 against 3.8e-7.
  constexpr int FOLD_TILES = PW ? 1 : 2;
  float acc[NT][NU];
  int tcount = 0;
  for (long j0 = js; j0 < je; j0 += TSL) {
    __syncthreads();
    // ---- stage source j0 + tid (j0 + lane per wavefront): A rows (BF16 splits in slot
    // order) and FP32 y, w
    {
      const long j = j0 + (PW ? lane : tid);
      float yf[D], w[DV];
      double n2 = 0.0;
      if (j < je) {
#pragma unroll
        for (int k = 0; k < D; k++) {
          yf[k] = (float)((p.y[j * D + k] - c[k]) * s);
          n2 += (double)yf[k] * (double)yf[k];
        }
#pragma unroll
        for (int k = 0; k < DV; k++) w[k] = (float)p.w[j * DV + k];
      } else {  // padding: zero weight, so K' = 1 contributes nothing
#pragma unroll
        for (int k = 0; k < D; k++) yf[k] = 0.f;
#pragma unroll
        for (int k = 0; k < DV; k++) w[k] = 0.f;
      }
      __bf16 yp[D][3], np[3];
#pragma unroll
      for (int k = 0; k < D; k++) bf_split3(yf[k], yp[k]);
      bf_split3((float)n2, np);
#pragma unroll
      for (int q4 = 0; q4 < 4; q4++) {
        bf16x8 row;
#pragma unroll
        for (int v8 = 0; v8 < 8; v8++) {
          const BfSlot sl = bf_slot<D>(q4 * 8 + v8);
          row[v8] = sl.kind == 1 ? np[sl.yp] : sl.kind == 2 ? yp[sl.k][sl.yp] : (__bf16)0.0f;
        }
        smA[q4 * TS + tid] = row;
      }
      if constexpr (DOT) {
        __bf16 wp[DV][3];
#pragma unroll
        for (int k = 0; k < DV; k++) bf_split3(w[k], wp[k]);
#pragma unroll
        for (int q4 = 0; q4 < 4; q4++) {
          bf16x8 row;
#pragma unroll
          for (int v8 = 0; v8 < 8; v8++) {
            const BfSlot sl = bf_dot_slot<DV>(q4 * 8 + v8);
            row[v8] = sl.kind == 2 ? wp[sl.k][sl.yp] : (__bf16)0.0f;
          }
          smW[q4 * TS + tid] = row;
        }
        if constexpr (NDOT == 2) {
          bf16x8 row;
#pragma unroll
          for (int v8 = 0; v8 < 8; v8++) {
            const BfSlot sl = bf_dot_slot<DV>(32 + v8);
            row[v8] = sl.kind == 2 ? wp[sl.k][sl.yp] : (__bf16)0.0f;
          }
          smW2[tid] = row;
        }
      }
      float *dv = smV + tid * NV;
#pragma unroll
      for (int k = 0; k < D; k++) dv[k] = yf[k];
#pragma unroll
      for (int k = 0; k < DV; k++) dv[D + k] = w[k];
    }
    __syncthreads();

    const long left = je - j0;
    const int ngrp = (int)(left < TSL ? (left + 15) / 16 : TSL / 16);
    // 16 sources of the tile; the first group after a fold (FIRST) sets the partial sums
    // instead of adding to zero (the same values: an FMA into 0 rounds as the product does)
    auto group = [&](int g, auto first) {
      constexpr bool FIRST = decltype(first)::value;
      // r' of sources 16 g + 4 h + v against each tile's column n (and a . w)
      const bf16x8 a = smA[h * TS + jb + 16 * g + n];
      floatx4 r[NT];
#pragma unroll
      for (int t = 0; t < NT; t++) r[t] = bf_mfma(a, bop[t]);
      floatx4 dt[DOT ? NT : 1];
      if constexpr (DOT) {
        const bf16x8 aw = smW[h * TS + jb + 16 * g + n];
#pragma unroll
        for (int t = 0; t < NT; t++) dt[t] = bf_mfma(aw, bdot[t][0]);
        if constexpr (NDOT == 2) {
          // slots 32..39 (every lane reads them; the B columns of lane groups h > 0 are zero)
          const bf16x8 aw2 = smW2[jb + 16 * g + n];
#pragma unroll
          for (int t = 0; t < NT; t++) dt[t] = bf_mfma_acc(aw2, bdot[t][1], dt[t]);
        }
      }
      // FP32 y, w of sources 16 g + 4 h + v, v = 0..3 (4 NV consecutive floats)
      float sv[4][NV];
      const float4 *src4 = reinterpret_cast<const float4 *>(smV + (jb + 16 * g + 4 * h) * NV);
#pragma unroll
      for (int q = 0; q < NV; q++) {
        const float4 f4 = src4[q];
        const float e4[4] = {f4.x, f4.y, f4.z, f4.w};
#pragma unroll
        for (int u = 0; u < 4; u++) sv[(4 * q + u) / NV][(4 * q + u) % NV] = e4[u];
      }
#pragma unroll
      for (int t = 0; t < NT; t++)
#pragma unroll
        for (int v4 = 0; v4 < 4; v4++) {
          const bool set = FIRST && v4 == 0;
          const float K = Prec<float>::kexp(r[t][v4]);
          const float *y = sv[v4];
          const float *w = sv[v4] + D;
          float kc = K;  // the Convolve weight: K, or K (u_i . u_j) for the varifold
          if constexpr (OP == OP_VARIFOLD || OP == OP_VARIFOLD_CONV) kc = K * dt[t][v4];
          if constexpr (OP != OP_GRAD1) {
#pragma unroll
            for (int k = 0; k < DV; k++) acc[t][k] = set ? kc * w[k] : acc[t][k] + kc * w[k];
          }
          if constexpr (OP != OP_CONV && OP != OP_VARIFOLD_CONV) {
            constexpr int G = (OP == OP_GRAD1) ? 0 : DV;
            const float tt = kc * dt[t][v4];
            if constexpr (OP == OP_GRAD1) acc[t][D] = set ? tt : acc[t][D] + tt;
#pragma unroll
            for (int l = 0; l < D; l++) acc[t][G + l] = set ? tt * y[l] : acc[t][G + l] + tt * y[l];
          }
        }
    };
    if (tcount % FOLD_TILES == 0)
      group(0, std::true_type{});
    else
      group(0, std::false_type{});
    for (int g = 1; g < ngrp; g++) group(g, std::false_type{});
    const bool fold = tcount % FOLD_TILES == FOLD_TILES - 1 || j0 + TSL >= je;
    tcount++;
    if (!fold) continue;
    // fold: sum each target tile's partials over the lane groups (sources 4h..4h+3 of every
    // group); lane (h, n) keeps target tile h
#pragma unroll
    for (int a = 0; a < NU; a++) {
      float mine = 0.f;
      if constexpr (NT == 4) {
        mine = fold4(acc[0][a], acc[1][a], acc[2][a], acc[3][a]);
      } else {
#pragma unroll
        for (int t = 0; t < NT; t++) {
          const float v = sum_lane_groups(acc[t][a]);
          mine = (t == h) ? v : mine;
        }
      }
      accd[a] += (double)mine;
    }
  }
  };
  if (per_wave)
    sweep(std::true_type{});
  else
    sweep(std::false_type{});

  // ---- lane (h, n) finishes and stores target tile h, column n (lane groups h >= NT hold
  // no tile)
  const long pos = wbase + 16 * h + n;
  if (h >= NT || pos >= gend) return;
  const long o = p.perm ? (long)p.perm[pos] : pos;
  float xm[D], am[NTE];
#pragma unroll
  for (int k = 0; k < D; k++) xm[k] = (float)((p.x[o * D + k] - c[k]) * s);
#pragma unroll
  for (int k = 0; k < NTE; k++) am[k] = OP == OP_CONV ? 0.f : (float)p.ta[o * DV + k];
  double mine[NA];
#pragma unroll
  for (int a = 0; a < NA; a++) mine[a] = accd[a];
  expand_finish<OP, D, DV, NA>(mine, xm, am);
  if (p.nsplit == 1) {
    finalize<OP, D, DV>(mine, p, o);
  } else {
    double *dst = p.partial + ((long)split * p.nx + pos) * NACC;
#pragma unroll
    for (int a = 0; a < NACC; a++) dst[a] = mine[a];
  }
}

}  // namespace dfm
