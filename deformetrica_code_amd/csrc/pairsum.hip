// pairsum.hip — Gaussian kernel sums on gfx950 (MI355X).
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
//
// FP64 path: coordinates scaled by 1/sigma, K = exp(-r^2) in FP64 — the arithmetic of the
// reference's ExactKernel.
// FP32C path: targets are ordered along a Morton curve (radix sort of 30-bit keys) so
// each wavefront's targets are spatially compact; each wavefront re-centres its targets
// and every staged source on the centre of its targets' bounding box in FP64, scales by
// sqrt(log2 e)/sigma and only then rounds to FP32, so the FP32 differences carry the
// rounding of the small centred coordinates rather than of the absolute ones.
// K = exp2(-r'^2) on the transcendental unit (v_exp_f32).  Each lane accumulates one
// source tile in FP32 and folds the tile sum into FP64 accumulators (cascaded,
// compensated summation: the FP32 rounding is bounded by the tile's partial sum, never by
// the running total).
//
// Work split: grid.x = target chunks of R*64, grid.y = source splits.  With one split the
// kernel writes the scaled results straight to the caller's output (through the sort
// permutation); with several it writes FP64 partials that k_reduce_splits sums in split
// order, so the result is deterministic.
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>
#include <utility>
#include <vector>

#include "dfm_hip.h"
#include "dfm_internal.h"

namespace dfm {

// ------------------------------------------------------------------------------------
// error state

static thread_local std::string g_last_error;

int set_error(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

const char *last_error() { return g_last_error.c_str(); }
void clear_error() { g_last_error.clear(); }

// ------------------------------------------------------------------------------------
// per-thread kernel timing: HIP events on the launch stream around the pair-sum kernel

struct ProfileState {
  bool on = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
};
static thread_local ProfileState g_profile;

void profile_enable(bool on) { g_profile.on = on; }

int profile_read(double *total_ms, long *launches) {
  double tot = 0.0;
  long cnt = 0;
  int rc = DFM_OK;
  for (auto &e : g_profile.events) {
    float ms = 0.f;
    if (rc == DFM_OK) {
      hipError_t err = hipEventSynchronize(e.second);
      if (err == hipSuccess) err = hipEventElapsedTime(&ms, e.first, e.second);
      if (err != hipSuccess) rc = set_error(DFM_ERR_HIP, "event timing: %s", hipGetErrorString(err));
    }
    tot += ms;
    cnt++;
    (void)hipEventDestroy(e.first);
    (void)hipEventDestroy(e.second);
  }
  g_profile.events.clear();
  if (total_ms) *total_ms = tot;
  if (launches) *launches = cnt;
  return rc;
}

// ------------------------------------------------------------------------------------
// precision traits

template <typename T>
struct Prec;

template <>
struct Prec<float> {
  static constexpr bool kCenter = true;
  static constexpr int kTile = 256;
  __device__ __forceinline__ static float kexp(float r2) { return __builtin_amdgcn_exp2f(-r2); }
  static double scale(double sigma) { return std::sqrt(1.4426950408889634074) / sigma; }
};

template <>
struct Prec<double> {
  static constexpr bool kCenter = false;
  static constexpr int kTile = 128;
  __device__ __forceinline__ static double kexp(double r2) { return exp(-r2); }
  static double scale(double sigma) { return 1.0 / sigma; }
};

// ------------------------------------------------------------------------------------
// op shapes

template <int OP, int D, int DV>
struct Shape {
  // per-target extra inputs: alpha (GRAD1, CONVGRAD1) or beta + eta (GRADDIFF)
  static constexpr int NTE = (OP == OP_GRAD1 || OP == OP_CONVGRAD1) ? DV
                             : (OP == OP_GRADDIFF)                  ? DV + D
                                                                    : 0;
  // per-source record: y, w (+ eta_j for GRADDIFF)
  static constexpr int NSRC = D + DV + (OP == OP_GRADDIFF ? D : 0);
  static constexpr int NSP = (NSRC + 3) / 4 * 4;
  static constexpr int NACC = OP == OP_CONV        ? DV
                              : OP == OP_GRAD1     ? D
                              : OP == OP_CONVGRAD1 ? DV + D
                              : OP == OP_VARLIN    ? DV * D
                              : OP == OP_GRADDIM   ? DV
                                                   : 2 * D;
};

// ------------------------------------------------------------------------------------
// per-pair update.  xi: scaled (centred) target; te: per-target extras; rec: staged source
// record [y (D) | w (DV) | eta_j (D)]; acc: the op's FP accumulators.

template <int OP, int D, int DV, int AUX, typename T>
__device__ __forceinline__ void pair_update(const T *xi, const T *te, const T *rec, T *acc) {
  T d[D];
  T r2 = T(0);
#pragma unroll
  for (int k = 0; k < D; k++) {
    d[k] = xi[k] - rec[k];
    r2 += d[k] * d[k];
  }
  const T K = Prec<T>::kexp(r2);
  const T *w = rec + D;
  if constexpr (OP == OP_CONV) {
#pragma unroll
    for (int k = 0; k < DV; k++) acc[k] += K * w[k];
  } else if constexpr (OP == OP_GRAD1 || OP == OP_CONVGRAD1) {
    constexpr int G = (OP == OP_CONVGRAD1) ? DV : 0;
    if constexpr (OP == OP_CONVGRAD1) {
#pragma unroll
      for (int k = 0; k < DV; k++) acc[k] += K * w[k];
    }
    T sa = te[0] * w[0];
#pragma unroll
    for (int k = 1; k < DV; k++) sa += te[k] * w[k];
    const T t = K * sa;
#pragma unroll
    for (int l = 0; l < D; l++) acc[G + l] += t * d[l];
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

// Scale the FP64 accumulators of one target and store them at original index `o`.
template <int OP, int D, int DV>
__device__ __forceinline__ void finalize(const double *a, const KParams &p, long o) {
  if constexpr (OP == OP_CONV) {
#pragma unroll
    for (int k = 0; k < DV; k++) p.out0[o * DV + k] = a[k];
  } else if constexpr (OP == OP_GRAD1) {
#pragma unroll
    for (int l = 0; l < D; l++) p.out0[o * D + l] = p.cg * a[l];
  } else if constexpr (OP == OP_CONVGRAD1) {
    if (p.out0) {
#pragma unroll
      for (int k = 0; k < DV; k++) p.out0[o * DV + k] = a[k];
    }
    if (p.out1) {
#pragma unroll
      for (int l = 0; l < D; l++) p.out1[o * D + l] = p.cg * a[DV + l];
    }
  } else if constexpr (OP == OP_VARLIN) {
#pragma unroll
    for (int k = 0; k < DV * D; k++) p.out0[o * DV * D + k] = p.cg * a[k];
  } else if constexpr (OP == OP_GRADDIM) {
#pragma unroll
    for (int k = 0; k < DV; k++) p.out0[o * DV + k] = p.cg * a[k];
  } else {
#pragma unroll
    for (int k = 0; k < D; k++) p.out0[o * D + k] = p.ch1 * a[k] + p.ch2 * a[D + k];
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

template <typename T, int OP, int D, int DV, int AUX, int R>
__global__ __launch_bounds__(64) void k_pairsum(KParams p) {
  using S = Shape<OP, D, DV>;
  constexpr int NSP = S::NSP;
  constexpr int NACC = S::NACC;
  constexpr int NTE = S::NTE > 0 ? S::NTE : 1;
  constexpr int TS = Prec<T>::kTile;
  constexpr int REC_PER_LANE = TS / 64;
  static_assert(NSP % 4 == 0, "records are whole 16-byte groups");

  __shared__ T sm[TS * NSP];

  const int lane = threadIdx.x;
  const long base = (long)blockIdx.x * (64 * R);
  const int split = blockIdx.y;

  // ---- targets -> registers
  long pos[R];
  long orig[R];
  bool valid[R];
  double xd[R][D];
#pragma unroll
  for (int r = 0; r < R; r++) {
    pos[r] = base + r * 64 + lane;
    valid[r] = pos[r] < p.nx;
    orig[r] = valid[r] ? (p.perm ? (long)p.perm[pos[r]] : pos[r]) : 0;
#pragma unroll
    for (int k = 0; k < D; k++) xd[r][k] = valid[r] ? p.x[orig[r] * D + k] : 0.0;
  }

  double c[D];
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
    }
  }

  const double s = p.s;
  T xt[R][D];
  T te[R][NTE];
#pragma unroll
  for (int r = 0; r < R; r++) {
#pragma unroll
    for (int k = 0; k < D; k++) xt[r][k] = (T)((xd[r][k] - c[k]) * s);
    if constexpr (OP == OP_GRAD1 || OP == OP_CONVGRAD1) {
#pragma unroll
      for (int k = 0; k < DV; k++) te[r][k] = valid[r] ? (T)p.ta[orig[r] * DV + k] : T(0);
    } else if constexpr (OP == OP_GRADDIFF) {
#pragma unroll
      for (int k = 0; k < DV; k++) te[r][k] = valid[r] ? (T)p.ta[orig[r] * DV + k] : T(0);
#pragma unroll
      for (int k = 0; k < D; k++) te[r][DV + k] = valid[r] ? (T)p.tb[orig[r] * D + k] : T(0);
    } else {
      te[r][0] = T(0);
    }
  }

  T acc[R][NACC];
  double accd[R][NACC];
#pragma unroll
  for (int r = 0; r < R; r++)
#pragma unroll
    for (int a = 0; a < NACC; a++) {
      acc[r][a] = T(0);
      accd[r][a] = 0.0;
    }

  // ---- sources, one LDS tile at a time
  const long js = (long)split * p.split_len;
  const long je = std::min(p.ny, js + p.split_len);
  // a padded source sits far outside the support of K: exp(-|x - far|^2) is exactly 0
  const T far = T(1.0e4);

  for (long j0 = js; j0 < je; j0 += TS) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < REC_PER_LANE; q++) {
      const long j = j0 + q * 64 + lane;
      T rec[NSP];
#pragma unroll
      for (int k = 0; k < NSP; k++) rec[k] = T(0);
      if (j < je) {
#pragma unroll
        for (int k = 0; k < D; k++) rec[k] = (T)((p.y[j * D + k] - c[k]) * s);
#pragma unroll
        for (int k = 0; k < DV; k++) rec[D + k] = (T)p.w[j * DV + k];
        if constexpr (OP == OP_GRADDIFF) {
#pragma unroll
          for (int k = 0; k < D; k++) rec[D + DV + k] = (T)p.se[j * D + k];
        }
      } else {
#pragma unroll
        for (int k = 0; k < D; k++) rec[k] = far;
      }
      T *dst = sm + (q * 64 + lane) * NSP;
#pragma unroll
      for (int k = 0; k < NSP; k++) dst[k] = rec[k];
    }
    __syncthreads();

    const long left = je - j0;
    const int cnt = (int)(left < TS ? ((left + 3) & ~3L) : TS);  // padded records are zero
    for (int jj = 0; jj < cnt; jj += 4) {
#pragma unroll
      for (int u = 0; u < 4; u++) {
        T rec[NSP];
        const T *srcp = sm + (jj + u) * NSP;
#pragma unroll
        for (int k = 0; k < NSP; k++) rec[k] = srcp[k];
#pragma unroll
        for (int r = 0; r < R; r++) pair_update<OP, D, DV, AUX, T>(xt[r], te[r], rec, acc[r]);
      }
    }
#pragma unroll
    for (int r = 0; r < R; r++)
#pragma unroll
      for (int a = 0; a < NACC; a++) {
        accd[r][a] += (double)acc[r][a];
        acc[r][a] = T(0);
      }
  }

  // ---- results
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
// Morton ordering of the targets (FP32C)

__global__ __launch_bounds__(256) void k_bbox_partial(const double *x, long n, int D,
                                                      double *part) {
  __shared__ double slo[3][256];
  __shared__ double shi[3][256];
  double lo[3] = {INFINITY, INFINITY, INFINITY};
  double hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x)
    for (int k = 0; k < D; k++) {
      const double v = x[i * D + k];
      lo[k] = fmin(lo[k], v);
      hi[k] = fmax(hi[k], v);
    }
  for (int k = 0; k < 3; k++) {
    slo[k][threadIdx.x] = lo[k];
    shi[k][threadIdx.x] = hi[k];
  }
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st)
      for (int k = 0; k < 3; k++) {
        slo[k][threadIdx.x] = fmin(slo[k][threadIdx.x], slo[k][threadIdx.x + st]);
        shi[k][threadIdx.x] = fmax(shi[k][threadIdx.x], shi[k][threadIdx.x + st]);
      }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int k = 0; k < 3; k++) {
      part[blockIdx.x * 6 + k] = slo[k][0];
      part[blockIdx.x * 6 + 3 + k] = shi[k][0];
    }
}

__device__ __forceinline__ uint32_t spread3(uint32_t v) {  // 10 bits -> every third bit
  v &= 0x3ffu;
  v = (v | (v << 16)) & 0x030000FFu;
  v = (v | (v << 8)) & 0x0300F00Fu;
  v = (v | (v << 4)) & 0x030C30C3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}
__device__ __forceinline__ uint32_t spread2(uint32_t v) {  // 15 bits -> every other bit
  v &= 0x7fffu;
  v = (v | (v << 8)) & 0x00FF00FFu;
  v = (v | (v << 4)) & 0x0F0F0F0Fu;
  v = (v | (v << 2)) & 0x33333333u;
  v = (v | (v << 1)) & 0x55555555u;
  return v;
}

__global__ __launch_bounds__(256) void k_morton_keys(const double *x, long n, int D,
                                                     const double *part, int nparts,
                                                     uint32_t *keys, uint32_t *vals) {
  double lo[3] = {INFINITY, INFINITY, INFINITY};
  double hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int b = 0; b < nparts; b++)
    for (int k = 0; k < 3; k++) {
      lo[k] = fmin(lo[k], part[b * 6 + k]);
      hi[k] = fmax(hi[k], part[b * 6 + 3 + k]);
    }
  double ext = 0.0;
  for (int k = 0; k < D; k++) ext = fmax(ext, hi[k] - lo[k]);
  const int bits = D == 3 ? 10 : 15;
  const double cells = (double)(1u << bits);
  const double inv = ext > 0.0 ? cells / ext : 0.0;
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t q[3] = {0, 0, 0};
  for (int k = 0; k < D; k++) {
    double f = (x[i * D + k] - lo[k]) * inv;
    f = fmin(fmax(f, 0.0), cells - 1.0);
    q[k] = (uint32_t)f;
  }
  keys[i] = D == 3 ? (spread3(q[0]) | (spread3(q[1]) << 1) | (spread3(q[2]) << 2))
                   : (spread2(q[0]) | (spread2(q[1]) << 1));
  vals[i] = (uint32_t)i;
}

__global__ __launch_bounds__(256) void k_fill(double *dst, long n, double v) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = v;
}

// ------------------------------------------------------------------------------------
// planning

static int targets_per_lane(int op, int prec, int D, int DV) {
  if (prec == DFM_FP64) return 1;
  switch (op) {
    case OP_VARLIN: return DV * D > 9 ? 1 : 2;
    case OP_GRADDIFF: return 2;
    default: return 4;
  }
}

static int tile_of(int prec) { return prec == DFM_FP64 ? Prec<double>::kTile : Prec<float>::kTile; }

static int nacc_of(int op, int D, int DV) {
  switch (op) {
    case OP_CONV: return DV;
    case OP_GRAD1: return D;
    case OP_CONVGRAD1: return DV + D;
    case OP_VARLIN: return DV * D;
    case OP_GRADDIM: return DV;
    default: return 2 * D;
  }
}

static size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

Plan make_plan(int op, int prec, int D, int DV, long nx, long ny) {
  Plan pl{};
  pl.R = targets_per_lane(op, prec, D, DV);
  pl.chunks = (nx + 64L * pl.R - 1) / (64L * pl.R);
  const int TS = tile_of(prec);
  const long tiles = (ny + TS - 1) / TS;
  // Aim for >= 4096 wavefronts (4 per SIMD on 256 CUs) and >= 4 source tiles per split.
  long want = pl.chunks > 0 ? (4096 + pl.chunks - 1) / pl.chunks : 1;
  long maxsplit = std::max(1L, tiles / 4);
  long ns = std::max(1L, std::min(want, maxsplit));
  ns = std::min(ns, 65535L);
  const long per = (tiles + ns - 1) / ns;
  pl.split_len = std::max(1L, per) * TS;
  pl.nsplit = (int)std::max(1L, (ny + pl.split_len - 1) / pl.split_len);
  size_t off = 0;
  if (prec == DFM_FP32C) {
    pl.off_bbox = off;
    off = align256(off + sizeof(double) * 6 * kBBoxParts);
    pl.off_keys_in = off;
    off = align256(off + sizeof(uint32_t) * nx);
    pl.off_keys_out = off;
    off = align256(off + sizeof(uint32_t) * nx);
    pl.off_vals_in = off;
    off = align256(off + sizeof(uint32_t) * nx);
    pl.off_perm = off;
    off = align256(off + sizeof(uint32_t) * nx);
    size_t tmp = 0;
    (void)rocprim::radix_sort_pairs((void *)nullptr, tmp, (uint32_t *)nullptr, (uint32_t *)nullptr,
                              (uint32_t *)nullptr, (uint32_t *)nullptr, (unsigned)nx, 0u,
                              30u, (hipStream_t)0);
    pl.sort_tmp_bytes = tmp;
    pl.off_sort_tmp = off;
    off = align256(off + tmp);
  }
  if (pl.nsplit > 1) {
    pl.off_partial = off;
    off = align256(off + sizeof(double) * (size_t)pl.nsplit * nx * nacc_of(op, D, DV));
  }
  pl.work_bytes = off;
  return pl;
}

// Sort targets along a Morton curve; writes perm into the workspace.
int sort_targets(hipStream_t st, const double *x, long nx, int D, char *work, const Plan &pl) {
  double *part = (double *)(work + pl.off_bbox);
  uint32_t *kin = (uint32_t *)(work + pl.off_keys_in);
  uint32_t *kout = (uint32_t *)(work + pl.off_keys_out);
  uint32_t *vin = (uint32_t *)(work + pl.off_vals_in);
  uint32_t *perm = (uint32_t *)(work + pl.off_perm);
  const int nparts = (int)std::min<long>(kBBoxParts, std::max(1L, (nx + 255) / 256));
  hipLaunchKernelGGL(k_bbox_partial, dim3(nparts), dim3(256), 0, st, x, nx, D, part);
  hipLaunchKernelGGL(k_morton_keys, dim3((unsigned)((nx + 255) / 256)), dim3(256), 0, st, x, nx,
                     D, part, nparts, kin, vin);
  size_t tmp = pl.sort_tmp_bytes;
  hipError_t e = rocprim::radix_sort_pairs((void *)(work + pl.off_sort_tmp), tmp, kin, kout, vin,
                                           perm, (unsigned)nx, 0u, 30u, st);
  if (e != hipSuccess) return set_error(DFM_ERR_HIP, "radix sort: %s", hipGetErrorString(e));
  return DFM_OK;
}

int fill(hipStream_t st, double *dst, long n, double v) {
  if (n <= 0 || !dst) return DFM_OK;
  hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, dst, n, v);
  return DFM_OK;
}

// ------------------------------------------------------------------------------------
// dispatch over (precision, op, D, DV, AUX)

template <typename T, int OP, int D, int DV, int AUX>
static void launch_one(hipStream_t st, const KParams &p, const Plan &pl) {
  dim3 grid((unsigned)pl.chunks, (unsigned)pl.nsplit);
  switch (pl.R) {
    case 1: hipLaunchKernelGGL((k_pairsum<T, OP, D, DV, AUX, 1>), grid, dim3(64), 0, st, p); break;
    case 2: hipLaunchKernelGGL((k_pairsum<T, OP, D, DV, AUX, 2>), grid, dim3(64), 0, st, p); break;
    default: hipLaunchKernelGGL((k_pairsum<T, OP, D, DV, AUX, 4>), grid, dim3(64), 0, st, p); break;
  }
  if (pl.nsplit > 1)
    hipLaunchKernelGGL((k_reduce_splits<OP, D, DV>), dim3((unsigned)((p.nx + 255) / 256)),
                       dim3(256), 0, st, p);
}

template <typename T, int OP, int D, int DV>
static int launch_aux(hipStream_t st, const KParams &p, const Plan &pl) {
  if constexpr (OP == OP_GRADDIM) {
    switch (p.dim) {
      case 0: launch_one<T, OP, D, DV, 0>(st, p, pl); break;
      case 1: launch_one<T, OP, D, DV, 1>(st, p, pl); break;
      default:
        if constexpr (D == 3) launch_one<T, OP, D, DV, 2>(st, p, pl);
        break;
    }
  } else {
    launch_one<T, OP, D, DV, 0>(st, p, pl);
  }
  return DFM_OK;
}

template <typename T, int OP, int D>
static int launch_dv(hipStream_t st, int DV, const KParams &p, const Plan &pl) {
  if constexpr (OP == OP_GRADDIFF) {
    if (DV != D) return set_error(DFM_ERR_DIM, "GradDiff needs dimvect == dimpoint");
    return launch_aux<T, OP, D, D>(st, p, pl);
  } else {
    switch (DV) {
      case 1: return launch_aux<T, OP, D, 1>(st, p, pl);
      case 2: return launch_aux<T, OP, D, 2>(st, p, pl);
      case 3: return launch_aux<T, OP, D, 3>(st, p, pl);
      case 4: return launch_aux<T, OP, D, 4>(st, p, pl);
      case 5: return launch_aux<T, OP, D, 5>(st, p, pl);
      case 6: return launch_aux<T, OP, D, 6>(st, p, pl);
    }
  }
  return set_error(DFM_ERR_DIM, "unsupported dimvect %d", DV);
}

template <typename T, int OP>
static int launch_d(hipStream_t st, int D, int DV, const KParams &p, const Plan &pl) {
  if (D == 2) return launch_dv<T, OP, 2>(st, DV, p, pl);
  if (D == 3) return launch_dv<T, OP, 3>(st, DV, p, pl);
  return set_error(DFM_ERR_DIM, "unsupported dimpoint %d", D);
}

template <typename T>
static int launch_op(hipStream_t st, int op, int D, int DV, const KParams &p, const Plan &pl) {
  switch (op) {
    case OP_CONV: return launch_d<T, OP_CONV>(st, D, DV, p, pl);
    case OP_GRAD1: return launch_d<T, OP_GRAD1>(st, D, DV, p, pl);
    case OP_CONVGRAD1: return launch_d<T, OP_CONVGRAD1>(st, D, DV, p, pl);
    case OP_VARLIN: return launch_d<T, OP_VARLIN>(st, D, DV, p, pl);
    case OP_GRADDIM: return launch_d<T, OP_GRADDIM>(st, D, DV, p, pl);
    case OP_GRADDIFF: return launch_d<T, OP_GRADDIFF>(st, D, DV, p, pl);
  }
  return set_error(DFM_ERR_ARG, "unknown op %d", op);
}

int launch_pairsum(hipStream_t st, int op, int prec, int D, int DV, KParams p,
                   const Plan &pl, char *work) {
  const double sigma = p.sigma;
  p.s = prec == DFM_FP64 ? Prec<double>::scale(sigma) : Prec<float>::scale(sigma);
  const double s2 = sigma * sigma;
  p.cg = -2.0 / (s2 * p.s);
  p.ch1 = 2.0 / s2;
  p.ch2 = -4.0 / (s2 * s2 * p.s * p.s);
  p.nsplit = pl.nsplit;
  p.split_len = pl.split_len;
  p.perm = nullptr;
  p.partial = pl.nsplit > 1 ? (double *)(work + pl.off_partial) : nullptr;
  if (prec == DFM_FP32C) {
    int rc = sort_targets(st, p.x, p.nx, D, work, pl);
    if (rc != DFM_OK) return rc;
    p.perm = (const uint32_t *)(work + pl.off_perm);
  }
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  if (g_profile.on) {
    if (hipEventCreate(&ev0) != hipSuccess || hipEventCreate(&ev1) != hipSuccess)
      return set_error(DFM_ERR_HIP, "hipEventCreate failed");
    (void)hipEventRecord(ev0, st);
  }
  int rc = prec == DFM_FP64 ? launch_op<double>(st, op, D, DV, p, pl)
                            : launch_op<float>(st, op, D, DV, p, pl);
  if (g_profile.on) {
    (void)hipEventRecord(ev1, st);
    g_profile.events.push_back({ev0, ev1});
  }
  if (rc != DFM_OK) return rc;
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(DFM_ERR_HIP, "kernel launch: %s", hipGetErrorString(e));
  return DFM_OK;
}

}  // namespace dfm
